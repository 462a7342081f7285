// Column kernels of the classes Q = 2048 .. 8192 (every mode and wave variant;
// Q = 8192 in the multichannel modes only).
#include "column_kernel.hpp"

namespace empi {

ColumnLauncher column_launcher_large(int lgQ, int mode, int variant, bool low) {
    switch (lgQ) {
    case 11: return launcher_for<2048>(mode, variant, low);
    case 12: return launcher_for<4096>(mode, variant, low);
    case 13: return (low ? nullptr : multichannel_launcher_for<8192>(mode, variant));
    default: return nullptr;
    }
}

}  // namespace empi
