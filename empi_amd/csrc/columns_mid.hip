// Column kernels of the classes Q = 128 .. 1024 (every mode and wave variant).
#include "column_kernel.hpp"

namespace empi {

ColumnLauncher column_launcher_mid(int lgQ, int mode, int variant, bool low) {
    switch (lgQ) {
    case 7: return launcher_for<128>(mode, variant, low);
    case 8: return launcher_for<256>(mode, variant, low);
    case 9: return launcher_for<512>(mode, variant, low);
    case 10: return launcher_for<1024>(mode, variant, low);
    default: return nullptr;
    }
}

}  // namespace empi
