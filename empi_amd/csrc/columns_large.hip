// Column kernels of the classes Q = 2048 .. 4096 (every mode and wave variant).
#include "column_kernel.hpp"

namespace empi {

ColumnLauncher column_launcher_large(int lgQ, int mode, int variant) {
    switch (lgQ) {
    case 11: return launcher_for<2048>(mode, variant);
    case 12: return launcher_for<4096>(mode, variant);
    default: return nullptr;
    }
}

}  // namespace empi
