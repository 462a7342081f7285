// Column kernels of the classes Q = 2 .. 64 (every mode and wave variant).
#include "column_kernel.hpp"

namespace empi {

ColumnLauncher column_launcher_small(int lgQ, int mode, int variant) {
    switch (lgQ) {
    case 1: return launcher_for<2>(mode, variant);
    case 2: return launcher_for<4>(mode, variant);
    case 3: return launcher_for<8>(mode, variant);
    case 4: return launcher_for<16>(mode, variant);
    case 5: return launcher_for<32>(mode, variant);
    case 6: return launcher_for<64>(mode, variant);
    default: return nullptr;
    }
}

}  // namespace empi
