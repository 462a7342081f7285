// Column kernels of the classes Q = 2 .. 64 (every mode and wave variant).
#include "column_kernel.hpp"

namespace empi {

ColumnLauncher column_launcher_small(int lgQ, int mode, int variant, bool low) {
    switch (lgQ) {
    case 1: return launcher_for<2>(mode, variant, low);
    case 2: return launcher_for<4>(mode, variant, low);
    case 3: return launcher_for<8>(mode, variant, low);
    case 4: return launcher_for<16>(mode, variant, low);
    case 5: return launcher_for<32>(mode, variant, low);
    case 6: return launcher_for<64>(mode, variant, low);
    default: return nullptr;
    }
}

}  // namespace empi
