// FFT-size classes of the column kernels (Q = 2 .. 8192 complex points per
// slot), their workgroup geometry, modes and tuning variants, and the launch
// table the host plan dispatches through.  The kernels are compiled in three
// translation units by class range; each exports one launcher lookup.
#pragma once

#include <hip/hip_runtime.h>

#include "dictionary.hpp"
#include "kernels.hpp"

namespace empi {

constexpr int kMaxQ = kMaxSubTransformMulti;   // 8192: multichannel modes only
constexpr int kMinLgQ = 1;    // Q = 2
constexpr int kMaxLgQ = 13;   // Q = 8192
constexpr int kNumClasses = kMaxLgQ - kMinLgQ + 1;
// The column kernels count re-evaluated entries into one copy of the
// per-class counters per XCD (workgroup b runs on XCD b mod 8), so that
// same-address atomics from different XCDs do not queue behind each other;
// the host sums the copies.
constexpr int kCounterCopies = 8;
constexpr int kCounterWords = 2 * kNumClasses * kCounterCopies;

// Workgroup geometry of class Q: NSL slots of Q complex points in LDS (4096
// points in at most 512 slots up to Q = 1024, two slots from Q = 2048 on),
// 16 points per thread (32 at Q = 8192, whose 16,384 points would otherwise
// take 1024 threads at 128 registers each).  Shared by the kernels and the
// host's grid and tile arithmetic.
constexpr int kClassPoints = 4096;
constexpr int class_slots(int Q) {
    return Q >= 2048 ? 2 : (kClassPoints / Q > 512 ? 512 : kClassPoints / Q);
}
constexpr int class_points_per_thread(int Q) { return Q >= 8192 ? 32 : 16; }
constexpr int class_threads(int Q) { return class_slots(Q) * Q / class_points_per_thread(Q); }

template <int Q>
struct Geo {
    static constexpr int NSL = class_slots(Q);
    static constexpr int NPTS = NSL * Q;
    static constexpr int EPT = class_points_per_thread(Q);
    static constexpr int NT = NPTS / EPT;
    static constexpr int NCELL = NSL / 2;
};

// column kernel modes: one channel; channel sum of products (MMP3 and
// multichannel workspaces); MMP1 constraint (three accumulators per bin)
enum { kModeSingle = 0, kModeSum = 1, kModeMmp1 = 2 };
// waves per SIMD the register allocator must allow in the multichannel modes
// (the default; tuning variants add one or two)
template <int Q, int MODE>
constexpr int column_waves() { return MODE == kModeSum ? 2 : 1; }

// Tuning variants of the column kernels: waves per SIMD the register
// allocator must allow (single channel 4, 3, 2; multichannel the mode's
// column_waves + 0, 1, 2), times two persistent grid sizes.
constexpr int kWaveVariants = 3;
constexpr int kGridVariants = 2;
constexpr int kTuneChoices = kWaveVariants * kGridVariants;
template <int Q, int V>
constexpr int variant_waves() { return V == 0 ? 4 : V == 1 ? 3 : 2; }

typedef void (*ColumnLauncher)(int grid, hipStream_t s, const ClassDev* classes, int cls, const ScaleDev* scales,
                               const UnitDev* units, const UnitState* state, const float* residual,
                               unsigned long long* cellbest, unsigned long long* tilebest,
                               unsigned long long* counter, float* celllow);

// launcher of class lgQ (log2 Q) in mode `mode`; `variant` selects the wave
// variant of the single-channel kernels, `low` the single-channel kernel that
// also writes the cells' low-bin bests (celllow); nullptr for a class that
// has no kernel in that mode (Q = 8192 single-channel).  Q = 2..64 /
// 128..1024 / 2048..8192.
ColumnLauncher column_launcher_small(int lgQ, int mode, int variant, bool low);
ColumnLauncher column_launcher_mid(int lgQ, int mode, int variant, bool low);
ColumnLauncher column_launcher_large(int lgQ, int mode, int variant, bool low);

}  // namespace empi
