// MI355X (gfx950) Matching Pursuit engine: unit and region kernels, batched
// plan and C ABI.
//
// Hot path (one MP iteration of every unit of a batch):
//   k_select    one workgroup per unit: argmax over the unit's cell bests,
//               FP64 refit of the chosen column, atom parameters, FP64
//               residual update and residual energy / stop test.
//   k_plan      per FFT-size class: the work list of tiles the atom touched.
//   k_columns<Q> (column_kernel.hpp, compiled in columns_*.hip) for every
//               class: re-evaluates the touched map cells and their bests.
// References (paths in the reference tree): gabor.cpp:75-133,143-209,
// 279-354; classes.cpp:26-87; mmp.cpp:9-45.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "columns.hpp"
#include "device_math.hpp"
#include "dictionary.hpp"
#include "empi_hip.h"
#include "kernels.hpp"
#include "mlp_internal.hpp"

using namespace empi;

// Scratch of the allocation-free entry point (empi_decompose_noalloc): every
// device buffer of its plan is carved from this array, one copy per GPU,
// held until the process ends.  Host code never passes the array's name to
// the device: it uses the address hipGetSymbolAddress returns per device.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

#define EMPI_HIP(call)                                  \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) {                         \
            std::fprintf(stderr, "empi_hip: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);           \
            return EMPI_DEVICE;                         \
        }                                               \
    } while (0)

namespace {

constexpr int kSelectThreads = 256;

// Affected tiles of every job of one class (one workgroup per class), one
// work item per job: columns with |t - tA| <= 3 (s + sA), widened by one
// column each side (the column kernel applies the exact test per cell), or
// every column for the full build.
constexpr int kPlanThreads = 1024;

// exclusive prefix of the items' tile counts; work[count] = total.  Wave w
// takes items w span .. (w+1) span - 1 (span a multiple of 64) in rounds of
// 64 consecutive items, so every load is coalesced, and scans each round
// across its lanes on top of the running total
__device__ void scan_items(const ClassDev& cd, int n, int* part) {
    constexpr int NW = kPlanThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int span = ((n + NW - 1) / NW + 63) & ~63;
    const int j0 = min(n, wv * span), j1 = min(n, j0 + span);
    int sum = 0;
    for (int j = j0 + lane; j < j1; j += 64) sum += cd.items[j].tiles;
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) part[wv] = sum;
    __syncthreads();
    int run = 0;
    for (int i = 0; i < wv; ++i) run += part[i];
    for (int j = j0; j < j1; j += 64) {
        const int t = j + lane < j1 ? cd.items[j + lane].tiles : 0;
        int incl = t;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (j + lane < j1) cd.work[j + lane] = run + incl - t;
        run += __shfl(incl, 63);
    }
    if (tid == 0) {
        int total = 0;
        for (int i = 0; i < NW; ++i) total += part[i];
        cd.work[n] = total;
    }
}

__global__ __launch_bounds__(kPlanThreads) void k_plan(const ClassDev* __restrict__ classes,
                                                       const UnitState* __restrict__ state,
                                                       const ScaleDev* __restrict__ scales, int full) {
    __shared__ int part[kPlanThreads / 64];   // per-wave totals of scan_items
    const ClassDev cd = classes[blockIdx.x];
    const int n = cd.njobs;
    for (int j = threadIdx.x; j < n; j += kPlanThreads) {
        const Job job = cd.jobs[j];
        const UnitState st = state[job.unit];
        const ScaleDev sd = scales[job.scale];
        WorkItem it{};
        it.job = j;
        it.atom = full ? 0 : 1;
        it.region = -1;
        it.tlo = 0;
        it.thi = sd.T - 1;
        if (st.active) {
            if (full) {
                it.tiles = job.tiles;
            } else {
                const double R = (sd.s + st.sA) * kGaussHalfWidth;
                const double scale = (sd.T - 1) / sd.tMax;
                long lo = static_cast<long>(ceil((st.tA - R) * scale)) - 1;
                long hi = static_cast<long>(floor((st.tA + R) * scale)) + 1;
                lo = lo < 0 ? 0 : lo;
                hi = hi > sd.T - 1 ? sd.T - 1 : hi;
                if (lo <= hi) {
                    const long c0 = lo * sd.G, c1 = hi * sd.G + sd.G - 1;
                    it.tile = static_cast<int>(c0 / cd.ncell);
                    it.tiles = static_cast<int>(c1 / cd.ncell) - it.tile + 1;
                    it.tlo = static_cast<int>(lo);
                    it.thi = static_cast<int>(hi);
                }
            }
        }
        cd.items[j] = it;
    }
    if (threadIdx.x == 0) *cd.count = n;
    __syncthreads();
    scan_items(cd, n, part);
}

// prefix over items appended by k_region_items (one workgroup per class)
__global__ __launch_bounds__(kPlanThreads) void k_scan_items(const ClassDev* __restrict__ classes) {
    __shared__ int part[kPlanThreads / 64];   // per-wave totals of scan_items
    const ClassDev cd = classes[blockIdx.x];
    const int n = min(*cd.count, cd.max_items);
    scan_items(cd, n, part);
}

// ------------------------------------------------------------ unit kernels

template <int NT>
__device__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// residual energy sum_c (sum_n x^2) / fs (base.hpp:39-58)
template <int NT>
__device__ double unit_energy(const double* x, int C, int N, double fs, double* red) {
    double total = 0.0;
    for (int c = 0; c < C; ++c) {
        double part = 0.0;
        for (int n = threadIdx.x; n < N; n += NT) {
            const double v = x[static_cast<long long>(c) * N + n];
            part += v * v;
        }
        total += block_sum<NT>(part, red) / fs;
    }
    return total;
}

// FP32 mirror of a unit's residual, scaled by the unit's power of two
__global__ __launch_bounds__(kSelectThreads) void k_mirror(const UnitDev* __restrict__ units,
                                                            const UnitState* __restrict__ state,
                                                            const double* __restrict__ residual,
                                                            float* __restrict__ res32) {
    const UnitDev ud = units[blockIdx.x];
    const float xs = state[blockIdx.x].xscale;
    const long long n = static_cast<long long>(ud.C) * ud.N;
    for (long long i = threadIdx.x; i < n; i += kSelectThreads)
        res32[ud.res_off + i] = static_cast<float>(residual[ud.res_off + i] * xs);
}

// residual <- input (FP64), energy, input scaling; one workgroup per unit
__global__ __launch_bounds__(kSelectThreads) void k_init(const UnitDev* __restrict__ units,
                                                          UnitState* __restrict__ state,
                                                          const float* __restrict__ input,
                                                          double* __restrict__ residual) {
    __shared__ double red[kSelectThreads];
    const UnitDev ud = units[blockIdx.x];
    double* x = residual + ud.res_off;
    for (int c = 0; c < ud.C; ++c) {
        for (int n = threadIdx.x; n < ud.N; n += kSelectThreads) {
            double v;
            if (ud.in_sum) {   // mmp.cpp:24-32: channel 0, then += channels 1..C-1
                v = input[ud.in_off + n];
                for (int k = 1; k < ud.in_sum; ++k) v += static_cast<double>(input[ud.in_off + static_cast<long long>(k) * ud.N + n]);
            } else {
                v = input[ud.in_off + static_cast<long long>(ud.in_channel + c) * ud.N + n];
            }
            x[static_cast<long long>(c) * ud.N + n] = v;
        }
    }
    __syncthreads();
    const double E = unit_energy<kSelectThreads>(x, ud.C, ud.N, ud.fs, red);
    if (threadIdx.x == 0) {
        UnitState s{};
        s.totalEnergy = E;
        s.residEnergy = E;
        s.iteration = 0;
        s.active = E != 0.0;
        s.status = E != 0.0 ? EMPI_OK : EMPI_SIGNAL_IS_EMPTY;
        // FP32 maps work on x * 2^e with the RMS near 1 (exact scaling)
        const double ms = E * ud.fs / (static_cast<double>(ud.C) * ud.N);
        int ex = 0;
        if (ms > 0) std::frexp(sqrt(ms), &ex);
        ex = max(-100, min(100, ex));
        s.xscale = ldexpf(1.0f, -ex);
        state[blockIdx.x] = s;
        ud.energies[0] = E;
    }
}

// FP64 value of map entry (k, f, t) of every channel (gabor.cpp:279-303 at one bin)
template <int NT>
__device__ void fit_column(const UnitDev& ud, const ScaleDev& sd, const double* x, int f, int t,
                           double* red, double2* out) {
    const double fs = sd.fs, s = sd.s;
    const double t0 = t * sd.tMax / (sd.T - 1);
    const double hw = kGaussHalfWidth * s;
    long l = lrint((t0 - hw) * fs), r = lrint((t0 + hw) * fs);
    const int iL = static_cast<int>(l > 0 ? l : 0);
    const int iR = static_cast<int>(r < sd.N - 1 ? r : sd.N - 1);
    const int W = min(iR - iL + 1, sd.Nfft);
    const double t0fixed = t0 - iL / fs;
    const double gw = sqrt(M_PI) / s;
    const double ga = gw * (1.0 / fs), gb = gw * t0fixed;
    const double norm = 1.0 / fs;
    const double ang = (f * fs / sd.Nfft) * (2 * M_PI * t0fixed);
    const double cr = cos(ang), si = sin(ang);
    const unsigned mask = static_cast<unsigned>(sd.Nfft - 1);
    for (int c0 = 0; c0 < ud.C; c0 += 4) {
        double pr[4] = {0, 0, 0, 0}, pi[4] = {0, 0, 0, 0};
        for (int n = threadIdx.x; n < W; n += NT) {
            const double xw = ga * n - gb;
            const double g = exp(-xw * xw);
            const unsigned idx = (static_cast<unsigned>(f) * static_cast<unsigned>(n)) & mask;
            double sn, cs;
            sincospi(-2.0 * static_cast<double>(idx) / sd.Nfft, &sn, &cs);
            const double gr = g * cs, gi = g * sn;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (c0 + q < ud.C) {
                    const double xv = x[static_cast<long long>(c0 + q) * sd.N + iL + n];
                    pr[q] += xv * gr;
                    pi[q] += xv * gi;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double Xr = block_sum<NT>(pr[q], red);
            const double Xi = block_sum<NT>(pi[q], red);
            if (threadIdx.x == 0 && c0 + q < ud.C) {
                const double vr = norm * Xr, vi = norm * Xi;
                out[c0 + q] = make_double2(vr * cr - vi * si, vr * si + vi * cr);
            }
        }
    }
    __syncthreads();
}

// Atom parameters from the (constrained) values (gabor.cpp:111-131, mmp.cpp:9-21).
__device__ void atom_params(const UnitDev& ud, const ScaleDev& sd, int k, int f, int t, double2* v,
                            empi_atom* out, int it) {
    const double fs = sd.fs, s = sd.s;
    const double normComplex = sqrt(M_SQRT2 / s);
    const double freqNyquist = 0.5 * fs;
    const double fv = f * fs / sd.Nfft;
    const bool isHigh = fv > 0.5 * freqNyquist;
    const double f4 = isHigh ? (freqNyquist - fv) : fv;
    const double expFactor = exp(-2 * M_PI * s * s * f4 * f4);
    const double tt = t * sd.tMax / (sd.T - 1);
    if (ud.constraint && !isHigh) {
        double dr = 0.0, di = 0.0;
        for (int c = 0; c < ud.C; ++c) {
            const double re = v[c].x, im = v[c].y;
            dr += re * re - im * im;
            di += re * im + im * re;
        }
        const double angle = 0.5 * atan2(di, dr);
        const double ca = cos(angle), sa = sin(angle);
        for (int c = 0; c < ud.C; ++c) {
            const double re = v[c].x, im = v[c].y;
            const double rr = hypot(re, im) * cos(atan2(im, re) - angle);
            v[c] = make_double2(rr * ca, rr * sa);
        }
    }
    for (int c = 0; c < ud.C; ++c) {
        const double re = v[c].x, im = v[c].y;
        const double phase = atan2(im, re);
        const double phase4 = isHigh ? (2 * M_PI * freqNyquist * tt - phase) : phase;
        const double normReal = M_SQRT2 * normComplex / sqrt(1.0 + cos(2 * phase4) * expFactor);
        const double mag = hypot(re, im);
        empi_atom a;
        a.params[0] = mag * normReal * sqrt(fs);
        a.params[1] = mag * normReal * normReal;
        a.params[2] = tt * fs;
        a.params[3] = s * fs;
        a.params[4] = fv * 2.0 / fs;
        a.params[5] = phase;
        a.scale = k;
        a.freq = f;
        a.time = t;
        a.reserved = 0;
        out[static_cast<long long>(c) * ud.atom_cstride + it] = a;
    }
}

// x -= atom waveform (gabor.cpp:143-187 with fit = false); all threads
template <int NT>
__device__ void subtract_waveform(const empi_atom& a, double fs, double* x, int N, float* x32, float xscale) {
    const double sA = a.params[3] / fs;
    const double fA = a.params[4] * fs / 2.0;
    const double tA = a.params[2] / fs;
    const double amplitude = a.params[1];
    const double omega = 2.0 * M_PI * fA;
    const double phase = a.params[5];
    const double hw = kGaussHalfWidth * sA;
    long l = lrint((tA - hw) * fs), r = lrint((tA + hw) * fs);
    const int iL = static_cast<int>(l > 0 ? l : 0);
    const int iR = static_cast<int>(r < N - 1 ? r : N - 1);
    const int Nw = 1 + iR - iL;
    const double gw = sqrt(M_PI) / sA;
    const double ga = gw * (1.0 / fs), gb = gw * (tA - iL / fs);
    for (int i = threadIdx.x; i < Nw; i += NT) {
        const double t = (i + iL) / fs - tA;
        double w = amplitude * cos(omega * t + phase);
        const double xw = ga * i - gb;
        w *= exp(-xw * xw);
        const double y = x[iL + i] - w;
        x[iL + i] = y;
        x32[iL + i] = static_cast<float>(y * xscale);
    }
}

// proposal regions (see the proposal section below)
constexpr int kBlock = 256;          // samples per region block
constexpr int kFeatBins = 64;
constexpr int kFeatDim = 80;         // network input
constexpr int kRegionThreads = 256;

struct RegionState {
    float rmax;                      // exact max product at the last evaluation
    float rbound;                    // bound on the current max (dirty regions)
    float rscore;                    // network score (predicted log max / residual)
    int flags;                       // bit 0 dirty, bit 1 evaluate now
    unsigned long long rbest;        // packed (product bits, ~bin) of the best cell
    int rcell;                       // best cell index within the scale
    float racc;                      // bound on |du| accumulated since the evaluation, all bins
    float rlow;                      // max over bins 1 <= j < JB of |u|^2 (the product at E = 0) at the evaluation
    float racc_lo;                   // bound on |du| accumulated since the evaluation, bins 1 <= j < JB
    int pad;
};

// mode bit 0: find the best atom; bit 1: subtract it and test the stop rule.
// split: multichannel units only record the chosen entry (sel_*); k_fit_mc,
// k_params_mc, k_sub_mc and k_stop_mc finish their iteration one workgroup
// per channel.
template <int NT>
__global__ __launch_bounds__(NT) void k_select(const UnitDev* __restrict__ units, UnitState* __restrict__ state,
                                               const ScaleDev* __restrict__ scales,
                                               const UnitScale* __restrict__ uscales,
                                               const unsigned long long* __restrict__ tilebest,
                                               double* __restrict__ residual, float* __restrict__ res32,
                                               int mode, const RegionState* __restrict__ regions, int split) {
    __shared__ double red[NT];
    __shared__ float redp[NT];
    __shared__ unsigned long long redk[NT];
    __shared__ empi_atom chosen;
    const int u = blockIdx.x;
    const int tid = threadIdx.x;
    UnitState st = state[u];
    if (!st.active) return;
    const UnitDev ud = units[u];
    const int it = st.iteration;

    // 1. argmax over the tile bests (over the clean regions' bests with the
    // proposal): product desc, then scale, bin, time asc
    SelBest b{-1.0f, ~0ull};
    if (regions) {
        for (int r = tid; r < ud.L * ud.B; r += NT) {
            const RegionState R = regions[ud.reg_off + r];
            if ((R.flags & 1) || R.rbest == 0ull) continue;
            const int l = r / ud.B;
            const int G = scales[ud.scale0 + l].G;
            const SelBest c = sel_key(__uint_as_float(static_cast<unsigned>(R.rbest >> 32)), l,
                                      0xFFFFFFFFu - static_cast<unsigned>(R.rbest & 0xFFFFFFFFu), R.rcell / G);
            if (sel_better(c, b)) b = c;
        }
    }
    for (int l = 0; l < (regions ? 0 : ud.L); ++l) {
        const UnitScale us = uscales[ud.us_off + l];
        const int G = scales[ud.scale0 + l].G;
        for (int i = tid; i < us.tiles; i += NT) {
            const SelBest c = sel_of_tile(tilebest[us.tile_off + i], l, i, us.ncell, G, kTileCellBits);
            if (sel_better(c, b)) b = c;
        }
    }
    redp[tid] = b.p;
    redk[tid] = b.key;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            SelBest o{redp[tid + s], redk[tid + s]};
            SelBest m{redp[tid], redk[tid]};
            if (sel_better(o, m)) {
                redp[tid] = o.p;
                redk[tid] = o.key;
            }
        }
        __syncthreads();
    }
    const float bp = redp[0];
    const unsigned long long bk = redk[0];
    __syncthreads();
    if (!(bp > 0.0f)) {
        if (tid == 0) {
            st.status = EMPI_ENERGY_MAP_IS_EMPTY;
            st.active = 0;
            state[u] = st;
        }
        return;
    }
    const int k = static_cast<int>(bk >> 52);
    const int f = static_cast<int>((bk >> 26) & ((1ull << 26) - 1));
    const int t = static_cast<int>(bk & ((1ull << 26) - 1));
    if (split && ud.C > 1) {
        if (tid == 0) {
            st.sel_k = k;
            st.sel_f = f;
            st.sel_t = t;
            state[u] = st;
        }
        return;
    }
    const ScaleDev sd = scales[ud.scale0 + k];
    double* x = residual + ud.res_off;

    // 2. FP64 refit of the chosen column, 3. parameters
    fit_column<NT>(ud, sd, x, f, t, red, ud.fitv);
    if (tid == 0) atom_params(ud, sd, k, f, t, ud.fitv, ud.atoms, it);
    __syncthreads();
    if (!(mode & 2)) {
        if (tid == 0) {
            st.iteration = it + 1;
            state[u] = st;
        }
        return;
    }
    // classes.cpp:57-59: the last atom is not subtracted
    if (it + 1 >= ud.iterMax) {
        if (tid == 0) {
            st.iteration = it + 1;
            st.active = 0;
            state[u] = st;
        }
        return;
    }
    // 4. residual update per channel, 5. energy and stop test (classes.cpp:61-70)
    for (int c = 0; c < ud.C; ++c) {
        if (tid == 0) chosen = ud.atoms[static_cast<long long>(c) * ud.atom_cstride + it];
        __syncthreads();
        subtract_waveform<NT>(chosen, ud.fs, x + static_cast<long long>(c) * ud.N, ud.N,
                              res32 + ud.res_off + static_cast<long long>(c) * ud.N, st.xscale);
        __syncthreads();
    }
    const double E = unit_energy<NT>(x, ud.C, ud.N, ud.fs, red);
    if (tid == 0) {
        const empi_atom& a = ud.atoms[it];
        st.iteration = it + 1;
        st.residEnergy = E;
        st.tA = a.params[2] / ud.fs;
        st.sA = a.params[3] / ud.fs;
        ud.energies[it + 1] = E;
        if (E / st.totalEnergy <= ud.residualEnergy) st.active = 0;
        state[u] = st;
    }
}

// ------------------------------------------------------------ multichannel units
//
// The rest of an iteration of the multichannel (MMP1/MMP3) units after
// k_select's argmax, over the plan's list of such units: the FP64 refit and
// the subtraction take one workgroup per (channel, unit) instead of one per
// unit; the channel energies are summed in the same order as unit_energy.

// FP64 map value of the chosen entry, channel blockIdx.x of unit mc[blockIdx.y]
template <int NT>
__global__ __launch_bounds__(NT) void k_fit_mc(const UnitDev* __restrict__ units, const UnitState* __restrict__ state,
                                               const ScaleDev* __restrict__ scales, const int* __restrict__ mc,
                                               const double* __restrict__ residual) {
    __shared__ double red[NT];
    const int u = mc[blockIdx.y], c = blockIdx.x;
    const UnitDev ud = units[u];
    const UnitState st = state[u];
    if (c >= ud.C || !st.active) return;
    const ScaleDev sd = scales[ud.scale0 + st.sel_k];
    UnitDev one = ud;
    one.C = 1;
    fit_column<NT>(one, sd, residual + ud.res_off + static_cast<long long>(c) * ud.N, st.sel_f, st.sel_t, red,
                   ud.fitv + c);
}

// atom parameters of every channel (MMP1 constraint included); the atom of
// the last iteration is not subtracted (classes.cpp:57-59)
__global__ void k_params_mc(const UnitDev* __restrict__ units, UnitState* __restrict__ state,
                            const ScaleDev* __restrict__ scales, const int* __restrict__ mc, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int u = mc[i];
    const UnitDev ud = units[u];
    UnitState st = state[u];
    if (!st.active) return;
    const ScaleDev sd = scales[ud.scale0 + st.sel_k];
    atom_params(ud, sd, st.sel_k, st.sel_f, st.sel_t, ud.fitv, ud.atoms, st.iteration);
    if (st.iteration + 1 >= ud.iterMax) {
        st.iteration += 1;
        st.active = 0;
        state[u] = st;
    }
}

// residual update of channel blockIdx.x and its energy sum_n x^2
template <int NT>
__global__ __launch_bounds__(NT) void k_sub_mc(const UnitDev* __restrict__ units, const UnitState* __restrict__ state,
                                               const int* __restrict__ mc, double* __restrict__ residual,
                                               float* __restrict__ res32, double* __restrict__ epart, int cmax) {
    __shared__ double red[NT];
    __shared__ empi_atom a;
    const int u = mc[blockIdx.y], c = blockIdx.x;
    const UnitDev ud = units[u];
    const UnitState st = state[u];
    if (c >= ud.C || !st.active) return;
    if (threadIdx.x == 0) a = ud.atoms[static_cast<long long>(c) * ud.atom_cstride + st.iteration];
    __syncthreads();
    const long long off = ud.res_off + static_cast<long long>(c) * ud.N;
    subtract_waveform<NT>(a, ud.fs, residual + off, ud.N, res32 + off, st.xscale);
    __syncthreads();
    double part = 0.0;
    for (int n = threadIdx.x; n < ud.N; n += NT) {
        const double v = residual[off + n];
        part += v * v;
    }
    const double e = block_sum<NT>(part, red);
    if (threadIdx.x == 0) epart[static_cast<long long>(blockIdx.y) * cmax + c] = e;
}

// residual energy and the stop test (classes.cpp:61-70)
__global__ void k_stop_mc(const UnitDev* __restrict__ units, UnitState* __restrict__ state, const int* __restrict__ mc,
                          int count, const double* __restrict__ epart, int cmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int u = mc[i];
    const UnitDev ud = units[u];
    UnitState st = state[u];
    if (!st.active) return;
    double E = 0.0;
    for (int c = 0; c < ud.C; ++c) E += epart[static_cast<long long>(i) * cmax + c] / ud.fs;
    const int it = st.iteration;
    const empi_atom& a = ud.atoms[it];
    st.iteration = it + 1;
    st.residEnergy = E;
    st.tA = a.params[2] / ud.fs;
    st.sA = a.params[3] / ud.fs;
    ud.energies[it + 1] = E;
    if (E / st.totalEnergy <= ud.residualEnergy) st.active = 0;
    state[u] = st;
}

// ------------------------------------------------------------ proposal
//
// Regions: (scale, 256-sample block) of one unit.  A region is clean while
// every cell in it holds its exact product; subtracting an atom makes the
// regions it overlaps (the reference's own rule) dirty.  Per iteration: the
// network scores the dirty regions whose bound exceeds the best clean product
// (the candidates) and the top-K are evaluated exactly; then
// every dirty region whose bound exceeds P* / (1 - delta), P* the best clean
// product, is evaluated too (the exact fallback).  A region's bound is the
// smaller of a Cauchy-Schwarz bound, csnorm * (the largest residual energy
// under one of its columns' windows), which holds for every cell, and the
// stale maximum grown by the accumulated atom-overlap bound (k_mark_dirty).

// affected time-index range of scale sd for the unit's last atom (widened by one)
__device__ inline bool atom_range(const ScaleDev& sd, const UnitState& st, int* lo, int* hi) {
    const double R = (sd.s + st.sA) * kGaussHalfWidth;
    const double scale = (sd.T - 1) / sd.tMax;
    long l = static_cast<long>(ceil((st.tA - R) * scale)) - 1;
    long h = static_cast<long>(floor((st.tA + R) * scale)) + 1;
    l = l < 0 ? 0 : l;
    h = h > sd.T - 1 ? sd.T - 1 : h;
    *lo = static_cast<int>(l);
    *hi = static_cast<int>(h);
    return l <= h;
}

// region bests from the tile bests and the low-bin bound seeds from the
// cells' low-bin bests: all regions, or those flagged "evaluate"; one wave
// per region.  A region is whole tiles of its scale (its first column is
// rounded to a tile boundary, and it ends where the next region starts or at
// the scale's end), and a tile key orders its cells as the region does, by
// cell best and then by the smallest cell, so the region's best is the best
// of its tiles' bests compared by the cell best they carry
// (pack_best(product, bin), recovered exactly from the key) and then by cell
__global__ __launch_bounds__(kRegionThreads) void k_regions_reduce(const UnitDev* __restrict__ units,
                                                                   const UnitState* __restrict__ state,
                                                                   const ScaleDev* __restrict__ scales,
                                                                   const RegionDev* __restrict__ regtab,
                                                                   const UnitScale* __restrict__ uscales,
                                                                   const unsigned long long* __restrict__ tilebest,
                                                                   const float* __restrict__ celllow,
                                                                   RegionState* __restrict__ regions, int all) {
    const int u = blockIdx.y;
    if (!all && !state[u].active) return;
    const UnitDev ud = units[u];
    const int waves = kRegionThreads / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nreg = ud.L * ud.B;
    for (int r = blockIdx.x * waves + wave; r < nreg; r += gridDim.x * waves) {
        RegionState* R = &regions[ud.reg_off + r];
        if (!all && !(R->flags & 2)) continue;
        const int l = r / ud.B;
        const RegionDev rt = regtab[ud.regtab_off + r];
        const ScaleDev& sd = scales[ud.scale0 + l];
        unsigned long long best = 0ull;
        int bcell = 0x7FFFFFFF;
        float low = 0.0f;
        if (rt.tlo <= rt.thi) {
            const long long base = ud.cell_off + sd.cells_before;
            const int c0 = rt.tlo * sd.G, c1 = rt.thi * sd.G + sd.G - 1;
            const UnitScale us = uscales[ud.us_off + l];
            const int ncell = us.ncell;
            for (int t = c0 / ncell + lane; t <= c1 / ncell; t += 64) {
                const unsigned long long k = tilebest[us.tile_off + t];
                const unsigned key = 0xFFFFFFFFu - static_cast<unsigned>(k);
                const unsigned bin = key >> kTileCellBits;
                const int c = t * ncell + static_cast<int>(key & ((1u << kTileCellBits) - 1u));
                const unsigned long long v = (k & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFu - bin);
                if (v > best || (v == best && c < bcell)) {
                    best = v;
                    bcell = c;
                }
            }
            for (int c = c0 + lane; c <= c1; c += 64) low = fmaxf(low, celllow[base + c]);
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const unsigned long long ov = shfl_xor_u64(best, off);
            const int oc = __shfl_xor(bcell, off);
            if (ov > best || (ov == best && oc < bcell)) {
                best = ov;
                bcell = oc;
            }
            low = fmaxf(low, __shfl_xor(low, off));
        }
        if (lane == 0) {
            R->rbest = best;
            R->rmax = __uint_as_float(static_cast<unsigned>(best >> 32));
            R->rcell = bcell;
            R->flags = 0;
            R->racc = 0.0f;
            R->racc_lo = 0.0f;
            R->rlow = low;
        }
    }
}

// Regions the unit's last atom A overlaps (the reference's rule) become
// dirty, and each accumulates a bound r on how far the subtractions since its
// evaluation move u = sqrt(K) v, v the map value (linear in the residual):
// the product is P = |u|^2 h(phi)^2, h = (1 + E cos 2 phi)^(-1/2)
// (gabor.cpp:62-70), so for bin j either
//   (A) sqrt P_new <= (|u_old| + r) / sqrt(1 - E_j), or
//   (B) sqrt P_new <= sqrt P_old + L(E_j) r, L the Lipschitz constant of |u| h
// (k_rows takes the smaller over bins 1 <= j < JB, (B) elsewhere; bins 0 and
// M are real, with a fixed denominator).  r = |a| sqrt(K) |v(g_A)| summed over
// the atoms, g_A the unit atom; by Poisson summation the discrete overlap of
// two sampled Gabor atoms is
//   sqrt(K) |v(g_A)| <= n_A n I / 2 (PG(fA - f) + PG(fA + f)),
//   I = s sA / sqrt(s^2 + sA^2) e^{-pi dt^2 / (s^2 + sA^2)},
//   PG(x) = sum_m e^{-pi sg^2 (x - m fs)^2}, sg^2 = s^2 sA^2 / (s^2 + sA^2),
// dt the distance from tA to the region's column centres, n = sqrt(2 sqrt 2 / s),
// n_A the atom's own norm.  racc takes the maximum over all bins f, racc_lo
// over the low bins, which lie within f_L = (JB - 1) fs / Nfft of 0 and of
// fs / 2: an atom far from both in frequency barely moves them.  The
// frequency decay holds for the untruncated sum over every sample n.  The
// map sums only n in [lo, hi] = [0, N-1] cut to the column's window (3 s)
// and the atom's support (3 sA); the product of the two envelopes is
// e^{-pi dt^2 / (s^2 + sA^2)} e^{-pi (n - c)^2 / sg_n^2}, a Gaussian of width
// sg_n = sg fs samples about c = fs (t sA^2 + tA s^2) / (s^2 + sA^2), so the
// samples cut off add at most I / sg_n times the tail sums
//   sum_{k >= 0} e^{-pi (D + k)^2 / sg_n^2} <= e^{-pi D^2 / sg_n^2} / (1 - e^{-2 pi D / sg_n^2}),
// D = c - lo + 1 below and hi + 1 - c above.  With lrint within 0.5 of its
// argument both D are concave in t, so the region's end columns give their
// minimum.  The envelope bound n_A n I PG(0) (every term in phase) caps the
// result, and is all that remains when a cut reaches the Gaussian's centre.
// 1% and 1e-9 cover rounding.
__global__ __launch_bounds__(kRegionThreads) void k_mark_dirty(const UnitDev* __restrict__ units,
                                                               const UnitState* __restrict__ state,
                                                               const ScaleDev* __restrict__ scales,
                                                               const RegionDev* __restrict__ regtab,
                                                               RegionState* __restrict__ regions) {
    const int u = blockIdx.x;
    const UnitState st = state[u];
    if (!st.active || st.iteration == 0) return;
    const UnitDev ud = units[u];
    const empi_atom& a = ud.atoms[st.iteration - 1];
    const double modulus = fabs(a.params[0]);
    const double nA = modulus > 0.0 ? sqrt(ud.fs) * fabs(a.params[1]) / modulus : 0.0;
    const double fA = a.params[4] * ud.fs * 0.5;
    const double sA = st.sA;
    const double coef = modulus * st.xscale;     // sqrt of the atom's product, FP32-map units
    for (int r = threadIdx.x; r < ud.L * ud.B; r += kRegionThreads) {
        const RegionDev rt = regtab[ud.regtab_off + r];
        if (rt.tlo > rt.thi) continue;
        const ScaleDev& sd = scales[ud.scale0 + r / ud.B];
        int lo, hi;
        if (!(atom_range(sd, st, &lo, &hi) && rt.thi >= lo && rt.tlo <= hi)) continue;
        RegionState* R = &regions[ud.reg_off + r];
        R->flags |= 1;
        const double s = sd.s;
        const double t0 = rt.tlo * sd.tMax / (sd.T - 1), t1 = rt.thi * sd.tMax / (sd.T - 1);
        const double dt = st.tA < t0 ? t0 - st.tA : (st.tA > t1 ? st.tA - t1 : 0.0);
        const double v = s * s + sA * sA;
        const double overlap = s * sA / sqrt(v) * exp(-M_PI * dt * dt / v);
        const double sg2 = s * s * sA * sA / v;
        const double fs = ud.fs, fN = 0.5 * fs;
        // tails cut off below lo and above hi, relative to I, at the region's end columns
        const double sgn = sqrt(sg2) * fs, sgn2 = sgn * sgn;
        // lrint(x) <= x + 0.5: lo from above, hi from below
        const double loA = fmax(0.0, (st.tA - kGaussHalfWidth * sA) * fs + 0.5);
        const double hiA = fmin(ud.N - 1.0, (st.tA + kGaussHalfWidth * sA) * fs - 0.5);
        auto tail = [&](double D) {
            return D > 0.0 ? exp(-M_PI * D * D / sgn2) / (1.0 - exp(-2.0 * M_PI * D / sgn2)) : HUGE_VAL;
        };
        double cut_lo = 0.0, cut_hi = 0.0;
        for (int e = 0; e < 2; ++e) {
            const double t = e ? t1 : t0;
            const double c = fs * (t * sA * sA + st.tA * s * s) / v;
            const double lo = fmax(loA, (t - kGaussHalfWidth * s) * fs + 0.5);
            const double hi = fmin(hiA, (t + kGaussHalfWidth * s) * fs - 0.5);
            cut_lo = fmax(cut_lo, tail(c - lo + 1.0));
            cut_hi = fmax(cut_hi, tail(hi + 1.0 - c));
        }
        const double cut = (cut_lo + cut_hi) / sgn;
        // sum_m e^{-pi sg2 (x - m fs)^2}; sg fs >= sqrt 2, so |m| <= 3 suffices
        auto pg = [&](double x) {
            double acc = 0.0;
            for (int m = -3; m <= 3; ++m) acc += exp(-M_PI * sg2 * (x - m * fs) * (x - m * fs));
            return acc;
        };
        const double base = 1.01 * nA * sqrt(2.0 * M_SQRT2 / s) * overlap;
        const double env = base * pg(0.0);
        const double fL = max(sd.JB - 1, 0) * fs / sd.Nfft;
        const double d_diff = fmin(fmax(0.0, fA - fL), fmax(0.0, fN - fL - fA));
        const double d_sum = fmax(0.0, fmin(fmin(fA, fs - fA - fL), fmin(fN - fA, fN - fL + fA)));
        // fmin drops the NaN of 0 * HUGE_VAL (an atom of zero modulus)
        const double rho_hi = fmin(env, base * (0.5 * (pg(0.0) + pg(fmin(fA, fN - fA))) + cut));
        const double rho_lo = fmin(env, base * (0.5 * (pg(d_diff) + pg(d_sum)) + cut));
        R->racc += static_cast<float>(coef * fmin(1.0, rho_hi + 1e-9));
        R->racc_lo += static_cast<float>(coef * fmin(1.0, rho_lo + 1e-9));
    }
}

constexpr int kSlices = 32;          // workgroups per unit of the region kernels

// prefix energies of the FP32 residual mirror, pre[n] = sum_{i<n} x_i^2 (FP64)
__global__ __launch_bounds__(kRegionThreads) void k_prefix(const UnitDev* __restrict__ units,
                                                           const UnitState* __restrict__ state,
                                                           const float* __restrict__ res32,
                                                           double* __restrict__ prefix,
                                                           const long long* __restrict__ prefix_off) {
    // wave w takes samples w span .. (w+1) span - 1 in coalesced rounds of 64,
    // scanned across the lanes on top of the running total
    constexpr int NW = kRegionThreads / 64;
    __shared__ double part[NW];
    const int u = blockIdx.x;
    if (!state[u].active) return;
    const UnitDev ud = units[u];
    const int N = ud.N;
    const float* x = res32 + ud.res_off;
    double* pre = prefix + prefix_off[u];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int span = ((N + NW - 1) / NW + 63) & ~63;
    const int n0 = min(N, wv * span), n1 = min(N, n0 + span);
    double sum = 0.0;
    for (int n = n0 + lane; n < n1; n += 64) sum += static_cast<double>(x[n]) * x[n];
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) part[wv] = sum;
    __syncthreads();
    double run = 0.0;
    for (int i = 0; i < wv; ++i) run += part[i];
    for (int n = n0; n < n1; n += 64) {
        const double e = n + lane < n1 ? static_cast<double>(x[n + lane]) * x[n + lane] : 0.0;
        double incl = e;
        for (int off = 1; off < 64; off <<= 1) {
            const double t = __shfl_up(incl, off);
            if (lane >= off) incl += t;
        }
        const double excl = __shfl_up(incl, 1);
        if (n + lane < n1) pre[n + lane] = run + (lane ? excl : 0.0);
        run += __shfl(incl, 63);
    }
    if (tid == 0) {
        double total = 0.0;
        for (int i = 0; i < NW; ++i) total += part[i];
        pre[N] = total;
    }
}

// log power spectra (64 pooled bins of a 256-point DFT) of the 256-sample
// blocks the last atom touched, or of every block on init; one wave per block,
// which transforms it by four radix-4 Stockham passes in its own LDS buffer
__global__ __launch_bounds__(kRegionThreads) void k_block_spectra(const UnitDev* __restrict__ units,
                                                                  const UnitState* __restrict__ state,
                                                                  const float* __restrict__ res32,
                                                                  float* __restrict__ blockfeat,
                                                                  const long long* __restrict__ feat_off,
                                                                  int init) {
    const int u = blockIdx.y;
    const UnitState st = state[u];
    if (!st.active) return;
    const UnitDev ud = units[u];
    const int N = ud.N;
    const float* x = res32 + ud.res_off;
    float* bf = blockfeat + feat_off[u];
    int b0 = 0, b1 = ud.B - 1;
    if (!init) {
        const double hw = kGaussHalfWidth * st.sA;
        b0 = max(0, static_cast<int>(lrint((st.tA - hw) * ud.fs)) / kBlock);
        b1 = min(ud.B - 1, static_cast<int>(lrint((st.tA + hw) * ud.fs)) / kBlock);
    }
    // e^{-2 pi i j / 256} from the hardware sine/cosine
    __shared__ float2 wtab[kBlock];
    constexpr int waves = kRegionThreads / 64;
    __shared__ float2 wbuf[waves][kBlock + kBlock / 16];   // one float2 of padding per 16
    for (int j = threadIdx.x; j < kBlock; j += kRegionThreads)
        wtab[j] = cis_rev(-static_cast<float>(j) * (1.0f / kBlock));
    __syncthreads();
    const int m = threadIdx.x & 63;
    float2* z = wbuf[threadIdx.x >> 6];
    auto pad = [](const int i) { return i + (i >> 4); };
    for (int b = b0 + blockIdx.x * waves + (threadIdx.x >> 6); b <= b1; b += gridDim.x * waves) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = b * kBlock + m + 64 * r;
            z[pad(m + 64 * r)] = make_float2(i < N ? x[i] : 0.0f, 0.0f);
        }
        wave_sync();
        // butterfly m of the pass at sub-length Ns: inputs m + 64 r, twiddles
        // e^{-2 pi i r k / (4 Ns)}, k = m mod Ns, outputs (m - k) 4 + k + r Ns
#pragma unroll
        for (int Ns = 1; Ns < kBlock; Ns *= 4) {
            const int k = m & (Ns - 1);
            float2 v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = z[pad(m + 64 * r)];
            wave_sync();
#pragma unroll
            for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], wtab[(r * k * (kBlock / (4 * Ns))) & (kBlock - 1)]);
            dft4(v);
#pragma unroll
            for (int r = 0; r < 4; ++r) z[pad((m - k) * 4 + k + r * Ns)] = v[r];
            wave_sync();
        }
        // bins 2m + 1 and 2m + 2
        const float2 a = z[pad(2 * m + 1)], c = z[pad(2 * m + 2)];
        bf[b * kFeatBins + m] = logf(1e-30f + (a.x * a.x + a.y * a.y + c.x * c.x + c.y * c.y));
        wave_sync();   // the next block overwrites the buffer
    }
}

// best product over the unit's clean regions
__global__ __launch_bounds__(kRegionThreads) void k_clean_best(const UnitDev* __restrict__ units,
                                                               const UnitState* __restrict__ state,
                                                               const RegionState* __restrict__ regions,
                                                               float* __restrict__ pbest) {
    __shared__ float red[kRegionThreads];
    const int u = blockIdx.x;
    if (!state[u].active) return;
    const UnitDev ud = units[u];
    const int tid = threadIdx.x;
    float best = 0.0f;
    for (int r = tid; r < ud.L * ud.B; r += kRegionThreads) {
        const RegionState* R = &regions[ud.reg_off + r];
        if (!(R->flags & 1)) best = fmaxf(best, R->rmax);
    }
    red[tid] = best;
    __syncthreads();
    for (int off = kRegionThreads / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] = fmaxf(red[tid], red[tid + off]);
        __syncthreads();
    }
    if (tid == 0) pbest[u] = red[0];
}

// Bounds of the dirty regions; those whose bound exceeds the best clean
// product / (1 - delta) are candidates (flag bit 2) and get a network row
// (every dirty region does when all_rows is set).  Row features: 64 pooled
// log powers of the region's block, log bound, log stale max, 8 scale
// features, block fill, log residual/total; every energy relative to the
// current residual energy.
__global__ __launch_bounds__(kRegionThreads) void k_rows(
    const UnitDev* __restrict__ units, const UnitState* __restrict__ state, const ScaleDev* __restrict__ scales,
    const RegionDev* __restrict__ regtab, RegionState* __restrict__ regions, const double* __restrict__ prefix,
    const long long* __restrict__ prefix_off, const float* __restrict__ blockfeat,
    const long long* __restrict__ feat_off, float* __restrict__ rows, int* __restrict__ rowmap,
    float* __restrict__ rowref, int* __restrict__ rowcount, int max_rows, const float* __restrict__ pbest,
    float delta, int all_rows) {
    const int u = blockIdx.y;
    const UnitState st = state[u];
    if (!st.active) return;
    const UnitDev ud = units[u];
    const int N = ud.N;
    const double* pre = prefix + prefix_off[u];
    const float* bf = blockfeat + feat_off[u];
    const double ref = pre[N] > 0.0 ? pre[N] : 1.0;
    const float lref = logf(static_cast<float>(ref));
    const float eref = static_cast<float>(st.totalEnergy * ud.fs) * st.xscale * st.xscale;
    const float limit = pbest[u] / (1.0f - delta);
    // Candidate rows are taken a wave at a time (a ballot and one atomic), so a
    // wave's candidates own consecutive rows in lane order; the wave then
    // stages them 16 at a time in LDS (the 64 block features by one coalesced
    // load per row, the 16 scalar features from their owners) and writes the
    // staged rows as one contiguous block, instead of each lane writing its
    // own 320-byte row across 80 scattered stores.
    constexpr int NW = kRegionThreads / 64;
    static_assert(kFeatBins == 64 && kFeatDim == kFeatBins + 16, "a row: one feature per lane, then 16 scalars");
    __shared__ float stage[NW][16][kFeatDim + 1];
    __shared__ int stage_b[NW][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nreg = ud.L * ud.B;
    for (int r0 = blockIdx.x * kRegionThreads; r0 < nreg; r0 += gridDim.x * kRegionThreads) {   // uniform rounds
        const int r = r0 + static_cast<int>(threadIdx.x);
        const int rc = min(r, nreg - 1);   // records of the round's last region for threads past it
        RegionState* R = &regions[ud.reg_off + rc];
        const int l = rc / ud.B, b = rc % ud.B;
        const ScaleDev& sd = scales[ud.scale0 + l];
        bool want = false;
        float bound = 0.0f;
        if (r < nreg && (R->flags & 1)) {
            const RegionDev rt = regtab[ud.regtab_off + r];
            // bins j >= JB and bins 0, M: (B) at E <= 1e-4 or a real value;
            // bins 1 <= j < JB: the smaller of (A) from their |u| seeds and (B)
            const float hi = sqrtf(R->rmax) + R->racc * sd.gH;
            const float loA = (sqrtf(R->rlow) + R->racc_lo) * sd.gL;
            const float loB = sqrtf(R->rmax) + R->racc_lo * sd.lL;
            const float lo = fminf(loA, loB);
            const float grow = fmaxf(hi * hi, lo * lo) * 1.0001f;
            // Cauchy-Schwarz per column: P(t, f) <= csnorm * (residual energy under
            // column t's window [iL, iR]), at the region's largest.  A region of
            // many narrow columns spans far more samples than any one window.
            // Only needed where the growth bound fails (the certification's limit
            // is never below this one).  Where the columns are many (more than 16
            // and more than the starts below), windows of Wmax + S samples at
            // starts S apart (S = Wmax / 2) cover every column's window
            // [iL, iL + Wmax) between the first and the last start.
            bound = grow;
            if (grow > limit || all_rows) {
                const double hw = kGaussHalfWidth * sd.s;
                auto win = [&](int t, int* iL, int* iR) {
                    const double tc = t * sd.tMax / (sd.T - 1);
                    const long wl = lrint((tc - hw) * sd.fs), wr = lrint((tc + hw) * sd.fs);
                    *iL = static_cast<int>(wl > 0 ? wl : 0);
                    *iR = static_cast<int>(wr < N - 1 ? wr : N - 1);
                };
                int il0, il1, ir0, ir1;
                win(rt.tlo, &il0, &ir0);
                win(rt.thi, &il1, &ir1);
                const int S = max(1, sd.Wmax / 2);
                double emax = 0.0;
                const int ncol = rt.thi - rt.tlo + 1;
                if (ncol > 16 && (il1 - il0) / S + 1 < ncol) {
#pragma unroll 4
                    for (int i = il0; i <= il1; i += S) emax = fmax(emax, pre[min(i + sd.Wmax + S, N)] - pre[i]);
                    emax = fmin(emax, pre[ir1 + 1] - pre[il0]);   // nor above the union of the windows
                } else {
#pragma unroll 4
                    for (int t = rt.tlo; t <= rt.thi; ++t) {
                        int iL, iR;
                        win(t, &iL, &iR);
                        emax = fmax(emax, pre[iR + 1] - pre[iL]);
                    }
                }
                bound = fminf(sd.csnorm * static_cast<float>(emax), grow);
            }
            R->rbound = bound;
            const bool candidate = bound > limit;
            R->flags = candidate ? (R->flags | 4) : (R->flags & ~4);
            want = candidate || all_rows;
        }
        const unsigned long long m = __ballot(want);
        if (m == 0ull) continue;   // uniform over the wave
        const int n = __popcll(m);
        const int lead = __ffsll(static_cast<long long>(m)) - 1;
        int base = 0;
        if (lane == lead) base = atomicAdd(rowcount, n);
        base = __shfl(base, lead);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        const int row = base + rank;
        if (want && row < max_rows) {
            rowmap[row] = static_cast<int>(ud.reg_off + r);
            rowref[row] = static_cast<float>(ref);
        }
        for (int g = 0; g < n; g += 16) {
            const int cnt = min(16, n - g);
            if (want && rank >= g && rank < g + cnt) {
                float* sc = stage[wv][rank - g] + kFeatBins;
                stage_b[wv][rank - g] = b;
                sc[0] = logf(1e-30f + bound / static_cast<float>(ref));
                sc[1] = logf(1e-30f + R->rmax / static_cast<float>(ref));
                const float ls = log2f(sd.s * sd.fs) * (1.0f / 16.0f);
                sc[2] = ls;
                sc[3] = ls * ls;
                sc[4] = __builtin_amdgcn_sinf(0.5f * ls);
                sc[5] = __builtin_amdgcn_cosf(0.5f * ls);
                sc[6] = __builtin_amdgcn_sinf(ls);
                sc[7] = __builtin_amdgcn_cosf(ls);
                sc[8] = __builtin_amdgcn_sinf(2.0f * ls);
                sc[9] = __builtin_amdgcn_cosf(2.0f * ls);
                sc[10] = static_cast<float>(min(kBlock, N - b * kBlock)) / kBlock;
                sc[11] = logf(1e-30f + static_cast<float>(ref) / (eref > 0.0f ? eref : 1.0f));
                sc[12] = sc[13] = sc[14] = sc[15] = 0.0f;
            }
            wave_sync();
            for (int i = 0; i < cnt; ++i) stage[wv][i][lane] = bf[stage_b[wv][i] * kFeatBins + lane] - lref;
            wave_sync();
            // rows base + g .. base + g + cnt - 1 (those below max_rows) are one block
            const long long first = static_cast<long long>(base) + g;
            const int keep = static_cast<int>(min(static_cast<long long>(cnt), max(0ll, max_rows - first)));
            for (int j = lane; j < keep * kFeatDim; j += 64) rows[first * kFeatDim + j] = stage[wv][j / kFeatDim][j % kFeatDim];
            wave_sync();
        }
    }
}

__global__ void k_scores_scatter(const float* __restrict__ y, const int* __restrict__ rowmap,
                                 const int* __restrict__ rowcount, int max_rows, RegionState* __restrict__ regions) {
    const int n = min(*rowcount, max_rows);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        regions[rowmap[i]].rscore = y[i];
}

// evaluate: the top-K dirty candidate regions by network score (by stale
// maximum without a network); every dirty region in COLLECT mode (all)
__global__ __launch_bounds__(kRegionThreads) void k_choose(const UnitDev* __restrict__ units,
                                                           const UnitState* __restrict__ state,
                                                           RegionState* __restrict__ regions, int K, int by_score,
                                                           int all) {
    __shared__ float bv[kRegionThreads];
    __shared__ int bi[kRegionThreads];
    const int u = blockIdx.x;
    if (!state[u].active) return;
    const UnitDev ud = units[u];
    const int tid = threadIdx.x;
    const int nreg = ud.L * ud.B;
    for (int r = tid; r < nreg; r += kRegionThreads) {
        RegionState* R = &regions[ud.reg_off + r];
        if ((R->flags & 1) && all) R->flags |= 2;
    }
    if (all) return;
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        float best = -INFINITY;
        int bidx = -1;
        for (int r = tid; r < nreg; r += kRegionThreads) {
            const RegionState* R = &regions[ud.reg_off + r];
            if ((R->flags & 7) != 5) continue;   // dirty candidate, not chosen yet
            const float v = by_score ? R->rscore : R->rmax;
            if (v > best || (v == best && r < bidx)) {
                best = v;
                bidx = r;
            }
        }
        bv[tid] = best;
        bi[tid] = bidx;
        __syncthreads();
        for (int off = kRegionThreads / 2; off > 0; off >>= 1) {
            if (tid < off) {
                const float ov = bv[tid + off];
                const int oi = bi[tid + off];
                if (oi >= 0 && (ov > bv[tid] || bi[tid] < 0 || (ov == bv[tid] && oi < bi[tid]))) {
                    bv[tid] = ov;
                    bi[tid] = oi;
                }
            }
            __syncthreads();
        }
        // every thread reads the winner before the next round rewrites bi
        const int chosen = bi[0];
        if (tid == 0 && chosen >= 0) regions[ud.reg_off + chosen].flags |= 2;
        __syncthreads();
        if (chosen < 0) break;
    }
}

__global__ void k_reset_counts(const ClassDev* __restrict__ classes, int n, int* __restrict__ extra) {
    const int i = threadIdx.x;
    if (i < n) *classes[i].count = 0;
    if (i == 0 && extra) *extra = 0;
}

// One work item per flagged region of unit blockIdx.x, appended to its
// class's list.  The items of each class are counted in LDS first and their
// range reserved with one global atomic per class and workgroup: per-region
// atomics on a dozen class counters serialise in L2.
__global__ __launch_bounds__(kRegionThreads) void k_region_items(const ClassDev* __restrict__ classes,
                                                                 const UnitDev* __restrict__ units,
                                                                 const UnitState* __restrict__ state,
                                                                 const ScaleDev* __restrict__ scales,
                                                                 const UnitScale* __restrict__ uscales,
                                                                 const RegionDev* __restrict__ regtab,
                                                                 const RegionState* __restrict__ regions,
                                                                 unsigned long long* __restrict__ stat) {
    constexpr int kSlots = 3 * kNumClasses;
    __shared__ int lcount[kSlots];
    __shared__ int lbase[kSlots];
    __shared__ unsigned nflag;
    const int u = blockIdx.x;
    if (!state[u].active) return;   // uniform over the workgroup
    const UnitDev ud = units[u];
    const int tid = threadIdx.x;
    const int nreg = ud.L * ud.B;
    for (int i = tid; i < kSlots; i += kRegionThreads) lcount[i] = 0;
    if (tid == 0) nflag = 0;
    __syncthreads();
    unsigned flagged = 0;
    for (int r = tid; r < nreg; r += kRegionThreads) {
        if (!(regions[ud.reg_off + r].flags & 2)) continue;
        ++flagged;
        const RegionDev rt = regtab[ud.regtab_off + r];
        if (rt.tlo <= rt.thi) atomicAdd(&lcount[uscales[ud.us_off + r / ud.B].cls_slot], 1);
    }
    if (flagged) atomicAdd(&nflag, flagged);
    __syncthreads();
    // slots with items are ClassDev entries of this plan
    for (int i = tid; i < kSlots; i += kRegionThreads) {
        lbase[i] = lcount[i] ? atomicAdd(classes[i].count, lcount[i]) : 0;
        lcount[i] = 0;
    }
    if (tid == 0 && stat && nflag) atomicAdd(stat, static_cast<unsigned long long>(nflag));
    __syncthreads();
    for (int r = tid; r < nreg; r += kRegionThreads) {
        if (!(regions[ud.reg_off + r].flags & 2)) continue;
        const RegionDev rt = regtab[ud.regtab_off + r];
        if (rt.tlo > rt.thi) continue;
        const int l = r / ud.B;
        const UnitScale us = uscales[ud.us_off + l];
        const ClassDev& cd = classes[us.cls_slot];
        const ScaleDev& sd = scales[ud.scale0 + l];
        WorkItem it{};
        it.job = us.job;
        it.tile = rt.tlo * sd.G / cd.ncell;
        it.tiles = (rt.thi * sd.G + sd.G - 1) / cd.ncell - it.tile + 1;
        it.tlo = rt.tlo;
        it.thi = rt.thi;
        it.atom = 0;
        it.region = static_cast<int>(ud.reg_off + r);
        const int pos = lbase[us.cls_slot] + atomicAdd(&lcount[us.cls_slot], 1);
        if (pos < cd.max_items) cd.items[pos] = it;
    }
}

// P* = best clean region; dirty regions whose bound exceeds P*/(1-delta) are evaluated
__global__ __launch_bounds__(kRegionThreads) void k_certify(const UnitDev* __restrict__ units,
                                                            const UnitState* __restrict__ state,
                                                            RegionState* __restrict__ regions, float delta,
                                                            unsigned long long* __restrict__ stats) {
    __shared__ float red[kRegionThreads];
    __shared__ int nviol, ndirty;   // this unit's flagged and dirty regions
    const int u = blockIdx.x;
    if (!state[u].active) return;
    const UnitDev ud = units[u];
    const int tid = threadIdx.x;
    const int nreg = ud.L * ud.B;
    float best = 0.0f;
    for (int r = tid; r < nreg; r += kRegionThreads) {
        const RegionState* R = &regions[ud.reg_off + r];
        if (!(R->flags & 1)) best = fmaxf(best, R->rmax);
    }
    red[tid] = best;
    if (tid == 0) nviol = ndirty = 0;
    __syncthreads();
    for (int off = kRegionThreads / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] = fmaxf(red[tid], red[tid + off]);
        __syncthreads();
    }
    const float limit = red[0] / (1.0f - delta);
    int mine = 0, dirty = 0;
    for (int r = tid; r < nreg; r += kRegionThreads) {
        RegionState* R = &regions[ud.reg_off + r];
        if (!(R->flags & 1)) continue;
        ++dirty;
        if (R->rbound > limit) {
            R->flags |= 2;
            ++mine;
        }
    }
    if (mine) atomicAdd(&nviol, mine);
    if (dirty) atomicAdd(&ndirty, dirty);
    __syncthreads();
    // one update of each statistic per unit (same-address atomics from every
    // wave of every unit queue behind each other)
    if (tid == 0) {
        atomicAdd(&stats[2], 1ull);
        if (nviol) {
            atomicAdd(&stats[0], 1ull);
            atomicAdd(&stats[1], static_cast<unsigned long long>(nviol));
        }
        if (ndirty) atomicAdd(&stats[3], static_cast<unsigned long long>(ndirty));
    }
}

// training rows: targets log(new exact max / residual) of the rows k_rows
// wrote, one row in `stride` kept (by a hash of region and residual energy)
__global__ void k_collect_rows(const float* __restrict__ rows, const int* __restrict__ rowmap,
                               const float* __restrict__ rowref, const int* __restrict__ rowcount, int max_rows,
                               const RegionState* __restrict__ regions, float* __restrict__ outX,
                               float* __restrict__ outT, unsigned long long* __restrict__ outCount,
                               unsigned long long capacity, unsigned stride) {
    const int n = min(*rowcount, max_rows);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (stride > 1) {
            unsigned h = static_cast<unsigned>(rowmap[i]) * 2654435761u ^ __float_as_uint(rowref[i]) * 2246822519u;
            h ^= h >> 15;
            h *= 2654435761u;
            if ((h >> 8) % stride != 0) continue;
        }
        const unsigned long long pos = atomicAdd(outCount, 1ull);
        if (pos >= capacity) continue;
        for (int q = 0; q < kFeatDim; ++q) outX[pos * kFeatDim + q] = rows[static_cast<long long>(i) * kFeatDim + q];
        outT[pos] = logf(1e-30f + regions[rowmap[i]].rmax / rowref[i]);
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_refit(const empi_atom* __restrict__ sum_atoms,
                                              const UnitState* __restrict__ sum_state,
                                              const float* __restrict__ input, long long in_off, int C,
                                              int N, double fs, double* __restrict__ scratch,
                                              empi_atom* __restrict__ out, int cstride) {
    __shared__ double red[NT];
    __shared__ empi_atom a;
    const int c = blockIdx.x;
    double* x = scratch + static_cast<long long>(c) * N;
    for (int n = threadIdx.x; n < N; n += NT) x[n] = input[in_off + static_cast<long long>(c) * N + n];
    __syncthreads();
    const int count = sum_state->iteration;
    for (int i = 0; i < count; ++i) {
        if (threadIdx.x == 0) a = sum_atoms[i];
        __syncthreads();
        const double sA = a.params[3] / fs, fA = a.params[4] * fs / 2.0, tA = a.params[2] / fs;
        const double omega = 2.0 * M_PI * fA, phase = a.params[5];
        const double hw = kGaussHalfWidth * sA;
        long l = lrint((tA - hw) * fs), r = lrint((tA + hw) * fs);
        const int iL = static_cast<int>(l > 0 ? l : 0);
        const int iR = static_cast<int>(r < N - 1 ? r : N - 1);
        const int Nw = 1 + iR - iL;
        const double gw = sqrt(M_PI) / sA;
        const double ga = gw * (1.0 / fs), gb = gw * (tA - iL / fs);
        double n2 = 0.0;
        for (int j = threadIdx.x; j < Nw; j += NT) {
            const double t = (j + iL) / fs - tA;
            const double xw = ga * j - gb;
            const double w = cos(omega * t + phase) * exp(-xw * xw);
            n2 += w * w;
        }
        const double norm = sqrt(block_sum<NT>(n2, red));
        double dot = 0.0;
        for (int j = threadIdx.x; j < Nw; j += NT) {
            const double t = (j + iL) / fs - tA;
            const double xw = ga * j - gb;
            const double w = cos(omega * t + phase) * exp(-xw * xw) / norm;
            dot += w * x[iL + j];
        }
        const double modulus = block_sum<NT>(dot, red);
        for (int j = threadIdx.x; j < Nw; j += NT) {
            const double t = (j + iL) / fs - tA;
            const double xw = ga * j - gb;
            const double w = cos(omega * t + phase) * exp(-xw * xw) / norm;
            x[iL + j] -= w * modulus;
        }
        if (threadIdx.x == 0) {
            empi_atom o = a;
            o.params[0] = modulus;
            o.params[1] = modulus / norm;
            out[static_cast<long long>(c) * cstride + i] = o;
        }
        __syncthreads();
    }
}

// subtract one given atom from channel c of a unit (per-call workspace API)
template <int NT>
__global__ __launch_bounds__(NT) void k_subtract_one(const UnitDev* __restrict__ units, UnitState* __restrict__ state,
                                                     const empi_atom* __restrict__ atom, int c,
                                                     double* __restrict__ residual, float* __restrict__ res32) {
    const UnitDev ud = units[0];
    __shared__ empi_atom a;
    if (threadIdx.x == 0) a = *atom;
    __syncthreads();
    const long long off = ud.res_off + static_cast<long long>(c) * ud.N;
    const float xs = state[0].xscale;
    subtract_waveform<NT>(a, ud.fs, residual + off, ud.N, res32 + off, xs);
    __syncthreads();
    // the host handed in the whole channel: refresh its FP32 mirror
    for (int n = threadIdx.x; n < ud.N; n += NT) res32[off + n] = static_cast<float>(residual[off + n] * xs);
    if (threadIdx.x == 0) {
        state[0].tA = a.params[2] / ud.fs;
        state[0].sA = a.params[3] / ud.fs;
    }
}

// ------------------------------------------------------------ host plan

// Device memory carved out of one fixed region, 256-byte aligned: the
// allocation-free entry point sets t_arena for the duration of its call, and
// every DevBuf::alloc on that thread then takes its memory from the region
// instead of hipMalloc.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    size_t wanted = 0;         // bytes the plan asked for, carved or not
    void* carve(size_t bytes) {
        const size_t off = (used + 255) & ~static_cast<size_t>(255);
        wanted = off + bytes > wanted ? off + bytes : wanted;
        if (off + bytes > size) {
            used = off + bytes;
            return nullptr;
        }
        used = off + bytes;
        return base + off;
    }
};
thread_local Arena* t_arena = nullptr;

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = false;        // from hipMalloc (else carved from an arena)
    ~DevBuf() {
        if (p && owned) (void)hipFree(p);
    }
    empi_status alloc(size_t count) {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        owned = false;
        n = count;
        if (!count) return EMPI_OK;
        if (t_arena) {
            void* q = t_arena->carve(sizeof(T) * count);
            if (!q) {
                n = 0;
                return EMPI_UNSUPPORTED;
            }
            p = static_cast<T*>(q);
            return EMPI_OK;
        }
        owned = true;
        hipError_t e = hipMalloc(&p, sizeof(T) * count);
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            owned = false;
            return e == hipErrorOutOfMemory ? EMPI_NO_MEMORY : EMPI_DEVICE;
        }
        return EMPI_OK;
    }
    empi_status upload(const T* h, size_t count, hipStream_t s) {
        if (!count) return EMPI_OK;
        EMPI_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
        return EMPI_OK;
    }
};

struct ClassJobs {
    std::vector<Job> jobs;
    long long tiles = 0;       // all tiles of the class (full build)
    int slot = -1;             // index in the device ClassDev array
    DevBuf<Job> d;
    DevBuf<WorkItem> items;    // [max_items]
    DevBuf<int> work;          // [max_items + 1]
    DevBuf<int> count;         // [1]
    int max_items = 0;         // jobs, or regions of the class in proposal mode
    long long region_items = 0;
};

struct ItemInfo {
    empi_item item;
    int dict;
    long long in_off;
    std::vector<int> units;   // unit indices
    int rows;                 // energy rows
    long long atom_off;       // atoms offset: [C][I]
    long long energy_off;     // energies offset: [rows][I+1]
    long long sum_atom_off;   // MMP2: atoms of the channel sum [I]
    long long scratch_off;    // MMP2: refit scratch [C][N]
};

// persistent grid: enough workgroups to fill 256 CUs a few times over (twice
// that with grid variant 1), a multiple of the 8 XCDs, which the kernel's
// work walk relies on
int column_grid(int Q, long long tiles, int gv) {
    const int nt = class_threads(Q);
    const long long cap = (nt >= 1024 ? 512 : nt >= 512 ? 1024 : 2048) << gv;
    const long long g = std::max(1ll, std::min(tiles, cap));
    return static_cast<int>((g + 7) / 8 * 8);
}

ColumnLauncher column_launcher(int mode, int cls, int choice, bool low = false) {
    const int lg = cls + kMinLgQ, v = choice % kWaveVariants;
    return lg <= 6 ? column_launcher_small(lg, mode, v, low)
         : lg <= 10 ? column_launcher_mid(lg, mode, v, low) : column_launcher_large(lg, mode, v, low);
}

// default tuning: the choice (wave variant + 3 * grid variant) per class that
// the autotuner picks on the 64 x 20 x 5120 SMP batch
const int kDefaultTune[kNumClasses] = {0, 0, 3, 5, 3, 3, 3, 3, 3, 3, 3, 3, 0};

int cells_per_wg(int Q) { return class_slots(Q) / 2; }

}  // namespace

// iterations the host may queue past the last one it has seen finish, and
// the ring of active-unit counts it reads them through
constexpr int kPollLag = 2;
constexpr int kPollRing = 4;

// number of active units, for the host's end-of-loop test
__global__ __launch_bounds__(256) void k_count_active(const UnitState* __restrict__ state, int U,
                                                      int* __restrict__ out) {
    __shared__ int n;
    if (threadIdx.x == 0) n = 0;
    __syncthreads();
    int c = 0;
    for (int u = threadIdx.x; u < U; u += 256) c += state[u].active != 0;
    if (c) atomicAdd(&n, c);
    __syncthreads();
    if (threadIdx.x == 0) *out = n;
}

struct empi_plan {
    std::vector<ItemInfo> items;
    std::vector<Dictionary> dicts;
    std::vector<int> dict_scale0;
    std::vector<UnitDev> units;
    std::vector<ScaleDev> scales_h;
    ClassJobs classes[3][kNumClasses];   // [mode][class]
    int tune[3][kNumClasses];            // column-kernel variant per (mode, class): wave + 3 * grid variant
    int maxIter = 0;
    long long total_in = 0;

    DevBuf<ScaleDev> d_scales;
    DevBuf<UnitDev> d_units;
    DevBuf<UnitState> d_state;
    DevBuf<double> d_res;
    DevBuf<float> d_res32;
    DevBuf<unsigned long long> d_cells;
    DevBuf<unsigned long long> d_tiles;
    DevBuf<UnitScale> d_uscales;
    DevBuf<ClassDev> d_classes;
    std::vector<UnitScale> uscales_h;
    int nclass_dev = 0;
    DevBuf<float2> d_etab;
    DevBuf<empi_atom> d_atoms;
    DevBuf<double> d_energies;
    DevBuf<double2> d_fitv;
    DevBuf<double> d_scratch;
    DevBuf<float> d_input;
    DevBuf<unsigned long long> d_counter;
    // multichannel units, whose selection step runs one workgroup per channel
    std::vector<int> mc_units;
    int mc_cmax = 0;
    DevBuf<int> d_mc;
    DevBuf<double> d_epart;               // [multichannel unit][channel] energy sums

    // proposal mode (single-channel units)
    std::vector<RegionDev> regtab_h;
    std::vector<int> dict_regtab_off, dict_B;
    long long nregions = 0;
    int prop_mode = EMPI_PROPOSAL_OFF;
    empi_mlp* net = nullptr;
    int top_k = 8;
    float delta = 1e-3f;
    bool prop_ready = false;
    int max_rows = 0;
    DevBuf<RegionDev> d_regtab;
    DevBuf<RegionState> d_regions;
    DevBuf<float> d_celllow;              // per cell: max over bins 1 <= j < JB of P (1 + E)/(1 - E)
    DevBuf<double> d_prefix;
    DevBuf<long long> d_prefix_off;
    DevBuf<float> d_blockfeat;
    DevBuf<long long> d_feat_off;
    DevBuf<float> d_rows;
    DevBuf<int> d_rowmap;
    DevBuf<float> d_rowref;
    DevBuf<int> d_rowcount;
    DevBuf<float> d_scores;
    DevBuf<unsigned long long> d_pstats;
    DevBuf<float> d_pbest;
    unsigned long long h_pstats[5] = {};
    DevBuf<float> d_trainX, d_trainT;
    DevBuf<unsigned long long> d_traincount;
    unsigned train_stride = 1;
    hipGraphExec_t prop_exec = nullptr;
    hipGraph_t prop_graph = nullptr;

    // host copies of the last run's results
    std::vector<empi_atom> h_atoms;
    std::vector<double> h_energies;
    std::vector<UnitState> h_state;
    unsigned long long h_counter[2 * kNumClasses] = {};          // summed over the XCD copies
    unsigned long long h_counter_copies[kCounterWords] = {};
    bool has_results = false;

    // one MP iteration (select, plan, every column class) captured once
    hipStream_t stream = nullptr;
    // allocation-free run: the caller's stream, no graphs, no stream of its own
    bool noalloc = false;
    // staged inputs (empi_plan_stage): two device buffers, one filled on the
    // upload stream while a run reads the other
    DevBuf<float> d_stage_buf[2];
    int stage_ready = -1;     // slot holding a staged input not yet run
    int stage_active = -1;    // slot the last run read
    std::mutex stage_mu;
    hipStream_t upload = nullptr;
    hipGraphExec_t iter_exec = nullptr;
    hipGraph_t iter_graph = nullptr;
    // end-of-loop detection: after each iteration the number of active units
    // goes to a pinned ring slot with an event (allocating runs only)
    DevBuf<int> d_nactive;
    int* h_nactive = nullptr;
    hipEvent_t poll_ev[kPollRing] = {};

    // optional per-class kernel timing (HIP events on the launch stream)
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[kNumClasses];
    size_t ev_used[kNumClasses] = {};
    double class_ms[kNumClasses] = {};
    long long class_launches[kNumClasses] = {};
    empi_plan() {
        for (int m = 0; m < 3; ++m)
            for (int c = 0; c < kNumClasses; ++c) tune[m][c] = m == kModeSingle ? kDefaultTune[c] : 0;
    }
    void drop_graphs() {
        if (iter_exec) (void)hipGraphExecDestroy(iter_exec);
        if (iter_graph) (void)hipGraphDestroy(iter_graph);
        if (prop_exec) (void)hipGraphExecDestroy(prop_exec);
        if (prop_graph) (void)hipGraphDestroy(prop_graph);
        iter_exec = prop_exec = nullptr;
        iter_graph = prop_graph = nullptr;
    }
    ~empi_plan() {
        if (iter_exec) (void)hipGraphExecDestroy(iter_exec);
        if (iter_graph) (void)hipGraphDestroy(iter_graph);
        if (prop_exec) (void)hipGraphExecDestroy(prop_exec);
        if (prop_graph) (void)hipGraphDestroy(prop_graph);
        if (stream && !noalloc) (void)hipStreamDestroy(stream);
        if (upload) (void)hipStreamDestroy(upload);
        if (h_nactive) (void)hipHostFree(h_nactive);
        for (hipEvent_t e : poll_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& v : ev)
            for (auto& e : v) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
    }
};

struct empi_ws {
    empi_plan plan;
    int C = 0, N = 0;
    double fs = 0;
    bool computed = false;
    bool pending = false;    // columns to re-evaluate before the next search
    DevBuf<empi_atom> d_one;
};

namespace {

int lg2(long long x) {
    int l = 0;
    while ((1ll << l) < x) ++l;
    return l;
}

empi_status dict_for(empi_plan* p, const empi_dict_params& dp, double fs, int N, int max_q, int* id) {
    for (size_t i = 0; i < p->dicts.size(); ++i) {
        const Dictionary& d = p->dicts[i];
        if (d.fs == fs && d.N == N && d.max_q == max_q && d.params.energyError == dp.energyError &&
            d.params.scaleMin == dp.scaleMin &&
            d.params.scaleMax == dp.scaleMax &&
            (d.params.freqMax == dp.freqMax || (std::isnan(d.params.freqMax) && std::isnan(dp.freqMax)))) {
            *id = static_cast<int>(i);
            return EMPI_OK;
        }
    }
    if (N > kMaxSamples) return EMPI_UNSUPPORTED;
    Dictionary d;
    empi_status st = build_dictionary(dp, fs, N, &d, max_q);
    if (st != EMPI_OK) return st;
    for (const Scale& sc : d.scales) {
        if (sc.Q > kMaxQ || sc.Nfft > kMaxNfft) return EMPI_UNSUPPORTED;
    }
    *id = static_cast<int>(p->dicts.size());
    p->dicts.push_back(std::move(d));
    return EMPI_OK;
}

// Lays out units, scales, jobs and buffers; items may already carry per-unit
// FP64 input (workspace mode) in which case in_sum/in_channel are unused.
empi_status plan_build(empi_plan* p, const empi_item* items, int count, int ws_constraint) {
    if (count <= 0 || !items) return EMPI_INVALID_ARGUMENT;
    long long in_off = 0, res_off = 0, cell_off = 0, atom_off = 0, energy_off = 0, scratch_off = 0;
    int max_c = 1;
    for (int i = 0; i < count; ++i) {
        const empi_item& it = items[i];
        if (it.channels <= 0 || it.samples < 2 || !(it.fs > 0) || it.iterationMax <= 0 || it.variant < EMPI_SMP ||
            it.variant > EMPI_MMP3)
            return EMPI_INVALID_ARGUMENT;
        const double residualEnergy = 1 - 0.01 * it.energyPercent;
        if (residualEnergy < 0 || residualEnergy >= 1.0) return EMPI_INVALID_ARGUMENT;
        // multichannel units (MMP1, MMP3 over several channels, constrained
        // workspaces) keep sub-transforms up to 8192 unfolded
        const bool multi = (it.variant == EMPI_MMP1 || it.variant == EMPI_MMP3) &&
                           (it.channels > 1 || it.variant == EMPI_MMP1 || ws_constraint);
        int did = 0;
        empi_status st =
            dict_for(p, it.dict, it.fs, it.samples, multi ? kMaxSubTransformMulti : kMaxSubTransform, &did);
        if (st != EMPI_OK) return st;
        ItemInfo info{};
        info.item = it;
        info.dict = did;
        info.in_off = in_off;
        const int C = it.channels, N = it.samples, I = it.iterationMax;
        info.rows = it.variant == EMPI_SMP ? C : 1;
        info.atom_off = atom_off;
        atom_off += static_cast<long long>(C) * I;
        info.energy_off = energy_off;
        energy_off += static_cast<long long>(info.rows) * (I + 1);
        info.sum_atom_off = -1;
        info.scratch_off = -1;
        if (it.variant == EMPI_MMP2) {
            info.sum_atom_off = atom_off;
            atom_off += I;
            info.scratch_off = scratch_off;
            scratch_off += static_cast<long long>(C) * N;
        }
        const Dictionary& d = p->dicts[did];
        auto add_unit = [&](int uc, int in_channel, int in_sum, long long aoff, int cstride, long long eoff,
                            int constraint) {
            UnitDev u{};
            u.res_off = res_off;
            res_off += static_cast<long long>(uc) * N;
            u.cell_off = cell_off;
            cell_off += d.cells;
            u.in_off = in_off;
            u.C = uc;
            u.N = N;
            u.in_channel = in_channel;
            u.in_sum = in_sum;
            u.scale0 = did;   // patched to a scale index below
            u.L = static_cast<int>(d.scales.size());
            u.constraint = constraint;
            u.iterMax = I;
            u.fs = it.fs;
            u.residualEnergy = residualEnergy;
            u.atoms = reinterpret_cast<empi_atom*>(aoff);      // patched to a device pointer below
            u.atom_cstride = cstride;
            u.energies = reinterpret_cast<double*>(eoff);      // patched below
            info.units.push_back(static_cast<int>(p->units.size()));
            p->units.push_back(u);
            max_c = std::max(max_c, uc);
        };
        if (it.variant == EMPI_SMP || (C == 1 && it.variant != EMPI_MMP2)) {
            if (it.variant == EMPI_SMP) {
                for (int c = 0; c < C; ++c)
                    add_unit(1, c, 0, info.atom_off + static_cast<long long>(c) * I, I,
                             info.energy_off + static_cast<long long>(c) * (I + 1), 0);
            } else {
                add_unit(1, 0, 0, info.atom_off, I, info.energy_off,
                         it.variant == EMPI_MMP1 ? 1 : ws_constraint);
            }
        } else if (it.variant == EMPI_MMP2) {
            add_unit(1, 0, C, info.sum_atom_off, I, info.energy_off, 0);
        } else {
            add_unit(C, 0, 0, info.atom_off, I, info.energy_off, it.variant == EMPI_MMP1 ? 1 : 0);
        }
        in_off += static_cast<long long>(C) * N;
        p->maxIter = std::max(p->maxIter, I);
        p->items.push_back(std::move(info));
    }
    p->total_in = in_off;

    // scales of every dictionary
    std::vector<float2> etab_h;
    for (size_t di = 0; di < p->dicts.size(); ++di) {
        const Dictionary& d = p->dicts[di];
        p->dict_scale0.push_back(static_cast<int>(p->scales_h.size()));
        long long before = 0;
        for (const Scale& sc : d.scales) {
            ScaleDev s{};
            s.s = sc.s;
            s.fs = d.fs;
            s.tMax = d.tMax;
            s.cells_before = before;
            before += static_cast<long long>(sc.T) * sc.G;
            s.N = d.N;
            s.Nfft = sc.Nfft;
            s.M = sc.Nfft / 2;
            s.T = sc.T;
            s.F = sc.F;
            s.Q = sc.Q;
            s.P = sc.P;
            s.G = sc.G;
            s.lgQ = lg2(sc.Q);
            s.lgNfft = lg2(sc.Nfft);
            s.alpha = static_cast<float>(sc.s * d.fs / sc.Nfft);
            s.Kp = static_cast<float>(2.0 * M_SQRT2 / (sc.s * d.fs));
            {
                // E = exp(-2 pi (j alpha)^2) is below 3e-25 from j alpha = 3 on: only
                // the first bins on either side of the spectrum need it
                const double alpha = sc.s * d.fs / sc.Nfft;
                const double Kp = 2.0 * M_SQRT2 / (sc.s * d.fs);
                const int JT = static_cast<int>(std::min<double>(sc.Nfft / 2 + 1, std::ceil(3.0 / alpha) + 1));
                s.JT = JT;
                s.invKp = static_cast<float>(1.0 / Kp);
                s.etab = reinterpret_cast<const float2*>(static_cast<uintptr_t>(etab_h.size()));   // patched below
                for (int j = 0; j < JT; ++j) {
                    const double x = 2.0 * M_PI * (j * alpha) * (j * alpha);
                    etab_h.push_back(make_float2(static_cast<float>(-std::expm1(-x) / Kp),
                                                 static_cast<float>(2.0 * std::exp(-x) / Kp)));
                }
            }
            s.gA = static_cast<float>(M_PI / ((sc.s * d.fs) * (sc.s * d.fs)));
            s.sgA = static_cast<float>(std::sqrt(M_PI * M_LOG2E) / (sc.s * d.fs));
            {
                // the proposal's region bound: bins with E_j > 1e-4 take the low-bin seeds
                const double alpha = sc.s * d.fs / sc.Nfft;
                const int M = sc.Nfft / 2;
                const double jb = std::ceil(std::sqrt(std::log(1e4) / (2.0 * M_PI)) / alpha);
                s.JB = static_cast<int>(std::min<double>(M + 1, std::min<double>(jb, s.JT)));
                const double eb = std::exp(-2.0 * M_PI * (s.JB * alpha) * (s.JB * alpha));
                const double e1 = M >= 2 ? std::exp(-2.0 * M_PI * alpha * alpha) : 0.0;
                // max over phi of |grad(|u| h)| = sqrt((1 + E^2 + 2 E c) / (1 + E c)^3), c = cos 2 phi
                auto lip = [](double E) {
                    return E >= 1.0 / 3.0 ? std::sqrt(32.0 / 27.0) / (1.0 - E * E) : 1.0 / std::sqrt(1.0 - E);
                };
                s.gH = static_cast<float>(lip(eb) * (1.0 + 1e-6));
                s.lL = static_cast<float>(lip(e1) * (1.0 + 1e-6));
                s.gL = static_cast<float>(1.0 / std::sqrt(1.0 - e1) * (1.0 + 1e-6));
            }
            s.Wmax = sc.Wmax;
            s.fold = sc.fold;
            s.csnorm = static_cast<float>(proposal_bound_factor(sc.s, d.fs, sc.Nfft));
            p->scales_h.push_back(s);
        }
    }

    // region tables: time indices of each (scale, 256-sample block); column t
    // sits at sample t (N-1)/(T-1) and belongs to block floor(of that / 256),
    // with each block's first column rounded to the nearest boundary of the
    // scale's column tiles (NCELL / G columns), so that evaluating a region
    // never transforms a tile of which only part is evaluated
    for (size_t di = 0; di < p->dicts.size(); ++di) {
        const Dictionary& d = p->dicts[di];
        const long long N = d.N;
        const int B = static_cast<int>((N + kBlock - 1) / kBlock);
        p->dict_regtab_off.push_back(static_cast<int>(p->regtab_h.size()));
        p->dict_B.push_back(B);
        for (const Scale& sc : d.scales) {
            const long long T = sc.T;
            const int cpw = cells_per_wg(sc.Q);
            const long long cols = cpw > sc.G ? cpw / sc.G : 1;   // powers of two
            auto first_t = [&](long long b) {
                const long long t = (b * kBlock * (T - 1) + (N - 2)) / (N - 1);
                return (t + cols / 2) / cols * cols;
            };
            for (int b = 0; b < B; ++b) {
                RegionDev r;
                r.tlo = static_cast<int>(std::min<long long>(first_t(b), T));
                r.thi = static_cast<int>(std::min<long long>(first_t(b + 1) - 1, T - 1));
                if (b == B - 1) r.thi = static_cast<int>(T - 1);
                // the region reduction takes a region's best from its tiles'
                // bests, so a region with columns must be whole tiles (up to
                // the scale's end)
                if (r.tlo <= r.thi) {
                    const long long c0 = static_cast<long long>(r.tlo) * sc.G;
                    const long long c1 = (static_cast<long long>(r.thi) + 1) * sc.G;
                    if (c0 % cpw != 0 || (c1 % cpw != 0 && r.thi != T - 1)) return EMPI_INTERNAL;
                }
                p->regtab_h.push_back(r);
            }
        }
    }

    // jobs per class, tile bests per (unit, scale)
    long long tile_off = 0;
    for (size_t ui = 0; ui < p->units.size(); ++ui) {
        UnitDev& u = p->units[ui];
        const int did = u.scale0;
        u.scale0 = p->dict_scale0[did];
        u.us_off = static_cast<long long>(p->uscales_h.size());
        u.B = p->dict_B[did];
        u.regtab_off = p->dict_regtab_off[did];
        u.reg_off = p->nregions;
        p->nregions += static_cast<long long>(u.L) * u.B;
        for (int l = 0; l < u.L; ++l) {
            const ScaleDev& s = p->scales_h[u.scale0 + l];
            const int cls = s.lgQ - kMinLgQ;
            const int mode = u.constraint ? kModeMmp1 : (u.C > 1 ? kModeSum : kModeSingle);
            if (!column_launcher(mode, cls, 0)) return EMPI_INTERNAL;   // no kernel for this class and mode
            ClassJobs& cj = p->classes[mode][cls];
            Job j{};
            j.unit = static_cast<int>(ui);
            j.scale = u.scale0 + l;
            j.cell_off = u.cell_off + s.cells_before;
            const int cpw = cells_per_wg(s.Q);
            j.tiles = static_cast<int>((static_cast<long long>(s.T) * s.G + cpw - 1) / cpw);
            j.tile_off = tile_off;
            tile_off += j.tiles;
            if (cj.tiles + j.tiles > INT_MAX) return EMPI_UNSUPPORTED;
            cj.tiles += j.tiles;
            UnitScale us{};
            us.tile_off = j.tile_off;
            us.tiles = j.tiles;
            us.ncell = cpw;
            us.cls_slot = mode * kNumClasses + cls;   // patched to the ClassDev slot below
            us.job = static_cast<int>(cj.jobs.size());
            cj.jobs.push_back(j);
            cj.region_items += u.B;
            p->uscales_h.push_back(us);
        }
    }

    // device buffers
    empi_status st;
    if ((st = p->d_scales.alloc(p->scales_h.size())) != EMPI_OK) return st;
    if ((st = p->d_units.alloc(p->units.size())) != EMPI_OK) return st;
    if ((st = p->d_state.alloc(p->units.size())) != EMPI_OK) return st;
    if ((st = p->d_res.alloc(static_cast<size_t>(res_off))) != EMPI_OK) return st;
    if ((st = p->d_res32.alloc(static_cast<size_t>(res_off))) != EMPI_OK) return st;
    if ((st = p->d_cells.alloc(static_cast<size_t>(cell_off))) != EMPI_OK) return st;
    if ((st = p->d_tiles.alloc(static_cast<size_t>(tile_off))) != EMPI_OK) return st;
    if ((st = p->d_uscales.alloc(p->uscales_h.size())) != EMPI_OK) return st;
    if ((st = p->d_etab.alloc(etab_h.size())) != EMPI_OK) return st;
    for (ScaleDev& sd : p->scales_h) sd.etab = p->d_etab.p + reinterpret_cast<uintptr_t>(sd.etab);
    if ((st = p->d_atoms.alloc(static_cast<size_t>(atom_off))) != EMPI_OK) return st;
    if ((st = p->d_energies.alloc(static_cast<size_t>(energy_off))) != EMPI_OK) return st;
    if ((st = p->d_fitv.alloc(p->units.size() * static_cast<size_t>(max_c))) != EMPI_OK) return st;
    if ((st = p->d_scratch.alloc(static_cast<size_t>(scratch_off))) != EMPI_OK) return st;
    if ((st = p->d_counter.alloc(kCounterWords)) != EMPI_OK) return st;
    for (size_t ui = 0; ui < p->units.size(); ++ui) {
        if (p->units[ui].C > 1) {
            p->mc_units.push_back(static_cast<int>(ui));
            p->mc_cmax = std::max(p->mc_cmax, p->units[ui].C);
        }
    }
    if ((st = p->d_mc.alloc(p->mc_units.size())) != EMPI_OK) return st;
    if ((st = p->d_epart.alloc(p->mc_units.size() * static_cast<size_t>(p->mc_cmax))) != EMPI_OK) return st;
    for (size_t ui = 0; ui < p->units.size(); ++ui) {
        UnitDev& u = p->units[ui];
        u.atoms = p->d_atoms.p + reinterpret_cast<long long>(u.atoms);
        u.energies = p->d_energies.p + reinterpret_cast<long long>(u.energies);
        u.fitv = p->d_fitv.p + ui * static_cast<size_t>(max_c);
    }
    hipStream_t s0 = nullptr;
    if ((st = p->d_scales.upload(p->scales_h.data(), p->scales_h.size(), s0)) != EMPI_OK) return st;
    if ((st = p->d_units.upload(p->units.data(), p->units.size(), s0)) != EMPI_OK) return st;
    if ((st = p->d_etab.upload(etab_h.data(), etab_h.size(), s0)) != EMPI_OK) return st;
    if ((st = p->d_mc.upload(p->mc_units.data(), p->mc_units.size(), s0)) != EMPI_OK) return st;
    std::vector<ClassDev> cds;
    for (int mode = 0; mode < 3; ++mode) {
        for (int c = 0; c < kNumClasses; ++c) {
            ClassJobs& cj = p->classes[mode][c];
            if (cj.jobs.empty()) continue;
            if ((st = cj.d.alloc(cj.jobs.size())) != EMPI_OK) return st;
            if ((st = cj.d.upload(cj.jobs.data(), cj.jobs.size(), s0)) != EMPI_OK) return st;
            if (cj.region_items > INT_MAX / 2) return EMPI_UNSUPPORTED;
            cj.max_items = std::max<int>(static_cast<int>(cj.jobs.size()), static_cast<int>(cj.region_items));
            if ((st = cj.items.alloc(static_cast<size_t>(cj.max_items))) != EMPI_OK) return st;
            if ((st = cj.work.alloc(static_cast<size_t>(cj.max_items) + 1)) != EMPI_OK) return st;
            if ((st = cj.count.alloc(1)) != EMPI_OK) return st;
            ClassDev cd{};
            cd.jobs = cj.d.p;
            cd.items = cj.items.p;
            cd.work = cj.work.p;
            cd.count = cj.count.p;
            cd.njobs = static_cast<int>(cj.jobs.size());
            cd.max_items = cj.max_items;
            cd.ncell = cells_per_wg(1 << (c + kMinLgQ));
            cd.cls = c;
            cj.slot = static_cast<int>(cds.size());
            cds.push_back(cd);
        }
    }
    p->nclass_dev = static_cast<int>(cds.size());
    if ((st = p->d_classes.alloc(cds.size())) != EMPI_OK) return st;
    if ((st = p->d_classes.upload(cds.data(), cds.size(), s0)) != EMPI_OK) return st;
    for (UnitScale& us : p->uscales_h)
        us.cls_slot = p->classes[us.cls_slot / kNumClasses][us.cls_slot % kNumClasses].slot;
    if ((st = p->d_uscales.upload(p->uscales_h.data(), p->uscales_h.size(), s0)) != EMPI_OK) return st;
    if ((st = p->d_regtab.alloc(p->regtab_h.size())) != EMPI_OK) return st;
    if ((st = p->d_regtab.upload(p->regtab_h.data(), p->regtab_h.size(), s0)) != EMPI_OK) return st;
    EMPI_HIP(hipStreamSynchronize(s0));
    return EMPI_OK;
}

// column kernels of every class over the current work lists
empi_status launch_class_columns(empi_plan* p, hipStream_t s) {
    for (int mc = 0; mc < 3 * kNumClasses; ++mc) {
        const int mode = mc / kNumClasses, c = mc % kNumClasses;
        ClassJobs& cj = p->classes[mode][c];
        if (cj.jobs.empty() || cj.tiles == 0) continue;
        std::pair<hipEvent_t, hipEvent_t>* evp = nullptr;
        if (p->profiling) {
            if (p->ev_used[c] == p->ev[c].size()) {
                std::pair<hipEvent_t, hipEvent_t> e;
                EMPI_HIP(hipEventCreate(&e.first));
                EMPI_HIP(hipEventCreate(&e.second));
                p->ev[c].push_back(e);
            }
            evp = &p->ev[c][p->ev_used[c]++];
            EMPI_HIP(hipEventRecord(evp->first, s));
        }
        const int choice = p->tune[mode][c];
        // in proposal mode the single-channel kernels also write the cells' low-bin bests
        const bool low = p->prop_mode != EMPI_PROPOSAL_OFF && mode == kModeSingle;
        column_launcher(mode, c, choice, low)(column_grid(1 << (c + kMinLgQ), cj.tiles, choice / kWaveVariants), s,
                                              p->d_classes.p, cj.slot, p->d_scales.p, p->d_units.p, p->d_state.p,
                                              p->d_res32.p, p->d_cells.p, p->d_tiles.p,
                                              p->d_counter.p, low ? p->d_celllow.p : nullptr);
        if (evp) EMPI_HIP(hipEventRecord(evp->second, s));
    }
    EMPI_HIP(hipGetLastError());
    return EMPI_OK;
}

// The selection step of one iteration: k_select for every unit; the
// multichannel units then finish it one workgroup per channel.
empi_status launch_select(empi_plan* p, hipStream_t s) {
    const int U = static_cast<int>(p->units.size());
    const int M = static_cast<int>(p->mc_units.size());
    hipLaunchKernelGGL(k_select<kSelectThreads>, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_scales.p, p->d_uscales.p, p->d_tiles.p, p->d_res.p, p->d_res32.p, 3, nullptr, M > 0 ? 1 : 0);
    if (M > 0) {
        const dim3 grid(p->mc_cmax, M);
        hipLaunchKernelGGL(k_fit_mc<kSelectThreads>, grid, dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p,
                           p->d_scales.p, p->d_mc.p, p->d_res.p);
        hipLaunchKernelGGL(k_params_mc, dim3((M + 63) / 64), dim3(64), 0, s, p->d_units.p, p->d_state.p,
                           p->d_scales.p, p->d_mc.p, M);
        hipLaunchKernelGGL(k_sub_mc<kSelectThreads>, grid, dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p,
                           p->d_mc.p, p->d_res.p, p->d_res32.p, p->d_epart.p, p->mc_cmax);
        hipLaunchKernelGGL(k_stop_mc, dim3((M + 63) / 64), dim3(64), 0, s, p->d_units.p, p->d_state.p, p->d_mc.p, M,
                           p->d_epart.p, p->mc_cmax);
    }
    EMPI_HIP(hipGetLastError());
    return EMPI_OK;
}

empi_status launch_all_columns(empi_plan* p, hipStream_t s, int full) {
    if (p->nclass_dev == 0) return EMPI_OK;
    hipLaunchKernelGGL(k_plan, dim3(p->nclass_dev), dim3(kPlanThreads), 0, s, p->d_classes.p, p->d_state.p,
                       p->d_scales.p, full);
    return launch_class_columns(p, s);
}

// ------------------------------------------------------------ proposal host side

empi_status proposal_prepare(empi_plan* p) {
    if (p->prop_ready) return EMPI_OK;
    for (const UnitDev& u : p->units)
        if (u.C != 1 || u.constraint) return EMPI_UNSUPPORTED;
    const size_t U = p->units.size();
    std::vector<long long> prefix_off(U), feat_off(U);
    long long pre = 0, feat = 0, rows = 0;
    for (size_t ui = 0; ui < U; ++ui) {
        const UnitDev& u = p->units[ui];
        prefix_off[ui] = pre;
        pre += u.N + 1;
        feat_off[ui] = feat;
        feat += static_cast<long long>(u.B) * kFeatBins;
        for (int l = 0; l < u.L; ++l) {
            for (int b = 0; b < u.B; ++b) {
                const RegionDev& r = p->regtab_h[u.regtab_off + l * u.B + b];
                rows += r.tlo <= r.thi;
            }
        }
    }
    if (rows > INT_MAX / kFeatDim) return EMPI_UNSUPPORTED;
    p->max_rows = static_cast<int>(rows);
    empi_status st;
    hipStream_t s0 = nullptr;
    if ((st = p->d_regions.alloc(static_cast<size_t>(p->nregions))) != EMPI_OK) return st;
    if ((st = p->d_celllow.alloc(p->d_cells.n)) != EMPI_OK) return st;
    if ((st = p->d_prefix.alloc(static_cast<size_t>(pre))) != EMPI_OK) return st;
    if ((st = p->d_prefix_off.alloc(U)) != EMPI_OK) return st;
    if ((st = p->d_blockfeat.alloc(static_cast<size_t>(feat))) != EMPI_OK) return st;
    if ((st = p->d_feat_off.alloc(U)) != EMPI_OK) return st;
    const size_t R = static_cast<size_t>(std::max(1, p->max_rows));
    if ((st = p->d_rows.alloc(R * kFeatDim)) != EMPI_OK) return st;
    if ((st = p->d_rowmap.alloc(R)) != EMPI_OK) return st;
    if ((st = p->d_rowref.alloc(R)) != EMPI_OK) return st;
    if ((st = p->d_scores.alloc(R)) != EMPI_OK) return st;
    if ((st = p->d_rowcount.alloc(1)) != EMPI_OK) return st;
    if ((st = p->d_pstats.alloc(5)) != EMPI_OK) return st;
    if ((st = p->d_pbest.alloc(U)) != EMPI_OK) return st;
    if ((st = p->d_prefix_off.upload(prefix_off.data(), U, s0)) != EMPI_OK) return st;
    if ((st = p->d_feat_off.upload(feat_off.data(), U, s0)) != EMPI_OK) return st;
    EMPI_HIP(hipStreamSynchronize(s0));
    p->prop_ready = true;
    return EMPI_OK;
}

// regions flagged "evaluate": work lists, column kernels, region bests
empi_status proposal_pass(empi_plan* p, hipStream_t s, unsigned long long* stat) {
    const int U = static_cast<int>(p->units.size());
    hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, s, p->d_classes.p, p->nclass_dev, nullptr);
    hipLaunchKernelGGL(k_region_items, dim3(U), dim3(kRegionThreads), 0, s, p->d_classes.p, p->d_units.p,
                       p->d_state.p, p->d_scales.p, p->d_uscales.p, p->d_regtab.p, p->d_regions.p, stat);
    hipLaunchKernelGGL(k_scan_items, dim3(p->nclass_dev), dim3(kPlanThreads), 0, s, p->d_classes.p);
    EMPI_HIP(hipGetLastError());
    empi_status st = launch_class_columns(p, s);
    if (st != EMPI_OK) return st;
    hipLaunchKernelGGL(k_regions_reduce, dim3(kSlices, U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_scales.p, p->d_regtab.p, p->d_uscales.p, p->d_tiles.p, p->d_celllow.p, p->d_regions.p, 0);
    EMPI_HIP(hipGetLastError());
    return EMPI_OK;
}

empi_status launch_features(empi_plan* p, hipStream_t s, int init) {
    const int U = static_cast<int>(p->units.size());
    hipLaunchKernelGGL(k_reset_counts, dim3(1), dim3(64), 0, s, p->d_classes.p, 0, p->d_rowcount.p);
    hipLaunchKernelGGL(k_prefix, dim3(U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p, p->d_res32.p,
                       p->d_prefix.p, p->d_prefix_off.p);
    hipLaunchKernelGGL(k_block_spectra, dim3(init ? kSlices : 4, U), dim3(kRegionThreads), 0, s, p->d_units.p,
                       p->d_state.p, p->d_res32.p, p->d_blockfeat.p, p->d_feat_off.p, init);
    hipLaunchKernelGGL(k_clean_best, dim3(U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_regions.p, p->d_pbest.p);
    hipLaunchKernelGGL(k_rows, dim3(kSlices, U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_scales.p, p->d_regtab.p, p->d_regions.p, p->d_prefix.p, p->d_prefix_off.p,
                       p->d_blockfeat.p, p->d_feat_off.p, p->d_rows.p, p->d_rowmap.p, p->d_rowref.p,
                       p->d_rowcount.p, p->max_rows, p->d_pbest.p, p->delta,
                       p->prop_mode == EMPI_PROPOSAL_COLLECT ? 1 : 0);
    EMPI_HIP(hipGetLastError());
    return EMPI_OK;
}

// one proposal iteration: select + subtract among clean regions, mark,
// features and bounds, scores, proposed pass, certification, fallback pass
empi_status launch_proposal_iteration(empi_plan* p, hipStream_t s) {
    const int U = static_cast<int>(p->units.size());
    empi_status st;
    hipLaunchKernelGGL(k_select<kSelectThreads>, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_scales.p, p->d_uscales.p, p->d_tiles.p, p->d_res.p, p->d_res32.p, 3, p->d_regions.p, 0);
    hipLaunchKernelGGL(k_mark_dirty, dim3(U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p, p->d_scales.p,
                       p->d_regtab.p, p->d_regions.p);
    EMPI_HIP(hipGetLastError());
    if ((st = launch_features(p, s, 0)) != EMPI_OK) return st;
    const bool collect = p->prop_mode == EMPI_PROPOSAL_COLLECT;
    const bool scored = !collect && p->net != nullptr;
    if (scored) {
        if ((st = empi_mlp_forward_counted(p->net, p->d_rows.p, p->d_rowcount.p, p->max_rows, p->d_scores.p, s)))
            return st;
        hipLaunchKernelGGL(k_scores_scatter, dim3(256), dim3(256), 0, s, p->d_scores.p, p->d_rowmap.p,
                           p->d_rowcount.p, p->max_rows, p->d_regions.p);
    }
    hipLaunchKernelGGL(k_choose, dim3(U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p, p->d_regions.p,
                       p->top_k, scored ? 1 : 0, collect ? 1 : 0);
    EMPI_HIP(hipGetLastError());
    if ((st = proposal_pass(p, s, p->d_pstats.p + 4)) != EMPI_OK) return st;
    if (collect && p->d_traincount.p) {
        hipLaunchKernelGGL(k_collect_rows, dim3(256), dim3(256), 0, s, p->d_rows.p, p->d_rowmap.p, p->d_rowref.p,
                           p->d_rowcount.p, p->max_rows, p->d_regions.p, p->d_trainX.p, p->d_trainT.p,
                           p->d_traincount.p, static_cast<unsigned long long>(p->d_trainT.n), p->train_stride);
    }
    hipLaunchKernelGGL(k_certify, dim3(U), dim3(kRegionThreads), 0, s, p->d_units.p, p->d_state.p, p->d_regions.p,
                       p->delta, p->d_pstats.p);
    EMPI_HIP(hipGetLastError());
    return proposal_pass(p, s, nullptr);
}

empi_status collect_profile(empi_plan* p) {
    for (int c = 0; c < kNumClasses; ++c) {
        p->class_ms[c] = 0.0;
        p->class_launches[c] = 0;
        for (size_t i = 0; i < p->ev_used[c]; ++i) {
            float ms = 0.0f;
            EMPI_HIP(hipEventElapsedTime(&ms, p->ev[c][i].first, p->ev[c][i].second));
            p->class_ms[c] += ms;
            p->class_launches[c] += 1;
        }
        p->ev_used[c] = 0;
    }
    return EMPI_OK;
}

// Autotuner: times every variant of every column-kernel class on the full
// map build of this plan's batch (HIP events on the plan's stream, best of
// `reps`) and keeps the fastest per class.
empi_status plan_autotune(empi_plan* p, const float* signal, int on_device, int reps, hipStream_t user) {
    empi_status st;
    EMPI_HIP(hipStreamSynchronize(user));
    if (!p->stream) EMPI_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    hipStream_t s = p->stream;
    const float* in = signal;
    if (!on_device) {
        if (!p->d_input.p || p->d_input.n < static_cast<size_t>(p->total_in)) {
            if ((st = p->d_input.alloc(static_cast<size_t>(p->total_in))) != EMPI_OK) return st;
        }
        if ((st = p->d_input.upload(signal, static_cast<size_t>(p->total_in), s)) != EMPI_OK) return st;
        in = p->d_input.p;
    }
    const int U = static_cast<int>(p->units.size());
    hipLaunchKernelGGL(k_init, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p, in, p->d_res.p);
    hipLaunchKernelGGL(k_mirror, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p, p->d_res.p,
                       p->d_res32.p);
    if (p->nclass_dev)
        hipLaunchKernelGGL(k_plan, dim3(p->nclass_dev), dim3(kPlanThreads), 0, s, p->d_classes.p, p->d_state.p,
                           p->d_scales.p, 1);
    EMPI_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    EMPI_HIP(hipEventCreate(&e0));
    EMPI_HIP(hipEventCreate(&e1));
    empi_status result = EMPI_OK;
    // every variant of every class on the current work lists, best of reps
    static thread_local float ms[2][3][kNumClasses][kTuneChoices];
    auto time_lists = [&](int pass) {
        for (int mode = 0; mode < 3 && result == EMPI_OK; ++mode) {
            for (int c = 0; c < kNumClasses && result == EMPI_OK; ++c) {
                ClassJobs& cj = p->classes[mode][c];
                if (cj.jobs.empty() || cj.tiles == 0) continue;
                // with the proposal on, the single-channel classes run (and are
                // timed as) their low-bin variants
                const bool low = p->prop_mode != EMPI_PROPOSAL_OFF && mode == kModeSingle && p->d_celllow.p;
                for (int v = 0; v < kTuneChoices && result == EMPI_OK; ++v) {
                    float best = INFINITY;
                    for (int r = 0; r < std::max(1, reps); ++r) {
                        if (hipEventRecord(e0, s) != hipSuccess) result = EMPI_DEVICE;
                        column_launcher(mode, c, v, low)(column_grid(1 << (c + kMinLgQ), cj.tiles, v / kWaveVariants),
                                                         s, p->d_classes.p, cj.slot, p->d_scales.p, p->d_units.p,
                                                         p->d_state.p, p->d_res32.p, p->d_cells.p, p->d_tiles.p,
                                                         nullptr, low ? p->d_celllow.p : nullptr);
                        float t = 0.0f;
                        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                            result = EMPI_DEVICE;
                        best = std::min(best, t);
                    }
                    ms[pass][mode][c][v] = best;
                }
            }
        }
    };
    // the full map build, then the strips one iteration re-evaluates (the
    // iteration's selection and work lists from a full build with the
    // current table); a class re-run on the same list recomputes the same cells
    time_lists(0);
    if (result == EMPI_OK) result = launch_class_columns(p, s);
    if (result == EMPI_OK) result = launch_select(p, s);
    if (result == EMPI_OK && p->nclass_dev)
        hipLaunchKernelGGL(k_plan, dim3(p->nclass_dev), dim3(kPlanThreads), 0, s, p->d_classes.p, p->d_state.p,
                           p->d_scales.p, 0);
    if (result == EMPI_OK) time_lists(1);
    // keep the variant with the least full build + iterationMax strips
    for (int mode = 0; mode < 3 && result == EMPI_OK; ++mode) {
        for (int c = 0; c < kNumClasses; ++c) {
            ClassJobs& cj = p->classes[mode][c];
            if (cj.jobs.empty() || cj.tiles == 0) continue;
            double best = INFINITY;
            for (int v = 0; v < kTuneChoices; ++v) {
                const double cost = ms[0][mode][c][v] + static_cast<double>(p->maxIter) * ms[1][mode][c][v];
                if (cost < best) {
                    best = cost;
                    p->tune[mode][c] = v;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->drop_graphs();
    return result;
}

empi_status plan_run(empi_plan* p, const float* signal, int on_device, hipStream_t user) {
    const float* in = signal;
    empi_status st;
    // work runs on the plan's stream (capturable); the caller's stream is drained
    // first.  An allocation-free run works on the caller's stream itself.
    EMPI_HIP(hipStreamSynchronize(user));
    if (p->noalloc) p->stream = user;
    else if (!p->stream) EMPI_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    hipStream_t s = p->stream;
    if (!signal) {
        // the input empi_plan_stage copied (its copy is complete)
        std::lock_guard<std::mutex> g(p->stage_mu);
        if (p->stage_ready < 0) return EMPI_INVALID_ARGUMENT;
        p->stage_active = p->stage_ready;
        p->stage_ready = -1;
        in = p->d_stage_buf[p->stage_active].p;
    } else if (!on_device) {
        if (!p->d_input.p || p->d_input.n < static_cast<size_t>(p->total_in)) {
            if ((st = p->d_input.alloc(static_cast<size_t>(p->total_in))) != EMPI_OK) return st;
        }
        if ((st = p->d_input.upload(signal, static_cast<size_t>(p->total_in), s)) != EMPI_OK) return st;
        in = p->d_input.p;
    }
    const int U = static_cast<int>(p->units.size());
    EMPI_HIP(hipMemsetAsync(p->d_counter.p, 0, sizeof(unsigned long long) * kCounterWords, s));
    EMPI_HIP(hipMemsetAsync(p->d_atoms.p, 0, sizeof(empi_atom) * p->d_atoms.n, s));
    for (int c = 0; c < kNumClasses; ++c) p->ev_used[c] = 0;
    {
        std::vector<double> nanv(p->d_energies.n, NAN);
        EMPI_HIP(hipMemcpyAsync(p->d_energies.p, nanv.data(), sizeof(double) * nanv.size(), hipMemcpyHostToDevice, s));
        EMPI_HIP(hipStreamSynchronize(s));
    }
    hipLaunchKernelGGL(k_init, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p, in, p->d_res.p);
    hipLaunchKernelGGL(k_mirror, dim3(U), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p, p->d_res.p,
                       p->d_res32.p);
    EMPI_HIP(hipGetLastError());
    if ((st = launch_all_columns(p, s, 1)) != EMPI_OK) return st;
    std::vector<UnitState> hs(U);
    const bool graph = !p->profiling && !p->noalloc && std::getenv("EMPI_NO_GRAPH") == nullptr;
    const bool prop = p->prop_mode != EMPI_PROPOSAL_OFF;
    if (prop) {
        EMPI_HIP(hipMemsetAsync(p->d_pstats.p, 0, sizeof(unsigned long long) * 5, s));
        hipLaunchKernelGGL(k_regions_reduce, dim3(kSlices, U), dim3(kRegionThreads), 0, s, p->d_units.p,
                           p->d_state.p, p->d_scales.p, p->d_regtab.p, p->d_uscales.p, p->d_tiles.p, p->d_celllow.p,
                           p->d_regions.p, 1);
        EMPI_HIP(hipGetLastError());
        if ((st = launch_features(p, s, 1)) != EMPI_OK) return st;
        if (graph && !p->prop_exec) {
            EMPI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            empi_status cst = launch_proposal_iteration(p, s);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (cst != EMPI_OK) return cst;
            if (e != hipSuccess) return EMPI_DEVICE;
            p->prop_graph = g;
            EMPI_HIP(hipGraphInstantiate(&p->prop_exec, g, nullptr, nullptr, 0));
        }
    }
    if (graph && !prop && !p->iter_exec) {
        // capture one iteration; every kernel reads its per-iteration state from HBM
        EMPI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        empi_status cst = launch_select(p, s);
        if (cst == EMPI_OK) cst = launch_all_columns(p, s, 0);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (cst != EMPI_OK) return cst;
        if (e != hipSuccess) return EMPI_DEVICE;
        p->iter_graph = g;
        EMPI_HIP(hipGraphInstantiate(&p->iter_exec, g, nullptr, nullptr, 0));
    }
    // End of the loop: an allocating run queues at most kPollLag iterations
    // past the last one whose active-unit count it has read (so the device
    // never waits for the host, and at most kPollLag empty iterations follow
    // the last stop); an allocation-free run reads the unit states every 8
    // iterations through pageable memory.
    const bool ring = !p->noalloc;
    if (ring) {
        if (!p->d_nactive.p && (st = p->d_nactive.alloc(1)) != EMPI_OK) return st;
        if (!p->h_nactive) {
            EMPI_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_nactive), sizeof(int) * kPollRing,
                                   hipHostMallocDefault));
            for (hipEvent_t& e : p->poll_ev) EMPI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }
    for (int it = 0; it < p->maxIter; ++it) {
        if (ring && it >= kPollLag) {
            const int slot = (it - kPollLag) % kPollRing;
            EMPI_HIP(hipEventSynchronize(p->poll_ev[slot]));
            if (p->h_nactive[slot] == 0) break;
        }
        // the last iteration's atom is not subtracted: no columns after it
        const bool last = it + 1 == p->maxIter;
        if (prop) {
            if (graph) {
                EMPI_HIP(hipGraphLaunch(p->prop_exec, s));
            } else if ((st = launch_proposal_iteration(p, s)) != EMPI_OK) {
                return st;
            }
        } else if (graph && !last) {
            EMPI_HIP(hipGraphLaunch(p->iter_exec, s));
        } else {
            if ((st = launch_select(p, s)) != EMPI_OK) return st;
            if (!last) {
                if ((st = launch_all_columns(p, s, 0)) != EMPI_OK) return st;
            }
        }
        if (ring) {
            const int slot = it % kPollRing;
            hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, s, p->d_state.p, U, p->d_nactive.p);
            EMPI_HIP(hipGetLastError());
            EMPI_HIP(hipMemcpyAsync(p->h_nactive + slot, p->d_nactive.p, sizeof(int), hipMemcpyDeviceToHost, s));
            EMPI_HIP(hipEventRecord(p->poll_ev[slot], s));
        } else if ((it & 7) == 7) {
            EMPI_HIP(hipMemcpyAsync(hs.data(), p->d_state.p, sizeof(UnitState) * U, hipMemcpyDeviceToHost, s));
            EMPI_HIP(hipStreamSynchronize(s));
            bool any = false;
            for (const UnitState& x : hs) any |= x.active != 0;
            if (!any) break;
        }
    }
    // MMP2 refits
    for (const ItemInfo& info : p->items) {
        if (info.sum_atom_off < 0) continue;
        const UnitDev& u = p->units[info.units[0]];
        (void)u;
        hipLaunchKernelGGL(k_refit<kSelectThreads>, dim3(info.item.channels), dim3(kSelectThreads), 0, s,
                           p->d_atoms.p + info.sum_atom_off, p->d_state.p + info.units[0], in, info.in_off,
                           info.item.channels, info.item.samples, info.item.fs, p->d_scratch.p + info.scratch_off,
                           p->d_atoms.p + info.atom_off, info.item.iterationMax);
        EMPI_HIP(hipGetLastError());
    }
    p->h_atoms.resize(p->d_atoms.n);
    p->h_energies.resize(p->d_energies.n);
    p->h_state.resize(U);
    EMPI_HIP(hipMemcpyAsync(p->h_atoms.data(), p->d_atoms.p, sizeof(empi_atom) * p->d_atoms.n, hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipMemcpyAsync(p->h_energies.data(), p->d_energies.p, sizeof(double) * p->d_energies.n,
                            hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipMemcpyAsync(p->h_state.data(), p->d_state.p, sizeof(UnitState) * U, hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipMemcpyAsync(p->h_counter_copies, p->d_counter.p, sizeof(unsigned long long) * kCounterWords,
                            hipMemcpyDeviceToHost, s));
    if (prop) {
        EMPI_HIP(hipMemcpyAsync(p->h_pstats, p->d_pstats.p, sizeof(unsigned long long) * 5, hipMemcpyDeviceToHost, s));
    } else {
        std::fill(p->h_pstats, p->h_pstats + 5, 0ull);
    }
    EMPI_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < 2 * kNumClasses; ++i) {
        p->h_counter[i] = 0;
        for (int x = 0; x < kCounterCopies; ++x) p->h_counter[i] += p->h_counter_copies[2 * kNumClasses * x + i];
    }
    if (p->profiling) {
        empi_status pst = collect_profile(p);
        if (pst != EMPI_OK) return pst;
    }
    p->has_results = true;
    return EMPI_OK;
}

}  // namespace

// ------------------------------------------------------------------- C ABI

// The argmax keys over a planted map: per cell the packed bin keys, per tile
// the tile keys, and the selection over tiles, in one workgroup (a test hook:
// planted maps are small).  Products [scale][t][g][b], bins [scale][g][b].
template <int NT>
__global__ __launch_bounds__(NT) void k_debug_argmax(int L, const int* __restrict__ T, const int* __restrict__ G,
                                                     const int* __restrict__ ncell, const int* __restrict__ B,
                                                     const long long* __restrict__ poff,
                                                     const long long* __restrict__ boff,
                                                     const float* __restrict__ products,
                                                     const int* __restrict__ bins, int* __restrict__ out,
                                                     float* __restrict__ best) {
    __shared__ float redp[NT];
    __shared__ unsigned long long redk[NT];
    const int tid = threadIdx.x;
    SelBest b{-1.0f, ~0ull};
    for (int l = 0; l < L; ++l) {
        const int cells = T[l] * G[l], nc = ncell[l];
        const int tiles = (cells + nc - 1) / nc;
        for (int i = tid; i < tiles; i += NT) {
            unsigned long long tb = 0ull;
            for (int c = 0; c < nc && i * nc + c < cells; ++c) {
                const int cg = i * nc + c, g = cg % G[l];
                unsigned long long cb = 0ull;
                for (int q = 0; q < B[l]; ++q)
                    cb = max(cb, pack_best(products[poff[l] + static_cast<long long>(cg) * B[l] + q],
                                           static_cast<unsigned>(bins[boff[l] + g * B[l] + q])));
                tb = max(tb, tile_key(cb, c, kTileCellBits));
            }
            const SelBest cand = sel_of_tile(tb, l, i, nc, G[l], kTileCellBits);
            if (sel_better(cand, b)) b = cand;
        }
    }
    redp[tid] = b.p;
    redk[tid] = b.key;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const SelBest o{redp[tid + s], redk[tid + s]}, m{redp[tid], redk[tid]};
            if (sel_better(o, m)) {
                redp[tid] = o.p;
                redk[tid] = o.key;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = static_cast<int>(redk[0] >> 52);
        out[1] = static_cast<int>((redk[0] >> 26) & ((1ull << 26) - 1));
        out[2] = static_cast<int>(redk[0] & ((1ull << 26) - 1));
        *best = redp[0];
    }
}

extern "C" {

const char* empi_strerror(int status) {
    switch (status) {
        case EMPI_OK: return "ok";
        case EMPI_SIGNAL_IS_EMPTY: return "signalIsEmpty";
        case EMPI_ENERGY_MAP_IS_EMPTY: return "energyMapIsEmpty";
        case EMPI_TOO_LONG: return "signalFileIsTooLongForThisDecomposition";
        case EMPI_INVALID_ARGUMENT: return "invalidArgument";
        case EMPI_NO_MEMORY: return "insufficientMemory";
        case EMPI_DEVICE: return "deviceError";
        case EMPI_UNSUPPORTED: return "unsupportedSize";
        default: return "internalError";
    }
}

const char* empi_version(void) { return "empi-mi355x 0.1 (gfx950)"; }

empi_status empi_dict_info(const empi_dict_params* dict, double fs, int samples, empi_scale_info* scales,
                           int max_scales, int* count, long long* entries) {
    if (!dict) return EMPI_INVALID_ARGUMENT;
    Dictionary d;
    empi_status st;
    try {
        st = build_dictionary(*dict, fs, samples, &d);
    } catch (const std::bad_alloc&) {
        return EMPI_NO_MEMORY;
    }
    if (st != EMPI_OK) return st;
    const int L = static_cast<int>(d.scales.size());
    for (int i = 0; i < L && i < max_scales && scales; ++i) {
        scales[i].scale = d.scales[i].s;
        scales[i].Nfft = d.scales[i].Nfft;
        scales[i].tCount = d.scales[i].T;
        scales[i].fCount = d.scales[i].F;
        scales[i].reserved = 0;
    }
    if (count) *count = L;
    if (entries) *entries = d.entries;
    return EMPI_OK;
}

empi_status empi_plan_create(const empi_item* items, int count, empi_plan** out) {
    if (!out) return EMPI_INVALID_ARGUMENT;
    *out = nullptr;
    empi_plan* p = new (std::nothrow) empi_plan();
    if (!p) return EMPI_NO_MEMORY;
    empi_status st;
    try {
        st = plan_build(p, items, count, 0);
    } catch (const std::bad_alloc&) {
        st = EMPI_NO_MEMORY;
    } catch (...) {
        st = EMPI_INTERNAL;
    }
    if (st != EMPI_OK) {
        delete p;
        return st;
    }
    *out = p;
    return EMPI_OK;
}

empi_status empi_plan_stage(empi_plan* p, const float* signal) {
    if (!p || !signal || p->noalloc) return EMPI_INVALID_ARGUMENT;
    int slot;
    {
        std::lock_guard<std::mutex> g(p->stage_mu);
        if (p->stage_ready >= 0) return EMPI_INVALID_ARGUMENT;   // one staged input at a time
        slot = p->stage_active == 0 ? 1 : 0;
    }
    try {
        DevBuf<float>& b = p->d_stage_buf[slot];
        empi_status st;
        if (!b.p || b.n < static_cast<size_t>(p->total_in)) {
            if ((st = b.alloc(static_cast<size_t>(p->total_in))) != EMPI_OK) return st;
        }
        if (!p->upload) EMPI_HIP(hipStreamCreateWithFlags(&p->upload, hipStreamNonBlocking));
        if ((st = b.upload(signal, static_cast<size_t>(p->total_in), p->upload)) != EMPI_OK) return st;
        EMPI_HIP(hipStreamSynchronize(p->upload));
    } catch (const std::bad_alloc&) {
        return EMPI_NO_MEMORY;
    }
    std::lock_guard<std::mutex> g(p->stage_mu);
    p->stage_ready = slot;
    return EMPI_OK;
}

empi_status empi_plan_run(empi_plan* plan, const float* signal, int signal_on_device, void* stream) {
    if (!plan || (!signal && signal_on_device)) return EMPI_INVALID_ARGUMENT;
    try {
        return plan_run(plan, signal, signal_on_device, static_cast<hipStream_t>(stream));
    } catch (const std::bad_alloc&) {
        return EMPI_NO_MEMORY;
    } catch (...) {
        return EMPI_INTERNAL;
    }
}

empi_status empi_plan_item_result(const empi_plan* p, int i, empi_atom* atoms, int* counts, double* energies) {
    if (!p || i < 0 || i >= static_cast<int>(p->items.size())) return EMPI_INVALID_ARGUMENT;
    if (!p->has_results) return EMPI_INVALID_ARGUMENT;
    const ItemInfo& info = p->items[i];
    const int C = info.item.channels, I = info.item.iterationMax;
    empi_status status = EMPI_OK;
    std::vector<int> cnt(C, 0);
    if (info.item.variant == EMPI_SMP) {
        for (int c = 0; c < C; ++c) {
            const UnitState& s = p->h_state[info.units[c]];
            cnt[c] = s.iteration;
            if (s.status != EMPI_OK && status == EMPI_OK) status = s.status;
        }
    } else {
        const UnitState& s = p->h_state[info.units[0]];
        for (int c = 0; c < C; ++c) cnt[c] = s.iteration;
        status = s.status;
    }
    if (counts) std::memcpy(counts, cnt.data(), sizeof(int) * C);
    if (atoms) std::memcpy(atoms, p->h_atoms.data() + info.atom_off, sizeof(empi_atom) * C * I);
    if (energies)
        std::memcpy(energies, p->h_energies.data() + info.energy_off, sizeof(double) * info.rows * (I + 1));
    return status;
}

namespace {

// device address of `scratch` on the current device, looked up once per device
empi_status scratch_base(void** out) {
    constexpr int kMaxDevices = 64;
    static void* base[kMaxDevices] = {};
    static std::mutex lock;
    int dev = 0;
    EMPI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return EMPI_DEVICE;
    std::lock_guard<std::mutex> g(lock);
    if (!base[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            std::fprintf(stderr, "empi_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                         hipGetErrorString(e));
            return EMPI_DEVICE;
        }
        base[dev] = ptr;
    }
    *out = base[dev];
    return EMPI_OK;
}

struct ArenaScope {
    explicit ArenaScope(Arena* a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

empi_status too_big(const Arena& a) {
    std::fprintf(stderr,
                 "empi_hip: empi_decompose_noalloc: this item needs %zu bytes of device scratch, the maximum is %zu "
                 "(unsupportedSize)\n", a.wanted, a.size);
    return EMPI_UNSUPPORTED;
}

}  // namespace

empi_status empi_decompose_noalloc(const empi_item* item, const float* signal, int signal_on_device, empi_atom* atoms,
                                   int* counts, double* energies, void* stream) {
    if (!item || !signal) return EMPI_INVALID_ARGUMENT;
    void* base = nullptr;
    empi_status st = scratch_base(&base);
    if (st != EMPI_OK) return st;
    Arena arena;
    arena.base = static_cast<char*>(base);
    arena.size = kScratchBytes;
    ArenaScope scope(&arena);
    try {
        empi_plan p;
        p.noalloc = true;
        st = plan_build(&p, item, 1, 0);
        if (st == EMPI_OK) st = plan_run(&p, signal, signal_on_device, static_cast<hipStream_t>(stream));
        if (st == EMPI_UNSUPPORTED && arena.wanted > arena.size) return too_big(arena);
        if (st != EMPI_OK) return st;
        return empi_plan_item_result(&p, 0, atoms, counts, energies);
    } catch (const std::bad_alloc&) {
        return EMPI_NO_MEMORY;
    } catch (...) {
        return EMPI_INTERNAL;
    }
}

size_t empi_plan_item_atom_count(const empi_plan* p, int i) {
    if (!p || i < 0 || i >= static_cast<int>(p->items.size())) return 0;
    return static_cast<size_t>(p->dicts[p->items[i].dict].entries);
}

double empi_plan_entries_evaluated(const empi_plan* p) {
    if (!p) return 0.0;
    double n = 0.0;
    for (int c = 0; c < kNumClasses; ++c) n += static_cast<double>(p->h_counter[2 * c]);
    return n;
}

empi_status empi_plan_set_profiling(empi_plan* p, int on) {
    if (!p) return EMPI_INVALID_ARGUMENT;
    p->profiling = on != 0;
    return EMPI_OK;
}

empi_status empi_plan_kernel_stats(const empi_plan* p, int cls, int* sub_length, long long* launches,
                                   double* total_ms, double* entries, double* entry_channels) {
    if (!p || cls < 0 || cls >= kNumClasses) return EMPI_INVALID_ARGUMENT;
    if (sub_length) *sub_length = 1 << (cls + kMinLgQ);
    if (launches) *launches = p->class_launches[cls];
    if (total_ms) *total_ms = p->class_ms[cls];
    if (entries) *entries = static_cast<double>(p->h_counter[2 * cls]);
    if (entry_channels) *entry_channels = static_cast<double>(p->h_counter[2 * cls + 1]);
    return EMPI_OK;
}

int empi_plan_kernel_classes(void) { return kNumClasses; }

void empi_plan_destroy(empi_plan* plan) { delete plan; }

double empi_proposal_bound_factor(double scale, double fs, int Nfft) {
    if (!(scale > 0) || !(fs > 0) || Nfft < 2) return NAN;
    return proposal_bound_factor(scale, fs, Nfft);
}

int empi_plan_tuning_size(void) { return 3 * kNumClasses; }

empi_status empi_plan_autotune(empi_plan* p, const float* signal, int signal_on_device, int repeats, void* stream) {
    if (!p || !signal || repeats < 0) return EMPI_INVALID_ARGUMENT;
    try {
        return plan_autotune(p, signal, signal_on_device, repeats, static_cast<hipStream_t>(stream));
    } catch (const std::bad_alloc&) {
        return EMPI_NO_MEMORY;
    }
}

empi_status empi_plan_get_tuning(const empi_plan* p, int* out) {
    if (!p || !out) return EMPI_INVALID_ARGUMENT;
    for (int m = 0; m < 3; ++m)
        for (int c = 0; c < kNumClasses; ++c) out[m * kNumClasses + c] = p->tune[m][c];
    return EMPI_OK;
}

empi_status empi_plan_set_tuning(empi_plan* p, const int* in) {
    if (!p || !in) return EMPI_INVALID_ARGUMENT;
    for (int m = 0; m < 3; ++m)
        for (int c = 0; c < kNumClasses; ++c) {
            const int v = in[m * kNumClasses + c];
            if (v < 0 || v >= kTuneChoices) return EMPI_INVALID_ARGUMENT;
        }
    for (int m = 0; m < 3; ++m)
        for (int c = 0; c < kNumClasses; ++c) p->tune[m][c] = in[m * kNumClasses + c];
    p->drop_graphs();
    return EMPI_OK;
}

int empi_proposal_feature_dim(void) { return kFeatDim; }

empi_status empi_plan_set_proposal(empi_plan* p, int mode, empi_mlp* net, int top_k, double delta) {
    if (!p || mode < EMPI_PROPOSAL_OFF || mode > EMPI_PROPOSAL_COLLECT || top_k < 0 || !(delta >= 0.0) ||
        !(delta < 1.0))
        return EMPI_INVALID_ARGUMENT;
    if (net && empi_mlp_input_dim(net) != kFeatDim) return EMPI_INVALID_ARGUMENT;
    if (mode != EMPI_PROPOSAL_OFF) {
        empi_status st = proposal_prepare(p);
        if (st != EMPI_OK) return st;
    }
    if (p->prop_exec) (void)hipGraphExecDestroy(p->prop_exec);
    if (p->prop_graph) (void)hipGraphDestroy(p->prop_graph);
    p->prop_exec = nullptr;
    p->prop_graph = nullptr;
    p->prop_mode = mode;
    p->net = net;
    p->top_k = top_k;
    p->delta = static_cast<float>(delta);
    return EMPI_OK;
}

empi_status empi_plan_proposal_stats(const empi_plan* p, unsigned long long out[5]) {
    if (!p || !out) return EMPI_INVALID_ARGUMENT;
    std::copy(p->h_pstats, p->h_pstats + 5, out);
    return EMPI_OK;
}

empi_status empi_plan_set_training_capacity(empi_plan* p, size_t rows, unsigned stride) {
    if (!p || stride == 0) return EMPI_INVALID_ARGUMENT;
    p->train_stride = stride;
    empi_status st;
    if ((st = p->d_trainX.alloc(rows * kFeatDim)) != EMPI_OK) return st;
    if ((st = p->d_trainT.alloc(rows)) != EMPI_OK) return st;
    if ((st = p->d_traincount.alloc(rows ? 1 : 0)) != EMPI_OK) return st;
    if (rows) EMPI_HIP(hipMemset(p->d_traincount.p, 0, sizeof(unsigned long long)));
    // the captured iteration holds the old pointers
    if (p->prop_exec) (void)hipGraphExecDestroy(p->prop_exec);
    if (p->prop_graph) (void)hipGraphDestroy(p->prop_graph);
    p->prop_exec = nullptr;
    p->prop_graph = nullptr;
    return EMPI_OK;
}

empi_status empi_plan_training_rows(empi_plan* p, float* X, float* target, size_t capacity, size_t* rows) {
    if (!p || !rows) return EMPI_INVALID_ARGUMENT;
    *rows = 0;
    if (!p->d_traincount.p) return EMPI_OK;
    unsigned long long n = 0;
    EMPI_HIP(hipMemcpy(&n, p->d_traincount.p, sizeof n, hipMemcpyDeviceToHost));
    n = std::min<unsigned long long>(n, p->d_trainT.n);
    n = std::min<unsigned long long>(n, capacity);
    if (n && X) EMPI_HIP(hipMemcpy(X, p->d_trainX.p, sizeof(float) * n * kFeatDim, hipMemcpyDeviceToHost));
    if (n && target) EMPI_HIP(hipMemcpy(target, p->d_trainT.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    EMPI_HIP(hipMemset(p->d_traincount.p, 0, sizeof(unsigned long long)));
    *rows = static_cast<size_t>(n);
    return EMPI_OK;
}

// ----------------------------------------------------------- argmax test hook

empi_status empi_debug_argmax(int scales, const int* T, const int* G, const int* ncell, const int* bins_per_cell,
                              const int* bins, const float* products, int out[3], float* best) {
    if (scales <= 0 || !T || !G || !ncell || !bins_per_cell || !bins || !products || !out || !best)
        return EMPI_INVALID_ARGUMENT;
    std::vector<long long> poff(scales + 1, 0), boff(scales + 1, 0);
    for (int l = 0; l < scales; ++l) {
        if (T[l] <= 0 || G[l] <= 0 || bins_per_cell[l] <= 0 || ncell[l] <= 0 || ncell[l] > (1 << kTileCellBits))
            return EMPI_INVALID_ARGUMENT;
        for (int q = 0; q < G[l] * bins_per_cell[l]; ++q) {
            const int k = bins[boff[l] + q];
            if (k < 0 || k >= (1 << (32 - kTileCellBits))) return EMPI_INVALID_ARGUMENT;
        }
        poff[l + 1] = poff[l] + static_cast<long long>(T[l]) * G[l] * bins_per_cell[l];
        boff[l + 1] = boff[l] + static_cast<long long>(G[l]) * bins_per_cell[l];
        if (T[l] >= (1 << 26)) return EMPI_INVALID_ARGUMENT;
    }
    DevBuf<int> dT, dG, dN, dB, dbins, dout;
    DevBuf<long long> dpo, dbo;
    DevBuf<float> dp, dbest;
    empi_status st;
    if ((st = dT.alloc(scales)) || (st = dG.alloc(scales)) || (st = dN.alloc(scales)) || (st = dB.alloc(scales)) ||
        (st = dbins.alloc(boff[scales])) || (st = dout.alloc(3)) || (st = dpo.alloc(scales)) ||
        (st = dbo.alloc(scales)) || (st = dp.alloc(poff[scales])) || (st = dbest.alloc(1)))
        return st;
    hipStream_t s = nullptr;
    if ((st = dT.upload(T, scales, s)) || (st = dG.upload(G, scales, s)) || (st = dN.upload(ncell, scales, s)) ||
        (st = dB.upload(bins_per_cell, scales, s)) || (st = dbins.upload(bins, boff[scales], s)) ||
        (st = dpo.upload(poff.data(), scales, s)) || (st = dbo.upload(boff.data(), scales, s)) ||
        (st = dp.upload(products, poff[scales], s)))
        return st;
    hipLaunchKernelGGL(k_debug_argmax<256>, dim3(1), dim3(256), 0, s, scales, dT.p, dG.p, dN.p, dB.p, dpo.p, dbo.p,
                       dp.p, dbins.p, dout.p, dbest.p);
    EMPI_HIP(hipGetLastError());
    EMPI_HIP(hipMemcpy(out, dout.p, 3 * sizeof(int), hipMemcpyDeviceToHost));
    EMPI_HIP(hipMemcpy(best, dbest.p, sizeof(float), hipMemcpyDeviceToHost));
    return EMPI_OK;
}

// ----------------------------------------------------------- workspace API

empi_status empi_ws_create(const empi_dict_params* dict, double fs, int channels, int samples, int constraint,
                           empi_ws** out) {
    if (!out || !dict) return EMPI_INVALID_ARGUMENT;
    *out = nullptr;
    if (channels <= 0 || samples < 2) return EMPI_INVALID_ARGUMENT;
    empi_ws* w = new (std::nothrow) empi_ws();
    if (!w) return EMPI_NO_MEMORY;
    empi_item it{};
    it.channels = channels;
    it.samples = samples;
    it.fs = fs;
    // one unit over all channels; the constraint is carried by the variant
    it.variant = channels == 1 ? (constraint ? EMPI_MMP1 : EMPI_MMP3) : (constraint ? EMPI_MMP1 : EMPI_MMP3);
    it.iterationMax = 1;
    it.energyPercent = 100.0;
    it.dict = *dict;
    empi_status st;
    try {
        st = plan_build(&w->plan, &it, 1, constraint ? 1 : 0);
        if (st == EMPI_OK) st = w->d_one.alloc(1);
    } catch (const std::bad_alloc&) {
        st = EMPI_NO_MEMORY;
    } catch (...) {
        st = EMPI_INTERNAL;
    }
    if (st != EMPI_OK) {
        delete w;
        return st;
    }
    w->C = channels;
    w->N = samples;
    w->fs = fs;
    *out = w;
    return EMPI_OK;
}

empi_status empi_ws_compute(empi_ws* w, const double* signal) {
    if (!w || !signal) return EMPI_INVALID_ARGUMENT;
    empi_plan* p = &w->plan;
    hipStream_t s = nullptr;
    EMPI_HIP(hipMemcpyAsync(p->d_res.p, signal, sizeof(double) * w->C * w->N, hipMemcpyHostToDevice, s));
    // state: active, energy only for the input scaling of the FP32 maps
    double E = 0.0;
    for (int c = 0; c < w->C; ++c) {
        double sum = 0.0;
        for (int n = 0; n < w->N; ++n) sum += signal[static_cast<size_t>(c) * w->N + n] * signal[static_cast<size_t>(c) * w->N + n];
        E += sum / w->fs;
    }
    UnitState us{};
    us.active = 1;
    us.totalEnergy = us.residEnergy = E;
    const double ms = E * w->fs / (static_cast<double>(w->C) * w->N);
    int ex = 0;
    if (ms > 0) std::frexp(std::sqrt(ms), &ex);
    ex = std::max(-100, std::min(100, ex));
    us.xscale = std::ldexp(1.0f, -ex);
    EMPI_HIP(hipMemcpyAsync(p->d_state.p, &us, sizeof us, hipMemcpyHostToDevice, s));
    EMPI_HIP(hipMemsetAsync(p->d_counter.p, 0, sizeof(unsigned long long) * kCounterWords, s));
    hipLaunchKernelGGL(k_mirror, dim3(1), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p, p->d_res.p,
                       p->d_res32.p);
    empi_status st = launch_all_columns(p, s, 1);
    if (st != EMPI_OK) return st;
    EMPI_HIP(hipStreamSynchronize(s));
    w->computed = true;
    w->pending = false;
    return EMPI_OK;
}

empi_status empi_ws_find_best(empi_ws* w, empi_atom* out) {
    if (!w || !out) return EMPI_INVALID_ARGUMENT;
    if (!w->computed) return EMPI_INTERNAL;   // "internalLogicError", gabor.cpp:305-310
    empi_plan* p = &w->plan;
    hipStream_t s = nullptr;
    empi_status st;
    if (w->pending) {
        if ((st = launch_all_columns(p, s, 0)) != EMPI_OK) return st;
        w->pending = false;
    }
    // the search writes iteration 0 of the unit's atom rows
    UnitState us;
    EMPI_HIP(hipMemcpyAsync(&us, p->d_state.p, sizeof us, hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipStreamSynchronize(s));
    us.iteration = 0;
    us.active = 1;
    us.status = EMPI_OK;
    EMPI_HIP(hipMemcpyAsync(p->d_state.p, &us, sizeof us, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_select<kSelectThreads>, dim3(1), dim3(kSelectThreads), 0, s, p->d_units.p, p->d_state.p,
                       p->d_scales.p, p->d_uscales.p, p->d_tiles.p, p->d_res.p, p->d_res32.p, 1, nullptr, 0);
    EMPI_HIP(hipGetLastError());
    EMPI_HIP(hipMemcpyAsync(&us, p->d_state.p, sizeof us, hipMemcpyDeviceToHost, s));
    const int I = p->items[0].item.iterationMax;
    std::vector<empi_atom> rows(static_cast<size_t>(w->C) * I);
    EMPI_HIP(hipMemcpyAsync(rows.data(), p->d_atoms.p, sizeof(empi_atom) * rows.size(), hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipStreamSynchronize(s));
    if (us.status != EMPI_OK) {
        us.status = EMPI_OK;
        us.active = 1;
        EMPI_HIP(hipMemcpy(p->d_state.p, &us, sizeof us, hipMemcpyHostToDevice));
        return EMPI_ENERGY_MAP_IS_EMPTY;
    }
    for (int c = 0; c < w->C; ++c) out[c] = rows[static_cast<size_t>(c) * I];
    return EMPI_OK;
}

size_t empi_ws_atom_count(const empi_ws* w) {
    return w ? static_cast<size_t>(w->plan.dicts[0].entries) : 0;
}

empi_status empi_ws_subtract(empi_ws* w, const empi_atom* atom, int channel, double* residual) {
    if (!w || !atom || !residual || channel < 0 || channel >= w->C) return EMPI_INVALID_ARGUMENT;
    if (!w->computed) return EMPI_INTERNAL;
    empi_plan* p = &w->plan;
    hipStream_t s = nullptr;
    double* dres = p->d_res.p + static_cast<size_t>(channel) * w->N;
    // a pending re-evaluation for a different atom must run before this one moves tA/sA
    if (w->pending) {
        UnitState us;
        EMPI_HIP(hipMemcpy(&us, p->d_state.p, sizeof us, hipMemcpyDeviceToHost));
        const double tA = atom->params[2] / w->fs, sA = atom->params[3] / w->fs;
        if (us.tA != tA || us.sA != sA) {
            empi_status st = launch_all_columns(p, s, 0);
            if (st != EMPI_OK) return st;
            w->pending = false;
        }
    }
    EMPI_HIP(hipMemcpyAsync(dres, residual, sizeof(double) * w->N, hipMemcpyHostToDevice, s));
    EMPI_HIP(hipMemcpyAsync(w->d_one.p, atom, sizeof(empi_atom), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_subtract_one<kSelectThreads>, dim3(1), dim3(kSelectThreads), 0, s, p->d_units.p,
                       p->d_state.p, w->d_one.p, channel, p->d_res.p, p->d_res32.p);
    EMPI_HIP(hipGetLastError());
    EMPI_HIP(hipMemcpyAsync(residual, dres, sizeof(double) * w->N, hipMemcpyDeviceToHost, s));
    EMPI_HIP(hipStreamSynchronize(s));
    w->pending = true;
    return EMPI_OK;
}

void empi_ws_destroy(empi_ws* w) { delete w; }

empi_status empi_subtract_atom_from_signal(double fs, empi_atom* atom, double* x, int N, int fit) {
    if (!atom || !x || N <= 0 || !(fs > 0)) return EMPI_INVALID_ARGUMENT;
    // gabor.cpp:143-187 on the host
    const double sA = atom->params[3] / fs;
    const double fA = atom->params[4] * fs / 2.0;
    const double tA = atom->params[2] / fs;
    const double amplitude = fit ? 1.0 : atom->params[1];
    const double omega = 2.0 * M_PI * fA;
    const double phase = atom->params[5];
    const long n = N;
    const double hw = kGaussHalfWidth * sA;
    long l = std::lrint((tA - hw) * fs), r = std::lrint((tA + hw) * fs);
    const int iL = static_cast<int>(l > 0 ? l : 0);
    const int iR = static_cast<int>(r < n - 1 ? r : n - 1);
    const int Nw = 1 + iR - iL;
    if (Nw <= 0) return EMPI_OK;
    std::vector<double> w(static_cast<size_t>(Nw));
    for (int i = 0; i < Nw; ++i) {
        const double t = (i + iL) / fs - tA;
        w[i] = amplitude * std::cos(omega * t + phase);
    }
    const double gw = std::sqrt(M_PI) / sA;
    const double ga = gw * (1.0 / fs), gb = gw * (tA - iL / fs);
    for (int i = 0; i < Nw; ++i) {
        const double xw = ga * i - gb;
        w[i] *= std::exp(-xw * xw);
    }
    if (fit) {
        double norm2 = 0.0;
        for (int i = 0; i < Nw; ++i) norm2 += w[i] * w[i];
        const double norm = std::sqrt(norm2);
        for (int i = 0; i < Nw; ++i) w[i] /= norm;
        double modulus = 0.0;
        for (int i = 0; i < Nw; ++i) modulus += w[i] * x[iL + i];
        for (int i = 0; i < Nw; ++i) w[i] *= modulus;
        atom->params[0] = modulus;
        atom->params[1] = modulus / norm;
    }
    for (int i = 0; i < Nw; ++i) x[iL + i] -= w[i];
    return EMPI_OK;
}

}  // extern "C"
