// The column kernel k_columns<Q, MODE, W>: re-evaluates the map cells the
// last atom touched -- window, Gaussian, packed real FFT in LDS (Stockham
// radix 16/8/4/2), mirror pairing, phase twiddle, product, per-cell and
// per-tile bests.  The complex map is never stored: a cell best is 8 bytes.
// Included only by the class-range translation units columns_*.hip.
// References (paths in the reference tree): gabor.cpp:75-133,143-209.
#pragma once

#include <hip/hip_runtime.h>

#include "columns.hpp"
#include "device_math.hpp"
#include "dictionary.hpp"
#include "kernels.hpp"

namespace empi {

// E-table entries staged in LDS per work item (scales with more fall back to
// HBM).  Q <= 16 keeps 64: its scales have Nfft <= 128 at energyError 0.01,
// so JT <= 33, and the 520 bytes saved keep the low-bin variant of Q = 16
// within four workgroups per CU (40,960 bytes each)
template <int Q>
constexpr int etab_lds() { return Q <= 16 ? 64 : 128; }

// LDS index with one float2 of padding per 8 points: the strided writes of
// the radix-16/8 passes then spread over all banks.  Classes Q <= 16 run one
// pass whose lanes each read 16 consecutive points: one float2 per 16 puts
// neighbouring lanes 17 points apart (and saves the LDS that lets Q = 16
// hold four workgroups per CU).
template <int Q>
constexpr int pad_shift() { return Q <= 16 ? 4 : 3; }
template <int Q>
__device__ __forceinline__ constexpr int lpad(int i) { return i + (i >> pad_shift<Q>()); }
template <int Q, int NPTS>
struct Lds {
    static constexpr int SIZE = NPTS + (NPTS >> pad_shift<Q>());
};

// The pass at sub-length Ns = 16 (classes whose first radix is 16) takes its
// twiddles w^r = e^{-2 pi i k r / (16 R)}, k < 16, r < R, from a table
// staged once per workgroup in LDS at [r][k]: a thread's k is the same in
// every tile, the 16 values of k fill the 32 banks once (rows [k][r] would
// put all sixteen on one bank), and R - 1 table reads replace one hardware
// sine/cosine and R - 2 complex products
template <int Q>
constexpr int tw16_radix() { return Q > 16 && pick_radix(Q) == 16 ? pick_radix(Q / 16) : 0; }
template <int Q>
struct Tw16 {
    static constexpr int SIZE = tw16_radix<Q>() > 0 ? 16 * tw16_radix<Q>() : 1;
};

// One Stockham autosort pass of radix R at sub-length Ns over all slots of
// Q points.  WL (wave-local): of NW = NT / 64 waves, wave w takes butterflies
// w NB/NW .. (w+1) NB/NW - 1, i.e. slots w NSL/NW .. (w+1) NSL/NW - 1 in
// every pass, so its reads and writes stay within its own slots and need no
// workgroup barrier.
template <int Q, int NT, int NPTS, int Ns, int R, bool WL>
__device__ inline void stockham_pass(float2* buf, const float2* tw16) {
    constexpr int BPS = Q / R;
    constexpr int NB = NPTS / R;
    constexpr int PER = NB / NT;
    static_assert(PER * NT == NB, "pass does not tile the workgroup");
    static_assert(!WL || (NB / (NT / 64)) % BPS == 0, "wave-local passes: whole slots per wave");
    const int tid = threadIdx.x;
    const int b0 = WL ? (tid & ~63) * PER + (tid & 63) : tid;
    constexpr int BSTEP = WL ? 64 : NT;
    float2 v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int bi = b0 + b * BSTEP;
        const int slot = bi / BPS, j = bi % BPS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = buf[lpad<Q>(slot * Q + j + r * BPS)];
    }
    if constexpr (WL) wave_sync(); else __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int bi = b0 + b * BSTEP;
        const int slot = bi / BPS, j = bi % BPS;
        int k = j & (Ns - 1);
        if constexpr (Ns == 16) {
            static_assert(R == tw16_radix<Q>(), "table radix");
            const float2* col = tw16 + k;
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], col[16 * r]);
        } else if constexpr (Ns > 1) {
            // twiddles are recomputed per pass rather than hoisted out of the
            // persistent loop: holding them for every pass costs more
            // registers than the hardware sine/cosine costs cycles
            asm volatile("" : "+v"(k));
            // w^r, w = e^{-2 pi i k / (Ns R)}, by products of the hardware w (depth <= 3)
            float2 w[R];
            w[1] = cis_rev(-static_cast<float>(k) * (1.0f / static_cast<float>(Ns * R)));
            if constexpr (R >= 4) {
                w[2] = cmul(w[1], w[1]);
                w[3] = cmul(w[2], w[1]);
            }
            if constexpr (R >= 8) {
                w[4] = cmul(w[2], w[2]);
                w[5] = cmul(w[4], w[1]);
                w[6] = cmul(w[3], w[3]);
                w[7] = cmul(w[4], w[3]);
            }
            if constexpr (R == 16) {
#pragma unroll
                for (int r = 8; r < 16; ++r) w[r] = cmul(w[r / 2], w[r - r / 2]);   // depth <= 4
            }
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r]);
        }
        dft<R>(v[b]);
        const int base = slot * Q + (j - k) * R + k;
        constexpr int SH = pad_shift<Q>();
        if constexpr (Ns % (1 << SH) == 0) {
            // 2^SH | r Ns: lpad(base + r Ns) = lpad(base) + r (Ns + Ns / 2^SH),
            // one address and constant offsets
            const int pb = lpad<Q>(base);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pb + r * (Ns + (Ns >> SH))] = v[b][r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lpad<Q>(base + r * Ns)] = v[b][r];
        }
    }
    if constexpr (WL) wave_sync(); else __syncthreads();
}

template <int Q, int NT, int NPTS, int Ns, bool WL>
__device__ inline void fft_passes(float2* buf, const float2* tw16) {
    if constexpr (Ns < Q) {
        constexpr int left = Q / Ns;
        constexpr int R = pick_radix(left);
        stockham_pass<Q, NT, NPTS, Ns, R, WL>(buf, tw16);
        fft_passes<Q, NT, NPTS, Ns * R, WL>(buf, tw16);
    }
}

// sub-transform indices of a cell: g = 0 -> {0, P/2}, g > 0 -> {g, P - g}; -1 = none
__device__ inline void cell_k2(int g, int P, int* ka, int* kb) {
    if (g == 0) {
        *ka = 0;
        *kb = P >= 2 ? P / 2 : -1;
    } else {
        *ka = g;
        *kb = P - g;
    }
}

// ------------------------------------------------------------ column kernel

struct CellInfo {
    int valid;      // re-evaluated by this pass (cells past T*G never are)
    int iL;
    int W;
    int dint;
    float dfrac;
    int ka, kb;     // sub-transforms of the cell's two slots, -1 = none
};   // 28 bytes: the Q = 16 and 32 classes fit four workgroups per CU

// bins a re-evaluated cell contributes per pass: k = P k1 + k2 (k1 < Q) below
// F for each of its sub-transforms k2, and the Nyquist bin k = M of a g = 0
// cell when F > M
__device__ inline unsigned cell_bins(const CellInfo& ci, int P, int Q, int F, int M) {
    auto below = [&](int k2) { return k2 < 0 || k2 >= F ? 0 : min(Q, (F - k2 + P - 1) / P); };
    return static_cast<unsigned>(below(ci.ka) + below(ci.kb) + (ci.ka == 0 && F > M ? 1 : 0));
}

// MODE kModeSingle: one channel, no constraint (SMP units, MMP2 sums);
// kModeSum: sum over channels of the products (MMP3); kModeMmp1: MMP1
// constraint on low-frequency bins, channel sum on high ones.
// LB (single-channel only, the proposal's evaluations): each re-evaluated
// cell also writes celllow = the maximum over its bins 1 <= j < JB of the
// product at E = 0, K' |X|^2, which bounds what a bin's product can become
// when a subtraction turns its phase (the proposal's region bound).
// Persistent: workgroups stride through the class's affected tiles.
template <int Q, int MODE, int W, bool LB>
__global__ __launch_bounds__(Geo<Q>::NT) __attribute__((amdgpu_waves_per_eu(W))) void k_columns(
    const ClassDev* __restrict__ classes, int cls, const ScaleDev* __restrict__ scales,
    const UnitDev* __restrict__ units, const UnitState* __restrict__ state,
    const float* __restrict__ residual, unsigned long long* __restrict__ cellbest,
    unsigned long long* __restrict__ tilebest,
    unsigned long long* __restrict__ counter, float* __restrict__ celllow) {
    using G_ = Geo<Q>;
    constexpr int NT = G_::NT, NPTS = G_::NPTS, NCELL = G_::NCELL, EPT = G_::EPT;
    constexpr bool MULTI = MODE != kModeSingle;
    static_assert(!(LB && MULTI), "low-bin bests are single-channel only");
    constexpr int NACC = MULTI ? EPT : 1;
    constexpr int NACC3 = MODE == kModeMmp1 ? EPT : 1;
    constexpr int SEG = 2 * Q < 64 ? 2 * Q : 64;          // lanes sharing a cell
    constexpr int GRP = 2 * Q >= NT ? 2 * Q / NT : 1;      // consecutive elements sharing a cell
    static_assert(EPT % GRP == 0, "element groups");
    static_assert(NCELL <= (1 << kTileCellBits), "cell index of a tile-best key");
    // Wave-local tiles (single channel, Q = 8..512, four waves): wave w owns
    // LDS cell positions w NCW .. (w+1) NCW - 1 and their slots from set-up
    // through load, passes and products to the cell bests, so within a tile
    // it orders its LDS accesses by itself.  Position w NCW + l holds the
    // tile's cell 4 l + w: neighbouring cells, which an atom re-evaluates
    // together, are dealt round-robin over the waves, and a wave none of
    // whose cells is re-evaluated skips its load, passes and products.  The one workgroup barrier per tile is the
    // tile-start vote; each wave leaves the maximum over its cells in tbw,
    // and thread 0 writes the tile best after the next tile's vote, from the
    // half of tbw that tile did not use.
    constexpr bool WL = !MULTI && Q >= 8 && Q <= 512;
    constexpr int NCW = WL ? NCELL / 4 : NCELL;
    static_assert(!WL || (NT == 256 && NCW >= 1 && NCW <= 64), "wave-local tiles");
    // Wave-local passes (Q = 8..4096, all modes): the Stockham passes run on
    // slots of QP points that each wave holds whole (a wave holds 64 EPT =
    // 1024 points).  Smaller Q pass over the Q-point slots themselves.  Q =
    // 2048 and 4096 take the first step of a decimation in frequency at load
    // time, in registers: with S = Q / QP, the S inputs j + QP r (r < S) of a
    // slot, which one thread loads, give
    // y_s[j] = w_Q^{js} sum_r x[j + QP r] e^{-2 pi i rs/S}, and the slot's
    // bin S k' + s is bin k' of the QP-point transform of y_s.  A workgroup
    // barrier then remains only after the load (where the tile is not
    // wave-local) and before the products.  Q = 8192 (multichannel) keeps
    // workgroup passes: the same step onto 2048-point slots spilled more
    // registers there and ran 10% slower.
    constexpr bool WLP = Q >= 8 && Q <= 4096;
    constexpr int QP = WLP && Q > 64 * EPT ? 64 * EPT : Q;
    constexpr int S = Q / QP;
    static_assert(!WLP || (NPTS / (NT / 64)) % QP == 0, "wave-local passes");
    static_assert(S == 1 || (Q / NT) % S == 0, "decimation step: a thread's pairs hold whole radix-S groups");
    __shared__ float2 buf[Lds<Q, NPTS>::SIZE];
    __shared__ CellInfo info[NCELL];
    __shared__ unsigned long long best[NCELL];
    __shared__ unsigned lowbest[LB ? NCELL : 1];   // FP32 bits of a non-negative maximum
    __shared__ unsigned long long tbest;
    __shared__ unsigned cnt_lds;
    constexpr int kEtabLds = etab_lds<Q>();
    __shared__ float2 etl[kEtabLds + 1];    // ((1 - E)/K', 2E/K') of the item's scale; [kEtabLds] = (1/K', 0)
    // windows longer than 2Q samples (Q = 4096 only): e^{-2 pi i k2 / P} of each cell's two slots
    __shared__ float2 ftw[Q == kMaxSubTransform ? 2 * NCELL : 1];
    __shared__ float2 tw16[Tw16<QP>::SIZE];
    __shared__ unsigned long long tbw[WL ? 8 : 1];   // per-wave tile maxima, two halves
    __shared__ unsigned long long evw[WL ? 4 : 1];   // per-wave entry counts at the end

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    if constexpr (tw16_radix<QP>() > 0) {
        constexpr int R16 = tw16_radix<QP>();
        for (int i = tid; i < 16 * R16; i += NT) {
            double sn, cs;
            sincospi(-2.0 * static_cast<double>((i / 16) * (i % 16)) / (16.0 * R16), &sn, &cs);
            tw16[i] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
        }
    }
    const ClassDev cd = classes[cls];
    const int nitems = min(*cd.count, cd.max_items);
    const int total = cd.work[nitems];
    unsigned long long evaluated = 0, evaluated_ch = 0;

    // XCD-aware walk: workgroups are dealt round-robin over the 8 XCDs, so
    // workgroup b runs on XCD b % 8; each XCD takes one contiguous eighth of
    // the work list, keeping neighbouring cells (overlapping windows) in one
    // L2, and each of its workgroups a contiguous run of that eighth.
    // Consecutive tiles then mostly belong to one work item: its item and
    // job records stay in registers, so a tile's unit state, scale and unit
    // records are loaded together instead of after a search and two
    // dependent loads, and the E table is read once per scale.
    constexpr int kXcds = 8;
    const int xcd = blockIdx.x % kXcds;
    const int per_xcd = gridDim.x / kXcds;
    const int chunk = (total + kXcds - 1) / kXcds;
    const int x_lo = min(total, xcd * chunk), x_hi = min(total, (xcd + 1) * chunk);
    const int run_len = (x_hi - x_lo + per_xcd - 1) / per_xcd;
    const int w_beg = min(x_hi, x_lo + static_cast<int>(blockIdx.x / kXcds) * run_len);
    const int w_end = min(x_hi, w_beg + run_len);
    int lo = 0;             // work[lo] <= w for every w this workgroup visits from here on
    int item_end = -1;      // first tile past the cached item
    int item_w0 = 0, etl_scale = -1;
    WorkItem item{};
    Job job{};
    int pend = -1, pend_half = 0, tb_half = 0;   // WL: tile-best slot awaiting its write, its half of tbw
    auto wl_tile_best = [&]() {
        const unsigned long long* tw = tbw + 4 * pend_half;
        return max(max(tw[0], tw[1]), max(tw[2], tw[3]));
    };
    for (int w = w_beg; w < w_end; ++w) {
        if (w >= item_end) {
            // item: last i with work[i] <= w.  w only grows, so gallop
            // forward from the last item before the binary search
            int step = 1;
            while (lo + step < nitems && cd.work[lo + step] <= w) {
                lo += step;
                step <<= 1;
            }
            int hi = min(lo + step, nitems) - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (cd.work[mid] <= w) lo = mid; else hi = mid - 1;
            }
            item = cd.items[lo];
            job = cd.jobs[item.job];
            item_w0 = cd.work[lo];
            item_end = cd.work[lo + 1];   // work has nitems + 1 entries, work[nitems] = total
        }
        const UnitState st = state[job.unit];
        const ScaleDev sd = scales[job.scale];
        const UnitDev ud = units[job.unit];
        const int tile = item.tile + (w - item_w0);
        const int cell0 = tile * NCELL;
        const int ncells = sd.T * sd.G;
        const double s = sd.s, fs = sd.fs;
        const int P = sd.P, G = sd.G, M = sd.M, F = sd.F, Nfft = sd.Nfft;

        // bins this thread's cells contribute to the evaluation count.
        // WL: lane l < NCW of wave w sets up position w NCW + l, cell 4 l + w
        unsigned mycnt = 0;
        int any = 0;
        const int c_own = WL ? (tid >> 6) * NCW + lane : tid;
        const int c_end = WL ? (lane < NCW ? c_own + 1 : 0) : NCELL;
        const int lc_own = WL ? lane * 4 + (tid >> 6) : tid;   // the tile's cell at c_own
        for (int c = c_own; c < c_end; c += NT) {
            CellInfo ci{};
            const int cell = cell0 + (WL ? lc_own : c);
            const bool exists = cell < ncells;
            ci.valid = exists;
            ci.kb = -1;
            if (exists) {
                const int t = cell / G;
                const int g = cell - t * G;
                const double t0 = t * sd.tMax / (sd.T - 1);
                ci.valid = t >= item.tlo && t <= item.thi;
                if (item.atom) ci.valid = ci.valid && fabs(t0 - st.tA) <= (s + st.sA) * kGaussHalfWidth;
                const double hw = kGaussHalfWidth * s;
                const long l = lrint((t0 - hw) * fs);
                const long r = lrint((t0 + hw) * fs);
                ci.iL = static_cast<int>(l > 0 ? l : 0);
                const int iR = static_cast<int>(r < sd.N - 1 ? r : sd.N - 1);
                ci.W = min(iR - ci.iL + 1, 2 * Q * sd.fold);
                const double d = (t0 - ci.iL / fs) * fs;
                const double dfl = floor(d);
                ci.dint = static_cast<int>(dfl);
                ci.dfrac = static_cast<float>(d - dfl);
                cell_k2(g, P, &ci.ka, &ci.kb);
            }
            if constexpr (Q == kMaxSubTransform) {
                ftw[2 * c] = cis_rev(-static_cast<float>(ci.ka) / static_cast<float>(P));
                ftw[2 * c + 1] = ci.kb >= 0 ? cis_rev(-static_cast<float>(ci.kb) / static_cast<float>(P))
                                            : make_float2(0.0f, 0.0f);
            }
            info[c] = ci;
            best[c] = 0ull;
            if constexpr (LB) lowbest[c] = 0u;
            any |= ci.valid;
            if (ci.valid) mycnt += cell_bins(ci, P, Q, F, M);
        }
        if constexpr (WL) {
            evaluated += mycnt;
        } else if (tid == 0) {
            cnt_lds = 0;
            tbest = 0ull;
        }
        if (job.scale != etl_scale) {
            // uniform.  The previous tile's readers have passed its last
            // barrier, except with wave-local tiles, whose waves may still be
            // in their products
            if constexpr (WL) __syncthreads();
            etl_scale = job.scale;
            for (int i = tid; i <= kEtabLds; i += NT) etl[i] = i < sd.JT ? sd.etab[i] : make_float2(sd.invKp, 0.0f);
        }
        const int any_all = __syncthreads_or(any);
        if constexpr (WL) {
            if (pend >= 0 && tid == 0) tilebest[pend] = wl_tile_best();
            pend = -1;
        }
        if (!any_all) continue;

        const float invN = ldexpf(1.0f, -sd.lgNfft);   // Nfft is a power of two
        const float invM = 2.0f * invN;
        const float Kq = 0.25f * sd.Kp;
        constexpr bool mmp1 = MODE == kModeMmp1;
        // WL: a wave none of whose cells is re-evaluated has no channel to load
        const int C = MULTI ? ud.C : (!WL || __ballot(any) != 0ull ? 1 : 0);

        float acc0[NACC], acc1[NACC3], acc2[NACC3];
        float accNyq = 0.0f;
#pragma unroll
        for (int i = 0; i < NACC; ++i) acc0[i] = 0.0f;
#pragma unroll
        for (int i = 0; i < NACC3; ++i) acc1[i] = acc2[i] = 0.0f;

        for (int ch = 0; ch < C; ++ch) {
            // load: packed pairs z[m] = w[2m] + i w[2m+1], pre-twiddled per slot.
            // The channel is read as a buffer resource: a sample at or past a
            // cell's window length (every sample of a cell not re-evaluated)
            // is read at an offset past the channel, which returns zero, so
            // the loads need no branch and the unrolled iterations overlap.
            const __amdgpu_buffer_rsrc_t xr =
                buffer_of(residual + ud.res_off + static_cast<long long>(ch) * ud.N, ud.N * 4);
            const unsigned oob = static_cast<unsigned>(ud.N) * 4u;
            const float sg = sd.sgA;
            // window samples 2mm and 2mm + 1 of a cell: x exp2(-(sgA u)^2),
            // u = n - dint - dfrac, with sgA (n - dint) - sgA dfrac in one fma
            auto window_pair = [&](const CellInfo& ci, const int lim, const float cg, const int mm) {
                const int n0 = 2 * mm;
                const unsigned o = static_cast<unsigned>(ci.iL + n0) * 4u;
                const float x0 = buffer_load(xr, n0 < lim ? o : oob);
                const float x1 = buffer_load(xr, n0 + 1 < lim ? o + 4u : oob);
                const float s0 = fmaf(sg, static_cast<float>(n0 - ci.dint), cg);
                const float s1 = s0 + sg;
                return make_float2(x0 * __builtin_amdgcn_exp2f(-(s0 * s0)), x1 * __builtin_amdgcn_exp2f(-(s1 * s1)));
            };
            constexpr int GL = Q >= NT ? Q / NT : 1;   // consecutive pairs sharing a cell
            // One cell per time position (G = 1, P <= 2): every cell has k2 =
            // 0 and P/2, and for GL = 1 a thread's pair index m = tid mod Q
            // is the same in all its cells, so the slot-b pre-twiddle is
            // computed once per tile
            // WL: wave w loads pairs 512 w .. 512 w + 511 (its own cells), 64
            // consecutive ones per step (Q = 512: the cell's first half, each
            // step's second pair NT further); a thread's m is then the same in
            // all its cells only for Q <= 64
            const bool g1 = GL == 1 && __builtin_amdgcn_readfirstlane(G) == 1;
            const int e0 = WL ? (tid & ~63) * (EPT / 2) + lane : tid;
            constexpr int ESTEP = WL ? 64 : GL * NT;
            constexpr bool TW1_ONCE = !WL || Q <= 64;
            float2 tw1 = make_float2(1.f, 0.f);
            if (g1 && TW1_ONCE)
                tw1 = cis_rev(-static_cast<float>(__umul24(static_cast<unsigned>(P / 2), static_cast<unsigned>(tid % Q))) *
                              invM);
#pragma unroll
            for (int gi = 0; gi < EPT / 2 / GL; ++gi) {
                const int c = (e0 + gi * ESTEP) / Q;
                const CellInfo ci = info[c];
                const int lim = ci.valid ? ci.W : 0;
                const float cg = -sg * ci.dfrac;
                const int m0 = (e0 + gi * ESTEP) % Q;
                // pre-twiddles w_M^{k2 m} from exact revolution counts (k2 m < M <
                // 2^24) at the group's first pair; the GL pairs of a group are NT
                // apart in one cell, so later ones advance by w_M^{k2 NT}
                const unsigned um0 = static_cast<unsigned>(m0);
                const unsigned uka = static_cast<unsigned>(ci.ka), ukb = static_cast<unsigned>(max(ci.kb, 0));
                float2 twa = make_float2(1.f, 0.f), twb = tw1;
                if (!g1) {
                    twa = cis_rev(-static_cast<float>(__umul24(uka, um0)) * invM);
                    twb = cis_rev(-static_cast<float>(__umul24(ukb, um0)) * invM);
                } else if constexpr (!TW1_ONCE) {
                    twb = cis_rev(-static_cast<float>(__umul24(static_cast<unsigned>(P / 2), um0)) * invM);
                }
                float2 sta = make_float2(1.f, 0.f), stb = make_float2(1.f, 0.f);
                if constexpr (GL > 1) {
                    sta = cis_rev(-static_cast<float>(__umul24(uka, static_cast<unsigned>(NT))) * invM);
                    stb = cis_rev(-static_cast<float>(__umul24(ukb, static_cast<unsigned>(NT))) * invM);
                }
                // slot a and b inputs of pair m, pre-twiddled by ta and tb
                auto load_pair = [&](const int m, const float2 ta_m, const float2 tb_m, float2& za, float2& zb) {
                    za = zb = make_float2(0.f, 0.f);
                    // pairs past the window (and every later section) are zero
                    if (2 * m < lim) {
                        float2 z = window_pair(ci, lim, cg, m);
                        float2 zin_b = z;
                        if constexpr (Q == kMaxSubTransform) {
                            // input m + f Q enters sub-transform k2 with e^{-2 pi i k2 f / P}
                            if (sd.fold > 1) {
                                // e^{-2 pi i k2 f / P} by products of e^{-2 pi i k2 / P},
                                // restarted every 16 sections from the exact revolution
                                // count (k2 f mod P) / P, so the error stays that of a
                                // few products however many sections a window spans
                                const float2 ta = ftw[2 * c], tb = ftw[2 * c + 1];
                                const float invP = 1.0f / static_cast<float>(P);   // P is a power of two
                                float2 wa = ta, wb = tb, sa = z, sb = z;
                                for (int f = 1; f < sd.fold; ++f) {
                                    if ((f & 15) == 0) {
                                        const unsigned uf = static_cast<unsigned>(f), mask = static_cast<unsigned>(P - 1);
                                        wa = cis_rev(-static_cast<float>((uka * uf) & mask) * invP);
                                        wb = ci.kb >= 0 ? cis_rev(-static_cast<float>((ukb * uf) & mask) * invP)
                                                        : make_float2(0.f, 0.f);
                                    } else if (f > 1) {
                                        wa = cmul(wa, ta);
                                        wb = cmul(wb, tb);
                                    }
                                    const float2 zf = window_pair(ci, lim, cg, m + f * Q);
                                    sa = cadd(sa, cmul(zf, wa));
                                    sb = cadd(sb, cmul(zf, wb));
                                }
                                z = sa;
                                zin_b = sb;
                            }
                        }
                        za = ci.ka == 0 ? z : cmul(z, ta_m);
                        if (ci.kb >= 0) zb = cmul(zin_b, tb_m);
                    }
                };
                if constexpr (S == 1) {
#pragma unroll 2
                    for (int q = 0; q < GL; ++q) {
                        const int m = m0 + q * NT;   // GL > 1: one cell, no wrap
                        float2 za, zb;
                        load_pair(m, twa, twb, za, zb);
                        const int pa = lpad<Q>((2 * c) * Q + m);
                        buf[pa] = za;
                        if constexpr (Q % (1 << pad_shift<Q>()) == 0)
                            buf[pa + Q + (Q >> pad_shift<Q>())] = zb;   // lpad(i + Q) = lpad(i) + lpad(Q)
                        else
                            buf[lpad<Q>((2 * c + 1) * Q + m)] = zb;
                        if constexpr (GL > 1) {
                            if (ci.ka != 0) twa = cmul(twa, sta);
                            twb = cmul(twb, stb);
                        }
                    }
                } else {
                    // the thread's pairs j + QP r, r < S, for its GL / S values
                    // of j (NT apart); pre-twiddles advance by w_M^{QP k2} along r
                    constexpr int GS = GL / S;
                    const float2 sra = cis_rev(-static_cast<float>(__umul24(uka, static_cast<unsigned>(QP))) * invM);
                    const float2 srb = cis_rev(-static_cast<float>(__umul24(ukb, static_cast<unsigned>(QP))) * invM);
#pragma unroll 1
                    for (int qq = 0; qq < GS; ++qq) {
                        const int j = m0 + qq * NT;
                        float2 va[S], vb[S];
                        float2 ta = twa, tb = twb;
#pragma unroll
                        for (int r = 0; r < S; ++r) {
                            load_pair(j + r * QP, ta, tb, va[r], vb[r]);
                            if (ci.ka != 0) ta = cmul(ta, sra);
                            tb = cmul(tb, srb);
                        }
                        dft<S>(va);
                        dft<S>(vb);
                        // w_Q^{js}: j / Q is an exact revolution count
                        const float2 d1 = cis_rev(-static_cast<float>(j) * (1.0f / static_cast<float>(Q)));
                        float2 ds = d1;
#pragma unroll
                        for (int r = 1; r < S; ++r) {
                            va[r] = cmul(va[r], ds);
                            vb[r] = cmul(vb[r], ds);
                            if (r + 1 < S) ds = cmul(ds, d1);
                        }
                        // y_s of slot a at sub-transform 2c S + s, of slot b at (2c + 1) S + s
                        const int pa = lpad<Q>((2 * c) * Q + j);
#pragma unroll
                        for (int r = 0; r < S; ++r) {
                            buf[pa + r * (QP + (QP >> pad_shift<Q>()))] = va[r];
                            buf[pa + (S + r) * (QP + (QP >> pad_shift<Q>()))] = vb[r];
                        }
                        if (ci.ka != 0) twa = cmul(twa, sta);
                        twb = cmul(twb, stb);
                    }
                }
            }
            if constexpr (WL) wave_sync(); else __syncthreads();
            fft_passes<QP, NT, NPTS, 1, WLP>(buf, tw16);
            if constexpr (WLP && !WL) __syncthreads();   // the products read both slots of a cell

            // LDS position of bin k1 of a slot's transform (sub-transform k1 mod S, bin k1 / S of it)
            auto xat = [](const int sl, const int k1) {
                if constexpr (S == 1) return lpad<Q>(sl * Q + k1);
                else return lpad<Q>((sl * S + (k1 & (S - 1))) * QP + (k1 >> ilog2c(S)));
            };

            // post: split the packed spectrum, phase-correct, product per bin.
            // element(rem, c, ci, i, run): bin `rem` (of 2Q) of cell c; i indexes
            // the multichannel accumulators.
            // product of bin k from its spectrum value X = Xr + i Xi
            // E-table entry of bin distance j < JT: from LDS, past kEtabLds
            // (scales whose table does not fit there) from global memory
            // through a buffer load, which the compiler cannot fold with the
            // LDS read into one generic-address load
            auto etab_at = [&](const int j) {
                float2 t = etl[min(j, kEtabLds)];
                if (sd.JT > kEtabLds && j >= kEtabLds) {
                    const __amdgpu_buffer_rsrc_t er = buffer_of(reinterpret_cast<const float*>(sd.etab), sd.JT * 8);
                    t = make_float2(buffer_load(er, static_cast<unsigned>(j) * 8u),
                                    buffer_load(er, static_cast<unsigned>(j) * 8u + 4u));
                }
                return t;
            };
            auto from_x = [&](const int k, const float Xr, const float Xi, const CellInfo& ci, const int i,
                              unsigned long long& run, float& low) {
                const bool high = 4 * k > Nfft;
                const int j = high ? M - k : k;
                const float num = Xr * Xr + Xi * Xi;
                // Beyond JT the window term E is zero and P = K' |X|^2: the
                // phase of u = X e^{2 pi i cyc} enters only the first bins
                // (and MMP1's constraint on the low half).
                const bool edge = j < sd.JT;
                float a = 0.0f, b = 0.0f;
                if (edge || (mmp1 && !high)) {
                    // j, dint < 2^24: the 24-bit product is exact modulo 2^32, and Nfft divides 2^32
                    const unsigned ci_int = __umul24(static_cast<unsigned>(j), static_cast<unsigned>(ci.dint)) &
                                            static_cast<unsigned>(Nfft - 1);
                    float cyc = static_cast<float>(ci_int) * invN + static_cast<float>(j) * ci.dfrac * invN;
                    if (high) cyc = -cyc - 0.5f * static_cast<float>(ci.iL & 1);
                    float sn, cs;
                    twiddle_rev(cyc, &sn, &cs);
                    a = Xr * cs - Xi * sn;
                    b = Xr * sn + Xi * cs;
                }
                float pe = num * sd.Kp;   // the product where E = 0
                if (edge) {
                    const float2 tt = etab_at(j);
                    const float den = fmaf(tt.x, num, tt.y * a * a);
                    pe = den > 0.0f ? num * num * __builtin_amdgcn_rcpf(den) : 0.0f;
                    if constexpr (LB)
                        if (j >= 1 && j < sd.JB) low = fmaxf(low, num * sd.Kp);
                }
                if constexpr (!MULTI) {
                    run = max(run, pack_best(pe, static_cast<unsigned>(k)));
                } else if (mmp1 && !high) {
                    if constexpr (mmp1) {
                        acc0[i] += a * a;
                        acc1[i] += b * b;
                        acc2[i] += a * b;
                    }
                } else {
                    acc0[i] += pe;
                }
            };
            auto element = [&](const int rem, const int c, const CellInfo& ci, const int i,
                               unsigned long long& run, float& low) {
                const int half = rem / Q, k1 = rem % Q;
                const int k2 = half ? ci.kb : ci.ka;
                const int k = static_cast<int>(__umul24(static_cast<unsigned>(P), static_cast<unsigned>(k1))) + k2;
                if (!(ci.valid && k2 >= 0 && k < F)) return;
                const int slot = 2 * c + half;
                const float2 Zk = buf[xat(slot, k1)];
                float2 Zp;
                if (k2 == 0) Zp = buf[xat(slot, (Q - k1) & (Q - 1))];
                else Zp = buf[xat((2 * k2 == P) ? slot : (slot ^ 1), Q - 1 - k1)];
                const float Ar = Zk.x + Zp.x, Ai = Zk.y - Zp.y;
                const float Br = Zk.x - Zp.x, Bi = Zk.y + Zp.y;
                const float2 wv = cis_rev(-static_cast<float>(k) * invN);
                const float pr = wv.x * Br - wv.y * Bi, pi = wv.x * Bi + wv.y * Br;
                const float Xr = 0.5f * (Ar + pi), Xi = 0.5f * (Ai - pr);
                from_x(k, Xr, Xi, ci, i, run, low);
            };
            // Two bins from one pair of LDS values (single-channel units):
            // bin k and its mirror M - k share A = Z[k] + conj Z[M-k],
            // B = Z[k] - conj Z[M-k] and w^k: X[k] = (A - i w^k B) / 2,
            // X[M-k] = conj(A + i w^k B) / 2.  Pair u < Q of cell c is (slot a, u)
            // with (slot b, Q-1-u) for g > 0; for g = 0, (a, u) with (a, Q-u) for
            // u < Q/2 and (b, u') with (b, Q-1-u'), u' = u - Q/2; u = 0 also takes
            // the self-mirrored (a, Q/2).  The Nyquist bin keeps its own path.
            // wv carries w_N^k of the pair: computed on the hardware sine and
            // cosine where `direct`, otherwise advanced by the caller.
            auto pair = [&](const int u, const int c, const CellInfo& ci, float2& wv, const bool direct,
                            unsigned long long& run, float& low) {
                if (!ci.valid) return;
                int sa, ka1, sb, kb1, k2;
                if (ci.ka != 0) {
                    sa = 2 * c; ka1 = u; sb = 2 * c + 1; kb1 = Q - 1 - u; k2 = ci.ka;
                } else if (u < Q / 2) {
                    sa = sb = 2 * c; ka1 = u; kb1 = (Q - u) & (Q - 1); k2 = 0;
                } else {
                    if (ci.kb < 0) return;
                    sa = sb = 2 * c + 1; ka1 = u - Q / 2; kb1 = Q - 1 - ka1; k2 = ci.kb;
                }
                const int k = static_cast<int>(__umul24(static_cast<unsigned>(P), static_cast<unsigned>(ka1))) + k2;
                const int km = M - k;
                const float2 Zk = buf[xat(sa, ka1)];
                const float2 Zp = buf[xat(sb, kb1)];
                const float Ar = Zk.x + Zp.x, Ai = Zk.y - Zp.y;
                const float Br = Zk.x - Zp.x, Bi = Zk.y + Zp.y;
                if (direct) wv = cis_rev(-static_cast<float>(k) * invN);
                const float pr = wv.x * Br - wv.y * Bi, pi = wv.x * Bi + wv.y * Br;
                // 2 X[k] and 2 X[M - k]: the factor 2 is exact, and every
                // product below is formed with the rounding of the unscaled one
                const float xr = Ar + pi, xi = Ai - pr, yr = Ar - pi, yi = -(Ai + pr);
                const float nx = fmaf(xr, xr, xi * xi), ny = fmaf(yr, yr, yi * yi);   // 4 |X|^2
                // k and M - k are both j = min(k, M - k) bins from the nearer end
                // of the spectrum
                const int j = min(k, km);
                if (j >= sd.JT) {
                    // E = 0 for both: P = K' |X|^2 = K'/4 |2X|^2
                    if (F > M) {
                        run = max(run, pack_nonneg(nx * Kq, static_cast<unsigned>(k)));
                        run = max(run, pack_nonneg(ny * Kq, static_cast<unsigned>(km)));
                    } else {
                        if (k < F) run = max(run, pack_nonneg(nx * Kq, static_cast<unsigned>(k)));
                        if (km < F) run = max(run, pack_nonneg(ny * Kq, static_cast<unsigned>(km)));
                    }
                } else {
                    // E > 0: P = |X|^4 / ((1 - E)/K' |X|^2 + 2E/K' a^2), a = Re(X e^{2 pi i cyc}).
                    // The low bin of the pair has cyc = j d / Nfft, the high one
                    // -cyc - (iL odd) / 2, so its a is +-Re(X conj e^{2 pi i cyc}):
                    // one sine and cosine and one E-table entry serve both bins,
                    // and the sign drops out of a^2.
                    // j, dint < 2^24: the 24-bit product is exact modulo 2^32, and Nfft divides 2^32
                    const unsigned ci_int = __umul24(static_cast<unsigned>(j), static_cast<unsigned>(ci.dint)) &
                                            static_cast<unsigned>(Nfft - 1);
                    const float cyc = static_cast<float>(ci_int) * invN + static_cast<float>(j) * ci.dfrac * invN;
                    float sn, cs;
                    twiddle_rev(cyc, &sn, &cs);
                    const float2 tt = etab_at(j);
                    const float sg = k < km ? -sn : sn;   // k is the low bin when k < M - k
                    const float ak = fmaf(xi, sg, xr * cs), am = fmaf(-yi, sg, yr * cs);
                    const float dk = fmaf(tt.x, nx, tt.y * ak * ak), dm = fmaf(tt.x, ny, tt.y * am * am);
                    const float pk = dk > 0.0f ? 0.25f * (nx * nx * __builtin_amdgcn_rcpf(dk)) : 0.0f;
                    const float pm = dm > 0.0f ? 0.25f * (ny * ny * __builtin_amdgcn_rcpf(dm)) : 0.0f;
                    const bool inK = k < F, inM = km < F && km != M;   // km = M: the DC bin's mirror
                    if (inK) run = max(run, pack_best(pk, static_cast<unsigned>(k)));
                    if (inM) run = max(run, pack_best(pm, static_cast<unsigned>(km)));
                    if constexpr (LB)
                        if (j >= 1 && j < sd.JB) {
                            if (inK) low = fmaxf(low, nx * Kq);
                            if (inM) low = fmaxf(low, ny * Kq);
                        }
                }
                if (ci.ka == 0 && u == 0 && M / 2 < F) {
                    // (a, Q/2): bin M/2, its own mirror (Z[k] = Z[M-k])
                    const float2 Zh = buf[xat(2 * c, Q / 2)];
                    const float2 wh = cis_rev(-static_cast<float>(M / 2) * invN);
                    const float hBi = Zh.y + Zh.y;                 // B = (0, 2 Im Zh), A = (2 Re Zh, 0)
                    const float hpr = -wh.y * hBi, hpi = wh.x * hBi;
                    from_x(M / 2, 0.5f * (Zh.x + Zh.x + hpi), -0.5f * hpr, ci, 0, run, low);
                }
            };
            constexpr bool CONTIG = !MULTI && Q >= 8 && Q <= 512;
            if constexpr (CONTIG) {
                // Small sub-transforms: a thread's EPT / 2 mirror pairs are
                // consecutive pairs of one cell, visited in a rotated order that
                // spreads the LDS banks, so each cell needs one cross-lane
                // reduction per tile.
                constexpr int TPC = 2 * Q / EPT;   // threads per cell
                const int c = tid / TPC;
                const CellInfo ci = info[c];
                const int u0 = (tid % TPC) * (EPT / 2);
                unsigned long long run = 0ull;
                float low = 0.0f;
                float2 wv;
#pragma unroll 2
                for (int s2 = 0; s2 < EPT / 2; ++s2) pair(u0 + ((s2 + tid) & (EPT / 2 - 1)), c, ci, wv, true, run, low);
#pragma unroll
                for (int off = TPC / 2; off >= 1; off >>= 1) run = max(run, shfl_xor_u64(run, off));
                if ((tid & (TPC - 1)) == 0 && run) atomicMax(&best[c], run);
                if constexpr (LB) {
#pragma unroll
                    for (int off = TPC / 2; off >= 1; off >>= 1) low = fmaxf(low, __shfl_xor(low, off));
                    if ((tid & (TPC - 1)) == 0 && low > 0.0f) atomicMax(&lowbest[c], __float_as_uint(low));
                }
            } else {
                // elements of one group share a cell: its info is read once
#pragma unroll MULTI ? EPT / GRP : 1
                for (int gi = 0; gi < EPT / GRP; ++gi) {
                    const int c = (tid + gi * GRP * NT) / (2 * Q);
                    const CellInfo ci = info[c];
                    unsigned long long run = 0ull;
                    float low = 0.0f;
                    if constexpr (MULTI) {
#pragma unroll
                        for (int q = 0; q < GRP; ++q) {
                            const int i = gi * GRP + q;
                            element((tid + i * NT) % (2 * Q), c, ci, i, run, low);
                            __builtin_amdgcn_sched_barrier(0);   // keep per-element live ranges short
                        }
                    } else {
                        if constexpr (GRP >= 2) {
                            // pair u = tid + q NT has k = P u + k2: consecutive pairs
                            // are Nfft / GRP bins apart, so w_N^k advances by
                            // e^{-2 pi i / GRP}; g = 0 cells change slot (and k2)
                            // at u = Q/2, i.e. q = GRP/4, where w_N^k restarts
                            const float2 wstep = cis_rev(-1.0f / static_cast<float>(GRP));
                            float2 wv = make_float2(1.f, 0.f);
#pragma unroll 2
                            for (int q = 0; q < GRP / 2; ++q) {
                                pair(tid + q * NT, c, ci, wv, q == 0 || q == GRP / 4, run, low);
                                wv = cmul(wv, wstep);
                            }
                        } else {
                            element((tid + gi * NT) % (2 * Q), c, ci, gi, run, low);
                        }
#pragma unroll
                        for (int off = SEG / 2; off >= 1; off >>= 1) run = max(run, shfl_xor_u64(run, off));
                        if ((lane & (SEG - 1)) == 0 && run) atomicMax(&best[c], run);
                        if constexpr (LB) {
#pragma unroll
                            for (int off = SEG / 2; off >= 1; off >>= 1) low = fmaxf(low, __shfl_xor(low, off));
                            if ((lane & (SEG - 1)) == 0 && low > 0.0f) atomicMax(&lowbest[c], __float_as_uint(low));
                        }
                    }
                }
            }
            // Nyquist bin k = M of g = 0 cells: X[M] = Re Z0 - Im Z0, u real, E = 1
            if (c_own < c_end && F > M) {
                const CellInfo& ci = info[c_own];
                if (ci.valid && ci.ka == 0) {
                    const float2 Z0 = buf[xat(2 * c_own, 0)];
                    const float a = Z0.x - Z0.y;
                    accNyq += a * a;
                }
            }
            // the next channel reloads buf; a single channel's last reads of
            // buf are ordered before the next tile by the barrier below
            if constexpr (MULTI) __syncthreads();
        }

        // final products of multi-channel units and per-cell maximum
        if constexpr (MULTI) {
#pragma unroll
            for (int gi = 0; gi < EPT / GRP; ++gi) {
                const int c = (tid + gi * GRP * NT) / (2 * Q);
                const CellInfo ci = info[c];
                unsigned long long run = 0ull;
#pragma unroll
                for (int q = 0; q < GRP; ++q) {
                    const int i = gi * GRP + q;
                    const int e = tid + i * NT;
                    const int rem = e % (2 * Q);
                    const int half = rem / Q, k1 = rem % Q;
                    const int k2 = half ? ci.kb : ci.ka;
                    const int k = static_cast<int>(__umul24(static_cast<unsigned>(P), static_cast<unsigned>(k1))) + k2;
                    if (!(ci.valid && k2 >= 0 && k < F)) continue;
                    const bool high = 4 * k > Nfft;
                    float pv;
                    if (mmp1 && !high) {
                        float Srr = acc0[i], Sii = 0.0f, Sri = 0.0f;
                        if constexpr (mmp1) {
                            Sii = acc1[i];
                            Sri = acc2[i];
                        }
                        const float Dr = Srr - Sii, Di = 2.0f * Sri;
                        const float Dm = hypotf(Dr, Di);
                        const float lam = 0.5f * (Srr + Sii + Dm);
                        float cos2 = 1.0f;
                        if (Dm > 0.0f) cos2 = Dr >= 0.0f ? (Dm + Dr) / (2.0f * Dm) : (Di * Di) / (2.0f * Dm * (Dm - Dr));
                        float2 tt = etl[min(k, kEtabLds)];
                        if (sd.JT > kEtabLds && k >= kEtabLds && k < sd.JT) tt = sd.etab[k];
                        const float T1 = tt.x, T2 = tt.y;
                        const float den = fmaf(T2, cos2, T1);
                        pv = den > 0.0f ? lam / den : 0.0f;
                    } else {
                        pv = acc0[i];
                    }
                    run = max(run, pack_best(pv, static_cast<unsigned>(k)));
                }
#pragma unroll
                for (int off = SEG / 2; off >= 1; off >>= 1) run = max(run, shfl_xor_u64(run, off));
                if ((lane & (SEG - 1)) == 0 && run) atomicMax(&best[c], run);
            }
        }
        if (c_own < c_end && F > M) {
            const CellInfo& ci = info[c_own];
            if (ci.valid && ci.ka == 0) {
                // E = 1 at Nyquist: |u|^4 / (2 E a^2 / K') = K' a^2 / 2 summed over channels
                atomicMax(&best[c_own], pack_best(0.5f * sd.Kp * accNyq, static_cast<unsigned>(M)));
            }
        }
        if constexpr (!WL) {
            if (mycnt) atomicAdd(&cnt_lds, mycnt);
            __syncthreads();
        }
        // cell bests out, tile best over re-evaluated and untouched cells
        auto cell_out = [&](const int c, const int lc) {   // LDS position c holds the tile's cell lc
            const CellInfo& ci = info[c];
            if (cell0 + lc >= ncells) return 0ull;
            unsigned long long cb;
            if (ci.valid) {
                cb = best[c];
                cellbest[job.cell_off + cell0 + lc] = cb;
                if constexpr (LB) celllow[job.cell_off + cell0 + lc] = __uint_as_float(lowbest[c]);
            } else {
                cb = cellbest[job.cell_off + cell0 + lc];
            }
            return tile_key(cb, lc, kTileCellBits);
        };
        if constexpr (WL) {
            wave_sync();
            unsigned long long tb = c_own < c_end ? cell_out(c_own, lc_own) : 0ull;
#pragma unroll
            for (int off = NCW / 2; off >= 1; off >>= 1) tb = max(tb, shfl_xor_u64(tb, off));
            if (lane == 0) tbw[4 * tb_half + (tid >> 6)] = tb;
            pend = job.tile_off + tile;
            pend_half = tb_half;
            tb_half ^= 1;
        } else if constexpr (NCELL <= 64) {
            // the first wave holds every cell: it reduces the tile best
            // itself, and it alone touches the cell records, the bests and
            // the counters again before the next tile's first barrier (the
            // other waves only restage the E table there, which this does
            // not read), so no barrier is needed here
            if (tid < 64) {
                unsigned long long tb = tid < NCELL ? cell_out(tid, tid) : 0ull;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) tb = max(tb, shfl_xor_u64(tb, off));
                if (tid == 0) {
                    tilebest[job.tile_off + tile] = tb;
                    evaluated += cnt_lds;
                    evaluated_ch += static_cast<unsigned long long>(cnt_lds) * (MULTI ? ud.C : 1);
                }
            }
        } else {
            unsigned long long tb = 0ull;
            for (int c = tid; c < NCELL; c += NT) tb = max(tb, cell_out(c, c));
            if (tb) atomicMax(&tbest, tb);
            __syncthreads();
            if (tid == 0) {
                tilebest[job.tile_off + tile] = tbest;
                evaluated += cnt_lds;
                evaluated_ch += static_cast<unsigned long long>(cnt_lds) * (MULTI ? ud.C : 1);
            }
            __syncthreads();
        }
    }
    // one counter update per workgroup, into its XCD's copy
    unsigned long long* const cnt = counter ? counter + 2 * kNumClasses * xcd + 2 * (ilog2c(Q) - 1) : nullptr;
    if constexpr (WL) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) evaluated += shfl_xor_u64(evaluated, off);
        if (lane == 0) evw[tid >> 6] = evaluated;
        __syncthreads();   // every wave has left its part of the last tile's best and its count
        if (tid == 0) {
            if (pend >= 0) tilebest[pend] = wl_tile_best();
            const unsigned long long n = evw[0] + evw[1] + evw[2] + evw[3];
            if (cnt && n) {
                atomicAdd(&cnt[0], n);
                atomicAdd(&cnt[1], n);
            }
        }
    } else if (tid == 0 && cnt && evaluated) {
        atomicAdd(&cnt[0], evaluated);
        atomicAdd(&cnt[1], evaluated_ch);
    }
}

template <int Q, int MODE, int W, bool LB>
void launch_columns(int grid, hipStream_t s, const ClassDev* classes, int cls, const ScaleDev* scales,
                    const UnitDev* units, const UnitState* state, const float* residual,
                    unsigned long long* cellbest, unsigned long long* tilebest, unsigned long long* counter,
                    float* celllow) {
    hipLaunchKernelGGL((k_columns<Q, MODE, W, LB>), dim3(grid), dim3(Geo<Q>::NT), 0, s, classes, cls, scales, units,
                       state, residual, cellbest, tilebest, counter, celllow);
}

// the launcher of class Q in a multichannel mode; variant v allows
// column_waves + v waves per SIMD
template <int Q, int MODE>
ColumnLauncher multichannel_variant(int variant) {
    constexpr int W = column_waves<Q, MODE>();
    if (variant == 1) return launch_columns<Q, MODE, W + 1, false>;
    if (variant == 2) return launch_columns<Q, MODE, W + 2, false>;
    return launch_columns<Q, MODE, W, false>;
}
template <int Q>
ColumnLauncher multichannel_launcher_for(int mode, int variant) {
    if (mode == kModeSum) return multichannel_variant<Q, kModeSum>(variant);
    if (mode == kModeMmp1) return multichannel_variant<Q, kModeMmp1>(variant);
    return nullptr;
}

// the launcher of class Q in one mode / wave variant; low: the single-channel
// kernel that also writes the cells' low-bin bests
template <int Q>
ColumnLauncher launcher_for(int mode, int variant, bool low) {
    if (mode != kModeSingle) return low ? nullptr : multichannel_launcher_for<Q>(mode, variant);
    if (low) {
        if (variant == 1) return launch_columns<Q, kModeSingle, variant_waves<Q, 1>(), true>;
        if (variant == 2) return launch_columns<Q, kModeSingle, variant_waves<Q, 2>(), true>;
        return launch_columns<Q, kModeSingle, variant_waves<Q, 0>(), true>;
    }
    if (variant == 1) return launch_columns<Q, kModeSingle, variant_waves<Q, 1>(), false>;
    if (variant == 2) return launch_columns<Q, kModeSingle, variant_waves<Q, 2>(), false>;
    return launch_columns<Q, kModeSingle, variant_waves<Q, 0>(), false>;
}

}  // namespace empi
