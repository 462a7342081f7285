/*
 * empi_oracle.h -- FP64 CPU restatement of empi's Matching Pursuit loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline of the MI355X engine.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it, and only as the checker; the
 * product path (empi_amd/) never links or calls it.
 *
 * Every function restates one piece of the reference (paths relative to the
 * reference tree):
 *   eo_dict_build     src/gabor.cpp:213-252, inc/fftw.hpp:161-168
 *   eo_rfft           inc/fftw.hpp:120-127,151-153 (r2c, unnormalised, forward)
 *   eo_column         src/gabor.cpp:279-303 (+ gaussize, gabor.cpp:31-40)
 *   eo_product        src/gabor.cpp:44-71, src/mmp.cpp:9-21 (MMP1 constraint)
 *   eo_heap_*         inc/heap.hpp:20-94
 *   eo_subtract       src/gabor.cpp:143-187
 *   eo_decompose      src/classes.cpp:26-87, src/mmp.cpp:23-45,
 *                     src/gabor.cpp:75-133,189-209,314-354
 *
 * What pins each function (tests/test_oracle*.py): dictionary sizes against
 * the closed-form table of the config-1 dictionary and the entry totals of
 * the other workloads; the FFT against a brute-force O(n^2) DFT and numpy;
 * columns against a brute-force windowed DFT; products against the closed
 * form of analytic Gabor atoms; the heap against the reference heap.hpp
 * compiled from its own source (oracle/_ref) on tie-heavy sequences; the
 * MMP1 constraint against its invariants; the loop at every iteration
 * against a brute-force argmax over fresh products of the residual and its
 * stored maps against a full rebuild (a planted narrow re-evaluation rule
 * fails both), energy monotonicity and least-squares invariants.
 */
#ifndef EMPI_ORACLE_H
#define EMPI_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EO_OK = 0,
    EO_SIGNAL_IS_EMPTY = 1,          /* "signalIsEmpty"      classes.cpp:37-39 */
    EO_ENERGY_MAP_IS_EMPTY = 2,      /* "energyMapIsEmpty"   gabor.cpp:108-110 */
    EO_TOO_LONG = 3,                 /* "signalFileIsTooLongForThisDecomposition" gabor.cpp:239-241 */
    EO_INVALID_ARGUMENT = 4,
    EO_NO_MEMORY = 5
};

enum { EO_SMP = 0, EO_MMP1 = 1, EO_MMP2 = 2, EO_MMP3 = 3 };

typedef struct {
    double s;        /* scale [s] */
    int Ngauss;      /* 2*lrint(3 s fs - 0.5) + 1 */
    int Nfft;        /* power of two */
    int T;           /* tCount */
    int F;           /* fCount */
} eo_scale;

typedef struct {
    double fs;
    double tMax;     /* (N-1)/fs */
    double root;
    int N;
    int L;           /* number of scales */
    long long entries;   /* sum of F*T */
    eo_scale* scales;
} eo_dict;

typedef struct {
    double params[6];    /* modulus, amplitude, t [samples], s [samples], f [x Nyquist], phase */
    int k, f, t;         /* scale index, frequency bin, time index */
    int pad;
} eo_atom;

typedef struct {
    int iterationMax;
    double residualEnergy;   /* 1 - energyPercent/100 */
    int variant;             /* EO_SMP .. EO_MMP3 */
} eo_settings;

int  eo_dict_build(double fs, int N, double energyError, double scaleMin,
                   double scaleMax, double freqMax, eo_dict* out);
void eo_dict_free(eo_dict* d);

/* r2c forward DFT X[k] = sum_n x[n] e^{-2 pi i k n / n}, k = 0..n/2, n a power of 2 >= 2 */
void eo_rfft(const double* x, int n, double* X /* 2*(n/2+1) interleaved */);

/* One map column of one channel: v[2*F] interleaved (re, im). */
void eo_column(const eo_dict* d, int k, const double* x, int t, double* v);

/* Product (selection energy) at (k, f, t) from the C complex map values in v
 * (interleaved); when constraint is set and the bin is not high-frequency the
 * MMP1 constraint is applied to v in place first. */
double eo_product(const eo_dict* d, int k, int f, int t, double* v, int C, int constraint);

/* MMP1 constraint on C complex values (interleaved, in place). */
void eo_mmp1_constraint(double* v, int C);

/* Atom parameters from the (constrained) channel values at (k, f, t). */
void eo_atom_params(const eo_dict* d, int k, int f, int t, const double* v, int C, eo_atom* out);

/* Residual update; fit=1 rewrites params[0..1] (least-squares refit). */
void eo_subtract(double fs, eo_atom* atom, double* x, int N, int fit);

/* ||w||^2 of the waveforms eo_subtract removes (fit = 0) for n atoms given by
 * their parameters params [n][6] on an N-sample signal. */
void eo_waveform_norm2(double fs, int N, const double* params, int n, double* out);

/* Indexed max-heap with the reference's sift rules. */
typedef struct eo_heap eo_heap;
eo_heap* eo_heap_new(const double* values, size_t n);
void     eo_heap_update(eo_heap* h, size_t key, double value);
void     eo_heap_peek(const eo_heap* h, size_t* key, double* value);
double   eo_heap_second(const eo_heap* h);   /* max of the root's children, -inf if none */
void     eo_heap_free(eo_heap* h);

/*
 * Full decomposition of one multichannel signal (FP64, signal [C][N]).
 * Output layout (caller-allocated, I = iterationMax):
 *   atoms    [C][I]      atoms[c*I + i]
 *   counts   [C]         atoms per channel
 *   energies [R][I+1]    residual energy before iteration i+1 (row 0 = total),
 *                        then after the last subtraction; NaN-padded.
 *   gaps     [R][I]      (best - second) / best of the FP64 products at each
 *                        iteration, the near-tie measure of the selection
 * R = C for SMP (one row per channel), 1 otherwise.
 * threads <= 0 keeps the OpenMP default.
 */
int eo_decompose(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                 eo_atom* atoms, int* counts, double* energies, double* gaps, int threads);

/* Same loop, stopping after iterLimit iterations (for a bounded CPU-baseline
 * sample); returns atoms found in *natoms. */
int eo_decompose_limited(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                         eo_atom* atoms, int* counts, double* energies, double* gaps,
                         int threads, int iterLimit);

const char* eo_strerror(int code);

/* ------------------------------------------------------------ loop pinning
 * Hooks for the tests that pin the lazy re-evaluation of the loop
 * (gabor.cpp:189-209, 314-354) against a from-scratch computation. */

/* The stored map of one scale: [C][F][T] complex, interleaved. */
typedef struct {
    const double* map;
    int C, F, T;
    int pad;
} eo_map_view;

/* Called once per MP iteration that subtracts, after the subtraction of every
 * channel and before the lazy index update of the next search: iteration =
 * atoms subtracted so far (1-based), row = the SMP channel (else 0), the
 * residual [C][N] and the stored maps of all L scales. */
typedef void (*eo_trace_fn)(void* user, int iteration, int row, const double* residual, int C, int N,
                            const eo_map_view* maps, int L);

/* eo_decompose_limited with a trace callback (fn may be NULL). */
int eo_decompose_traced(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                        eo_atom* atoms, int* counts, double* energies, double* gaps,
                        int threads, int iterLimit, eo_trace_fn fn, void* user);

/* eo_decompose_traced that follows a given path through near-ties: follow
 * [R][I][3] holds (k, f, t) per row and iteration (R as for energies).
 * Where that entry differs from the oracle's own best and its product is at
 * least (1 - tol) of the best, the loop takes it instead (a near-tie branch;
 * followed [R][I] records 1 there); otherwise the loop keeps its own best. */
int eo_decompose_follow(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                        eo_atom* atoms, int* counts, double* energies, double* gaps,
                        int threads, int iterLimit, eo_trace_fn fn, void* user,
                        const int* follow, double tol, int* followed);

/* Products of every entry of scale k computed from fresh columns of x [C][N]
 * (column + product only: no stored map, heap or lazy index), out[f*T + t]. */
void eo_fresh_products(const eo_dict* d, int k, const double* x, int C, int constraint, double* out);

/* Planted mistake for the tests: on != 0 narrows subtractAtom's
 * re-evaluation rule from |t - tA| <= 3 (s + sA) to |t - tA| <= 3 s. */
void eo_plant_narrow_dirty_rule(int on);

#ifdef __cplusplus
}
#endif

#endif
