// Optimal Gabor dictionary, host side (FP64).
//
// Restates GaborWorkspaceBuilder::prepareWorkspace (src/gabor.cpp:213-252)
// expression by expression, so scales, FFT sizes and grid sizes are
// bit-identical to the reference's: lrint is round-half-even under the
// default rounding mode, exactly as the reference calls it.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "empi_hip.h"

namespace empi {

constexpr double kGaussHalfWidth = 3.0;        // src/gabor.cpp:18
constexpr double kMinScaleInSamples = 2.0;     // src/gabor.cpp:23
// Sub-transforms of the column kernels (complex points): up to 4096 for
// single-channel units, 8192 for multichannel ones (their channel loop keeps
// one workgroup per CU either way).  Longer windows fold onto 4096 points,
// one section per 8192 samples.
constexpr int kMaxSubTransform = 4096;
constexpr int kMaxSubTransformMulti = 8192;
// Sizes the column kernels compute exactly: bin indices, window offsets and
// the revolution counts of their twiddles are integers below 2^24 (24-bit
// products, exact FP32 conversions, bins in 24 bits of a tile-best key), so
// FFTs up to 2^24 points and signals up to 2^24 samples.
constexpr int kMaxNfft = 1 << 24;
constexpr int kMaxSamples = 1 << 24;

struct Scale {
    double s;        // scale [s]
    int Ngauss;
    int Nfft;        // power of two
    int T;           // time grid size
    int F;           // frequency grid size
    // launch geometry of the fused column kernels
    int Wmax;        // longest window (samples) over the time grid
    int Q;           // sub-transform length (complex points)
    int fold;        // Q-point sections of the longest window's packed input
    int P;           // Nfft/2 / Q sub-transforms per column
    int G;           // cells per column = max(1, P/2)
};

struct Dictionary {
    double fs = 0, tMax = 0, root = 0;
    int N = 0;
    int max_q = kMaxSubTransform;   // longest unfolded sub-transform
    empi_dict_params params{};
    std::vector<Scale> scales;
    long long entries = 0;       // sum F*T
    long long cells = 0;         // sum T*G
};

// Window [iL, iR] of column t of scale s (src/gabor.cpp:284-286).
inline void column_window(double fs, int N, double tMax, int T, double s, int t, int* iL, int* iR,
                          double* t0fixed) {
    const long n = N;
    const double t0 = t * tMax / (T - 1);
    const double hw = kGaussHalfWidth * s;
    long l = std::lrint((t0 - hw) * fs);
    long r = std::lrint((t0 + hw) * fs);
    *iL = static_cast<int>(l > 0 ? l : 0);
    *iR = static_cast<int>(r < n - 1 ? r : n - 1);
    *t0fixed = t0 - *iL / fs;
}

// Builds the dictionary; returns EMPI_OK, EMPI_INVALID_ARGUMENT or EMPI_TOO_LONG.
// max_q: the longest sub-transform used unfolded (kMaxSubTransform or
// kMaxSubTransformMulti); longer inputs fold onto kMaxSubTransform points.
empi_status build_dictionary(const empi_dict_params& p, double fs, int N, Dictionary* out,
                             int max_q = kMaxSubTransform);

// Factor c of the proposal's energy bound P <= c * (residual energy under the
// column's window) for every cell of a scale.  With g the chosen real atom
// (continuous unit norm), P = (sum_n x_n g_n)^2 / fs <= (window energy) *
// sum_n g_n^2 / fs by Cauchy-Schwarz.  By Poisson summation the discrete norm
// exceeds the continuous one, (s / 2 sqrt 2)(1 + cos(2 phi4) E(f4)), by at most
// the aliases of the squared envelope at multiples of fs (2.2 e^{-pi (s fs)^2 / 2}
// each side) and of its cos(2 omega t) part at fs - 2 f4 >= fs / 2
// (e^{-pi (s fs)^2 / 8}); the continuous norm is at least 1 - E at bin 1 (bins 0
// and Nyquist are real-valued, norm 2).  A 1e-4 margin covers FP32 rounding.
double proposal_bound_factor(double s, double fs, int Nfft);

}  // namespace empi
