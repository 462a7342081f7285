// Optimal Gabor dictionary (src/gabor.cpp:213-252), plus the launch geometry
// the fused column kernels need for each scale.
#include "dictionary.hpp"

#include <climits>

namespace empi {

static long pow2_at_least(long x) {   // inc/fftw.hpp:161-168 (minimum 2)
    long y = 2;
    while (y < x) y *= 2;
    return y;
}

empi_status build_dictionary(const empi_dict_params& p, double fs, int N, Dictionary* out, int max_q) {
    *out = Dictionary{};
    out->max_q = max_q;
    if (!(fs > 0) || N < 2 || !(p.energyError > 0 && p.energyError < 1)) return EMPI_INVALID_ARGUMENT;
    const double energyError = p.energyError;
    const double scaleMin = std::fmax(p.scaleMin, kMinScaleInSamples / fs);
    const double scaleMax = std::fmin(p.scaleMax, N / fs);
    const double root = std::sqrt(-2.0 / M_PI * std::log(1.0 - energyError));
    const double aDenomSqrt = 1.0 - energyError;
    const double aNominPart = energyError * (2.0 - energyError) * (energyError * energyError - 2 * energyError + 2);
    const double a = (1.0 + std::sqrt(aNominPart)) / (aDenomSqrt * aDenomSqrt);
    const double dl = std::log(a);
    const double tMax = (N - 1) / fs;
    const double lMin = std::log(scaleMin);
    const double lMax = std::log(scaleMax);
    long lCount = std::lrint((lMax - lMin) / dl + 0.5) + 1;
    if (lCount < 0) lCount = 0;

    out->fs = fs;
    out->tMax = tMax;
    out->root = root;
    out->N = N;
    out->params = p;
    for (long i = 0; i < lCount; ++i) {
        // One scale divides 0 by 0 in the reference (a NaN scale); it is scaleMin here.
        const double s = lCount > 1 ? std::exp(lMin + static_cast<double>(i) * (lMax - lMin) / static_cast<double>(lCount - 1))
                                    : scaleMin;
        const double dt = root * s;
        const double df = root / s;
        const double hwGabor = kGaussHalfWidth * s;
        const long Ngauss = 2 * std::lrint(hwGabor * fs - 0.5) + 1;
        const long alt = std::lrint(fs / df + 0.5);
        const long Nfft = pow2_at_least(Ngauss > alt ? Ngauss : alt);
        const long tCount = std::lrint(tMax / dt + 0.5) + 1;
        if ((Nfft > tCount ? Nfft : tCount) > static_cast<long>(INT_MAX)) {
            *out = Dictionary{};
            return EMPI_TOO_LONG;
        }
        long fCount = Nfft / 2 + 1;
        if (std::isfinite(p.freqMax) && p.freqMax > 0) {
            const long fc = std::lrint(p.freqMax / fs * Nfft + 0.5) + 1;
            if (fc < fCount) fCount = fc;
        }
        Scale sc{};
        sc.s = s;
        sc.Ngauss = static_cast<int>(Ngauss);
        sc.Nfft = static_cast<int>(Nfft);
        sc.T = static_cast<int>(tCount);
        sc.F = static_cast<int>(fCount);

        // Longest window over the grid; the packed real transform needs
        // ceil(W/2) complex inputs, rounded up to a power of two (>= 2) up to
        // max_q.  Longer inputs fold onto kMaxSubTransform points: the input
        // m = m1 + Q f enters sub-transform k2 as z[m] e^{-2 pi i k2 f / P},
        // summed over f.
        int wmax = 0;
        for (int t = 0; t < sc.T; ++t) {
            int iL, iR;
            double t0f;
            column_window(fs, N, tMax, sc.T, s, t, &iL, &iR, &t0f);
            if (iR - iL + 1 > wmax) wmax = iR - iL + 1;
        }
        const int M = sc.Nfft / 2;
        if (wmax > sc.Nfft) wmax = sc.Nfft;
        sc.Wmax = wmax;
        int q = 1;
        while (q < (wmax + 1) / 2) q *= 2;
        if (q < 2) q = 2;
        if (q > M) q = M;
        if (q > max_q) q = kMaxSubTransform;
        sc.Q = q;
        sc.fold = ((wmax + 1) / 2 + q - 1) / q;
        if (sc.fold < 1) sc.fold = 1;
        sc.P = M / q;
        sc.G = sc.P >= 2 ? sc.P / 2 : 1;
        out->scales.push_back(sc);
        out->entries += static_cast<long long>(sc.F) * sc.T;
        out->cells += static_cast<long long>(sc.T) * sc.G;
    }
    return EMPI_OK;
}

double proposal_bound_factor(double s, double fs, int Nfft) {
    const double sfs = s * fs;
    const double a1 = sfs / Nfft;
    const double den = -std::expm1(-2.0 * M_PI * a1 * a1);
    const double alias = std::exp(-M_PI * sfs * sfs / 8.0) + 4.4 * std::exp(-M_PI * sfs * sfs / 2.0);
    return (1.0 + alias / std::max(den, 1e-30)) * (1.0 + 1e-4);
}

}  // namespace empi
