/*
 * empi_oracle.c -- FP64 CPU restatement of empi's Matching Pursuit loop.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle + CPU baseline); the product
 * path never links it.  Citations are reference-tree paths.
 *
 * Parallelism mirrors the reference: OpenMP `for schedule(dynamic,1)` over
 * scales in the full build, in the lazy per-scale argmax and in the
 * post-subtraction column re-evaluation (gabor.cpp:77-82, 89-107, 198-208).
 * Cross-scale exact ties resolve to the lowest scale index (the reference's
 * single-thread order; with several threads the reference itself is
 * nondeterministic there).
 *
 * The FFT is this file's own FP64 radix-2 transform (FFTW3, which the
 * reference uses, is absent from this image).
 */
#include "empi_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define EO_SQRT2 1.41421356237309504880

/* gabor.cpp:18,23 */
static const double GAUSS_HALF_WIDTH = 3.0;
static const double MIN_SCALE_IN_SAMPLES = 2.0;

const char* eo_strerror(int code) {
    switch (code) {
        case EO_OK: return "ok";
        case EO_SIGNAL_IS_EMPTY: return "signalIsEmpty";
        case EO_ENERGY_MAP_IS_EMPTY: return "energyMapIsEmpty";
        case EO_TOO_LONG: return "signalFileIsTooLongForThisDecomposition";
        case EO_INVALID_ARGUMENT: return "invalidArgument";
        case EO_NO_MEMORY: return "insufficientMemory";
        default: return "internalError";
    }
}

/* ---------------------------------------------------------------- dictionary */

/* smallest power of two >= x, minimum 2 (fftw.hpp:161-168) */
static long pow2_at_least(long x) {
    long y = 2;
    while (y < x) y *= 2;
    return y;
}

/* gabor.cpp:213-252 */
int eo_dict_build(double fs, int N, double energyError, double scaleMinCfg,
                  double scaleMaxCfg, double freqMax, eo_dict* out) {
    memset(out, 0, sizeof *out);
    if (!(fs > 0) || N < 2 || !(energyError > 0 && energyError < 1)) return EO_INVALID_ARGUMENT;
    double scaleMin = fmax(scaleMinCfg, MIN_SCALE_IN_SAMPLES / fs);
    double scaleMax = fmin(scaleMaxCfg, N / fs);
    double root = sqrt(-2.0 / M_PI * log(1.0 - energyError));
    double aDenomSqrt = 1.0 - energyError;
    double aNominPart = energyError * (2.0 - energyError) * (energyError * energyError - 2 * energyError + 2);
    double a = (1.0 + sqrt(aNominPart)) / (aDenomSqrt * aDenomSqrt);
    double dl = log(a);
    double tMax = (N - 1) / fs;
    double lMin = log(scaleMin);
    double lMax = log(scaleMax);
    long lCount = lrint((lMax - lMin) / dl + 0.5) + 1;
    if (lCount < 0) lCount = 0;

    out->fs = fs;
    out->tMax = tMax;
    out->root = root;
    out->N = N;
    out->L = (int)lCount;
    out->scales = (eo_scale*)calloc(lCount > 0 ? (size_t)lCount : 1, sizeof(eo_scale));
    if (!out->scales) return EO_NO_MEMORY;
    long long entries = 0;
    for (long i = 0; i < lCount; ++i) {
        /* lCount == 1 divides 0 by 0 in the reference (a NaN scale); here the
         * single scale is scaleMin. */
        double s = lCount > 1 ? exp(lMin + (double)i * (lMax - lMin) / (double)(lCount - 1)) : scaleMin;
        double dt = root * s;
        double df = root / s;
        double hwGabor = GAUSS_HALF_WIDTH * s;
        long Ngauss = 2 * lrint(hwGabor * fs - 0.5) + 1;
        long alt = lrint(fs / df + 0.5);
        long Nfft = pow2_at_least(Ngauss > alt ? Ngauss : alt);
        long tCount = lrint(tMax / dt + 0.5) + 1;
        if ((Nfft > tCount ? Nfft : tCount) > (long)INT_MAX) {
            eo_dict_free(out);
            return EO_TOO_LONG;
        }
        long fCount = Nfft / 2 + 1;
        if (isfinite(freqMax) && freqMax > 0) {
            long fc = lrint(freqMax / fs * Nfft + 0.5) + 1;
            if (fc < fCount) fCount = fc;
        }
        out->scales[i].s = s;
        out->scales[i].Ngauss = (int)Ngauss;
        out->scales[i].Nfft = (int)Nfft;
        out->scales[i].T = (int)tCount;
        out->scales[i].F = (int)fCount;
        entries += (long long)fCount * tCount;
    }
    out->entries = entries;
    return EO_OK;
}

void eo_dict_free(eo_dict* d) {
    free(d->scales);
    d->scales = NULL;
    d->L = 0;
}

/* time grid t_i = i tMax/(T-1) (gabor.cpp:262-264) and frequency grid f_j = j fs/Nfft (l.265-267) */
static double t_of(const eo_dict* d, int k, int ti) {
    return ti * d->tMax / (d->scales[k].T - 1);
}
static double f_of(const eo_dict* d, int k, int fi) {
    return fi * d->fs / d->scales[k].Nfft;
}

/* ---------------------------------------------------------------------- FFT */

typedef struct {
    int n;          /* real length (power of 2) */
    int m;          /* complex length n/2 */
    double* tw;     /* e^{-2 pi i j / m}, j < m/2, interleaved */
    double* twr;    /* e^{-2 pi i k / n}, k <= m, interleaved */
    int* rev;       /* bit reversal of m */
} eo_fft_plan;

static int ilog2i(int x) { int l = 0; while ((1 << l) < x) ++l; return l; }

static int plan_init(eo_fft_plan* p, int n) {
    memset(p, 0, sizeof *p);
    p->n = n;
    p->m = n / 2;
    int m = p->m;
    p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)(m / 2 + 1));
    p->twr = (double*)malloc(sizeof(double) * 2 * (size_t)(m + 1));
    p->rev = (int*)malloc(sizeof(int) * (size_t)m);
    if (!p->tw || !p->twr || !p->rev) return EO_NO_MEMORY;
    for (int j = 0; j <= m / 2; ++j) {
        double ang = -2.0 * M_PI * (double)j / (double)m;
        p->tw[2 * j] = cos(ang);
        p->tw[2 * j + 1] = sin(ang);
    }
    for (int k = 0; k <= m; ++k) {
        double ang = -2.0 * M_PI * (double)k / (double)n;
        p->twr[2 * k] = cos(ang);
        p->twr[2 * k + 1] = sin(ang);
    }
    int lg = ilog2i(m);
    for (int j = 0; j < m; ++j) {
        int r = 0;
        for (int b = 0; b < lg; ++b) if (j & (1 << b)) r |= 1 << (lg - 1 - b);
        p->rev[j] = r;
    }
    return EO_OK;
}

static void plan_free(eo_fft_plan* p) {
    free(p->tw); free(p->twr); free(p->rev);
    memset(p, 0, sizeof *p);
}

/* in-place forward complex FFT of length p->m (radix-2, decimation in time) */
static void cfft(const eo_fft_plan* p, double* z) {
    int m = p->m;
    for (int j = 0; j < m; ++j) {
        int r = p->rev[j];
        if (r > j) {
            double a = z[2 * j], b = z[2 * j + 1];
            z[2 * j] = z[2 * r]; z[2 * j + 1] = z[2 * r + 1];
            z[2 * r] = a; z[2 * r + 1] = b;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        int half = len >> 1;
        int step = m / len;
        for (int i = 0; i < m; i += len) {
            for (int j = 0; j < half; ++j) {
                double wr = p->tw[2 * j * step], wi = p->tw[2 * j * step + 1];
                double* u = z + 2 * (i + j);
                double* v = z + 2 * (i + j + half);
                double xr = v[0] * wr - v[1] * wi;
                double xi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - xr; v[1] = u[1] - xi;
                u[0] += xr; u[1] += xi;
            }
        }
    }
}

/* r2c through a half-length complex FFT of z[j] = x[2j] + i x[2j+1]:
 * X[k] = (Z[k] + conj Z[m-k])/2 - i W^k (Z[k] - conj Z[m-k])/2, W = e^{-2 pi i/n}. */
static void rfft_plan(const eo_fft_plan* p, const double* x, double* X, double* z) {
    int n = p->n, m = p->m;
    if (n == 2) {
        X[0] = x[0] + x[1]; X[1] = 0;
        X[2] = x[0] - x[1]; X[3] = 0;
        return;
    }
    for (int j = 0; j < m; ++j) { z[2 * j] = x[2 * j]; z[2 * j + 1] = x[2 * j + 1]; }
    cfft(p, z);
    for (int k = 0; k <= m; ++k) {
        int a = k % m, b = (m - k) % m;
        double zr = z[2 * a], zi = z[2 * a + 1];
        double cr = z[2 * b], ci = -z[2 * b + 1];
        double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
        double dr = 0.5 * (zr - cr), di = 0.5 * (zi - ci);
        double wr = p->twr[2 * k], wi = p->twr[2 * k + 1];
        /* -i * W * d */
        double tr = wr * dr - wi * di, ti = wr * di + wi * dr;
        X[2 * k] = er + ti;
        X[2 * k + 1] = ei - tr;
    }
    X[1] = 0.0;
    X[2 * m + 1] = 0.0;
}

void eo_rfft(const double* x, int n, double* X) {
    eo_fft_plan p;
    if (plan_init(&p, n) != EO_OK) { plan_free(&p); return; }
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)(n / 2 + 1));
    if (z) rfft_plan(&p, x, X, z);
    free(z);
    plan_free(&p);
}

/* ------------------------------------------------------------------ columns */

/* x[i] *= exp(-(a i - b)^2), a = sqrt(pi) step / width, b = sqrt(pi) center / width (gabor.cpp:31-40) */
static void gaussize(double* buf, int n, double step, double center, double width) {
    const double w = sqrt(M_PI) / width;
    const double a = w * step;
    const double b = w * center;
    for (int i = 0; i < n; ++i) {
        const double xw = a * i - b;
        buf[i] *= exp(-xw * xw);
    }
}

/* window bounds of column t (gabor.cpp:284-286) */
static void window_of(const eo_dict* d, int k, int t, int* iL, int* iR, double* t0fixed) {
    const long N = d->N;
    const double t0 = t_of(d, k, t);
    const double hw = GAUSS_HALF_WIDTH * d->scales[k].s;
    long l = lrint((t0 - hw) * d->fs);
    long r = lrint((t0 + hw) * d->fs);
    *iL = (int)(l > 0 ? l : 0);
    *iR = (int)(r < N - 1 ? r : N - 1);
    *t0fixed = t0 - *iL / d->fs;
}

/* column with a caller-provided plan and scratch (buf: n reals, X: 2(m+1), z: 2(m+1)) */
static void column_plan(const eo_dict* d, int k, const eo_fft_plan* p, const double* x, int t,
                        double* v, double* buf, double* X, double* z) {
    const eo_scale* sc = &d->scales[k];
    int iL, iR;
    double t0fixed;
    window_of(d, k, t, &iL, &iR, &t0fixed);
    memset(buf, 0, sizeof(double) * (size_t)sc->Nfft);
    int W = iR - iL + 1;
    if (W > sc->Nfft) W = sc->Nfft;   /* the reference writes past its buffer here; never hit by valid dictionaries */
    for (int i = 0; i < W; ++i) buf[i] = x[iL + i];
    gaussize(buf, W, 1.0 / d->fs, t0fixed, sc->s);
    rfft_plan(p, buf, X, z);
    const double norm = 1.0 / d->fs;
    const double mult = 2 * M_PI * t0fixed;
    for (int fi = 0; fi < sc->F; ++fi) {
        double ang = f_of(d, k, fi) * mult;
        double c = cos(ang), s = sin(ang);
        double xr = norm * X[2 * fi], xi = norm * X[2 * fi + 1];
        v[2 * fi] = xr * c - xi * s;
        v[2 * fi + 1] = xr * s + xi * c;
    }
}

void eo_column(const eo_dict* d, int k, const double* x, int t, double* v) {
    const eo_scale* sc = &d->scales[k];
    eo_fft_plan p;
    if (plan_init(&p, sc->Nfft) != EO_OK) { plan_free(&p); return; }
    double* buf = (double*)malloc(sizeof(double) * (size_t)sc->Nfft);
    double* X = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
    if (buf && X && z) column_plan(d, k, &p, x, t, v, buf, X, z);
    free(buf); free(X); free(z);
    plan_free(&p);
}

/* ----------------------------------------------------------------- products */

/* mmp.cpp:9-21: D = sum v^2, theta = arg(D)/2, v <- polar(|v| cos(arg v - theta), theta) */
void eo_mmp1_constraint(double* v, int C) {
    double dr = 0.0, di = 0.0;
    for (int c = 0; c < C; ++c) {
        double re = v[2 * c], im = v[2 * c + 1];
        dr += re * re - im * im;
        di += re * im + im * re;
    }
    double angle = 0.5 * atan2(di, dr);
    double ca = cos(angle), sa = sin(angle);
    for (int c = 0; c < C; ++c) {
        double re = v[2 * c], im = v[2 * c + 1];
        double r = hypot(re, im) * cos(atan2(im, re) - angle);
        v[2 * c] = r * ca;
        v[2 * c + 1] = r * sa;
    }
}

typedef struct {
    double normComplex, freqNyquist, f, expFactor;
    int isHighFreq;
} computer;

/* GaborComputer ctor (gabor.cpp:44-52) */
static computer computer_of(const eo_dict* d, int k, int fi) {
    computer g;
    double s = d->scales[k].s;
    g.normComplex = sqrt(EO_SQRT2 / s);
    g.freqNyquist = 0.5 * d->fs;
    g.f = f_of(d, k, fi);
    g.isHighFreq = g.f > 0.5 * g.freqNyquist;
    double f4 = g.isHighFreq ? (g.freqNyquist - g.f) : g.f;
    g.expFactor = exp(-2 * M_PI * s * s * f4 * f4);
    return g;
}

/* GaborComputer::compute after the buffer gather (gabor.cpp:59-70) */
static double computer_square(const computer* g, const eo_dict* d, double t, double* v, int C, int constraint) {
    if (constraint && !g->isHighFreq) eo_mmp1_constraint(v, C);
    double square = 0.0;
    double sfs = sqrt(d->fs);
    for (int c = 0; c < C; ++c) {
        double phase = atan2(v[2 * c + 1], v[2 * c]);
        double phase4 = g->isHighFreq ? (2 * M_PI * g->freqNyquist * t - phase) : phase;
        double normReal = EO_SQRT2 * g->normComplex / sqrt(1.0 + cos(2 * phase4) * g->expFactor);
        double moduli = hypot(v[2 * c], v[2 * c + 1]) * normReal * sfs;
        square += moduli * moduli;
    }
    return square;
}

double eo_product(const eo_dict* d, int k, int fi, int ti, double* v, int C, int constraint) {
    computer g = computer_of(d, k, fi);
    return computer_square(&g, d, t_of(d, k, ti), v, C, constraint);
}

/* gabor.cpp:111-131 */
void eo_atom_params(const eo_dict* d, int k, int fi, int ti, const double* v, int C, eo_atom* out) {
    computer g = computer_of(d, k, fi);
    double t = t_of(d, k, ti);
    for (int c = 0; c < C; ++c) {
        double re = v[2 * c], im = v[2 * c + 1];
        double phase = atan2(im, re);
        double phase4 = g.isHighFreq ? (2 * M_PI * g.freqNyquist * t - phase) : phase;
        double normReal = EO_SQRT2 * g.normComplex / sqrt(1.0 + cos(2 * phase4) * g.expFactor);
        double mag = hypot(re, im);
        eo_atom* a = &out[c];
        a->params[0] = mag * normReal * sqrt(d->fs);
        a->params[1] = mag * normReal * normReal;
        a->params[2] = t * d->fs;
        a->params[3] = d->scales[k].s * d->fs;
        a->params[4] = g.f * 2.0 / d->fs;
        a->params[5] = phase;
        a->k = k; a->f = fi; a->t = ti; a->pad = 0;
    }
}

/* --------------------------------------------------------------------- heap */

struct eo_heap {
    size_t size;
    double* heap;
    size_t* heap_from_key;
    size_t* key_from_heap;
};

static void heap_swap(eo_heap* h, size_t i, size_t n) {
    double tv = h->heap[i]; h->heap[i] = h->heap[n]; h->heap[n] = tv;
    size_t ki = h->key_from_heap[i], kn = h->key_from_heap[n];
    h->key_from_heap[i] = kn; h->key_from_heap[n] = ki;
    size_t t = h->heap_from_key[kn]; h->heap_from_key[kn] = h->heap_from_key[ki]; h->heap_from_key[ki] = t;
}

/* heap.hpp:41-70: strict comparisons; the right child wins only when it is
 * strictly greater than the left one. */
static void heap_sift(eo_heap* h, size_t i, double value, int allow_parent) {
    const size_t n = h->size;
    if (allow_parent && i > 0 && value > h->heap[(i - 1) / 2]) {
        do {
            size_t p = (i - 1) / 2;
            heap_swap(h, i, p);
            i = p;
        } while (i > 0 && value > h->heap[(i - 1) / 2]);
        return;
    }
    for (;;) {
        size_t a = 2 * i + 1, b = a + 1;
        int takeB = b < n && value < h->heap[b] && h->heap[b] > h->heap[a];
        int takeA = a < n && value < h->heap[a];
        if (takeB) { heap_swap(h, i, b); i = b; }
        else if (takeA) { heap_swap(h, i, a); i = a; }
        else break;
    }
}

eo_heap* eo_heap_new(const double* values, size_t n) {
    eo_heap* h = (eo_heap*)calloc(1, sizeof *h);
    if (!h) return NULL;
    h->size = n;
    h->heap = (double*)malloc(sizeof(double) * (n ? n : 1));
    h->heap_from_key = (size_t*)malloc(sizeof(size_t) * (n ? n : 1));
    h->key_from_heap = (size_t*)malloc(sizeof(size_t) * (n ? n : 1));
    if (!h->heap || !h->heap_from_key || !h->key_from_heap) { eo_heap_free(h); return NULL; }
    for (size_t i = 0; i < n; ++i) {
        h->heap[i] = values[i];
        h->heap_from_key[i] = h->key_from_heap[i] = i;
    }
    /* Floyd build from n/2-1 down to 0 (heap.hpp:31-39); n <= 1 needs no sifting */
    if (n >= 2) {
        size_t i = n / 2 - 1;
        for (;;) {
            heap_sift(h, i, h->heap[i], 0);
            if (i == 0) break;
            --i;
        }
    }
    return h;
}

void eo_heap_update(eo_heap* h, size_t key, double value) {
    size_t i = h->heap_from_key[key];
    h->heap[i] = value;
    heap_sift(h, i, value, 1);
}

void eo_heap_peek(const eo_heap* h, size_t* key, double* value) {
    *key = h->key_from_heap[0];
    *value = h->heap[0];
}

double eo_heap_second(const eo_heap* h) {
    double s = -INFINITY;
    if (h->size > 1) s = h->heap[1];
    if (h->size > 2 && h->heap[2] > s) s = h->heap[2];
    return s;
}

void eo_heap_free(eo_heap* h) {
    if (!h) return;
    free(h->heap); free(h->heap_from_key); free(h->key_from_heap);
    free(h);
}

/* -------------------------------------------------------------- subtraction */

/* gabor.cpp:143-187 */
void eo_subtract(double fs, eo_atom* atom, double* x, int Nint, int fit) {
    const double sA = atom->params[3] / fs;
    const double fA = atom->params[4] * fs / 2.0;
    const double tA = atom->params[2] / fs;
    const double amplitude = fit ? 1.0 : atom->params[1];
    const double omega = 2.0 * M_PI * fA;
    const double phase = atom->params[5];
    const long N = Nint;
    const double hw = GAUSS_HALF_WIDTH * sA;
    long l = lrint((tA - hw) * fs), r = lrint((tA + hw) * fs);
    int iL = (int)(l > 0 ? l : 0);
    int iR = (int)(r < N - 1 ? r : N - 1);
    const int Nw = 1 + iR - iL;
    if (Nw <= 0) return;
    double* w = (double*)malloc(sizeof(double) * (size_t)Nw);
    if (!w) return;
    for (int i = 0; i < Nw; ++i) {
        double t = (i + iL) / fs - tA;
        w[i] = amplitude * cos(omega * t + phase);
    }
    gaussize(w, Nw, 1.0 / fs, tA - iL / fs, sA);
    if (fit) {
        double norm2 = 0.0;
        for (int i = 0; i < Nw; ++i) norm2 += w[i] * w[i];
        double norm = sqrt(norm2);
        for (int i = 0; i < Nw; ++i) w[i] /= norm;
        double modulus = 0.0;
        for (int i = 0; i < Nw; ++i) modulus += w[i] * x[iL + i];
        for (int i = 0; i < Nw; ++i) w[i] *= modulus;
        atom->params[0] = modulus;
        atom->params[1] = modulus / norm;
    }
    for (int i = 0; i < Nw; ++i) x[iL + i] -= w[i];
    free(w);
}

void eo_waveform_norm2(double fs, int N, const double* params, int n, double* out) {
    double* x = (double*)calloc(N > 0 ? (size_t)N : 1, sizeof(double));
    for (int i = 0; i < n; ++i) {
        out[i] = NAN;
        if (!x) continue;
        eo_atom a;
        memset(&a, 0, sizeof a);
        for (int j = 0; j < 6; ++j) a.params[j] = params[6 * (size_t)i + j];
        eo_subtract(fs, &a, x, N, 0);
        double sum = 0.0;
        for (int m = 0; m < N; ++m) {
            sum += x[m] * x[m];
            x[m] = 0.0;
        }
        out[i] = sum;
    }
    free(x);
}

/* ---------------------------------------------------------------- workspace */

typedef struct {
    int k;                 /* scale index */
    int C, F, T;
    double* map;           /* [C][F][T] complex, interleaved */
    eo_heap* heap;
    unsigned char* mark;   /* [T] */
    int marked;
    eo_fft_plan plan;
    double *buf, *X, *z, *gather, *col;
} ws_map;

typedef struct {
    const eo_dict* d;
    int C;
    int constraint;
    ws_map* maps;
} workspace;

static double* map_at(ws_map* m, int c, int f, int t) {
    return m->map + 2 * (((size_t)c * m->F + f) * m->T + t);
}

static void ws_free(workspace* w) {
    if (!w->maps) return;
    for (int k = 0; k < w->d->L; ++k) {
        ws_map* m = &w->maps[k];
        free(m->map); free(m->mark); free(m->buf); free(m->X); free(m->z); free(m->gather); free(m->col);
        eo_heap_free(m->heap);
        plan_free(&m->plan);
    }
    free(w->maps);
    w->maps = NULL;
}

static int ws_init(workspace* w, const eo_dict* d, int C, int constraint) {
    w->d = d; w->C = C; w->constraint = constraint;
    w->maps = (ws_map*)calloc(d->L > 0 ? (size_t)d->L : 1, sizeof(ws_map));
    if (!w->maps) return EO_NO_MEMORY;
    for (int k = 0; k < d->L; ++k) {
        ws_map* m = &w->maps[k];
        const eo_scale* sc = &d->scales[k];
        m->k = k; m->C = C; m->F = sc->F; m->T = sc->T;
        m->map = (double*)calloc((size_t)2 * C * sc->F * sc->T, sizeof(double));
        m->mark = (unsigned char*)calloc((size_t)sc->T, 1);
        m->buf = (double*)malloc(sizeof(double) * (size_t)sc->Nfft);
        m->X = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
        m->z = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
        m->gather = (double*)malloc(sizeof(double) * 2 * (size_t)C);
        m->col = (double*)malloc(sizeof(double) * 2 * (size_t)sc->F);
        if (!m->map || !m->mark || !m->buf || !m->X || !m->z || !m->gather || !m->col ||
            plan_init(&m->plan, sc->Nfft) != EO_OK) {
            ws_free(w);
            return EO_NO_MEMORY;
        }
    }
    return EO_OK;
}

/* GaborWorkspaceMap::compute(signal, c, t) (gabor.cpp:279-303) */
static void map_column(workspace* w, ws_map* m, const double* x, int c, int t) {
    const eo_scale* sc = &w->d->scales[m->k];
    column_plan(w->d, m->k, &m->plan, x, t, m->col, m->buf, m->X, m->z);
    for (int f = 0; f < sc->F; ++f) {
        double* dst = map_at(m, c, f, t);
        dst[0] = m->col[2 * f];
        dst[1] = m->col[2 * f + 1];
    }
    if (m->heap) { m->mark[t] = 1; m->marked = 1; }
}

static double map_product(workspace* w, ws_map* m, const computer* g, int f, int t) {
    for (int c = 0; c < w->C; ++c) {
        double* src = map_at(m, c, f, t);
        m->gather[2 * c] = src[0];
        m->gather[2 * c + 1] = src[1];
    }
    return computer_square(g, w->d, t_of(w->d, m->k, t), m->gather, w->C, w->constraint);
}

/* GaborWorkspaceIndex ctor (gabor.cpp:327-339): products in f-major key order */
static int map_build_index(workspace* w, ws_map* m) {
    size_t n = (size_t)m->F * m->T;
    double* data = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!data) return EO_NO_MEMORY;
    size_t key = 0;
    for (int f = 0; f < m->F; ++f) {
        computer g = computer_of(w->d, m->k, f);
        for (int t = 0; t < m->T; ++t) data[key++] = map_product(w, m, &g, f, t);
    }
    eo_heap_free(m->heap);
    m->heap = eo_heap_new(data, n);
    free(data);
    memset(m->mark, 0, (size_t)m->T);
    m->marked = 0;
    return m->heap ? EO_OK : EO_NO_MEMORY;
}

/* GaborWorkspaceIndex::update (gabor.cpp:314-325) */
static void map_update(workspace* w, ws_map* m) {
    if (!m->marked) return;
    for (int f = 0; f < m->F; ++f) {
        computer g = computer_of(w->d, m->k, f);
        for (int t = 0; t < m->T; ++t) {
            if (!m->mark[t]) continue;
            size_t key = (size_t)f * m->T + t;
            eo_heap_update(m->heap, key, map_product(w, m, &g, f, t));
        }
    }
    memset(m->mark, 0, (size_t)m->T);
    m->marked = 0;
}

/* GaborWorkspace::compute (gabor.cpp:75-83) */
static int ws_compute(workspace* w, const double* signal /*[C][N]*/) {
    int status = EO_OK;
    const int L = w->d->L;
    const int N = w->d->N;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < L; ++k) {
        ws_map* m = &w->maps[k];
        eo_heap_free(m->heap);
        m->heap = NULL;
        for (int t = 0; t < m->T; ++t)
            for (int c = 0; c < w->C; ++c)
                map_column(w, m, signal + (size_t)c * N, c, t);
        if (map_build_index(w, m) != EO_OK) {
            #pragma omp critical
            status = EO_NO_MEMORY;
        }
    }
    return status;
}

/* GaborWorkspace::findBestMatch (gabor.cpp:85-133); also reports the
 * runner-up product for the near-tie measure.  follow (k, f, t) or NULL: an
 * entry to take instead of the best one when its product is at least
 * (1 - tol) of the best (*followed is set then). */
static int ws_find_best(workspace* w, eo_atom* out, double* gap, const int* follow, double tol, int* followed) {
    const int L = w->d->L;
    double* top = (double*)malloc(sizeof(double) * (L > 0 ? (size_t)L : 1));
    double* second = (double*)malloc(sizeof(double) * (L > 0 ? (size_t)L : 1));
    size_t* keys = (size_t*)malloc(sizeof(size_t) * (L > 0 ? (size_t)L : 1));
    if (!top || !second || !keys) { free(top); free(second); free(keys); return EO_NO_MEMORY; }
    #pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < L; ++k) {
        ws_map* m = &w->maps[k];
        map_update(w, m);
        eo_heap_peek(m->heap, &keys[k], &top[k]);
        second[k] = eo_heap_second(m->heap);
    }
    int bestK = -1;
    double bestV = 0.0;
    for (int k = 0; k < L; ++k) {
        if (top[k] > bestV) { bestV = top[k]; bestK = k; }
    }
    if (bestK < 0) { free(top); free(second); free(keys); return EO_ENERGY_MAP_IS_EMPTY; }
    double runner = second[bestK];
    for (int k = 0; k < L; ++k) if (k != bestK && top[k] > runner) runner = top[k];
    if (gap) *gap = (bestV - (runner > -INFINITY ? runner : 0.0)) / bestV;

    ws_map* m = &w->maps[bestK];
    int f = (int)(keys[bestK] / (size_t)m->T);
    int t = (int)(keys[bestK] % (size_t)m->T);
    if (followed) *followed = 0;
    if (follow) {
        const int fk = follow[0], ff = follow[1], ft = follow[2];
        if (fk >= 0 && fk < L && ff >= 0 && ff < w->maps[fk].F && ft >= 0 && ft < w->maps[fk].T &&
            (fk != bestK || ff != f || ft != t)) {
            computer gf = computer_of(w->d, fk, ff);
            const double pf = map_product(w, &w->maps[fk], &gf, ff, ft);
            if (pf >= bestV * (1.0 - tol)) {
                bestK = fk;
                m = &w->maps[fk];
                f = ff;
                t = ft;
                if (followed) *followed = 1;
            }
        }
    }
    computer g = computer_of(w->d, bestK, f);
    for (int c = 0; c < w->C; ++c) {
        double* src = map_at(m, c, f, t);
        m->gather[2 * c] = src[0];
        m->gather[2 * c + 1] = src[1];
    }
    if (w->constraint && !g.isHighFreq) eo_mmp1_constraint(m->gather, w->C);
    eo_atom_params(w->d, bestK, f, t, m->gather, w->C, out);
    free(top); free(second); free(keys);
    return EO_OK;
}

/* planted mistake for the tests (eo_plant_narrow_dirty_rule) */
static int g_narrow_dirty_rule = 0;

void eo_plant_narrow_dirty_rule(int on) { g_narrow_dirty_rule = on != 0; }

/* GaborWorkspace::subtractAtom (gabor.cpp:189-209) */
static void ws_subtract(workspace* w, eo_atom* atom, double* residual, int c) {
    const eo_dict* d = w->d;
    const double sA = g_narrow_dirty_rule ? 0.0 : atom->params[3] / d->fs;
    const double tA = atom->params[2] / d->fs;
    eo_subtract(d->fs, atom, residual, d->N, 0);
    const int L = d->L;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < L; ++k) {
        ws_map* m = &w->maps[k];
        for (int t = 0; t < m->T; ++t) {
            if (fabs(t_of(d, k, t) - tA) <= (d->scales[k].s + sA) * GAUSS_HALF_WIDTH)
                map_column(w, m, residual, c, t);
        }
    }
}

/* the stored maps of every scale, for a trace callback */
static void trace_call(workspace* w, eo_trace_fn fn, void* user, int iteration, int row, const double* residual) {
    const int L = w->d->L;
    eo_map_view* views = (eo_map_view*)calloc(L > 0 ? (size_t)L : 1, sizeof(eo_map_view));
    if (!views) return;
    for (int k = 0; k < L; ++k) {
        views[k].map = w->maps[k].map;
        views[k].C = w->maps[k].C;
        views[k].F = w->maps[k].F;
        views[k].T = w->maps[k].T;
    }
    fn(user, iteration, row, residual, w->C, w->d->N, views, L);
    free(views);
}

void eo_fresh_products(const eo_dict* d, int k, const double* x, int C, int constraint, double* out) {
    const eo_scale* sc = &d->scales[k];
    eo_fft_plan p;
    double* buf = (double*)malloc(sizeof(double) * (size_t)sc->Nfft);
    double* X = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)(sc->Nfft / 2 + 1));
    double* cols = (double*)malloc(sizeof(double) * 2 * (size_t)sc->F * (size_t)C);
    double* v = (double*)malloc(sizeof(double) * 2 * (size_t)C);
    if (plan_init(&p, sc->Nfft) == EO_OK && buf && X && z && cols && v) {
        for (int t = 0; t < sc->T; ++t) {
            for (int c = 0; c < C; ++c)
                column_plan(d, k, &p, x + (size_t)c * d->N, t, cols + 2 * (size_t)c * sc->F, buf, X, z);
            for (int f = 0; f < sc->F; ++f) {
                for (int c = 0; c < C; ++c) {
                    v[2 * c] = cols[2 * ((size_t)c * sc->F + f)];
                    v[2 * c + 1] = cols[2 * ((size_t)c * sc->F + f) + 1];
                }
                out[(size_t)f * sc->T + t] = eo_product(d, k, f, t, v, C, constraint);
            }
        }
    }
    plan_free(&p);
    free(buf); free(X); free(z); free(cols); free(v);
}

/* SingleSignal/MultiSignal::computeEnergy (base.hpp:39-58) */
static double energy_of(const double* x, int C, int N, double fs) {
    double total = 0.0;
    for (int c = 0; c < C; ++c) {
        double sum = 0.0;
        for (int i = 0; i < N; ++i) sum += x[(size_t)c * N + i] * x[(size_t)c * N + i];
        total += sum / fs;
    }
    return total;
}

/* Decomposition::compute (classes.cpp:26-73) on a C-channel signal.  atoms
 * [C][I] with row stride I; energies / gaps one row. */
typedef struct {
    eo_trace_fn trace;
    void* user;
    int row;
    const int* follow;     /* [I][3] of this row, or NULL */
    double tol;
    int* followed;         /* [I] of this row, or NULL */
} loop_ctl;

static int mp_loop(workspace* w, const double* signal, int C, const eo_settings* st, int iterLimit,
                   eo_atom* atoms, int I, int* count, double* energies, double* gaps, const loop_ctl* ctl) {
    const eo_dict* d = w->d;
    const int N = d->N;
    int status = ws_compute(w, signal);
    if (status != EO_OK) return status;
    double* residue = (double*)malloc(sizeof(double) * (size_t)C * N);
    eo_atom* best = (eo_atom*)malloc(sizeof(eo_atom) * (size_t)C);
    if (!residue || !best) { free(residue); free(best); return EO_NO_MEMORY; }
    memcpy(residue, signal, sizeof(double) * (size_t)C * N);
    const double totalEnergy = energy_of(signal, C, N, d->fs);
    double residueEnergy = totalEnergy;
    *count = 0;
    if (totalEnergy == 0) { free(residue); free(best); return EO_SIGNAL_IS_EMPTY; }
    energies[0] = totalEnergy;
    int limit = st->iterationMax < iterLimit ? st->iterationMax : iterLimit;
    for (int iteration = 1; iteration <= limit; ++iteration) {
        double gap = NAN;
        int fl = 0;
        status = ws_find_best(w, best, &gap, ctl->follow ? ctl->follow + 3 * (size_t)(iteration - 1) : NULL,
                              ctl->tol, &fl);
        if (ctl->followed) ctl->followed[iteration - 1] = fl;
        if (status != EO_OK) break;
        gaps[iteration - 1] = gap;
        for (int c = 0; c < C; ++c) atoms[(size_t)c * I + iteration - 1] = best[c];
        *count = iteration;
        if (iteration == st->iterationMax) break;
        for (int c = 0; c < C; ++c) ws_subtract(w, &best[c], residue + (size_t)c * N, c);
        if (ctl->trace) trace_call(w, ctl->trace, ctl->user, iteration, ctl->row, residue);
        residueEnergy = energy_of(residue, C, N, d->fs);
        energies[iteration] = residueEnergy;
        if (residueEnergy / totalEnergy <= st->residualEnergy) break;
    }
    free(residue);
    free(best);
    return status;
}

int eo_decompose_limited(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                         eo_atom* atoms, int* counts, double* energies, double* gaps,
                         int threads, int iterLimit) {
    return eo_decompose_traced(d, signal, C, st, atoms, counts, energies, gaps, threads, iterLimit, NULL, NULL);
}

int eo_decompose_traced(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                        eo_atom* atoms, int* counts, double* energies, double* gaps,
                        int threads, int iterLimit, eo_trace_fn trace, void* user) {
    return eo_decompose_follow(d, signal, C, st, atoms, counts, energies, gaps, threads, iterLimit, trace, user,
                               NULL, 0.0, NULL);
}

int eo_decompose_follow(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                        eo_atom* atoms, int* counts, double* energies, double* gaps,
                        int threads, int iterLimit, eo_trace_fn trace, void* user,
                        const int* follow, double tol, int* followed) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int I = st->iterationMax;
    const int N = d->N;
    if (C <= 0 || I <= 0) return EO_INVALID_ARGUMENT;
    int rows = st->variant == EO_SMP ? C : 1;
    for (size_t i = 0; i < (size_t)rows * (I + 1); ++i) energies[i] = NAN;
    for (size_t i = 0; i < (size_t)rows * I; ++i) gaps[i] = NAN;
    for (int c = 0; c < C; ++c) counts[c] = 0;
    if (followed)
        for (size_t i = 0; i < (size_t)rows * I; ++i) followed[i] = 0;
    loop_ctl ctl = {trace, user, 0, follow, tol, followed};
    workspace w;
    int status = EO_OK;
    if (st->variant == EO_SMP) {
        /* SmpDecomposition::compute (classes.cpp:77-87): channels in turn */
        status = ws_init(&w, d, 1, 0);
        if (status != EO_OK) return status;
        for (int c = 0; c < C && status == EO_OK; ++c) {
            loop_ctl cc = ctl;
            cc.row = c;
            cc.follow = follow ? follow + 3 * (size_t)c * I : NULL;
            cc.followed = followed ? followed + (size_t)c * I : NULL;
            status = mp_loop(&w, signal + (size_t)c * N, 1, st, iterLimit, atoms + (size_t)c * I, I,
                             &counts[c], energies + (size_t)c * (I + 1), gaps + (size_t)c * I, &cc);
        }
        ws_free(&w);
    } else if (st->variant == EO_MMP2) {
        /* Mmp2Decomposition::compute (mmp.cpp:23-45) */
        double* sum = (double*)malloc(sizeof(double) * (size_t)N);
        double* residue = (double*)malloc(sizeof(double) * (size_t)C * N);
        if (!sum || !residue) { free(sum); free(residue); return EO_NO_MEMORY; }
        memcpy(sum, signal, sizeof(double) * (size_t)N);
        for (int c = 1; c < C; ++c)
            for (int i = 0; i < N; ++i) sum[i] += signal[(size_t)c * N + i];
        status = ws_init(&w, d, 1, 0);
        if (status == EO_OK) {
            int n = 0;
            status = mp_loop(&w, sum, 1, st, iterLimit, atoms, I, &n, energies, gaps, &ctl);
            ws_free(&w);
            if (status == EO_OK) {
                memcpy(residue, signal, sizeof(double) * (size_t)C * N);
                /* atoms[0..n) hold the sum's atoms; refit per channel */
                eo_atom* sumAtoms = (eo_atom*)malloc(sizeof(eo_atom) * (size_t)(n ? n : 1));
                if (!sumAtoms) status = EO_NO_MEMORY;
                else {
                    memcpy(sumAtoms, atoms, sizeof(eo_atom) * (size_t)n);
                    for (int c = 0; c < C; ++c) {
                        for (int i = 0; i < n; ++i) {
                            eo_atom a = sumAtoms[i];
                            eo_subtract(d->fs, &a, residue + (size_t)c * N, N, 1);
                            atoms[(size_t)c * I + i] = a;
                        }
                        counts[c] = n;
                    }
                    free(sumAtoms);
                }
            }
        }
        free(sum);
        free(residue);
    } else if (st->variant == EO_MMP1 || st->variant == EO_MMP3) {
        status = ws_init(&w, d, C, st->variant == EO_MMP1);
        if (status != EO_OK) return status;
        int n = 0;
        status = mp_loop(&w, signal, C, st, iterLimit, atoms, I, &n, energies, gaps, &ctl);
        for (int c = 0; c < C; ++c) counts[c] = n;
        ws_free(&w);
    } else {
        return EO_INVALID_ARGUMENT;
    }
    return status;
}

int eo_decompose(const eo_dict* d, const double* signal, int C, const eo_settings* st,
                 eo_atom* atoms, int* counts, double* energies, double* gaps, int threads) {
    return eo_decompose_limited(d, signal, C, st, atoms, counts, energies, gaps, threads, INT_MAX);
}
