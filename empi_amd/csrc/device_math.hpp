// Complex arithmetic, the small forward DFTs of the Stockham passes and the
// hardware sine/cosine twiddles, shared by the column and region kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace empi {

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ inline float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// forward DFTs, w = e^{-2 pi i / R}
__device__ inline void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
__device__ inline void dft4(float2* v) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}
__device__ inline void dft8(float2* v) {
    const float r = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e);
    dft4(o);
    // o_k *= e^{-i pi k / 4}
    float2 o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
    float2 o2 = cmul_mi(o[2]);
    float2 o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}
// 16 points as two interleaved 8-point transforms and the e^{-2 pi i k / 16} butterflies
__device__ inline void dft16(float2* v) {
    const float c1 = 0.92387953251128675613f;   // cos(pi / 8)
    const float s1 = 0.38268343236508977173f;   // sin(pi / 8)
    const float r = 0.70710678118654752440f;
    float2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft8(e);
    dft8(o);
    float2 t[8];
    t[0] = o[0];
    t[1] = make_float2(c1 * o[1].x + s1 * o[1].y, c1 * o[1].y - s1 * o[1].x);
    t[2] = make_float2(r * (o[2].x + o[2].y), r * (o[2].y - o[2].x));
    t[3] = make_float2(s1 * o[3].x + c1 * o[3].y, s1 * o[3].y - c1 * o[3].x);
    t[4] = cmul_mi(o[4]);
    t[5] = make_float2(c1 * o[5].y - s1 * o[5].x, -(s1 * o[5].y + c1 * o[5].x));
    t[6] = make_float2(r * (o[6].y - o[6].x), -r * (o[6].x + o[6].y));
    t[7] = make_float2(s1 * o[7].y - c1 * o[7].x, -(c1 * o[7].y + s1 * o[7].x));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = cadd(e[k], t[k]);
        v[k + 8] = csub(e[k], t[k]);
    }
}
template <int R>
__device__ inline void dft(float2* v) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 8) dft8(v);
    else dft16(v);
}

// radix of the Stockham pass at sub-length Ns: 8 where the remaining length
// is a power of 8, otherwise a 16 first so that the passes stay at
// ceil(log2(left) / 3) or fewer (2048 = 16 * 16 * 8, 1024 = 16 * 8 * 8);
// 4096 = 16^3 (three passes, measured 2% faster than 8^4)
constexpr int pick_radix(int left) {
    return left <= 2 ? left : left == 4 ? 4 : (left == 8 || left == 32) ? 8
         : left == 4096 ? 16 : (ilog2c(left) % 3 == 0 ? 8 : 16);
}

// e^{2 pi i rev} from the hardware sine/cosine (argument in revolutions;
// absolute error 1.3e-7 for |rev| <= 1.5, measured on gfx950)
__device__ inline float2 cis_rev(float rev) {
    return make_float2(__builtin_amdgcn_cosf(rev), __builtin_amdgcn_sinf(rev));
}

__device__ inline unsigned long long pack_best(float p, unsigned bin) {
    if (!(p > 0.0f)) p = 0.0f;   // NaN and negatives never win
    return (static_cast<unsigned long long>(__float_as_uint(p)) << 32) | (0xFFFFFFFFu - bin);
}
// pack_best of a value known to be >= 0 (a sum of squares times a positive
// constant): the same bits without the clamp
__device__ inline unsigned long long pack_nonneg(float p, unsigned bin) {
    return (static_cast<unsigned long long>(__float_as_uint(p)) << 32) | (0xFFFFFFFFu - bin);
}

// The argmax keys, shared by the column kernels, the selection and the
// argmax test hook.  Every level orders by product, and exact ties by the
// smallest scale, then bin, then time index:
//   cell best  (product bits << 32) | ~bin                   (pack_best)
//   tile best  (product bits << 32) | ~(bin << cell_bits | cell in tile)
//   selection  product desc, then (scale << 52 | bin << 26 | time) asc
__device__ inline unsigned long long tile_key(unsigned long long cell_best, int cell, int cell_bits) {
    const unsigned bin = 0xFFFFFFFFu - static_cast<unsigned>(cell_best);
    const unsigned key = (bin << cell_bits) | static_cast<unsigned>(cell);
    return (cell_best & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFu - key);
}
struct SelBest {
    float p;
    unsigned long long key;   // (scale << 52) | (bin << 26) | t
};
__device__ inline bool sel_better(const SelBest& a, const SelBest& b) {
    return a.p > b.p || (a.p == b.p && a.key < b.key);
}
__device__ inline SelBest sel_key(float p, int scale, unsigned bin, int t) {
    return SelBest{p, (static_cast<unsigned long long>(scale) << 52) | (static_cast<unsigned long long>(bin) << 26) |
                          static_cast<unsigned long long>(t)};
}
// candidate of tile `tile` (ncell cells, G per time index) of scale `scale`
__device__ inline SelBest sel_of_tile(unsigned long long tile_best, int scale, int tile, int ncell, int G,
                                      int cell_bits) {
    const unsigned key = 0xFFFFFFFFu - static_cast<unsigned>(tile_best & 0xFFFFFFFFu);
    const int cell = tile * ncell + static_cast<int>(key & ((1u << cell_bits) - 1u));
    return sel_key(__uint_as_float(static_cast<unsigned>(tile_best >> 32)), scale, key >> cell_bits, cell / G);
}

// A float array as a raw buffer resource (stride 0, 32-bit data format): a
// load at a byte offset at or past `bytes` returns zero instead of touching
// memory.  p and bytes must be the same in every lane of the wave: they are
// taken from the first active lane, so that the resource sits in scalar
// registers (a resource the compiler cannot prove uniform costs a loop over
// lanes at every load).
__device__ inline __amdgpu_buffer_rsrc_t buffer_of(const float* p, int bytes) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a >> 32));
    float* u = reinterpret_cast<float*>((static_cast<unsigned long long>(hi) << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, 0, __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ inline float buffer_load(__amdgpu_buffer_rsrc_t r, unsigned byte_offset) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(byte_offset), 0, 0));
}

// e^{2 pi i cyc} from the hardware sine/cosine, whose argument is in
// revolutions (absolute error 1.3e-7 for |cyc| <= 1.5, measured on gfx950)
__device__ inline void twiddle_rev(float cyc, float* sn, float* cs) {
    *sn = __builtin_amdgcn_sinf(cyc);
    *cs = __builtin_amdgcn_cosf(cyc);
}

// Orders one wave's LDS accesses across its lanes.  LDS serves a wave's
// instructions in issue order, so only the compiler must be kept from moving
// accesses across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const unsigned lo = __shfl_xor(static_cast<unsigned>(v), m);
    const unsigned hi = __shfl_xor(static_cast<unsigned>(v >> 32), m);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

}  // namespace empi
