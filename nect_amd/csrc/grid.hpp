// Multiresolution hash-grid encoding, device side.
//
// Specification (shared with the oracle; tiny-cuda-nn's HashGrid conventions
// re-stated, since tiny-cuda-nn is not part of the reference tree):
//   p_d = fmaf(scale, x_d, 0.5) (one rounding),  c_d = floor(p_d),  f_d = p_d - c_d
//   corner k (bit d of k set -> c_d + 1) has weight prod_d (f_d | 1 - f_d)
//   index = dense sum c_d * res^d while stride <= level_size, else
//           xor_d c_d * prime_d (uint32), then % level_size.
#pragma once

#include "common.hpp"

namespace nect {

__device__ constexpr uint32_t hash_prime(int d) {
    return d == 0 ? 1u : d == 1 ? 2654435761u : d == 2 ? 805459861u : 3674653429u;
}

// bfloat16 is carried as raw uint16 bits.
struct bf16_t {
    uint16_t bits;
};

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(_Float16 v) { return (float)v; }
__device__ __forceinline__ float to_f32(bf16_t v) { return __uint_as_float((uint32_t)v.bits << 16); }
__device__ __forceinline__ float to_f32(__bf16 v) { return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, v) << 16); }

template <typename T>
__device__ __forceinline__ T from_f32(float v);
template <>
__device__ __forceinline__ float from_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ _Float16 from_f32<_Float16>(float v) { return (_Float16)v; }
// Round to nearest, ties to even, subnormals kept, NaN stays NaN: gfx950's
// v_cvt_pk_bf16_f32 does it in one instruction (the integer rounding sequence
// it replaces took six and a branch for NaN).
template <>
__device__ __forceinline__ bf16_t from_f32<bf16_t>(float v) {
    bf16_t r;
    r.bits = __builtin_bit_cast(uint16_t, (__bf16)v);
    return r;
}

// v of lane + OFF for lanes whose lane + OFF lies in their own row of 16 lanes
// (a DPP row shift: no data goes through the LDS crossbar, as a shuffle's
// does); the other lanes get their own v.
template <int OFF>
__device__ __forceinline__ float row_down(float v) {
    static_assert(OFF >= 1 && OFF <= 15, "row shifts move 1..15 lanes");
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(x, x, 0x100 | OFF, 0xf, 0xf, false));
}

// x of lane - OFF, likewise within the lane's row of 16 lanes.
template <int OFF>
__device__ __forceinline__ int row_up(int x) {
    static_assert(OFF >= 1 && OFF <= 15, "row shifts move 1..15 lanes");
    return __builtin_amdgcn_update_dpp(x, x, 0x110 | OFF, 0xf, 0xf, false);
}

// Two values in one instruction: a in the low half, b in the high half.
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Position of one coordinate on a level: integer corner and fraction.  The
// position scale*x + 0.5 is rounded once (a fused multiply-add), as
// tiny-cuda-nn's HashGrid computes it.
__device__ __forceinline__ void grid_pos(float scale, float x, uint32_t &c, float &f) {
    const float p = fmaf(scale, x, 0.5f);
    const float fl = floorf(p);
    c = (uint32_t)(int)fl;
    f = p - fl;
}

template <int D>
__device__ __forceinline__ uint32_t grid_index(uint32_t res, uint32_t size, const uint32_t (&c)[D]) {
    uint64_t stride = 1;
    uint32_t index = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (stride > size) break;
        index += c[d] * (uint32_t)stride;
        stride *= res;
    }
    if (stride > size) {
        uint32_t h = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) h ^= c[d] * hash_prime(d);
        index = h;
    }
    return index % size;
}

// Corner k of a sample on one level: its entry (within the level) and weight.
template <int D>
__device__ __forceinline__ void grid_corner(uint32_t res, uint32_t size, const uint32_t (&c)[D], const float (&f)[D],
                                            int k, uint32_t &idx, float &w) {
    uint32_t cc[D];
    w = 1.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if ((k >> d) & 1) {
            w = w * f[d];
            cc[d] = c[d] + 1;
        } else {
            w = w * (1.0f - f[d]);
            cc[d] = c[d];
        }
    }
    idx = grid_index<D>(res, size, cc);
}

template <typename T, int F>
struct Entry {
    T v[F];
};

// Constants of one level.
struct LevelInfo {
    float scale;
    uint32_t res, size, offset;
};

__device__ __forceinline__ LevelInfo level_info(const NectGridDesc &g, int l) {
    return LevelInfo{g.scale[l], g.resolution[l], g.level_size[l], g.offset[l]};
}

// Level groups.  With G groups, level l belongs to group (l / G) even ? l % G
// : G - 1 - l % G, so group g holds levels g, 2G-1-g, 2G+g, ... (for L = 16,
// G = 8: {g, 15-g}, one coarse and one fine level).  A grouped launch gives
// block b the levels of group b % G: consecutive workgroups are dealt to the
// XCDs in turn, so with G = 8 each XCD touches the tables of two levels only
// and keeps them in its own L2.
__host__ __device__ __forceinline__ int group_level(int grp, int j, int G) {
    return j * G + ((j & 1) ? G - 1 - grp : grp);
}
__host__ __device__ __forceinline__ int level_groups(int n_levels) { return n_levels < 8 ? n_levels : 8; }

// Features per plane of the group-planar layout (levels per group times F;
// defined for a multiple of 8 levels).
__host__ __device__ __forceinline__ int grouped_vpp(const NectGridDesc &g) { return g.n_levels / 8 * g.n_features; }

// Weighted sum over the corners k with x-bit cx of one level (float, corners
// in increasing k).  The interpolated feature is the sum of the cx = 0 and
// cx = 1 halves, added in that order; every encoder in this library forms it
// that way, so a feature is the same bit pattern whichever kernel made it.
template <int D, int F, typename T>
__device__ __forceinline__ void encode_half(const LevelInfo &lv, const uint32_t (&c)[D], const float (&f)[D],
                                            const T *__restrict__ table, int cx, float (&out)[F]) {
    const Entry<T, F> *lvl = reinterpret_cast<const Entry<T, F> *>(table) + lv.offset;
#pragma unroll
    for (int j = 0; j < F; ++j) out[j] = 0.0f;
#pragma unroll
    for (int q = 0; q < (1 << (D - 1)); ++q) {
        uint32_t idx;
        float w;
        grid_corner<D>(lv.res, lv.size, c, f, cx | (q << 1), idx, w);
        const Entry<T, F> e = lvl[idx];
#pragma unroll
        for (int j = 0; j < F; ++j) out[j] = fmaf(w, to_f32(e.v[j]), out[j]);
    }
}

// Interpolated features of one level (float accumulation).
template <int D, int F, typename T>
__device__ __forceinline__ void encode_level(const LevelInfo &lv, const float (&x)[D], const T *__restrict__ table,
                                             float (&out)[F]) {
    uint32_t c[D];
    float f[D];
#pragma unroll
    for (int d = 0; d < D; ++d) grid_pos(lv.scale, x[d], c[d], f[d]);
    float odd[F];
    encode_half<D, F, T>(lv, c, f, table, 0, out);
    encode_half<D, F, T>(lv, c, f, table, 1, odd);
#pragma unroll
    for (int j = 0; j < F; ++j) out[j] = out[j] + odd[j];
}

// Scatter of one level's feature gradient into the float gradient table.
template <int D, int F>
__device__ __forceinline__ void scatter_level(const LevelInfo &lv, const float (&x)[D], const float (&dout)[F],
                                              float *__restrict__ grad) {
    uint32_t c[D];
    float f[D];
#pragma unroll
    for (int d = 0; d < D; ++d) grid_pos(lv.scale, x[d], c[d], f[d]);
    float *lvl = grad + (size_t)lv.offset * F;
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
        uint32_t idx;
        float w;
        grid_corner<D>(lv.res, lv.size, c, f, k, idx, w);
#pragma unroll
        for (int j = 0; j < F; ++j) unsafeAtomicAdd(lvl + (size_t)idx * F + j, w * dout[j]);
    }
}

// Workgroup barrier ordering LDS only.  __syncthreads() also orders global
// memory, which makes every wave wait for its outstanding global stores and
// atomics to be performed (for the scatter's atomics: far past the L2, behind
// every other CU's); a loop that only reuses LDS needs its LDS traffic settled,
// and the global writes may complete whenever they do.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace nect
