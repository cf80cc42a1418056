// Hash-grid encoder kernels and entry points.
//
// Work layout: one 64-lane wave takes a tile of 32 samples.  Both kernels are
// shaped by how gfx950 serves scattered memory instructions: a wave
// instruction costs one request per distinct 64-byte block its lanes touch
// (measured: float atomics sustain ~21 G requests/s whether a request carries
// 1 or 16 dwords, and wherever the data lives), so lanes are arranged for the
// two x-neighbour corners of a sample -- adjacent table entries -- to be in
// the same instruction.
#include <cmath>
#include <type_traits>

#include "grid.hpp"
#include "sampler.hpp"

namespace nect {

// ----------------------------------------------------------- position sources

template <int D>
struct PointSource {
    const float *x;
    int64_t n;
    int64_t stride;
    // 0: no such slot, 1: live slot without a sample, 2: sample.
    __device__ __forceinline__ int get(int64_t s, float (&p)[D]) const {
        if (s >= n) return 0;
        const float *r = x + s * stride;
#pragma unroll
        for (int d = 0; d < D; ++d) p[d] = r[d];
        return 2;
    }
};

// Samples of the fused training path: slot 32*tile + j of the tile list.
template <int D>
struct RaySource {
    const NectRay *rays;
    const int32_t *tile_offsets;
    const int32_t *tile_ray;
    const int32_t *n_tiles;
    float t;  // 4th coordinate (time) for D == 4
    __device__ __forceinline__ int get(int64_t slot, float (&p)[D]) const {
        const int64_t tile = slot / kTile;
        if (tile >= n_tiles[0]) return 0;
        const int r = tile_ray[tile];
        const int i = (int)(tile - tile_offsets[r]) * kTile + (int)(slot % kTile);
        const NectRay d = rays[r];
        float q[3];
        const bool ok = ray_sample(d, i, q);
#pragma unroll
        for (int k = 0; k < D && k < 3; ++k) p[k] = q[k];
        if (D == 4) p[D - 1] = t;
        return ok ? 2 : 1;
    }
};

// ------------------------------------------------------------------- kernels

template <int D>
__global__ __launch_bounds__(256) void k_grid_index(NectGridDesc g, const float *__restrict__ x, int64_t n,
                                                    int64_t x_stride, uint32_t *__restrict__ index,
                                                    float *__restrict__ weight) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = t / g.n_levels;
    const int l = (int)(t - s * g.n_levels);
    if (s >= n) return;
    float xs[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xs[d] = x[s * x_stride + d];
    uint32_t c[D];
    float f[D];
#pragma unroll
    for (int d = 0; d < D; ++d) grid_pos(g.scale[l], xs[d], c[d], f[d]);
    const int64_t base = t << D;
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
        uint32_t idx;
        float w;
        grid_corner<D>(g.resolution[l], g.level_size[l], c, f, k, idx, w);
        index[base + k] = idx;
        weight[base + k] = w;
    }
}

// Encoder, pair layout: lanes 2s and 2s+1 hold sample s and gather the
// corners with x-bit 0 and 1; x-neighbour entries are adjacent in the table
// (dense levels, and hashed levels unless x = 7 mod 8, the first hash prime
// being 1), so one load instruction fetches both from the same 64-byte block.
// The halves are added across the lane pair (cx = 0 half first).
// With GROUPED and vpp > 0 the output is group-planar: plane `grp` holds, per
// slot, the vpp = (L/G)*F features of the group's levels in group order.
template <int D, int F, typename T, typename Src, bool GROUPED>
__global__ __launch_bounds__(256) void k_grid_encode_pair(NectGridDesc g, const T *__restrict__ table, Src src,
                                                          int64_t n_slots, T *__restrict__ out, int64_t out_stride,
                                                          int n_groups, int vpp) {
    static_assert(!std::is_reference<Src>::value, "kernel arguments must be passed by value");
    const int lane = threadIdx.x & (kWave - 1);
    const int cx = lane & 1;
    const int grp = GROUPED ? (int)(blockIdx.x % (unsigned)n_groups) : 0;
    const int64_t block = GROUPED ? (int64_t)(blockIdx.x / (unsigned)n_groups) : (int64_t)blockIdx.x;
    const int64_t tile = (block * blockDim.x + threadIdx.x) / kWave;
    const int64_t slot = tile * kTile + (lane >> 1);
    float p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p[d] = 0.0f;
    const int st = slot < n_slots ? src.get(slot, p) : 0;
    // both lanes of a pair see the same st: the pair exchange below needs both active
    if (__ballot(st != 0) == 0) return;
    const bool ok = st == 2;
    const int64_t row = st != 0 ? slot : 0;
    const bool planar = GROUPED && vpp > 0;
    T *o = planar ? out + (int64_t)grp * n_slots * vpp + row * vpp : out + row * out_stride;
    for (int jl = 0;; ++jl) {
        const int l = GROUPED ? group_level(grp, jl, n_groups) : jl;
        if (l >= g.n_levels) break;
        const LevelInfo lv = level_info(g, l);
        uint32_t c[D];
        float f[D];
#pragma unroll
        for (int d = 0; d < D; ++d) grid_pos(lv.scale, p[d], c[d], f[d]);
        float v[F];
        encode_half<D, F, T>(lv, c, f, table, cx, v);
        float sum[F];
#pragma unroll
        for (int j = 0; j < F; ++j) {
            const float other = __shfl_xor(v[j], 1, kWave);
            sum[j] = cx == 0 ? v[j] + other : other + v[j];
        }
        if (st != 0 && cx == (l & 1)) {
            T *dst = o + (planar ? jl * F : l * F);
#pragma unroll
            for (int j = 0; j < F; ++j) dst[j] = from_f32<T>(ok ? sum[j] : 0.0f);
        }
    }
}

// Sum of v over samples m..end of this lane's pass (lanes 4m..4m+3 hold sample m);
// complete at the run's first sample.
__device__ __forceinline__ float run_sum_quad(float v, int m, int end) {
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        const float t = __shfl_down(v, 4 * off, kWave);
        if (m + off <= end) v += t;
    }
    return v;
}

// Backward, quad layout: lane 4m + 2cx + fq of a pass handles sample m of the
// pass (16 samples per pass, two passes per 32-sample tile), corners with
// x-bit cx and the feature half fq.  An atomic instruction then carries, for
// each sample, both x-neighbour corners and all features of the two entries,
// which share one 64-byte block -- one request -- unless the pair straddles a
// block boundary (dense levels: 1 in 8; hashed levels: x = 7 mod 8, the first
// hash prime being 1).  Four instructions per level and pass cover the 2^D
// corners (D = 3).
template <int D, int F, typename Src, bool GROUPED>
__global__ __launch_bounds__(256) void k_grid_backward_quad(NectGridDesc g, Src src, int64_t n_slots,
                                                            const float *__restrict__ dout, int64_t dout_stride,
                                                            float *__restrict__ grad, int n_groups) {
    static_assert(!std::is_reference<Src>::value, "kernel arguments must be passed by value");
    constexpr int FH = F / 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int m = lane >> 2;
    const int cx = (lane >> 1) & 1;
    const int fq = lane & 1;
    const int grp = GROUPED ? (int)(blockIdx.x % (unsigned)n_groups) : 0;
    const int64_t block = GROUPED ? (int64_t)(blockIdx.x / (unsigned)n_groups) : (int64_t)blockIdx.x;
    const int64_t tile = (block * blockDim.x + threadIdx.x) / kWave;
    float p[2][D];
    bool live[2], prev_live[2];
    const float *go[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int64_t slot = tile * kTile + 16 * r + m;
#pragma unroll
        for (int d = 0; d < D; ++d) p[r][d] = 0.0f;
        live[r] = slot < n_slots && src.get(slot, p[r]) == 2;
        go[r] = dout + (live[r] ? slot : 0) * dout_stride + fq * FH;
    }
    // from here on every lane of the wave stays active (cross-lane sums)
    if (__ballot(live[0] || live[1]) == 0) return;
#pragma unroll
    for (int r = 0; r < 2; ++r) prev_live[r] = __shfl_up((int)live[r], 4, kWave) != 0;
    const int bit = 4 * m + 3;
    for (int jl = 0;; ++jl) {
        const int l = GROUPED ? group_level(grp, jl, n_groups) : jl;
        if (l >= g.n_levels) break;
        const LevelInfo lv = level_info(g, l);
        float *lvl = grad + (size_t)lv.offset * F + fq * FH;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            uint32_t c[D];
            float f[D];
#pragma unroll
            for (int d = 0; d < D; ++d) grid_pos(lv.scale, p[r][d], c[d], f[d]);
            bool head = m == 0 || live[r] != prev_live[r];
#pragma unroll
            for (int d = 0; d < D; ++d) head |= (uint32_t)__shfl_up((int)c[d], 4, kWave) != c[d];
            float gv[FH];
#pragma unroll
            for (int j = 0; j < FH; ++j) gv[j] = live[r] ? go[r][l * F + j] : 0.0f;
            const uint64_t heads = __ballot(head);
            const bool merge = heads != ~0ull;  // the same on every lane
            int end = 15;
            if (merge) {
                const uint64_t after = bit == kWave - 1 ? 0ull : heads & (~0ull << (bit + 1));
                end = after ? (__builtin_ctzll(after) >> 2) - 1 : 15;
            }
#pragma unroll
            for (int q = 0; q < (1 << (D - 1)); ++q) {
                uint32_t idx;
                float w;
                grid_corner<D>(lv.res, lv.size, c, f, cx | (q << 1), idx, w);
                float v[FH];
#pragma unroll
                for (int j = 0; j < FH; ++j) v[j] = w * gv[j];
                if (merge) {
#pragma unroll
                    for (int j = 0; j < FH; ++j) v[j] = run_sum_quad(v[j], m, end);
                }
                if (head && live[r]) {
#pragma unroll
                    for (int j = 0; j < FH; ++j)
                        if (v[j] != 0.0f) unsafeAtomicAdd(lvl + (size_t)idx * F + j, v[j]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- dispatch

inline unsigned blocks_for_slots(int64_t n_slots) {
    const int64_t waves = (n_slots + kTile - 1) / kTile;
    return (unsigned)((waves + 3) / 4);  // 4 waves per 256-thread block
}

// Launch schedules.  The encoder's gathers are L2-bound, so it takes the
// level-grouped schedule (each XCD reads two levels' tables); the scatter's
// atomics are performed past the L2 and gain nothing from it, so it takes one
// wave per tile over all levels, which reads each dL/dfeature row once.
template <typename Src3, typename Src4>
static int32_t launch_encode(const NectGridDesc &g, int32_t dtype, const void *table, const Src3 &s3, const Src4 &s4,
                             int64_t n_slots, void *out, int64_t out_stride, int vpp, hipStream_t st) {
    const int G = level_groups(g.n_levels);
    const dim3 grid(blocks_for_slots(n_slots) * G), block(256);
#define NECT_ENC(D, F, T, S)                                                                                 \
    hipLaunchKernelGGL((k_grid_encode_pair<D, F, T, S, true>), grid, block, 0, st, g, (const T *)table, s##D, \
                       n_slots, (T *)out, out_stride, G, vpp)
#define NECT_ENC_T(D, F)                                                                    \
    do {                                                                                    \
        if (dtype == NECT_F32) NECT_ENC(D, F, float, Src##D);                               \
        else if (dtype == NECT_F16) NECT_ENC(D, F, _Float16, Src##D);                       \
        else NECT_ENC(D, F, bf16_t, Src##D);                                                \
    } while (0)
    if (g.n_dims == 3) {
        if (g.n_features == 2) NECT_ENC_T(3, 2);
        else NECT_ENC_T(3, 4);
    } else {
        if (g.n_features == 2) NECT_ENC_T(4, 2);
        else NECT_ENC_T(4, 4);
    }
#undef NECT_ENC_T
#undef NECT_ENC
    return check_launch("k_grid_encode_pair");
}

template <typename Src3, typename Src4>
static int32_t launch_backward(const NectGridDesc &g, const Src3 &s3, const Src4 &s4, int64_t n_slots,
                               const float *dout, int64_t dout_stride, float *grad, hipStream_t st) {
    const dim3 grid(blocks_for_slots(n_slots)), block(256);
#define NECT_BWD(D, F)                                                                                        \
    hipLaunchKernelGGL((k_grid_backward_quad<D, F, Src##D, false>), grid, block, 0, st, g, s##D, n_slots, dout, \
                       dout_stride, grad, 1)
    if (g.n_dims == 3) {
        if (g.n_features == 2) NECT_BWD(3, 2);
        else NECT_BWD(3, 4);
    } else {
        if (g.n_features == 2) NECT_BWD(4, 2);
        else NECT_BWD(4, 4);
    }
#undef NECT_BWD
    return check_launch("k_grid_backward_quad");
}

int32_t check_grid(const NectGridDesc *g) {
    NECT_REQUIRE(g != nullptr, "grid descriptor is NULL");
    NECT_REQUIRE(g->n_dims == 3 || g->n_dims == 4, "grid n_dims must be 3 or 4");
    NECT_REQUIRE(g->n_features == 2 || g->n_features == 4, "grid n_features_per_level must be 2 or 4");
    NECT_REQUIRE(g->n_levels >= 1 && g->n_levels <= NECT_MAX_LEVELS, "grid n_levels out of range");
    for (int l = 0; l < g->n_levels; ++l) NECT_REQUIRE(g->level_size[l] > 0, "grid level of size 0");
    return 0;
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_grid_init(NectGridDesc *desc, int32_t n_dims, int32_t n_levels, int32_t n_features,
                                  int32_t log2_hashmap_size, int32_t base_resolution, float per_level_scale) {
    clear_error();
    NECT_REQUIRE(desc != nullptr, "desc is NULL");
    NECT_REQUIRE(n_dims >= 2 && n_dims <= 4, "n_dims must be 2..4");
    NECT_REQUIRE(n_levels >= 1 && n_levels <= NECT_MAX_LEVELS, "n_levels must be 1..32");
    NECT_REQUIRE(n_features == 1 || n_features == 2 || n_features == 4 || n_features == 8,
                 "n_features_per_level must be 1, 2, 4 or 8");
    NECT_REQUIRE(log2_hashmap_size >= 1 && log2_hashmap_size <= 28, "log2_hashmap_size must be 1..28");
    NECT_REQUIRE(base_resolution >= 1, "base_resolution must be >= 1");
    NECT_REQUIRE(per_level_scale >= 1.0f, "per_level_scale must be >= 1");
    NectGridDesc g = {};
    g.n_dims = n_dims;
    g.n_levels = n_levels;
    g.n_features = n_features;
    g.log2_hashmap_size = log2_hashmap_size;
    g.base_resolution = base_resolution;
    g.per_level_scale = per_level_scale;
    const uint64_t T = (uint64_t)1 << log2_hashmap_size;
    uint64_t offset = 0;
    for (int l = 0; l < n_levels; ++l) {
        const double sc = (double)base_resolution * std::pow((double)per_level_scale, (double)l) - 1.0;
        const float scf = (float)sc;
        const double res = std::ceil((double)scf) + 1.0;
        NECT_REQUIRE(res < 4294967295.0, "grid resolution overflows uint32");
        const uint32_t r = (uint32_t)res;
        const double dense = std::pow((double)r, (double)n_dims);
        uint64_t size;
        if (dense > (double)T) {
            size = T;
        } else {
            uint64_t d = 1;
            for (int k = 0; k < n_dims; ++k) d *= r;
            size = ((d + 7) / 8) * 8;
            if (size > T) size = T;
        }
        g.scale[l] = scf;
        g.resolution[l] = r;
        g.level_size[l] = (uint32_t)size;
        g.offset[l] = (uint32_t)offset;
        offset += size;
        NECT_REQUIRE(offset * (uint64_t)n_features < ((uint64_t)1 << 32), "grid too large (> 2^32 parameters)");
    }
    g.offset[n_levels] = (uint32_t)offset;
    g.n_params = (int64_t)offset * n_features;
    *desc = g;
    return 0;
}

extern "C" int32_t nect_grid_index(const NectGridDesc *desc, const float *x, int64_t n, int64_t x_stride,
                                   uint32_t *index, float *weight, void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(x && index && weight, "NULL pointer");
    NECT_REQUIRE(x_stride >= desc->n_dims, "x_stride smaller than n_dims");
    const int64_t threads = n * desc->n_levels;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    if (desc->n_dims == 3)
        hipLaunchKernelGGL(k_grid_index<3>, grid, block, 0, as_stream(stream), *desc, x, n, x_stride, index, weight);
    else
        hipLaunchKernelGGL(k_grid_index<4>, grid, block, 0, as_stream(stream), *desc, x, n, x_stride, index, weight);
    return check_launch("k_grid_index");
}

extern "C" int32_t nect_grid_encode(const NectGridDesc *desc, int32_t dtype, const void *table, const float *x,
                                    int64_t n, int64_t x_stride, void *out, int64_t out_stride, void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(dtype >= NECT_F32 && dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(table && x && out, "NULL pointer");
    NECT_REQUIRE(x_stride >= desc->n_dims, "x_stride smaller than n_dims");
    NECT_REQUIRE(out_stride >= (int64_t)desc->n_levels * desc->n_features, "out_stride smaller than L*F");
    PointSource<3> s3{x, n, x_stride};
    PointSource<4> s4{x, n, x_stride};
    return launch_encode(*desc, dtype, table, s3, s4, n, out, out_stride, 0, as_stream(stream));
}

extern "C" int32_t nect_grid_backward(const NectGridDesc *desc, const float *x, int64_t n, int64_t x_stride,
                                      const float *dout, int64_t dout_stride, float *grad, void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(x && dout && grad, "NULL pointer");
    NECT_REQUIRE(x_stride >= desc->n_dims, "x_stride smaller than n_dims");
    NECT_REQUIRE(dout_stride >= (int64_t)desc->n_levels * desc->n_features, "dout_stride smaller than L*F");
    PointSource<3> s3{x, n, x_stride};
    PointSource<4> s4{x, n, x_stride};
    return launch_backward(*desc, s3, s4, n, dout, dout_stride, grad, as_stream(stream));
}

extern "C" int32_t nect_grid_encode_rays(const NectGridDesc *desc, int32_t dtype, const void *table,
                                         const NectRay *rays, const int32_t *tile_offsets, const int32_t *tile_ray,
                                         const int32_t *n_tiles, int64_t max_tiles, float t, void *out,
                                         int64_t out_stride, int32_t grouped, void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(dtype >= NECT_F32 && dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(table && rays && tile_offsets && tile_ray && n_tiles && out, "NULL pointer");
    NECT_REQUIRE(max_tiles >= 1, "max_tiles must be >= 1");
    int vpp = 0;
    if (grouped) {
        NECT_REQUIRE(desc->n_levels % 8 == 0, "the grouped feature layout needs a multiple of 8 levels");
        vpp = grouped_vpp(*desc);
    } else {
        NECT_REQUIRE(out_stride >= (int64_t)desc->n_levels * desc->n_features, "out_stride smaller than L*F");
    }
    RaySource<3> s3{rays, tile_offsets, tile_ray, n_tiles, t};
    RaySource<4> s4{rays, tile_offsets, tile_ray, n_tiles, t};
    return launch_encode(*desc, dtype, table, s3, s4, max_tiles * kTile, out, out_stride, vpp, as_stream(stream));
}

extern "C" int32_t nect_grid_backward_rays(const NectGridDesc *desc, const NectRay *rays,
                                           const int32_t *tile_offsets, const int32_t *tile_ray,
                                           const int32_t *n_tiles, int64_t max_tiles, float t, const float *dout,
                                           int64_t dout_stride, float *grad, void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(rays && tile_offsets && tile_ray && n_tiles && dout && grad, "NULL pointer");
    NECT_REQUIRE(max_tiles >= 1, "max_tiles must be >= 1");
    NECT_REQUIRE(dout_stride >= (int64_t)desc->n_levels * desc->n_features, "dout_stride smaller than L*F");
    RaySource<3> s3{rays, tile_offsets, tile_ray, n_tiles, t};
    RaySource<4> s4{rays, tile_offsets, tile_ray, n_tiles, t};
    return launch_backward(*desc, s3, s4, max_tiles * kTile, dout, dout_stride, grad, as_stream(stream));
}
