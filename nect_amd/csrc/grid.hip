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

// An unsigned type of BYTES bytes (one level's features in the staged encoder).
template <int BYTES>
struct ChunkOf;
template <>
struct ChunkOf<4> { using type = uint32_t; };
template <>
struct ChunkOf<8> { using type = uint2; };
template <>
struct ChunkOf<16> { using type = uint4; };

// Encoder, pair layout: lanes 2s and 2s+1 hold sample s and gather the
// corners with x-bit 0 and 1; x-neighbour entries are adjacent in the table
// (dense levels, and hashed levels unless x = 7 mod 8, the first hash prime
// being 1), so one load instruction fetches both from the same 64-byte block.
// The halves are added across the lane pair (cx = 0 half first).
// With GROUPED and vpp > 0 the output is group-planar: plane `grp` holds, per
// slot, the vpp = (L/G)*F features of the group's levels in group order.
// STAGE (one level group, rows of out_stride elements, a PointSource: every
// slot below n_slots present): the wave's 32 rows are assembled in `stage`
// (LDS, kTile * L * F elements per wave) and written back one level's F values
// per lane, each row's L * F values by consecutive lanes -- where a store per
// level leaves the L2 as a partial line of its own once the gathers have
// evicted the row's line.
// The body, for workgroup index bx (the kernels below give it blockIdx.x).
template <int D, int F, typename T, typename Src, bool GROUPED, bool STAGE = false>
__device__ __forceinline__ void encode_pair_tile(const NectGridDesc &g, const T *__restrict__ table, const Src &src,
                                                 int64_t n_slots, T *__restrict__ out, int64_t out_stride,
                                                 int n_groups, int vpp, unsigned bx, T *stage = nullptr) {
    static_assert(!STAGE || std::is_same<Src, PointSource<D>>::value,
                  "the staged write-back copies every row below n_slots");
    const int lane = threadIdx.x & (kWave - 1);
    const int cx = lane & 1;
    const int grp = GROUPED ? (int)(bx % (unsigned)n_groups) : 0;
    const int64_t block = GROUPED ? (int64_t)(bx / (unsigned)n_groups) : (int64_t)bx;
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
    const int pitch = g.n_levels * F;
    T *o = STAGE ? stage + (threadIdx.x >> 1) * pitch
                 : (planar ? out + (int64_t)grp * n_slots * vpp + row * vpp : out + row * out_stride);
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
            T *dst = o + (planar && !STAGE ? jl * F : l * F);
#pragma unroll
            for (int j = 0; j < F; ++j) dst[j] = from_f32<T>(ok ? sum[j] : 0.0f);
        }
    }
    if constexpr (STAGE) {
        // the wave's own LDS rows: LDS operations of one wave complete in order
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        using Chunk = typename ChunkOf<F * (int)sizeof(T)>::type;
        const T *ws = stage + (threadIdx.x / kWave) * kTile * pitch;
        const int64_t row0 = tile * kTile;
        for (int i = lane; i < kTile * g.n_levels; i += kWave) {
            const int r = i / g.n_levels;
            if (row0 + r >= n_slots) break;
            const int c = (i - r * g.n_levels) * F;
            *reinterpret_cast<Chunk *>(out + (row0 + r) * out_stride + c) =
                *reinterpret_cast<const Chunk *>(ws + r * pitch + c);
        }
    }
}

template <int D, int F, typename T, typename Src, bool GROUPED>
__global__ __launch_bounds__(256) void k_grid_encode_pair(NectGridDesc g, const T *__restrict__ table, Src src,
                                                          int64_t n_slots, T *__restrict__ out, int64_t out_stride,
                                                          int n_groups, int vpp) {
    static_assert(!std::is_reference<Src>::value, "kernel arguments must be passed by value");
    encode_pair_tile<D, F, T, Src, GROUPED>(g, table, src, n_slots, out, out_stride, n_groups, vpp, blockIdx.x);
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
// The quad layout's atomic adds of one corner pair: lane 4m + 2cx + fq adds its
// FH = F/2 values v (features fq FH .. of the entry idx of x corner cx) at
// base + idx F, base being the level's first float plus fq FH, when `act`.
// With four features, lane L and lane L ^ 32 (the same corner and feature half
// of samples m and m ^ 8) trade their second values, so that each instruction
// carries all four features of both x corners of 8 samples -- one 64-byte block,
// one request per sample -- where two instructions of 16 half-entries each would
// request every block twice.  Every lane of the wave must call it.
template <int F>
__device__ __forceinline__ void add_corner_pair(float *__restrict__ base, uint32_t idx, const float (&v)[F / 2],
                                                bool act, int lane) {
    constexpr int FH = F / 2;
    if constexpr (FH == 2) {
        const bool actx = __shfl_xor((int)act, 32, kWave) != 0;
        const uint32_t idxx = (uint32_t)__shfl_xor((int)idx, 32, kWave);
        const float v1x = __shfl_xor(v[1], 32, kWave);
        float *const own = base + (size_t)idx * F;
        float *const other = base + (size_t)idxx * F + 1;
        const bool lo = lane < 32;
        // samples 0..7 of the pass: lanes < 32 their first values, lanes >= 32 their second ones
        if (lo ? act : actx) {
            const float val = lo ? v[0] : v1x;
            if (val != 0.0f) unsafeAtomicAdd(lo ? own : other, val);
        }
        // samples 8..15
        if (lo ? actx : act) {
            const float val = lo ? v1x : v[0];
            if (val != 0.0f) unsafeAtomicAdd(lo ? other : own, val);
        }
    } else {
        if (act) {
#pragma unroll
            for (int j = 0; j < FH; ++j)
                if (v[j] != 0.0f) unsafeAtomicAdd(base + (size_t)idx * F + j, v[j]);
        }
    }
}

template <int D, int F, typename Src, bool GROUPED>
__device__ __forceinline__ void backward_quad_tile(const NectGridDesc &g, const Src &src, int64_t n_slots,
                                                   const float *__restrict__ dout, int64_t dout_stride,
                                                   float *__restrict__ grad, int n_groups, unsigned bx) {
    constexpr int FH = F / 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int m = lane >> 2;
    const int cx = (lane >> 1) & 1;
    const int fq = lane & 1;
    const int grp = GROUPED ? (int)(bx % (unsigned)n_groups) : 0;
    const int64_t block = GROUPED ? (int64_t)(bx / (unsigned)n_groups) : (int64_t)bx;
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
                add_corner_pair<F>(lvl, idx, v, head && live[r], lane);
            }
        }
    }
}

template <int D, int F, typename Src, bool GROUPED>
__global__ __launch_bounds__(256) void k_grid_backward_quad(NectGridDesc g, Src src, int64_t n_slots,
                                                            const float *__restrict__ dout, int64_t dout_stride,
                                                            float *__restrict__ grad, int n_groups) {
    static_assert(!std::is_reference<Src>::value, "kernel arguments must be passed by value");
    backward_quad_tile<D, F, Src, GROUPED>(g, src, n_slots, dout, dout_stride, grad, n_groups, blockIdx.x);
}

// Several grids over their own input columns in one launch (blockIdx.y = the
// grid): the nested encodings of a tiny-cuda-nn Composite (QuadCubes,
// HyperCubes), every part with the same D and F.
constexpr int kMaxParts = NECT_MAX_GRID_PARTS;
struct GridParts {
    NectGridDesc g[kMaxParts];
    int64_t table_off[kMaxParts];   // elements of the tables vector
    int32_t col0[kMaxParts];        // first input column
    int32_t out0[kMaxParts];        // first output column
};

// A workgroup encodes every level of its samples for one grid, so the grid's
// columns of a feature row are written by one workgroup.  The level-group
// schedule of the single-grid encoder (levels dealt to XCDs) would split each
// row into pieces written from 8 XCDs' L2s: on the QuadCubes preset (4 grids x
// 23 levels x 4 FP16 features into rows of 392) the step took 74.5 ms against
// 71.0 with whole pieces.  Grids stay the slowest launch index (blockIdx.y), so
// the workgroups in flight gather from one grid's tables: one workgroup
// encoding all grids of its samples measured 74.1 ms.  The staged write-back
// (STAGE) then took the launch's writes from 15.1 to 3.9 GiB for 3.7 GB of
// features, and the encoder from 9.06 to 5.75 ms (step 67.8 ms).
// STAGE: rows assembled in LDS (dynamic, 4 * kTile * L * F elements) and
// written back whole per grid (the launch checks the alignment it needs).
template <int D, int F, typename T, bool STAGE>
__global__ __launch_bounds__(256) void k_grid_encode_parts(GridParts P, const T *__restrict__ tables,
                                                           const float *__restrict__ x, int64_t n, int64_t x_stride,
                                                           T *__restrict__ out, int64_t out_stride) {
    extern __shared__ __align__(16) char smem_raw[];
    const int k = blockIdx.y;
    const PointSource<D> src{x + P.col0[k], n, x_stride};
    encode_pair_tile<D, F, T, PointSource<D>, true, STAGE>(P.g[k], tables + P.table_off[k], src, n,
                                                            out + P.out0[k], out_stride, 1, 0, blockIdx.x,
                                                            reinterpret_cast<T *>(smem_raw));
}

template <int D, int F>
__global__ __launch_bounds__(256) void k_grid_backward_parts(GridParts P, const float *__restrict__ x, int64_t n,
                                                             int64_t x_stride, const float *__restrict__ dout,
                                                             int64_t dout_stride, float *__restrict__ grad) {
    const int k = blockIdx.y;
    const PointSource<D> src{x + P.col0[k], n, x_stride};
    backward_quad_tile<D, F, PointSource<D>, false>(P.g[k], src, n, dout + P.out0[k], dout_stride,
                                                     grad + P.table_off[k], 1, blockIdx.x);
}

// Backward over ray groups, with block aggregation.  The unit of work is 64
// consecutive rays of `order` (spatial neighbours: nect_ray_order groups a
// batch by 8x8 detector patches) times one 16-sample pass of their tile j.
// At the coarse and middle levels the unit's samples fall into far fewer
// 64-byte blocks of the table than they would issue requests one ray at a
// time (the ray-local run sums only merge samples of one ray), so a level with
// plan.rounds[l] = R > 0 is summed in an LDS table of blocks keyed by the
// block's global index and flushed with one atomic request per touched block
// (16 lanes, one per dword), after every 16/R samples of the 64 rays: finer
// levels touch more blocks per sample and take more, smaller rounds to fit
// the table.  Blocks that find no LDS slot within kAggProbes probes go
// straight to memory, so every plan is correct.  Levels with R <= 0 are
// scattered directly in the quad layout of k_grid_backward_quad: R = 0 puts
// 16 samples of one ray in an instruction (runs in one cell summed first),
// R = -1 one sample of 16 neighbouring rays (with the patch order a 4x4 block
// of pixels, whose x-neighbour entries share 64-byte blocks at the middle
// levels).  Up to 8 levels' dL/dfeature of a sample are loaded at once.
// The aggregated levels run one lane per sample (each lane computes its
// position once and handles all corners; only runs along a ray are summed
// across lanes) and issue a sample's table lookups together.  One launch
// serves both: workgroups in their LDS-bound phase overlap others in their
// request-bound phase.
constexpr int kUnitRays = 64;
constexpr int kUnitSamples = 16;
constexpr int kRaysPerWave = kUnitRays / 4;
constexpr int kAggLog2Slots = 9;  // 512 blocks: 32 KiB of values, four workgroups per CU
constexpr int kAggSlots = 1 << kAggLog2Slots;
constexpr int kAggProbes = 16;
constexpr int kLevelChunk = 8;
constexpr uint32_t kEmptyKey = 0xffffffffu;

struct AggPlan {
    // 0: direct, one ray per instruction; -1: direct, 16 rays per instruction;
    // 1, 2, 4, 8, 16: LDS rounds
    int8_t rounds[NECT_MAX_LEVELS];
    // grid dimension d reads sample coordinate ((cmap >> 2d) & 3) ^ d of
    // (z, y, x, t): 0 is the identity (a nested grid of a composite encoding
    // over, e.g., (y, x, t) has its own map)
    uint32_t cmap;
};

// dL/dfeatures as the grouped scatter reads it: rows of `stride` elements of
// `dtype` (NECT_F32, NECT_F16 or NECT_BF16), each value multiplied by *scale
// when scale is set (a 16-bit network backward's scaled gradient and the
// reciprocal of its power-of-two scale: the products are exact, so the values
// equal those of the unscaled float32 copy, which is then never written).
struct DoutView {
    const void *p;
    int64_t stride;
    int32_t dtype;
    const float *scale;
};

__device__ __forceinline__ float dout_at(const DoutView &d, float sc, int64_t i) {
    if (d.dtype == NECT_F16) return (float)static_cast<const _Float16 *>(d.p)[i] * sc;
    if (d.dtype == NECT_BF16) return to_f32(static_cast<const bf16_t *>(d.p)[i]) * sc;
    return static_cast<const float *>(d.p)[i] * sc;
}

// The grid input of a sample at (z, y, x) = q and time t under the column map.
template <int D>
__device__ __forceinline__ void place_sample(uint32_t cmap, const float (&q)[3], float t, float (&p)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t k = ((cmap >> (2 * d)) & 3u) ^ (uint32_t)d;
        p[d] = k == 0 ? q[0] : (k == 1 ? q[1] : (k == 2 ? q[2] : t));
    }
}

// One step of the run sums along a ray in the one-lane-per-sample layout: lane
// i adds the values of lane i + OFF of its row when `add`.  ROWS: read them
// with DPP row shifts (VALU) rather than shuffles (the LDS crossbar).  Row
// shifts pay where the aggregated levels carry many values per sample (four
// features: QuadCubes' scatter 9.57 -> 9.38 ms per grid; four dimensions, 16
// corners: the 4-D two-feature scatter 46.6 -> 45.6 ms) and cost time where
// they carry few (the 3-D two-feature grid of the 256^3 workload: 19.01 ->
// 19.43 ms, the VALU slots and DPP wait states outweighing the LDS traffic).
template <int OFF, bool ROWS, int K, int F>
__device__ __forceinline__ void row_run_step(float (&v)[K][F], bool add) {
    float tmp[K][F];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int f = 0; f < F; ++f) tmp[k][f] = ROWS ? row_down<OFF>(v[k][f]) : __shfl_down(v[k][f], OFF, kWave);
    if (add) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int f = 0; f < F; ++f) v[k][f] += tmp[k][f];
    }
}

// MODE 0: the direct levels only (no table, full occupancy); 1: the
// aggregated levels only; 2: both, the aggregated levels first.
template <int D, int F, int MODE>
__global__ __launch_bounds__(256, MODE != 0 ? 4 : 1) void k_grid_backward_groups(NectGridDesc g, const NectRay *__restrict__ rays,
                                                              const int32_t *__restrict__ tile_offsets,
                                                              const int32_t *__restrict__ order, int num_rays,
                                                              int n_passes, float t,
                                                              DoutView dv, float *__restrict__ grad, AggPlan plan) {
    constexpr int FH = F / 2;
    constexpr bool AGG = MODE != 0;
    __shared__ uint32_t keys[AGG ? kAggSlots : 1];
    __shared__ float vals[AGG ? kAggSlots * 16 : 1];
    __shared__ NectRay ray_of[kUnitRays];
    __shared__ int32_t slot0[kUnitRays];
    const int grp = (int)(blockIdx.x / (unsigned)n_passes);
    const int pass = (int)(blockIdx.x % (unsigned)n_passes);
    const int j = pass >> 1;
    const int first = kTile * j + kUnitSamples * (pass & 1);  // first sample index of the unit's pass
    // rays of the unit: descriptor and slot of their pass (-1: no sample there)
    int32_t my_ray = -1;
    if (threadIdx.x < kUnitRays) {
        const int i = grp * kUnitRays + threadIdx.x;
        int32_t s0 = -1;
        if (i < num_rays) {
            const int r = order ? order[i] : i;
            const int32_t a = tile_offsets[r];
            const NectRay d = rays[r];
            if (j < tile_offsets[r + 1] - a && first < d.count) {
                s0 = (a + j) * kTile + kUnitSamples * (pass & 1);
                my_ray = r;
                ray_of[threadIdx.x] = d;
            }
        }
        slot0[threadIdx.x] = s0;
    }
    if (AGG) {
        for (int i = threadIdx.x; i < kAggSlots; i += blockDim.x) keys[i] = kEmptyKey;
        for (int i = threadIdx.x; i < kAggSlots * 16; i += blockDim.x) vals[i] = 0.0f;
    }
    if (__syncthreads_or(my_ray >= 0) == 0) return;
    const float gsc = dv.scale ? dv.scale[0] : 1.0f;
    // dword w of the gradient lies in 64-byte block (w + lead) >> 4
    const int64_t lead = (int64_t)(((uintptr_t)grad & 63) >> 2);
    const int64_t n_words = (int64_t)g.offset[g.n_levels] * F;
    // adds v to dword w: in the LDS block table when w's block has (or finds)
    // a slot within kAggProbes probes, else directly
    auto agg_add = [&](int64_t w, float v) {
        const uint32_t key = (uint32_t)((w + lead) >> 4);
        uint32_t h = (key * 2654435761u) >> (32 - kAggLog2Slots);
#pragma unroll 1
        for (int probe = 0; probe < kAggProbes; ++probe) {
            uint32_t k = keys[h];
            if (k == kEmptyKey) k = atomicCAS(&keys[h], kEmptyKey, key);
            if (k == kEmptyKey || k == key) {
                atomicAdd(&vals[h * 16 + (uint32_t)((w + lead) & 15)], v);
                return;
            }
            h = (h + 1) & (kAggSlots - 1);
        }
        unsafeAtomicAdd(grad + w, v);
    };
    // one request per touched block (lanes 16b..16b+15 take a slot's dwords);
    // leaves the table empty
    auto agg_flush = [&]() {
        const int word = threadIdx.x & 15;
        for (int sl = threadIdx.x >> 4; sl < kAggSlots; sl += blockDim.x >> 4) {
            const uint32_t key = keys[sl];
            if (key != kEmptyKey) {
                const int64_t w = (int64_t)key * 16 + word - lead;
                const float v = vals[sl * 16 + word];
                if (v != 0.0f && w >= 0 && w < n_words) unsafeAtomicAdd(grad + w, v);
                vals[sl * 16 + word] = 0.0f;
            }
            // the 16 lanes of the slot read its key in the instruction above
            if (word == 0) keys[sl] = kEmptyKey;
        }
    };
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int m = lane >> 2;
    const int cx = (lane >> 1) & 1;
    const int fq = lane & 1;
    const int bit = 4 * m + 3;
    // sample m of ray i (samples lo..hi-1 of the pass count, the others are
    // missing): position and liveness
    auto sample_of = [&](int i, int lo, int hi, float (&p)[D]) {
        float q[3];
        const bool live = ray_sample(ray_of[i], first + m, q) && m >= lo && m < hi;
        place_sample<D>(plan.cmap, q, t, p);
        return live;
    };
    // quad-layout scatter of one level, straight to memory: lanes 4m.. hold
    // sample m of one ray (runs: consecutive samples in one cell are summed
    // first) or one sample of 16 rays (no runs)
    auto scatter_level = [&](const LevelInfo &lv, const float (&p)[D], bool live, bool prev_live, const float *gv,
                             bool runs) {
        uint32_t c[D];
        float f[D];
#pragma unroll
        for (int d = 0; d < D; ++d) grid_pos(lv.scale, p[d], c[d], f[d]);
        bool head = !runs || m == 0 || live != prev_live;
        if (runs) {
#pragma unroll
            for (int d = 0; d < D; ++d) head |= (uint32_t)__shfl_up((int)c[d], 4, kWave) != c[d];
        }
        const uint64_t heads = __ballot(head);
        const bool merge = heads != ~0ull;
        int end = 15;
        if (merge) {
            const uint64_t after = bit == kWave - 1 ? 0ull : heads & (~0ull << (bit + 1));
            end = after ? (__builtin_ctzll(after) >> 2) - 1 : 15;
        }
#pragma unroll
        for (int qq = 0; qq < (1 << (D - 1)); ++qq) {
            uint32_t idx;
            float w;
            grid_corner<D>(lv.res, lv.size, c, f, cx | (qq << 1), idx, w);
            float v[FH];
#pragma unroll
            for (int jj = 0; jj < FH; ++jj) v[jj] = w * gv[jj];
            if (merge) {
#pragma unroll
                for (int jj = 0; jj < FH; ++jj) v[jj] = run_sum_quad(v[jj], m, end);
            }
            add_corner_pair<F>(grad + (int64_t)lv.offset * F + fq * FH, idx, v, head && live, lane);
        }
    };
    bool direct = false;
    for (int l = 0; l < g.n_levels; ++l) direct |= plan.rounds[l] <= 0;
    if (MODE != 0) {
        // Aggregated levels, one lane per sample: lane 16 rr + ma holds sample
        // ma of ray 4 it + rr of the wave (it = 0..3); the LDS table merges
        // what the quad layout's x-pairs and run sums merge for the HBM
        // requests, so each lane computes its position once and handles all
        // 2^D corners, and only runs along a ray are summed across lanes (they
        // cut same-address LDS atomics at the coarse levels).
        const int ma = lane & 15;
        const int rr = lane >> 4;
        constexpr bool kRows = F == 4 || D == 4;
        // slot of table block `key`, inserting it when absent; -1 when no slot
        // is found within kAggProbes probes
        auto find_slot = [&](uint32_t key) -> int {
            uint32_t h = (key * 2654435761u) >> (32 - kAggLog2Slots);
#pragma unroll 1
            for (int probe = 0; probe < kAggProbes; ++probe) {
                uint32_t k = keys[h];
                if (k == kEmptyKey) k = atomicCAS(&keys[h], kEmptyKey, key);
                if (k == kEmptyKey || k == key) return (int)h;
                h = (h + 1) & (kAggSlots - 1);
            }
            return -1;
        };
        // adds v[0..F) to the F dwords from w0 (one table entry), using `slot`
        // when it holds their block
        auto add_entry = [&](int64_t w0, const float *v, uint32_t key, int slot) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                if (v[f] == 0.0f) continue;
                const int64_t w = w0 + f;
                if (slot >= 0 && (uint32_t)((w + lead) >> 4) == key)
                    atomicAdd(&vals[slot * 16 + (int)((w + lead) & 15)], v[f]);
                else
                    agg_add(w, v[f]);
            }
        };
        for (int l = 0; l < g.n_levels; ++l) {
            const int R = plan.rounds[l];
            if (R <= 0) continue;
            const LevelInfo lv = level_info(g, l);
            const int span = kUnitSamples / R;
            for (int h = 0; h < kUnitSamples; h += span) {
#pragma unroll 1
                for (int it = 0; it < 4; ++it) {
                    const int i = wave * kRaysPerWave + 4 * it + rr;
                    const int32_t s0 = slot0[i];
                    float q[3] = {0.0f, 0.0f, 0.0f};
                    const bool live = s0 >= 0 && ray_sample(ray_of[i], first + ma, q) && ma >= h && ma < h + span;
                    if (__ballot(live) == 0) continue;
                    float pa[D];
                    place_sample<D>(plan.cmap, q, t, pa);
                    float gv[F];
                    const int64_t go = (int64_t)(live ? s0 + ma : 0) * dv.stride + l * F;
#pragma unroll
                    for (int f = 0; f < F; ++f) gv[f] = live ? dout_at(dv, gsc, go + f) : 0.0f;
                    uint32_t c[D];
                    float fr[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) grid_pos(lv.scale, pa[d], c[d], fr[d]);
                    // lane 16 rr (ma = 0) is a head whatever its row shift reads
                    auto prev = [&](int x) { return kRows ? row_up<1>(x) : __shfl_up(x, 1, kWave); };
                    const bool prev_live = prev((int)live) != 0;
                    bool head = ma == 0 || live != prev_live;
#pragma unroll
                    for (int d = 0; d < D; ++d) head |= (uint32_t)prev((int)c[d]) != c[d];
                    const uint64_t heads = __ballot(head);
                    const bool merge = heads != ~0ull;
                    int end = kWave - 1;
                    if (merge) {
                        const uint64_t after = lane == kWave - 1 ? 0ull : heads & (~0ull << (lane + 1));
                        end = after ? __builtin_ctzll(after) - 1 : kWave - 1;
                    }
                    // the 2^D corners in chunks of 8: indices and values, run sums,
                    // then the chunk's table lookups issued together (one LDS
                    // round trip when the blocks are already in the table)
                    constexpr int K = 8;
#pragma unroll
                    for (int k0 = 0; k0 < (1 << D); k0 += K) {
                        uint32_t ix[K];
                        float v[K][F];
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            float w;
                            grid_corner<D>(lv.res, lv.size, c, fr, k0 + k, ix[k], w);
#pragma unroll
                            for (int f = 0; f < F; ++f) v[k][f] = w * gv[f];
                        }
                        if (merge) {
                            // a run never leaves its ray's row of 16 lanes (lane
                            // 16 rr is always a head), so row shifts serve
                            row_run_step<1, kRows>(v, lane + 1 <= end);
                            row_run_step<2, kRows>(v, lane + 2 <= end);
                            row_run_step<4, kRows>(v, lane + 4 <= end);
                            row_run_step<8, kRows>(v, lane + 8 <= end);
                        }
                        if (head && live) {
                            int64_t a[K];
                            uint32_t key[K], hs[K], got[K];
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                a[k] = ((int64_t)lv.offset + ix[k]) * F;
                                key[k] = (uint32_t)((a[k] + lead) >> 4);
                                hs[k] = (key[k] * 2654435761u) >> (32 - kAggLog2Slots);
                            }
#pragma unroll
                            for (int k = 0; k < K; ++k) got[k] = keys[hs[k]];
                            int slot[K];
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                if (got[k] == key[k]) slot[k] = (int)hs[k];
                                else if ((k & 1) && key[k] == key[k - 1]) slot[k] = slot[k - 1];
                                else slot[k] = find_slot(key[k]);
                            }
#pragma unroll
                            for (int k = 0; k < K; ++k) add_entry(a[k], v[k], key[k], slot[k]);
                        }
                    }
                }
                lds_barrier();
                agg_flush();
                lds_barrier();
            }
        }
    }
    if (MODE == 1 || !direct) return;
    bool by_ray = false, across = false;
    for (int l = 0; l < g.n_levels; ++l) {
        by_ray |= plan.rounds[l] == 0;
        across |= plan.rounds[l] < 0;
    }
    if (by_ray) {
        for (int k = 0; k < kRaysPerWave; ++k) {
            const int i = wave * kRaysPerWave + k;
            const int32_t s0 = slot0[i];
            if (s0 < 0) continue;
            float p[D];
            const bool live = sample_of(i, 0, kUnitSamples, p);
            const bool prev_live = __shfl_up((int)live, 4, kWave) != 0;
            const int64_t go = (int64_t)(live ? s0 + m : 0) * dv.stride + fq * FH;
            for (int c0 = 0; c0 < g.n_levels; c0 += kLevelChunk) {
                float gv[kLevelChunk][FH];
#pragma unroll
                for (int u = 0; u < kLevelChunk; ++u) {
                    const int l = c0 + u;
                    const bool want = live && l < g.n_levels && plan.rounds[l < g.n_levels ? l : 0] == 0;
#pragma unroll
                    for (int jj = 0; jj < FH; ++jj) gv[u][jj] = want ? dout_at(dv, gsc, go + l * F + jj) : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < kLevelChunk; ++u) {
                    const int l = c0 + u;
                    if (l >= g.n_levels) break;
                    if (plan.rounds[l] != 0) continue;
                    scatter_level(level_info(g, l), p, live, prev_live, gv[u], true);
                }
            }
        }
    }
    if (across) {
        // lanes 4m..4m+3: ray m of the wave (with the patch order, the 16 rays
        // of a full patch quarter: 4x4 neighbouring pixels), one sample index at
        // a time, so an instruction's x-neighbour entries of neighbouring rays
        // share 64-byte blocks at the middle levels
        const int i = wave * kRaysPerWave + m;
        const int32_t s0 = slot0[i];
        for (int k = 0; k < kUnitSamples; ++k) {
            float q[3] = {0.0f, 0.0f, 0.0f};
            const bool live = s0 >= 0 && ray_sample(ray_of[i], first + k, q);
            if (__ballot(live) == 0) continue;
            float p[D];
            place_sample<D>(plan.cmap, q, t, p);
            const int64_t go = (int64_t)(live ? s0 + k : 0) * dv.stride + fq * FH;
            for (int c0 = 0; c0 < g.n_levels; c0 += kLevelChunk) {
                float gv[kLevelChunk][FH];
#pragma unroll
                for (int u = 0; u < kLevelChunk; ++u) {
                    const int l = c0 + u;
                    const bool want = live && l < g.n_levels && plan.rounds[l < g.n_levels ? l : 0] < 0;
#pragma unroll
                    for (int jj = 0; jj < FH; ++jj) gv[u][jj] = want ? dout_at(dv, gsc, go + l * F + jj) : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < kLevelChunk; ++u) {
                    const int l = c0 + u;
                    if (l >= g.n_levels) break;
                    if (plan.rounds[l] >= 0) continue;
                    scatter_level(level_info(g, l), p, live, false, gv[u], false);
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

static int32_t launch_backward_groups(const NectGridDesc &g, const NectRay *rays, const int32_t *tile_offsets,
                                      int num_rays, const int32_t *order, int n_passes, float t, const DoutView &dv,
                                      float *grad, const AggPlan &plan, bool split, hipStream_t st) {
    const int64_t groups = (num_rays + kUnitRays - 1) / kUnitRays;
    NECT_REQUIRE(groups * n_passes < ((int64_t)1 << 31), "too many ray groups");
    bool any_agg = false, any_direct = false;
    for (int l = 0; l < g.n_levels; ++l) {
        any_agg |= plan.rounds[l] > 0;
        any_direct |= plan.rounds[l] <= 0;
    }
    const dim3 grid((unsigned)(groups * n_passes)), block(256);
#define NECT_BWDG(D, F, M)                                                                                        \
    hipLaunchKernelGGL((k_grid_backward_groups<D, F, M>), grid, block, 0, st, g, rays, tile_offsets, order,        \
                       num_rays, n_passes, t, dv, grad, plan)
#define NECT_BWDG_DF(M)                                \
    do {                                               \
        if (g.n_dims == 3) {                           \
            if (g.n_features == 2) NECT_BWDG(3, 2, M); \
            else NECT_BWDG(3, 4, M);                   \
        } else {                                       \
            if (g.n_features == 2) NECT_BWDG(4, 2, M); \
            else NECT_BWDG(4, 4, M);                   \
        }                                              \
    } while (0)
    if (!any_agg) {
        NECT_BWDG_DF(0);
    } else if (!split || !any_direct) {
        NECT_BWDG_DF(2);
    } else {
        NECT_BWDG_DF(1);
        if (check_launch("k_grid_backward_groups")) return 1;
        NECT_BWDG_DF(0);
    }
#undef NECT_BWDG_DF
#undef NECT_BWDG
    return check_launch("k_grid_backward_groups");
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

static int32_t backward_ray_groups(const NectGridDesc *desc, const NectRay *rays, const int32_t *tile_offsets,
                                   int32_t num_rays, const int32_t *order, int32_t max_points_per_ray, float t,
                                   const DoutView &dv, float *grad, const int32_t *agg_rounds, int32_t split,
                                   uint32_t cmap, void *stream);

extern "C" int32_t nect_grid_backward_ray_groups(const NectGridDesc *desc, const NectRay *rays,
                                                 const int32_t *tile_offsets, int32_t num_rays, const int32_t *order,
                                                 int32_t max_points_per_ray, float t, const float *dout,
                                                 int64_t dout_stride, float *grad, const int32_t *agg_rounds,
                                                 int32_t split, void *stream) {
    clear_error();
    const DoutView dv = {dout, dout_stride, NECT_F32, nullptr};
    return backward_ray_groups(desc, rays, tile_offsets, num_rays, order, max_points_per_ray, t, dv, grad,
                               agg_rounds, split, 0u, stream);
}

extern "C" int32_t nect_grid_backward_ray_groups_cols(const NectGridDesc *desc, const int32_t *src_cols,
                                                      const NectRay *rays, const int32_t *tile_offsets,
                                                      int32_t num_rays, const int32_t *order,
                                                      int32_t max_points_per_ray, float t, int32_t dout_dtype,
                                                      const void *dout, int64_t dout_stride, const float *dout_scale,
                                                      float *grad, const int32_t *agg_rounds, int32_t split,
                                                      void *stream) {
    clear_error();
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(src_cols != nullptr, "src_cols is NULL");
    NECT_REQUIRE(dout_dtype == NECT_F32 || dout_dtype == NECT_F16 || dout_dtype == NECT_BF16,
                 "dout_dtype must be NECT_F32, NECT_F16 or NECT_BF16");
    uint32_t cmap = 0;
    for (int d = 0; d < desc->n_dims; ++d) {
        NECT_REQUIRE(src_cols[d] >= 0 && src_cols[d] <= 3, "src_cols entries must be 0 (z), 1 (y), 2 (x) or 3 (t)");
        cmap |= ((uint32_t)src_cols[d] ^ (uint32_t)d) << (2 * d);
    }
    const DoutView dv = {dout, dout_stride, dout_dtype, dout_scale};
    return backward_ray_groups(desc, rays, tile_offsets, num_rays, order, max_points_per_ray, t, dv, grad,
                               agg_rounds, split, cmap, stream);
}

static int32_t backward_ray_groups(const NectGridDesc *desc, const NectRay *rays, const int32_t *tile_offsets,
                                   int32_t num_rays, const int32_t *order, int32_t max_points_per_ray, float t,
                                   const DoutView &dv, float *grad, const int32_t *agg_rounds, int32_t split,
                                   uint32_t cmap, void *stream) {
    if (check_grid(desc)) return 1;
    NECT_REQUIRE(num_rays >= 0, "num_rays must be >= 0");
    if (num_rays == 0) return 0;
    NECT_REQUIRE(rays && tile_offsets && dv.p && grad, "NULL pointer");
    NECT_REQUIRE(max_points_per_ray >= 1, "max_points_per_ray must be >= 1");
    NECT_REQUIRE(dv.stride >= (int64_t)desc->n_levels * desc->n_features, "dout_stride smaller than L*F");
    AggPlan plan = {};
    plan.cmap = cmap;
    for (int l = 0; l < desc->n_levels && agg_rounds; ++l) {
        const int32_t r = agg_rounds[l];
        NECT_REQUIRE(r == -1 || r == 0 || r == 1 || r == 2 || r == 4 || r == 8 || r == 16,
                     "agg_rounds entries must be -1, 0, 1, 2, 4, 8 or 16");
        plan.rounds[l] = (int8_t)r;
    }
    const int n_passes = 2 * ((max_points_per_ray + kTile - 1) / kTile);
    return launch_backward_groups(*desc, rays, tile_offsets, num_rays, order, n_passes, t, dv, grad, plan,
                                  split != 0, as_stream(stream));
}

static int32_t gather_parts(int32_t n_parts, const NectGridDesc *descs, const int64_t *table_offsets,
                            const int32_t *in_cols, const int32_t *out_cols, int64_t x_stride, int64_t out_stride,
                            GridParts &P) {
    NECT_REQUIRE(n_parts >= 1 && n_parts <= kMaxParts, "n_parts must be 1..NECT_MAX_GRID_PARTS");
    NECT_REQUIRE(descs && table_offsets && in_cols && out_cols, "NULL part array");
    P = GridParts{};
    for (int k = 0; k < n_parts; ++k) {
        if (check_grid(&descs[k])) return 1;
        NECT_REQUIRE(descs[k].n_dims == descs[0].n_dims && descs[k].n_features == descs[0].n_features,
                     "the parts of one launch must share n_dims and n_features");
        NECT_REQUIRE(table_offsets[k] >= 0, "negative table offset");
        NECT_REQUIRE(in_cols[k] >= 0 && in_cols[k] + descs[k].n_dims <= x_stride, "input columns outside x_stride");
        NECT_REQUIRE(out_cols[k] >= 0 && out_cols[k] + (int64_t)descs[k].n_levels * descs[k].n_features <= out_stride,
                     "output columns outside the row stride");
        P.g[k] = descs[k];
        P.table_off[k] = table_offsets[k];
        P.col0[k] = in_cols[k];
        P.out0[k] = out_cols[k];
    }
    return 0;
}

extern "C" int32_t nect_grid_encode_parts(int32_t n_parts, const NectGridDesc *descs, const int64_t *table_offsets,
                                          const int32_t *in_cols, const int32_t *out_cols, int32_t dtype,
                                          const void *tables, const float *x, int64_t n, int64_t x_stride, void *out,
                                          int64_t out_stride, void *stream) {
    clear_error();
    NECT_REQUIRE(dtype >= NECT_F32 && dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    GridParts P;
    if (gather_parts(n_parts, descs, table_offsets, in_cols, out_cols, x_stride, out_stride, P)) return 1;
    if (n == 0) return 0;
    NECT_REQUIRE(tables && x && out, "NULL pointer");
    const dim3 grid(blocks_for_slots(n), (unsigned)n_parts), block(256);
    const hipStream_t st = as_stream(stream);
    const int D = P.g[0].n_dims, F = P.g[0].n_features;
    // the staged write-back stores one level's F values at a time: rows, grid
    // offsets and the base must keep those aligned
    const int64_t esz = dtype == NECT_F32 ? 4 : 2, chunk = F * esz;
    bool stage = ((uintptr_t)out % chunk) == 0 && (out_stride * esz) % chunk == 0;
    int max_levels = 0;
    for (int k = 0; k < n_parts; ++k) {
        stage &= (P.out0[k] * esz) % chunk == 0;
        max_levels = std::max(max_levels, P.g[k].n_levels);
    }
    const size_t lds = stage ? (size_t)4 * kTile * max_levels * F * esz : 0;
#define NECT_ENCP(DD, FF, T)                                                                                      \
    do {                                                                                                          \
        if (stage)                                                                                                \
            hipLaunchKernelGGL((k_grid_encode_parts<DD, FF, T, true>), grid, block, lds, st, P, (const T *)tables, \
                               x, n, x_stride, (T *)out, out_stride);                                            \
        else                                                                                                      \
            hipLaunchKernelGGL((k_grid_encode_parts<DD, FF, T, false>), grid, block, 0, st, P, (const T *)tables,  \
                               x, n, x_stride, (T *)out, out_stride);                                            \
    } while (0)
#define NECT_ENCP_T(DD, FF)                                     \
    do {                                                        \
        if (dtype == NECT_F32) NECT_ENCP(DD, FF, float);        \
        else if (dtype == NECT_F16) NECT_ENCP(DD, FF, _Float16); \
        else NECT_ENCP(DD, FF, bf16_t);                         \
    } while (0)
    if (D == 3) {
        if (F == 2) NECT_ENCP_T(3, 2);
        else NECT_ENCP_T(3, 4);
    } else {
        if (F == 2) NECT_ENCP_T(4, 2);
        else NECT_ENCP_T(4, 4);
    }
#undef NECT_ENCP_T
#undef NECT_ENCP
    return check_launch("k_grid_encode_parts");
}

extern "C" int32_t nect_grid_backward_parts(int32_t n_parts, const NectGridDesc *descs, const int64_t *table_offsets,
                                            const int32_t *in_cols, const int32_t *out_cols, const float *x,
                                            int64_t n, int64_t x_stride, const float *dout, int64_t dout_stride,
                                            float *grad, void *stream) {
    clear_error();
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    GridParts P;
    if (gather_parts(n_parts, descs, table_offsets, in_cols, out_cols, x_stride, dout_stride, P)) return 1;
    if (n == 0) return 0;
    NECT_REQUIRE(x && dout && grad, "NULL pointer");
    const dim3 grid(blocks_for_slots(n), (unsigned)n_parts), block(256);
    const hipStream_t st = as_stream(stream);
    const int D = P.g[0].n_dims, F = P.g[0].n_features;
#define NECT_BWDP(DD, FF) \
    hipLaunchKernelGGL((k_grid_backward_parts<DD, FF>), grid, block, 0, st, P, x, n, x_stride, dout, dout_stride, grad)
    if (D == 3) {
        if (F == 2) NECT_BWDP(3, 2);
        else NECT_BWDP(3, 4);
    } else {
        if (F == 2) NECT_BWDP(4, 2);
        else NECT_BWDP(4, 4);
    }
#undef NECT_BWDP
    return check_launch("k_grid_backward_parts");
}
