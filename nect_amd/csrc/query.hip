// Fused export-query kernel: voxel coordinates generated in-kernel, hash-grid
// encoding into LDS (lane pairs gather the two x-neighbour corners, as the
// training encoder does, and add the halves in the same order, so features are
// bit-identical to nect_grid_encode's), MLP forward on MFMA, one value per voxel.
#include "mlp.hpp"

namespace nect {

// torch.linspace(a, b, n)[i] in float32 as torch computes it on the GPU.
__device__ __forceinline__ float linspace_at(float a, float b, int n, int64_t i) {
    if (n == 1) return a;
    const float step = (b - a) / (float)(n - 1);
    const int64_t half = n / 2;
    if (i < half) return fmaf(step, (float)i, a);
    return fmaf(-step, (float)(n - 1 - i), b);
}

__device__ __forceinline__ float query_coord(const NectQueryJob &J, int k, int64_t i) {
    return J.axis[k] ? J.axis[k][i] : linspace_at(J.a[k], J.b[k], J.n[k], i);
}

// Voxel value from the network output: optional rounding of mu to half, then
// ((mu / l_max) * span) + lo with the division and rounding torch applies.
__device__ __forceinline__ float rescale_value(const NectQueryJob &J, float mu) {
    if (J.mu_dtype == NECT_F16) mu = (float)(_Float16)mu;
    if (J.rescale == NECT_RESCALE_DIVIDE) return ((mu / J.l_max) * J.span) + J.lo;
    const float inv = 1.0f / J.l_max;
    if (J.rescale == NECT_RESCALE_RECIPROCAL) return ((mu * inv) * J.span) + J.lo;
    _Float16 h = (_Float16)mu;
    h = (_Float16)((float)h * inv);
    h = (_Float16)((float)h * J.span);
    h = (_Float16)((float)h + J.lo);
    return (float)h;
}

// Voxel (iz, iy, ix) at traversal position v of a job's grid (row order, or
// bricks of 2^bz x 2^by x 2^bx voxels over the grid padded to whole bricks);
// false for positions in the padding or past the grid.  The host keeps the
// brick count below 2^31, so the brick arithmetic runs in 32 bits.
__device__ __forceinline__ bool query_voxel(const NectQueryJob &J, int64_t v, int32_t &iz, int32_t &iy, int32_t &ix) {
    if (J.brick == 0) {
        const int64_t nyx = (int64_t)J.n[1] * J.n[2];
        const int64_t z = v / nyx;
        const int64_t r = v - z * nyx;
        const int64_t y = r / J.n[2];
        iz = (int32_t)min(z, (int64_t)J.n[0]);
        iy = (int32_t)y;
        ix = (int32_t)(r - y * J.n[2]);
        return z < J.n[0];
    }
    const uint32_t lz = min((J.brick >> 8) & 15, 4), ly = min((J.brick >> 4) & 15, 4), lx = min(J.brick & 15, 4);
    const uint32_t ib = lz + ly + lx;
    const uint32_t b = (uint32_t)(v >> ib);
    const uint32_t in = (uint32_t)v & ((1u << ib) - 1u);
    const uint32_t nbx = ((uint32_t)J.n[2] + (1u << lx) - 1u) >> lx;
    const uint32_t nby = ((uint32_t)J.n[1] + (1u << ly) - 1u) >> ly;
    const uint32_t plane = nby * nbx;
    const uint32_t bz = b / plane;
    const uint32_t r = b - bz * plane;
    const uint32_t by = r / nbx;
    const uint32_t bx = r - by * nbx;
    iz = (int32_t)((bz << lz) + (in >> (ly + lx)));
    iy = (int32_t)((by << ly) + ((in >> lx) & ((1u << ly) - 1u)));
    ix = (int32_t)((bx << lx) + (in & ((1u << lx) - 1u)));
    return iz < J.n[0] && iy < J.n[1] && ix < J.n[2];
}

// Workgroups are dealt to the XCDs in turn (workgroup b runs on XCD b mod 8).
// Renumbering them so that XCD k holds the k-th eighth of each round of tiles
// keeps neighbouring bricks -- which share hash-table entries -- in one L2.
__device__ __forceinline__ int64_t xcd_block(unsigned b, unsigned n) {
    if (n % kXcds != 0) return b;
    return (int64_t)(b % kXcds) * (n / kXcds) + b / kXcds;
}

// The gathers are latency-bound: a register budget of three waves per SIMD
// (the compiler fits it in ~125 VGPRs without spills, against ~190 unbounded)
// lets three 4-wave workgroups share a CU, the most their LDS allows.
template <typename T, int D, int F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_query(MlpArgs a, NectGridDesc g, const T *__restrict__ table,
                                               const NectQueryJob *__restrict__ jobs,
                                               const int64_t *__restrict__ prefix, int n_jobs, int64_t n_tiles,
                                               int out_f16, int x_stride) {
    extern __shared__ __align__(16) char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    stage_weights<T, false>(a, lds);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = lane & 31, h = lane >> 5;
    const int waves = blockDim.x >> 6;
    T *X = lds + a.L.total + wave * kTile * x_stride;
    const int64_t vb = xcd_block(blockIdx.x, gridDim.x);
    for (int64_t tile = vb * waves + wave; tile < n_tiles; tile += (int64_t)gridDim.x * waves) {
        __asm__ volatile("" ::: "memory");
        int j = 0;
        while (j + 1 < n_jobs && prefix[j + 1] <= tile) ++j;
        const NectQueryJob J = jobs[j];
        const int64_t first = (tile - prefix[j]) * kTile;
        // output: lane s (both halves) holds voxel s of the tile
        int32_t oz, oy, ox;
        const bool rv = first + s < J.count && query_voxel(J, J.v0 + first + s, oz, oy, ox);
        // encoding: lanes 2e and 2e+1 gather the x-bit 0 and 1 corners of voxel e of the tile
        {
            const int e = lane >> 1, cx = lane & 1;
            int32_t iz, iy, ix;
            const bool ev = first + e < J.count && query_voxel(J, J.v0 + first + e, iz, iy, ix);
            if (!ev) iz = iy = ix = 0;  // padding: any in-range voxel (never stored)
            float p[D];
            p[0] = query_coord(J, 0, iz);
            p[1] = query_coord(J, 1, iy);
            p[2] = query_coord(J, 2, ix);
            if (D == 4) p[D - 1] = J.t;
            T *xe = X + e * x_stride;
            for (int l = 0; l < g.n_levels; ++l) {
                const LevelInfo lv = level_info(g, l);
                uint32_t c[D];
                float f[D];
#pragma unroll
                for (int d = 0; d < D; ++d) grid_pos(lv.scale, p[d], c[d], f[d]);
                float half_sum[F];
                encode_half<D, F, T>(lv, c, f, table, cx, half_sum);
#pragma unroll
                for (int q = 0; q < F; ++q) {
                    const float other = __shfl_xor(half_sum[q], 1, kWave);
                    const float sum = cx == 0 ? half_sum[q] + other : other + half_sum[q];
                    if (cx == (l & 1)) xe[l * F + q] = from_f32<T>(sum);
                }
            }
        }
        T *xr = X + s * x_stride;
        wave_lds_sync();
        f32x16 zl[2];
        hidden_forward_last<T>(a, lds, xr, rv, true, s, h, zl);
        const float zo = output_pre_last<T>(a, lds, 0, h, zl);
        const float y = rescale_value(J, act_fwd_slow(a.d.output_activation, zo));
        if (h == 0 && rv) {
            const int64_t o = ((int64_t)oz * J.n[1] + oy) * J.n[2] + ox;
            if (out_f16) static_cast<_Float16 *>(J.out)[o] = (_Float16)y;
            else static_cast<float *>(J.out)[o] = y;
        }
        wave_lds_sync();
    }
}

int32_t check_grid(const NectGridDesc *g);

static int query_cus() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 256;
    return v;
}

// Persistent grid: as many workgroups as are resident at once (so the tiles in
// flight stay a contiguous stretch of the traversal), a multiple of the XCD
// count for xcd_block.
static unsigned query_blocks(const void *kernel, int threads, size_t lds_bytes, int64_t needed) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds_bytes) != hipSuccess ||
        per_cu <= 0)
        per_cu = 2;
    const int64_t resident = (int64_t)query_cus() * per_cu;
    int64_t b = std::min<int64_t>(needed, resident);
    b = (b + kXcds - 1) / kXcds * kXcds;
    return (unsigned)std::max<int64_t>(b, kXcds);
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_query(const NectGridDesc *grid, const NectMlpDesc *mlp, int32_t dtype, const void *weights,
                              const void *table, const NectQueryJob *jobs, const int64_t *tile_prefix, int32_t n_jobs,
                              int64_t n_tiles, int32_t out_dtype, void *stream) {
    clear_error();
    if (check_grid(grid)) return 1;
    NECT_REQUIRE(mlp != nullptr, "mlp descriptor is NULL");
    NECT_REQUIRE(dtype >= NECT_F32 && dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(out_dtype == NECT_F32 || out_dtype == NECT_F16, "out_dtype must be NECT_F32 or NECT_F16");
    NECT_REQUIRE(mlp->width == kWidth && mlp->n_hidden_layers >= 1 && mlp->n_hidden_layers <= NECT_MLP_MAX_HIDDEN,
                 "unsupported MLP shape");
    NECT_REQUIRE(mlp->n_output == 1, "queries need a one-output MLP");
    NECT_REQUIRE(mlp->n_input == grid->n_levels * grid->n_features, "MLP input must equal the grid's L*F");
    NECT_REQUIRE(mlp->n_input_padded <= 128, "padded input width must be <= 128");
    NECT_REQUIRE(n_jobs >= 1 && weights && table && jobs && tile_prefix, "bad arguments");
    if (n_tiles <= 0) return 0;
    const int esz = dtype == NECT_F32 ? 4 : 2;
    const int waves = dtype == NECT_F32 ? 2 : 4;
    MlpArgs a = {};
    a.d = *mlp;
    a.L = mlp_lds(*mlp, esz, false, waves);
    a.w = weights;
    const int x_stride = mlp->n_input_padded + 16 / esz;
    const size_t bytes = ((size_t)a.L.total + (size_t)waves * kTile * x_stride) * esz;
    const int64_t needed = (n_tiles + waves - 1) / waves;
    const hipStream_t st = as_stream(stream);
    const int out_f16 = out_dtype == NECT_F16;
    const int D = grid->n_dims, F = grid->n_features;
#define NECT_Q(T, DD, FF)                                                                                           \
    do {                                                                                                            \
        auto k = k_query<T, DD, FF>;                                                                                \
        if (bytes > 64 * 1024 &&                                                                                    \
            hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) { \
            set_error("k_query: cannot reserve LDS");                                                               \
            return 1;                                                                                               \
        }                                                                                                           \
        const unsigned blocks = query_blocks((const void *)k, 64 * waves, bytes, needed);                           \
        hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * waves), bytes, st, a, *grid, (const T *)table, jobs,          \
                           tile_prefix, n_jobs, n_tiles, out_f16, x_stride);                                        \
        return check_launch("k_query");                                                                             \
    } while (0)
#define NECT_Q_DF(T)                           \
    do {                                       \
        if (D == 3 && F == 2) NECT_Q(T, 3, 2); \
        if (D == 3 && F == 4) NECT_Q(T, 3, 4); \
        if (D == 4 && F == 2) NECT_Q(T, 4, 2); \
        NECT_Q(T, 4, 4);                       \
    } while (0)
    if (dtype == NECT_F32) NECT_Q_DF(float);
    if (dtype == NECT_F16) NECT_Q_DF(_Float16);
    NECT_Q_DF(bf16_t);
#undef NECT_Q_DF
#undef NECT_Q
}
