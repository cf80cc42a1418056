// Ray sampler kernels and their C-ABI entry points.
//
// nect_sample       -> reference-layout points [B, S, 3] (ct_sampling.cu:578-648)
// nect_ray_setup    -> 64-byte ray descriptors for the fused path
// nect_ray_tiles    -> 32-sample tiles per ray (offsets + owner map)
// nect_ray_positions-> slot-ordered positions/validity (parity helper)
#include <cmath>

#include "sampler.hpp"

namespace nect {

GeomConst make_geom_const(const NectGeometry &g, float angle_rad) {
    GeomConst k;
    k.g = g;
    k.c = (float)std::cos((double)angle_rad);
    k.s = (float)std::sin((double)angle_rad);
    k.has_rot = (g.rot_det[0] != 0 || g.rot_det[1] != 0 || g.rot_det[2] != 0) ? 1 : 0;
    // Detector rotation entries, float products of float cos/sin
    // (ct_sampling.cu:202-216).
    const float alpha = g.rot_det[2], beta = g.rot_det[1], gamma = g.rot_det[0];
    const float ca = (float)std::cos(-(double)alpha), cb = (float)std::cos(-(double)beta);
    const float cg = (float)std::cos(-(double)gamma);
    const float sa = (float)std::sin(-(double)alpha), sb = (float)std::sin(-(double)beta);
    const float sg = (float)std::sin(-(double)gamma);
    k.R[0] = ca * sb * sg - sa * cg;
    k.R[1] = ca * sb * cg + sa * sg;
    k.R[2] = sa * sb * sg + ca * cg;
    k.R[3] = sa * sb * cg - ca * sg;
    k.R[4] = cb * sg;
    k.R[5] = cb * cg;
    k.par_offset = std::sqrt(g.s_voxel[2] * g.s_voxel[2] + g.s_voxel[1] * g.s_voxel[1]);
    k.hx = (g.s_voxel[2] - g.d_voxel[2] / 2) / 2;
    k.hy = (g.s_voxel[1] - g.d_voxel[1] / 2) / 2;
    k.hz = (g.s_voxel[0] - g.d_voxel[0] / 2) / 2;
    k.two_r = 2 * g.radius;
    k.z_lo = 0 + g.remove_top;
    k.z_hi = 1 - g.remove_bottom;
    k.n_pixels = g.n_det[0] * g.n_det[1];
    return k;
}

// Pixel of ray slot t, or -1 when the reference leaves the slot empty
// (ct_sampling.cu:130-148; the int64 -> int narrowing is the reference's).
__device__ __forceinline__ bool slot_pixel(const GeomConst &k, const int64_t *ray_index, int64_t n_ray_index,
                                           int t, int start, int &ray) {
    if ((int64_t)t + start >= n_ray_index) return false;
    ray = (int)ray_index[t + start];
    return ray < k.n_pixels;
}

constexpr int kSampleRaysPerBlock = 64;

// One block = 64 rays.  Wave 0 describes the rays into LDS; then all 256
// threads stream the block's [64, S, 3] output, consecutive threads writing
// consecutive floats.
__global__ __launch_bounds__(256) void k_sample_points(GeomConst k, const int64_t *__restrict__ ray_index,
                                                       int64_t n_ray_index, int num_rays, int start, int S,
                                                       float step_mm, int uniform,
                                                       const float *__restrict__ jitter,
                                                       const float *__restrict__ det_off,
                                                       float *__restrict__ points,
                                                       float *__restrict__ distances,
                                                       int32_t *__restrict__ counts) {
    __shared__ NectRay rays[kSampleRaysPerBlock];
    const int r0 = blockIdx.x * kSampleRaysPerBlock;
    const int nr = min(kSampleRaysPerBlock, num_rays - r0);
    if (threadIdx.x < kSampleRaysPerBlock) {
        NectRay d;
        d.count = 0;
        d.distance = 0.0f;
        d.start[0] = d.start[1] = d.start[2] = 0.0;
        d.step[0] = d.step[1] = d.step[2] = 0.0;
        const int t = r0 + threadIdx.x;
        int ray;
        if (threadIdx.x < nr && slot_pixel(k, ray_index, n_ray_index, t, start, ray)) {
            describe_ray(k, ray, det_off[2 * t], det_off[2 * t + 1], jitter[t], S, step_mm, uniform, d);
        }
        rays[threadIdx.x] = d;
        if (threadIdx.x < nr) {
            distances[t] = d.distance;
            if (counts) counts[t] = d.count;
        }
    }
    __syncthreads();
    const int row = 3 * S;
    const int64_t total = (int64_t)nr * row;
    float *out = points + (int64_t)r0 * row;
    int r = (int)(threadIdx.x / row);
    int e = (int)(threadIdx.x - r * row);
    for (int64_t g = threadIdx.x; g < total; g += blockDim.x) {
        const int i = e / 3;
        const int ax = e - 3 * i;
        const NectRay &d = rays[r];
        float v = 0.0f;
        if (i < d.count) v = clamp01_nan(coord(d.start[2 - ax], d.step[2 - ax], i));
        out[g] = v;
        e += blockDim.x;
        while (e >= row) {
            e -= row;
            ++r;
        }
    }
}

__global__ __launch_bounds__(256) void k_ray_setup(GeomConst k, const int64_t *__restrict__ ray_index,
                                                   int64_t n_ray_index, int num_rays, int start, int S,
                                                   float step_mm, int uniform, const float *__restrict__ jitter,
                                                   const float *__restrict__ det_off, NectRay *__restrict__ rays) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_rays) return;
    NectRay d;
    d.count = 0;
    d.distance = 0.0f;
    d.start[0] = d.start[1] = d.start[2] = 0.0;
    d.step[0] = d.step[1] = d.step[2] = 0.0;
    d.pad[0] = d.pad[1] = 0;
    int ray;
    if (slot_pixel(k, ray_index, n_ray_index, t, start, ray)) {
        describe_ray(k, ray, det_off[2 * t], det_off[2 * t + 1], jitter[t], S, step_mm, uniform, d);
    }
    rays[t] = d;
}

constexpr int kScanThreads = 1024;

// Exclusive scan of tiles-per-ray in three launches: per-block totals,
// a scan of the totals, then each block's own scan plus its offset.  The
// block totals live in tile_ray until k_tile_owner overwrites it.
__device__ __forceinline__ int32_t ray_tiles_of(const NectRay *rays, int num_rays, int r) {
    return r < num_rays ? (rays[r].count + kTile - 1) / kTile : 0;
}

// Inclusive scan of v over the block (kScanThreads threads).
__device__ __forceinline__ int32_t block_scan(int32_t v, int32_t *sums) {
    sums[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int32_t u = (threadIdx.x >= (unsigned)off) ? sums[threadIdx.x - off] : 0;
        __syncthreads();
        sums[threadIdx.x] += u;
        __syncthreads();
    }
    return sums[threadIdx.x];
}

__global__ __launch_bounds__(kScanThreads) void k_tile_block_sums(const NectRay *__restrict__ rays, int num_rays,
                                                                  int32_t *__restrict__ block_sums) {
    __shared__ int32_t sums[kScanThreads];
    const int r = blockIdx.x * kScanThreads + threadIdx.x;
    const int32_t inc = block_scan(ray_tiles_of(rays, num_rays, r), sums);
    if (threadIdx.x == kScanThreads - 1) block_sums[blockIdx.x] = inc;
}

// Exclusive scan of block_sums[0..n) in place (one block).
__global__ __launch_bounds__(kScanThreads) void k_tile_scan_sums(int32_t *__restrict__ block_sums, int n) {
    __shared__ int32_t sums[kScanThreads];
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    int32_t s = 0;
    for (int i = b; i < e; ++i) s += block_sums[i];
    int32_t run = block_scan(s, sums) - s;
    for (int i = b; i < e; ++i) {
        const int32_t c = block_sums[i];
        block_sums[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_tile_apply(const NectRay *__restrict__ rays, int num_rays,
                                                             const int32_t *__restrict__ block_offsets,
                                                             int32_t *__restrict__ tile_offsets,
                                                             int32_t *__restrict__ n_tiles) {
    __shared__ int32_t sums[kScanThreads];
    const int r = blockIdx.x * kScanThreads + threadIdx.x;
    const int32_t v = ray_tiles_of(rays, num_rays, r);
    const int32_t inc = block_scan(v, sums) + block_offsets[blockIdx.x];
    if (r < num_rays) tile_offsets[r] = inc - v;
    if (r == num_rays - 1) {
        tile_offsets[num_rays] = inc;
        n_tiles[0] = inc;
    }
}

__global__ __launch_bounds__(256) void k_tile_owner(const int32_t *__restrict__ tile_offsets, int num_rays,
                                                    int64_t max_tiles, int32_t *__restrict__ tile_ray) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_rays) return;
    const int32_t a = tile_offsets[r];
    const int32_t b = tile_offsets[r + 1];
    for (int32_t t = a; t < b && t < max_tiles; ++t) tile_ray[t] = r;
}

__global__ __launch_bounds__(256) void k_ray_positions(const NectRay *__restrict__ rays,
                                                       const int32_t *__restrict__ tile_offsets,
                                                       const int32_t *__restrict__ tile_ray,
                                                       const int32_t *__restrict__ n_tiles, int64_t max_tiles,
                                                       float *__restrict__ pos, int32_t *__restrict__ valid) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= max_tiles * kTile) return;
    const int64_t tile = slot / kTile;
    float p[3] = {0.0f, 0.0f, 0.0f};
    bool ok = false;
    if (tile < n_tiles[0]) {
        const int r = tile_ray[tile];
        const int i = (int)(tile - tile_offsets[r]) * kTile + (int)(slot % kTile);
        const NectRay d = rays[r];
        ok = ray_sample(d, i, p);
        if (i >= d.count) p[0] = p[1] = p[2] = 0.0f;
    }
    pos[3 * slot + 0] = p[0];
    pos[3 * slot + 1] = p[1];
    pos[3 * slot + 2] = p[2];
    valid[slot] = ok ? 1 : 0;
}

// ------------------------------------------------------------------ ray order
// A counting sort of a batch's rays by the 8x8 detector patch of their pixel
// (rays without a pixel last): groups of 64 consecutive rays of the order are
// then spatial neighbours, whose samples share table blocks at the coarse and
// middle levels of the grid (used by the aggregating scatter).  Inside a patch
// the rays follow their position: the four 4x4 quarters in turn, each row by
// row, so that 16 consecutive rays of a full patch are a 4x4 block of pixels.
// A patch that holds a pixel twice keeps an arbitrary order.
constexpr int kPatch = 8;

struct PatchKey {
    int cols, pcols, n_cells;  // detector columns, patch columns, patches
    // patch of ray t (n_cells: no pixel) and its position in the patch (0..63)
    __device__ __forceinline__ int of(const int64_t *ray_index, int64_t n_ray_index, int start, int t,
                                      int &pos) const {
        pos = 0;
        if ((int64_t)t + start >= n_ray_index) return n_cells;
        const int64_t pix = ray_index[t + start];
        if (pix < 0) return n_cells;
        const int64_t row = pix / cols, col = pix % cols;
        const int64_t key = (row / kPatch) * pcols + col / kPatch;
        if (key >= n_cells) return n_cells;
        const int r = (int)(row % kPatch), c = (int)(col % kPatch);
        pos = ((r >> 2) * 2 + (c >> 2)) * 16 + (r & 3) * 4 + (c & 3);
        return (int)key;
    }
};

// Workspace: offsets[n] (counts, then their exclusive scan), counts[n],
// cursors[n] (int32) and masks[n] (uint64, pixels present per patch).
struct OrderWs {
    int32_t *offsets, *counts, *cursors;
    unsigned long long *masks;
};

static OrderWs order_ws(void *ws, int n) {
    OrderWs w;
    w.masks = (unsigned long long *)ws;
    w.offsets = (int32_t *)(w.masks + n);
    w.counts = w.offsets + n;
    w.cursors = w.counts + n;
    return w;
}

__global__ __launch_bounds__(256) void k_order_zero(OrderWs w, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        w.offsets[i] = 0;
        w.cursors[i] = 0;
        w.masks[i] = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_order_count(PatchKey pk, const int64_t *__restrict__ ray_index,
                                                     int64_t n_ray_index, int num_rays, int start, OrderWs w) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_rays) return;
    int pos;
    const int key = pk.of(ray_index, n_ray_index, start, t, pos);
    atomicAdd(&w.offsets[key], 1);
    if (key < pk.n_cells) atomicOr(&w.masks[key], 1ull << pos);
}

// counts <- offsets; offsets <- exclusive scan (one block).
__global__ __launch_bounds__(kScanThreads) void k_order_scan(OrderWs w, int n) {
    __shared__ int32_t sums[kScanThreads];
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    int32_t s = 0;
    for (int i = b; i < e; ++i) s += w.offsets[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int32_t v = (threadIdx.x >= (unsigned)off) ? sums[threadIdx.x - off] : 0;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t run = sums[threadIdx.x] - s;
    for (int i = b; i < e; ++i) {
        const int32_t c = w.offsets[i];
        w.counts[i] = c;
        w.offsets[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(256) void k_order_place(PatchKey pk, const int64_t *__restrict__ ray_index,
                                                     int64_t n_ray_index, int num_rays, int start, OrderWs w,
                                                     int32_t *__restrict__ order) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_rays) return;
    int pos;
    const int key = pk.of(ray_index, n_ray_index, start, t, pos);
    int slot;
    if (key < pk.n_cells && __popcll(w.masks[key]) == w.counts[key]) {
        // every pixel of the patch at most once: rank by position
        slot = w.offsets[key] + __popcll(w.masks[key] & ((1ull << pos) - 1ull));
    } else {
        slot = w.offsets[key] + atomicAdd(&w.cursors[key], 1);
    }
    order[slot] = t;
}

static PatchKey patch_key(const NectGeometry &g) {
    PatchKey pk;
    pk.cols = g.n_det[1];
    pk.pcols = (g.n_det[1] + kPatch - 1) / kPatch;
    pk.n_cells = pk.pcols * ((g.n_det[0] + kPatch - 1) / kPatch);
    return pk;
}

static int32_t check_geometry(const NectGeometry *g) {
    NECT_REQUIRE(g != nullptr, "geometry is NULL");
    NECT_REQUIRE(g->kind >= NECT_CONE_VOXEL && g->kind <= NECT_PARALLEL_CYLINDRICAL, "unknown geometry kind");
    NECT_REQUIRE(g->n_det[0] > 0 && g->n_det[1] > 0, "nDetector must be positive");
    NECT_REQUIRE((int64_t)g->n_det[0] * g->n_det[1] < (int64_t)1 << 31, "nDetector rows*cols overflows int32");
    return 0;
}

static int32_t check_rays(const NectGeometry *g, const int64_t *ray_index, int64_t n_ray_index, int32_t num_rays,
                          int32_t start, int32_t S, const float *jitter, const float *det_off) {
    if (check_geometry(g)) return 1;
    NECT_REQUIRE(num_rays >= 0, "num_rays must be >= 0");
    NECT_REQUIRE(S >= 1, "num_points_per_ray must be >= 1");
    NECT_REQUIRE(start >= 0, "starting_ray_index must be >= 0");
    NECT_REQUIRE(n_ray_index >= 0, "random_ray_index length must be >= 0");
    NECT_REQUIRE(num_rays == 0 || (ray_index && jitter && det_off), "NULL input pointer");
    return 0;
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_sample(const NectGeometry *geometry, float angle_rad, const int64_t *ray_index,
                               int64_t n_ray_index, int32_t num_rays, int32_t starting_ray_index,
                               int32_t num_points_per_ray, float max_ray_distance_per_point,
                               int32_t uniform_ray_spacing, const float *jitter, const float *detector_offset,
                               float *points, float *distances, int32_t *counts, void *stream) {
    clear_error();
    if (check_rays(geometry, ray_index, n_ray_index, num_rays, starting_ray_index, num_points_per_ray, jitter,
                   detector_offset))
        return 1;
    NECT_REQUIRE((int64_t)num_rays * num_points_per_ray * 3 < ((int64_t)1 << 40), "output too large");
    if (num_rays == 0) return 0;
    NECT_REQUIRE(points && distances, "NULL output pointer");
    const GeomConst k = make_geom_const(*geometry, angle_rad);
    const int blocks = (num_rays + kSampleRaysPerBlock - 1) / kSampleRaysPerBlock;
    hipLaunchKernelGGL(k_sample_points, dim3(blocks), dim3(256), 0, as_stream(stream), k, ray_index, n_ray_index,
                       num_rays, starting_ray_index, num_points_per_ray, max_ray_distance_per_point,
                       uniform_ray_spacing, jitter, detector_offset, points, distances, counts);
    return check_launch("k_sample_points");
}

extern "C" int32_t nect_ray_setup(const NectGeometry *geometry, float angle_rad, const int64_t *ray_index,
                                  int64_t n_ray_index, int32_t num_rays, int32_t starting_ray_index,
                                  int32_t num_points_per_ray, float max_ray_distance_per_point,
                                  int32_t uniform_ray_spacing, const float *jitter, const float *detector_offset,
                                  NectRay *rays, void *stream) {
    clear_error();
    if (check_rays(geometry, ray_index, n_ray_index, num_rays, starting_ray_index, num_points_per_ray, jitter,
                   detector_offset))
        return 1;
    if (num_rays == 0) return 0;
    NECT_REQUIRE(rays, "NULL output pointer");
    const GeomConst k = make_geom_const(*geometry, angle_rad);
    hipLaunchKernelGGL(k_ray_setup, dim3((num_rays + 255) / 256), dim3(256), 0, as_stream(stream), k, ray_index,
                       n_ray_index, num_rays, starting_ray_index, num_points_per_ray, max_ray_distance_per_point,
                       uniform_ray_spacing, jitter, detector_offset, rays);
    return check_launch("k_ray_setup");
}

extern "C" int32_t nect_ray_tiles(const NectRay *rays, int32_t num_rays, int64_t max_tiles, int32_t *tile_offsets,
                                  int32_t *tile_ray, int32_t *n_tiles, void *stream) {
    clear_error();
    NECT_REQUIRE(num_rays >= 1, "num_rays must be >= 1");
    NECT_REQUIRE(rays && tile_offsets && tile_ray && n_tiles, "NULL pointer");
    NECT_REQUIRE(max_tiles >= 1 && max_tiles < ((int64_t)1 << 31) / kTile, "max_tiles out of range");
    const int nb = (num_rays + kScanThreads - 1) / kScanThreads;
    NECT_REQUIRE(max_tiles >= nb, "max_tiles smaller than the scan's block count");
    hipStream_t st = as_stream(stream);
    int32_t *block_sums = tile_ray;  // scratch until k_tile_owner writes the owners
    hipLaunchKernelGGL(k_tile_block_sums, dim3(nb), dim3(kScanThreads), 0, st, rays, num_rays, block_sums);
    if (check_launch("k_tile_block_sums")) return 1;
    hipLaunchKernelGGL(k_tile_scan_sums, dim3(1), dim3(kScanThreads), 0, st, block_sums, nb);
    if (check_launch("k_tile_scan_sums")) return 1;
    hipLaunchKernelGGL(k_tile_apply, dim3(nb), dim3(kScanThreads), 0, st, rays, num_rays, block_sums, tile_offsets,
                       n_tiles);
    if (check_launch("k_tile_apply")) return 1;
    hipLaunchKernelGGL(k_tile_owner, dim3((num_rays + 255) / 256), dim3(256), 0, st, tile_offsets, num_rays,
                       max_tiles, tile_ray);
    return check_launch("k_tile_owner");
}

extern "C" int32_t nect_ray_positions(const NectRay *rays, const int32_t *tile_offsets, const int32_t *tile_ray,
                                      const int32_t *n_tiles, int64_t max_tiles, float *pos, int32_t *valid,
                                      void *stream) {
    clear_error();
    NECT_REQUIRE(rays && tile_offsets && tile_ray && n_tiles && pos && valid, "NULL pointer");
    NECT_REQUIRE(max_tiles >= 1 && max_tiles < ((int64_t)1 << 31) / kTile, "max_tiles out of range");
    const int64_t slots = max_tiles * kTile;
    hipLaunchKernelGGL(k_ray_positions, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, as_stream(stream),
                       rays, tile_offsets, tile_ray, n_tiles, max_tiles, pos, valid);
    return check_launch("k_ray_positions");
}

extern "C" int64_t nect_ray_order_workspace(const NectGeometry *geometry) {
    clear_error();
    if (check_geometry(geometry)) return -1;
    return 20 * ((int64_t)patch_key(*geometry).n_cells + 1);
}

extern "C" int32_t nect_ray_order(const NectGeometry *geometry, const int64_t *ray_index, int64_t n_ray_index,
                                  int32_t num_rays, int32_t starting_ray_index, int32_t *order, void *workspace,
                                  int64_t workspace_bytes, void *stream) {
    clear_error();
    if (check_geometry(geometry)) return 1;
    NECT_REQUIRE(num_rays >= 0 && starting_ray_index >= 0 && n_ray_index >= 0, "negative size or start");
    if (num_rays == 0) return 0;
    NECT_REQUIRE(ray_index && order && workspace, "NULL pointer");
    NECT_REQUIRE(((uintptr_t)workspace & 7) == 0, "workspace must be 8-byte aligned");
    const PatchKey pk = patch_key(*geometry);
    const int n = pk.n_cells + 1;
    NECT_REQUIRE(workspace_bytes >= 20 * (int64_t)n, "workspace smaller than nect_ray_order_workspace()");
    hipStream_t st = as_stream(stream);
    const OrderWs w = order_ws(workspace, n);
    hipLaunchKernelGGL(k_order_zero, dim3((n + 255) / 256), dim3(256), 0, st, w, n);
    if (check_launch("k_order_zero")) return 1;
    const dim3 grid((num_rays + 255) / 256), block(256);
    hipLaunchKernelGGL(k_order_count, grid, block, 0, st, pk, ray_index, n_ray_index, num_rays, starting_ray_index,
                       w);
    if (check_launch("k_order_count")) return 1;
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(kScanThreads), 0, st, w, n);
    if (check_launch("k_order_scan")) return 1;
    hipLaunchKernelGGL(k_order_place, grid, block, 0, st, pk, ray_index, n_ray_index, num_rays, starting_ray_index,
                       w, order);
    return check_launch("k_order_place");
}
