/*
 * nect_amd — C ABI of the MI355X (gfx950) kernels behind NeCT's training and
 * query hot path.
 *
 * Conventions
 *   - Every pointer argument is a device pointer unless stated otherwise;
 *     descriptors (NectGeometry, NectGridDesc, NectMlpDesc) are host structs.
 *   - `stream` is a hipStream_t passed as void*; every call is asynchronous on
 *     that stream and never synchronises the device.
 *   - Return value 0 = success; nonzero = failure, with a message available
 *     from nect_last_error() (thread-local).  Arguments are validated on the
 *     host before anything is launched.
 *   - Inputs are borrowed; outputs are caller-allocated.  Scratch memory is a
 *     caller-provided workspace whose size a *_workspace_size() call reports.
 *
 * Coordinate conventions follow the reference: detector tuples are
 * (rows, cols), volume tuples (z, y, x), sample coordinates are written as
 * (z, y, x) in [0, 1].
 */
#ifndef NECT_AMD_H
#define NECT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NECT_ABI_VERSION 2

/* ---------------------------------------------------------------- errors */

/* Message of the last failing call on this thread ("" if none). */
const char *nect_last_error(void);
/* NECT_ABI_VERSION of the loaded library. */
int32_t nect_abi_version(void);

/* -------------------------------------------------------------- geometry */

enum NectGeometryKind {
    NECT_CONE_VOXEL = 0,            /* ConeGeometryVoxel        ct_sampling.cu:73-86  */
    NECT_PARALLEL_VOXEL = 1,        /* ParallelGeometryVoxel    ct_sampling.cu:88-98  */
    NECT_CONE_CYLINDRICAL = 2,      /* ConeGeometryCylindrical  ct_sampling.cu:100-115 */
    NECT_PARALLEL_CYLINDRICAL = 3   /* ParallelGeometryCylindrical ct_sampling.cu:54-66 */
};

/* One struct for the four reference geometry structs (ct_sampling.cu:7-115);
 * all fields keep the reference's float/int widths.  Fields a kind does not
 * use are ignored. */
typedef struct NectGeometry {
    int32_t kind;           /* NectGeometryKind */
    int32_t n_det[2];       /* nDetector (rows, cols), after binning */
    float d_det[2];         /* dDetector (mm) */
    float off_det[2];       /* offDetector (mm) */
    float rot_det[3];       /* rotDetector (roll, pitch, yaw) radians */
    float cor;              /* COR (mm) */
    float dso;              /* distance_source_to_origin (cone) */
    float dod;              /* distance_origin_to_detector (cone) */
    float s_voxel[3];       /* sVoxel (z, y, x) mm (voxel kinds) */
    float d_voxel[3];       /* dVoxel (z, y, x) mm (voxel kinds) */
    float radius;           /* object_radius (cylindrical kinds) */
    float height;           /* object_height (cylindrical kinds) */
    float remove_top;       /* remove_factor_top (cylindrical kinds) */
    float remove_bottom;    /* remove_factor_bottom (cylindrical kinds) */
} NectGeometry;

/*
 * Replaces ct_sampling.sample (nect/sampling/ct_sampling.cu:606-648, bound at
 * :756-807) followed by Projector's clamp_(0, 1) (nect/sampling/projector.py:127).
 *
 * For ray slot t < num_rays the detector pixel is
 * ray_index[starting_ray_index + t] (slots past n_ray_index or with a pixel
 * >= rows*cols are left empty, as the reference does).  jitter[num_rays] is
 * the per-ray stratification offset in [-0.5, 0.5) and detector_offset
 * [num_rays, 2] the sub-pixel (u, v) offset; the reference draws both with
 * torch::rand inside the call, here the caller passes them so that every
 * implementation sees the same numbers.
 *
 * Outputs: points [num_rays, S, 3] float (z, y, x), zero where no sample was
 * written; distances [num_rays] float (0 for a miss); counts [num_rays] int32
 * (samples written per ray; may be NULL).
 */
int32_t nect_sample(const NectGeometry *geometry, float angle_rad, const int64_t *ray_index,
                    int64_t n_ray_index, int32_t num_rays, int32_t starting_ray_index,
                    int32_t num_points_per_ray, float max_ray_distance_per_point,
                    int32_t uniform_ray_spacing, const float *jitter, const float *detector_offset,
                    float *points, float *distances, int32_t *counts, void *stream);

/*
 * Per-ray descriptor of the fused path: the same rays as nect_sample, but
 * instead of S materialised points it stores the normalised start point and
 * step (double, x/y/z) from which sample i is recomputed bit-identically as
 * float(start + i*step), clamped to [0, 1].
 */
typedef struct NectRay {
    double start[3];        /* x, y, z */
    double step[3];         /* x, y, z */
    int32_t count;          /* samples on the ray (0 for a miss) */
    float distance;         /* per-sample length (mm), 0 for a miss */
    int32_t pad[2];
} NectRay;

/* Same inputs as nect_sample; fills rays[num_rays]. No reference counterpart:
 * it is nect_sample with the [num_rays, S, 3] output replaced by a 64-byte
 * descriptor per ray. */
int32_t nect_ray_setup(const NectGeometry *geometry, float angle_rad, const int64_t *ray_index,
                       int64_t n_ray_index, int32_t num_rays, int32_t starting_ray_index,
                       int32_t num_points_per_ray, float max_ray_distance_per_point,
                       int32_t uniform_ray_spacing, const float *jitter, const float *detector_offset,
                       NectRay *rays, void *stream);

/*
 * Groups the samples of each ray into tiles of 32 consecutive samples:
 * tile_offsets[num_rays + 1] (exclusive scan of ceil(count/32)),
 * tile_ray[max_tiles] (owning ray of each tile), n_tiles[1] (total).  Sample
 * slot s = 32*tile + j of ray r holds sample index
 * 32*(tile - tile_offsets[r]) + j.  max_tiles must be >= num_rays*ceil(S/32).
 */
int32_t nect_ray_tiles(const NectRay *rays, int32_t num_rays, int64_t max_tiles, int32_t *tile_offsets,
                       int32_t *tile_ray, int32_t *n_tiles, void *stream);

/*
 * A permutation order[num_rays] of the batch's rays (those nect_ray_setup
 * describes from ray_index[starting_ray_index + t]) that groups them by 8x8
 * patch of the detector (row-major patches; rays without a pixel last).
 * Inside a patch the rays follow their pixel: the four 4x4 quarters in turn
 * (top-left, top-right, bottom-left, bottom-right), each row by row, so 16
 * consecutive rays of a full patch are a 4x4 block; a patch holding a pixel
 * twice keeps an arbitrary order.  workspace: nect_ray_order_workspace()
 * bytes of device memory, 8-byte aligned.  No reference counterpart (the reference's batches are in
 * random order, projector.py:93-130); it only schedules work.
 */
int64_t nect_ray_order_workspace(const NectGeometry *geometry);
int32_t nect_ray_order(const NectGeometry *geometry, const int64_t *ray_index, int64_t n_ray_index,
                       int32_t num_rays, int32_t starting_ray_index, int32_t *order, void *workspace,
                       int64_t workspace_bytes, void *stream);

/* Sample positions of the fused path in slot order: pos[max_tiles*32, 3]
 * (z, y, x), valid[max_tiles*32] (1 for a sample the reference trainer would
 * keep, base_trainer.py:523-528).  Parity/debug helper. */
int32_t nect_ray_positions(const NectRay *rays, const int32_t *tile_offsets, const int32_t *tile_ray,
                           const int32_t *n_tiles, int64_t max_tiles, float *pos, int32_t *valid,
                           void *stream);


/* ------------------------------------------------------------ hash grid */

#define NECT_MAX_LEVELS 32

enum NectDtype { NECT_F32 = 0, NECT_F16 = 1, NECT_BF16 = 2 };

/*
 * Multiresolution hash-grid layout (tiny-cuda-nn "HashGrid" encoding as
 * configured by nect/config.py:86-112 and consumed by nect/network/networks.py:
 * 28-65; the encoding itself lives in the un-vendored tiny-cuda-nn, so these
 * rules are this library's specification):
 *   scale[l]      = (float)(base_resolution * per_level_scale^l - 1)  (double, then rounded)
 *   resolution[l] = ceil(scale[l]) + 1
 *   level_size[l] = min(next_multiple_of_8(resolution^D), 2^log2_hashmap_size)
 *   offset[l]     = sum of level_size below l (entries); parameters are
 *                   offset[L] * n_features floats, level-major, entry-major,
 *                   features contiguous.
 * Sample x in [0,1]^D: p_d = fmaf(scale, x_d, 0.5) (one rounding), corner c_d = floor(p_d)
 * (+1), weight = product over d of (frac or 1-frac) in float, d = 0..D-1.
 * Index: dense sum c_d * res^d while the stride stays <= level_size,
 * otherwise the xor of c_d * prime_d (primes 1, 2654435761, 805459861,
 * 3674653429) in uint32; finally modulo level_size.
 */
typedef struct NectGridDesc {
    int32_t n_dims;                         /* 2..4 */
    int32_t n_levels;                       /* 1..NECT_MAX_LEVELS */
    int32_t n_features;                     /* per level: 1, 2, 4 or 8 */
    int32_t log2_hashmap_size;
    int32_t base_resolution;
    float per_level_scale;
    float scale[NECT_MAX_LEVELS];
    uint32_t resolution[NECT_MAX_LEVELS];
    uint32_t level_size[NECT_MAX_LEVELS];
    uint32_t offset[NECT_MAX_LEVELS + 1];
    int64_t n_params;                       /* offset[L] * n_features */
} NectGridDesc;

/* Host only: fills *desc (no device work). */
int32_t nect_grid_init(NectGridDesc *desc, int32_t n_dims, int32_t n_levels, int32_t n_features,
                       int32_t log2_hashmap_size, int32_t base_resolution, float per_level_scale);

/* Parity helper: corner indices (within the level) and weights,
 * index/weight [n, n_levels, 2^D].  x is [n, D] float with row stride x_stride. */
int32_t nect_grid_index(const NectGridDesc *desc, const float *x, int64_t n, int64_t x_stride, uint32_t *index,
                        float *weight, void *stream);

/* Forward encoding (replaces tcnn Encoding.forward for a HashGrid, reached
 * from nect/network/networks.py:52-65): out[n, L*F] (row stride out_stride
 * elements) in `dtype`, from table[n_params] in the same dtype.  The
 * interpolation accumulates in float. */
int32_t nect_grid_encode(const NectGridDesc *desc, int32_t dtype, const void *table, const float *x, int64_t n,
                         int64_t x_stride, void *out, int64_t out_stride, void *stream);

/* Backward (tcnn Encoding backward): grad[n_params] (float) +=
 * sum over samples and corners of weight * dout[n, L*F] (float, row stride
 * dout_stride), by float atomics. */
int32_t nect_grid_backward(const NectGridDesc *desc, const float *x, int64_t n, int64_t x_stride,
                           const float *dout, int64_t dout_stride, float *grad, void *stream);

#define NECT_MAX_GRID_PARTS 5

/*
 * The nested hash grids of a tiny-cuda-nn "Composite" encoding (QuadCubes and
 * HyperCubes, nect/network/networks.py:392-508) in one launch: part k encodes
 * input columns [in_cols[k], in_cols[k] + D) of x[n, x_stride] (float) with
 * grid descs[k] and its table at tables + table_offsets[k] (elements, dtype),
 * writing output columns [out_cols[k], out_cols[k] + L_k*F) of
 * out[n, out_stride] (dtype).  All parts share n_dims and n_features; at most
 * NECT_MAX_GRID_PARTS parts.  Identity parts hold no grid: the caller copies
 * their columns.  descs, table_offsets, in_cols, out_cols are host arrays.
 * Each part's features equal nect_grid_encode's for that grid bit for bit.
 */
int32_t nect_grid_encode_parts(int32_t n_parts, const NectGridDesc *descs, const int64_t *table_offsets,
                               const int32_t *in_cols, const int32_t *out_cols, int32_t dtype, const void *tables,
                               const float *x, int64_t n, int64_t x_stride, void *out, int64_t out_stride,
                               void *stream);

/* Backward of nect_grid_encode_parts: grad (float, laid out as the tables)
 * += the scatter of dout[n, dout_stride] (float) for every part, by float
 * atomics, as nect_grid_backward does for one grid. */
int32_t nect_grid_backward_parts(int32_t n_parts, const NectGridDesc *descs, const int64_t *table_offsets,
                                 const int32_t *in_cols, const int32_t *out_cols, const float *x, int64_t n,
                                 int64_t x_stride, const float *dout, int64_t dout_stride, float *grad,
                                 void *stream);

/* ------------------------------------------------------------------ MLP */

enum NectActivation {
    NECT_ACT_NONE = 0, NECT_ACT_RELU = 1, NECT_ACT_LEAKY_RELU = 2, NECT_ACT_EXPONENTIAL = 3, NECT_ACT_SINE = 4,
    NECT_ACT_SIGMOID = 5, NECT_ACT_SQUAREPLUS = 6, NECT_ACT_SOFTPLUS = 7, NECT_ACT_TANH = 8
};

#define NECT_MLP_MAX_HIDDEN 8

/*
 * Bias-free fully fused MLP (tiny-cuda-nn "FullyFusedMLP"/"CutlassMLP" as
 * configured by nect/config.py:125-132): n_hidden_layers hidden layers of
 * `width` (= 64) neurons.  Weights, row-major [out][in] per layer,
 * concatenated: W0 [width][n_input_padded], W1..W(H-1) [width][width],
 * Wout [n_output_padded][width]; paddings are multiples of 16.  Inputs beyond
 * n_input are taken as zero.  Matrix products run on MFMA in the weights'
 * dtype with float accumulation; activations are applied in float.
 * 1..NECT_MLP_MAX_HIDDEN hidden layers (the reference's presets use 2-6) and
 * padded inputs of 16..128; the float32 backward takes at most 3 hidden layers
 * and 64 inputs (its LDS image), inputs above 64 in the backward need a None,
 * ReLU or LeakyReLU hidden activation.
 */
typedef struct NectMlpDesc {
    int32_t n_input;
    int32_t n_input_padded;
    int32_t n_output;
    int32_t n_output_padded;
    int32_t width;
    int32_t n_hidden_layers;
    int32_t activation;          /* NectActivation of the hidden layers */
    int32_t output_activation;   /* NectActivation of the output */
    int64_t n_params;
} NectMlpDesc;

/* Host only. */
int32_t nect_mlp_init(NectMlpDesc *desc, int32_t n_input, int32_t n_output, int32_t width, int32_t n_hidden_layers,
                      int32_t activation, int32_t output_activation);

/* y[n, n_output] (float, row stride y_stride) = MLP(x[n, n_input]) with x and
 * weights in `dtype`; x row stride x_stride elements.
 * x_groups: NULL for that row layout; otherwise x is the group-planar feature
 * layout that nect_grid_encode_rays writes for this grid (grouped != 0) over
 * n rows, and x_stride is ignored.
 * n_tiles_dev, if not NULL, is a device int32 holding a count of 32-row tiles
 * that bounds n at run time. */
int32_t nect_mlp_forward(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x, int64_t n,
                         int64_t x_stride, const NectGridDesc *x_groups, const int32_t *n_tiles_dev, float *y,
                         int64_t y_stride, void *stream);

/* Bytes of workspace nect_mlp_backward needs for n rows. */
int64_t nect_mlp_backward_workspace(const NectMlpDesc *desc, int32_t dtype, int64_t n);

/* Backward of nect_mlp_forward for dy[n, n_output] (float), x as there:
 * dweights[n_params] (float) += grad_scale * dL/dW, and, if dx != NULL,
 * dx[n, n_input] (float, row stride dx_stride, row layout) = grad_scale * dL/dx.
 * dy is multiplied by the loss scale before the low-precision products and
 * the results by grad_scale / loss scale afterwards; the loss scale is
 * *loss_scale_dev (device float) when that pointer is not NULL, else
 * loss_scale. */
int32_t nect_mlp_backward(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x, int64_t n,
                          int64_t x_stride, const NectGridDesc *x_groups, const int32_t *n_tiles_dev,
                          const float *dy, int64_t dy_stride,
                          float loss_scale, const float *loss_scale_dev, float grad_scale, float *dweights,
                          float *dx, int64_t dx_stride, void *workspace, int64_t workspace_bytes, void *stream);

/*
 * nect_mlp_backward without a workspace argument and without allocating: the
 * per-workgroup partial weight gradients go to a 16 MiB device array of the
 * library (one copy per GPU, looked up at the first call on that GPU and held
 * until the process ends).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Networks with more than 4194304
 * parameters are rejected; one whose partials of the usual workgroup count do
 * not fit (n_params * workgroups > 4194304 floats) runs with fewer workgroups,
 * so its dweights may differ from nect_mlp_backward's in the order of float
 * additions.  Calls that overlap on one GPU (on different streams) share the
 * array and race: serialise them.
 */
int32_t nect_mlp_backward_noalloc(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x,
                                  int64_t n, int64_t x_stride, const NectGridDesc *x_groups,
                                  const int32_t *n_tiles_dev, const float *dy, int64_t dy_stride, float loss_scale,
                                  const float *loss_scale_dev, float grad_scale, float *dweights, float *dx,
                                  int64_t dx_stride, void *stream);

/* --------------------------------------------------- fused training path */

/* nect_grid_encode over the samples of the fused path: slot 32*tile + j of
 * the tile list (nect_ray_tiles) -> out[max_tiles*32, L*F] (row stride
 * out_stride elements).  Slots without a kept sample are written as zeros.
 * For a 4-D grid the fourth coordinate is `t` (the view's time).
 * grouped != 0 (n_levels a multiple of 8) selects the group-planar layout
 * instead: 8 planes of [max_tiles*32][V], V = n_levels/8 * F, plane g holding
 * the features of levels g, 15-g, 16+g, 31-g, ... (in that order) -- each
 * XCD then writes whole rows of its own plane; out_stride is ignored. */
int32_t nect_grid_encode_rays(const NectGridDesc *desc, int32_t dtype, const void *table, const NectRay *rays,
                              const int32_t *tile_offsets, const int32_t *tile_ray, const int32_t *n_tiles,
                              int64_t max_tiles, float t, void *out, int64_t out_stride, int32_t grouped,
                              void *stream);

/* nect_grid_backward over the samples of the fused path (dout float,
 * [max_tiles*32, L*F], row stride dout_stride). */
int32_t nect_grid_backward_rays(const NectGridDesc *desc, const NectRay *rays, const int32_t *tile_offsets,
                                const int32_t *tile_ray, const int32_t *n_tiles, int64_t max_tiles, float t,
                                const float *dout, int64_t dout_stride, float *grad, void *stream);

/*
 * The same gradient as nect_grid_backward_rays (up to the order of float
 * additions), computed over groups of 64 consecutive rays of `order` (a
 * permutation of 0..num_rays-1 from nect_ray_order, or NULL for batch order)
 * and one 16-sample half tile at a time.  agg_rounds (host array, n_levels
 * entries; NULL: all 0) plans each level: 0 scatters it directly, 16
 * samples of one ray per atomic instruction; -1 directly, one sample of 16
 * neighbouring rays per instruction (with the patch order: a 4x4 block of
 * pixels, whose x-neighbour entries share 64-byte blocks at the middle
 * levels); R in {1, 2, 4, 8, 16} sums the group's contributions per 64-byte block of the
 * table in LDS (512 blocks) and adds each touched block with one atomic
 * request, after every 16/R samples of the group's rays.  split != 0 runs the
 * aggregated and the direct levels as two launches (the second without the
 * LDS table), else as one.  Every plan is correct (blocks that find no LDS
 * slot go straight to memory); aggregation pays for levels whose cells are
 * not much smaller than the spacing of neighbouring rays, with R large enough
 * for the blocks of 64 x 16/R samples to fit.  max_points_per_ray is the
 * num_points_per_ray given to nect_ray_setup (no ray holds more samples).
 * No reference counterpart: a schedule of the tiny-cuda-nn HashGrid backward
 * scatter.
 */
int32_t nect_grid_backward_ray_groups(const NectGridDesc *desc, const NectRay *rays, const int32_t *tile_offsets,
                                      int32_t num_rays, const int32_t *order, int32_t max_points_per_ray, float t,
                                      const float *dout, int64_t dout_stride, float *grad,
                                      const int32_t *agg_rounds, int32_t split, void *stream);

/*
 * nect_grid_backward_ray_groups for a nested grid of a Composite encoding
 * whose inputs are a selection of the sample's (z, y, x, t): grid dimension d
 * reads coordinate src_cols[d] (0 z, 1 y, 2 x, 3 t; n_dims entries), e.g.
 * {1, 2, 3} for QuadCubes' (y, x, t) grid (nect/network/networks.py:437-449).
 * dout points at the grid's first feature column of dL/dfeatures, elements of
 * dout_dtype (NECT_F32, NECT_F16 or NECT_BF16); dout_stride is the row length
 * of all features (elements).  When dout_scale is not NULL every value read is
 * multiplied by *dout_scale (a device float): a 16-bit network backward hands
 * over its loss-scaled dL/dfeatures with the reciprocal of the power-of-two
 * scale, whose products are exact, so the scatter sees the values of the
 * unscaled float32 copy without that copy being written.  grad points at the
 * grid's table gradient.
 */
int32_t nect_grid_backward_ray_groups_cols(const NectGridDesc *desc, const int32_t *src_cols, const NectRay *rays,
                                           const int32_t *tile_offsets, int32_t num_rays, const int32_t *order,
                                           int32_t max_points_per_ray, float t, int32_t dout_dtype,
                                           const void *dout, int64_t dout_stride, const float *dout_scale,
                                           float *grad, const int32_t *agg_rounds, int32_t split, void *stream);

/*
 * Calibration of the ceiling the scatter works against: n_threads threads
 * (a multiple of 64) each add 1.0f per_thread times to table, a device array
 * of n_blocks 64-byte blocks (16 floats each); lanes_per_block (1, 2, 4, 8 or
 * 16) adjacent lanes of a wave share a hash-drawn block, each on its own
 * dword, so every wave instruction carries 64 / lanes_per_block requests.
 * Requests per second = n_threads * per_thread / lanes_per_block / time.
 * No reference counterpart: a measurement kernel.
 */
int32_t nect_calibrate_atomics(float *table, int64_t n_blocks, int32_t lanes_per_block, int32_t per_thread,
                               int64_t n_threads, uint32_t seed, void *stream);

/*
 * Weight gradient of a wide network layer (tiny-cuda-nn's backward GEMM for
 * networks the fused kernels do not cover, configured at nect/config.py:125-132):
 * dw[i * ld_dw + j] += scale * sum over k < n_samples of dz[k * ld_dz + i] * h[k * ld_h + j]
 * for i < n_out, j < n_in; dz and h in `dtype` (NECT_F16 or NECT_BF16), dw
 * float32, accumulated with float atomics (the addition order varies).
 */
int32_t nect_wgrad(int32_t dtype, const void *dz, int64_t ld_dz, const void *h, int64_t ld_h, int64_t n_samples,
                   int32_t n_out, int32_t n_in, float scale, float *dw, int64_t ld_dw, void *stream);

/*
 * One product of a wide network layer with its activation fused (the
 * tiny-cuda-nn network's forward layer, or its backward through a layer, for
 * the widths the fused kernels do not cover, nect/config.py:125-132):
 *     c[i * ldc + j] = epi( sum over l < k of a[i * lda + l] * bt[j * ldb + l] )
 * for i < n, j < m, in `dtype` (NECT_F16 or NECT_BF16) with float32
 * accumulation and one rounding.  epilogue 0: none; 1 (forward): the
 * piecewise-linear activation z > 0 ? z : slope * z (slope 0: ReLU, 0.01:
 * LeakyReLU, 1: none); 2 (backward): h[i * ldh + j] > 0 ? z : slope * z with h
 * the layer's saved activations (act'(z) from the sign of act(z)).  m is 128 or
 * 384, k a multiple of 16 up to 384 with (m + 64) (k + 8) 16-bit elements
 * fitting in 160 KB; rows of a and bt start 16-byte aligned, rows of c and h
 * 8-byte aligned.
 */
int32_t nect_gemm_act(int32_t dtype, const void *a, int64_t lda, const void *bt, int64_t ldb, int64_t n, int32_t k,
                      int32_t m, int32_t epilogue, float slope, const void *h, int64_t ldh, void *c, int64_t ldc,
                      void *stream);

/* NECT_LOSS_GIVEN: `target` holds dL/dy_pred per ray (an upstream gradient, e.g.
 * of the continuous-scanning mean over sub-angles); dmu = target * distance /
 * l_max and the loss is 0. */
enum NectLoss { NECT_LOSS_L1 = 0, NECT_LOSS_L2 = 1, NECT_LOSS_GIVEN = 2 };

/*
 * Line integral and loss of BaseTrainer.fit (nect/trainers/base_trainer.py:
 * 545-568, loss from nect/config.py:481-497): for ray r,
 *   y_pred[r] = (sum of mu over the kept samples of r) * (distance_r / l_max)
 * and loss = mean over the num_rays rays of |y_pred - target| (L1) or
 * (y_pred - target)^2 (L2).  Also writes dmu[slot] = dloss/dmu for every slot
 * of the ray's tiles (0 for slots without a kept sample).  mu and dmu are
 * [max_tiles*32] float in slot order.  The per-ray sum runs in a fixed order
 * (deterministic).  loss[1] (float) receives the mean loss; `partials` is a
 * scratch array of ceil(num_rays/4) floats.
 */
int32_t nect_integral_loss(const NectRay *rays, const int32_t *tile_offsets, int32_t num_rays, const float *mu,
                           const float *target, float l_max, int32_t loss, float *y_pred, float *dmu, float *partials,
                           float *loss_out, void *stream);

/*
 * One dense Adam step (torch.optim.Adam semantics, as NeCT builds it at
 * nect/config.py:400-414 and steps it at base_trainer.py:653) over n
 * parameters, step number `step` (>= 1):
 *   g = clamp(g, -clip_value, clip_value) if clip_value > 0 (the
 *       clip_grad_value_ of base_trainer.py:651-652);
 *   g += weight_decay * p; m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 *   p -= lr/(1-b1^step) * m / (sqrt(v)/sqrt(1-b2^step) + eps).
 * The gradient is read and then zeroed.  If `shadow` is not NULL the updated
 * parameters are also written there in shadow_dtype (the low-precision copy
 * the kernels read).
 */
int32_t nect_adam(float *params, float *grad, float *m, float *v, void *shadow, int32_t shadow_dtype, int64_t n,
                  float lr, float beta1, float beta2, float eps, float weight_decay, int32_t step, float clip_value,
                  void *stream);

/* dst[n] (dtype) = src[n] (float) — refreshes a low-precision copy. */
int32_t nect_cast(const float *src, void *dst, int32_t dtype, int64_t n, void *stream);

/*
 * Mixed-precision overflow handling of a 16-bit training step: the
 * precision="16-mixed" the reference trains with (nect/trainers/
 * base_trainer.py:106-117, Lightning's torch.amp.GradScaler: a step whose
 * gradient holds an Inf or NaN is skipped, the loss scale is halved, and it is
 * doubled after growth_interval finite steps in a row).  The state lives in
 * device memory (caller-allocated, 32 bytes, 4-byte aligned) so that no call
 * synchronises the host.  Its first field is the loss scale: pass its address
 * as loss_scale_dev of nect_mlp_backward.
 */
typedef struct NectAmpState {
    float scale;            /* loss scale */
    int32_t growth_tracker; /* finite steps since the scale last changed */
    int32_t step;           /* optimiser steps taken (skipped steps are not counted) */
    int32_t found_inf;      /* set by nect_grad_check, cleared by nect_adam_amp */
    int32_t skipped;        /* steps skipped so far */
    float step_size;        /* internal: lr / (1 - beta1^step) of the step in flight */
    float inv_sqrt_bc2;     /* internal: 1 / sqrt(1 - beta2^step) of the step in flight */
    int32_t pad;
} NectAmpState;

/* state = {init_scale, 0, 0, 0, 0, ...} (a device write on `stream`). */
int32_t nect_amp_init(NectAmpState *state, float init_scale, void *stream);

/* Sets state->found_inf if any of grad[n] is not finite (GradScaler's
 * unscale-and-check, the gradient here being unscaled already). */
int32_t nect_grad_check(const float *grad, int64_t n, NectAmpState *state, void *stream);

/*
 * nect_adam with the step decided on the device: if state->found_inf is set
 * the parameters, m, v and the shadow copy are left untouched (the gradient
 * is still zeroed), the scale is multiplied by backoff_factor and `skipped`
 * counts the step; otherwise Adam runs with step number state->step + 1 and
 * the scale is multiplied by growth_factor after growth_interval finite steps
 * in a row.  found_inf is cleared.
 */
int32_t nect_adam_amp(float *params, float *grad, float *m, float *v, void *shadow, int32_t shadow_dtype, int64_t n,
                      float lr, float beta1, float beta2, float eps, float weight_decay, NectAmpState *state,
                      float growth_factor, float backoff_factor, int32_t growth_interval, float clip_value,
                      void *stream);

/* -------------------------------------------------------------- queries */

/* How an export turns the network output mu into a voxel value (the three
 * operations of the reference, each rounded, out = ((mu / l_max) * span) + lo,
 * nect/static_export.py:138-140, nect/dynamic_export.py:535-537,
 * nect/trainers/base_trainer.py:442-444): */
enum NectRescale {
    NECT_RESCALE_DIVIDE = 0,          /* float32, true division: torch on CPU tensors (static export) */
    NECT_RESCALE_RECIPROCAL = 1,      /* float32, mu * fl(1/l_max): torch on GPU tensors (dynamic export) */
    NECT_RESCALE_RECIPROCAL_F16 = 2   /* as 1, each result rounded to half: a half tensor on the GPU
                                         (create_volume of a half-precision model) */
};

/*
 * One chunk of an export query (nect/static_export.py:106-141,
 * nect/dynamic_export.py:519-537, nect/trainers/base_trainer.py:402-450):
 * positions [v0, v0 + count) of the traversal of an n[0] x n[1] x n[2]
 * (z, y, x) grid.  With brick = 0 the traversal is row order, position
 * v = (iz*n[1] + iy)*n[2] + ix.  Otherwise brick = (bz << 8) | (by << 4) | bx
 * (each 0..4) names bricks of 2^bz x 2^by x 2^bx voxels that tile the grid
 * padded up to whole bricks: position v is voxel v mod 2^(bz+by+bx) (row order
 * inside the brick) of brick v >> (bz+by+bx) (row order over the bricks);
 * positions in the padding are skipped.  Neighbouring voxels are then
 * evaluated together, so their shared hash-table entries are read once from
 * HBM; the traversal length is the padded voxel count.  Every voxel is written
 * at out[(iz*n[1] + iy)*n[2] + ix] (out is the whole grid's output, in the
 * launch's out_dtype).  The coordinate along axis k is axis[k][i_k]
 * when axis[k] is not NULL (a device array of n[k] floats, e.g. made by
 * torch.linspace on the CPU as the static export makes them), otherwise
 * torch.linspace(a[k], b[k], n[k])[i_k] as torch computes it on the GPU
 * (float32: step = (b - a)/(n - 1); fma(step, i, a) below n/2,
 * fma(-step, n-1-i, b) from there on).  The model input is (z, y, x), plus t
 * for a 4-D grid.  mu is the network output, first rounded to half if
 * mu_dtype is NECT_F16 (the output of a half-precision network), then
 * rescaled as `rescale` says.
 */
typedef struct NectQueryJob {
    int32_t n[3];
    float a[3];
    float b[3];
    float t;
    float l_max;
    float span;
    float lo;
    int32_t mu_dtype;       /* NECT_F32 or NECT_F16 */
    int32_t rescale;        /* NectRescale */
    int32_t brick;          /* traversal: 0 = row order, else (bz << 8) | (by << 4) | bx */
    const float *axis[3];   /* optional explicit coordinates (z, y, x) */
    int64_t v0;
    int64_t count;
    void *out;
} NectQueryJob;

/*
 * Evaluates n_jobs query chunks in one launch (the scheduling quantum of the
 * continuous-batching export server): hash-grid encoding and MLP forward fused
 * per 32-voxel tile, features kept in LDS.  jobs: device array of n_jobs
 * NectQueryJob; tile_prefix: device int64[n_jobs + 1], exclusive scan of
 * ceil(count/32); n_tiles = tile_prefix[n_jobs].  weights/table: the model's
 * parameters in `dtype`; the MLP must have one output.  A voxel's value depends
 * only on its coordinates and the parameters (no atomics, no reductions), so
 * results do not depend on how requests are chunked or mixed.
 */
int32_t nect_query(const NectGridDesc *grid, const NectMlpDesc *mlp, int32_t dtype, const void *weights,
                   const void *table, const NectQueryJob *jobs, const int64_t *tile_prefix, int32_t n_jobs,
                   int64_t n_tiles, int32_t out_dtype, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NECT_AMD_H */
