// Fully fused MLP kernels (forward, backward, weight-gradient reduction) and
// their C-ABI entry points.  Layout conventions are described in mlp.hpp.
#include <mutex>
#include <string>

#include "mlp_bwd.hpp"

namespace nect {

template <typename T>
__device__ __forceinline__ void mlp_forward_body(MlpArgs a) {
    extern __shared__ __align__(16) char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    stage_weights<T, false>(a, lds);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = lane & 31, h = lane >> 5;
    const int64_t n_tiles = tile_bound(a);
    const int waves = blockDim.x >> 6;
    for (int64_t tile = (int64_t)blockIdx.x * waves + wave; tile < n_tiles; tile += (int64_t)gridDim.x * waves) {
        // keeps the weight fragments in LDS instead of hoisting them into registers
        __asm__ volatile("" ::: "memory");
        const int64_t row = tile * kTile + s;
        const bool rv = row < a.n;
        f32x16 zl[2];
        hidden_forward_last<T>(a, lds, input_row<T>(a, rv ? row : 0), rv, a.x_vec != 0, s, h, zl);
        for (int o = 0; o < a.d.n_output; ++o) {
            const float zo = output_pre_last<T>(a, lds, o, h, zl);
            if (h == 0 && rv) a.y[row * a.y_stride + o] = act_fwd_slow(a.d.output_activation, zo);
        }
    }
}

// 16-bit operands: a register budget of three waves per SIMD (168 VGPRs, a few
// spills) beats two waves without spills -- 0.69 -> 0.55 ms on the 256^3 cone
// workload's 10.6 M samples.  Float32 would spill far more and keeps two.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_mlp_forward(MlpArgs a) {
    mlp_forward_body<T>(a);
}

__global__ __launch_bounds__(256) void k_mlp_forward_f32(MlpArgs a) { mlp_forward_body<float>(a); }

// dw[i] += scale * sum over slots of ws[slot][i]: 64 parameters per workgroup,
// each of its four waves sums every fourth slot, then the four partial sums
// are added in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ ws, int64_t slots, int64_t n_params,
                                                    float grad_scale, float loss_scale,
                                                    const float *__restrict__ loss_scale_dev, float *__restrict__ dw) {
    __shared__ float part[4][64];
    const int t = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + t;
    float acc = 0.0f;
    if (i < n_params) {
#pragma unroll 8
        for (int64_t k = g; k < slots; k += 4) acc += ws[k * n_params + i];
    }
    part[g][t] = acc;
    __syncthreads();
    if (g == 0 && i < n_params) {
        const float scale = grad_scale / (loss_scale_dev ? loss_scale_dev[0] : loss_scale);
        dw[i] += ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t])) * scale;
    }
}

// ------------------------------------------------------------------ host side

static int device_cus() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cached[dev] = v;
    }
    return cached[dev];
}

static int elem_bytes(int32_t dtype) { return dtype == NECT_F32 ? 4 : 2; }

constexpr size_t kLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950

// Whether every row start is 16-byte aligned (fragment loads of 16 bytes).
static int32_t vector_rows(const void *x, int64_t stride, int32_t dtype) {
    const int64_t row_bytes = stride * elem_bytes(dtype);
    return ((uintptr_t)x % 16 == 0 && row_bytes % 16 == 0) ? 1 : 0;
}

// f32 images are twice as large: one wave per block keeps them within LDS.
// float32 images are twice as large: two waves per block keep the LDS image within the CU.  The backward's
// count must equal bwd_waves<T>().
static int waves_per_block(int32_t dtype, bool backward) { (void)backward; return dtype == NECT_F32 ? 2 : 4; }

static int32_t check_mlp(const NectMlpDesc *d, int32_t dtype) {
    NECT_REQUIRE(d != nullptr, "mlp descriptor is NULL");
    NECT_REQUIRE(dtype >= NECT_F32 && dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(d->width == kWidth, "only width 64 is implemented");
    NECT_REQUIRE(d->n_hidden_layers >= 1 && d->n_hidden_layers <= NECT_MLP_MAX_HIDDEN, "n_hidden_layers must be 1..8");
    NECT_REQUIRE(d->n_input_padded % 16 == 0 && d->n_input_padded <= 128 && d->n_input_padded >= 16,
                 "padded input width must be 16..128");
    NECT_REQUIRE(d->n_output >= 1 && d->n_output <= 16 && d->n_output_padded == 16, "n_output must be 1..16");
    return 0;
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_mlp_init(NectMlpDesc *desc, int32_t n_input, int32_t n_output, int32_t width,
                                 int32_t n_hidden_layers, int32_t activation, int32_t output_activation) {
    clear_error();
    NECT_REQUIRE(desc != nullptr, "desc is NULL");
    NECT_REQUIRE(n_input >= 1 && n_output >= 1, "n_input and n_output must be >= 1");
    NECT_REQUIRE(width == kWidth, "only width 64 is implemented");
    NECT_REQUIRE(n_hidden_layers >= 1 && n_hidden_layers <= NECT_MLP_MAX_HIDDEN, "n_hidden_layers must be 1..8");
    NECT_REQUIRE(activation >= NECT_ACT_NONE && activation <= NECT_ACT_TANH, "unknown activation");
    NECT_REQUIRE(output_activation >= NECT_ACT_NONE && output_activation <= NECT_ACT_TANH, "unknown output activation");
    NectMlpDesc d = {};
    d.n_input = n_input;
    d.n_input_padded = (n_input + 15) / 16 * 16;
    d.n_output = n_output;
    d.n_output_padded = (n_output + 15) / 16 * 16;
    d.width = width;
    d.n_hidden_layers = n_hidden_layers;
    d.activation = activation;
    d.output_activation = output_activation;
    d.n_params = (int64_t)width * d.n_input_padded + (int64_t)(n_hidden_layers - 1) * width * width +
                 (int64_t)d.n_output_padded * width;
    *desc = d;
    return 0;
}

// Grouped input: checks the grid against the MLP and fills the layout fields.
static int32_t set_grouped_input(MlpArgs &a, const NectGridDesc *x_groups, int64_t n) {
    if (!x_groups) return 0;
    NECT_REQUIRE(x_groups->n_levels % 8 == 0, "grouped input needs a multiple of 8 levels");
    NECT_REQUIRE(x_groups->n_levels * x_groups->n_features == a.d.n_input &&
                     a.d.n_input == a.d.n_input_padded,
                 "grouped input must be the grid's L*F features, a multiple of 16");
    a.x_vpp = grouped_vpp(*x_groups);
    a.x_G = 8;
    a.x_F = x_groups->n_features;
    a.x_plane = n * a.x_vpp;
    return 0;
}

extern "C" int32_t nect_mlp_forward(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x,
                                    int64_t n, int64_t x_stride, const NectGridDesc *x_groups,
                                    const int32_t *n_tiles_dev, float *y, int64_t y_stride, void *stream) {
    clear_error();
    if (check_mlp(desc, dtype)) return 1;
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(weights && x && y, "NULL pointer");
    NECT_REQUIRE(x_groups || x_stride >= desc->n_input, "x_stride smaller than n_input");
    NECT_REQUIRE(y_stride >= desc->n_output, "y_stride smaller than n_output");
    const int waves = waves_per_block(dtype, false);
    MlpArgs a = {};
    a.d = *desc;
    a.L = mlp_lds(*desc, elem_bytes(dtype), false, waves);
    a.w = weights;
    a.x = x;
    a.n = n;
    a.x_stride = x_stride;
    a.x_vec = x_groups ? 0 : vector_rows(x, x_stride, dtype);
    if (set_grouped_input(a, x_groups, n)) return 1;
    a.n_tiles_dev = n_tiles_dev;
    a.y = y;
    a.y_stride = y_stride;
    const int64_t tiles = (n + kTile - 1) / kTile;
    const unsigned blocks = (unsigned)std::min<int64_t>((tiles + waves - 1) / waves, (int64_t)device_cus() * 8);
    const size_t bytes = (size_t)a.L.total * elem_bytes(dtype);
    const hipStream_t st = as_stream(stream);
    NECT_REQUIRE(bytes <= kLdsBytes, "network too large for the LDS image of the forward pass");
    if (dtype == NECT_F32) return launch_with_lds(k_mlp_forward_f32, blocks, 64 * waves, bytes, st, a, "k_mlp_forward");
    if (dtype == NECT_F16)
        return launch_with_lds(k_mlp_forward<_Float16>, blocks, 64 * waves, bytes, st, a, "k_mlp_forward");
    return launch_with_lds(k_mlp_forward<bf16_t>, blocks, 64 * waves, bytes, st, a, "k_mlp_forward");
}

// Persistent workgroups: as many as fit on the device at once (two per CU when the LDS image allows).
static unsigned backward_blocks(const NectMlpDesc *d, int32_t dtype, int64_t n) {
    const int waves = waves_per_block(dtype, true);
    const int64_t groups = ((n + kTile - 1) / kTile + waves - 1) / waves;
    const MlpLds L = mlp_lds(*d, elem_bytes(dtype), true, waves);
    const bool two = dtype != NECT_F32 && d->n_hidden_layers <= 2;  // bwd_two_per_simd
    const int per_cu = two && (size_t)L.total * elem_bytes(dtype) * 2 <= 160 * 1024 ? 2 : 1;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(groups, (int64_t)device_cus() * per_cu));
}

extern "C" int64_t nect_mlp_backward_workspace(const NectMlpDesc *desc, int32_t dtype, int64_t n) {
    if (!desc || n <= 0) return 0;
    return (int64_t)backward_blocks(desc, dtype, n) * desc->n_params * (int64_t)sizeof(float);
}

// The backward with its per-workgroup partials in `ws` (blocks slots of
// n_params floats, each written in full by its workgroup) and their reduction.
static int32_t mlp_backward_impl(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x, int64_t n,
                                 int64_t x_stride, const NectGridDesc *x_groups, const int32_t *n_tiles_dev,
                                 const float *dy, int64_t dy_stride, float loss_scale, const float *loss_scale_dev,
                                 float grad_scale, float *dweights, float *dx, int64_t dx_stride, float *ws,
                                 unsigned blocks, void *stream) {
    NECT_REQUIRE(x_groups || x_stride >= desc->n_input, "x_stride smaller than n_input");
    NECT_REQUIRE(dy_stride >= desc->n_output, "dy_stride smaller than n_output");
    NECT_REQUIRE(!dx || dx_stride >= desc->n_input, "dx_stride smaller than n_input");
    NECT_REQUIRE(loss_scale_dev || loss_scale > 0.0f, "loss_scale must be positive");
    const int waves = waves_per_block(dtype, true);
    MlpArgs a = {};
    a.d = *desc;
    a.L = mlp_lds(*desc, elem_bytes(dtype), true, waves);
    a.w = weights;
    a.x = x;
    a.n = n;
    a.x_stride = x_stride;
    a.x_vec = x_groups ? 0 : vector_rows(x, x_stride, dtype);
    if (set_grouped_input(a, x_groups, n)) return 1;
    a.n_tiles_dev = n_tiles_dev;
    a.dy = dy;
    a.dy_stride = dy_stride;
    a.loss_scale = loss_scale;
    a.loss_scale_dev = loss_scale_dev;
    a.grad_scale = grad_scale;
    a.dx = dx;
    a.dx_stride = dx_stride;
    a.ws = ws;
    const size_t bytes = (size_t)a.L.total * elem_bytes(dtype);
    const hipStream_t st = as_stream(stream);
    const int H = desc->n_hidden_layers;
    const int NQ0 = a.L.PT / 32;
    NECT_REQUIRE(dtype != NECT_F32 || H <= 3, "float32 backward supports at most 3 hidden layers (LDS image)");
    NECT_REQUIRE(bytes <= kLdsBytes, "network too large for the LDS image of the backward pass");
    int32_t rc = 1;
    if (dtype == NECT_F32) rc = launch_mlp_backward_f32(a, H, NQ0, act_is_pwl(desc->activation), blocks, waves, bytes, st);
    else if (dtype == NECT_F16) rc = launch_mlp_backward_f16(a, H, NQ0, act_is_pwl(desc->activation), blocks, waves, bytes, st);
    else rc = launch_mlp_backward_bf16(a, H, NQ0, act_is_pwl(desc->activation), blocks, waves, bytes, st);
    if (rc) return rc;
    const int64_t slots = (int64_t)blocks;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)((desc->n_params + 63) / 64)), dim3(256), 0, st, a.ws, slots,
                       desc->n_params, grad_scale, loss_scale, loss_scale_dev, dweights);
    return check_launch("k_mlp_reduce");
}

// ------------------------------------------------ allocation-free backward

// The partials of nect_mlp_backward_noalloc: a device array of the library,
// one copy per GPU, held until the process ends.  16 MiB: the 2x64 network
// (7168 parameters) keeps its usual 512 workgroups; larger networks run with
// fewer (kNoAllocFloats / n_params).
constexpr int64_t kNoAllocFloats = (int64_t)4 << 20;
static __device__ __align__(256) float noalloc_scratch[kNoAllocFloats];

// Device address of this GPU's noalloc_scratch (the array's name is a host
// stand-in in host code); looked up once per device with that device current.
static float *noalloc_scratch_address() {
    static std::mutex mu;
    static float *addr[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("nect_mlp_backward_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("nect_mlp_backward_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        addr[dev] = static_cast<float *>(p);
    }
    return addr[dev];
}

extern "C" int32_t nect_mlp_backward(const NectMlpDesc *desc, int32_t dtype, const void *weights, const void *x,
                                     int64_t n, int64_t x_stride, const NectGridDesc *x_groups,
                                     const int32_t *n_tiles_dev, const float *dy,
                                     int64_t dy_stride, float loss_scale, const float *loss_scale_dev,
                                     float grad_scale, float *dweights, float *dx,
                                     int64_t dx_stride, void *workspace, int64_t workspace_bytes, void *stream) {
    clear_error();
    if (check_mlp(desc, dtype)) return 1;
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(weights && x && dy && dweights && workspace, "NULL pointer");
    NECT_REQUIRE(workspace_bytes >= nect_mlp_backward_workspace(desc, dtype, n), "workspace too small");
    return mlp_backward_impl(desc, dtype, weights, x, n, x_stride, x_groups, n_tiles_dev, dy, dy_stride, loss_scale,
                             loss_scale_dev, grad_scale, dweights, dx, dx_stride, static_cast<float *>(workspace),
                             backward_blocks(desc, dtype, n), stream);
}

extern "C" int32_t nect_mlp_backward_noalloc(const NectMlpDesc *desc, int32_t dtype, const void *weights,
                                             const void *x, int64_t n, int64_t x_stride,
                                             const NectGridDesc *x_groups, const int32_t *n_tiles_dev,
                                             const float *dy, int64_t dy_stride, float loss_scale,
                                             const float *loss_scale_dev, float grad_scale, float *dweights,
                                             float *dx, int64_t dx_stride, void *stream) {
    clear_error();
    if (check_mlp(desc, dtype)) return 1;
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(weights && x && dy && dweights, "NULL pointer");
    NECT_REQUIRE(desc->n_params <= kNoAllocFloats,
                 "nect_mlp_backward_noalloc: the network has more parameters than its scratch holds (4194304)");
    float *ws = noalloc_scratch_address();
    if (!ws) return 1;
    const int64_t fit = kNoAllocFloats / desc->n_params;
    const unsigned blocks = (unsigned)std::min<int64_t>(backward_blocks(desc, dtype, n), fit);
    return mlp_backward_impl(desc, dtype, weights, x, n, x_stride, x_groups, n_tiles_dev, dy, dy_stride, loss_scale,
                             loss_scale_dev, grad_scale, dweights, dx, dx_stride, ws, blocks, stream);
}
