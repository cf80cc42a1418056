// Training-step kernels: per-ray line integral + loss, and the fused Adam step.
#include <cmath>

#include "grid.hpp"
#include "sampler.hpp"

namespace nect {

constexpr int kRaysPerBlock = 4;  // one wave per ray

// One wave per ray: sum mu over the ray's kept samples in a fixed order, form
// the projection, its loss term and dloss/dmu, and broadcast dloss/dmu over
// the ray's slots.
__global__ __launch_bounds__(256) void k_integral_loss(const NectRay *__restrict__ rays,
                                                       const int32_t *__restrict__ tile_offsets, int num_rays,
                                                       const float *__restrict__ mu, const float *__restrict__ target,
                                                       float l_max, int loss, float *__restrict__ y_pred,
                                                       float *__restrict__ dmu, float *__restrict__ partials) {
    __shared__ float part[kRaysPerBlock];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * kRaysPerBlock + wave;
    float term = 0.0f;
    if (r < num_rays) {
        const NectRay d = rays[r];
        const int64_t s0 = (int64_t)tile_offsets[r] * kTile, s1 = (int64_t)tile_offsets[r + 1] * kTile;
        float acc = 0.0f;
        for (int64_t sl = s0 + lane; sl < s1; sl += 64) {
            float p[3];
            if (ray_sample(d, (int)(sl - s0), p)) acc += mu[sl];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        const float dl = d.distance / l_max;
        const float yp = acc * dl;
        const float diff = yp - target[r];
        float gy;
        if (loss == NECT_LOSS_L2) {
            term = diff * diff;
            gy = 2.0f * diff / (float)num_rays;
        } else if (loss == NECT_LOSS_GIVEN) {
            gy = target[r];
        } else {
            term = fabsf(diff);
            gy = (float)((diff > 0.0f) - (diff < 0.0f)) / (float)num_rays;
        }
        const float g = gy * dl;
        if (lane == 0) y_pred[r] = yp;
        for (int64_t sl = s0 + lane; sl < s1; sl += 64) {
            float p[3];
            dmu[sl] = ray_sample(d, (int)(sl - s0), p) ? g : 0.0f;
        }
    }
    if (lane == 0) part[wave] = term;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

constexpr int kReduceThreads = 1024;

// Fixed-order sum of the block partials -> mean loss.
__global__ __launch_bounds__(kReduceThreads) void k_loss_final(const float *__restrict__ partials, int n, int num_rays,
                                                               float *__restrict__ loss_out) {
    __shared__ float buf[kReduceThreads];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += kReduceThreads) acc += partials[i];
    buf[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kReduceThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) buf[threadIdx.x] += buf[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_out[0] = buf[0] / (float)num_rays;
}

// Dense Adam.  With `amp` set the step is decided on the device: a step whose
// gradient was found not finite only zeroes the gradient, and the bias
// corrections come from the state (written by k_amp_prepare).
template <typename T>
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, float *__restrict__ grad, float *__restrict__ m,
                                              float *__restrict__ v, T *__restrict__ shadow, int64_t n, float step_size,
                                              float beta1, float beta2, float inv_sqrt_bc2, float eps,
                                              float weight_decay, float clip, const NectAmpState *__restrict__ amp) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    bool skip = false;
    if (amp) {
        skip = amp->found_inf != 0;
        step_size = amp->step_size;
        inv_sqrt_bc2 = amp->inv_sqrt_bc2;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (skip) {
            grad[i] = 0.0f;
            continue;
        }
        float g = grad[i];
        float w = p[i];
        if (clip > 0.0f) g = fminf(fmaxf(g, -clip), clip);
        if (weight_decay != 0.0f) g = fmaf(weight_decay, w, g);
        const float mi = fmaf(beta1, m[i], (1.0f - beta1) * g);
        const float vi = fmaf(beta2, v[i], (1.0f - beta2) * g * g);
        const float denom = sqrtf(vi) * inv_sqrt_bc2 + eps;
        w = w - step_size * (mi / denom);
        m[i] = mi;
        v[i] = vi;
        p[i] = w;
        grad[i] = 0.0f;
        if (shadow) shadow[i] = from_f32<T>(w);
    }
}

// ------------------------------------------------------ mixed precision

__global__ void k_amp_init(NectAmpState *st, float scale) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        NectAmpState s = {};
        s.scale = scale;
        *st = s;
    }
}

// found_inf |= any element of g is not finite (16-byte loads where the pointer
// allows it, four of them in flight per thread: with one, a read stream of a
// billion-parameter gradient ran at 4.6 TB/s).
__global__ __launch_bounds__(256) void k_grad_check(const float *__restrict__ g, int64_t n, NectAmpState *st) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    int64_t head = 0;
    if (((uintptr_t)g & 15) == 0) {
        const int64_t n4 = n / 4;
        const float4 *g4 = reinterpret_cast<const float4 *>(g);
        for (int64_t i = t0; i < n4; i += 4 * stride) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t j = i + u * stride;
                q[u] = j < n4 ? g4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                bad |= !(isfinite(q[u].x) && isfinite(q[u].y) && isfinite(q[u].z) && isfinite(q[u].w));
        }
        head = n4 * 4;
    }
    for (int64_t i = head + t0; i < n; i += stride) bad |= !isfinite(g[i]);
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(&st->found_inf, 1);
}

// Bias corrections of the step in flight (step number state->step + 1), as the
// host computes them for nect_adam.
__global__ void k_amp_prepare(NectAmpState *st, float lr, float beta1, float beta2) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double step = (double)st->step + 1.0;
        const double bc1 = 1.0 - pow((double)beta1, step);
        const double bc2 = 1.0 - pow((double)beta2, step);
        st->step_size = (float)((double)lr / bc1);
        st->inv_sqrt_bc2 = (float)(1.0 / sqrt(bc2));
    }
}

// GradScaler.update: back off after a non-finite gradient, grow after
// `interval` finite steps in a row; count the step.
__global__ void k_amp_update(NectAmpState *st, float growth, float backoff, int interval) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (st->found_inf) {
            st->scale = st->scale * backoff;
            st->growth_tracker = 0;
            st->skipped = st->skipped + 1;
        } else {
            st->step = st->step + 1;
            const int tr = st->growth_tracker + 1;
            if (tr >= interval) {
                st->scale = st->scale * growth;
                st->growth_tracker = 0;
            } else {
                st->growth_tracker = tr;
            }
        }
        st->found_inf = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_cast(const float *__restrict__ src, T *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = from_f32<T>(src[i]);
}

static unsigned stream_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::min<int64_t>(b, 256 * 64);
}

static int32_t launch_adam(float *params, float *grad, float *m, float *v, void *shadow, int32_t shadow_dtype,
                           int64_t n, float step_size, float beta1, float beta2, float inv_sqrt_bc2, float eps,
                           float weight_decay, float clip, const NectAmpState *amp, hipStream_t st) {
    const dim3 grid(stream_blocks(n)), block(256);
    if (!shadow || shadow_dtype == NECT_F32)
        hipLaunchKernelGGL(k_adam<float>, grid, block, 0, st, params, grad, m, v, (float *)shadow, n, step_size, beta1,
                           beta2, inv_sqrt_bc2, eps, weight_decay, clip, amp);
    else if (shadow_dtype == NECT_F16)
        hipLaunchKernelGGL(k_adam<_Float16>, grid, block, 0, st, params, grad, m, v, (_Float16 *)shadow, n, step_size,
                           beta1, beta2, inv_sqrt_bc2, eps, weight_decay, clip, amp);
    else
        hipLaunchKernelGGL(k_adam<bf16_t>, grid, block, 0, st, params, grad, m, v, (bf16_t *)shadow, n, step_size,
                           beta1, beta2, inv_sqrt_bc2, eps, weight_decay, clip, amp);
    return check_launch("k_adam");
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_integral_loss(const NectRay *rays, const int32_t *tile_offsets, int32_t num_rays,
                                      const float *mu, const float *target, float l_max, int32_t loss, float *y_pred,
                                      float *dmu, float *partials, float *loss_out, void *stream) {
    clear_error();
    NECT_REQUIRE(num_rays >= 1, "num_rays must be >= 1");
    NECT_REQUIRE(rays && tile_offsets && mu && target && y_pred && dmu && partials && loss_out, "NULL pointer");
    NECT_REQUIRE(l_max > 0.0f, "l_max must be positive");
    NECT_REQUIRE(loss == NECT_LOSS_L1 || loss == NECT_LOSS_L2 || loss == NECT_LOSS_GIVEN, "unknown loss");
    const int blocks = (num_rays + kRaysPerBlock - 1) / kRaysPerBlock;
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_integral_loss, dim3(blocks), dim3(64 * kRaysPerBlock), 0, st, rays, tile_offsets, num_rays,
                       mu, target, l_max, loss, y_pred, dmu, partials);
    if (check_launch("k_integral_loss")) return 1;
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kReduceThreads), 0, st, partials, blocks, num_rays, loss_out);
    return check_launch("k_loss_final");
}

extern "C" int32_t nect_adam(float *params, float *grad, float *m, float *v, void *shadow, int32_t shadow_dtype,
                             int64_t n, float lr, float beta1, float beta2, float eps, float weight_decay, int32_t step,
                             float clip_value, void *stream) {
    clear_error();
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(params && grad && m && v, "NULL pointer");
    NECT_REQUIRE(step >= 1, "step must be >= 1");
    NECT_REQUIRE(shadow_dtype >= NECT_F32 && shadow_dtype <= NECT_BF16, "unknown dtype");
    const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
    const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
    const float step_size = (float)((double)lr / bc1);
    const float inv_sqrt_bc2 = (float)(1.0 / std::sqrt(bc2));
    return launch_adam(params, grad, m, v, shadow, shadow_dtype, n, step_size, beta1, beta2, inv_sqrt_bc2, eps,
                       weight_decay, clip_value, nullptr, as_stream(stream));
}

extern "C" int32_t nect_amp_init(NectAmpState *state, float init_scale, void *stream) {
    clear_error();
    NECT_REQUIRE(state != nullptr, "NULL state");
    NECT_REQUIRE(((uintptr_t)state & 3) == 0, "state must be 4-byte aligned");
    NECT_REQUIRE(init_scale > 0.0f && std::isfinite(init_scale), "init_scale must be positive and finite");
    hipLaunchKernelGGL(k_amp_init, dim3(1), dim3(64), 0, as_stream(stream), state, init_scale);
    return check_launch("k_amp_init");
}

extern "C" int32_t nect_grad_check(const float *grad, int64_t n, NectAmpState *state, void *stream) {
    clear_error();
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    NECT_REQUIRE(state != nullptr, "NULL state");
    if (n == 0) return 0;
    NECT_REQUIRE(grad != nullptr, "NULL gradient");
    hipLaunchKernelGGL(k_grad_check, dim3(stream_blocks((n + 3) / 4)), dim3(256), 0, as_stream(stream), grad, n, state);
    return check_launch("k_grad_check");
}

extern "C" int32_t nect_adam_amp(float *params, float *grad, float *m, float *v, void *shadow, int32_t shadow_dtype,
                                 int64_t n, float lr, float beta1, float beta2, float eps, float weight_decay,
                                 NectAmpState *state, float growth_factor, float backoff_factor,
                                 int32_t growth_interval, float clip_value, void *stream) {
    clear_error();
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    NECT_REQUIRE(state != nullptr, "NULL state");
    NECT_REQUIRE(params && grad && m && v, "NULL pointer");
    NECT_REQUIRE(shadow_dtype >= NECT_F32 && shadow_dtype <= NECT_BF16, "unknown dtype");
    NECT_REQUIRE(growth_factor >= 1.0f && backoff_factor > 0.0f && backoff_factor <= 1.0f && growth_interval >= 1,
                 "growth_factor >= 1, 0 < backoff_factor <= 1 and growth_interval >= 1 required");
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_amp_prepare, dim3(1), dim3(64), 0, st, state, lr, beta1, beta2);
    if (check_launch("k_amp_prepare")) return 1;
    if (n > 0 && launch_adam(params, grad, m, v, shadow, shadow_dtype, n, 0.0f, beta1, beta2, 0.0f, eps, weight_decay,
                             clip_value, state, st))
        return 1;
    hipLaunchKernelGGL(k_amp_update, dim3(1), dim3(64), 0, st, state, growth_factor, backoff_factor, growth_interval);
    return check_launch("k_amp_update");
}

extern "C" int32_t nect_cast(const float *src, void *dst, int32_t dtype, int64_t n, void *stream) {
    clear_error();
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(src && dst, "NULL pointer");
    const hipStream_t st = as_stream(stream);
    const dim3 grid(stream_blocks(n)), block(256);
    if (dtype == NECT_F32) hipLaunchKernelGGL(k_cast<float>, grid, block, 0, st, src, (float *)dst, n);
    else if (dtype == NECT_F16) hipLaunchKernelGGL(k_cast<_Float16>, grid, block, 0, st, src, (_Float16 *)dst, n);
    else if (dtype == NECT_BF16) hipLaunchKernelGGL(k_cast<bf16_t>, grid, block, 0, st, src, (bf16_t *)dst, n);
    else {
        set_error("unknown dtype");
        return 1;
    }
    return check_launch("k_cast");
}
