// Fully fused MLP on gfx950 MFMA: types, layouts and the shared per-tile code.
//
// Orientation.  A wave computes Z^T = W * A^T for a tile of 32 samples with
// v_mfma_f32_32x32x16_{f16,bf16} (or 32x32x2_f32): samples are the N
// dimension, neurons the M dimension.  The 32x32 accumulator of lane
// (s, h) = (lane % 32, lane / 32) holds sample s and neuron rows
//     R(j, h) = 8*(j/4) + 4*h + (j%4),   j = 0..15,
// which is exactly the B-operand shape the next layer needs once the next
// layer's weight columns are permuted to the same row order: activations
// never leave registers between layers.  The permutation is applied once per
// block while staging the weights into LDS, so every A fragment is one
// contiguous LDS read.
#pragma once

#include "grid.hpp"

namespace nect {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kWidth = 64;

template <typename T>
struct MF;

template <>
struct MF<_Float16> {
    static constexpr int Kw = 8;  // K elements per lane per MFMA
    using frag = f16x8;
    __device__ static __forceinline__ f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ frag load(const _Float16 *p) { return *reinterpret_cast<const frag *>(p); }
    __device__ static __forceinline__ frag pack(const float *v) {
        frag f;
#pragma unroll
        for (int m = 0; m < 8; ++m) f[m] = (_Float16)v[m];
        return f;
    }
};

template <>
struct MF<bf16_t> {
    static constexpr int Kw = 8;
    using frag = bf16x8;
    __device__ static __forceinline__ f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ frag load(const bf16_t *p) {
        return __builtin_bit_cast(frag, *reinterpret_cast<const u16x8 *>(p));
    }
    __device__ static __forceinline__ frag pack(const float *v) {
        u16x8 u;
#pragma unroll
        for (int m = 0; m < 8; ++m) u[m] = from_f32<bf16_t>(v[m]).bits;
        return __builtin_bit_cast(frag, u);
    }
};

template <>
struct MF<float> {
    static constexpr int Kw = 1;
    using frag = float;
    __device__ static __forceinline__ f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ frag load(const float *p) { return *p; }
    __device__ static __forceinline__ frag pack(const float *v) { return v[0]; }
};

// Accumulator row of register j in lane half h.
__host__ __device__ constexpr int acc_row(int j, int h) { return 8 * (j / 4) + 4 * h + (j % 4); }

// LDS column position -> weight column (input neuron) for a 64-wide layer
// whose input arrives in accumulator layout.
template <int Kw>
__host__ __device__ constexpr int perm64(int pos) {
    const int p = pos / 32, r = pos % 32;
    const int j0 = (r / (2 * Kw)) * Kw, h = (r / Kw) % 2, m = r % Kw;
    return 32 * p + acc_row(j0 + m, h);
}

// Element offsets of the LDS image (type T) of one block.
struct MlpLds {
    int P, PT, H, NO;      // padded input, input rounded to 32, hidden layers, padded output
    int pad;               // row padding (elements) = 16 bytes
    int sw, sp, st;        // row strides: 64-wide, P-wide, transpose scratch (32 samples)
    int w0, wh, wo, w0t, wht, scratch;
    int scratch_per_wave;  // TA (64 rows) + TB (max(64, PT) rows)
    int waves;
    int total;             // elements
};

inline MlpLds mlp_lds(const NectMlpDesc &d, int elem_bytes, bool backward, int waves) {
    MlpLds L;
    L.P = d.n_input_padded;
    L.PT = (L.P + 31) / 32 * 32;
    L.H = d.n_hidden_layers;
    L.NO = d.n_output_padded;
    L.pad = 16 / elem_bytes;
    L.sw = kWidth + L.pad;
    L.sp = L.P + L.pad;
    L.st = kTile + L.pad;
    L.waves = waves;
    int o = 0;
    L.w0 = o; o += kWidth * L.sp;
    L.wh = o; o += (L.H - 1) * kWidth * L.sw;
    L.wo = o; o += L.NO * L.sw;
    L.w0t = o; if (backward) o += L.PT * L.sw;
    L.wht = o; if (backward) o += (L.H - 1) * kWidth * L.sw;
    L.scratch = o;
    L.scratch_per_wave = backward ? (kWidth + (L.PT > kWidth ? L.PT : kWidth)) * L.st : 0;
    o += waves * L.scratch_per_wave;
    L.total = o;
    return L;
}

// Global offsets (elements) of the weight matrices.
__host__ __device__ inline int64_t w_off_hidden(const NectMlpDesc &d, int k) {  // W_k, k >= 1
    return (int64_t)kWidth * d.n_input_padded + (int64_t)(k - 1) * kWidth * kWidth;
}
__host__ __device__ inline int64_t w_off_out(const NectMlpDesc &d) {
    return (int64_t)kWidth * d.n_input_padded + (int64_t)(d.n_hidden_layers - 1) * kWidth * kWidth;
}

// ------------------------------------------------------------- activations

__device__ __forceinline__ float act_fwd_inline(int a, float z) {
    switch (a) {
        case NECT_ACT_RELU: return z > 0.0f ? z : 0.0f;
        case NECT_ACT_LEAKY_RELU: return z > 0.0f ? z : 0.01f * z;
        case NECT_ACT_EXPONENTIAL: return expf(z);
        case NECT_ACT_SINE: return sinf(z);
        case NECT_ACT_SIGMOID: return 1.0f / (1.0f + expf(-z));
        case NECT_ACT_SQUAREPLUS: return 0.5f * (z + sqrtf(z * z + 4.0f));
        case NECT_ACT_SOFTPLUS: return z > 20.0f ? z : log1pf(expf(z));
        case NECT_ACT_TANH: return tanhf(z);
        default: return z;
    }
}

// d act / d z at pre-activation z.
__device__ __forceinline__ float act_bwd_inline(int a, float z) {
    switch (a) {
        case NECT_ACT_RELU: return z > 0.0f ? 1.0f : 0.0f;
        case NECT_ACT_LEAKY_RELU: return z > 0.0f ? 1.0f : 0.01f;
        case NECT_ACT_EXPONENTIAL: return expf(z);
        case NECT_ACT_SINE: return cosf(z);
        case NECT_ACT_SIGMOID: {
            const float s = 1.0f / (1.0f + expf(-z));
            return s * (1.0f - s);
        }
        case NECT_ACT_SQUAREPLUS: return 0.5f * (1.0f + z / sqrtf(z * z + 4.0f));
        case NECT_ACT_SOFTPLUS: return 1.0f / (1.0f + expf(-z));
        case NECT_ACT_TANH: {
            const float t = tanhf(z);
            return 1.0f - t * t;
        }
        default: return 1.0f;
    }
}

// Out-of-line versions keep the transcendental cases out of unrolled loops.
__device__ __noinline__ float act_fwd_slow(int a, float z) { return act_fwd_inline(a, z); }
__device__ __noinline__ float act_bwd_slow(int a, float z) { return act_bwd_inline(a, z); }

// Activation of one 16-value accumulator fragment (a is wave-uniform).
__device__ __forceinline__ void act_fwd16(int a, const f32x16 &z, float (&o)[16]) {
    if (a == NECT_ACT_RELU) {
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? z[j] : 0.0f;
    } else if (a == NECT_ACT_LEAKY_RELU) {
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? z[j] : 0.01f * z[j];
    } else if (a == NECT_ACT_NONE) {
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j];
    } else {
        for (int j = 0; j < 16; ++j) o[j] = act_fwd_slow(a, z[j]);
    }
}

// o[j] *= act'(z[j]).
__device__ __forceinline__ void act_bwd16(int a, const f32x16 &z, float (&o)[16]) {
    if (a == NECT_ACT_RELU) {
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? o[j] : 0.0f;
    } else if (a == NECT_ACT_LEAKY_RELU) {
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? o[j] : 0.01f * o[j];
    } else if (a != NECT_ACT_NONE) {
        for (int j = 0; j < 16; ++j) o[j] *= act_bwd_slow(a, z[j]);
    }
}

// Orders this wave's LDS writes before its later LDS reads (waves of a block
// run independent tile loops, so there is no block barrier inside the loop).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace nect
