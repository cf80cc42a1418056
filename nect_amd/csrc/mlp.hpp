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
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

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
        u32x4 u;
#pragma unroll
        for (int m = 0; m < 4; ++m) u[m] = pack_bf16x2(v[2 * m], v[2 * m + 1]);
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

// Row strides (elements of T) of the LDS image: 64-wide rows and 32-sample
// transpose rows, each padded by 16 bytes.  Compile-time in device code, so that
// the offsets of a lane's rows fold into the LDS instructions' immediates; the
// host's MlpLds (sw, st) holds the same values.
template <typename T>
constexpr int lds_sw() { return kWidth + 16 / (int)sizeof(T); }
template <typename T>
constexpr int lds_st() { return kTile + 16 / (int)sizeof(T); }

// Element offsets of the LDS image (type T) of one block.
struct MlpLds {
    int P, PT, H, NO;      // padded input, input rounded to 32, hidden layers, padded output
    int pad;               // row padding (elements) = 16 bytes
    int sw, sp, st;        // row strides: 64-wide, P-wide, transpose scratch (32 samples)
    int w0, wh, wo, w0t, wht, scratch;
    int scratch_per_wave;  // TA (64 rows) + TB (max(64, PT) rows)
    int cmap;              // backward: int32 [PT], row-layout column of input column i (-1 past n_input)
    int waves;
    int w0t_staged;        // backward: W0 transposed in LDS (else dL/dx reads W0 column-wise)
    int total;             // elements
};

// LDS per workgroup on gfx950.
constexpr int kLdsLimit = 160 * 1024;

inline MlpLds mlp_lds(const NectMlpDesc &d, int elem_bytes, bool backward, int waves, bool w0t_staged = true) {
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
    L.w0t_staged = backward && w0t_staged;
    L.w0t = o; if (L.w0t_staged) o += L.PT * L.sw;
    L.wht = o; if (backward) o += (L.H - 1) * kWidth * L.sw;
    L.scratch = o;
    L.scratch_per_wave = backward ? (kWidth + (L.PT > kWidth ? L.PT : kWidth)) * L.st : 0;
    o += waves * L.scratch_per_wave;
    // every region above is a whole number of 16-byte rows, so the map is 16-byte aligned
    L.cmap = o; if (backward) o += L.PT * 4 / elem_bytes;
    L.total = o;
    // a wide first layer may not leave room for its transposed copy: drop it
    if (backward && w0t_staged && (int64_t)o * elem_bytes > kLdsLimit) return mlp_lds(d, elem_bytes, true, waves, false);
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

// Piecewise-linear hidden activations (None, ReLU, LeakyReLU) as one inline
// form with the slope below zero; kernels instantiated for them make no calls.
__host__ __device__ inline bool act_is_pwl(int a) {
    return a == NECT_ACT_NONE || a == NECT_ACT_RELU || a == NECT_ACT_LEAKY_RELU;
}
__host__ __device__ inline float act_pwl_slope(int a) {
    return a == NECT_ACT_NONE ? 1.0f : (a == NECT_ACT_LEAKY_RELU ? 0.01f : 0.0f);
}

template <bool PWL>
__device__ __forceinline__ void act_fwd16t(int a, const f32x16 &z, float (&o)[16]) {
    if constexpr (PWL) {
        const float slope = act_pwl_slope(a);
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? z[j] : (slope == 0.0f ? 0.0f : slope * z[j]);
    } else {
        act_fwd16(a, z, o);
    }
}

template <bool PWL>
__device__ __forceinline__ void act_bwd16t(int a, const f32x16 &z, float (&o)[16]) {
    if constexpr (PWL) {
        const float slope = act_pwl_slope(a);
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = z[j] > 0.0f ? o[j] : (slope == 0.0f ? 0.0f : slope * o[j]);
    } else {
        act_bwd16(a, z, o);
    }
}

// Orders this wave's LDS writes before its later LDS reads (waves of a block
// run independent tile loops, so there is no block barrier inside the loop).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------- shared per-tile code

struct MlpArgs {
    NectMlpDesc d;
    MlpLds L;
    const void *w;
    const void *x;
    int64_t n;
    int64_t x_stride;
    int32_t x_vec;    // rows are 16-byte aligned: whole fragments can be loaded at once
    // Grouped input (the level-grouped encoder's layout), when x_vpp > 0: G planes of
    // [rows][x_vpp]; input column i of a row is element i % x_vpp of plane i / x_vpp and holds
    // feature in_col(i) of the row layout (x_G, x_F: the grid's level groups and features).
    int32_t x_vpp;
    int32_t x_G;
    int32_t x_F;
    int64_t x_plane;  // plane stride (elements)
    const int32_t *n_tiles_dev;
    float *y;
    int64_t y_stride;
    const float *dy;
    int64_t dy_stride;
    float loss_scale;
    const float *loss_scale_dev;  // if set, the loss scale is read from device memory
    float grad_scale;
    float *dx;
    int64_t dx_stride;
    float *ws;        // per-wave weight-gradient partials [slots][n_params]
};

// Row-layout feature index of input column i (identity unless the input is grouped).
__host__ __device__ inline int in_col(const MlpArgs &a, int i) {
    if (a.x_vpp == 0) return i;
    const int r = i % a.x_vpp;
    return group_level(i / a.x_vpp, r / a.x_F, a.x_G) * a.x_F + r % a.x_F;
}

template <typename T>
__device__ __forceinline__ T zero_of() {
    return from_f32<T>(0.0f);
}

// Cooperative copy of the weights into the block's LDS image.
template <typename T, bool BWD>
__device__ void stage_weights(const MlpArgs &a, T *lds) {
    constexpr int Kw = MF<T>::Kw;
    const MlpLds &L = a.L;
    const T *W = static_cast<const T *>(a.w);
    const int P = L.P;
    for (int e = threadIdx.x; e < kWidth * P; e += blockDim.x) {
        const int o = e / P, c = e - o * P;
        lds[L.w0 + o * L.sp + c] = W[o * P + in_col(a, c)];
    }
    for (int k = 1; k < L.H; ++k) {
        const T *Wk = W + w_off_hidden(a.d, k);
        T *dst = lds + L.wh + (k - 1) * kWidth * lds_sw<T>();
        for (int e = threadIdx.x; e < kWidth * kWidth; e += blockDim.x) {
            const int o = e >> 6, pos = e & 63;
            dst[o * lds_sw<T>() + pos] = Wk[o * kWidth + perm64<Kw>(pos)];
        }
    }
    const T *Wo = W + w_off_out(a.d);
    for (int e = threadIdx.x; e < L.NO * kWidth; e += blockDim.x) {
        const int o = e >> 6, c = e & 63;
        lds[L.wo + o * lds_sw<T>() + c] = Wo[o * kWidth + c];
    }
    if (BWD && L.w0t_staged) {
        for (int e = threadIdx.x; e < L.PT * kWidth; e += blockDim.x) {
            const int i = e >> 6, pos = e & 63;
            lds[L.w0t + i * lds_sw<T>() + pos] = (i < P) ? W[perm64<Kw>(pos) * P + in_col(a, i)] : zero_of<T>();
        }
    }
    if (BWD) {
        for (int k = 1; k < L.H; ++k) {
            const T *Wk = W + w_off_hidden(a.d, k);
            T *dst = lds + L.wht + (k - 1) * kWidth * lds_sw<T>();
            for (int e = threadIdx.x; e < kWidth * kWidth; e += blockDim.x) {
                const int c = e >> 6, pos = e & 63;
                dst[c * lds_sw<T>() + pos] = Wk[perm64<Kw>(pos) * kWidth + c];
            }
        }
    }
}

// Start of row `row` of the input (grouped: its elements in plane 0).
template <typename T>
__device__ __forceinline__ const T *input_row(const MlpArgs &a, int64_t row) {
    return static_cast<const T *>(a.x) + row * (a.x_vpp ? (int64_t)a.x_vpp : a.x_stride);
}

// Kw grouped-input elements starting at column col (a multiple of Kw): whole
// pieces of min(vpp, Kw) elements from consecutive planes.
template <typename T>
__device__ __forceinline__ typename MF<T>::frag load_grouped(const MlpArgs &a, const T *xrow, int col) {
    constexpr int Kw = MF<T>::Kw;
    if constexpr (Kw == 1) {
        return MF<T>::load(xrow + (col / a.x_vpp) * a.x_plane + col % a.x_vpp);
    } else {
        static_assert(Kw == 8 && sizeof(T) == 2, "16-bit fragments of eight elements");
        uint32_t w[4];
        const int vpp = a.x_vpp;
        if (vpp >= 8) {
            const uint4 q = *reinterpret_cast<const uint4 *>(xrow + (col / vpp) * a.x_plane + col % vpp);
            w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
        } else if (vpp == 4) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint2 q = *reinterpret_cast<const uint2 *>(xrow + (col / 4 + k) * a.x_plane);
                w[2 * k] = q.x;
                w[2 * k + 1] = q.y;
            }
        } else {  // vpp == 2
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = *reinterpret_cast<const uint32_t *>(xrow + (col / 2 + k) * a.x_plane);
        }
        const u32x4 v = {w[0], w[1], w[2], w[3]};
        return __builtin_bit_cast(typename MF<T>::frag, v);
    }
}

// Kw input elements of row `row` starting at column col (zero past n_input or n).
template <typename T>
__device__ __forceinline__ typename MF<T>::frag load_input(const T *xrow, int col, int n_input, bool rv, bool vec) {
    constexpr int Kw = MF<T>::Kw;
    if (vec && rv && col + Kw <= n_input) return MF<T>::load(xrow + col);
    float v[Kw];
#pragma unroll
    for (int m = 0; m < Kw; ++m) v[m] = (rv && col + m < n_input) ? to_f32(xrow[col + m]) : 0.0f;
    return MF<T>::pack(v);
}

// Pre-activations of the first hidden layer for the input row xrow of this lane
// (global or LDS memory; vec: 16-byte fragment loads are allowed).
template <typename T>
__device__ __forceinline__ void first_layer(const MlpArgs &a, const T *lds, const T *xrow, bool rv, bool vec, int s,
                                            int h, f32x16 &c0, f32x16 &c1) {
    constexpr int Kw = MF<T>::Kw;
    const MlpLds &L = a.L;
    const float zeros16[16] = {};
    const int halfP = L.P / 2;
    c0 = f32x16{};
    c1 = f32x16{};
    for (int t = 0; t < halfP / Kw; ++t) {
        const int col = h * halfP + t * Kw;
        typename MF<T>::frag b;
        if (a.x_vpp && xrow != nullptr) {
            if (rv) b = load_grouped<T>(a, xrow, col);
            else b = MF<T>::pack(zeros16);
        } else {
            b = load_input<T>(xrow, col, a.d.n_input, rv, vec);
        }
        c0 = MF<T>::mma(MF<T>::load(lds + L.w0 + s * L.sp + col), b, c0);
        c1 = MF<T>::mma(MF<T>::load(lds + L.w0 + (32 + s) * L.sp + col), b, c1);
    }
}

// Pre-activations of hidden layer k+1 from the B fragments b[p][j0/Kw] of
// layer k's activations (accumulator row order).
template <typename T>
__device__ __forceinline__ void next_layer(const MlpArgs &a, const T *lds, int k, int s, int h,
                                           const typename MF<T>::frag (&b)[2][16 / MF<T>::Kw], f32x16 &n0,
                                           f32x16 &n1) {
    constexpr int Kw = MF<T>::Kw;
    const MlpLds &L = a.L;
    const T *Wk = lds + L.wh + (k - 1) * kWidth * lds_sw<T>();
    n0 = f32x16{};
    n1 = f32x16{};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
        for (int j0 = 0; j0 < 16; j0 += Kw) {
            const int pos = 32 * p + 2 * j0 + Kw * h;
            n0 = MF<T>::mma(MF<T>::load(Wk + s * lds_sw<T>() + pos), b[p][j0 / Kw], n0);
            n1 = MF<T>::mma(MF<T>::load(Wk + (32 + s) * lds_sw<T>() + pos), b[p][j0 / Kw], n1);
        }
    }
}

// Activation fragments (B operands of the next layer) of pre-activations z.
template <typename T, bool PWL>
__device__ __forceinline__ void act_frags(int act, const f32x16 (&z)[2], typename MF<T>::frag (&b)[2][16 / MF<T>::Kw]) {
    constexpr int Kw = MF<T>::Kw;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        float v[16];
        act_fwd16t<PWL>(act, z[p], v);
#pragma unroll
        for (int j0 = 0; j0 < 16; j0 += Kw) b[p][j0 / Kw] = MF<T>::pack(&v[j0]);
    }
}

// Pre-activations of the hidden layers for the input row xrow of this lane.
template <typename T, int H, bool PWL = false>
__device__ __forceinline__ void hidden_forward_from(const MlpArgs &a, const T *lds, const T *xrow, bool rv, bool vec,
                                                    int s, int h, f32x16 (&z)[H][2]) {
    first_layer<T>(a, lds, xrow, rv, vec, s, h, z[0][0], z[0][1]);
#pragma unroll
    for (int k = 1; k < H; ++k) {
        typename MF<T>::frag b[2][16 / MF<T>::Kw];
        act_frags<T, PWL>(a.d.activation, z[k - 1], b);
        next_layer<T>(a, lds, k, s, h, b, z[k][0], z[k][1]);
    }
}

// The same, keeping only the last layer's pre-activations and, of the layers
// before it, the activation fragments (what the backward reads of them).
template <typename T, int H, bool PWL>
__device__ __forceinline__ void hidden_forward_packed(const MlpArgs &a, const T *lds, const T *xrow, bool rv, bool vec,
                                                      int s, int h,
                                                      typename MF<T>::frag (&za)[H > 1 ? H - 1 : 1][2][16 / MF<T>::Kw],
                                                      f32x16 (&zl)[2]) {
    first_layer<T>(a, lds, xrow, rv, vec, s, h, zl[0], zl[1]);
#pragma unroll
    for (int k = 1; k < H; ++k) {
        act_frags<T, PWL>(a.d.activation, zl, za[k - 1]);
        next_layer<T>(a, lds, k, s, h, za[k - 1], zl[0], zl[1]);
    }
}

// Pre-activations of the last hidden layer for the input row xrow of this lane;
// the depth is read at run time (no layer's values are kept but the last).
template <typename T, bool PWL = false>
__device__ __forceinline__ void hidden_forward_last(const MlpArgs &a, const T *lds, const T *xrow, bool rv, bool vec,
                                                    int s, int h, f32x16 (&zl)[2]) {
    first_layer<T>(a, lds, xrow, rv, vec, s, h, zl[0], zl[1]);
    for (int k = 1; k < a.d.n_hidden_layers; ++k) {
        typename MF<T>::frag b[2][16 / MF<T>::Kw];
        act_frags<T, PWL>(a.d.activation, zl, b);
        next_layer<T>(a, lds, k, s, h, b, zl[0], zl[1]);
    }
}

// Pre-activation of output o from the last hidden layer's pre-activations
// (both lane halves receive the full sum).
template <typename T, bool PWL = false>
__device__ __forceinline__ float output_pre_last(const MlpArgs &a, const T *lds, int o, int h, const f32x16 (&zl)[2]) {
    const T *wo = lds + a.L.wo + o * lds_sw<T>();
    float acc = 0.0f;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        float v[16];
        act_fwd16t<PWL>(a.d.activation, zl[p], v);
#pragma unroll
        for (int j = 0; j < 16; ++j) acc = fmaf(to_f32(wo[32 * p + acc_row(j, h)]), v[j], acc);
    }
    return acc + __shfl_xor(acc, 32);
}

// Pre-activation of output o (both lane halves receive the full sum).
template <typename T, int H, bool PWL = false>
__device__ __forceinline__ float output_pre(const MlpArgs &a, const T *lds, int o, int h, const f32x16 (&z)[H][2]) {
    const T *wo = lds + a.L.wo + o * lds_sw<T>();
    float acc = 0.0f;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        float v[16];
        act_fwd16t<PWL>(a.d.activation, z[H - 1][p], v);
#pragma unroll
        for (int j = 0; j < 16; ++j) acc = fmaf(to_f32(wo[32 * p + acc_row(j, h)]), v[j], acc);
    }
    return acc + __shfl_xor(acc, 32);
}

__device__ __forceinline__ int64_t tile_bound(const MlpArgs &a) {
    int64_t t = (a.n + kTile - 1) / kTile;
    if (a.n_tiles_dev) {
        const int64_t d = a.n_tiles_dev[0];
        t = d < t ? d : t;
    }
    return t;
}

}  // namespace nect
