// Backward pass of the fully fused MLP: the kernel template, shared by the
// per-dtype translation units that instantiate it (mlp_bwd_*.hip), and their
// launchers.
#pragma once

#include <string>

#include "mlp.hpp"

namespace nect {

template <typename Kernel>
inline int32_t launch_with_lds(Kernel k, unsigned blocks, int threads, size_t bytes, hipStream_t st, const MlpArgs &a,
                               const char *name) {
    if (bytes > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) {
            set_error(std::string(name) + ": cannot reserve LDS: " + hipGetErrorString(e));
            return 1;
        }
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(threads), bytes, st, a);
    return check_launch(name);
}

// Accumulates a 32-row x 32-col tile of sum_s TA[rowA + i][s] * TB[rowB + c][s].
template <typename T>
__device__ __forceinline__ f32x16 outer_tile(const T *TA, int rowA, const T *TB, int rowB, int st, int s, int h,
                                             f32x16 acc) {
    constexpr int Kw = MF<T>::Kw;
#pragma unroll
    for (int k0 = 0; k0 < kTile; k0 += 2 * Kw) {
        const typename MF<T>::frag A = MF<T>::load(TA + (rowA + s) * st + k0 + Kw * h);
        const typename MF<T>::frag B = MF<T>::load(TB + (rowB + s) * st + k0 + Kw * h);
        acc = MF<T>::mma(A, B, acc);
    }
    return acc;
}

// Stores accumulator-layout values (rows 32p + R(j,h), column s) as T rows of TB.
template <typename T>
__device__ __forceinline__ void put_rows(T *TB, int st, int s, int h, int p, const float (&v)[16]) {
#pragma unroll
    for (int j = 0; j < 16; ++j) TB[(32 * p + acc_row(j, h)) * st + s] = from_f32<T>(v[j]);
}

// Y^T = W^T * G for G in accumulator layout g[p][j] (rows permuted), W^T rows from LDS (permuted columns).
template <typename T>
__device__ __forceinline__ f32x16 transpose_product(const T *WT, int sw, int rowq, int s, int h,
                                                    const float (&g)[2][16]) {
    constexpr int Kw = MF<T>::Kw;
    f32x16 acc = {};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
        for (int j0 = 0; j0 < 16; j0 += Kw) {
            const typename MF<T>::frag b = MF<T>::pack(&g[p][j0]);
            acc = MF<T>::mma(MF<T>::load(WT + (rowq + s) * sw + 32 * p + 2 * j0 + Kw * h), b, acc);
        }
    }
    return acc;
}

// As transpose_product for W0 when its transposed copy is not staged: the A
// fragment of input row i (= rowq + s) and hidden positions pos..pos+Kw-1 is
// read column-wise from W0 in LDS (W0^T[i][pos] = W0[perm64(pos)][i]).
template <typename T>
__device__ __forceinline__ f32x16 transpose_product_w0(const T *W0, int sp, int rowq, int s, int h,
                                                       const float (&g)[2][16]) {
    constexpr int Kw = MF<T>::Kw;
    f32x16 acc = {};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
        for (int j0 = 0; j0 < 16; j0 += Kw) {
            const typename MF<T>::frag b = MF<T>::pack(&g[p][j0]);
            const int pos = 32 * p + 2 * j0 + Kw * h;
            typename MF<T>::frag av;
            if constexpr (Kw == 1) {
                av = W0[perm64<Kw>(pos) * sp + rowq + s];
            } else {
                T e[Kw];
#pragma unroll
                for (int m = 0; m < Kw; ++m) e[m] = W0[perm64<Kw>(pos + m) * sp + rowq + s];
                __builtin_memcpy(&av, e, sizeof(av));
            }
            acc = MF<T>::mma(av, b, acc);
        }
    }
    return acc;
}

// Waves per workgroup of the backward pass (float32 images are twice as large).
template <typename T>
constexpr int bwd_waves() { return sizeof(T) == 4 ? 2 : 4; }

// Weight-gradient blocks of the backward pass: every layer's dW is cut into
// 32x32 blocks (rows: output neurons, columns: inputs), enumerated
//   Wo: (q = 0, 1) | W_{H-1}: (p, q) in {0,1}^2 | ... | W_1 | W0: (p, q), q < NQ0
// and block b is accumulated by wave b % NW of the workgroup.
template <int H, int NQ0, int NW>
struct GradBlocks {
    static constexpr int NO = 2;
    static constexpr int NH = 4 * (H - 1);
    static constexpr int N0 = 2 * NQ0;
    static constexpr int N = NO + NH + N0;
    static constexpr int per_wave = (N + NW - 1) / NW;
    // layer of block b: H = output layer, k = hidden layer k (1..H-1), 0 = first layer
    __host__ __device__ static constexpr int layer(int b) {
        return b < NO ? H : (b < NO + NH ? H - 1 - (b - NO) / 4 : 0);
    }
    __host__ __device__ static constexpr int p(int b) {
        return b < NO ? 0 : (b < NO + NH ? ((b - NO) % 4) / 2 : (b - NO - NH) / NQ0);
    }
    __host__ __device__ static constexpr int q(int b) {
        return b < NO ? b : (b < NO + NH ? (b - NO) % 2 : (b - NO - NH) % NQ0);
    }
};

// Accumulates the blocks of `layer` owned by this wave over the NW tiles
// staged in LDS (K = 32 NW samples).
template <typename T, int H, int NQ0>
__device__ __forceinline__ void grad_blocks(const MlpLds &L, const T *scratch, int layer, int wave, int s, int h,
                                            f32x16 (&acc)[GradBlocks<H, NQ0, bwd_waves<T>()>::per_wave]) {
    constexpr int NW = bwd_waves<T>();
    using GB = GradBlocks<H, NQ0, NW>;
#pragma unroll
    for (int b = 0; b < GB::N; ++b) {
        if (GB::layer(b) != layer || b % NW != wave) continue;
#pragma unroll
        for (int t = 0; t < NW; ++t) {
            const T *TA = scratch + t * L.scratch_per_wave;
            const T *TB = TA + kWidth * lds_st<T>();
            acc[b / NW] = outer_tile<T>(TA, 32 * GB::p(b), TB, 32 * GB::q(b), lds_st<T>(), s, h, acc[b / NW]);
        }
    }
}

// Backward pass.  The NW waves of a workgroup take NW tiles at a time: each
// wave recomputes its tile's forward pass and carries dL/dz back through the
// layers in registers (MFMA against the transposed weights), staging each
// layer's dL/dz and input activations in LDS; after a barrier, each wave
// accumulates its own weight-gradient blocks over all NW tiles.  Partial sums
// per workgroup go to the workspace and are reduced by k_mlp_reduce.
// Two waves per SIMD for 16-bit types and up to two hidden layers; deeper
// networks, and float32 (whose LDS image allows one workgroup per CU anyway),
// take the whole register file of one wave per SIMD and run without spills.
template <typename T>
constexpr bool bwd_two_per_simd(int H) { return sizeof(T) == 2 && H <= 2; }

template <typename T, int H, int NQ0, bool PWL>
__global__ __launch_bounds__(256, bwd_two_per_simd<T>(H) ? 2 : 1) void k_mlp_backward(MlpArgs a) {
    constexpr int NW = bwd_waves<T>();
    using GB = GradBlocks<H, NQ0, NW>;
    extern __shared__ __align__(16) char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    stage_weights<T, true>(a, lds);
    const MlpLds &L = a.L;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar branches
    const int s = lane & 31, h = lane >> 5;
    T *scratch = lds + L.scratch;
    T *TA = scratch + wave * L.scratch_per_wave;
    T *TB = TA + kWidth * lds_st<T>();
    const float loss_scale = a.loss_scale_dev ? a.loss_scale_dev[0] : a.loss_scale;
    const float out_scale = a.grad_scale / loss_scale;
    // output rows past n_output_padded stay zero in TA for the output blocks
    for (int e = lane; e < (32 - L.NO) * lds_st<T>(); e += 64) TA[L.NO * lds_st<T>() + e] = zero_of<T>();
    // dL/dx columns: the grouped input's column map, evaluated once (its integer divisions by run-time
    // sizes would otherwise cost dozens of instructions per element of every tile)
    int *cmap = reinterpret_cast<int *>(lds + L.cmap);
    for (int i = threadIdx.x; i < L.PT; i += blockDim.x) cmap[i] = i < a.d.n_input ? in_col(a, i) : -1;
    __syncthreads();

    f32x16 acc[GB::per_wave];
#pragma unroll
    for (int i = 0; i < GB::per_wave; ++i) acc[i] = f32x16{};

    const int64_t n_tiles = tile_bound(a);
    for (int64_t group = blockIdx.x; group * NW < n_tiles; group += gridDim.x) {
        // keeps the weight fragments in LDS instead of hoisting them into registers
        __asm__ volatile("" ::: "memory");
        // the lane coordinates are made opaque per iteration so that the many LDS addresses derived from them
        // are formed where they are used rather than hoisted (and spilled) across the loop
        int s = lane & 31, h = lane >> 5;
        __asm__ volatile("" : "+v"(s), "+v"(h));
        T *TA = scratch + wave * L.scratch_per_wave;
        T *TB = TA + kWidth * lds_st<T>();
        const int64_t tile = group * NW + wave;
        const int64_t row = tile * kTile + s;
        const bool rv = tile < n_tiles && row < a.n;
        // Piecewise-linear activations in 16-bit: the backward needs, of the layers before the last, only the
        // activations in T (what the weight-gradient products read anyway) and their sign, so those layers are
        // kept as packed activation fragments, halving their registers.
        constexpr bool PACK = PWL && sizeof(T) == 2 && H > 1;
        f32x16 z[H][2];
        typename MF<T>::frag za[H > 1 ? H - 1 : 1][2][16 / MF<T>::Kw];
        if constexpr (PACK) {
            hidden_forward_packed<T, H, PWL>(a, lds, input_row<T>(a, rv ? row : 0), rv, a.x_vec != 0, s, h, za,
                                             z[H - 1]);
        } else {
            hidden_forward_from<T, H, PWL>(a, lds, input_row<T>(a, rv ? row : 0), rv, a.x_vec != 0, s, h, z);
        }

        // ---- output layer: dz_out (rows of TA), activations of the last hidden layer (rows of TB)
        float g[2][16];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int j = 0; j < 16; ++j) g[p][j] = 0.0f;
        for (int o = 0; o < L.NO; ++o) {
            float go = 0.0f;
            if (o < a.d.n_output) {
                const float zo = output_pre<T, H, PWL>(a, lds, o, h, z);
                if (rv) go = a.dy[row * a.dy_stride + o] * loss_scale * act_bwd_inline(a.d.output_activation, zo);
                const T *wo = lds + L.wo + o * lds_sw<T>();
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int j = 0; j < 16; ++j) g[p][j] = fmaf(to_f32(wo[32 * p + acc_row(j, h)]), go, g[p][j]);
            }
            if (h == 0) TA[o * lds_st<T>() + s] = from_f32<T>(go);
        }
        {
            float v[16];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                act_fwd16t<PWL>(a.d.activation, z[H - 1][p], v);
                put_rows<T>(TB, lds_st<T>(), s, h, p, v);
            }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) act_bwd16t<PWL>(a.d.activation, z[H - 1][p], g[p]);
        __syncthreads();
        grad_blocks<T, H, NQ0>(L, scratch, H, wave, s, h, acc);
        __syncthreads();

        // ---- hidden layers H-1 .. 1
#pragma unroll
        for (int k = H - 1; k >= 1; --k) {
            float v[2][16];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                put_rows<T>(TA, lds_st<T>(), s, h, p, g[p]);
                if constexpr (PACK) {
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        v[p][m] = to_f32(za[k - 1][p][0][m]);
                        v[p][8 + m] = to_f32(za[k - 1][p][1][m]);
                    }
                } else {
                    act_fwd16t<PWL>(a.d.activation, z[k - 1][p], v[p]);
                }
                put_rows<T>(TB, lds_st<T>(), s, h, p, v[p]);
            }
            const T *WT = lds + L.wht + (k - 1) * kWidth * lds_sw<T>();
            float ng[2][16];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const f32x16 d = transpose_product<T>(WT, lds_sw<T>(), 32 * q, s, h, g);
#pragma unroll
                for (int j = 0; j < 16; ++j) ng[q][j] = d[j];
                if constexpr (PACK) {
                    // act'(z) from the sign of act(z) (the activation is increasing and act(0) = 0)
                    const float slope = act_pwl_slope(a.d.activation);
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        ng[q][j] = v[q][j] > 0.0f ? ng[q][j] : (slope == 0.0f ? 0.0f : slope * ng[q][j]);
                } else {
                    act_bwd16t<PWL>(a.d.activation, z[k - 1][q], ng[q]);
                }
            }
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int j = 0; j < 16; ++j) g[p][j] = ng[p][j];
            __syncthreads();
            grad_blocks<T, H, NQ0>(L, scratch, k, wave, s, h, acc);
            __syncthreads();
        }

        // ---- first layer: dL/dz0 and the inputs, then dx
#pragma unroll
        for (int p = 0; p < 2; ++p) put_rows<T>(TA, lds_st<T>(), s, h, p, g[p]);
        {
            const T *xrow = input_row<T>(a, rv ? row : 0);
            const int halfP = L.P / 2;
            if (a.x_vpp) {
                // grouped input: whole fragments (the columns are all features, P = n_input)
                constexpr int Kw = MF<T>::Kw;
                for (int c = h * halfP; c < (h + 1) * halfP; c += Kw) {
                    T v[Kw];
                    if (rv) {
                        const typename MF<T>::frag f = load_grouped<T>(a, xrow, c);
                        __builtin_memcpy(v, &f, sizeof(v));
                    } else {
#pragma unroll
                        for (int m = 0; m < Kw; ++m) v[m] = zero_of<T>();
                    }
#pragma unroll
                    for (int m = 0; m < Kw; ++m) TB[(c + m) * lds_st<T>() + s] = v[m];
                }
            } else {
                for (int i = h * halfP; i < (h + 1) * halfP; ++i)
                    TB[i * lds_st<T>() + s] = (rv && i < a.d.n_input) ? xrow[i] : zero_of<T>();
            }
            for (int i = L.P + h; i < L.PT; i += 2) TB[i * lds_st<T>() + s] = zero_of<T>();
        }
        if (a.dx) {
#pragma unroll
            for (int q = 0; q < NQ0; ++q) {
                const f32x16 d = L.w0t_staged ? transpose_product<T>(lds + L.w0t, lds_sw<T>(), 32 * q, s, h, g)
                                               : transpose_product_w0<T>(lds + L.w0, L.sp, 32 * q, s, h, g);
                if (rv) {
                    float *dxr = a.dx + row * a.dx_stride;
                    // registers 4u..4u+3 hold input columns 32q + 8u + 4h + (0..3): one 16-byte map read
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int4 c = *reinterpret_cast<const int4 *>(cmap + 32 * q + 8 * u + 4 * h);
                        if (c.x >= 0) dxr[c.x] = d[4 * u + 0] * out_scale;
                        if (c.y >= 0) dxr[c.y] = d[4 * u + 1] * out_scale;
                        if (c.z >= 0) dxr[c.z] = d[4 * u + 2] * out_scale;
                        if (c.w >= 0) dxr[c.w] = d[4 * u + 3] * out_scale;
                    }
                }
            }
        }
        __syncthreads();
        grad_blocks<T, H, NQ0>(L, scratch, 0, wave, s, h, acc);
        __syncthreads();
    }

    // ---- this workgroup's partial weight gradients (every parameter of the slot is written once)
    float *out = a.ws + (int64_t)blockIdx.x * a.d.n_params;
    const int P = L.P;
#pragma unroll
    for (int b = 0; b < GB::N; ++b) {
        if (b % NW != wave) continue;
        const f32x16 &v = acc[b / NW];
        const int layer = GB::layer(b), p = GB::p(b), q = GB::q(b);
        if (layer == H) {
            float *wo = out + w_off_out(a.d);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int o = acc_row(j, h);
                if (o < L.NO) wo[o * kWidth + 32 * q + s] = v[j];
            }
        } else if (layer == 0) {
            const int i = 32 * q + s;
            if (i < P) {
                const int c = in_col(a, i);
#pragma unroll
                for (int j = 0; j < 16; ++j) out[(32 * p + acc_row(j, h)) * P + c] = v[j];
            }
        } else {
            float *wk = out + w_off_hidden(a.d, layer);
#pragma unroll
            for (int j = 0; j < 16; ++j) wk[(32 * p + acc_row(j, h)) * kWidth + 32 * q + s] = v[j];
        }
    }
}

// Launchers, one per weight dtype (mlp_bwd_f16*.hip, mlp_bwd_bf16*.hip, mlp_bwd_f32.hip); each returns 0 or
// records an error.  H: hidden layers, NQ0: input blocks of 32, pwl: piecewise-linear hidden activation.
int32_t launch_mlp_backward_f16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                hipStream_t st);
int32_t launch_mlp_backward_bf16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                 hipStream_t st);
int32_t launch_mlp_backward_f32(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                hipStream_t st);
int32_t launch_mlp_backward_wide_f16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                     size_t bytes, hipStream_t st);
int32_t launch_mlp_backward_wide_bf16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                      size_t bytes, hipStream_t st);

// Launches k_mlp_backward<T, H, NQ0, PWL> for the run-time (H, NQ0, pwl) if H is in [H0, H1]; WIDE
// instantiates inputs of 65..128 (NQ0 = 3, 4) for piecewise-linear activations, else NQ0 = 1, 2.
template <typename T, int H0, int H1, bool WIDE = false>
int32_t launch_mlp_backward_range(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                  hipStream_t st) {
    if constexpr (H0 <= H1) {
        if (H == H0) {
            const int threads = 64 * waves;
            if constexpr (WIDE) {
                if (!pwl) {
                    set_error("k_mlp_backward: inputs wider than 64 need a None/ReLU/LeakyReLU hidden activation");
                    return 1;
                }
                if (NQ0 == 3)
                    return launch_with_lds(k_mlp_backward<T, H0, 3, true>, blocks, threads, bytes, st, a, "k_mlp_backward");
                return launch_with_lds(k_mlp_backward<T, H0, 4, true>, blocks, threads, bytes, st, a, "k_mlp_backward");
            } else {
                if (NQ0 == 1)
                    return pwl ? launch_with_lds(k_mlp_backward<T, H0, 1, true>, blocks, threads, bytes, st, a, "k_mlp_backward")
                               : launch_with_lds(k_mlp_backward<T, H0, 1, false>, blocks, threads, bytes, st, a, "k_mlp_backward");
                return pwl ? launch_with_lds(k_mlp_backward<T, H0, 2, true>, blocks, threads, bytes, st, a, "k_mlp_backward")
                           : launch_with_lds(k_mlp_backward<T, H0, 2, false>, blocks, threads, bytes, st, a, "k_mlp_backward");
            }
        }
        return launch_mlp_backward_range<T, H0 + 1, H1, WIDE>(a, H, NQ0, pwl, blocks, waves, bytes, st);
    } else {
        set_error("k_mlp_backward: no instantiation for this depth and dtype");
        return 1;
    }
}

}  // namespace nect
