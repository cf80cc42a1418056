// Fully fused MLP kernels (forward, backward, weight-gradient reduction) and
// their C-ABI entry points.  Layout conventions are described in mlp.hpp.
#include "mlp.hpp"

namespace nect {

template <typename T, int H>
__global__ __launch_bounds__(256) void k_mlp_forward(MlpArgs a) {
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
        f32x16 z[H][2];
        hidden_forward<T, H>(a, lds, row, s, h, z);
        for (int o = 0; o < a.d.n_output; ++o) {
            const float zo = output_pre<T, H>(a, lds, o, h, z);
            if (h == 0 && row < a.n) a.y[row * a.y_stride + o] = act_fwd_slow(a.d.output_activation, zo);
        }
    }
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
            const T *TB = TA + kWidth * L.st;
            acc[b / NW] = outer_tile<T>(TA, 32 * GB::p(b), TB, 32 * GB::q(b), L.st, s, h, acc[b / NW]);
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
    T *TB = TA + kWidth * L.st;
    const float loss_scale = a.loss_scale_dev ? a.loss_scale_dev[0] : a.loss_scale;
    const float out_scale = a.grad_scale / loss_scale;
    // output rows past n_output_padded stay zero in TA for the output blocks
    for (int e = lane; e < (32 - L.NO) * L.st; e += 64) TA[L.NO * L.st + e] = zero_of<T>();
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
        T *TB = TA + kWidth * L.st;
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
                const T *wo = lds + L.wo + o * L.sw;
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int j = 0; j < 16; ++j) g[p][j] = fmaf(to_f32(wo[32 * p + acc_row(j, h)]), go, g[p][j]);
            }
            if (h == 0) TA[o * L.st + s] = from_f32<T>(go);
        }
        {
            float v[16];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                act_fwd16t<PWL>(a.d.activation, z[H - 1][p], v);
                put_rows<T>(TB, L.st, s, h, p, v);
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
                put_rows<T>(TA, L.st, s, h, p, g[p]);
                if constexpr (PACK) {
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        v[p][m] = to_f32(za[k - 1][p][0][m]);
                        v[p][8 + m] = to_f32(za[k - 1][p][1][m]);
                    }
                } else {
                    act_fwd16t<PWL>(a.d.activation, z[k - 1][p], v[p]);
                }
                put_rows<T>(TB, L.st, s, h, p, v[p]);
            }
            const T *WT = lds + L.wht + (k - 1) * kWidth * L.sw;
            float ng[2][16];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const f32x16 d = transpose_product<T>(WT, L.sw, 32 * q, s, h, g);
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
        for (int p = 0; p < 2; ++p) put_rows<T>(TA, L.st, s, h, p, g[p]);
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
                    for (int m = 0; m < Kw; ++m) TB[(c + m) * L.st + s] = v[m];
                }
            } else {
                for (int i = h * halfP; i < (h + 1) * halfP; ++i)
                    TB[i * L.st + s] = (rv && i < a.d.n_input) ? xrow[i] : zero_of<T>();
            }
            for (int i = L.P + h; i < L.PT; i += 2) TB[i * L.st + s] = zero_of<T>();
        }
        if (a.dx) {
#pragma unroll
            for (int q = 0; q < NQ0; ++q) {
                const f32x16 d = transpose_product<T>(lds + L.w0t, L.sw, 32 * q, s, h, g);
                if (rv) {
                    float *dxr = a.dx + row * a.dx_stride;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int i = 32 * q + acc_row(j, h);
                        if (i < a.d.n_input) dxr[in_col(a, i)] = d[j] * out_scale;
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
    NECT_REQUIRE(d->n_hidden_layers >= 1 && d->n_hidden_layers <= NECT_MLP_MAX_HIDDEN, "n_hidden_layers must be 1..3");
    NECT_REQUIRE(d->n_input_padded % 16 == 0 && d->n_input_padded <= 64 && d->n_input_padded >= 16,
                 "padded input width must be 16..64");
    NECT_REQUIRE(d->n_output >= 1 && d->n_output <= 16 && d->n_output_padded == 16, "n_output must be 1..16");
    return 0;
}

template <typename Kernel>
static int32_t launch_with_lds(Kernel k, unsigned blocks, int threads, size_t bytes, hipStream_t st, const MlpArgs &a,
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

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_mlp_init(NectMlpDesc *desc, int32_t n_input, int32_t n_output, int32_t width,
                                 int32_t n_hidden_layers, int32_t activation, int32_t output_activation) {
    clear_error();
    NECT_REQUIRE(desc != nullptr, "desc is NULL");
    NECT_REQUIRE(n_input >= 1 && n_output >= 1, "n_input and n_output must be >= 1");
    NECT_REQUIRE(width == kWidth, "only width 64 is implemented");
    NECT_REQUIRE(n_hidden_layers >= 1 && n_hidden_layers <= NECT_MLP_MAX_HIDDEN, "n_hidden_layers must be 1..3");
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
    const int H = desc->n_hidden_layers;
#define NECT_FWD(T, HH) \
    return launch_with_lds(k_mlp_forward<T, HH>, blocks, 64 * waves, bytes, st, a, "k_mlp_forward")
#define NECT_FWD_H(T)                 \
    do {                              \
        if (H == 1) NECT_FWD(T, 1);   \
        if (H == 2) NECT_FWD(T, 2);   \
        NECT_FWD(T, 3);               \
    } while (0)
    if (dtype == NECT_F32) NECT_FWD_H(float);
    if (dtype == NECT_F16) NECT_FWD_H(_Float16);
    NECT_FWD_H(bf16_t);
#undef NECT_FWD_H
#undef NECT_FWD
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
    NECT_REQUIRE(x_groups || x_stride >= desc->n_input, "x_stride smaller than n_input");
    NECT_REQUIRE(dy_stride >= desc->n_output, "dy_stride smaller than n_output");
    NECT_REQUIRE(!dx || dx_stride >= desc->n_input, "dx_stride smaller than n_input");
    NECT_REQUIRE(loss_scale_dev || loss_scale > 0.0f, "loss_scale must be positive");
    NECT_REQUIRE(workspace_bytes >= nect_mlp_backward_workspace(desc, dtype, n), "workspace too small");
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
    a.ws = static_cast<float *>(workspace);
    const unsigned blocks = backward_blocks(desc, dtype, n);
    const size_t bytes = (size_t)a.L.total * elem_bytes(dtype);
    const hipStream_t st = as_stream(stream);
    const int H = desc->n_hidden_layers;
    const int NQ0 = a.L.PT / 32;
    int32_t rc;
#define NECT_BWD(T, HH, Q)                                                                                      \
    rc = act_is_pwl(desc->activation)                                                                           \
             ? launch_with_lds(k_mlp_backward<T, HH, Q, true>, blocks, 64 * waves, bytes, st, a, "k_mlp_backward") \
             : launch_with_lds(k_mlp_backward<T, HH, Q, false>, blocks, 64 * waves, bytes, st, a, "k_mlp_backward")
#define NECT_BWD_Q(T, HH)                 \
    do {                                  \
        if (NQ0 == 1) NECT_BWD(T, HH, 1); \
        else NECT_BWD(T, HH, 2);          \
    } while (0)
#define NECT_BWD_H(T)                        \
    do {                                     \
        if (H == 1) NECT_BWD_Q(T, 1);        \
        else if (H == 2) NECT_BWD_Q(T, 2);   \
        else NECT_BWD_Q(T, 3);               \
    } while (0)
    if (dtype == NECT_F32) NECT_BWD_H(float);
    else if (dtype == NECT_F16) NECT_BWD_H(_Float16);
    else NECT_BWD_H(bf16_t);
#undef NECT_BWD_H
#undef NECT_BWD_Q
#undef NECT_BWD
    if (rc) return rc;
    const int64_t slots = (int64_t)blocks;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)((desc->n_params + 63) / 64)), dim3(256), 0, st, a.ws, slots,
                       desc->n_params, grad_scale, loss_scale, loss_scale_dev, dweights);
    return check_launch("k_mlp_reduce");
}
