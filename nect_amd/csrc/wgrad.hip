// Weight gradient of a wide network layer: dW[i][j] += scale * sum_k dZ[k][i] * H[k][j]
// over the samples k -- a GEMM whose output is small (at most a few hundred rows
// and columns) and whose reduction runs over millions of samples.  Library GEMMs
// parallelise such shapes over the output and leave most of the device idle;
// here the samples are split: each workgroup reduces one chunk of samples for
// one 128 x 128 output tile on MFMA and adds its partial tile to dW with float
// atomics (16 consecutive floats of a row per 64-byte request).
#include <algorithm>

#include "mlp_bwd.hpp"

namespace nect {

constexpr int kWgTile = 128;                 // output tile (rows and columns)
constexpr int kWgSteps = kTile;              // samples staged per step
constexpr int kWgStride = kWgSteps + 8;      // LDS row stride (elements): 80-byte rows keep 16-byte alignment

// LDS image [kWgTile][kWgStride] of X[k0 .. k0+31][c0 .. c0+127], transposed so
// that a row holds one column's 32 samples (zero outside the matrix).  Thread t
// reads column t % 128 for samples 16 * (t / 128) .. +15: each load instruction
// of a wave reads 64 consecutive elements of one row.  The loads (wg_load) and
// the LDS stores (wg_store) are separate so that the next step's loads can be
// in flight while the current step's products run.
template <typename T>
struct WgStage {
    T v[16];
};

template <typename T>
__device__ __forceinline__ void wg_load(WgStage<T> &r, const T *__restrict__ X, int64_t ldx, int64_t K, int cols,
                                        int64_t k0, int c0) {
    const int c = threadIdx.x % kWgTile;
    const int kg = threadIdx.x / kWgTile;
    const bool cv = c0 + c < cols;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int64_t k = k0 + 16 * kg + u;
        r.v[u] = (cv && k < K) ? X[k * ldx + c0 + c] : zero_of<T>();
    }
}

template <typename T>
__device__ __forceinline__ void wg_store(T *lds, const WgStage<T> &r) {
    const int c = threadIdx.x % kWgTile;
    const int kg = threadIdx.x / kWgTile;
    uint4 w0, w1;
    __builtin_memcpy(&w0, r.v, sizeof(w0));
    __builtin_memcpy(&w1, r.v + 8, sizeof(w1));
    T *dst = lds + c * kWgStride + 16 * kg;  // 16-byte aligned: rows of 80 bytes, offsets of 32
    *reinterpret_cast<uint4 *>(dst) = w0;
    *reinterpret_cast<uint4 *>(dst + 8) = w1;
}

template <typename T>
__global__ __launch_bounds__(256) void k_wgrad(const T *__restrict__ A, int64_t lda, const T *__restrict__ B,
                                               int64_t ldb, int64_t K, int m, int n, int64_t chunk, float scale,
                                               float *__restrict__ C, int64_t ldc) {
    __shared__ __align__(16) T TA[kWgTile * kWgStride];
    __shared__ __align__(16) T TB[kWgTile * kWgStride];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int s = lane & 31, h = lane >> 5;
    const int wm = wave >> 1, wn = wave & 1;  // the wave's 64 x 64 quarter of the tile
    const int i0 = blockIdx.y * kWgTile, j0 = blockIdx.z * kWgTile;
    const int64_t kb = (int64_t)blockIdx.x * chunk;
    const int64_t ke = kb + chunk < K ? kb + chunk : K;
    f32x16 acc[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[p][q] = f32x16{};
    WgStage<T> ra, rb;
    wg_load<T>(ra, A, lda, ke, m, kb, i0);
    wg_load<T>(rb, B, ldb, ke, n, kb, j0);
    for (int64_t k0 = kb; k0 < ke; k0 += kWgSteps) {
        wg_store<T>(TA, ra);
        wg_store<T>(TB, rb);
        __syncthreads();
        // the next step's samples, loaded while this step's products run
        if (k0 + kWgSteps < ke) {
            wg_load<T>(ra, A, lda, ke, m, k0 + kWgSteps, i0);
            wg_load<T>(rb, B, ldb, ke, n, k0 + kWgSteps, j0);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                acc[p][q] = outer_tile<T>(TA, 64 * wm + 32 * p, TB, 64 * wn + 32 * q, kWgStride, s, h, acc[p][q]);
        __syncthreads();
    }
    // accumulator register j of lane (s, h): row acc_row(j, h), column s of its 32 x 32 block
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int col = j0 + 64 * wn + 32 * q + s;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = i0 + 64 * wm + 32 * p + acc_row(j, h);
                if (row < m && col < n) unsafeAtomicAdd(C + (int64_t)row * ldc + col, scale * acc[p][q][j]);
            }
        }
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_wgrad(int32_t dtype, const void *dz, int64_t ld_dz, const void *h, int64_t ld_h,
                              int64_t n_samples, int32_t n_out, int32_t n_in, float scale, float *dw, int64_t ld_dw,
                              void *stream) {
    clear_error();
    NECT_REQUIRE(dtype == NECT_F16 || dtype == NECT_BF16, "dtype must be NECT_F16 or NECT_BF16");
    NECT_REQUIRE(n_samples >= 0 && n_out >= 1 && n_in >= 1, "bad sizes");
    NECT_REQUIRE(ld_dz >= n_out && ld_h >= n_in && ld_dw >= n_in, "leading dimensions smaller than the rows");
    if (n_samples == 0) return 0;
    NECT_REQUIRE(dz && h && dw, "NULL pointer");
    const int64_t mt = (n_out + kWgTile - 1) / kWgTile, nt = (n_in + kWgTile - 1) / kWgTile;
    NECT_REQUIRE(mt <= 65535 && nt <= 65535, "output too large");
    // about 4096 workgroups over the device, at least 256 samples each, whole steps of 32
    int64_t chunk = (n_samples * mt * nt + 4095) / 4096;
    chunk = std::max<int64_t>(256, (chunk + kWgSteps - 1) / kWgSteps * kWgSteps);
    const int64_t chunks = (n_samples + chunk - 1) / chunk;
    NECT_REQUIRE(chunks < ((int64_t)1 << 31), "too many samples");
    const dim3 grid((unsigned)chunks, (unsigned)mt, (unsigned)nt), block(256);
    const hipStream_t st = as_stream(stream);
    if (dtype == NECT_F16)
        hipLaunchKernelGGL(k_wgrad<_Float16>, grid, block, 0, st, (const _Float16 *)dz, ld_dz, (const _Float16 *)h,
                           ld_h, n_samples, n_out, n_in, chunk, scale, dw, ld_dw);
    else
        hipLaunchKernelGGL(k_wgrad<bf16_t>, grid, block, 0, st, (const bf16_t *)dz, ld_dz, (const bf16_t *)h, ld_h,
                           n_samples, n_out, n_in, chunk, scale, dw, ld_dw);
    return check_launch("k_wgrad");
}
