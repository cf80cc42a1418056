// One layer of a wide network (more than 64 neurons) as an MFMA GEMM with its
// activation in the epilogue:  C[n][j] = epi( sum_k A[n][k] * Bt[j][k] ),
// 16-bit operands, float32 accumulation, one rounding of the result.
//
// The layers of a wide network are skinny products -- millions of rows, K and M
// of at most a few hundred -- so a layer is bound by reading A and writing C.
// A library GEMM followed by separate activation kernels moves each activation
// tensor three more times; here the forward activation (epilogue 1) or the
// backward's act'(z) mask from the saved activations H (epilogue 2) is applied
// to the accumulators before the single store.
//
// Schedule: persistent workgroups of eight waves.  Bt ([M][K], the weights in the
// orientation the product reads) is staged into LDS once per workgroup; each
// tile of 64 rows of A is loaded with 16-byte coalesced loads into registers
// one tile ahead (the next tile's loads are in flight while the current one's
// MFMAs run), stored to LDS, and multiplied.  The MFMAs compute C^T: Bt rows
// are the A operand and rows of A the B operand, so a lane's accumulators hold
// one row of C (lane s: row 32 wr + s) and output columns
// 32 b + 8 u + 4 h + (0..3) -- four consecutive 16-bit values, one 8-byte store
// (and, for the mask, one 8-byte load of H, issued before the MFMAs).  Wave
// (wr, wc), wr < 2, wc < 4, owns rows 32 wr .. +31 and columns wc M/4 ..
// +M/4-1 of the tile as NB = M/128 blocks of 32x32 (eight waves rather than
// four: the 384-wide first layer's LDS image leaves room for one workgroup per
// CU, which then still keeps two waves per SIMD in flight).  The barriers order LDS only: stores of C and the loads of
// the next tile stay in flight across them.
#include <algorithm>

#include "mlp.hpp"

namespace nect {

constexpr int kGemmRows = 64;                           // rows of A per tile
constexpr int kGemmPad = 8;                             // LDS row padding (elements): 16 bytes
constexpr int kGemmMaxK = 384;
// 16-byte A chunks per thread and tile: CH = 64 K / 8 / 512 rounded up (2 for K <= 128, 6 for K <= 384)

enum : int32_t { kEpiNone = 0, kEpiAct = 1, kEpiMask = 2 };

template <typename T>
__device__ __forceinline__ float epi_value(float v, float hv, int epi, float slope) {
    if (epi == kEpiNone) return v;
    const float key = epi == kEpiAct ? v : hv;
    return key > 0.0f ? v : (slope == 0.0f ? 0.0f : slope * v);
}

template <typename T, int NB, int CH>
__global__ __launch_bounds__(512) void k_gemm_act(const T *__restrict__ A, int64_t lda, const T *__restrict__ Bt,
                                                  int64_t ldb, int64_t N, int K, T *__restrict__ C, int64_t ldc,
                                                  int epi, float slope, const T *__restrict__ H, int64_t ldh) {
    static_assert(sizeof(T) == 2, "16-bit operands");
    constexpr int M = 128 * NB;
    extern __shared__ __align__(16) char smem_raw[];
    const int ks = K + kGemmPad;  // LDS row stride (elements), a multiple of 8
    T *Bs = reinterpret_cast<T *>(smem_raw);
    T *As = Bs + M * ks;
    const int kc = K / 8;  // 16-byte chunks per row
    for (int c = threadIdx.x; c < M * kc; c += blockDim.x) {
        const int r = c / kc, q = c - r * kc;
        *reinterpret_cast<uint4 *>(Bs + r * ks + 8 * q) = *reinterpret_cast<const uint4 *>(Bt + r * ldb + 8 * q);
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int s = lane & 31, h = lane >> 5;
    const int wr = wave >> 2, wc = wave & 3;
    const int64_t n_tiles = (N + kGemmRows - 1) / kGemmRows;
    const int chunks = kGemmRows * kc;
    // two register buffers: a tile's loads are issued two tiles before it is multiplied
    uint4 pre0[CH], pre1[CH];
    auto load_tile = [&](uint4 (&pre)[CH], int64_t tile) {
        const int64_t r0 = tile * kGemmRows;
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int c = (int)threadIdx.x + 512 * i;
            if (c < chunks) {
                const int r = c / kc, q = c - r * kc;
                pre[i] = r0 + r < N ? *reinterpret_cast<const uint4 *>(A + (r0 + r) * lda + 8 * q)
                                    : uint4{0u, 0u, 0u, 0u};
            }
        }
    };
    auto store_tile = [&](const uint4 (&pre)[CH]) {
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int c = (int)threadIdx.x + 512 * i;
            if (c < chunks) {
                const int r = c / kc, q = c - r * kc;
                *reinterpret_cast<uint4 *>(As + r * ks + 8 * q) = pre[i];
            }
        }
    };
    const int64_t G = gridDim.x;
    if ((int64_t)blockIdx.x < n_tiles) load_tile(pre0, blockIdx.x);
    if ((int64_t)blockIdx.x + G < n_tiles) load_tile(pre1, blockIdx.x + G);
    // tiles blockIdx.x + G * i: even i from pre0, odd i from pre1, each buffer refilled with the tile two ahead
    for (int64_t it = 0;; ++it) {
        const int64_t tile = (int64_t)blockIdx.x + G * it;
        if (tile >= n_tiles) break;
        if (it & 1) {
            store_tile(pre1);
            lds_barrier();
            if (tile + 2 * G < n_tiles) load_tile(pre1, tile + 2 * G);
        } else {
            store_tile(pre0);
            lds_barrier();  // the tile (and, the first time, Bt) is in LDS
            if (tile + 2 * G < n_tiles) load_tile(pre0, tile + 2 * G);
        }
        const int64_t row = tile * kGemmRows + 32 * wr + s;
        const bool rv = row < N;
        // the mask's saved activations, loaded before the products so their latency overlaps them
        uint2 hq[NB][4];
        if (epi == kEpiMask) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    hq[b][u] = rv ? *reinterpret_cast<const uint2 *>(H + row * ldh + wc * (M / 4) + 32 * b + 8 * u + 4 * h)
                                  : uint2{0u, 0u};
        }
        f32x16 acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = f32x16{};
        const T *arow = As + (32 * wr + s) * ks + 8 * h;
        const T *brow = Bs + (wc * (M / 4) + s) * ks + 8 * h;
        for (int k0 = 0; k0 < K; k0 += 16) {
            const typename MF<T>::frag x = MF<T>::load(arow + k0);
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[b] = MF<T>::mma(MF<T>::load(brow + 32 * b * ks + k0), x, acc[b]);
        }
        // accumulator register 4u + m of lane (s, h) in block b: output column 32 b + 8 u + 4 h + m of row `row`
        if (rv) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    T hv[4] = {};
                    if (epi == kEpiMask) __builtin_memcpy(hv, &hq[b][u], sizeof(hv));
                    T o[4];
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        o[m] = from_f32<T>(epi_value<T>(acc[b][4 * u + m], to_f32(hv[m]), epi, slope));
                    uint2 w;
                    __builtin_memcpy(&w, o, sizeof(w));
                    *reinterpret_cast<uint2 *>(C + row * ldc + wc * (M / 4) + 32 * b + 8 * u + 4 * h) = w;
                }
        }
        lds_barrier();  // every wave is done with As before the next tile overwrites it
    }
}

template <typename T, int NB, int CH>
static int32_t launch_gemm_act(const void *a, int64_t lda, const void *bt, int64_t ldb, int64_t n, int32_t k, void *c,
                               int64_t ldc, int32_t epi, float slope, const void *h, int64_t ldh, hipStream_t st) {
    constexpr int M = 128 * NB;
    const size_t bytes = (size_t)(M + kGemmRows) * (size_t)(k + kGemmPad) * sizeof(T);
    const auto kern = k_gemm_act<T, NB, CH>;
    if (bytes > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) {
            set_error(std::string("k_gemm_act: cannot reserve LDS: ") + hipGetErrorString(e));
            return 1;
        }
    }
    int dev = 0, cus = 256, per_cu = 1;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, 512, bytes) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int64_t tiles = (n + kGemmRows - 1) / kGemmRows;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)cus * per_cu));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(512), bytes, st, static_cast<const T *>(a), lda,
                       static_cast<const T *>(bt), ldb, n, k, static_cast<T *>(c), ldc, epi, slope,
                       static_cast<const T *>(h), ldh);
    return check_launch("k_gemm_act");
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_gemm_act(int32_t dtype, const void *a, int64_t lda, const void *bt, int64_t ldb, int64_t n,
                                 int32_t k, int32_t m, int32_t epilogue, float slope, const void *h, int64_t ldh,
                                 void *c, int64_t ldc, void *stream) {
    clear_error();
    NECT_REQUIRE(dtype == NECT_F16 || dtype == NECT_BF16, "dtype must be NECT_F16 or NECT_BF16");
    NECT_REQUIRE(m == 128 || m == 384, "m must be 128 or 384");
    NECT_REQUIRE(k >= 16 && k <= kGemmMaxK && k % 16 == 0, "k must be a multiple of 16 in 16..384");
    NECT_REQUIRE((size_t)(m + kGemmRows) * (size_t)(k + kGemmPad) * 2 <= 160 * 1024,
                 "the weights and a 64-row tile of A do not fit in LDS for this (m, k)");
    NECT_REQUIRE(m == 128 || k <= 128, "m = 384 needs k <= 128");
    NECT_REQUIRE(epilogue == kEpiNone || epilogue == kEpiAct || epilogue == kEpiMask, "unknown epilogue");
    NECT_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    NECT_REQUIRE(a && bt && c, "NULL pointer");
    NECT_REQUIRE(epilogue != kEpiMask || h != nullptr, "the mask epilogue needs h");
    NECT_REQUIRE(lda >= k && ldb >= k && ldc >= m && (epilogue != kEpiMask || ldh >= m), "leading dimensions too small");
    NECT_REQUIRE(((uintptr_t)a | (uintptr_t)bt) % 16 == 0 && (lda * 2) % 16 == 0 && (ldb * 2) % 16 == 0,
                 "a and bt rows must start 16-byte aligned");
    NECT_REQUIRE((uintptr_t)c % 8 == 0 && ldc % 4 == 0 && (epilogue != kEpiMask || ((uintptr_t)h % 8 == 0 && ldh % 4 == 0)),
                 "c and h rows must start 8-byte aligned");
    const hipStream_t st = as_stream(stream);
    // (m, k <= 128) and (128, k <= 384): the register buffer of the next tile holds 2 or 6 chunks per thread
    if (dtype == NECT_F16) {
        if (m == 384) return launch_gemm_act<_Float16, 3, 2>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
        if (k <= 128) return launch_gemm_act<_Float16, 1, 2>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
        return launch_gemm_act<_Float16, 1, 6>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
    }
    if (m == 384) return launch_gemm_act<bf16_t, 3, 2>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
    if (k <= 128) return launch_gemm_act<bf16_t, 1, 2>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
    return launch_gemm_act<bf16_t, 1, 6>(a, lda, bt, ldb, n, k, c, ldc, epilogue, slope, h, ldh, st);
}
