// Calibration kernel: the device's float-atomic request rate, the ceiling the
// hash-gradient scatter works against.  Each lane adds 1.0f to a dword of a
// 64-byte block; `lanes_per_block` adjacent lanes share a block (dwords
// 0..lanes_per_block-1 of it), so one wave instruction carries
// 64 / lanes_per_block requests.  Blocks are drawn by a hash of (wave, group,
// step) over the table, so nothing is cached between instructions.
#include "common.hpp"

namespace nect {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void k_atomic_probe(float *__restrict__ table, uint32_t n_blocks, int lanes_per_block,
                                                      int per_thread, int64_t n_threads, uint32_t seed) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_threads) return;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (uint32_t)(tid >> 6);
    const uint32_t group = (uint32_t)(lane / lanes_per_block);
    const uint32_t off = (uint32_t)(lane % lanes_per_block);
    uint32_t h = mix32(seed ^ mix32(wave * 64u + group));
    for (int k = 0; k < per_thread; ++k) {
        h = mix32(h + 0x9e3779b9u);
        unsafeAtomicAdd(table + (size_t)(h % n_blocks) * 16 + off, 1.0f);
    }
}

}  // namespace nect

using namespace nect;

extern "C" int32_t nect_calibrate_atomics(float *table, int64_t n_blocks, int32_t lanes_per_block, int32_t per_thread,
                                          int64_t n_threads, uint32_t seed, void *stream) {
    clear_error();
    NECT_REQUIRE(table != nullptr, "table is NULL");
    NECT_REQUIRE(n_blocks >= 1 && n_blocks <= 0xffffffffll, "n_blocks must be in [1, 2^32)");
    NECT_REQUIRE(lanes_per_block == 1 || lanes_per_block == 2 || lanes_per_block == 4 || lanes_per_block == 8 ||
                     lanes_per_block == 16,
                 "lanes_per_block must be 1, 2, 4, 8 or 16");
    NECT_REQUIRE(per_thread >= 1 && n_threads >= 1 && n_threads % kWave == 0,
                 "per_thread >= 1 and n_threads a positive multiple of 64");
    const int64_t blocks = (n_threads + 255) / 256;
    NECT_REQUIRE(blocks < ((int64_t)1 << 31), "too many threads");
    hipLaunchKernelGGL(k_atomic_probe, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), table,
                       (uint32_t)n_blocks, lanes_per_block, per_thread, n_threads, seed);
    return check_launch("k_atomic_probe");
}
