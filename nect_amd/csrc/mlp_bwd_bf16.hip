// k_mlp_backward instantiations: bfloat16 weights, 1-4 hidden layers.
#include "mlp_bwd.hpp"

namespace nect {

int32_t launch_mlp_backward_bf16_hi(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                   size_t bytes, hipStream_t st);

int32_t launch_mlp_backward_bf16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                hipStream_t st) {
    if (NQ0 > 2) return launch_mlp_backward_wide_bf16(a, H, NQ0, pwl, blocks, waves, bytes, st);
    if (H > 4) return launch_mlp_backward_bf16_hi(a, H, NQ0, pwl, blocks, waves, bytes, st);
    return launch_mlp_backward_range<bf16_t, 1, 4>(a, H, NQ0, pwl, blocks, waves, bytes, st);
}

}  // namespace nect
