// k_mlp_backward instantiations: bfloat16 weights, inputs of 65..128 features, 1-8 hidden layers,
// piecewise-linear hidden activations.
#include "mlp_bwd.hpp"

namespace nect {

int32_t launch_mlp_backward_wide_bf16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                    size_t bytes, hipStream_t st) {
    return launch_mlp_backward_range<bf16_t, 1, NECT_MLP_MAX_HIDDEN, true>(a, H, NQ0, pwl, blocks, waves, bytes, st);
}

}  // namespace nect
