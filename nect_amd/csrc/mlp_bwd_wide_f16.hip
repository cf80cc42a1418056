// k_mlp_backward instantiations: float16 weights, inputs of 65..128 features, 1-8 hidden layers,
// piecewise-linear hidden activations.
#include "mlp_bwd.hpp"

namespace nect {

int32_t launch_mlp_backward_wide_f16(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                    size_t bytes, hipStream_t st) {
    return launch_mlp_backward_range<_Float16, 1, NECT_MLP_MAX_HIDDEN, true>(a, H, NQ0, pwl, blocks, waves, bytes, st);
}

}  // namespace nect
