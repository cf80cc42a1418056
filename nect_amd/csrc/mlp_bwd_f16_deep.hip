// k_mlp_backward instantiations: float16 weights, 5-8 hidden layers.
#include "mlp_bwd.hpp"

namespace nect {

int32_t launch_mlp_backward_f16_hi(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves,
                                   size_t bytes, hipStream_t st) {
    return launch_mlp_backward_range<_Float16, 5, NECT_MLP_MAX_HIDDEN>(a, H, NQ0, pwl, blocks, waves, bytes, st);
}

}  // namespace nect
