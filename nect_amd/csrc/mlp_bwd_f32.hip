// k_mlp_backward instantiations: float32 weights, 1-3 hidden layers (the LDS image limit).
#include "mlp_bwd.hpp"

namespace nect {

int32_t launch_mlp_backward_f32(const MlpArgs &a, int H, int NQ0, bool pwl, unsigned blocks, int waves, size_t bytes,
                                hipStream_t st) {
    if (NQ0 > 2) {
        set_error("float32 backward: inputs wider than 64 are not supported (LDS image)");
        return 1;
    }
    return launch_mlp_backward_range<float, 1, 3>(a, H, NQ0, pwl, blocks, waves, bytes, st);
}

}  // namespace nect
