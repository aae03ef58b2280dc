// GPU octree construction (placeholder until the builder kernels land).
#include "svo_internal.hpp"

namespace svo {

int build_mesh(const float *, int64_t, int, const int32_t *, const int32_t *, int64_t, int, const svo_build_opts *,
               svo_octree &) {
    return fail(SVO_ERR_UNSUPPORTED, "GPU mesh builder not available in this build");
}

int build_grid(const uint32_t *, int, int, int, const svo_build_opts *, svo_octree &) {
    return fail(SVO_ERR_UNSUPPORTED, "GPU grid builder not available in this build");
}

}  // namespace svo
