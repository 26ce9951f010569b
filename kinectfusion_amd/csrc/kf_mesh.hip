// placeholder, replaced below
#include "kf_internal.hpp"
extern "C" {
kf_status kf_mesh_extract(kf_context*, float, kf_mesh**) { return KF_ERR_UNSUPPORTED; }
kf_status kf_mesh_size(const kf_mesh*, int64_t*, int64_t*) { return KF_ERR_UNSUPPORTED; }
kf_status kf_mesh_copy(const kf_mesh*, float*, float*, float*, uint32_t*) { return KF_ERR_UNSUPPORTED; }
void kf_mesh_free(kf_mesh*) {}
}
