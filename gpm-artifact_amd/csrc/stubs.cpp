// Entry points declared in gpm.h whose engines are not built yet report
// GPM_EUNSUPPORTED instead of being absent, so the ABI is complete.
#include "common.hpp"

using gpm::fail;

extern "C" {

int gpm_vertex_features(gpm_graph*, gpm_graph*, float*) {
  return fail(GPM_EUNSUPPORTED, "gpm_vertex_features: not built yet");
}

int gpm_mlp_predict(const float*, uint32_t, const float*, const float*, const float*,
                    const float*, const float*, const float*, const float*, const float*,
                    float*) {
  return fail(GPM_EUNSUPPORTED, "gpm_mlp_predict: not built yet");
}

}  // extern "C"
