// Backward entry point (filled in with the hand-written backward blend kernel).
#include <hip/hip_runtime.h>
#include "../../include/fuzzy_ellipses_amd.h"

extern "C" fe_status fe_backward(void* stream, const fe_scene* scenes, int32_t n_scenes,
                                 const fe_view* views, int32_t n_views, const fe_options* opt,
                                 const fe_view_cotangent* cot, const fe_scene_grad* grads,
                                 void* workspace, size_t workspace_bytes) {
  (void)stream; (void)scenes; (void)n_scenes; (void)views; (void)n_views; (void)opt;
  (void)cot; (void)grads; (void)workspace; (void)workspace_bytes;
  return FE_ERR_INVALID_ARGUMENT;
}
