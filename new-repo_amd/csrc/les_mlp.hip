// Learned SGS surrogate entry points (filled in by the MLP kernels).
#include <hip/hip_runtime.h>

#include "les_state.h"

extern "C" {
int les_sgs_samples(void* d_x, void* d_y, long max_samples, long* n_written, int stride) {
  (void)d_x; (void)d_y; (void)max_samples; (void)stride;
  if (n_written) *n_written = 0;
  les_set_error("learned SGS not built");
  return 1;
}
int les_mlp_set_weights(const float* w, long nweights) {
  (void)w; (void)nweights;
  les_set_error("learned SGS not built");
  return 1;
}
int les_mlp_layout(long* nweights) {
  if (nweights) *nweights = 0;
  return 1;
}
int les_mlp_judge_stats(double* out) {
  (void)out;
  return 1;
}
}
