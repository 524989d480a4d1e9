// Psiformer pipeline (wavefunction/backbone/psiformer.py) -- constructed in
// psiformer.hip once available.
#include "model.h"

int jq_model::init_psiformer(const float*, size_t) {
  return jq_fail(JQ_EINVAL, "Psiformer is not built in this library version");
}

int jq_model::forward_psiformer(const float*, int, bool, hipStream_t) {
  return jq_fail(JQ_EINVAL, "Psiformer is not built in this library version");
}
