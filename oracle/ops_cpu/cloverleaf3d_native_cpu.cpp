// ORACLE — test infrastructure only. The device time-step control of the CloverLeaf
// applications' captured-step mode (native/clover_dt.hip) has no host counterpart: on the
// host backend the captured mode is unavailable and these calls fail.
#include <stdexcept>

extern "C" {
void *clover_dt_state_alloc(void) { throw std::runtime_error("captured steps need the GPU"); }
void clover_dt_state_free(void *) {}
void clover_dt_state_set(void *, double, double, double) {
  throw std::runtime_error("captured steps need the GPU");
}
void clover_dt_state_get(void *, double *) { throw std::runtime_error("captured steps need the GPU"); }
void clover_dt_launch(const void *, void *, void *, double, double, double, double) {
  throw std::runtime_error("captured steps need the GPU");
}
}
