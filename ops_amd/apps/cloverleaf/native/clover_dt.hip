// Time-step control on the device, for captured CloverLeaf steps (2D and 3D).
//
// The host logic that follows calc_dt's MIN reduction (apps/c/CloverLeaf/timestep.cpp and
// apps/c/CloverLeaf_3D/time_step.cpp: dt = MIN(MIN(MIN(local_dt, g_big), dtold * dtrise),
// dtmax), then time += dt at the end of the step) evaluated by one lane, which writes dt
// into the kernels' `dt` constant. Later launches of the same step read the new value, so
// a whole step replays from a hipGraph without a host round trip.
//
// State (device doubles): [0] dtold, [1] time, [2] error flag (1: dt fell below dtmin),
// [3] dt.
#include <hip/hip_runtime.h>

#include "ops_hip_rt.h"

namespace {

__device__ inline double dmin(double a, double b) { return (a < b) ? (a) : (b); }

__global__ void clover_dt_kernel(const double *local_dt, double *dt_const, double *st,
                                 double g_big, double dtrise, double dtmax, double dtmin) {
  if (threadIdx.x != 0) return;
  double dtlp = dmin(*local_dt, g_big);
  if (dtlp < dtmin) st[2] = 1.0;
  double dt_temp = g_big;
  if (dtlp <= dt_temp) dt_temp = dtlp;
  dt_temp = dmin(dmin(dt_temp, (st[0] * dtrise)), dtmax);
  if (dt_temp < dtmin) st[2] = 1.0;
  *dt_const = dt_temp;
  st[3] = dt_temp;
  st[0] = dt_temp;
  st[1] = st[1] + dt_temp;
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw OPSException(OPS_RUNTIME_ERROR, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

}  // namespace

extern "C" void *clover_dt_state_alloc(void) {
  void *p = nullptr;
  check(hipMalloc(&p, 4 * sizeof(double)), "clover_dt_state_alloc");
  return p;
}

extern "C" void clover_dt_state_free(void *st) {
  if (st) (void)hipFree(st);
}

// st <- {dtold, time, 0, dt}; synchronous with the runtime stream
extern "C" void clover_dt_state_set(void *st, double dtold, double time, double dt) {
  double v[4] = {dtold, time, 0.0, dt};
  hipStream_t s = (hipStream_t)ops_hip_get_stream();
  check(hipMemcpyAsync(st, v, sizeof v, hipMemcpyHostToDevice, s), "clover_dt_state_set");
  check(hipStreamSynchronize(s), "clover_dt_state_set");
}

// out4 <- st after everything queued on the runtime stream
extern "C" void clover_dt_state_get(void *st, double *out4) {
  hipStream_t s = (hipStream_t)ops_hip_get_stream();
  check(hipMemcpyAsync(out4, st, 4 * sizeof(double), hipMemcpyDeviceToHost, s),
        "clover_dt_state_get");
  check(hipStreamSynchronize(s), "clover_dt_state_get");
}

extern "C" void clover_dt_launch(const void *local_dt, void *dt_const, void *st, double g_big,
                                 double dtrise, double dtmax, double dtmin) {
  hipLaunchKernelGGL(clover_dt_kernel, dim3(1), dim3(64), 0,
                     (hipStream_t)ops_hip_get_stream(), (const double *)local_dt,
                     (double *)dt_const, (double *)st, g_big, dtrise, dtmax, dtmin);
  check(hipGetLastError(), "clover_dt_launch");
}
