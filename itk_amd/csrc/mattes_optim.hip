// The device side of the RegularStep optimizer loop: after each batched
// evaluation, one step per registration on the device (no host round trip),
// with the fast path's map and bounds recomputed for the new parameters, so a
// captured chunk of iterations runs evaluation -> step -> evaluation ... on
// the GPU alone.  The step and the map are the host loop's own functions
// (mattes_optim.h, mattes_affine.h), so the iterates are the host loop's.
#include <hip/hip_runtime.h>

#include "mattes_optim.h"

namespace mmi {

namespace {

// One member per workgroup; its first lane does the scalar work (a few
// hundred FP64 operations: the step is not worth spreading over lanes).  The
// state is copied to registers (every index is a compile-time constant for
// the dimension's P = D*D + D parameters).
template <int D>
__global__ void __launch_bounds__(64) regular_step_kernel(StepItem* __restrict__ items,
                                                         const StepCtl* __restrict__ ctl,
                                                         int32_t* __restrict__ halted) {
  constexpr int P = D * D + D;
  const int r = blockIdx.x;
  if (threadIdx.x != 0 || halted[r]) return;
  StepItem& it = items[r];
  const DevResult& res = *it.result;
  if (res.exc_overflow) {  // the host redoes this evaluation on the exact path
    it.handback = kStepListOverflow;
    halted[r] = 1;
    return;
  }
  mmi_regular_step_state s = it.s;
  // the host loop keeps the last derivative as the previous one before it
  // evaluates (GetValueAndDerivative overwrites the gradient)
  for (int j = 0; j < 12; ++j) s.previous_gradient[j] = s.gradient[j];
  if (res.status != 0) {
    s.stop = MMI_STOP_METRIC_ERROR;
    s.status = MMI_ERR_DEGENERATE;
    it.s = s;
    it.dstatus = res.status;
    halted[r] = 1;
    return;
  }
  for (int j = 0; j < P; ++j) s.gradient[j] = it.deriv[j];
  s.value = res.value;
  double params[P];
  for (int j = 0; j < P; ++j) params[j] = it.params[j];
  const mmi_regular_step_options o = ctl->o;
  const int stop = regular_step_advance(s, o, params, P);
  s.iteration += 1;
  const int steps = it.steps + 1;
  it.steps = steps;
  if (stop != MMI_STOP_RUNNING) {
    s.stop = stop;
    it.s = s;
    halted[r] = 1;
    return;
  }
  for (int j = 0; j < P; ++j) it.params[j] = params[j];
  // the host loop's check before its next evaluation, when it has one
  const int max_steps = ctl->max_steps;
  if (steps < max_steps && s.iteration >= o.number_of_iterations) {
    s.stop = MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS;
    it.s = s;
    halted[r] = 1;
    return;
  }
  it.s = s;
  if (steps >= max_steps) {
    halted[r] = 1;
    return;
  }
  // the next evaluation's moving transform, map and bounds, in place
  if (!affine_fast_update_d<D>(it.in, params, it.ea->mtf, *it.fa)) {
    it.handback = kStepNotCertified;
    halted[r] = 1;
  }
}

}  // namespace

hipError_t launch_regular_step(StepItem* items, const StepCtl* ctl, int32_t* halted, int n, int dim, hipStream_t st) {
  if (dim == 2)
    hipLaunchKernelGGL(regular_step_kernel<2>, dim3(n), dim3(64), 0, st, items, ctl, halted);
  else
    hipLaunchKernelGGL(regular_step_kernel<3>, dim3(n), dim3(64), 0, st, items, ctl, halted);
  return hipGetLastError();
}

}  // namespace mmi
