// RegularStepGradientDescentOptimizerv4's iteration as __host__ __device__
// code, and the device loop that runs it between batched evaluations: the
// host loop (mmi_batch_regular_step) and the step kernel (mattes_optim.hip)
// advance a registration with the same operations, so the iterates of the two
// loops are the same bits.  Compiled without FMA contraction (as every source
// of the library).
#pragma once

#include "itk_mattes_hip.h"
#include "mattes_affine.h"

namespace mmi {

// CompensatedSummation (itkCompensatedSummation.hxx:38-48).
struct Kahan {
  double sum = 0.0, c = 0.0;
  __host__ __device__ void add(double x) {
    const double y = x - c;
    const double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
};

// RegularStepGradientDescentOptimizerv4::AdvanceOneStep
// (itkRegularStepGradientDescentOptimizerv4.hxx:78-171) on s.gradient, the
// derivative just evaluated; returns the stop condition it met
// (MMI_STOP_RUNNING when it updated the P parameters).
__host__ __device__ inline int regular_step_advance(mmi_regular_step_state& s, const mmi_regular_step_options& o,
                                                    double* params, int P) {
  if (!s.scales_identity && P > 0) {
    // one scale per parameter (mmi_batch_regular_step checks n_scales == P)
    for (int j = 0; j < P; ++j) {
      const double f = 1.0 / s.scales[j];
      s.gradient[j] = s.gradient[j] * f;
      s.previous_gradient[j] = s.previous_gradient[j] * f;
    }
  }
  Kahan g2;
  for (int j = 0; j < P; ++j) {
    const double w = s.gradient[j];
    g2.add(w * w);
  }
  const double magnitude = sqrt(g2.sum);
  s.gradient_magnitude = magnitude;
  if (magnitude < o.gradient_magnitude_tolerance) return MMI_STOP_GRADIENT_MAGNITUDE_TOLERANCE;
  Kahan sp;
  for (int j = 0; j < P; ++j) {
    const double a = s.gradient[j], b = s.previous_gradient[j];
    sp.add(a * b);
  }
  if (sp.sum < 0) s.relaxation *= o.relaxation_factor;
  const double step = s.relaxation * o.learning_rate;
  if (step < o.minimum_step_length) return MMI_STOP_STEP_TOO_SMALL;
  for (int j = 0; j < P; ++j) s.gradient[j] = s.gradient[j] * s.relaxation * o.learning_rate;
  const double factor = 1.0 / magnitude;
  // Transform::UpdateTransformParameters (itkTransform.hxx:71-115)
  for (int j = 0; j < P; ++j) params[j] = factor == 1.0 ? params[j] + s.gradient[j] : params[j] + s.gradient[j] * factor;
  return MMI_STOP_RUNNING;
}

// The loop's settings, in device memory (so a captured chunk of iterations
// serves every call).
struct StepCtl {
  mmi_regular_step_options o;
  int32_t max_steps;  // evaluations per member in this call
  int32_t n;
};

// Why a member left the device loop for the host's.
enum StepHandback : int32_t {
  kStepOnDevice = 0,
  kStepListOverflow = 1,  // the evaluation's exception list overflowed: the host redoes it (state untouched)
  kStepNotCertified = 2,  // after the step the fast path's bounds do not hold: the host evaluates next
};

// One registration of the device loop.
struct StepItem {
  mmi_regular_step_state s;
  double params[12];
  AffineMapInputs in;
  FastArgs* fa;  // its staged fast-path arguments (the map is rewritten after each step)
  EvalArgs* ea;  // its staged evaluation arguments (the moving transform, for listed samples)
  const DevResult* result;
  const double* deriv;
  int32_t P;
  int32_t steps;     // evaluations in this call
  int32_t handback;  // StepHandback
  int32_t dstatus;   // the evaluation's status when it ended with a metric error (1, 2)
};

hipError_t launch_regular_step(StepItem* items, const StepCtl* ctl, int32_t* halted, int n, int dim, hipStream_t st);

}  // namespace mmi
