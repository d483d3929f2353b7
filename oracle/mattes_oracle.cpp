// CPU FP64 restatement of ITK 6.0.0's dense Mattes mutual-information v4 metric
// and of the filters that prepare its inputs.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  It is built by
// oracle/Makefile with -ffp-contract=off so that every expression rounds in the
// order the reference's C++ does on x86-64 (SSE2 doubles, no FMA).
//
// Reference files (paths relative to the ITK source root):
//   Mattes   = Modules/Registration/Metricsv4/include/itkMattesMutualInformationImageToImageMetricv4.hxx
//   MattesT  = Modules/Registration/Metricsv4/include/itkMattesMutualInformationImageToImageMetricv4GetValueAndDerivativeThreader.hxx
//   I2IM     = Modules/Registration/Metricsv4/include/itkImageToImageMetricv4.hxx
//   ThrBase  = Modules/Registration/Metricsv4/include/itkImageToImageMetricv4GetValueAndDerivativeThreaderBase.hxx
//   LinInterp= Modules/Core/ImageFunction/include/itkLinearInterpolateImageFunction.h
//   ImgFunc  = Modules/Core/ImageFunction/include/itkImageFunction.h / .hxx
//   ImgBase  = Modules/Core/Common/include/itkImageBase.h
//   MOTB     = Modules/Core/Transform/include/itkMatrixOffsetTransformBase.hxx
//   BSplTf   = Modules/Core/Transform/include/itkBSplineTransform.hxx
//   BSplW    = Modules/Core/Common/include/itkBSplineInterpolationWeightFunction.hxx
//   Kern     = Modules/Core/Common/include/itkBSplineKernelFunction.h
//   DKern    = Modules/Core/Common/include/itkBSplineDerivativeKernelFunction.h
//   CentDiff = Modules/Core/ImageFunction/include/itkCentralDifferenceImageFunction.hxx
//   RecGauss = Modules/Filtering/Smoothing/include/itkRecursiveGaussianImageFilter.hxx
//   RecSep   = Modules/Filtering/ImageFilterBase/include/itkRecursiveSeparableImageFilter.hxx
//   GradRG   = Modules/Filtering/ImageGradient/include/itkGradientRecursiveGaussianImageFilter.hxx
//   Kahan    = Modules/Core/Common/include/itkCompensatedSummation.hxx
//   Thr      = Modules/Registration/Metricsv4/include/itkImageToImageMetricv4GetValueAndDerivativeThreader.hxx
//   MaskSO   = Modules/Core/SpatialObjects/include/itkImageMaskSpatialObject.hxx
//   RegMethod= Modules/Registration/RegistrationMethodsv4/include/itkImageRegistrationMethodv4.hxx
//   MTGen    = Modules/Core/Common/src/itkMersenneTwisterRandomVariateGenerator.cxx (and include/...h)
//   RandIt   = Modules/Core/Common/include/itkImageRandomConstIteratorWithIndex.hxx (and .h)
//
// Parity pins (tests/test_oracle.py): the reference test's known metric
// values and its finite-difference derivative check
// (Modules/Registration/Metricsv4/test/itkMattesMutualInformationImageToImageMetricv4Test.cxx:349-427),
// closed forms of the kernels, brute-force joint histograms on tiny images,
// invariances of the metric, and closed forms of the recursive Gaussian.

#include "mattes_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

namespace {

void set_err(char* err, int errlen, const std::string& msg) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", msg.c_str());
  }
}

// Math::Floor<int64>(double) = RoundHalfIntegerToEven_64(2x - 0.5) >> 1
// (Modules/Core/Common/include/itkMathDetail.h:393-396).
inline int64_t itk_floor(double x) {
  return (int64_t)std::nearbyint(2.0 * x - 0.5) >> 1;
}

// BSplineKernelFunction<3>::Evaluate (Kern:153-175).
inline double bspline3(double u) {
  const double a = std::fabs(u);
  if (a < 1.0) {
    const double s = a * a;
    return (4.0 - 6.0 * s + 3.0 * s * a) / 6.0;
  }
  if (a < 2.0) {
    const double s = a * a;
    return (8.0 - 12.0 * a + 6.0 * s - s * a) / 6.0;
  }
  return 0.0;
}

// BSplineDerivativeKernelFunction<3>::Evaluate (DKern:157-180).
inline double bspline3_deriv(double u) {
  if (u >= 0.0 && u < 1.0) return -2.0 * u + 1.5 * u * u;
  if (u > -1.0 && u < 0.0) return -2.0 * u - 1.5 * u * u;
  if (u >= 1.0 && u < 2.0) return -2.0 + 2.0 * u - 0.5 * u * u;
  if (u > -2.0 && u <= -1.0) return 2.0 + 2.0 * u + 0.5 * u * u;
  return 0.0;
}

struct Image {
  const orc_geometry* g;
  const void* px;
  int f64;
  double at(int64_t x, int64_t y, int64_t z) const {
    const int64_t ix = x - g->start[0], iy = y - g->start[1], iz = z - g->start[2];
    const int64_t off = ix + g->size[0] * (iy + g->size[1] * iz);
    return f64 ? ((const double*)px)[off] : (double)((const float*)px)[off];
  }
};

// ImageBase::TransformIndexToPhysicalPoint (ImgBase:594-604): origin first,
// then += M[i][j]*index[j] in j order.
inline void index_to_point(const orc_geometry* g, const int64_t* idx, double* p) {
  const int D = g->dim;
  for (int i = 0; i < D; ++i) {
    p[i] = g->origin[i];
    for (int j = 0; j < D; ++j) p[i] += g->index_to_physical[i * D + j] * (double)idx[j];
  }
}

// ImageBase::TransformPhysicalPointToContinuousIndex (ImgBase:517-531) with
// Matrix*Vector summation order (Modules/Core/Common/include/itkMatrix.hxx:30-41).
inline void point_to_cindex(const orc_geometry* g, const double* p, double* ci) {
  const int D = g->dim;
  double v[3];
  for (int k = 0; k < D; ++k) v[k] = p[k] - g->origin[k];
  for (int r = 0; r < D; ++r) {
    double sum = 0.0;
    for (int c = 0; c < D; ++c) sum += g->physical_to_index[r * D + c] * v[c];
    ci[r] = sum;
  }
}

// ImageFunction::IsInsideBuffer(ContinuousIndex) (ImgFunc.h:159-170) with the
// bounds set in ImageFunction::SetInputImage (ImgFunc.hxx:50-67).
inline bool inside_buffer(const orc_geometry* g, const double* ci) {
  for (int j = 0; j < g->dim; ++j) {
    const int64_t endIdx = g->start[j] + g->size[j] - 1;
    const double lo = (double)g->start[j] - 0.5;
    const double hi = (double)endIdx + 0.5;
    if (!(ci[j] >= lo && ci[j] < hi)) return false;
  }
  return true;
}

// Math::RoundHalfIntegerUp<int64>(double) as built on x86-64 (SSE2 rounding of
// 2x + 0.5 to even, then a halving shift; itkMathDetail.h:382-385).
inline int64_t itk_round(double x) {
  return (int64_t)std::nearbyint(2.0 * x + 0.5) >> 1;
}

// ImageBase::TransformPhysicalPointToIndex (ImgBase:465-479: per row,
// sum += M[i][j] * (p[j] - origin[j]) from zero, then RoundHalfIntegerUp) and
// ImageRegion::IsInside(Index) on the image's region.  Indices whose doubled
// value does not fit int64 lie outside every region.
inline bool physical_point_to_index(const orc_geometry* g, const double* p, int64_t* idx) {
  const int D = g->dim;
  bool inside = true;
  for (int i = 0; i < D; ++i) {
    double sum = 0.0;
    for (int j = 0; j < D; ++j) sum += g->physical_to_index[i * D + j] * (p[j] - g->origin[j]);
    if (!(std::fabs(sum) < 1.0e18)) return false;
    idx[i] = itk_round(sum);
    inside = inside && idx[i] >= g->start[i] && idx[i] < g->start[i] + g->size[i];
  }
  return inside;
}

inline int64_t buffer_offset(const orc_geometry* g, const int64_t* idx) {
  const int D = g->dim;
  const int64_t z = D == 3 ? idx[2] - g->start[2] : 0;
  return (idx[0] - g->start[0]) + g->size[0] * ((idx[1] - g->start[1]) + g->size[1] * z);
}

// ImageMaskSpatialObject::IsInsideInObjectSpace with UseMaskValue off
// (MaskSO:36-48); the object-to-world transform is the
// identity, so world and object points coincide (itkSpatialObject.hxx:186-190).
struct Mask {
  const orc_geometry* g = nullptr;
  const uint8_t* px = nullptr;
  bool inside(const double* p) const {
    int64_t idx[3] = {0, 0, 0};
    if (!physical_point_to_index(g, p, idx)) return false;
    return px[buffer_offset(g, idx)] != 0;
  }
};

// LinearInterpolateImageFunction::EvaluateOptimized, 2-D (LinInterp:157-224) and
// 3-D (LinInterp:226-471).  The reference branches on distance <= 0 and on
// the upper neighbour leaving the buffer; in every branch the result equals
// the full multilinear formula with that axis' distance forced to zero,
// because v + (w - v) * 0 == v exactly.  The oracle keeps the reference's
// branches so that it rounds identically even for non-finite inputs.
template <typename Fetch>
double interp_scalar(const orc_geometry* g, const double* ci, Fetch fetch) {
  const int D = g->dim;
  int64_t b[3] = {0, 0, 0};
  double d[3] = {0, 0, 0};
  int64_t endIdx[3] = {0, 0, 0};
  for (int j = 0; j < D; ++j) {
    b[j] = std::max(itk_floor(ci[j]), g->start[j]);
    d[j] = ci[j] - (double)b[j];
    endIdx[j] = g->start[j] + g->size[j] - 1;
  }
  if (D == 2) {
    const double v00 = fetch(b[0], b[1], 0);
    if (d[0] <= 0. && d[1] <= 0.) return v00;
    if (d[1] <= 0.) {
      if (b[0] + 1 > endIdx[0]) return v00;
      const double v10 = fetch(b[0] + 1, b[1], 0);
      return v00 + (v10 - v00) * d[0];
    } else if (d[0] <= 0.) {
      if (b[1] + 1 > endIdx[1]) return v00;
      const double v01 = fetch(b[0], b[1] + 1, 0);
      return v00 + (v01 - v00) * d[1];
    }
    if (b[0] + 1 > endIdx[0]) {
      if (b[1] + 1 > endIdx[1]) return v00;
      const double v01 = fetch(b[0], b[1] + 1, 0);
      return v00 + (v01 - v00) * d[1];
    }
    const double v10 = fetch(b[0] + 1, b[1], 0);
    const double vx0 = v00 + (v10 - v00) * d[0];
    if (b[1] + 1 > endIdx[1]) return vx0;
    const double v11 = fetch(b[0] + 1, b[1] + 1, 0);
    const double v01 = fetch(b[0], b[1] + 1, 0);
    const double vx1 = v01 + (v11 - v01) * d[0];
    return vx0 + (vx1 - vx0) * d[1];
  }
  // 3-D: the reference's branch tree collapses to per-axis "active" flags.
  const bool a0 = d[0] > 0. && b[0] + 1 <= endIdx[0];
  const bool a1 = d[1] > 0. && b[1] + 1 <= endIdx[1];
  const bool a2 = d[2] > 0. && b[2] + 1 <= endIdx[2];
  const double v000 = fetch(b[0], b[1], b[2]);
  if (!a0 && !a1 && !a2) return v000;
  // Each reference branch lerps x first, then y, then z over the active axes,
  // with exactly the expressions below (see LinInterp:262-467).
  auto lerp = [](double a, double c, double t) { return a + (c - a) * t; };
  const int64_t x1 = a0 ? b[0] + 1 : b[0];
  const int64_t y1 = a1 ? b[1] + 1 : b[1];
  const int64_t z1 = a2 ? b[2] + 1 : b[2];
  if (!a2) {
    if (!a1) return lerp(v000, fetch(x1, b[1], b[2]), d[0]);
    if (!a0) return lerp(v000, fetch(b[0], y1, b[2]), d[1]);
    const double vx00 = lerp(v000, fetch(x1, b[1], b[2]), d[0]);
    const double vx10 = lerp(fetch(b[0], y1, b[2]), fetch(x1, y1, b[2]), d[0]);
    return lerp(vx00, vx10, d[1]);
  }
  if (!a1) {
    if (!a0) return lerp(v000, fetch(b[0], b[1], z1), d[2]);
    const double vx00 = lerp(v000, fetch(x1, b[1], b[2]), d[0]);
    const double vx01 = lerp(fetch(b[0], b[1], z1), fetch(x1, b[1], z1), d[0]);
    return lerp(vx00, vx01, d[2]);
  }
  if (!a0) {
    const double v0x0 = lerp(v000, fetch(b[0], y1, b[2]), d[1]);
    const double v0x1 = lerp(fetch(b[0], b[1], z1), fetch(b[0], y1, z1), d[1]);
    return lerp(v0x0, v0x1, d[2]);
  }
  const double vx00 = lerp(v000, fetch(x1, b[1], b[2]), d[0]);
  const double vx10 = lerp(fetch(b[0], y1, b[2]), fetch(x1, y1, b[2]), d[0]);
  const double vxx0 = lerp(vx00, vx10, d[1]);
  const double vx01 = lerp(fetch(b[0], b[1], z1), fetch(x1, b[1], z1), d[0]);
  const double vx11 = lerp(fetch(b[0], y1, z1), fetch(x1, y1, z1), d[0]);
  const double vxx1 = lerp(vx01, vx11, d[1]);
  return lerp(vxx0, vxx1, d[2]);
}

// The interpolator's double assigned to the pixel type (ImgMetric:318,369):
// float rounds, double keeps, an integral type truncates toward zero.
inline double round_pixel(double v, int f64, int integer) {
  return integer ? std::trunc(v) : (f64 ? v : (double)(float)v);
}

// B-spline transform (BSplTf:478-590, weights BSplW:43-94).
struct BSplineEval {
  bool inside = false;
  double weights[64];
  int64_t node[64];
  int nw = 0;
};

// BSplineTransform::InsideValidRegion (BSplTf:476-506), including the
// 4-ULP snap of indices equal to the upper limit.
bool bspline_inside_valid(const orc_geometry* grid, double* ci) {
  const double minLimit = 0.5 * (double)(3 - 1);
  for (int j = 0; j < grid->dim; ++j) {
    const double maxLimit = (double)grid->size[j] - 0.5 * (double)(3 - 1) - 1.0;
    // Math::FloatAlmostEqual(ci, maxLimit, 4) (Modules/Core/Common/include/itkMath.h:331-356)
    bool almost = false;
    const double absDiff = std::fabs(ci[j] - maxLimit);
    if (absDiff <= 0.1 * std::numeric_limits<double>::epsilon()) {
      almost = true;
    } else if (std::signbit(ci[j]) == std::signbit(maxLimit)) {
      int64_t a, c;
      std::memcpy(&a, &ci[j], 8);
      std::memcpy(&c, &maxLimit, 8);
      // FloatIEEE::AsULP maps negatives to a monotone integer line.
      auto asulp = [](int64_t i) { return i < 0 ? (int64_t)(0x8000000000000000ULL) - i : i; };
      int64_t u = asulp(a) - asulp(c);
      if (u < 0) u = -u;
      almost = u <= 4;
    }
    if (almost) {
      int64_t bits;
      std::memcpy(&bits, &maxLimit, 8);
      bits -= 6;  // Math::FloatAddULP(maxLimit, -6)
      std::memcpy(&ci[j], &bits, 8);
    } else if (ci[j] >= maxLimit) {
      return false;
    } else if (ci[j] < minLimit) {
      return false;
    }
  }
  return true;
}

void bspline_weights(const orc_transform* tf, const double* p, BSplineEval* e) {
  const orc_geometry* grid = &tf->grid;
  const int D = grid->dim;
  double ci[3];
  point_to_cindex(grid, p, ci);
  e->inside = bspline_inside_valid(grid, ci);
  if (!e->inside) return;
  int64_t start[3] = {0, 0, 0};
  double w1[3][4];
  for (int j = 0; j < D; ++j) {
    start[j] = itk_floor(ci[j] + 0.5 - 3 / 2.0);
    double x = ci[j] - (double)start[j];
    for (int k = 0; k <= 3; ++k) {
      w1[j][k] = bspline3(x);
      x -= 1.0;
    }
  }
  const int nw = D == 2 ? 16 : 64;
  e->nw = nw;
  for (int k = 0; k < nw; ++k) {
    const int off[3] = {k & 3, (k >> 2) & 3, (k >> 4) & 3};
    double w = 1.0;
    for (int j = 0; j < D; ++j) w *= w1[j][off[j]];
    e->weights[k] = w;
    int64_t lin = start[D - 1] + off[D - 1];
    for (int j = D - 2; j >= 0; --j) lin = lin * grid->size[j] + (start[j] + off[j]);
    e->node[k] = lin;
  }
}

int64_t grid_nodes(const orc_geometry* g) {
  int64_t n = 1;
  for (int j = 0; j < g->dim; ++j) n *= g->size[j];
  return n;
}

// Transform::TransformPoint for the three kinds the driver uses.
// Affine: MatrixOffsetTransformBase::TransformPoint = m_Matrix * point + m_Offset (MOTB:124-129).
// B-spline: BSplineTransform::TransformPoint (BSplTf:508-590).
void tf_point(const orc_transform* tf, int D, const double* in, double* out, BSplineEval* be) {
  if (tf->kind == ORC_TF_IDENTITY) {
    for (int i = 0; i < D; ++i) out[i] = in[i];
    return;
  }
  if (tf->kind == ORC_TF_AFFINE) {
    for (int r = 0; r < D; ++r) {
      double sum = 0.0;
      for (int c = 0; c < D; ++c) sum += tf->matrix[r * D + c] * in[c];
      out[r] = sum + tf->offset[r];
    }
    return;
  }
  BSplineEval local;
  BSplineEval* e = be ? be : &local;
  bspline_weights(tf, in, e);
  if (!e->inside) {
    for (int i = 0; i < D; ++i) out[i] = in[i];
    return;
  }
  const int64_t pdim = grid_nodes(&tf->grid);
  for (int j = 0; j < D; ++j) out[j] = 0.0;
  for (int k = 0; k < e->nw; ++k)
    for (int j = 0; j < D; ++j) out[j] += e->weights[k] * tf->coeffs[j * pdim + e->node[k]];
  for (int j = 0; j < D; ++j) out[j] += in[j];
}

int64_t tf_num_params(const orc_transform* tf, int D) {
  if (tf->kind == ORC_TF_AFFINE) return D * D + D;
  if (tf->kind == ORC_TF_BSPLINE) return D * grid_nodes(&tf->grid);
  return 0;
}

// CentralDifferenceImageFunction::EvaluateSpecialized at a point (CentDiff:252-318),
// with its default LinearInterpolateImageFunction (double output, not rounded).
void central_difference(const Image& mov, const double* p, double* grad) {
  const orc_geometry* g = mov.g;
  const int D = g->dim;
  double n1[3], n2[3];
  for (int i = 0; i < D; ++i) n1[i] = n2[i] = p[i];
  auto fetch = [&](int64_t x, int64_t y, int64_t z) { return mov.at(x, y, z); };
  for (int dim = 0; dim < D; ++dim) {
    const double offset = 0.5 * g->spacing[dim];
    double ci[3];
    n1[dim] = p[dim] - offset;
    point_to_cindex(g, n1, ci);
    if (!inside_buffer(g, ci)) {
      grad[dim] = 0.0;
      n1[dim] = p[dim];
      n2[dim] = p[dim];
      continue;
    }
    n2[dim] = p[dim] + offset;
    point_to_cindex(g, n2, ci);
    if (!inside_buffer(g, ci)) {
      grad[dim] = 0.0;
      n1[dim] = p[dim];
      n2[dim] = p[dim];
      continue;
    }
    const double delta = n2[dim] - n1[dim];
    if (delta > 10.0 * std::numeric_limits<double>::epsilon()) {
      double c2[3], c1[3];
      point_to_cindex(g, n2, c2);
      point_to_cindex(g, n1, c1);
      grad[dim] = (interp_scalar(g, c2, fetch) - interp_scalar(g, c1, fetch)) / delta;
    } else {
      grad[dim] = 0.0;
    }
    n1[dim] = p[dim];
    n2[dim] = p[dim];
  }
}

struct ThreadState {
  std::vector<double> pdf;      // [B*B], row = fixed bin
  std::vector<double> fmarg;    // [B]
  std::vector<double> dpdf;     // [B*B*P]: index (f*B + m)*P + mu (reference layout {P,B,B}, P fastest)
  int64_t nvalid = 0;
};

struct Problem {
  const orc_geometry* virt;
  Image fixed, moving;
  const double* mgrad;
  const orc_transform* ftf;
  const orc_transform* mtf;
  const orc_options* opt;
  double fBinSize, fNormMin, mBinSize, mNormMin;
  int64_t P;
  orc_result* res;
  Mask fmask, mmask;
};

// One virtual sample: ThrBase:174-273 (ProcessVirtualPoint), I2IM:289-373
// (TransformAndEvaluateFixed/MovingPoint), MattesT:215-387 (ProcessPoint).
// Masks: TransformAndEvaluateFixedPoint / MovingPoint test the mask at the
// mapped point before the buffer (I2IM:298-309, 351-362).
void process_sample(const Problem& pb, const double* vp, int64_t lin, ThreadState& ts) {
  const int D = pb.virt->dim;
  const int B = pb.opt->bins;

  // Fixed side.
  double fp[3], ci[3];
  tf_point(pb.ftf, D, vp, fp, nullptr);
  if (pb.fmask.g && !pb.fmask.inside(fp)) return;
  point_to_cindex(pb.fixed.g, fp, ci);
  if (!inside_buffer(pb.fixed.g, ci)) return;
  auto ffetch = [&](int64_t x, int64_t y, int64_t z) { return pb.fixed.at(x, y, z); };
  const double fval = round_pixel(interp_scalar(pb.fixed.g, ci, ffetch), pb.fixed.f64, pb.opt->pixel_integer);

  // Moving side.
  double mp[3];
  BSplineEval be;
  tf_point(pb.mtf, D, vp, mp, &be);
  // CompositeTransform::TransformPoint: the earlier queue entries after it,
  // in reverse queue order (itkCompositeTransform.hxx:60-71)
  for (int st = pb.mtf->n_initial - 1; st >= 0; --st) {
    double q[3];
    tf_point(&pb.mtf->initial[st], D, mp, q, nullptr);
    for (int i = 0; i < D; ++i) mp[i] = q[i];
  }
  if (pb.mmask.g && !pb.mmask.inside(mp)) return;
  point_to_cindex(pb.moving.g, mp, ci);
  if (!inside_buffer(pb.moving.g, ci)) return;
  auto mfetch = [&](int64_t x, int64_t y, int64_t z) { return pb.moving.at(x, y, z); };
  const double mval = round_pixel(interp_scalar(pb.moving.g, ci, mfetch), pb.moving.f64, pb.opt->pixel_integer);
  if (pb.res->moving_value) pb.res->moving_value[lin] = mval;

  double grad[3] = {0, 0, 0};
  const bool deriv = pb.opt->compute_derivative != 0;
  if (deriv) {
    if (pb.opt->gradient_source == 1) {
      // Gradient image (CovariantVector<double,D>) through a linear interpolator (I2IM:407-425).
      for (int c = 0; c < D; ++c) {
        auto gfetch = [&](int64_t x, int64_t y, int64_t z) {
          const orc_geometry* g = pb.moving.g;
          const int64_t off = (x - g->start[0]) + g->size[0] * ((y - g->start[1]) + g->size[1] * (z - g->start[2]));
          return pb.mgrad[off * D + c];
        };
        grad[c] = interp_scalar(pb.moving.g, ci, gfetch);
      }
    } else {
      central_difference(pb.moving, mp, grad);
    }
  }

  // ProcessPoint (MattesT:238-245): intensity outside [TrueMin, TrueMax] is dropped.
  if (mval < pb.opt->moving_min) return;
  if (mval > pb.opt->moving_max) return;

  const double mterm = mval / pb.mBinSize - pb.mNormMin;
  int64_t mIdx = (int64_t)mterm;
  if (mIdx < 2) mIdx = 2;
  else if (mIdx > (int64_t)B - 3) mIdx = B - 3;
  int64_t pdfMovingIndex = mIdx - 1;
  const int64_t pdfMovingIndexMax = mIdx + 2;

  // ComputeSingleFixedImageParzenWindowIndex (Mattes:565-597).
  const double fterm = fval / pb.fBinSize - pb.fNormMin;
  int64_t fIdx = (int64_t)fterm;
  if (fIdx < 2) fIdx = 2;
  else if (fIdx > (int64_t)B - 3) fIdx = B - 3;

  ts.fmarg[fIdx] += 1;
  if (pb.res->fixed_bin) pb.res->fixed_bin[lin] = (int16_t)fIdx;
  if (pb.res->moving_bin) pb.res->moving_bin[lin] = (int16_t)mIdx;

  double arg = (double)pdfMovingIndex - mterm;
  double* pdfRow = ts.pdf.data() + fIdx * B;

  // Parameter Jacobian at the virtual point (MOTB:563-591, BSplTf:594-650).
  double v[3] = {0, 0, 0};
  if (deriv && pb.mtf->kind == ORC_TF_AFFINE)
    for (int d = 0; d < D; ++d) v[d] = vp[d] - pb.mtf->center[d];
  const int64_t pdim = pb.mtf->kind == ORC_TF_BSPLINE ? grid_nodes(&pb.mtf->grid) : 0;

  // Composite moving transform: ComputeJacobianWithRespectToParameters of the
  // CompositeTransform (itkCompositeTransform.hxx:467-560): the moving
  // transform's Jacobian at the virtual point, left-multiplied by the
  // position Jacobian (the matrix) of each earlier queue entry, last to
  // first; then the threader's inner product with the gradient
  // (MattesT:361-371).  Only the Jacobian's non-zero columns are formed (the
  // others stay zero through the products).
  std::vector<double> jcomp;  // [D][ncol]
  std::vector<int64_t> jcol;  // parameter index of each column
  if (deriv && pb.mtf->n_initial > 0) {
    if (pb.mtf->kind == ORC_TF_AFFINE) {
      for (int64_t c = 0; c < D * D + D; ++c) jcol.push_back(c);
    } else if (be.inside) {
      for (int k = 0; k < be.nw; ++k)
        for (int d = 0; d < D; ++d) jcol.push_back(be.node[k] + d * pdim);
    }
    const size_t nc = jcol.size();
    jcomp.assign((size_t)D * nc, 0.0);
    for (size_t c = 0; c < nc; ++c) {
      for (int r = 0; r < D; ++r) {
        double j = 0.0;
        if (pb.mtf->kind == ORC_TF_AFFINE) {
          const int64_t mu = jcol[c];
          if (mu < D * D) j = (mu / D == r) ? v[mu % D] : 0.0;
          else j = (mu - D * D == r) ? 1.0 : 0.0;
        } else {
          j = ((int)(c % D) == r) ? be.weights[c / D] : 0.0;
        }
        jcomp[r * nc + c] = j;
      }
    }
    for (int st = pb.mtf->n_initial - 1; st >= 0; --st) {
      const double* M = pb.mtf->initial[st].matrix;
      for (size_t c = 0; c < nc; ++c) {
        double temp[3];
        for (int r = 0; r < D; ++r) {
          temp[r] = 0.0;
          for (int k = 0; k < D; ++k) temp[r] += M[r * D + k] * jcomp[k * nc + c];
        }
        for (int r = 0; r < D; ++r) jcomp[r * nc + c] = temp[r];
      }
    }
  }

  while (pdfMovingIndex <= pdfMovingIndexMax) {
    pdfRow[pdfMovingIndex] += bspline3(arg);
    if (deriv) {
      const double db = bspline3_deriv(arg);
      double* dp = ts.dpdf.data() + (fIdx * B + pdfMovingIndex) * pb.P;
      if (pb.mtf->n_initial > 0) {
        const size_t nc = jcol.size();
        for (size_t c = 0; c < nc; ++c) {
          double inner = 0.0;
          for (int d = 0; d < D; ++d) inner += jcomp[d * nc + c] * grad[d];
          dp[jcol[c]] += inner * db;
        }
      } else if (pb.mtf->kind == ORC_TF_AFFINE) {
        // innerProduct = sum_dim J[dim][mu] * grad[dim]; only one row of J is
        // non-zero per mu, so the sum equals that single product exactly.
        for (int b = 0; b < D; ++b)
          for (int d = 0; d < D; ++d) dp[b * D + d] += (v[d] * grad[b]) * db;
        for (int d = 0; d < D; ++d) dp[D * D + d] += (1.0 * grad[d]) * db;
      } else if (pb.mtf->kind == ORC_TF_BSPLINE && be.inside) {
        for (int k = 0; k < be.nw; ++k)
          for (int d = 0; d < D; ++d) dp[be.node[k] + d * pdim] += (be.weights[k] * grad[d]) * db;
      }
    }
    arg += 1.0;
    ++pdfMovingIndex;
  }
  ts.nvalid++;
}

}  // namespace

extern "C" {

int64_t orc_itk_floor(double x) { return itk_floor(x); }
double orc_bspline3(double u) { return bspline3(u); }
double orc_bspline3_derivative(double u) { return bspline3_deriv(u); }

int orc_bin_parameters(double tmin, double tmax, int32_t bins, double* bin_size, double* norm_min, char* err,
                       int errlen) {
  // Mattes:250-317.
  const int padding = 2;
  if (tmax < tmin) {
    set_err(err, errlen, "No samples were observed in the analysis region.");
    return 1;
  }
  if (tmax - tmin <= std::numeric_limits<double>::epsilon()) {
    set_err(err, errlen, "have the same intensity value");
    return 1;
  }
  *bin_size = (tmax - tmin) / (double)(bins - 2 * padding);
  *norm_min = tmin / *bin_size - (double)padding;
  return 0;
}

int64_t orc_parzen_index(double value, double bin_size, double norm_min, int32_t bins) {
  const double term = value / bin_size - norm_min;
  int64_t p = (int64_t)term;
  if (p < 2) p = 2;
  else if (p > (int64_t)bins - 3) p = bins - 3;
  return p;
}

int orc_interpolate(const orc_geometry* g, const void* pixels, int32_t f64, const double* point, double* out) {
  double ci[3];
  point_to_cindex(g, point, ci);
  if (!inside_buffer(g, ci)) return 0;
  Image im{g, pixels, f64};
  auto fetch = [&](int64_t x, int64_t y, int64_t z) { return im.at(x, y, z); };
  *out = interp_scalar(g, ci, fetch);
  return 1;
}

void orc_transform_point(const orc_transform* tf, int32_t dim, const double* in, double* out) {
  tf_point(tf, dim, in, out, nullptr);
}

int64_t orc_itk_round(double x) { return itk_round(x); }

int orc_physical_point_to_index(const orc_geometry* g, const double* point, int64_t* idx) {
  return physical_point_to_index(g, point, idx) ? 1 : 0;
}

// MattesMI::Initialize (Mattes:97-220): running min / max from the pixel
// type's max and NonpositiveMin over the selected pixels.
int64_t orc_intensity_range(const orc_geometry* g, const void* pixels, int32_t pixel_f64,
                            const orc_geometry* mask_geom, const uint8_t* mask, const double* points,
                            int64_t n_points, const orc_transform* fixed_tf, double* lo, double* hi) {
  const int D = g->dim;
  const Image img{g, pixels, pixel_f64};
  double mn = pixel_f64 ? std::numeric_limits<double>::max() : (double)std::numeric_limits<float>::max();
  double mx = -mn;
  int64_t taken = 0;
  auto take = [&](const int64_t* idx) {
    const double v = img.at(idx[0], D > 1 ? idx[1] : 0, D > 2 ? idx[2] : 0);
    mn = mn < v ? mn : v;
    mx = mx > v ? mx : v;
    ++taken;
  };
  if (points) {
    // Mattes:106-133: virtual point -> fixed transform -> rounded index
    for (int64_t i = 0; i < n_points; ++i) {
      double fp[3] = {0, 0, 0};
      int64_t idx[3] = {0, 0, 0};
      tf_point(fixed_tf, D, points + i * D, fp, nullptr);
      if (physical_point_to_index(g, fp, idx)) take(idx);
    }
  } else {
    Mask m{mask_geom, mask};
    const int64_t sz = D == 3 ? g->size[2] : 1;
    int64_t idx[3] = {0, 0, 0};
    for (int64_t z = 0; z < sz; ++z)
      for (int64_t y = 0; y < g->size[1]; ++y)
        for (int64_t x = 0; x < g->size[0]; ++x) {
          idx[0] = g->start[0] + x;
          idx[1] = g->start[1] + y;
          idx[2] = D == 3 ? g->start[2] + z : 0;
          if (m.g) {  // Mattes:159-172, 195-207
            double p[3];
            index_to_point(g, idx, p);
            if (!m.inside(p)) continue;
          }
          take(idx);
        }
  }
  *lo = mn;
  *hi = mx;
  return taken;
}

}  // extern "C"

namespace {

// MersenneTwisterRandomVariateGenerator (MTGen.cxx): the state is seeded and
// reloaded at once (InitializeWithoutMutexLocking, :175-193), reload() walks
// the state with the two loops of :195-222 and twist() of :53-57, and
// GetIntegerVariate tempers (:226-240).
struct MtGen {
  uint32_t state[624];
  uint32_t* next = nullptr;
  int left = 0;

  static uint32_t twist(uint32_t m, uint32_t s0, uint32_t s1) {
    const uint32_t mix = (s0 & 0x80000000u) | (s1 & 0x7fffffffu);
    return m ^ (mix >> 1) ^ (uint32_t)(-(int32_t)(s1 & 1u) & 0x9908b0df);
  }
  void reload() {
    uint32_t* p = state;
    for (int i = 624 - 397; i--; ++p) *p = twist(p[397], p[0], p[1]);
    for (int i = 397; --i; ++p) *p = twist(p[397 - 624], p[0], p[1]);
    *p = twist(p[397 - 624], p[0], state[0]);
    left = 624;
    next = state;
  }
  void seed(uint32_t s) {
    state[0] = s;
    for (uint32_t i = 1; i < 624; ++i) state[i] = 1812433253u * (state[i - 1] ^ (state[i - 1] >> 30)) + i;
    reload();
  }
  uint32_t integer() {
    if (left == 0) reload();
    --left;
    uint32_t s1 = *next++;
    s1 ^= (s1 >> 11);
    s1 ^= (s1 << 7) & 0x9d2c5680u;
    s1 ^= (s1 << 15) & 0xefc60000u;
    return s1 ^ (s1 >> 18);
  }
  // MTGen.h:203-226
  double open_range() { return ((double)integer() + 0.5) / (double)(1ull << 32); }
  double open_upper_range() { return (double)integer() / (double)(1ull << 32); }
  // MTGen.h:364-375 (mean 0, variance 1)
  double normal() {
    const double mean = 0.0, variance = 1.0;
    const double r = std::sqrt(-2.0 * std::log(1.0 - open_range()) * variance);
    const double phi = 2.0 * 3.14159265358979323846 * open_upper_range();
    return mean + r * std::cos(phi);
  }
};

}  // namespace

extern "C" {

void orc_mt_integers(uint32_t seed, uint32_t* out, int64_t n) {
  MtGen g;
  g.seed(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g.integer();
}

void orc_mt_normals(uint32_t seed, double* out, int64_t n) {
  MtGen g;
  g.seed(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g.normal();
}

// RegMethod:867-1024 for a single image metric.
int64_t orc_metric_sample_points(const orc_geometry* virt, int32_t strategy, double percentage, uint32_t* seed,
                                 const orc_geometry* mask_geom, const uint8_t* mask, double* out) {
  if (!virt || !seed || (strategy != 1 && strategy != 2) || !(percentage > 0.0 && percentage <= 1.0)) return -1;
  const int D = virt->dim;
  int64_t npix = 1;
  for (int j = 0; j < D; ++j) npix *= virt->size[j];
  double oneThird[3];
  for (int d = 0; d < D; ++d) oneThird[d] = virt->spacing[d] / 3.0;
  MtGen randomizer;
  randomizer.seed((*seed)++);
  Mask fmask{mask_geom, mask};
  int64_t index = 0;
  auto add = [&](const int64_t* idx) {
    double point[3];
    index_to_point(virt, idx, point);
    for (int d = 0; d < D; ++d) point[d] += randomizer.normal() * oneThird[d];
    if (!fmask.g || fmask.inside(point)) {
      for (int d = 0; d < D; ++d) out[index * D + d] = point[d];
      ++index;
    }
  };
  if (strategy == 1) {
    const auto sampleCount = (unsigned long)std::ceil(1.0 / percentage);
    unsigned long count = sampleCount;
    int64_t idx[3] = {0, 0, 0};
    const int64_t sz = D == 3 ? virt->size[2] : 1;
    for (int64_t z = 0; z < sz; ++z)
      for (int64_t y = 0; y < virt->size[1]; ++y)
        for (int64_t x = 0; x < virt->size[0]; ++x) {
          if (count == sampleCount) {
            count = 0;
            idx[0] = virt->start[0] + x;
            idx[1] = virt->start[1] + y;
            idx[2] = D == 3 ? virt->start[2] + z : 0;
            add(idx);
          }
          ++count;
        }
  } else {
    const auto sampleCount = (unsigned long)((float)npix * percentage);
    MtGen iter;  // ImageRandomConstIteratorWithIndex's generator, reseeded
    iter.seed((uint32_t)(int32_t)(*seed)++);
    // GoToBegin / operator++ each jump (RandIt.hxx:47-68, RandIt.h:161-197)
    auto jump = [&](int64_t* idx) {
      auto position = (int64_t)(iter.open_range() * ((double)npix - 0.5));
      for (int d = 0; d < D; ++d) {
        const int64_t residual = position % virt->size[d];
        idx[d] = residual + virt->start[d];
        position -= residual;
        position /= virt->size[d];
      }
    };
    int64_t idx[3] = {0, 0, 0};
    jump(idx);
    for (unsigned long done = 0; done < sampleCount; ++done) {
      add(idx);
      jump(idx);
    }
  }
  return index;
}

int orc_mattes_evaluate(const orc_geometry* virt, const orc_geometry* fixed_geom, const void* fixed_pixels,
                        const orc_geometry* moving_geom, const void* moving_pixels, const double* moving_gradient,
                        const orc_transform* fixed_tf, const orc_transform* moving_tf, const orc_options* opt,
                        orc_result* res, char* err, int errlen) {
  return orc_mattes_evaluate_ex(virt, fixed_geom, fixed_pixels, moving_geom, moving_pixels, moving_gradient, fixed_tf,
                                moving_tf, opt, nullptr, res, err, errlen);
}

int orc_mattes_evaluate_ex(const orc_geometry* virt, const orc_geometry* fixed_geom, const void* fixed_pixels,
                           const orc_geometry* moving_geom, const void* moving_pixels,
                           const double* moving_gradient, const orc_transform* fixed_tf,
                           const orc_transform* moving_tf, const orc_options* opt, const orc_sampling* smp,
                           orc_result* res, char* err, int errlen) {
  const int D = virt->dim;
  const int B = opt->bins;
  if (D != 2 && D != 3) {
    set_err(err, errlen, "dimension must be 2 or 3");
    return 1;
  }
  if (B < 5) {
    set_err(err, errlen, "at least 5 histogram bins are required");
    return 1;
  }
  Problem pb;
  pb.virt = virt;
  pb.fixed = Image{fixed_geom, fixed_pixels, opt->pixel_f64};
  pb.moving = Image{moving_geom, moving_pixels, opt->pixel_f64};
  pb.mgrad = moving_gradient;
  pb.ftf = fixed_tf;
  pb.mtf = moving_tf;
  pb.opt = opt;
  pb.res = res;
  const bool sparse = smp && smp->points;
  if (smp) {
    pb.fmask = Mask{smp->fixed_mask_geom, smp->fixed_mask};
    pb.mmask = Mask{smp->moving_mask_geom, smp->moving_mask};
  }
  if (orc_bin_parameters(opt->fixed_min, opt->fixed_max, B, &pb.fBinSize, &pb.fNormMin, err, errlen)) return 1;
  if (orc_bin_parameters(opt->moving_min, opt->moving_max, B, &pb.mBinSize, &pb.mNormMin, err, errlen)) return 1;
  pb.P = tf_num_params(moving_tf, D);
  const bool deriv = opt->compute_derivative != 0;
  if (deriv && opt->gradient_source == 1 && !moving_gradient) {
    set_err(err, errlen, "gradient image required");
    return 1;
  }

  int64_t N = 1;
  for (int j = 0; j < D; ++j) N *= virt->size[j];
  if (sparse) N = smp->n_points;
  if (res->fixed_bin)
    for (int64_t i = 0; i < N; ++i) res->fixed_bin[i] = -1;
  if (res->moving_bin)
    for (int64_t i = 0; i < N; ++i) res->moving_bin[i] = -1;
  if (res->moving_value)
    for (int64_t i = 0; i < N; ++i) res->moving_value[i] = std::numeric_limits<double>::quiet_NaN();

  // Slabs along the slowest axis (ThreadedImageRegionPartitioner), or
  // contiguous runs of the point set (ThreadedIndexedContainerPartitioner).
  const int64_t slow = sparse ? N : virt->size[D - 1];
  int T = std::max(1, opt->threads);
  if ((int64_t)T > slow) T = (int)slow;
  const double dpdfBytes = deriv ? (double)B * B * (double)pb.P * 8.0 : 0.0;
  while (T > 1 && dpdfBytes * T > 4e9) --T;
  std::vector<ThreadState> ts(T);
  for (auto& s : ts) {
    s.pdf.assign((size_t)B * B, 0.0);
    s.fmarg.assign(B, 0.0);
    if (deriv) s.dpdf.assign((size_t)B * B * (size_t)pb.P, 0.0);
  }
  auto work = [&](int t) {
    const int64_t z0 = slow * t / T, z1 = slow * (t + 1) / T;
    int64_t idx[3];
    double vp[3];
    const int64_t sx = virt->size[0], sy = D == 3 ? virt->size[1] : 1;
    if (sparse) {
      // sparse threader (Thr:44-63): the listed point, in list order
      for (int64_t i = z0; i < z1; ++i) {
        for (int k = 0; k < D; ++k) vp[k] = smp->points[i * D + k];
        process_sample(pb, vp, i, ts[t]);
      }
      return;
    }
    for (int64_t s = z0; s < z1; ++s) {
      if (D == 2) {
        for (int64_t x = 0; x < sx; ++x) {
          idx[0] = virt->start[0] + x;
          idx[1] = virt->start[1] + s;
          index_to_point(virt, idx, vp);
          process_sample(pb, vp, x + sx * s, ts[t]);
        }
      } else {
        for (int64_t y = 0; y < sy; ++y)
          for (int64_t x = 0; x < sx; ++x) {
            idx[0] = virt->start[0] + x;
            idx[1] = virt->start[1] + y;
            idx[2] = virt->start[2] + s;
            index_to_point(virt, idx, vp);
            process_sample(pb, vp, x + sx * (y + sy * s), ts[t]);
          }
      }
    }
  };
  if (T == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }

  // AfterThreadedExecution (MattesT:413-463) + GetValueCommonAfterThreadedExecution (Mattes:507-542).
  int64_t nvalid = 0;
  for (auto& s : ts) nvalid += s.nvalid;
  std::vector<double>& pdf = ts[0].pdf;
  std::vector<double>& fm = ts[0].fmarg;
  for (int t = 1; t < T; ++t) {
    for (size_t i = 0; i < pdf.size(); ++i) pdf[i] += ts[t].pdf[i];
    for (int i = 0; i < B; ++i) fm[i] += ts[t].fmarg[i];
  }
  // CompensatedSummation (Kahan:38-48).
  double ksum = 0.0, kcomp = 0.0;
  for (size_t i = 0; i < pdf.size(); ++i) {
    const double in = pdf[i] - kcomp;
    const double tmp = ksum + in;
    kcomp = (tmp - ksum) - in;
    ksum = tmp;
  }
  const double jointPDFSum = ksum;
  res->n_valid = nvalid;
  res->joint_pdf_sum = jointPDFSum;

  std::vector<double>& dpdf = ts[0].dpdf;
  if (deriv) {
    for (int t = 1; t < T; ++t)
      for (size_t i = 0; i < dpdf.size(); ++i) dpdf[i] += ts[t].dpdf[i];
    if (nvalid > 0) {
      const double nFactor = -1.0 / (pb.mBinSize * (double)nvalid);
      for (auto& x : dpdf) x *= nFactor;
    }
  }

  // ComputeResults (Mattes:353-499).
  if (nvalid == 0) {
    set_err(err, errlen,
            "All samples map outside moving image buffer. The images do not sufficiently overlap. They need to be "
            "initialized to have more overlap before this metric will work. For instance, you can align the image "
            "centers by translation.");
    return 1;
  }
  if (jointPDFSum < std::numeric_limits<double>::epsilon()) {
    set_err(err, errlen, "Joint PDF summed to zero");
    return 1;
  }
  const double normalizationFactor = 1.0 / jointPDFSum;
  for (auto& x : pdf) x = x * normalizationFactor;
  for (auto& x : fm) x = x * normalizationFactor;
  std::vector<double> mm(B, 0.0);
  for (int m = 0; m < B; ++m) {
    double acc = 0.0;
    for (int f = 0; f < B; ++f) acc += pdf[(size_t)f * B + m];
    mm[m] = acc;
  }
  const double closeToZero = std::numeric_limits<double>::epsilon();
  std::vector<double> der(deriv ? pb.P : 0, 0.0);
  double sum = 0.0;
  for (int f = 0; f < B; ++f) {
    const double fmv = fm[f];
    if (fmv > closeToZero) {
      const double logf = std::log(fmv);
      for (int m = 0; m < B; ++m) {
        const double mmv = mm[m];
        const double jp = pdf[(size_t)f * B + m];
        if (mmv > closeToZero && jp > closeToZero) {
          const double pRatio = std::log(jp / mmv);
          sum += jp * (pRatio - logf);
          if (deriv) {
            const double* dp = dpdf.data() + ((size_t)f * B + m) * pb.P;
            for (int64_t mu = 0; mu < pb.P; ++mu) der[mu] += dp[mu] * pRatio;
          }
        }
      }
    }
  }
  res->value = -1.0 * sum;
  if (deriv && res->derivative)
    for (int64_t mu = 0; mu < pb.P; ++mu) res->derivative[mu] = der[mu];
  if (res->joint_pdf) std::memcpy(res->joint_pdf, pdf.data(), sizeof(double) * B * B);
  if (res->fixed_marginal) std::memcpy(res->fixed_marginal, fm.data(), sizeof(double) * B);
  if (res->moving_marginal) std::memcpy(res->moving_marginal, mm.data(), sizeof(double) * B);
  return 0;
}

// ---------------------------------------------------------------------------
// Recursive Gaussian (Deriche-style IIR with the Farneback 2006 parameters).
// ---------------------------------------------------------------------------

struct RecCoeffs {
  double N0, N1, N2, N3, D1, D2, D3, D4, M1, M2, M3, M4, BN1, BN2, BN3, BN4, BM1, BM2, BM3, BM4;
};

// RecursiveGaussianImageFilter::SetUp (RecGauss:56-190) for orders 0 and 1,
// ComputeNCoefficients (RecGauss:193-233), ComputeDCoefficients (:237-262),
// ComputeRemainingCoefficients (:266-300).  The parameter table is the
// published one (Farneback & Westin 2006, Table 3) and must be identical.
static int rec_setup(double sigma, double spacing, int order, bool normalizeAcrossScale, RecCoeffs* c) {
  double direction = 1.0;
  if (spacing < 0.0) {
    direction = -1.0;
    spacing *= -1;
  }
  if (spacing < 1e-8) return 1;
  const double sigmad = sigma / spacing;
  double across = 1.0;
  const double A1[3] = {1.3530, -0.6724, -1.3563};
  const double B1[3] = {1.8151, -3.4327, 5.2318};
  const double W1 = 0.6681, L1 = -1.3932;
  const double A2[3] = {-0.3531, 0.6724, 0.3446};
  const double B2[3] = {0.0902, 0.6100, -2.2355};
  const double W2 = 2.0787, L2 = -1.3732;
  double SD, DD, ED;
  {
    const double Cos1 = std::cos(W1 / sigmad), Cos2 = std::cos(W2 / sigmad);
    const double Exp1 = std::exp(L1 / sigmad), Exp2 = std::exp(L2 / sigmad);
    c->D4 = Exp1 * Exp1 * Exp2 * Exp2;
    c->D3 = -2 * Cos1 * Exp1 * Exp2 * Exp2;
    c->D3 += -2 * Cos2 * Exp2 * Exp1 * Exp1;
    c->D2 = 4 * Cos2 * Cos1 * Exp1 * Exp2;
    c->D2 += Exp1 * Exp1 + Exp2 * Exp2;
    c->D1 = -2 * (Exp2 * Cos2 + Exp1 * Cos1);
    SD = 1.0 + c->D1 + c->D2 + c->D3 + c->D4;
    DD = c->D1 + 2 * c->D2 + 3 * c->D3 + 4 * c->D4;
    ED = c->D1 + 4 * c->D2 + 9 * c->D3 + 16 * c->D4;
  }
  (void)ED;
  auto ncoef = [&](double a1, double b1, double a2, double b2, double& SN, double& DN) {
    const double Sin1 = std::sin(W1 / sigmad), Sin2 = std::sin(W2 / sigmad);
    const double Cos1 = std::cos(W1 / sigmad), Cos2 = std::cos(W2 / sigmad);
    const double Exp1 = std::exp(L1 / sigmad), Exp2 = std::exp(L2 / sigmad);
    c->N0 = a1 + a2;
    c->N1 = Exp2 * (b2 * Sin2 - (a2 + 2 * a1) * Cos2);
    c->N1 += Exp1 * (b1 * Sin1 - (a1 + 2 * a2) * Cos1);
    c->N2 = (a1 + a2) * Cos2 * Cos1;
    c->N2 -= b1 * Cos2 * Sin1 + b2 * Cos1 * Sin2;
    c->N2 *= 2 * Exp1 * Exp2;
    c->N2 += a2 * Exp1 * Exp1 + a1 * Exp2 * Exp2;
    c->N3 = Exp2 * Exp1 * Exp1 * (b2 * Sin2 - a2 * Cos2);
    c->N3 += Exp1 * Exp2 * Exp2 * (b1 * Sin1 - a1 * Cos1);
    SN = c->N0 + c->N1 + c->N2 + c->N3;
    DN = c->N1 + 2 * c->N2 + 3 * c->N3;
  };
  double SN, DN;
  bool symmetric;
  if (order == 0) {
    ncoef(A1[0], B1[0], A2[0], B2[0], SN, DN);
    const double alpha0 = 2 * SN / SD - c->N0;
    c->N0 *= across / alpha0;
    c->N1 *= across / alpha0;
    c->N2 *= across / alpha0;
    c->N3 *= across / alpha0;
    symmetric = true;
  } else if (order == 1) {
    if (normalizeAcrossScale) across = sigma;
    ncoef(A1[1], B1[1], A2[1], B2[1], SN, DN);
    double alpha1 = 2 * (SN * DD - DN * SD) / (SD * SD);
    alpha1 *= direction;
    c->N0 *= across / alpha1;
    c->N1 *= across / alpha1;
    c->N2 *= across / alpha1;
    c->N3 *= across / alpha1;
    symmetric = false;
  } else {
    return 1;
  }
  if (symmetric) {
    c->M1 = c->N1 - c->D1 * c->N0;
    c->M2 = c->N2 - c->D2 * c->N0;
    c->M3 = c->N3 - c->D3 * c->N0;
    c->M4 = -c->D4 * c->N0;
  } else {
    c->M1 = -(c->N1 - c->D1 * c->N0);
    c->M2 = -(c->N2 - c->D2 * c->N0);
    c->M3 = -(c->N3 - c->D3 * c->N0);
    c->M4 = c->D4 * c->N0;
  }
  const double SN2 = c->N0 + c->N1 + c->N2 + c->N3;
  const double SM = c->M1 + c->M2 + c->M3 + c->M4;
  const double SD2 = 1.0 + c->D1 + c->D2 + c->D3 + c->D4;
  c->BN1 = c->D1 * SN2 / SD2;
  c->BN2 = c->D2 * SN2 / SD2;
  c->BN3 = c->D3 * SN2 / SD2;
  c->BN4 = c->D4 * SN2 / SD2;
  c->BM1 = c->D1 * SM / SD2;
  c->BM2 = c->D2 * SM / SD2;
  c->BM3 = c->D3 * SM / SD2;
  c->BM4 = c->D4 * SM / SD2;
  return 0;
}

// RecursiveSeparableImageFilter::FilterDataArray (RecSep:62-143).
static void filter_line(const RecCoeffs& c, double* outs, const double* data, double* scratch, int64_t ln) {
  auto E = [](double& o, double a1, double b1, double a2, double b2, double a3, double b3, double a4, double b4) {
    o = a1 * b1 + a2 * b2 + a3 * b3 + a4 * b4;
  };
  auto S = [](double& o, double a1, double b1, double a2, double b2, double a3, double b3, double a4, double b4) {
    o -= a1 * b1 + a2 * b2 + a3 * b3 + a4 * b4;
  };
  const double v1 = data[0];
  E(outs[0], v1, c.N0, v1, c.N1, v1, c.N2, v1, c.N3);
  E(outs[1], data[1], c.N0, v1, c.N1, v1, c.N2, v1, c.N3);
  E(outs[2], data[2], c.N0, data[1], c.N1, v1, c.N2, v1, c.N3);
  E(outs[3], data[3], c.N0, data[2], c.N1, data[1], c.N2, v1, c.N3);
  S(outs[0], v1, c.BN1, v1, c.BN2, v1, c.BN3, v1, c.BN4);
  S(outs[1], outs[0], c.D1, v1, c.BN2, v1, c.BN3, v1, c.BN4);
  S(outs[2], outs[1], c.D1, outs[0], c.D2, v1, c.BN3, v1, c.BN4);
  S(outs[3], outs[2], c.D1, outs[1], c.D2, outs[0], c.D3, v1, c.BN4);
  for (int64_t i = 4; i < ln; ++i) {
    E(outs[i], data[i], c.N0, data[i - 1], c.N1, data[i - 2], c.N2, data[i - 3], c.N3);
    S(outs[i], outs[i - 1], c.D1, outs[i - 2], c.D2, outs[i - 3], c.D3, outs[i - 4], c.D4);
  }
  const double v2 = data[ln - 1];
  E(scratch[ln - 1], v2, c.M1, v2, c.M2, v2, c.M3, v2, c.M4);
  E(scratch[ln - 2], data[ln - 1], c.M1, v2, c.M2, v2, c.M3, v2, c.M4);
  E(scratch[ln - 3], data[ln - 2], c.M1, data[ln - 1], c.M2, v2, c.M3, v2, c.M4);
  E(scratch[ln - 4], data[ln - 3], c.M1, data[ln - 2], c.M2, data[ln - 1], c.M3, v2, c.M4);
  S(scratch[ln - 1], v2, c.BM1, v2, c.BM2, v2, c.BM3, v2, c.BM4);
  S(scratch[ln - 2], scratch[ln - 1], c.D1, v2, c.BM2, v2, c.BM3, v2, c.BM4);
  S(scratch[ln - 3], scratch[ln - 2], c.D1, scratch[ln - 1], c.D2, v2, c.BM3, v2, c.BM4);
  S(scratch[ln - 4], scratch[ln - 3], c.D1, scratch[ln - 2], c.D2, scratch[ln - 1], c.D3, v2, c.BM4);
  for (int64_t i = ln - 4; i > 0; i--) {
    E(scratch[i - 1], data[i], c.M1, data[i + 1], c.M2, data[i + 2], c.M3, data[i + 3], c.M4);
    S(scratch[i - 1], scratch[i], c.D1, scratch[i + 1], c.D2, scratch[i + 2], c.D3, scratch[i + 3], c.D4);
  }
  for (int64_t i = 0; i < ln; ++i) outs[i] += scratch[i];
}

// Filter every line of a double volume along one axis, in place.
static void filter_axis(const orc_geometry* g, double* vol, int axis, const RecCoeffs& c, int threads) {
  const int D = g->dim;
  int64_t sz[3] = {g->size[0], D > 1 ? g->size[1] : 1, D > 2 ? g->size[2] : 1};
  const int64_t ln = sz[axis];
  int64_t stride = 1;
  for (int j = 0; j < axis; ++j) stride *= sz[j];
  const int64_t nlines = sz[0] * sz[1] * sz[2] / ln;
  auto work = [&](int t, int T) {
    std::vector<double> in(ln), out(ln), scr(ln);
    for (int64_t l = t; l < nlines; l += T) {
      // line l enumerates the other two axes; base offset:
      const int64_t lo = l % stride, hi = l / stride;
      const int64_t base = lo + hi * stride * ln;
      for (int64_t i = 0; i < ln; ++i) in[i] = vol[base + i * stride];
      filter_line(c, out.data(), in.data(), scr.data(), ln);
      for (int64_t i = 0; i < ln; ++i) vol[base + i * stride] = out[i];
    }
  };
  const int T = std::max(1, threads);
  if (T == 1) {
    work(0, 1);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(work, t, T);
    for (auto& x : th) x.join();
  }
}

static double axis_spacing(const orc_geometry* g, int axis) { return g->spacing[axis]; }

int orc_gradient_recursive_gaussian(const orc_geometry* g, const void* pixels, int32_t f64, double sigma,
                                    double* out, int32_t threads) {
  // GradientRecursiveGaussianImageFilter::GenerateData (GradRG:159-280):
  // derivative along `dim`, zero-order smoothing along the other axes in
  // increasing order, divide by spacing[dim], then rotate by the direction.
  const int D = g->dim;
  int64_t n = 1;
  for (int j = 0; j < D; ++j) {
    n *= g->size[j];
    if (g->size[j] < 4) return 1;
  }
  std::vector<double> vol(n);
  for (int dim = 0; dim < D; ++dim) {
    for (int64_t i = 0; i < n; ++i) vol[i] = f64 ? ((const double*)pixels)[i] : (double)((const float*)pixels)[i];
    RecCoeffs c;
    if (rec_setup(sigma, axis_spacing(g, dim), 1, true, &c)) return 1;
    filter_axis(g, vol.data(), dim, c, threads);
    for (int other = 0; other < D; ++other) {
      if (other == dim) continue;
      RecCoeffs c0;
      if (rec_setup(sigma, axis_spacing(g, other), 0, true, &c0)) return 1;
      filter_axis(g, vol.data(), other, c0, threads);
    }
    const double sp = axis_spacing(g, dim);
    for (int64_t i = 0; i < n; ++i) out[i * D + dim] = vol[i] / sp;
  }
  // TransformLocalVectorToPhysicalVector (ImgBase:636-653).
  const double* dir = g->direction;
  for (int64_t i = 0; i < n; ++i) {
    double in[3], o[3];
    for (int j = 0; j < D; ++j) in[j] = out[i * D + j];
    for (int r = 0; r < D; ++r) {
      double s = 0.0;
      for (int cc = 0; cc < D; ++cc) s += dir[r * D + cc] * in[cc];
      o[r] = s;
    }
    for (int j = 0; j < D; ++j) out[i * D + j] = o[j];
  }
  return 0;
}

// SmoothingRecursiveGaussianImageFilter (itkSmoothingRecursiveGaussianImageFilter.hxx:30-70,
// 119-131, 199-260): the first filter runs along the last axis, then one
// filter per axis 0 .. dim-2, each zero order with that axis' sigma; the
// mini-pipeline is FP64 and the cast to the output pixel type comes last.
int orc_smooth_recursive_gaussian(const orc_geometry* g, const void* pixels, int32_t f64, const double* sigma,
                                  void* out, int32_t threads) {
  const int D = g->dim;
  int64_t n = 1;
  for (int j = 0; j < D; ++j) {
    n *= g->size[j];
    if (g->size[j] < 4) return 1;
  }
  std::vector<double> vol(n);
  for (int64_t i = 0; i < n; ++i) vol[i] = f64 ? ((const double*)pixels)[i] : (double)((const float*)pixels)[i];
  for (int step = 0; step < D; ++step) {
    const int axis = step == 0 ? D - 1 : step - 1;
    RecCoeffs c;
    if (rec_setup(sigma[axis], axis_spacing(g, axis), 0, false, &c)) return 1;
    filter_axis(g, vol.data(), axis, c, threads);
  }
  for (int64_t i = 0; i < n; ++i) {
    if (f64) ((double*)out)[i] = vol[i];
    else ((float*)out)[i] = (float)vol[i];
  }
  return 0;
}

}  // extern "C"
