// Per-sample evaluators shared by the exact and fast Mattes MI kernels.
#pragma once

#include "mattes_kernels.h"

namespace mmi {

constexpr double kFixedPointScale = 4294967296.0;  // 2^32

template <int D>
struct SampleOut {
  int fbin;
  int mbin;
  double mterm;
  double vp[3];  // virtual point (the Jacobian is taken there)
  double mp[3];  // mapped moving point
  Cell cell;     // moving interpolation cell
  BSplineSupport<D> sup;
};

// Fixed transform: identity or affine.
template <int D>
__device__ __forceinline__ void fixed_point(const DevTransform& t, const double* in, double* out) {
  if (t.kind == 1) {
    transform_point<D, 1>(t, in, out, nullptr);
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) out[i] = in[i];
  }
}

// Virtual point of sample s: a listed point (sparse threader,
// itkImageToImageMetricv4GetValueAndDerivativeThreader.hxx:44-63) or the
// physical point of voxel s (dense threader, same file :26-42).
template <int D>
__device__ __forceinline__ void virtual_point(const EvalArgs& a, uint32_t s, double* vp) {
  if (a.points) {
#pragma unroll
    for (int k = 0; k < D; ++k) vp[k] = a.points[(int64_t)s * D + k];
  } else {
    int32_t idx[3];
    sample_index<D>(a.virt, s, idx);
    index_to_point<D>(a.virt, idx, vp);
  }
}

// The interpolator's double assigned to the image's pixel type
// (itkImageToImageMetricv4.hxx:318,369): a float rounds, a double keeps it, an
// integer type truncates toward zero.
__device__ __forceinline__ double pixel_value(int32_t kind, double v) {
  return kind == kPixelDouble ? v : (kind == kPixelInteger ? trunc(v) : (double)(float)v);
}

// TransformAndEvaluateFixedPoint (itkImageToImageMetricv4.hxx:289-320),
// exact FP64: false when the fixed point is masked out or outside the fixed
// buffer, else the fixed value assigned to the pixel type.
template <int D, typename P>
__device__ __forceinline__ bool fixed_value_exact(const EvalArgs& a, const double* vp, double* fval) {
  double fp[3], ci[3];
  fixed_point<D>(a.ftf, vp, fp);
  // the fixed mask at the mapped fixed point (itkImageToImageMetricv4.hxx:298-309)
  if (a.fixed_mask && !mask_inside<D>(*a.fixed_mask, fp)) return false;
  point_to_cindex<D>(a.fixed, fp, ci);
  if (!inside_buffer<D>(a.fixed, ci)) return false;
  const Cell fc = make_cell<D>(a.fixed, ci);
  *fval = pixel_value(a.fixed_kind, interp_exact<D, P>(static_cast<const typename P::pixel*>(a.fixed_px), fc));
  return true;
}

// The fixed Parzen window index of a fixed value
// (ComputeSingleFixedImageParzenWindowIndex,
// itkMattesMutualInformationImageToImageMetricv4.hxx:571-593, called at
// itkMattesMutualInformationImageToImageMetricv4GetValueAndDerivativeThreader.hxx:269-270).
__device__ __forceinline__ int fixed_bin_exact(const EvalArgs& a, double fval) {
  const double fterm = fval / a.fixed_bin_size - a.fixed_norm_min;
  return fterm < 2.0 ? 2 : (fterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)fterm);
}

// ProcessVirtualPoint + Mattes ProcessPoint up to the bins
// (itkImageToImageMetricv4GetValueAndDerivativeThreaderBase.hxx:174-273,
// itkImageToImageMetricv4.hxx:289-373,
// itkMattesMutualInformationImageToImageMetricv4GetValueAndDerivativeThreader.hxx:215-275),
// exact FP64.  MK is the moving transform kind (1 affine, 2 B-spline); P the
// image storage format.
template <int D, int MK, typename P>
__device__ __forceinline__ bool eval_sample(const EvalArgs& a, const double* vp, SampleOut<D>& o) {
#pragma unroll
  for (int k = 0; k < D; ++k) o.vp[k] = vp[k];
  double ci[3];
  double fval;
  if (!fixed_value_exact<D, P>(a, o.vp, &fval)) return false;

  transform_point<D, MK>(a.mtf, o.vp, o.mp, &o.sup);
  // CompositeTransform::TransformPoint: the queue's earlier transforms after
  // the moving one, last to first (itkCompositeTransform.hxx:60-71), each as
  // MatrixOffsetTransformBase::TransformPoint (matrix * point + offset)
  for (int st = a.n_init - 1; st >= 0; --st) {
    const double* m = a.init_stages + 12 * st;
    double q[3];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double sum = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) sum += m[r * D + c] * o.mp[c];
      q[r] = sum + m[9 + r];
    }
#pragma unroll
    for (int r = 0; r < D; ++r) o.mp[r] = q[r];
  }
  // the moving mask at the mapped moving point (itkImageToImageMetricv4.hxx:351-362)
  if (a.moving_mask && !mask_inside<D>(*a.moving_mask, o.mp)) return false;
  point_to_cindex<D>(a.moving, o.mp, ci);
  if (!inside_buffer<D>(a.moving, ci)) return false;
  o.cell = make_cell<D>(a.moving, ci);
  const double mval =
      pixel_value(a.moving_kind, interp_exact<D, P>(static_cast<const typename P::pixel*>(a.moving_px), o.cell));

  if (mval < a.moving_min) return false;
  if (mval > a.moving_max) return false;
  const double mterm = mval / a.moving_bin_size - a.moving_norm_min;
  const int mb = mterm < 2.0 ? 2 : (mterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)mterm);
  o.fbin = fixed_bin_exact(a, fval);
  o.mbin = mb;
  o.mterm = mterm;
  return true;
}

// Central differences of the moving image at a mapped point
// (CentralDifferenceImageFunction::EvaluateSpecialized,
// Modules/Core/ImageFunction/include/itkCentralDifferenceImageFunction.hxx:252-318).
template <int D, typename P>
__device__ __forceinline__ void central_difference(const EvalArgs& a, const double* p, double* grad) {
  const typename P::pixel* img = static_cast<const typename P::pixel*>(a.moving_px);
  double n1[3], n2[3];
#pragma unroll
  for (int i = 0; i < D; ++i) n1[i] = n2[i] = p[i];
#pragma unroll
  for (int dim = 0; dim < D; ++dim) {
    const double offset = 0.5 * a.moving.spacing[dim];
    double c1[3], c2[3];
    n1[dim] = p[dim] - offset;
    point_to_cindex<D>(a.moving, n1, c1);
    bool ok = inside_buffer<D>(a.moving, c1);
    if (ok) {
      n2[dim] = p[dim] + offset;
      point_to_cindex<D>(a.moving, n2, c2);
      ok = inside_buffer<D>(a.moving, c2);
    }
    if (!ok) {
      grad[dim] = 0.0;
    } else {
      const double delta = n2[dim] - n1[dim];
      if (delta > 10.0 * 2.220446049250313e-16) {
        const double v2 = interp_exact<D, P>(img, make_cell<D>(a.moving, c2));
        const double v1 = interp_exact<D, P>(img, make_cell<D>(a.moving, c1));
        grad[dim] = (v2 - v1) / delta;
      } else {
        grad[dim] = 0.0;
      }
    }
    n1[dim] = p[dim];
    n2[dim] = p[dim];
  }
}

// ---------------------------------------------------------------------------
// Parzen weights
// ---------------------------------------------------------------------------

__device__ __forceinline__ float bspline3f(float u) {
  const float a = fabsf(u);
  if (a < 1.0f) return fmaf(a * a, fmaf(0.5f, a, -1.0f), 2.0f / 3.0f);
  if (a < 2.0f) {
    const float t = 2.0f - a;
    return t * t * t * (1.0f / 6.0f);
  }
  return 0.0f;
}

__device__ __forceinline__ float bspline3_derivf(float u) {
  if (u >= 0.0f && u < 1.0f) return u * fmaf(1.5f, u, -2.0f);
  if (u > -1.0f && u < 0.0f) return -u * fmaf(1.5f, u, 2.0f);
  if (u >= 1.0f && u < 2.0f) return fmaf(u, fmaf(-0.5f, u, 2.0f), -2.0f);
  if (u > -2.0f && u <= -1.0f) return fmaf(u, fmaf(0.5f, u, 2.0f), 2.0f);
  return 0.0f;
}

// The four Parzen weights B3(arg0 + k) as 2^32 fixed point, summing to exactly
// 2^32 (partition of unity; the rounding remainder goes to weight 1, which is
// >= 1/6 whenever the window is not clamped).
__device__ __forceinline__ void parzen_fixed(float arg0, unsigned int* q) {
  float w[4];
  float arg = arg0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    w[k] = bspline3f(arg);
    arg += 1.0f;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = (unsigned int)rintf(w[k] * 4294967296.0f);
  // arg0 in (-2, -1]: unclamped window, the weights are a partition of unity
  if (arg0 <= -1.0f && arg0 > -2.0f) q[1] = 0u - q[0] - q[2] - q[3];
}

// sum_k B3'(arg0 + k) * pRatio[f][m - 1 + k] (MattesT:329-378 arguments).
__device__ __forceinline__ float parzen_deriv_weight(float arg0, const float* prow) {
  float w = 0.0f, arg = arg0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    w = fmaf(bspline3_derivf(arg), prow[k], w);
    arg += 1.0f;
  }
  return w;
}

// Unclamped Parzen window (arg0 = -1 - t, t in [0, 1)): the four cubic
// B-spline weights and derivatives in closed form, branch-free.  They equal
// B3(arg0 + k) and B3'(arg0 + k) of the piecewise definitions above.
__device__ __forceinline__ void parzen_weights_t(float t, float* w) {
  const float u = 1.0f - t;
  const float t2 = t * t, t3 = t2 * t;
  w[0] = u * u * u * (1.0f / 6.0f);
  w[1] = fmaf(t2, fmaf(0.5f, t, -1.0f), 2.0f / 3.0f);  // (3t^3 - 6t^2 + 4) / 6
  w[2] = fmaf(u * u, fmaf(0.5f, u, -1.0f), 2.0f / 3.0f);  // B3(1 - t)
  w[3] = t3 * (1.0f / 6.0f);
}

__device__ __forceinline__ float parzen_deriv_weight_t(float t, const float* prow) {
  const float u = 1.0f - t;
  const float d0 = 0.5f * u * u;
  const float d1 = t * fmaf(-1.5f, t, 2.0f);
  const float d2 = u * fmaf(1.5f, u, -2.0f);
  const float d3 = -0.5f * t * t;
  return fmaf(d0, prow[0], fmaf(d1, prow[1], fmaf(d2, prow[2], d3 * prow[3])));
}

// Fixed-point Parzen weights for either window: closed form when unclamped
// (the FP32 products w 2^32 of weights above 2^-8 are integers already, so
// the conversion truncates without rounding first: below that it moves a
// weight by less than 2^-32).
__device__ __forceinline__ void parzen_fixed_fast(float arg0, unsigned int* q) {
  if (arg0 <= -1.0f && arg0 > -2.0f) {
    float w[4];
    parzen_weights_t(-1.0f - arg0, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = (unsigned int)(w[k] * 4294967296.0f);
    q[1] = 0u - q[0] - q[2] - q[3];
  } else {
    parzen_fixed(arg0, q);
  }
}

__device__ __forceinline__ float parzen_deriv_weight_fast(float arg0, const float* prow) {
  if (arg0 <= -1.0f && arg0 > -2.0f) return parzen_deriv_weight_t(-1.0f - arg0, prow);
  return parzen_deriv_weight(arg0, prow);
}

// Sum of N per-thread values over a 256-thread workgroup, written by thread 0
// to out[0..N): wavefront shuffles, then the 4 wavefront results in LDS, in a
// fixed order.  scratch holds 4 * N doubles.
template <int N>
__device__ __forceinline__ void block_reduce_array(double (&v)[N], double* scratch, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double x = v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) scratch[wid * N + i] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double r = 0.0;
    for (int w = 0; w < 4; ++w) r += scratch[w * N + threadIdx.x];
    out[threadIdx.x] = r;
  }
}

}  // namespace mmi
