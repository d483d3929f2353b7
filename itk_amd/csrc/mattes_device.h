// Device-side building blocks of the Mattes MI v4 hot path (gfx950).
//
// The coordinate path (index -> physical point -> transform -> continuous
// index -> linear interpolation -> pixel rounding -> Parzen bin) runs in FP64
// in exactly the operation order of the reference C++ (this file is compiled
// with -ffp-contract=off), so Parzen bin indices and valid-sample counts are
// bit-identical to ITK's.  Everything that only feeds the derivative (image
// gradient interpolation, per-sample weights) runs in FP32/FP64 without that
// constraint.  Index arithmetic is 32-bit: images and virtual domains hold
// fewer than 2^31 voxels (checked on the host).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmi {

// Unsigned division by a runtime constant: multiply-high and shifts
// (Granlund & Montgomery 1994), valid for every 32-bit numerator.
struct FastDiv {
  uint32_t d, m, s1, s2;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) ++l;
  FastDiv f;
  f.d = d;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 0 ? l - 1 : 0;
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
  const uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Geometry of an image or B-spline grid, with the ImageFunction bounds
// precomputed (Modules/Core/ImageFunction/include/itkImageFunction.hxx:50-67).
struct DevGeom {
  int32_t dim;
  int32_t size[3];
  int32_t start[3];
  int32_t end[3];  // start + size - 1
  int32_t sy;      // size[0]
  int32_t sz;      // size[0] * size[1]
  double origin[3];
  double spacing[3];
  double i2p[9];
  double p2i[9];
  double cilo[3];  // start - 0.5
  double cihi[3];  // end + 0.5
  FastDiv div_x;   // division by size[0]
  FastDiv div_y;   // division by size[1]
};

struct DevTransform {
  int32_t kind;  // 0 identity, 1 affine, 2 B-spline
  int32_t pad_;
  double m[9];
  double off[3];
  double center[3];
  DevGeom grid;
  const double* coeffs;  // B-spline parameters on the device, dimension-major
  int32_t pdim;          // nodes per dimension
  int32_t pad2_;
};

// Virtual index of linear sample s (x fastest).
template <int D>
__device__ __forceinline__ void sample_index(const DevGeom& v, uint32_t s, int32_t* idx) {
  const uint32_t r = fdiv(s, v.div_x);
  idx[0] = v.start[0] + (int32_t)(s - r * (uint32_t)v.size[0]);
  if (D == 2) {
    idx[1] = v.start[1] + (int32_t)r;
    idx[2] = 0;
  } else {
    const uint32_t z = fdiv(r, v.div_y);
    idx[1] = v.start[1] + (int32_t)(r - z * (uint32_t)v.size[1]);
    idx[2] = v.start[2] + (int32_t)z;
  }
}

// Math::Floor<int64>(double) = RoundHalfIntegerToEven(2x - 0.5) >> 1
// (Modules/Core/Common/include/itkMathDetail.h:393-396).  Not a plain floor
// for tiny negative x, so it is reproduced literally.  Callers only pass
// indices of points already inside a buffer or grid (|x| < 2^30).
__device__ __forceinline__ int32_t itk_floor(double x) {
  return ((int32_t)__builtin_rint(2.0 * x - 0.5)) >> 1;
}

// Math::RoundHalfIntegerUp<int64>(double) = RoundHalfIntegerToEven(2x + 0.5)
// >> 1 (itkMathDetail.h:232-235, the form of x86-64 builds).  Callers pass
// |x| < 2^30.
__device__ __forceinline__ int32_t itk_round(double x) {
  return ((int32_t)__builtin_rint(2.0 * x + 0.5)) >> 1;
}

// BSplineKernelFunction<3> (Modules/Core/Common/include/itkBSplineKernelFunction.h:153-175).
__device__ __forceinline__ double bspline3(double u) {
  const double a = fabs(u);
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

// ImageBase::TransformIndexToPhysicalPoint (itkImageBase.h:594-604).
template <int D>
__device__ __forceinline__ void index_to_point(const DevGeom& g, const int32_t* idx, double* p) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double s = g.origin[i];
#pragma unroll
    for (int j = 0; j < D; ++j) s += g.i2p[i * D + j] * (double)idx[j];
    p[i] = s;
  }
}

// ImageBase::TransformPhysicalPointToContinuousIndex (itkImageBase.h:517-531)
// with Matrix * Vector summation order (itkMatrix.hxx:30-41).
template <int D>
__device__ __forceinline__ void point_to_cindex(const DevGeom& g, const double* p, double* ci) {
  double v[D];
#pragma unroll
  for (int k = 0; k < D; ++k) v[k] = p[k] - g.origin[k];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) s += g.p2i[r * D + c] * v[c];
    ci[r] = s;
  }
}

// An image mask as ImageMaskSpatialObject::IsInsideInObjectSpace sees a
// point (Modules/Core/SpatialObjects/include/itkImageMaskSpatialObject.hxx:36-48,
// identity object-to-world transform): the rounded index
// (ImageBase::TransformPhysicalPointToIndex, itkImageBase.h:465-479) inside
// the mask's buffer and the mask pixel there not 0.
struct DevMask {
  DevGeom g;
  const uint8_t* px;
};

// ImageBase::TransformPhysicalPointToIndex(point, index) (itkImageBase.h:465-499):
// whether the rounded index is in the buffer, and its buffer offset.
template <int D>
__device__ __forceinline__ bool rounded_index_inside(const DevGeom& g, const double* p, int32_t* off) {
  double ci[3];
  point_to_cindex<D>(g, p, ci);
  int32_t k[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < D; ++i) {
    // indices outside [start - 1, end + 1] round outside the buffer (and stay
    // in range for the rounding); NaN fails the test
    if (!(ci[i] > (double)g.start[i] - 1.0 && ci[i] < (double)g.end[i] + 1.0)) return false;
    k[i] = itk_round(ci[i]);
    if (k[i] < g.start[i] || k[i] > g.end[i]) return false;
  }
  *off = (k[0] - g.start[0]) + g.sy * (k[1] - g.start[1]) + (D == 3 ? g.sz * (k[2] - g.start[2]) : 0);
  return true;
}

template <int D>
__device__ __forceinline__ bool mask_inside(const DevMask& m, const double* p) {
  int32_t off = 0;
  return rounded_index_inside<D>(m.g, p, &off) && m.px[off] != 0;
}

// ImageFunction::IsInsideBuffer(ContinuousIndex) (itkImageFunction.h:159-170).
template <int D>
__device__ __forceinline__ bool inside_buffer(const DevGeom& g, const double* ci) {
  bool in = true;
#pragma unroll
  for (int j = 0; j < D; ++j) in = in && (ci[j] >= g.cilo[j] && ci[j] < g.cihi[j]);
  return in;
}

// The interpolation cell of LinearInterpolateImageFunction::EvaluateOptimized
// (itkLinearInterpolateImageFunction.h:157-471).  The reference skips an axis
// when its distance is <= 0 or when the upper neighbour leaves the buffer;
// duplicating the lower corner on such an axis makes the full multilinear
// formula return the same bits (v + (v - v) * d == v), so one branch-free
// formula reproduces every branch.  Only called for points inside the buffer.
struct Cell {
  int32_t base;
  int32_t ox, oy, oz;  // offsets of the upper corners (0 on inactive axes)
  double d[3];
};

template <int D>
__device__ __forceinline__ Cell make_cell(const DevGeom& g, const double* ci) {
  Cell c;
  int32_t b[3] = {0, 0, 0};
  bool act[3] = {false, false, false};
  c.d[2] = 0.0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const int32_t f = itk_floor(ci[j]);
    b[j] = f > g.start[j] ? f : g.start[j];
    c.d[j] = ci[j] - (double)b[j];
    act[j] = (c.d[j] > 0.0) && (b[j] + 1 <= g.end[j]);
  }
  c.base = (b[0] - g.start[0]) + g.sy * (b[1] - g.start[1]) + (D == 3 ? g.sz * (b[2] - g.start[2]) : 0);
  c.ox = act[0] ? 1 : 0;
  c.oy = act[1] ? g.sy : 0;
  c.oz = (D == 3 && act[2]) ? g.sz : 0;
  return c;
}

__device__ __forceinline__ double lerp_exact(double a, double b, double t) { return a + (b - a) * t; }

// Image storage formats.  Each gives the pixel and gradient-texel types and
// their exact conversion to FP32 (every 16-bit value is exact in FP32).
// Texels hold (moving value, dI/dx, dI/dy, dI/dz): float4 for FP32 storage,
// four 16-bit halves packed in a uint2 for BF16 and FP16 storage (x = value |
// dI/dx << 16, y = dI/dy | dI/dz << 16).
// Element i of a buffer below 4 GiB through a 32-bit byte offset from the
// wavefront-uniform base: the load takes the base in scalar registers and the
// offset in one vector register, with no 64-bit address arithmetic per load
// (the fast path's images; the host checks their size).
template <typename T>
__device__ __forceinline__ T at32(const T* p, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (uint32_t)sizeof(T));
}
// The same with the byte offset given (a cell's corners then cost one add
// each, the scaling done once per sample).
template <typename T>
__device__ __forceinline__ T atb(const T* p, uint32_t b) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + b);
}

struct PxF32 {
  using pixel = float;
  using texel = float4;
  static __device__ __forceinline__ float ld(const pixel* p, int32_t i) { return p[i]; }
  static __device__ __forceinline__ float4 ldt(const texel* t, int32_t i) { return t[i]; }
  static __device__ __forceinline__ float ld32(const pixel* p, uint32_t i) { return at32(p, i); }
  static __device__ __forceinline__ float4 ldt32(const texel* t, uint32_t i) { return at32(t, i); }
  static __device__ __forceinline__ float ldb(const pixel* p, uint32_t b) { return atb(p, b); }
  static __device__ __forceinline__ float4 ldtb(const texel* t, uint32_t b) { return atb(t, b); }
};

struct PxBF16 {
  using pixel = uint16_t;
  using texel = uint2;
  static __device__ __forceinline__ float cv(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
  static __device__ __forceinline__ float4 cvt(uint2 u) {
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u), __uint_as_float(u.y << 16),
                       __uint_as_float(u.y & 0xFFFF0000u));
  }
  static __device__ __forceinline__ float ld(const pixel* p, int32_t i) { return cv(p[i]); }
  static __device__ __forceinline__ float4 ldt(const texel* t, int32_t i) { return cvt(t[i]); }
  static __device__ __forceinline__ float ld32(const pixel* p, uint32_t i) { return cv(at32(p, i)); }
  static __device__ __forceinline__ float4 ldt32(const texel* t, uint32_t i) { return cvt(at32(t, i)); }
  static __device__ __forceinline__ float ldb(const pixel* p, uint32_t b) { return cv(atb(p, b)); }
  static __device__ __forceinline__ float4 ldtb(const texel* t, uint32_t b) { return cvt(atb(t, b)); }
};

struct PxF16 {
  using pixel = uint16_t;
  using texel = uint2;
  static __device__ __forceinline__ float h2f(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }
  static __device__ __forceinline__ float4 cvt(uint2 u) {
    return make_float4(h2f(u.x & 0xFFFFu), h2f(u.x >> 16), h2f(u.y & 0xFFFFu), h2f(u.y >> 16));
  }
  static __device__ __forceinline__ float ld(const pixel* p, int32_t i) { return h2f(p[i]); }
  static __device__ __forceinline__ float4 ldt(const texel* t, int32_t i) { return cvt(t[i]); }
  static __device__ __forceinline__ float ld32(const pixel* p, uint32_t i) { return h2f(at32(p, i)); }
  static __device__ __forceinline__ float4 ldt32(const texel* t, uint32_t i) { return cvt(at32(t, i)); }
  static __device__ __forceinline__ float ldb(const pixel* p, uint32_t b) { return h2f(atb(p, b)); }
  static __device__ __forceinline__ float4 ldtb(const texel* t, uint32_t b) { return cvt(atb(t, b)); }
};

// Exact FP64 multilinear value of a float image (the reference's RealType is
// double for float pixels).
template <int D, typename P>
__device__ __forceinline__ double interp_exact(const typename P::pixel* __restrict__ img, const Cell& c) {
  const double v000 = (double)P::ld(img, c.base);
  const double v100 = (double)P::ld(img, c.base + c.ox);
  const double v010 = (double)P::ld(img, c.base + c.oy);
  const double v110 = (double)P::ld(img, c.base + c.oy + c.ox);
  const double vx00 = lerp_exact(v000, v100, c.d[0]);
  const double vx10 = lerp_exact(v010, v110, c.d[0]);
  const double vxx0 = lerp_exact(vx00, vx10, c.d[1]);
  if (D == 2) return vxx0;
  const double v001 = (double)P::ld(img, c.base + c.oz);
  const double v101 = (double)P::ld(img, c.base + c.oz + c.ox);
  const double v011 = (double)P::ld(img, c.base + c.oz + c.oy);
  const double v111 = (double)P::ld(img, c.base + c.oz + c.oy + c.ox);
  const double vx01 = lerp_exact(v001, v101, c.d[0]);
  const double vx11 = lerp_exact(v011, v111, c.d[0]);
  const double vxx1 = lerp_exact(vx01, vx11, c.d[1]);
  return lerp_exact(vxx0, vxx1, c.d[2]);
}

// Gradient at a moving cell: only the gradient components of the texels are
// interpolated here (FP32), with the same cell as the value.
// U32: 32-bit byte-offset loads (fast path; texel buffer below 4 GiB).
template <int D, typename P, bool U32 = false>
__device__ __forceinline__ float3 interp_gradient3(const typename P::texel* __restrict__ tex, const Cell& c) {
  const float t0 = (float)c.d[0], t1 = (float)c.d[1], t2 = (float)c.d[2];
  auto L = [&](int32_t i) { return U32 ? P::ldt32(tex, (uint32_t)i) : P::ldt(tex, i); };
  const float4 a000 = L(c.base);
  const float4 a100 = L(c.base + c.ox);
  const float4 a010 = L(c.base + c.oy);
  const float4 a110 = L(c.base + c.oy + c.ox);
  const float gx0 = fmaf(a100.y - a000.y, t0, a000.y), gx1 = fmaf(a110.y - a010.y, t0, a010.y);
  const float gy0 = fmaf(a100.z - a000.z, t0, a000.z), gy1 = fmaf(a110.z - a010.z, t0, a010.z);
  const float gz0 = fmaf(a100.w - a000.w, t0, a000.w), gz1 = fmaf(a110.w - a010.w, t0, a010.w);
  float3 r = make_float3(fmaf(gx1 - gx0, t1, gx0), fmaf(gy1 - gy0, t1, gy0), fmaf(gz1 - gz0, t1, gz0));
  if (D == 3) {
    const float4 a001 = L(c.base + c.oz);
    const float4 a101 = L(c.base + c.oz + c.ox);
    const float4 a011 = L(c.base + c.oz + c.oy);
    const float4 a111 = L(c.base + c.oz + c.oy + c.ox);
    const float hx0 = fmaf(a101.y - a001.y, t0, a001.y), hx1 = fmaf(a111.y - a011.y, t0, a011.y);
    const float hy0 = fmaf(a101.z - a001.z, t0, a001.z), hy1 = fmaf(a111.z - a011.z, t0, a011.z);
    const float hz0 = fmaf(a101.w - a001.w, t0, a001.w), hz1 = fmaf(a111.w - a011.w, t0, a011.w);
    const float s0 = fmaf(hx1 - hx0, t1, hx0), s1 = fmaf(hy1 - hy0, t1, hy0), s2 = fmaf(hz1 - hz0, t1, hz0);
    r.x = fmaf(s0 - r.x, t2, r.x);
    r.y = fmaf(s1 - r.y, t2, r.y);
    r.z = fmaf(s2 - r.z, t2, r.z);
  }
  return r;
}

// Math::FloatAlmostEqual(x, y, 4 ulps, 0.1 eps) (itkMath.h:331-356).
__device__ __forceinline__ bool float_almost_equal(double x, double y) {
  if (fabs(x - y) <= 0.1 * 2.220446049250313e-16) return true;
  if (signbit(x) != signbit(y)) return false;
  auto asulp = [](int64_t i) -> int64_t { return i < 0 ? (int64_t)(0x8000000000000000ULL) - i : i; };
  int64_t u = asulp(__double_as_longlong(x)) - asulp(__double_as_longlong(y));
  if (u < 0) u = -u;
  return u <= 4;
}

// B-spline transform support of one point: BSplineTransform::TransformPoint
// (Modules/Core/Transform/include/itkBSplineTransform.hxx:476-590) with
// BSplineInterpolationWeightFunction::Evaluate
// (Modules/Core/Common/include/itkBSplineInterpolationWeightFunction.hxx:43-94).
template <int D>
struct BSplineSupport {
  bool inside;
  int32_t start[3];
  double w1[3][4];
};

// One axis of BSplineTransform::InsideValidRegion (itkBSplineTransform.hxx:478-506):
// false when the continuous grid index c is outside [minLimit, maxLimit); c is
// moved 6 ulp below maxLimit when it is within 4 ulp of it.  The region test
// is the conjunction of the axes.
__device__ __forceinline__ bool bspline_axis_inside(const DevGeom& grid, int j, double& c) {
  const double minLimit = 0.5 * (double)(3 - 1);
  const double maxLimit = (double)grid.size[j] - 0.5 * (double)(3 - 1) - 1.0;
  if (float_almost_equal(c, maxLimit)) {
    c = __longlong_as_double(__double_as_longlong(maxLimit) - 6);  // FloatAddULP(maxLimit, -6)
    return true;
  }
  if (c >= maxLimit) return false;
  if (c < minLimit) return false;
  return true;
}

// One axis of BSplineInterpolationWeightFunction::Evaluate
// (itkBSplineInterpolationWeightFunction.hxx:43-94): support start and the
// four 1-D cubic weights.
__device__ __forceinline__ int32_t bspline_axis_weights(double c, double* w) {
  const int32_t start = itk_floor(c + 0.5 - 3 / 2.0);
  double x = c - (double)start;
#pragma unroll
  for (int k = 0; k <= 3; ++k) {
    w[k] = bspline3(x);
    x -= 1.0;
  }
  return start;
}

template <int D>
__device__ __forceinline__ void bspline_support(const DevGeom& grid, const double* p, BSplineSupport<D>& s) {
  double ci[3];
  point_to_cindex<D>(grid, p, ci);
  s.inside = true;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    if (!s.inside) break;
    s.inside = bspline_axis_inside(grid, j, ci[j]);
  }
  if (!s.inside) return;
#pragma unroll
  for (int j = 0; j < D; ++j) s.start[j] = bspline_axis_weights(ci[j], s.w1[j]);
}

// Weight and node index of support point k (x fastest), as the reference's
// weights[k] = 1 * w0 * w1 (* w2) and coefficient scanline order.
template <int D>
__device__ __forceinline__ double bspline_weight(const BSplineSupport<D>& s, int k) {
  double w = 1.0;
  w *= s.w1[0][k & 3];
  w *= s.w1[1][(k >> 2) & 3];
  if (D == 3) w *= s.w1[2][(k >> 4) & 3];
  return w;
}

template <int D>
__device__ __forceinline__ int32_t bspline_node(const DevGeom& grid, const BSplineSupport<D>& s, int k) {
  int32_t n = s.start[0] + (k & 3);
  n += grid.size[0] * (s.start[1] + ((k >> 2) & 3));
  if (D == 3) n += grid.size[0] * grid.size[1] * (s.start[2] + ((k >> 4) & 3));
  return n;
}

// Transform::TransformPoint for identity, affine (MatrixOffsetTransformBase,
// itkMatrixOffsetTransformBase.hxx:124-129) and cubic B-spline.  KIND is the
// transform kind when known at compile time, -1 to dispatch on t.kind.
template <int D, int KIND = -1>
__device__ __forceinline__ void transform_point(const DevTransform& t, const double* in, double* out,
                                                BSplineSupport<D>* sup) {
  const int kind = KIND >= 0 ? KIND : t.kind;
  if (kind == 1) {
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) s += t.m[r * D + c] * in[c];
      out[r] = s + t.off[r];
    }
    return;
  }
  if (kind == 2) {
    bspline_support<D>(t.grid, in, *sup);
    if (!sup->inside) {
#pragma unroll
      for (int i = 0; i < D; ++i) out[i] = in[i];
      return;
    }
    // the support's nodes in the reference's order (x fastest), summed in
    // that order; the four x nodes of a support row are adjacent in memory and
    // their loads are issued together, several rows ahead of the sums
    double acc[3] = {0.0, 0.0, 0.0};
    constexpr int NR = D == 3 ? 16 : 4;
    const int32_t gx = t.grid.size[0], gxy = t.grid.size[0] * t.grid.size[1];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int32_t row = sup->start[0] + gx * (sup->start[1] + (r & 3)) + (D == 3 ? gxy * (sup->start[2] + (r >> 2)) : 0);
      double c[3][4];
#pragma unroll
      for (int j = 0; j < D; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) c[j][i] = t.coeffs[j * t.pdim + row + i];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double w = bspline_weight<D>(*sup, r * 4 + i);
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] += w * c[j][i];
      }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) out[j] = acc[j] + in[j];
    return;
  }
#pragma unroll
  for (int i = 0; i < D; ++i) out[i] = in[i];
}

}  // namespace mmi
