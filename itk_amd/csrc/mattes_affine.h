// The transform-dependent arguments of the certified affine fast path, as
// __host__ __device__ code: the host's set-up and the device's optimizer loop
// (mattes_optim.hip) compute them with the same operations, so an evaluation
// after a device-side parameter update has the bits of one set up on the host.
// Compiled without FMA contraction (as every source of the library).
#pragma once

#include "mattes_kernels.h"

namespace mmi {

// What the map and its bounds need besides the transform's parameters; fixed
// for a level (virtual domain, moving image geometry and intensity range).
struct AffineMapInputs {
  int32_t dim;
  int32_t rows;  // 1: the row walks' fixed-point parameters (FastRows) are kept up to date
  int32_t walk;  // rows per walk (rows == 1)
  int32_t bins;
  int32_t v_start[3], v_size[3];
  int32_t m_start[3], m_size[3];
  double v_origin[3], v_i2p[9];
  double m_origin[3], m_p2i[9];
  double center[3];  // the affine transform's centre (its fixed parameters)
  double Mm;         // largest |moving intensity bound|
  double moving_bin_size, moving_norm_min;
};

__host__ __device__ inline double amax(double a, double b) { return a > b ? a : b; }

// The functions below are templates on the dimension D, so that the device's
// loops unroll and its arrays stay in registers; the runtime-D wrappers serve
// the host.  The operations and their order do not depend on the unrolling.

// MatrixOffsetTransformBase::SetParameters + ComputeOffset
// (itkMatrixOffsetTransformBase.hxx:518-556, 615-631): matrix row-major, then
// offset = translation + centre - matrix * centre, summed in that order.
template <int D>
__host__ __device__ inline void affine_from_parameters_d(const double* params, const double* center,
                                                         DevTransform& t) {
  t.kind = 1;
  for (int i = 0; i < D * D; ++i) t.m[i] = params[i];
  for (int i = 0; i < D; ++i) {
    double off = params[D * D + i] + center[i];
    for (int j = 0; j < D; ++j) off -= params[i * D + j] * center[j];
    t.off[i] = off;
    t.center[i] = center[i];
  }
}

inline void affine_from_parameters(int D, const double* params, const double* center, DevTransform& t) {
  if (D == 2)
    affine_from_parameters_d<2>(params, center, t);
  else
    affine_from_parameters_d<3>(params, center, t);
}

// Continuous index in geometry (g_origin, g_p2i) of the point T(o_v + I2P_v
// idx), T(x) = M x + off, as one affine map of the virtual index
// ci = A idx + b, with a bound e on |FP64 A idx + b - the reference's chain
// of roundings| (index -> point -> transform -> continuous index): 64 ulp of
// the largest intermediate magnitude over the domain's corners.
template <int D>
__host__ __device__ inline void affine_index_map_d(const int32_t* v_start, const int32_t* v_size,
                                                   const double* v_origin, const double* v_i2p, const double* M,
                                                   const double* off, const double* g_origin, const double* g_p2i,
                                                   double* A, double* b, double* e) {
  double MI[D * D], w[D];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s += M[i * D + k] * v_i2p[k * D + j];
      MI[i * D + j] = s;
    }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s += g_p2i[i * D + k] * MI[k * D + j];
      A[i * D + j] = s;
    }
  for (int i = 0; i < D; ++i) {
    double s = off[i] - g_origin[i];
    for (int k = 0; k < D; ++k) s += M[i * D + k] * v_origin[k];
    w[i] = s;
  }
  for (int i = 0; i < D; ++i) {
    double s = 0.0;
    for (int k = 0; k < D; ++k) s += g_p2i[i * D + k] * w[k];
    b[i] = s;
  }
  double rowA = 0.0;
  for (int i = 0; i < D * D; ++i) rowA = amax(rowA, fabs(A[i]));
  double S = 0.0;
  for (int corner = 0; corner < (1 << D); ++corner) {
    double idx[D], vp[D], mp[D], mag = 0.0;
    for (int k = 0; k < D; ++k) idx[k] = (double)(v_start[k] + ((corner >> k) & 1 ? v_size[k] - 1 : 0));
    for (int i = 0; i < D; ++i) {
      double s = fabs(v_origin[i]);
      vp[i] = v_origin[i];
      for (int j = 0; j < D; ++j) {
        vp[i] += v_i2p[i * D + j] * idx[j];
        s += fabs(v_i2p[i * D + j] * idx[j]);
      }
      mag = amax(mag, s);
    }
    for (int i = 0; i < D; ++i) {
      double s = fabs(off[i]);
      mp[i] = off[i];
      for (int j = 0; j < D; ++j) {
        mp[i] += M[i * D + j] * vp[j];
        s += fabs(M[i * D + j] * vp[j]);
      }
      mag = amax(mag, s);
    }
    for (int i = 0; i < D; ++i) {
      double s = 0.0;
      for (int j = 0; j < D; ++j) s += fabs(g_p2i[i * D + j]) * (fabs(mp[j]) + fabs(g_origin[j]));
      mag = amax(mag, s);
    }
    S = amax(S, mag * (1.0 + rowA));
  }
  *e = 64.0 * S * 2.220446049250313e-16 + 1e-300;
}

inline void affine_index_map(int D, const int32_t* v_start, const int32_t* v_size, const double* v_origin,
                             const double* v_i2p, const double* M, const double* off, const double* g_origin,
                             const double* g_p2i, double* A, double* b, double* e) {
  if (D == 2)
    affine_index_map_d<2>(v_start, v_size, v_origin, v_i2p, M, off, g_origin, g_p2i, A, b, e);
  else
    affine_index_map_d<3>(v_start, v_size, v_origin, v_i2p, M, off, g_origin, g_p2i, A, b, e);
}

// The row walks' steps and buffer thresholds in 32.32 fixed point for the map
// A (FastRows; walk, sizes and divisions set by the host), and e widened by
// the walk's error: the FP64 start is within e of the reference's chain and
// within 3 roundings of S (< e / 8) of the exact map, its conversion
// truncates below 2^-32, each step's rounded increment is within 2^-33 of the
// map's y column.  Thresholds in integers: with ce = ceil(e 2^32), the
// certainly-outside bounds are start - 1/2 - e and end + 1/2 + e rounded
// outward, the certainly-inside ones start - 1/2 + e and end + 1/2 - e
// rounded inward, one more unit of margin each.
template <int D>
__host__ __device__ inline void rows_thresholds_d(const int32_t* m_start, const int32_t* m_size, int32_t walk,
                                                  const double* A, double* e, FastRows& R) {
  const double ew = 1.125 * *e + (double)(walk + 2) * 2.3283064365386963e-10;
  const long long ce = (long long)ceil(ew * 4294967296.0);
  const long long half = 1ll << 31;
  for (int k = 0; k < D; ++k) {
    R.qy[k] = (long long)rint(A[k * D + 1] * 4294967296.0);
    const long long lo = (long long)m_start[k] * 4294967296ll - half;             // start - 1/2
    const long long hi = (long long)(m_start[k] + m_size[k]) * 4294967296ll - half;  // end + 1/2
    const long long lo_out = lo - ce - 1, hi_out = hi + ce + 1;
    const long long lo_in = lo + ce + 1, hi_in = hi - ce - 1;
    R.lo_out[k] = lo_out;
    R.span_out[k] = (unsigned long long)(hi_out - lo_out);
    R.lo_in[k] = lo_in;
    R.span_in[k] = hi_in > lo_in ? (unsigned long long)(hi_in - lo_in) : 0ull;
    if (R.cd_on) {
      // central differences: both half-step points inside iff the index is in
      // [start, end - 1); the reference rounds the offset point and its index
      // once more each (a few ulp of S, < e / 8), so the band is 2 e wide
      const long long ce2 = (long long)ceil(2.0 * ew * 4294967296.0) + 1;
      const long long a0 = (long long)m_start[k] * 4294967296ll;
      const long long a1 = (long long)(m_start[k] + m_size[k] - 1) * 4294967296ll;
      const long long in_lo = a0 + ce2, in_hi = a1 - ce2;
      const long long out_lo = a0 - ce2, out_hi = a1 + ce2;
      R.cd_lo_in[k] = in_lo;
      R.cd_span_in[k] = in_hi > in_lo ? (unsigned long long)(in_hi - in_lo) : 0ull;
      R.cd_lo_out[k] = out_lo;
      R.cd_span_out[k] = (unsigned long long)(out_hi - out_lo);
    }
  }
  *e = ew;
}

// Bounds of the moving value (ev) and moving Parzen term (bt_m) for a map
// error e_ci, for FP32 interpolation of values of magnitude <= Mm.
__host__ __device__ inline void moving_value_bounds(double Mm, double moving_bin_size, double moving_norm_min, int B,
                                                    double e_ci, double* ev, double* bt_m) {
  *ev = Mm * 1.9073486328125e-06 + 12.0 * Mm * e_ci;  // 2^-19
  const double u = 2.384185791015625e-07;             // 2^-22
  *bt_m = 2.0 * ((*ev + Mm * u) / moving_bin_size + u * (Mm / moving_bin_size + fabs(moving_norm_min) + B));
}

// New parameters of the affine moving transform: the device transform and the
// fast path's map, error bound, row-walk thresholds and moving bounds.  false:
// the certified path does not hold for these parameters (the bounds exceed
// what it takes; t and F are then partly updated).
template <int D>
__host__ __device__ inline bool affine_fast_update_d(const AffineMapInputs& in, const double* params,
                                                     DevTransform& t, FastArgs& F) {
  double M[D * D], off[D], A[D * D], bb[D];
  for (int i = 0; i < D * D; ++i) M[i] = params[i];
  for (int i = 0; i < D; ++i) {
    double o = params[D * D + i] + in.center[i];
    for (int j = 0; j < D; ++j) o -= M[i * D + j] * in.center[j];
    off[i] = o;
  }
  double e = 0.0;
  affine_index_map_d<D>(in.v_start, in.v_size, in.v_origin, in.v_i2p, M, off, in.m_origin, in.m_p2i, A, bb, &e);
  t.kind = 1;
  for (int i = 0; i < D * D; ++i) t.m[i] = M[i];
  for (int i = 0; i < D; ++i) {
    t.off[i] = off[i];
    t.center[i] = in.center[i];
  }
  for (int i = 0; i < D * D; ++i) F.A[i] = A[i];
  for (int i = 0; i < D; ++i) F.b[i] = bb[i];
  if (in.rows) {
    // the walk's fixed-point steps stay far below 2^31 voxels
    for (int i = 0; i < D * D; ++i)
      if (!(fabs(A[i]) < 1048576.0)) return false;
    rows_thresholds_d<D>(in.m_start, in.m_size, in.walk, A, &e, F.rows);
  }
  F.e_ci = e;
  double ev = 0.0, bt_m = 0.0;
  moving_value_bounds(in.Mm, in.moving_bin_size, in.moving_norm_min, in.bins, e, &ev, &bt_m);
  F.ev = (float)ev;
  F.bt_m = (float)bt_m;
  return e <= 1e-6 && bt_m <= 0.05;
}

inline bool affine_fast_update(const AffineMapInputs& in, const double* params, DevTransform& t, FastArgs& F) {
  return in.dim == 2 ? affine_fast_update_d<2>(in, params, t, F) : affine_fast_update_d<3>(in, params, t, F);
}

}  // namespace mmi
