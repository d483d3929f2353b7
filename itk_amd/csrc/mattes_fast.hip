// Certified fast path of the Mattes MI v4 evaluation (affine moving transform,
// identity fixed transform, fixed samples on fixed pixel centres).
//
// Per sample: the moving continuous index is one FP64 affine map of the
// virtual index (|error| <= e_ci against the reference's chain of
// index -> point -> transform -> continuous index), interpolation and Parzen
// terms run in FP32 with precomputed error bounds.  A sample is "certain" when
// its inside test, intensity-range test and both bin indices provably equal
// the reference's; every other sample goes to a per-workgroup exception list
// and is evaluated afterwards on the exact FP64 path (mattes_eval.h), so the
// bins and the valid count are bit-identical to the reference's.
//
// Kernels, affine: pass1_fast (histogram of certain samples; each workgroup
// then evaluates its own listed samples exactly into the same LDS histogram)
// -> finalize (mattes_kernels.hip) -> pass2_fast (derivative moments of
// certain samples, then the workgroup's listed samples exactly, sorted so the
// sums are in a fixed order) -> fast_derivative (fixed-order reduction,
// index-space moments mapped to the affine parameters).  B-spline:
// bspline_rows -> pass1_fast -> exceptions_pass1 (listed samples exactly,
// eight lists per workgroup, their results kept per list slot) -> finalize
// -> pass2 (per-sample factors; its tail writes the listed samples' factors
// from the kept results) -> separable contraction (mattes_prep.hip).
//
// The hot loops read a compact argument block so that the uniform values they
// use fit the scalar register file; the exact evaluations take the full
// EvalArgs.
#include "mattes_eval.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace mmi {

// Samples of local workgroup lb of lg (one evaluation's share of a launch).
template <int D>
__device__ __forceinline__ void run_range(const FastArgs& f, uint32_t lb, uint32_t lg, uint32_t& begin, uint32_t& end,
                                          uint32_t& step) {
  if (f.xcd_chunks) {
    // workgroup b runs on XCD b % 8 (lb and b agree mod 8: shares start at
    // multiples of 8): give each XCD one contiguous share
    const uint32_t xcd = lb % kXcd;
    const uint32_t lbx = lb / kXcd;
    const uint32_t bpx = lg / kXcd;
    const uint32_t chunk = (f.n + kXcd - 1) / kXcd;
    const uint32_t c0 = xcd * chunk;
    begin = c0 + lbx * blockDim.x + threadIdx.x;
    end = min(f.n, c0 + chunk);
    step = bpx * blockDim.x;
  } else {
    begin = lb * blockDim.x + threadIdx.x;
    end = f.n;
    step = lg * blockDim.x;
  }
}

template <int D>
__device__ __forceinline__ void fast_index(const FastArgs& f, uint32_t s, int32_t* idx) {
  const uint32_t r = fdiv(s, f.vdx);
  idx[0] = f.vstart[0] + (int32_t)(s - r * (uint32_t)f.vsize0);
  if (D == 2) {
    idx[1] = f.vstart[1] + (int32_t)r;
    idx[2] = 0;
  } else {
    const uint32_t z = fdiv(r, f.vdy);
    idx[1] = f.vstart[1] + (int32_t)(r - z * (uint32_t)f.vsize1);
    idx[2] = f.vstart[2] + (int32_t)z;
  }
}

// Bin of an FP32 Parzen term, clamped to [2, B-3] like the reference; sets
// *uncertain when the reference's FP64 term could fall on the other side of
// an unclamped bin edge (|term - n| <= bound for an edge n in [3, B-3]).
__device__ __forceinline__ int fast_bin(float term, float bound, int B, bool* uncertain) {
  const float n = rintf(term);
  *uncertain = *uncertain | ((fabsf(term - n) <= bound) & (n >= 3.0f) & (n <= (float)(B - 3)));
  return (int)fminf(fmaxf(term, 2.0f), (float)(B - 3));
}

// 0: not counted, 1: counted (bins, arg0, cell set), 2: uncertain.  Straight
// line code: the cell is clamped into the buffer so every load is in bounds,
// and the decisions are combined at the end.
// Moving continuous index of a virtual sample.  MK 1: one FP64 affine map of
// the index.  MK 2: the virtual point, the B-spline displacement from the
// per-row coefficient lines (sum over the four x nodes) and the moving
// physical-to-index map; *bsin tells whether the sample is inside the
// transform's valid region (else the displacement is zero).
template <int D, int MK>
__device__ __forceinline__ void fast_cindex(const FastArgs& f, const int32_t* idx, double* ci, bool* bsin) {
  if constexpr (MK == 2) {
    // diagonal virtual geometry: the off-diagonal terms of the reference's
    // sum are exact zeros, so one product per axis gives the same point
    double vp[3], mp[3];
#pragma unroll
    for (int k = 0; k < D; ++k) vp[k] = f.vorg[k] + f.vsp[k] * (double)idx[k];
    const int xl = idx[0] - f.vstart[0], yl = idx[1] - f.vstart[1], zl = D == 3 ? idx[2] - f.vstart[2] : 0;
    // The three flags, the weights and the coefficient line are read without
    // waiting for one another: an index outside the transform's region has
    // support start 0 and zero weights, and every virtual row has its line,
    // so the reads are in bounds for every sample; the flags then select.
    const uint32_t in_x = f.bs_inside[0][xl], in_y = f.bs_inside[1][yl];
    const uint32_t in_z = D == 3 ? f.bs_inside[2][zl] : 1u;
    const bool in = (in_x & in_y & in_z) != 0u;
    *bsin = in;
    const double* w = f.bs_w[0] + 4 * xl;
    const double* row = f.bs_rows + ((int64_t)(zl * f.vsize1 + yl) * D) * f.bs_gx + f.bs_start[0][xl];
    double wv[4], rv[3][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) wv[k] = w[k];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int k = 0; k < 4; ++k) rv[d][k] = row[d * f.bs_gx + k];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double disp = wv[0] * rv[d][0] + wv[1] * rv[d][1] + wv[2] * rv[d][2] + wv[3] * rv[d][3];
      mp[d] = in ? disp + vp[d] : vp[d];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < D; ++j) s += f.mp2i[k * D + j] * (mp[j] - f.morg[j]);
      ci[k] = s;
    }
  } else {
    *bsin = true;
    const double x = (double)idx[0], y = (double)idx[1], z = (double)idx[2];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double c = f.b[k];
      if (D == 3) c = fma(f.A[k * 3 + 2], z, c);
      c = fma(f.A[k * D + 1], y, c);
      ci[k] = fma(f.A[k * D + 0], x, c);
    }
  }
}

// ImageMaskSpatialObject test of a fast-path sample: the mask's continuous
// index from its affine map, rounded half up; certain when it lies more than
// e from every half-integer (then the reference's rounding gives the same
// index).  Sets *out when the mask certainly excludes the sample, *uncertain
// when the rounding is not certain.
template <int D>
__device__ __forceinline__ void fast_mask(const FastMask& m, double x, double y, double z, bool* out,
                                          bool* uncertain) {
  if (!m.px) return;  // uniform
  bool in = true, unc = false;
  int32_t off = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double c = m.b[k];
    if (D == 3) c = fma(m.A[k * 3 + 2], z, c);
    c = fma(m.A[k * D + 1], y, c);
    c = fma(m.A[k * D + 0], x, c);
    const double r = floor(c + 0.5);
    const double fr = c - (r - 0.5);
    unc = unc | !((fr > m.e) & (fr < 1.0 - m.e));
    const int32_t ri = (int32_t)fmin(fmax(r, -1.0e9), 1.0e9);
    in = in & (ri >= m.start[k]) & (ri <= m.end[k]);
    off += (ri - m.start[k]) * (k == 0 ? 1 : (k == 1 ? m.sy : m.sz));
  }
  if (in) in = m.px[off] != 0;
  *uncertain = *uncertain | unc;
  *out = *out | (!unc & !in);
}

// FP32 multilinear value of the moving image at a cell (corner element offsets
// cb + {0, cx} + {0, cy} + {0, cz}, weights t0..t2; every offset non-negative
// and inside the buffer).  TEX: gather the moving texels (value and gradient
// channels) instead of the moving values and return the interpolated gradient
// as well; the value is interpolated with the same operations either way, and
// every fast-path kernel interpolates through this function.
// The multilinear arithmetic of every fast-path interpolation, on corner
// values already loaded (cXYZ: corner with upper x, y, z flags).  TEX: the
// corners are texels (value, dI/dx, dI/dy, dI/dz); the value channel takes the
// same operations as plain values, and the gradient channels those of
// interp_gradient3.
__device__ __forceinline__ float mlerp(float c000, float c100, float c010, float c110, float t0, float t1) {
  const float vx00 = fmaf(c100 - c000, t0, c000);
  const float vx10 = fmaf(c110 - c010, t0, c010);
  return fmaf(vx10 - vx00, t1, vx00);
}

template <int D>
__device__ __forceinline__ float mlerp_value(float c000, float c100, float c010, float c110, float c001, float c101,
                                             float c011, float c111, float t0, float t1, float t2) {
  float v = mlerp(c000, c100, c010, c110, t0, t1);
  if (D == 3) {
    const float vx01 = fmaf(c101 - c001, t0, c001);
    const float vx11 = fmaf(c111 - c011, t0, c011);
    const float vxx1 = fmaf(vx11 - vx01, t1, vx01);
    v = fmaf(vxx1 - v, t2, v);
  }
  return v;
}

template <int D>
__device__ __forceinline__ float mlerp_texels(const float4& a000, const float4& a100, const float4& a010,
                                              const float4& a110, const float4& a001, const float4& a101,
                                              const float4& a011, const float4& a111, float t0, float t1, float t2,
                                              float3* grad) {
  const float vx00 = fmaf(a100.x - a000.x, t0, a000.x);
  const float vx10 = fmaf(a110.x - a010.x, t0, a010.x);
  float v = fmaf(vx10 - vx00, t1, vx00);
  const float gx0 = fmaf(a100.y - a000.y, t0, a000.y), gx1 = fmaf(a110.y - a010.y, t0, a010.y);
  const float gy0 = fmaf(a100.z - a000.z, t0, a000.z), gy1 = fmaf(a110.z - a010.z, t0, a010.z);
  const float gz0 = fmaf(a100.w - a000.w, t0, a000.w), gz1 = fmaf(a110.w - a010.w, t0, a010.w);
  float3 g = make_float3(fmaf(gx1 - gx0, t1, gx0), fmaf(gy1 - gy0, t1, gy0), fmaf(gz1 - gz0, t1, gz0));
  if (D == 3) {
    const float vx01 = fmaf(a101.x - a001.x, t0, a001.x);
    const float vx11 = fmaf(a111.x - a011.x, t0, a011.x);
    const float vxx1 = fmaf(vx11 - vx01, t1, vx01);
    v = fmaf(vxx1 - v, t2, v);
    const float hx0 = fmaf(a101.y - a001.y, t0, a001.y), hx1 = fmaf(a111.y - a011.y, t0, a011.y);
    const float hy0 = fmaf(a101.z - a001.z, t0, a001.z), hy1 = fmaf(a111.z - a011.z, t0, a011.z);
    const float hz0 = fmaf(a101.w - a001.w, t0, a001.w), hz1 = fmaf(a111.w - a011.w, t0, a011.w);
    const float s0 = fmaf(hx1 - hx0, t1, hx0), s1 = fmaf(hy1 - hy0, t1, hy0), s2 = fmaf(hz1 - hz0, t1, hz0);
    g.x = fmaf(s0 - g.x, t2, g.x);
    g.y = fmaf(s1 - g.y, t2, g.y);
    g.z = fmaf(s2 - g.z, t2, g.z);
  }
  *grad = g;
  return v;
}

// FP32 multilinear value of the moving image at a cell (corner element offsets
// cb + {0, cx} + {0, cy} + {0, cz}, weights t0..t2; every offset non-negative
// and inside the buffer).  TEX: gather the moving texels (value and gradient
// channels) instead of the moving values and return the interpolated gradient
// as well; the value is interpolated with the same operations either way, and
// every fast-path kernel interpolates through mlerp_value / mlerp_texels.
template <int D, typename P, bool TEX, bool BYTES = false>
__device__ __forceinline__ float fast_interp(const FastArgs& f, uint32_t cb, uint32_t cx, uint32_t cy, uint32_t cz,
                                             float t0, float t1, float t2, float3* grad) {
  // BYTES: the offsets are byte offsets into the array read
#define LDT(o) (BYTES ? P::ldtb(tex, (o)) : P::ldt32(tex, (o)))
#define LDP(o) (BYTES ? P::ldb(img, (o)) : P::ld32(img, (o)))
  if constexpr (TEX) {
    const typename P::texel* tex = static_cast<const typename P::texel*>(f.texels);
    const float4 a000 = LDT(cb), a100 = LDT(cb + cx);
    const float4 a010 = LDT(cb + cy), a110 = LDT(cb + cy + cx);
    float4 a001 = a000, a101 = a000, a011 = a000, a111 = a000;
    if (D == 3) {
      a001 = LDT(cb + cz);
      a101 = LDT(cb + cz + cx);
      a011 = LDT(cb + cz + cy);
      a111 = LDT(cb + cz + cy + cx);
    }
    return mlerp_texels<D>(a000, a100, a010, a110, a001, a101, a011, a111, t0, t1, t2, grad);
  } else {
    const typename P::pixel* img = static_cast<const typename P::pixel*>(f.moving_px);
    const float v000 = LDP(cb), v100 = LDP(cb + cx);
    const float v010 = LDP(cb + cy), v110 = LDP(cb + cy + cx);
    float v001 = 0.0f, v101 = 0.0f, v011 = 0.0f, v111 = 0.0f;
    if (D == 3) {
      v001 = LDP(cb + cz);
      v101 = LDP(cb + cz + cx);
      v011 = LDP(cb + cz + cy);
      v111 = LDP(cb + cz + cy + cx);
    }
    return mlerp_value<D>(v000, v100, v010, v110, v001, v101, v011, v111, t0, t1, t2);
  }
#undef LDT
#undef LDP
}

// The intensity part of a fast-path classification, after the coordinate
// part's out / uncertain flags: the moving value's range test, the moving and
// fixed Parzen bins with their certainty.  0: not counted, 1: counted, 2:
// uncertain.
__device__ __forceinline__ int fast_bins(const FastArgs& f, float v, float fv, bool out, bool uncertain, int& fbin,
                                         int& mbin, float& arg0, float* mterm_out) {
  out = out | (v < f.mmin - f.ev) | (v > f.mmax + f.ev);
  uncertain = uncertain | (v < f.mmin + f.ev) | (v > f.mmax - f.ev);
  const float mterm = fmaf(v, f.inv_mbs, -f.mnm);
  mbin = fast_bin(mterm, f.bt_m, f.bins, &uncertain);
  arg0 = (float)(mbin - 1) - mterm;
  if (mterm_out) *mterm_out = mterm;
  const float fterm = fmaf(fv, f.inv_fbs, -f.fnm);
  fbin = fast_bin(fterm, f.bt_f, f.bins, &uncertain);
  return out ? 0 : (uncertain ? 2 : 1);
}

// TEX: gather the moving texels (value and gradient channels) instead of the
// moving values, and return the interpolated gradient and the moving Parzen
// term as well; the value is interpolated with the same operations either way.
template <int D, typename P, int MK, bool TEX = false, bool MASK = false>
__device__ __forceinline__ int fast_classify(const FastArgs& f, const int32_t* idx, int& fbin, int& mbin, float& arg0,
                                             Cell& cell, bool* bsin, float3* grad = nullptr,
                                             float* mterm_out = nullptr) {
  double ci[3];
  fast_cindex<D, MK>(f, idx, ci, bsin);
  bool out = false, uncertain = false;
  int32_t bb[3] = {0, 0, 0};
  bool act[3] = {false, false, false};
  cell.d[2] = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double lo = (double)f.cilo[k], hi = (double)f.cihi[k];
    out = out | (ci[k] < lo - f.e_ci) | (ci[k] >= hi + f.e_ci);
    uncertain = uncertain | (ci[k] < lo + f.e_ci) | (ci[k] >= hi - f.e_ci);
    // moving cell (the reference's branch rules as duplicated corners)
    // floor, not ITK's Math::Floor: they differ only within an ulp below an
    // integer, where either cell holds ci and the interpolant is continuous
    // (its slope bound is part of ev)
    const double cc = fmin(fmax(ci[k], lo), hi);
    int32_t fl = (int32_t)floor(cc);
    fl = fl < f.mstart[k] ? f.mstart[k] : (fl > f.mend[k] ? f.mend[k] : fl);
    bb[k] = fl;
    cell.d[k] = ci[k] - (double)fl;
    act[k] = (cell.d[k] > 0.0) & (fl + 1 <= f.mend[k]);
  }
  if constexpr (MASK) {
    const double x = (double)idx[0], y = (double)idx[1], z = (double)idx[2];
    fast_mask<D>(f.fmask, x, y, z, &out, &uncertain);
    fast_mask<D>(f.mmask, x, y, z, &out, &uncertain);
  }
  cell.base = (bb[0] - f.mstart[0]) + f.msy * (bb[1] - f.mstart[1]) + (D == 3 ? f.msz * (bb[2] - f.mstart[2]) : 0);
  cell.ox = act[0] ? 1 : 0;
  cell.oy = act[1] ? f.msy : 0;
  cell.oz = (D == 3 && act[2]) ? f.msz : 0;
  const float t0 = (float)cell.d[0], t1 = (float)cell.d[1], t2 = (float)cell.d[2];
  // every cell offset is non-negative and below the buffer size
  const uint32_t cb = (uint32_t)cell.base, cx = (uint32_t)cell.ox, cy = (uint32_t)cell.oy, cz = (uint32_t)cell.oz;
  const float v = fast_interp<D, P, TEX>(f, cb, cx, cy, cz, t0, t1, t2, grad);
  const typename P::pixel* fimg = static_cast<const typename P::pixel*>(f.fixed_px);
  float fv;
  if (f.fixed_interp) {
    // fixed cell from the per-axis tables (the reference's continuous index
    // along each axis, computed on the host); FP32 multilinear value
    const int4 tx = f.ftab[0][idx[0] - f.vstart[0]];
    const int4 ty = f.ftab[1][idx[1] - f.vstart[1]];
    const int4 tz = D == 3 ? f.ftab[2][idx[2] - f.vstart[2]] : make_int4(0, 0, 0, 1);
    out = out | ((tx.w & ty.w & tz.w) == 0);
    const uint32_t fb = (uint32_t)(tx.x + ty.x + tz.x);
    const uint32_t fx = (uint32_t)tx.y, fy = (uint32_t)ty.y, fz = (uint32_t)tz.y;
    const float u0 = __int_as_float(tx.z), u1 = __int_as_float(ty.z);
    const float f000 = P::ld32(fimg, fb), f100 = P::ld32(fimg, fb + fx);
    const float f010 = P::ld32(fimg, fb + fy), f110 = P::ld32(fimg, fb + fy + fx);
    const float fx00 = fmaf(f100 - f000, u0, f000);
    const float fx10 = fmaf(f110 - f010, u0, f010);
    fv = fmaf(fx10 - fx00, u1, fx00);
    if (D == 3) {
      const float u2 = __int_as_float(tz.z);
      const float f001 = P::ld32(fimg, fb + fz), f101 = P::ld32(fimg, fb + fz + fx);
      const float f011 = P::ld32(fimg, fb + fz + fy), f111 = P::ld32(fimg, fb + fz + fy + fx);
      const float fx01 = fmaf(f101 - f001, u0, f001);
      const float fx11 = fmaf(f111 - f011, u0, f011);
      const float fxx1 = fmaf(fx11 - fx01, u1, fx01);
      fv = fmaf(fxx1 - fv, u2, fv);
    }
  } else {
    // the pixel the fixed sample sits on (verified on the host)
    const int32_t fl =
        (idx[0] + f.foff[0]) + f.fsy * (idx[1] + f.foff[1]) + (D == 3 ? f.fsz * (idx[2] + f.foff[2]) : 0);
    fv = P::ld32(fimg, (uint32_t)fl);
  }
  return fast_bins(f, v, fv, out, uncertain, fbin, mbin, arg0, TEX ? mterm_out : nullptr);
}

// The uncertain samples a pass-1 workgroup listed, on the exact FP64 path,
// into the workgroup's LDS histogram (copy 0) and valid count,
// at the end of pass 1 (affine: the exact path's registers do not raise the
// kernel's count; the samples' data is still in L2).
template <int D, typename P, int MK>
__device__ __forceinline__ void pass1_exceptions(const EvalArgs& a, const uint32_t* list, uint32_t ne,
                                                             unsigned long long* hist, unsigned int* nvalid) {
  const int B = a.bins;
  for (uint32_t t = threadIdx.x; t < ne; t += blockDim.x) {
    const uint32_t s = list[t];
    double vp[3];
    virtual_point<D>(a, s, vp);
    SampleOut<D> o;
    const bool ok = eval_sample<D, MK, P>(a, vp, o);
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)o.fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)o.mbin : (int16_t)-1;
    }
    if (!ok) continue;
    atomicAdd(nvalid, 1u);
    unsigned int q[4];
    parzen_fixed((float)((double)(o.mbin - 1) - o.mterm), q);
    unsigned long long* row = hist + o.fbin * B + (o.mbin - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) atomicAdd(&row[k], (unsigned long long)q[k]);
  }
}

// A pass-1 workgroup's accumulators in LDS: `copies` joint histograms (1, 2 or
// 4; lane i takes copy i mod copies), the valid count and the exception count.
// The fixed marginal is not accumulated: each counted sample's Parzen weights
// sum to 2^32, and finalize takes the counts from the histogram's rows.
struct P1Lds {
  unsigned long long* hist0;  // [copies][B][B]
  unsigned int* counters;     // [0] valid count, [1] exception count
  unsigned long long* hist;   // this lane's copy
  int copies;
};

__device__ __forceinline__ P1Lds pass1_begin(int B, int copies) {
  extern __shared__ unsigned long long lds_u64[];
  const int BB = B * B;
  P1Lds L;
  L.copies = copies;
  L.hist0 = lds_u64;
  L.counters = reinterpret_cast<unsigned int*>(lds_u64 + L.copies * BB);
  const int mine = (int)(threadIdx.x & (uint32_t)(L.copies - 1));
  L.hist = L.hist0 + mine * BB;
  for (int i = threadIdx.x; i < L.copies * BB; i += blockDim.x) L.hist0[i] = 0ull;
  if (threadIdx.x < 2) L.counters[threadIdx.x] = 0u;
  __syncthreads();
  return L;
}

// One classified sample s of pass 1 (c: 0 not counted, 1 counted, 2
// uncertain): its byte for pass 2, its list slot when uncertain, its recorded
// bins, and the Parzen weights of a counted sample in the LDS histogram.
__device__ __forceinline__ void pass1_add(const FastArgs& f, const P1Lds& L, uint32_t lb, uint32_t s, int c, int fbin,
                                          int mbin, float arg0, unsigned int& nvalid) {
  if (f.cls) f.cls[s] = (uint8_t)(c == 1 ? (uint32_t)fbin : (c == 2 ? kClsListed : kClsSkip));
  if (c == 2) {
    // listed for the exact path (here and in pass 2)
    const unsigned int slot = atomicAdd(&L.counters[1], 1u);
    if (slot < (unsigned int)kExcCap) f.exc_list[lb * kExcCap + slot] = s;
    return;
  }
  if (f.sample_fbin) {
    f.sample_fbin[s] = c == 1 ? (int16_t)fbin : (int16_t)-1;
    f.sample_mbin[s] = c == 1 ? (int16_t)mbin : (int16_t)-1;
  }
  if (c == 0) return;
  ++nvalid;
  unsigned int q[4];
  parzen_fixed_fast(arg0, q);
  // mbin is clamped to [2, B-3], so the 4 bins stay in the row; a zero
  // weight (clamped windows only) adds nothing
  unsigned long long* row = L.hist + __umul24((uint32_t)fbin, (uint32_t)f.bins) + (mbin - 1);
#pragma unroll
  for (int k = 0; k < 4; ++k) atomicAdd(&row[k], (unsigned long long)q[k]);
}

// End of pass 1: the valid count, the workgroup's listed samples on the exact
// path (affine), the LDS accumulators into the evaluation's.
template <int D, typename P, int MK>
__device__ __forceinline__ void pass1_end(const FastArgs& f, const EvalArgs& a, const P1Lds& L, uint32_t lb,
                                          unsigned int nvalid) {
  const int B = f.bins;
  const int BB = B * B;
  unsigned int v = nvalid;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&L.counters[0], v);
  __syncthreads();
  const unsigned int ne = L.counters[1];
  if constexpr (MK == 1) {
    // the list this workgroup wrote (visible to it after the barrier)
    if (ne)
      pass1_exceptions<D, P, MK>(a, f.exc_list + lb * kExcCap, min(ne, (unsigned int)kExcCap), L.hist0,
                                 &L.counters[0]);
    __syncthreads();
  }
  for (int i = threadIdx.x; i < BB; i += blockDim.x) {
    unsigned long long h = L.hist0[i];
    for (int c = 1; c < L.copies; ++c) h += L.hist0[c * BB + i];
    if (h) atomicAdd(&f.hist[i], h);
  }
  if (threadIdx.x == 0) {
    if (L.counters[0]) atomicAdd(f.nvalid, (unsigned long long)L.counters[0]);
    f.exc_count[lb] = ne;
    if (ne > (unsigned int)kExcCap) atomicOr(f.exc_overflow, 1u);
  }
}

// MASK: the affine path with image masks (kernels of their own, so that the
// unmasked kernels keep their registers).
template <int D, typename P, int MK, bool MASK = false>
__device__ __forceinline__ void pass1_fast_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  const P1Lds L = pass1_begin(f.bins, f.hist_copies);
  unsigned int nvalid = 0;
  uint32_t s, end, step;
  run_range<D>(f, lb, lg, s, end, step);
  for (; s < end; s += step) {
    int32_t idx[3];
    fast_index<D>(f, s, idx);
    int fbin = 0, mbin = 0;
    float arg0 = 0.0f;
    Cell cell;
    bool bsin;
    const int c = fast_classify<D, P, MK, false, MASK>(f, idx, fbin, mbin, arg0, cell, &bsin);
    pass1_add(f, L, lb, s, c, fbin, mbin, arg0, nvalid);
  }
  pass1_end<D, P, MK>(f, a, L, lb, nvalid);
}

// ---------------------------------------------------------------------------
// Row walks (affine moving transform, dense sampling, no masks, fixed samples
// on fixed pixel centres).  A wavefront takes one item at a time: 64
// consecutive x of `walk` consecutive rows.  Each lane maps its first sample
// with the FP64 affine map of fast_cindex, converts the continuous index to
// 32.32 fixed point and steps it down the column by the rounded y column of
// the map; the host's bound e_ci covers the map, the conversion and the steps.
// Per sample, the cell, the interpolation weights and the buffer tests are a
// few integer operations on the fixed-point index.
// ---------------------------------------------------------------------------

// The wavefront's items: XCD-contiguous shares (as run_range), the
// workgroup's wavefronts on consecutive items.  Wavefront-uniform.
__device__ __forceinline__ void rows_range(const FastArgs& f, uint32_t lb, uint32_t lg, uint32_t& it, uint32_t& end,
                                           uint32_t& step) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nw = blockDim.x >> 6;
  const uint32_t items = f.rows.items;
  if (f.xcd_chunks) {
    const uint32_t xcd = lb % kXcd, lbx = lb / kXcd, bpx = lg / kXcd;
    const uint32_t chunk = (items + kXcd - 1) / kXcd;
    const uint32_t c0 = xcd * chunk;
    it = c0 + lbx * nw + wave;
    end = min(items, c0 + chunk);
    step = bpx * nw;
  } else {
    it = lb * nw + wave;
    end = items;
    step = lg * nw;
  }
}

template <int D>
struct RowWalk {
  long long q[3];   // moving continuous index, 32.32 fixed point (floor in the high word)
  uint32_t s;       // sample index (virtual buffer order)
  uint32_t fp;      // fixed pixel byte offset
  int32_t x, y, z;  // virtual index of the lane's sample (y advances)
  int32_t rows;     // samples of the walk (wavefront-uniform)
  bool active;      // x inside the virtual row
};

// c in 32.32 fixed point, truncated (error below 2^-32); |c| clamped to 2^30
// (such samples are outside every buffer the fast path takes).
__device__ __forceinline__ long long to_fixed32(double c) {
  c = fmin(fmax(c, -1073741824.0), 1073741824.0);
  const double fl = floor(c);
  const double fr = fmin((c - fl) * 4294967296.0, 4294967295.0);
  return (long long)(((unsigned long long)(uint32_t)(int32_t)fl << 32) | (unsigned long long)(uint32_t)fr);
}

template <int D, typename P>
__device__ __forceinline__ void rows_start(const FastArgs& f, uint32_t item, RowWalk<D>& w) {
  const FastRows& R = f.rows;
  const uint32_t r = fdiv(item, R.dxseg);
  const uint32_t xs = item - r * R.dxseg.d;
  const uint32_t zl = D == 3 ? fdiv(r, R.dyblk) : 0u;
  const uint32_t yb = r - zl * R.dyblk.d;
  const int32_t xl = (int32_t)(xs * 64u + (threadIdx.x & 63u));
  const int32_t y0 = (int32_t)(yb * (uint32_t)R.walk);
  w.active = xl < R.nx;
  w.rows = min(R.walk, R.ny - y0);
  w.x = f.vstart[0] + xl;
  w.y = f.vstart[1] + y0;
  w.z = D == 3 ? f.vstart[2] + (int32_t)zl : 0;
  const int32_t idx[3] = {w.x, w.y, w.z};
  double ci[3];
  bool bsin;
  fast_cindex<D, 1>(f, idx, ci, &bsin);
#pragma unroll
  for (int k = 0; k < D; ++k) w.q[k] = to_fixed32(ci[k]);
  w.s = (uint32_t)xl + (uint32_t)R.nx * ((uint32_t)y0 + (uint32_t)R.ny * zl);
  w.fp = (uint32_t)((w.x + f.foff[0]) + f.fsy * (w.y + f.foff[1]) + (D == 3 ? f.fsz * (w.z + f.foff[2]) : 0)) *
         (uint32_t)sizeof(typename P::pixel);
}

template <int D, typename P>
__device__ __forceinline__ void rows_step(const FastArgs& f, RowWalk<D>& w) {
#pragma unroll
  for (int k = 0; k < D; ++k) w.q[k] += f.rows.qy[k];
  w.s += (uint32_t)f.rows.nx;
  w.fp += (uint32_t)f.fsy * (uint32_t)sizeof(typename P::pixel);
  ++w.y;
}

// The cell of a row-walk sample: buffer tests against the fixed-point
// thresholds, the cell from the high word (clamped into the buffer, the upper
// corner duplicated outside [start, end - 1] as in the reference's branch
// rules), the weights from the low word.  cb and off[] are byte offsets into
// the moving values (E bytes per element) or texels.
template <int D, uint32_t E>
__device__ __forceinline__ void rows_cell(const FastArgs& f, const RowWalk<D>& w, bool& out, bool& uncertain,
                                          uint32_t& cb, uint32_t* off, float* t) {
  const FastRows& R = f.rows;
  int32_t fl[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const long long q = w.q[k];
    out = out | ((unsigned long long)(q - R.lo_out[k]) >= R.span_out[k]);
    uncertain = uncertain | ((unsigned long long)(q - R.lo_in[k]) >= R.span_in[k]);
    const int32_t I = (int32_t)(q >> 32);
    fl[k] = min(max(I, f.mstart[k]), f.mend[k]);
    const bool act = (uint32_t)(I - f.mstart[k]) < (uint32_t)(f.mend[k] - f.mstart[k]);
    const uint32_t stride = k == 0 ? E : (k == 1 ? (uint32_t)f.msy * E : (uint32_t)f.msz * E);
    off[k] = act ? stride : 0u;
    t[k] = (float)(uint32_t)q * 2.3283064365386963e-10f;  // 2^-32
  }
  // 24-bit products: the host checks the moving extents and strides
  cb = (uint32_t)(fl[0] - f.mstart[0]) + __umul24((uint32_t)f.msy, (uint32_t)(fl[1] - f.mstart[1]));
  if (D == 3) cb += __umul24((uint32_t)f.msz, (uint32_t)(fl[2] - f.mstart[2]));
  cb *= E;
}

// fast_classify of a row-walk sample: the cell (rows_cell); interpolation and
// bins as fast_classify.
template <int D, typename P, bool TEX = false>
__device__ __forceinline__ int rows_classify(const FastArgs& f, const RowWalk<D>& w, int& fbin, int& mbin,
                                             float& arg0, float3* grad = nullptr) {
  // byte offsets into the moving values (or texels): scaled once per sample
  constexpr uint32_t E = TEX ? (uint32_t)sizeof(typename P::texel) : (uint32_t)sizeof(typename P::pixel);
  bool out = false, uncertain = false;
  uint32_t cb = 0u, off[3] = {0u, 0u, 0u};
  float t[3] = {0.0f, 0.0f, 0.0f};
  rows_cell<D, E>(f, w, out, uncertain, cb, off, t);
  const float v = fast_interp<D, P, TEX, true>(f, cb, off[0], off[1], off[2], t[0], t[1], t[2], grad);
  // pass 2 (TEX) takes the fixed bin from pass 1's byte: no fixed read
  const float fv = TEX ? 0.0f : P::ldb(static_cast<const typename P::pixel*>(f.fixed_px), w.fp);
  return fast_bins(f, v, fv, out, uncertain, fbin, mbin, arg0, nullptr);
}

template <int D, typename P>
__device__ __forceinline__ void pass1_rows_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  const P1Lds L = pass1_begin(f.bins, f.hist_copies);
  unsigned int nvalid = 0;
  uint32_t it, iend, istep;
  rows_range(f, lb, lg, it, iend, istep);
  for (; it < iend; it += istep) {
    RowWalk<D> w;
    rows_start<D, P>(f, it, w);
    for (int j = 0; j < w.rows; ++j, rows_step<D, P>(f, w)) {
      if (!w.active) continue;
      int fbin = 0, mbin = 0;
      float arg0 = 0.0f;
      const int c = rows_classify<D, P>(f, w, fbin, mbin, arg0);
      pass1_add(f, L, lb, w.s, c, fbin, mbin, arg0, nvalid);
    }
  }
  pass1_end<D, P, 1>(f, a, L, lb, nvalid);
}

// ---------------------------------------------------------------------------
// Row walks with the exact fixed bins (FastArgs::fbin8).  The fixed side of a
// sample is one byte, and the moving value needs no range test (the moving
// range is the min/max of every value the interpolation reads).  The moving
// values and texels are read from copies with a one-voxel replicated border
// (FastArgs::mpad, tpad), so a sample certainly inside the moving buffer
// needs no clamped cell: its lower corner lies in [start - 1, end] and the
// border holds the duplicates of the reference's branch rules.  The moving
// continuous index is monotone along a walk, so the walk's first and last
// samples give each lane's class: every sample certainly inside, or every
// sample certainly outside (one axis below or above the buffer at both
// ends), or neither.  A wavefront without a lane of the third class takes a
// loop with no buffer test (outside lanes idle); the others test each sample
// (the inside test, then, for the few that fail it, the outside test).
// Either loop gives a sample the same class, value and bins, in both passes.
// ---------------------------------------------------------------------------

// Parzen weights of an unclamped window, t = mterm - mbin in [0, 1): bins
// mbin - 1 .. mbin + 2 take B3(1 + t), B3(t), B3(1 - t), B3(2 - t), that is
// u^3/6, (3t^3 - 6t^2 + 4)/6, (3u^3 - 6u^2 + 4)/6, t^3/6 with u = 1 - t, as
// 2^32 fixed point truncated; the second takes the remainder, so the four sum
// to exactly 2^32.
__device__ __forceinline__ void parzen_fixed_t(float t, unsigned int* q) {
  constexpr float C = 4294967296.0f;
  const float u = 1.0f - t;
  const float t2 = t * t, u2 = u * u;
  q[0] = (unsigned int)(u2 * (u * (C / 6.0f)));
  q[2] = (unsigned int)fmaf(u2, fmaf(0.5f * C, u, -C), (2.0f / 3.0f) * C);
  q[3] = (unsigned int)(t2 * (t * (C / 6.0f)));
  q[1] = 0u - q[0] - q[2] - q[3];
}

// The moving bin of a fast value, its certainty (fast_bin) and the Parzen
// fraction t = mterm - mbin (exact by Sterbenz's lemma for mterm in [mbin,
// mbin + 1)); t leaves [0, 1) only where the term was clamped.
__device__ __forceinline__ int fb_bin(const FastArgs& f, float v, bool* uncertain, float* t) {
  const float mterm = fmaf(v, f.inv_mbs, -f.mnm);
  const int mbin = fast_bin(mterm, f.bt_m, f.bins, uncertain);
  *t = mterm - (float)mbin;
  return mbin;
}

// The buffer tests of the padded walks, in fixed point relative to the padded
// origin (start - 1 on every axis): [in_lo, in_lo + in_span) certainly inside,
// below out_lo or at out_hi and above certainly outside.
template <int D>
struct PadTests {
  long long in_lo[3], out_lo[3], out_hi[3];
  unsigned long long in_span[3];
  long long shift[3];  // (start - 1) * 2^32
};

template <int D>
__device__ __forceinline__ PadTests<D> pad_tests(const FastArgs& f) {
  PadTests<D> T;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    T.shift[k] = (long long)((unsigned long long)(uint32_t)(f.mstart[k] - 1) << 32);
    T.in_lo[k] = f.rows.lo_in[k] - T.shift[k];
    T.in_span[k] = f.rows.span_in[k];
    T.out_lo[k] = f.rows.lo_out[k] - T.shift[k];
    T.out_hi[k] = T.out_lo[k] + (long long)f.rows.span_out[k];
  }
  return T;
}

// The lane's walk relative to the padded origin and its class: 0 every sample
// certainly inside, 1 every sample certainly outside, 2 neither.
template <int D>
__device__ __forceinline__ int pad_walk(const FastArgs& f, const PadTests<D>& T, const RowWalk<D>& w, long long* q) {
  bool in = true, out = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    q[k] = w.q[k] - T.shift[k];
    const long long q1 = q[k] + (long long)(w.rows - 1) * f.rows.qy[k];
    in = in & ((unsigned long long)(q[k] - T.in_lo[k]) < T.in_span[k]) &
         ((unsigned long long)(q1 - T.in_lo[k]) < T.in_span[k]);
    out = out | ((q[k] < T.out_lo[k]) & (q1 < T.out_lo[k])) | ((q[k] >= T.out_hi[k]) & (q1 >= T.out_hi[k]));
  }
  return in ? 0 : (out ? 1 : 2);
}

template <int D>
__device__ __forceinline__ bool pad_inside(const PadTests<D>& T, const long long* q) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < D; ++k) in = in & ((unsigned long long)(q[k] - T.in_lo[k]) < T.in_span[k]);
  return in;
}

template <int D>
__device__ __forceinline__ bool pad_outside(const PadTests<D>& T, const long long* q) {
  bool out = false;
#pragma unroll
  for (int k = 0; k < D; ++k) out = out | (q[k] < T.out_lo[k]) | (q[k] >= T.out_hi[k]);
  return out;
}

// The cell of a certainly inside sample in the padded copy: byte offset of its
// lower corner and the weights.
template <int D, uint32_t E>
__device__ __forceinline__ uint32_t pad_cell(const FastArgs& f, const long long* q, float* t) {
  uint32_t cb = (uint32_t)(q[0] >> 32) + __umul24((uint32_t)f.psy, (uint32_t)(q[1] >> 32));
  if (D == 3) cb += __umul24((uint32_t)f.psz, (uint32_t)(q[2] >> 32));
#pragma unroll
  for (int k = 0; k < D; ++k) t[k] = (float)(uint32_t)q[k] * 2.3283064365386963e-10f;  // 2^-32
  return cb * E;
}

// The loads of one certainly inside sample of a padded walk: its fixed bin,
// its cell's eight corners (TEX: texels) and weights.
template <int D, typename P, bool TEX>
struct PadLoad {
  using V = typename std::conditional<TEX, float4, float>::type;
  V c[8];
  float t[3];
  uint32_t fb;
};

template <int D, typename P, bool TEX>
__device__ __forceinline__ void pad_load(const FastArgs& f, uint32_t s, const long long* q, uint32_t sy, uint32_t sz,
                                         PadLoad<D, P, TEX>& L) {
  using T = typename std::conditional<TEX, typename P::texel, typename P::pixel>::type;
  constexpr uint32_t E = (uint32_t)sizeof(T);
  L.fb = f.fbin8[s];
  L.t[0] = L.t[1] = L.t[2] = 0.0f;
  const uint32_t cb = pad_cell<D, E>(f, q, L.t);
  const char* p00 = static_cast<const char*>(TEX ? f.tpad : f.mpad) + cb;
  const char* p10 = p00 + sy;
  auto ld = [](const char* p, uint32_t o) {
    if constexpr (TEX)
      return P::ldtb(reinterpret_cast<const T*>(p), o);
    else
      return P::ldb(reinterpret_cast<const T*>(p), o);
  };
  L.c[0] = ld(p00, 0u);
  L.c[1] = ld(p00, E);
  L.c[2] = ld(p10, 0u);
  L.c[3] = ld(p10, E);
  if (D == 3) {
    const char* p01 = p00 + sz;
    const char* p11 = p10 + sz;
    L.c[4] = ld(p01, 0u);
    L.c[5] = ld(p01, E);
    L.c[6] = ld(p11, 0u);
    L.c[7] = ld(p11, E);
  } else {
    L.c[4] = L.c[5] = L.c[6] = L.c[7] = L.c[0];
  }
}

// The multilinear value (TEX: and gradient) of loaded corners.
template <int D, typename P, bool TEX>
__device__ __forceinline__ float pad_value(const PadLoad<D, P, TEX>& L, float3* grad) {
  if constexpr (TEX)
    return mlerp_texels<D>(L.c[0], L.c[1], L.c[2], L.c[3], L.c[4], L.c[5], L.c[6], L.c[7], L.t[0], L.t[1], L.t[2],
                           grad);
  else
    return mlerp_value<D>(L.c[0], L.c[1], L.c[2], L.c[3], L.c[4], L.c[5], L.c[6], L.c[7], L.t[0], L.t[1], L.t[2]);
}

// pass1_add with the Parzen fraction t instead of arg0.
__device__ __forceinline__ void pass1_add_t(const FastArgs& f, const P1Lds& L, uint32_t lb, uint32_t s, int c,
                                           uint32_t fbin, int mbin, float t, unsigned int& nvalid) {
  if (f.cls) f.cls[s] = (uint8_t)(c == 1 ? fbin : (c == 2 ? kClsListed : kClsSkip));
  if (c == 2) {
    const unsigned int slot = atomicAdd(&L.counters[1], 1u);
    if (slot < (unsigned int)kExcCap) f.exc_list[lb * kExcCap + slot] = s;
    return;
  }
  if (f.sample_fbin) {
    f.sample_fbin[s] = c == 1 ? (int16_t)fbin : (int16_t)-1;
    f.sample_mbin[s] = c == 1 ? (int16_t)mbin : (int16_t)-1;
  }
  if (c == 0) return;
  ++nvalid;
  unsigned int q[4];
  if (t >= 0.0f && t < 1.0f)
    parzen_fixed_t(t, q);
  else
    parzen_fixed(-1.0f - t, q);
  // mbin is clamped to [2, B-3], so the 4 bins stay in the row
  unsigned long long* row = L.hist + __umul24(fbin, (uint32_t)f.bins) + (mbin - 1);
#pragma unroll
  for (int k = 0; k < 4; ++k) atomicAdd(&row[k], (unsigned long long)q[k]);
}

// Pass 1 of a certainly inside sample from its loads: value, bins, class,
// accumulation.
template <int D, typename P>
__device__ __forceinline__ void pass1_pad_compute(const FastArgs& f, const P1Lds& L, uint32_t lb, uint32_t s,
                                                  const PadLoad<D, P, false>& x, unsigned int& nvalid) {
  const float v = pad_value<D, P, false>(x, nullptr);
  bool unc = false;
  float t;
  const int mbin = fb_bin(f, v, &unc, &t);
  pass1_add_t(f, L, lb, s, x.fb == kClsSkip ? 0 : (unc ? 2 : 1), x.fb, mbin, t, nvalid);
}

template <int D, typename P>
__device__ __forceinline__ void pass1_pad_sample(const FastArgs& f, const P1Lds& L, uint32_t lb, uint32_t s,
                                                 const long long* q, uint32_t sy, uint32_t sz, unsigned int& nvalid) {
  PadLoad<D, P, false> x;
  pad_load<D, P, false>(f, s, q, sy, sz, x);
  pass1_pad_compute<D, P>(f, L, lb, s, x, nvalid);
}

template <int D, typename P>
__device__ __forceinline__ void pass1_rows_fb_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  const P1Lds L = pass1_begin(f.bins, f.hist_copies);
  unsigned int nvalid = 0;
  constexpr uint32_t E = (uint32_t)sizeof(typename P::pixel);
  const uint32_t sy = (uint32_t)f.psy * E, sz = (uint32_t)f.psz * E;
  const uint32_t nx = (uint32_t)f.rows.nx;
  const PadTests<D> T = pad_tests<D>(f);
  uint32_t it, iend, istep;
  rows_range(f, lb, lg, it, iend, istep);
  for (; it < iend; it += istep) {
    RowWalk<D> w;
    rows_start<D, P>(f, it, w);
    long long q[3] = {0, 0, 0};
    const int cls = w.active ? pad_walk<D>(f, T, w, q) : 1;
    uint32_t s = w.s;
    if (__all(cls != 2)) {
      if (cls == 0) {
        for (int j = 0; j < w.rows; ++j, s += nx) {
          pass1_pad_sample<D, P>(f, L, lb, s, q, sy, sz, nvalid);
#pragma unroll
          for (int k = 0; k < D; ++k) q[k] += f.rows.qy[k];
        }
      } else if (w.active && (f.cls || f.sample_fbin)) {
        // a walk outside the buffer: its records say not counted
        for (int j = 0; j < w.rows; ++j, s += nx) pass1_add_t(f, L, lb, s, 0, 0u, 0, 0.0f, nvalid);
      }
    } else {
      for (int j = 0; j < w.rows; ++j, s += nx) {
        if (w.active) {
          if (pad_inside<D>(T, q))
            pass1_pad_sample<D, P>(f, L, lb, s, q, sy, sz, nvalid);
          else
            pass1_add_t(f, L, lb, s, (pad_outside<D>(T, q) || f.fbin8[s] == kClsSkip) ? 0 : 2, 0u, 0, 0.0f, nvalid);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) q[k] += f.rows.qy[k];
      }
    }
  }
  pass1_end<D, P, 1>(f, a, L, lb, nvalid);
}

// Exceptions of kExcListsPerGroup consecutive pass-1 lists per workgroup:
// the lists' counts and prefix in LDS, then one listed sample per thread.
constexpr int kExcListsPerGroup = 8;

struct ExcLists {
  uint32_t first;                      // first list of the workgroup
  uint32_t pre[kExcListsPerGroup + 1];  // prefix of the clamped list lengths
};

__device__ __forceinline__ void exc_lists_setup(const FastArgs& f, uint32_t lb, ExcLists& e) {
  if (threadIdx.x == 0) {
    e.first = lb * kExcListsPerGroup;
    uint32_t acc = 0;
    for (int l = 0; l < kExcListsPerGroup; ++l) {
      e.pre[l] = acc;
      acc += min(f.exc_count[e.first + l], (uint32_t)kExcCap);
    }
    e.pre[kExcListsPerGroup] = acc;
  }
  __syncthreads();
}

// List slot (list * kExcCap + position) of the workgroup's t-th listed sample.
__device__ __forceinline__ uint32_t exc_lists_slot(const ExcLists& e, uint32_t t) {
  int l = 0;
  while (t >= e.pre[l + 1]) ++l;
  return (e.first + l) * kExcCap + (t - e.pre[l]);
}

__device__ __forceinline__ uint32_t exc_lists_sample(const FastArgs& f, const ExcLists& e, uint32_t t) {
  return f.exc_list[exc_lists_slot(e, t)];
}

// Exact FP64 evaluation of the listed samples of pass 1, as a kernel of its
// own (B-spline: its exact path would raise pass 1's register count if it ran
// in pass 1's tail).  Histogram bins take global atomics (integer sums, so the
// order does not matter); the valid count is summed in LDS first.
template <int D, typename P, int MK>
__device__ __forceinline__ void exceptions_pass1_body(const EvalArgs& a, const FastArgs& f, uint32_t lb) {
  __shared__ ExcLists e;
  __shared__ unsigned int n_ok;
  const int B = a.bins;
  if (threadIdx.x == 0) n_ok = 0u;
  exc_lists_setup(f, lb, e);
  const uint32_t ne = e.pre[kExcListsPerGroup];
  if (ne == 0) return;  // uniform over the workgroup
  for (uint32_t t = threadIdx.x; t < ne; t += blockDim.x) {
    const uint32_t s = exc_lists_sample(f, e, t);
    double vp[3];
    virtual_point<D>(a, s, vp);
    SampleOut<D> o;
    const bool ok = eval_sample<D, MK, P>(a, vp, o);
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)o.fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)o.mbin : (int16_t)-1;
    }
    if constexpr (MK == 2) {
      // what pass 2 needs of this sample, at its list slot
      ExcResult r;
      r.fbin = -1;
      r.mbin = 0;
      r.arg = 0.0f;
      r.g[0] = r.g[1] = r.g[2] = 0.0f;
      if (ok && o.sup.inside) {
        r.fbin = o.fbin;
        r.mbin = o.mbin;
        r.arg = (float)((double)(o.mbin - 1) - o.mterm);
        const float3 g = interp_gradient3<D, P>(static_cast<const typename P::texel*>(a.texels), o.cell);
        r.g[0] = g.x;
        r.g[1] = g.y;
        r.g[2] = g.z;
      }
      f.exc_res[exc_lists_slot(e, t)] = r;
    }
    if (!ok) continue;
    atomicAdd(&n_ok, 1u);
    unsigned int q[4];
    parzen_fixed((float)((double)(o.mbin - 1) - o.mterm), q);
    unsigned long long* row = a.hist + o.fbin * B + (o.mbin - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (q[k]) atomicAdd(&row[k], (unsigned long long)q[k]);
  }
  __syncthreads();
  if (threadIdx.x == 0 && n_ok) atomicAdd(a.nvalid, (unsigned long long)n_ok);
}

// Derivative moments in index space, per sample:
//   M[b][j] += w * g_b * idx_j,  M1[b] += w * g_b
// The physical affine derivative is J(M) = V M + c0 M1 (fast_derivative_kernel).
template <int D>
__device__ __forceinline__ void add_moments(double* acc, float w, float3 g, const int32_t* idx) {
  const double gw[3] = {(double)(w * g.x), (double)(w * g.y), (double)(w * g.z)};
#pragma unroll
  for (int b = 0; b < D; ++b) {
#pragma unroll
    for (int j = 0; j < D; ++j) acc[b * D + j] = fma(gw[b], (double)idx[j], acc[b * D + j]);
    acc[D * D + b] += gw[b];
  }
}

template <int D>
__device__ __forceinline__ void add_moments_d(double* acc, float w, const double* g, const int32_t* idx) {
#pragma unroll
  for (int b = 0; b < D; ++b) {
    const double gw = (double)w * g[b];
#pragma unroll
    for (int j = 0; j < D; ++j) acc[b * D + j] = fma(gw, (double)idx[j], acc[b * D + j]);
    acc[D * D + b] += gw;
  }
}

// Exception samples of one pass-1 workgroup, in sample order (bitonic sort of
// the list in LDS), summed in a fixed tree: deterministic.  CD: the gradient
// by central differences of the moving image (else the gradient image), and
// the calling lane's nd deferred samples at defer (pass2_rows_cd_body) as
// well, in the list's order.
template <int D, typename P, int MK, bool CD = false>
__device__ __forceinline__ void exceptions_pass2_body(const EvalArgs& a, const FastArgs& f, uint32_t lb,
                                                      uint32_t lg, const uint32_t* defer = nullptr,
                                                      uint32_t nd = 0) {
  __shared__ uint32_t keys[256];
  __shared__ double scratch[4 * 12];
  constexpr int PA = D * D + D;
  const uint32_t ne = min(f.exc_count[lb], (uint32_t)f.exc_cap);
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  if (ne > 0) {
    keys[threadIdx.x] = threadIdx.x < ne ? f.exc_list[lb * f.exc_cap + threadIdx.x] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2; k <= 256; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        const uint32_t i = threadIdx.x, l = i ^ j;
        if (l > i) {
          const uint32_t ki = keys[i], kl = keys[l];
          const bool up = (i & k) == 0;
          if ((ki > kl) == up) {
            keys[i] = kl;
            keys[l] = ki;
          }
        }
        __syncthreads();
      }
    }
    if (threadIdx.x < ne) {
      // the fast path samples the virtual voxels: the moments need the index
      const uint32_t s = keys[threadIdx.x];
      int32_t idx[3];
      sample_index<D>(a.virt, s, idx);
      double vp[3];
      index_to_point<D>(a.virt, idx, vp);
      SampleOut<D> o;
      if (eval_sample<D, 1, P>(a, vp, o)) {
        const float w = parzen_deriv_weight((float)((double)(o.mbin - 1) - o.mterm),
                                            f.pratio_f + o.fbin * a.bins + (o.mbin - 1));
        if (w != 0.0f) {
          if constexpr (CD) {
            double g[3] = {0.0, 0.0, 0.0};
            central_difference<D, P>(a, o.mp, g);
            add_moments_d<D>(acc, w, g, idx);
          } else {
            add_moments<D>(acc, w, interp_gradient3<D, P>(static_cast<const typename P::texel*>(a.texels), o.cell),
                           idx);
          }
        }
      }
    }
  }
  if constexpr (CD) {
    for (uint32_t k = 0; k < nd; ++k) {
      const uint32_t s = defer[k];
      int32_t idx[3];
      sample_index<D>(a.virt, s, idx);
      double vp[3];
      index_to_point<D>(a.virt, idx, vp);
      SampleOut<D> o;
      if (!eval_sample<D, 1, P>(a, vp, o)) continue;
      const float w = parzen_deriv_weight((float)((double)(o.mbin - 1) - o.mterm),
                                          f.pratio_f + o.fbin * a.bins + (o.mbin - 1));
      if (w == 0.0f) continue;
      double g[3] = {0.0, 0.0, 0.0};
      central_difference<D, P>(a, o.mp, g);
      add_moments_d<D>(acc, w, g, idx);
    }
  }
  block_reduce_array<PA>(acc, scratch, f.partials + (int64_t)(lg + lb) * PA);
}

// TEX: the moving values come from the texels' value channel, gathered with
// the gradient channels (one load per corner instead of two); the same
// values, so the same classification and results.
// CLS: pass 1's byte per sample says which samples count and their fixed
// bins, so pass 2 skips the others before any gather and drops its bound
// tests and fixed-image reads (the classifier's unused results are not
// computed); the moving side is evaluated as in pass 1, so the bits agree.
template <int D, typename P, int MK, bool TEX = false, bool CLS = false, bool MASK = false>
__device__ __forceinline__ void pass2_fast_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  extern __shared__ float lds_f32[];
  const int B = f.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  float* pr = lds_f32 + 128;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = f.pratio_f[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  uint32_t s, end, step;
  run_range<D>(f, lb, lg, s, end, step);
  for (; s < end; s += step) {
    uint32_t cb = 0;
    if constexpr (CLS) {
      cb = f.cls[s];
      if (cb >= kClsListed) {
        if constexpr (MK == 2) {
          if (cb == kClsSkip) {
#pragma unroll
            for (int d = 0; d < D; ++d) f.bspline_g[(int64_t)d * f.n + s] = 0.0f;
          }
        }
        continue;
      }
    }
    int32_t idx[3];
    fast_index<D>(f, s, idx);
    int fbin = 0, mbin = 0;
    float arg0 = 0.0f;
    Cell cell;
    bool bsin;
    float3 gt = make_float3(0.0f, 0.0f, 0.0f);
    float mterm_unused;
    int c = TEX ? fast_classify<D, P, MK, true, MASK>(f, idx, fbin, mbin, arg0, cell, &bsin, &gt, &mterm_unused)
                : fast_classify<D, P, MK, false, MASK>(f, idx, fbin, mbin, arg0, cell, &bsin);
    if constexpr (CLS) {
      c = 1;
      fbin = (int)cb;
    }
    auto gradient = [&]() {
      return TEX ? gt : interp_gradient3<D, P, true>(static_cast<const typename P::texel*>(f.texels), cell);
    };
    if constexpr (MK == 2) {
      // per-sample factors of the separable B-spline derivative, for every
      // sample (uncertain ones are rewritten by the exception kernel)
      float gw[3] = {0.0f, 0.0f, 0.0f};
      if (c == 1 && bsin) {
        const float w = parzen_deriv_weight_fast(arg0, pr + fbin * B + (mbin - 1));
        if (w != 0.0f) {
          const float3 g = gradient();
          gw[0] = (float)((double)w * (double)g.x);
          gw[1] = (float)((double)w * (double)g.y);
          gw[2] = (float)((double)w * (double)g.z);
        }
      }
      // listed samples are written by the tail below from pass 1's results
      if (c != 2) {
#pragma unroll
        for (int d = 0; d < D; ++d) f.bspline_g[(int64_t)d * f.n + s] = gw[d];
      }
    } else {
      if (c != 1) continue;
      const float w = parzen_deriv_weight_fast(arg0, pr + fbin * B + (mbin - 1));
      if (w == 0.0f) continue;
      add_moments<D>(acc, w, gradient(), idx);
    }
  }
  if constexpr (MK == 1) {
    block_reduce_array<PA>(acc, scratch, f.partials + (int64_t)lb * PA);
    // this workgroup's listed samples, exactly, into their own partial row
    exceptions_pass2_body<D, P, 1>(a, f, lb, lg);
  } else {
    // this workgroup's listed samples (the same partition as pass 1): their
    // per-sample factors from the exact results pass 1's exception kernel
    // stored, with the exact path's weight and product
    const uint32_t ne = min(f.exc_count[lb], (uint32_t)kExcCap);
    for (uint32_t t = threadIdx.x; t < ne; t += blockDim.x) {
      const uint32_t slot = lb * kExcCap + t;
      const uint32_t s = f.exc_list[slot];
      const ExcResult r = f.exc_res[slot];
      float gw[3] = {0.0f, 0.0f, 0.0f};
      if (r.fbin >= 0) {
        const float w = parzen_deriv_weight(r.arg, f.pratio_f + r.fbin * B + (r.mbin - 1));
        if (w != 0.0f) {
#pragma unroll
          for (int d = 0; d < D; ++d) gw[d] = (float)((double)w * (double)r.g[d]);
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) f.bspline_g[(int64_t)d * f.n + s] = gw[d];
    }
  }
}

// Pass 2 over row walks (after a row-walk pass 1 that wrote its byte per
// sample, same items per workgroup): per counted sample the moving value and
// gradient from the texels, the Parzen derivative weight w, and the
// index-space moments.  Along a walk x and z are fixed and y advances, so a
// lane sums S_b = sum w g_b and Y_b = sum y w g_b in FP32 and adds
// M[b][x] += x S_b, M[b][y] += Y_b, M[b][z] += z S_b, M1[b] += S_b in FP64 at
// the end of the walk.
template <int D, typename P>
__device__ __forceinline__ void pass2_rows_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  extern __shared__ float lds_f32[];
  const int B = f.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  float* pr = lds_f32 + 128;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = f.pratio_f[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  uint32_t it, iend, istep;
  rows_range(f, lb, lg, it, iend, istep);
  for (; it < iend; it += istep) {
    RowWalk<D> w;
    rows_start<D, P>(f, it, w);
    float sg[3] = {0.0f, 0.0f, 0.0f}, sy[3] = {0.0f, 0.0f, 0.0f};
    float yf = (float)w.y;
    for (int j = 0; j < w.rows; ++j, rows_step<D, P>(f, w), yf += 1.0f) {
      if (!w.active) continue;
      const uint32_t cb = f.cls[w.s];
      if (cb >= kClsListed) continue;
      int fbin, mbin;
      float arg0;
      float3 g;
      rows_classify<D, P, true>(f, w, fbin, mbin, arg0, &g);
      const float wt = parzen_deriv_weight_fast(arg0, pr + (int)cb * B + (mbin - 1));
      const float gw[3] = {wt * g.x, wt * g.y, wt * g.z};
#pragma unroll
      for (int b = 0; b < D; ++b) {
        sg[b] += gw[b];
        sy[b] = fmaf(yf, gw[b], sy[b]);
      }
    }
    const double x = (double)w.x, z = (double)w.z;
#pragma unroll
    for (int b = 0; b < D; ++b) {
      acc[b * D + 0] = fma((double)sg[b], x, acc[b * D + 0]);
      acc[b * D + 1] += (double)sy[b];
      if (D == 3) acc[b * D + 2] = fma((double)sg[b], z, acc[b * D + 2]);
      acc[D * D + b] += (double)sg[b];
    }
  }
  block_reduce_array<PA>(acc, scratch, f.partials + (int64_t)lb * PA);
  // this workgroup's listed samples, exactly, into their own partial row
  exceptions_pass2_body<D, P, 1>(a, f, lb, lg);
}

// Parzen derivative weight from the fraction t (parzen_deriv_weight_fast).
__device__ __forceinline__ float parzen_deriv_weight_of_t(float t, const float* prow) {
  if (t >= 0.0f && t < 1.0f) return parzen_deriv_weight_t(t, prow);
  return parzen_deriv_weight(-1.0f - t, prow);
}

// Pass 2 of a certainly inside sample from its loads: when pass 1 counted it
// as certain, its Parzen derivative weight times the interpolated gradient
// into the walk's FP32 sums (pass2_rows_body).  The fixed bin's byte is only
// tested once the value is interpolated, and a sample that does not count
// adds an exact zero: the byte and the texels are then in flight together
// (a test before the value would let the texel loads wait for the byte).
// Every sample here is certainly inside the moving buffer, so its texel loads
// are in bounds whatever its byte says.
template <int D, typename P>
__device__ __forceinline__ void pass2_pad_compute(const FastArgs& f, const float* pr, const PadLoad<D, P, true>& x,
                                                  float yf, float* sg, float* sy_) {
  float3 g;
  const float v = pad_value<D, P, true>(x, &g);
  bool unc = false;
  float t;
  const int mbin = fb_bin(f, v, &unc, &t);
  const bool counted = x.fb != kClsSkip && !unc;
  // a row of the table for every sample (mbin is within [2, B - 3])
  const uint32_t row = counted ? x.fb : 0u;
  const float wt = parzen_deriv_weight_of_t(t, pr + __umul24(row, (uint32_t)f.bins) + (mbin - 1));
  const float gw[3] = {counted ? wt * g.x : 0.0f, counted ? wt * g.y : 0.0f, counted ? wt * g.z : 0.0f};
#pragma unroll
  for (int b = 0; b < D; ++b) {
    sg[b] += gw[b];
    sy_[b] = fmaf(yf, gw[b], sy_[b]);
  }
}

// Pass 2 over row walks with the exact fixed bins: it classifies each sample
// as pass 1 did (the same walks, classes and arithmetic), so it needs no byte
// from pass 1; counted certain samples add their moments as pass2_rows_body.
template <int D, typename P>
__device__ __forceinline__ void pass2_rows_fb_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  extern __shared__ float lds_f32[];
  const int B = f.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  float* pr = lds_f32 + 128;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = f.pratio_f[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  constexpr uint32_t E = (uint32_t)sizeof(typename P::texel);
  const uint32_t sy = (uint32_t)f.psy * E, sz = (uint32_t)f.psz * E;
  const uint32_t nx = (uint32_t)f.rows.nx;
  const PadTests<D> T = pad_tests<D>(f);
  uint32_t it, iend, istep;
  rows_range(f, lb, lg, it, iend, istep);
  for (; it < iend; it += istep) {
    RowWalk<D> w;
    rows_start<D, P>(f, it, w);
    long long q[3] = {0, 0, 0};
    const int cls = w.active ? pad_walk<D>(f, T, w, q) : 1;
    float sg[3] = {0.0f, 0.0f, 0.0f}, sy_[3] = {0.0f, 0.0f, 0.0f};
    float yf = (float)w.y;
    uint32_t s = w.s;
    if (__all(cls != 2)) {
      if (cls == 0) {
        for (int j = 0; j < w.rows; ++j, s += nx, yf += 1.0f) {
          PadLoad<D, P, true> x;
          pad_load<D, P, true>(f, s, q, sy, sz, x);
          pass2_pad_compute<D, P>(f, pr, x, yf, sg, sy_);
#pragma unroll
          for (int k = 0; k < D; ++k) q[k] += f.rows.qy[k];
        }
      }
    } else {
      for (int j = 0; j < w.rows; ++j, s += nx, yf += 1.0f) {
        if (w.active && pad_inside<D>(T, q)) {
          PadLoad<D, P, true> x;
          pad_load<D, P, true>(f, s, q, sy, sz, x);
          pass2_pad_compute<D, P>(f, pr, x, yf, sg, sy_);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) q[k] += f.rows.qy[k];
      }
    }
    // the walk's sums into the thread's FP64 moments (x and z are the walk's)
    const double x = (double)w.x, z = (double)w.z;
#pragma unroll
    for (int b = 0; b < D; ++b) {
      acc[b * D + 0] = fma((double)sg[b], x, acc[b * D + 0]);
      acc[b * D + 1] += (double)sy_[b];
      if (D == 3) acc[b * D + 2] = fma((double)sg[b], z, acc[b * D + 2]);
      acc[D * D + b] += (double)sg[b];
    }
  }
  block_reduce_array<PA>(acc, scratch, f.partials + (int64_t)lb * PA);
  // this workgroup's listed samples, exactly, into their own partial row
  exceptions_pass2_body<D, P, 1>(a, f, lb, lg);
}

// Pass 2 over row walks with central differences of the moving image
// (CentralDifferenceImageFunction::EvaluateSpecialized,
// itkCentralDifferenceImageFunction.hxx:252-318) on a moving geometry with a
// diagonal physical-to-index map.  Along axis b the reference's points at
// -/+ half a spacing are the sample's continuous index -/+ 1/2 along index
// axis b (the other coordinates unchanged); their cells along b take the
// positions P0, P1, P3 with P1 the upper (t >= 1/2) or lower (t < 1/2) corner
// of the sample's own cell, so the sample's 2^D corner values and 2^(D-1)
// more per axis give the value (interpolated exactly as pass 1 does: the same
// bins) and the D gradient components, each (v+ - v-) * sign / spacing.  A
// sample whose half-step points are not certainly inside or outside the
// buffer is deferred to the lane's region of cd_defer and taken by the tail
// kernel on the exact FP64 path (eval_sample, central_difference), in the
// lane's own order, so the results do not depend on timing.
template <int D, typename P>
__device__ __forceinline__ void pass2_rows_cd_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  extern __shared__ float lds_f32[];
  const int B = f.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  float* pr = lds_f32 + 128;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = f.pratio_f[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  const FastRows& R = f.rows;
  const typename P::pixel* img = static_cast<const typename P::pixel*>(f.moving_px);
  // this lane's region of deferred samples (at most one per row of each of
  // its walks)
  const uint32_t lane_cap = f.cd_wave_cap >> 6;
  uint32_t* defer = f.cd_defer + (int64_t)(lb * 256u + threadIdx.x) * lane_cap;
  uint32_t nd = 0;
  uint32_t it, iend, istep;
  rows_range(f, lb, lg, it, iend, istep);
  for (; it < iend; it += istep) {
    RowWalk<D> w;
    rows_start<D, P>(f, it, w);
    float sg[3] = {0.0f, 0.0f, 0.0f}, sy[3] = {0.0f, 0.0f, 0.0f};
    float yf = (float)w.y;
    for (int j = 0; j < w.rows; ++j, rows_step<D, P>(f, w), yf += 1.0f) {
      if (!w.active) continue;
      // Every load of the sample is issued before any test of loaded data
      // (pass 1's byte, the weight): the cells are clamped into the buffer,
      // so the loads are in bounds for every sample, and a sample that does
      // not count is dropped at the end.
      const uint32_t cb = f.cls[w.s];
      // the sample's cell and weights, as rows_classify
      int32_t I[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
      uint32_t off[3] = {0u, 0u, 0u}, stride[3] = {1u, (uint32_t)f.msy, (uint32_t)f.msz};
      float t[3] = {0.0f, 0.0f, 0.0f}, th[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const long long q = w.q[k];
        I[k] = (int32_t)(q >> 32);
        fl[k] = min(max(I[k], f.mstart[k]), f.mend[k]);
        const bool act = (uint32_t)(I[k] - f.mstart[k]) < (uint32_t)(f.mend[k] - f.mstart[k]);
        off[k] = act ? stride[k] : 0u;
        t[k] = (float)(uint32_t)q * 2.3283064365386963e-10f;
        th[k] = (float)((uint32_t)q ^ 0x80000000u) * 2.3283064365386963e-10f;  // fraction of the index + 1/2
      }
      uint32_t cbase = (uint32_t)(fl[0] - f.mstart[0]) + __umul24((uint32_t)f.msy, (uint32_t)(fl[1] - f.mstart[1]));
      if (D == 3) cbase += __umul24((uint32_t)f.msz, (uint32_t)(fl[2] - f.mstart[2]));
      // corner values c[i] at cbase + (i&1) off0 + (i>>1&1) off1 + (i>>2&1) off2
      float c[8];
#pragma unroll
      for (int i = 0; i < (1 << D); ++i)
        c[i] = P::ld32(img, cbase + ((i & 1) ? off[0] : 0u) + ((i & 2) ? off[1] : 0u) + ((i & 4) ? off[2] : 0u));
      // the third position along each axis b: P3 = clamp(I + 1) + act(I + 1)
      // when t >= 1/2, P0 = clamp(I - 1) else; its slab over the other axes
      // (e[b][0] at the cell's lower corners, then + off of the first other
      // axis, + off of the second, + both)
      bool upper[3] = {false, false, false};
      float e[3][4];
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const long long q = w.q[b];
        upper[b] = (uint32_t)q >= 0x80000000u;
        const int32_t Ie = upper[b] ? I[b] + 1 : I[b] - 1;
        const int32_t fe = min(max(Ie, f.mstart[b]), f.mend[b]);
        const int32_t pe = upper[b] ? fe + ((uint32_t)(Ie - f.mstart[b]) < (uint32_t)(f.mend[b] - f.mstart[b]) ? 1 : 0)
                                    : fe;
        const uint32_t ebase = (uint32_t)((int32_t)cbase + (pe - fl[b]) * (int32_t)stride[b]);
        if (D == 2) {
          const int o = 1 - b;
          e[b][0] = P::ld32(img, ebase);
          e[b][1] = P::ld32(img, ebase + off[o]);
          e[b][2] = e[b][3] = 0.0f;
        } else {
          const int o1 = b == 0 ? 1 : 0, o2 = b == 2 ? 1 : 2;  // the other axes
          e[b][0] = P::ld32(img, ebase);
          e[b][1] = P::ld32(img, ebase + off[o1]);
          e[b][2] = P::ld32(img, ebase + off[o2]);
          e[b][3] = P::ld32(img, ebase + off[o1] + off[o2]);
        }
      }
      // the value with fast_interp's operations (pass 1's bits)
      float v;
      {
        const float vx00 = fmaf(c[1] - c[0], t[0], c[0]);
        const float vx10 = fmaf(c[3] - c[2], t[0], c[2]);
        v = fmaf(vx10 - vx00, t[1], vx00);
        if (D == 3) {
          const float vx01 = fmaf(c[5] - c[4], t[0], c[4]);
          const float vx11 = fmaf(c[7] - c[6], t[0], c[6]);
          const float vxx1 = fmaf(vx11 - vx01, t[1], vx01);
          v = fmaf(vxx1 - v, t[2], v);
        }
      }
      const float mterm = fmaf(v, f.inv_mbs, -f.mnm);
      const int mbin = (int)fminf(fmaxf(mterm, 2.0f), (float)(B - 3));
      const float arg0 = (float)(mbin - 1) - mterm;
      // listed and not counted samples read row 0 of the table and are dropped
      const bool listed = cb >= kClsListed;
      const float wt = parzen_deriv_weight_fast(arg0, pr + (int)(listed ? 0u : cb) * B + (mbin - 1));
      const bool counted = !listed && wt != 0.0f;
      float g[3] = {0.0f, 0.0f, 0.0f};
      bool inb[3] = {false, false, false};
      bool any_unc = false;
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const long long q = w.q[b];
        inb[b] = (unsigned long long)(q - R.cd_lo_in[b]) < R.cd_span_in[b];
        const bool unc = ((unsigned long long)(q - R.cd_lo_out[b]) < R.cd_span_out[b]) & !inb[b];
        any_unc = any_unc | unc;
        // slabs over the other axes at the low, high and third positions
        float slo, shi, se;
        if (D == 2) {
          const int o = 1 - b;
          const int bit_b = 1 << b, bit_o = 1 << o;
          slo = fmaf(c[bit_o] - c[0], t[o], c[0]);
          shi = fmaf(c[bit_b | bit_o] - c[bit_b], t[o], c[bit_b]);
          se = fmaf(e[b][1] - e[b][0], t[o], e[b][0]);
        } else {
          const int o1 = b == 0 ? 1 : 0, o2 = b == 2 ? 1 : 2;  // the other axes
          const int b1 = 1 << o1, b2 = 1 << o2, bb = 1 << b;
          const float l0 = fmaf(c[b1] - c[0], t[o1], c[0]);
          const float l1 = fmaf(c[b1 | b2] - c[b2], t[o1], c[b2]);
          slo = fmaf(l1 - l0, t[o2], l0);
          const float h0 = fmaf(c[bb | b1] - c[bb], t[o1], c[bb]);
          const float h1 = fmaf(c[bb | b1 | b2] - c[bb | b2], t[o1], c[bb | b2]);
          shi = fmaf(h1 - h0, t[o2], h0);
          const float m0 = fmaf(e[b][1] - e[b][0], t[o1], e[b][0]);
          const float m1 = fmaf(e[b][3] - e[b][2], t[o1], e[b][2]);
          se = fmaf(m1 - m0, t[o2], m0);
        }
        const float vm = upper[b] ? fmaf(shi - slo, th[b], slo) : fmaf(slo - se, th[b], se);
        const float vp = upper[b] ? fmaf(se - shi, th[b], shi) : fmaf(shi - slo, th[b], slo);
        g[b] = (vp - vm) * R.cd_scale[b];
      }
      if (!counted) continue;
      if (any_unc) {
        // half-step points near the buffer's edge: after the walks
        if (nd < lane_cap) defer[nd] = w.s;
        ++nd;
        continue;
      }
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const float gw = inb[b] ? wt * g[b] : 0.0f;
        sg[b] += gw;
        sy[b] = fmaf(yf, gw, sy[b]);
      }
    }
    const double x = (double)w.x, z = (double)w.z;
#pragma unroll
    for (int b = 0; b < D; ++b) {
      acc[b * D + 0] = fma((double)sg[b], x, acc[b * D + 0]);
      acc[b * D + 1] += (double)sy[b];
      if (D == 3) acc[b * D + 2] = fma((double)sg[b], z, acc[b * D + 2]);
      acc[D * D + b] += (double)sg[b];
    }
  }
  block_reduce_array<PA>(acc, scratch, f.partials + (int64_t)lb * PA);
  // this workgroup's listed samples and each lane's deferred ones go to the
  // exact tail kernel (its FP64 central differences would raise this
  // kernel's registers)
  if (nd > lane_cap) atomicOr(f.exc_overflow, 1u);
  f.cd_nd[lb * 256u + threadIdx.x] = min(nd, lane_cap);
}

// The exact tail of pass2_rows_cd_body's workgroup lb: its listed samples and
// its lanes' deferred samples, into partial row lg + lb.
template <int D, typename P>
__device__ __forceinline__ void pass2_cd_tail_body(const FastArgs& f, const EvalArgs& a, uint32_t lb, uint32_t lg) {
  const uint32_t lane_cap = f.cd_wave_cap >> 6;
  const uint32_t t = lb * 256u + threadIdx.x;
  exceptions_pass2_body<D, P, 1, true>(a, f, lb, lg, f.cd_defer + (int64_t)t * lane_cap, f.cd_nd[t]);
}

// Fixed-order reduction of 2 * nblocks rows of index-space moments, then the
// map to the affine derivative:
//   d[b*D + d]   = nFactor * (sum_j V[d][j] M[b][j] + c0[d] M1[b])
//   d[D*D + b]   = nFactor * M1[b]
// with V the virtual index-to-physical matrix and c0 = virtual origin - centre.
template <int D>
__device__ __forceinline__ void fast_derivative_body(const double* __restrict__ partials, int rows, const FastMap& map,
                                                     const DevResult* __restrict__ result, double* __restrict__ deriv,
                                                     char* host_out = nullptr) {
  constexpr int PA = D * D + D;
  __shared__ double scratch[4 * PA];
  __shared__ double mom[PA];
  double acc[PA];
#pragma unroll
  for (int mu = 0; mu < PA; ++mu) acc[mu] = 0.0;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) {
#pragma unroll
    for (int mu = 0; mu < PA; ++mu) acc[mu] += partials[(int64_t)r * PA + mu];
  }
  block_reduce_array<PA>(acc, scratch, mom);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nf = result->n_factor;
    for (int b = 0; b < D; ++b) {
      const double m1 = mom[D * D + b];
      for (int d = 0; d < D; ++d) {
        double s = map.c0[d] * m1;
        for (int j = 0; j < D; ++j) s += map.V[d * D + j] * mom[b * D + j];
        deriv[b * D + d] = s * nf;
      }
      deriv[D * D + b] = m1 * nf;
    }
    if (host_out) {
      // the result block and the derivative straight into the caller's
      // pinned block (its layout: DevResult, 16 bytes, the derivative), so
      // no copy follows the kernel
      DevResult* hr = reinterpret_cast<DevResult*>(host_out);
      *hr = *result;
      double* hd = reinterpret_cast<double*>(host_out + sizeof(DevResult) + 16);
      for (int k = 0; k < PA; ++k) hd[k] = deriv[k];
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------
// Kernels: one evaluation per launch, or a batch of evaluations per launch
// (workgroup ranges per evaluation in `start`, multiples of 8 so that the
// XCD of a workgroup and of its local index agree).
// ---------------------------------------------------------------------------

__device__ __forceinline__ int batch_item(const int32_t* __restrict__ start, int n, uint32_t block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((uint32_t)start[mid] <= block) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int D, typename P, int MK>
__global__ void __launch_bounds__(256) pass1_fast_kernel(FastArgs f, EvalArgs a) {
  pass1_fast_body<D, P, MK>(f, a, blockIdx.x, gridDim.x);
}

// Batched kernels: each member runs the body it runs alone (row walks or
// the sample-strided loop; a workgroup-uniform branch), so its results are the
// bits of its own evaluation.
template <int D, typename P>
__global__ void __launch_bounds__(256) pass1_fast_batch(const FastArgs* __restrict__ fa,
                                                        const EvalArgs* __restrict__ ea,
                                                        const int32_t* __restrict__ start, int n,
                                                        const int32_t* __restrict__ halted) {
  const int r = batch_item(start, n, blockIdx.x);
  if (halted && halted[r]) return;
  if (fa[r].rows.on)
    fa[r].fbin8 ? pass1_rows_fb_body<D, P>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r])
                : pass1_rows_body<D, P>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r]);
  else
    pass1_fast_body<D, P, 1>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r]);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_cls_batch(const FastArgs* __restrict__ fa,
                                                       const EvalArgs* __restrict__ ea,
                                                       const int32_t* __restrict__ start, int n,
                                                       const int32_t* __restrict__ halted) {
  const int r = batch_item(start, n, blockIdx.x);
  if (halted && halted[r]) return;
  if (fa[r].rows.on)
    fa[r].fbin8 ? pass2_rows_fb_body<D, P>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r])
                : pass2_rows_body<D, P>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r]);
  else
    pass2_fast_body<D, P, 1, true, true>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r]);
}

// The exact fixed Parzen bin of every virtual voxel (the fixed side of
// eval_sample: fixed point, fixed mask, buffer test, FP64 interpolation, pixel
// assignment, ComputeSingleFixedImageParzenWindowIndex), or kClsSkip where
// the fixed side does not count the voxel.  Once per level: the fixed image,
// its range, the virtual domain and the identity fixed transform do not change
// between evaluations.
template <int D, typename P>
__global__ void __launch_bounds__(256) fixed_bins_kernel(EvalArgs a, uint8_t* __restrict__ out, uint32_t n) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    double vp[3];
    virtual_point<D>(a, s, vp);
    double fval;
    out[s] = fixed_value_exact<D, P>(a, vp, &fval) ? (uint8_t)fixed_bin_exact(a, fval) : (uint8_t)kClsSkip;
  }
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass1_rows_kernel(FastArgs f, EvalArgs a) {
  if (f.fbin8)
    pass1_rows_fb_body<D, P>(f, a, blockIdx.x, gridDim.x);
  else
    pass1_rows_body<D, P>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_rows_kernel(FastArgs f, EvalArgs a) {
  if (f.fbin8)
    pass2_rows_fb_body<D, P>(f, a, blockIdx.x, gridDim.x);
  else
    pass2_rows_body<D, P>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_rows_cd_kernel(FastArgs f, EvalArgs a) {
  pass2_rows_cd_body<D, P>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_cd_tail_kernel(FastArgs f, EvalArgs a) {
  pass2_cd_tail_body<D, P>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P, int MK>
__global__ void __launch_bounds__(256) exceptions_pass1_kernel(EvalArgs a, FastArgs f) {
  exceptions_pass1_body<D, P, MK>(a, f, blockIdx.x);
}

template <int D, typename P, int MK>
__global__ void __launch_bounds__(256) pass2_fast_kernel(FastArgs f, EvalArgs a) {
  pass2_fast_body<D, P, MK>(f, a, blockIdx.x, gridDim.x);
}

// Affine with image masks: pass 1, and the pass 2 that classifies again (the
// one after pass 1's byte per sample needs no mask test).
template <int D, typename P>
__global__ void __launch_bounds__(256) pass1_mask_kernel(FastArgs f, EvalArgs a) {
  pass1_fast_body<D, P, 1, true>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_mask_kernel(FastArgs f, EvalArgs a) {
  pass2_fast_body<D, P, 1, false, false, true>(f, a, blockIdx.x, gridDim.x);
}

// Pass 2 after a pass 1 that wrote its byte per sample, with the moving values
// interpolated from the texels it gathers for the gradient (one load per
// corner; the byte per sample keeps the registers under 128 with it).
template <int D, typename P, int MK>
__global__ void __launch_bounds__(256) pass2_cls_kernel(FastArgs f, EvalArgs a) {
  pass2_fast_body<D, P, MK, true, true>(f, a, blockIdx.x, gridDim.x);
}

template <int D, typename P>
__global__ void __launch_bounds__(256) pass2_fast_batch(const FastArgs* __restrict__ fa,
                                                        const EvalArgs* __restrict__ ea,
                                                        const int32_t* __restrict__ start, int n,
                                                        const int32_t* __restrict__ halted) {
  const int r = batch_item(start, n, blockIdx.x);
  if (halted && halted[r]) return;
  pass2_fast_body<D, P, 1>(fa[r], ea[r], blockIdx.x - start[r], start[r + 1] - start[r]);
}

template <int D>
__global__ void __launch_bounds__(256) fast_derivative_kernel(const double* __restrict__ partials, int rows,
                                                              FastMap map, const DevResult* __restrict__ result,
                                                              double* __restrict__ deriv, char* host_out) {
  fast_derivative_body<D>(partials, rows, map, result, deriv, host_out);
}

template <int D>
__global__ void __launch_bounds__(256) fast_derivative_batch(const BatchDeriv* __restrict__ items,
                                                             const int32_t* __restrict__ halted) {
  if (halted && halted[blockIdx.x]) return;
  const BatchDeriv& it = items[blockIdx.x];
  fast_derivative_body<D>(it.partials, it.rows, it.map, it.result, it.deriv);
}

// Zeroes one evaluation's histogram, marginal, count and overflow flag per
// workgroup (a kernel rather than memsets, so the sequence can be captured).
__global__ void __launch_bounds__(256) batch_clear(const FastArgs* __restrict__ fa,
                                                   const int32_t* __restrict__ halted) {
  if (halted && halted[blockIdx.x]) return;
  const FastArgs& f = fa[blockIdx.x];
  const int BB = f.bins * f.bins;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) f.hist[i] = 0ull;
  if (threadIdx.x == 0) {
    *f.nvalid = 0ull;
    *f.exc_overflow = 0u;
  }
}

// Value, count, status (+16 on exception-list overflow) and derivative of
// every evaluation of the batch into one buffer, for one copy to the host.
__global__ void batch_gather(const BatchOut* __restrict__ items, double* __restrict__ out, int stride) {
  const BatchOut& it = items[blockIdx.x];
  double* o = out + (int64_t)blockIdx.x * stride;
  if (threadIdx.x == 0) {
    o[0] = it.result->value;
    o[1] = (double)it.result->n_valid;
    o[2] = (double)(it.result->status + (it.result->exc_overflow ? 16 : 0));
  }
  if (it.deriv)
    for (int i = threadIdx.x; i < it.P; i += blockDim.x) o[3 + i] = it.deriv[i];
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------

#define MMI_DISPATCH_FAST(KERNEL, storage, dim, grid, lds, st, ...)                                  \
  do {                                                                                              \
    if ((storage) == kStorageBF16) {                                                                \
      if ((dim) == 2)                                                                               \
        hipLaunchKernelGGL((KERNEL<2, PxBF16>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);       \
      else                                                                                          \
        hipLaunchKernelGGL((KERNEL<3, PxBF16>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);       \
    } else if ((storage) == kStorageF16) {                                                          \
      if ((dim) == 2)                                                                               \
        hipLaunchKernelGGL((KERNEL<2, PxF16>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);        \
      else                                                                                          \
        hipLaunchKernelGGL((KERNEL<3, PxF16>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);        \
    } else {                                                                                        \
      if ((dim) == 2)                                                                               \
        hipLaunchKernelGGL((KERNEL<2, PxF32>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);        \
      else                                                                                          \
        hipLaunchKernelGGL((KERNEL<3, PxF32>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);        \
    }                                                                                               \
  } while (0)

// Single-evaluation kernels also dispatch on the moving transform kind.
#define MMI_DISPATCH_FAST_MK(KERNEL, mk, storage, dim, grid, lds, st, ...)                          \
  do {                                                                                              \
    if ((mk) == 2) {                                                                                \
      if ((storage) == kStorageBF16) {                                                              \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxBF16, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);  \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxBF16, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);  \
      } else if ((storage) == kStorageF16) {                                                        \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxF16, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxF16, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
      } else {                                                                                      \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxF32, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxF32, 2>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
      }                                                                                             \
    } else {                                                                                        \
      if ((storage) == kStorageBF16) {                                                              \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxBF16, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);  \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxBF16, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);  \
      } else if ((storage) == kStorageF16) {                                                        \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxF16, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxF16, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
      } else {                                                                                      \
        if ((dim) == 2)                                                                             \
          hipLaunchKernelGGL((KERNEL<2, PxF32, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
        else                                                                                        \
          hipLaunchKernelGGL((KERNEL<3, PxF32, 1>), dim3(grid), dim3(256), lds, st, __VA_ARGS__);   \
      }                                                                                             \
    }                                                                                               \
  } while (0)

hipError_t launch_fast_pass1(const FastArgs& f, const EvalArgs& a, int dim, int grid, hipStream_t st) {
  const size_t lds = pass1_lds_bytes_with(f.bins, f.hist_copies);
  if (f.fmask.px || f.mmask.px) {
    if (f.mk != 1) return hipErrorInvalidValue;
    MMI_DISPATCH_FAST(pass1_mask_kernel, f.storage, dim, grid, lds, st, f, a);
  } else if (f.rows.on) {
    MMI_DISPATCH_FAST(pass1_rows_kernel, f.storage, dim, grid, lds, st, f, a);
  } else {
    MMI_DISPATCH_FAST_MK(pass1_fast_kernel, f.mk, f.storage, dim, grid, lds, st, f, a);
  }
  return hipGetLastError();
}

hipError_t launch_fixed_bins(const EvalArgs& a, uint8_t* out, uint32_t n, int grid, hipStream_t st) {
  MMI_DISPATCH_FAST(fixed_bins_kernel, a.storage, a.virt.dim, grid, 0, st, a, out, n);
  return hipGetLastError();
}

hipError_t launch_fast_exceptions1(const EvalArgs& a, const FastArgs& f, int grid, hipStream_t st) {
  // grid (the pass-1 workgroups) is a multiple of 8 = kExcListsPerGroup
  MMI_DISPATCH_FAST_MK(exceptions_pass1_kernel, f.mk, f.storage, a.virt.dim, grid / kExcListsPerGroup, 0, st, a, f);
  return hipGetLastError();
}

hipError_t launch_fast_pass2(const FastArgs& f, const EvalArgs& a, int dim, int grid, hipStream_t st) {
  const size_t lds = ((size_t)f.bins * f.bins + 128) * 4;
  // a row-walk pass 1 is followed by a pass 2 that reads its byte per sample
  if (f.rows.on && !f.cls && !f.fbin8) return hipErrorInvalidValue;
  if (f.rows.on && f.rows.cd_on) {
    if (!f.cls) return hipErrorInvalidValue;
    MMI_DISPATCH_FAST(pass2_rows_cd_kernel, f.storage, dim, grid, lds, st, f, a);
    MMI_DISPATCH_FAST(pass2_cd_tail_kernel, f.storage, dim, grid, 0, st, f, a);
  } else if (f.rows.on)
    MMI_DISPATCH_FAST(pass2_rows_kernel, f.storage, dim, grid, lds, st, f, a);
  else if (f.cls)
    MMI_DISPATCH_FAST_MK(pass2_cls_kernel, f.mk, f.storage, dim, grid, lds, st, f, a);
  else if (f.fmask.px || f.mmask.px)
    MMI_DISPATCH_FAST(pass2_mask_kernel, f.storage, dim, grid, lds, st, f, a);
  else
    MMI_DISPATCH_FAST_MK(pass2_fast_kernel, f.mk, f.storage, dim, grid, lds, st, f, a);
  return hipGetLastError();
}

hipError_t launch_fast_derivative(const double* partials, int rows, const FastMap& map, int dim,
                                  const DevResult* r, double* deriv, char* host_out, hipStream_t st) {
  if (dim == 2)
    hipLaunchKernelGGL(fast_derivative_kernel<2>, dim3(1), dim3(256), 0, st, partials, rows, map, r, deriv, host_out);
  else
    hipLaunchKernelGGL(fast_derivative_kernel<3>, dim3(1), dim3(256), 0, st, partials, rows, map, r, deriv, host_out);
  return hipGetLastError();
}

hipError_t launch_fast_batch(const BatchLaunch& b, hipStream_t st) {
  const int B = b.bins_max;
  hipLaunchKernelGGL(batch_clear, dim3(b.n), dim3(256), 0, st, b.fa, b.halted);
  const size_t lds1 = (size_t)b.lds1;
  MMI_DISPATCH_FAST(pass1_fast_batch, b.storage, b.dim, b.total_blocks, lds1, st, b.fa, b.ea, b.start, b.n,
                    b.halted);
  hipError_t e = launch_finalize_batch(b.fz, b.n, B, b.halted, st);
  if (e != hipSuccess) return e;
  if (b.want_derivative) {
    const size_t lds2 = ((size_t)B * B + 128) * 4;
    if (b.cls)
      MMI_DISPATCH_FAST(pass2_cls_batch, b.storage, b.dim, b.total_blocks, lds2, st, b.fa, b.ea, b.start, b.n,
                        b.halted);
    else
      MMI_DISPATCH_FAST(pass2_fast_batch, b.storage, b.dim, b.total_blocks, lds2, st, b.fa, b.ea, b.start, b.n,
                        b.halted);
    if (b.dim == 2)
      hipLaunchKernelGGL(fast_derivative_batch<2>, dim3(b.n), dim3(256), 0, st, b.dv, b.halted);
    else
      hipLaunchKernelGGL(fast_derivative_batch<3>, dim3(b.n), dim3(256), 0, st, b.dv, b.halted);
  }
  if (b.out) hipLaunchKernelGGL(batch_gather, dim3(b.n), dim3(64), 0, st, b.items, b.out, b.out_stride);
  return hipGetLastError();
}

}  // namespace mmi
