// Mattes mutual information v4, dense value + derivative, on gfx950.
//
// One evaluation = three kernels on the context stream:
//   pass 1   every virtual sample -> fixed/moving Parzen bins; 4 cubic B-spline
//            Parzen weights into a per-workgroup joint histogram in LDS held as
//            64-bit fixed point (weight * 2^32, the four weights of a sample
//            summing to exactly 2^32), fixed-marginal and valid counts; LDS
//            partials merged into HBM with 64-bit integer atomics.  Integer
//            accumulation makes the histogram independent of scheduling order.
//   finalize one workgroup: FP64 normalisation, marginals, the entropy sum
//            (MattesMutualInformationImageToImageMetricv4::ComputeResults,
//            itkMattesMutualInformationImageToImageMetricv4.hxx:353-499) and
//            the table pRatio[f][m] = log(p / pM) (0 where the reference skips
//            the bin).
//   pass 2   the derivative without the reference's P x B x B dPDF buffer:
//            d[mu] = nFactor * sum_s (J_s^T gradM_s)[mu] *
//                    sum_k B3'(arg_{s,k}) * pRatio[f_s][m_s - 1 + k]
//            which is the reference's sum_{f,m} dPDF[f,m,mu] * pRatio[f,m]
//            regrouped by sample (MattesT:355-373 + Mattes:478-495).
//
// Two sample evaluators share these passes: the exact path (FP64 in the
// reference's operation order, any transform) and, for affine moving
// transforms, a certified fast path (FP64 continuous index, FP32 interpolation
// and Parzen terms) that falls back to the exact path for every sample whose
// inside test, intensity-range test or bin index it cannot prove equal.
//
// Compiled with -ffp-contract=off: the FP64 coordinate path must round like
// the reference's x86-64 build so that bins match bit for bit.
#include "mattes_kernels.h"

#include <hip/hip_runtime.h>

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

// Evaluation modes of the pass kernels.
constexpr int kExactAffine = 0;   // FP64 reference path, affine moving transform
constexpr int kFastAffine = 1;    // certified fast path, affine moving transform
constexpr int kExactBSpline = 2;  // FP64 reference path, cubic B-spline moving transform

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

// ProcessVirtualPoint + Mattes ProcessPoint up to the bins
// (ThrBase:174-273, I2IM:289-373, MattesT:215-275), exact FP64.  MK is the
// moving transform kind (1 affine, 2 B-spline).
template <int D, int MK>
__device__ __forceinline__ bool eval_sample(const EvalArgs& a, const int32_t* idx, SampleOut<D>& o) {
  index_to_point<D>(a.virt, idx, o.vp);
  double fp[3], ci[3];
  fixed_point<D>(a.ftf, o.vp, fp);
  point_to_cindex<D>(a.fixed, fp, ci);
  if (!inside_buffer<D>(a.fixed, ci)) return false;
  const Cell fc = make_cell<D>(a.fixed, ci);
  const double fval = (double)(float)interp_exact<D>(a.fixed_px, fc);

  transform_point<D, MK>(a.mtf, o.vp, o.mp, &o.sup);
  point_to_cindex<D>(a.moving, o.mp, ci);
  if (!inside_buffer<D>(a.moving, ci)) return false;
  o.cell = make_cell<D>(a.moving, ci);
  const double mval = (double)(float)interp_exact<D>(a.moving_px, o.cell);

  if (mval < a.moving_min) return false;
  if (mval > a.moving_max) return false;
  const double mterm = mval / a.moving_bin_size - a.moving_norm_min;
  const int mb = mterm < 2.0 ? 2 : (mterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)mterm);
  const double fterm = fval / a.fixed_bin_size - a.fixed_norm_min;
  const int fb = fterm < 2.0 ? 2 : (fterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)fterm);
  o.fbin = fb;
  o.mbin = mb;
  o.mterm = mterm;
  return true;
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

// ---------------------------------------------------------------------------
// Certified fast path (affine moving transform)
// ---------------------------------------------------------------------------

template <int D>
struct FastSample {
  int fbin;
  int mbin;
  float arg0;  // (mbin - 1) - mterm
  Cell cell;   // moving interpolation cell
  int32_t idx[3];
};

// FP32 multilinear value with the reference's duplicated-corner convention.
template <int D>
__device__ __forceinline__ float interp_fast(const float* __restrict__ img, const Cell& c, const float* t) {
  const float v000 = img[c.base], v100 = img[c.base + c.ox];
  const float v010 = img[c.base + c.oy], v110 = img[c.base + c.oy + c.ox];
  const float vx00 = fmaf(v100 - v000, t[0], v000);
  const float vx10 = fmaf(v110 - v010, t[0], v010);
  const float vxx0 = fmaf(vx10 - vx00, t[1], vx00);
  if (D == 2) return vxx0;
  const float v001 = img[c.base + c.oz], v101 = img[c.base + c.oz + c.ox];
  const float v011 = img[c.base + c.oz + c.oy], v111 = img[c.base + c.oz + c.oy + c.ox];
  const float vx01 = fmaf(v101 - v001, t[0], v001);
  const float vx11 = fmaf(v111 - v011, t[0], v011);
  const float vxx1 = fmaf(vx11 - vx01, t[1], vx01);
  return fmaf(vxx1 - vxx0, t[2], vxx0);
}

// Bin of an FP32 Parzen term; *uncertain when the reference's FP64 term could
// fall on the other side of an unclamped bin edge.
__device__ __forceinline__ int certified_bin(float term, float bound, int B, bool* uncertain) {
  if (term < 2.0f) {
    *uncertain = *uncertain || (term >= 3.0f - bound);
    return 2;
  }
  if (term >= (float)(B - 2)) {
    *uncertain = *uncertain || (term < (float)(B - 3) + bound);
    return B - 3;
  }
  const float n = rintf(term);
  if (n >= 3.0f && n <= (float)(B - 3) && fabsf(term - n) <= bound) *uncertain = true;
  return (int)term;
}

template <int D>
__device__ __forceinline__ bool fast_sample(const EvalArgs& a, uint32_t s, FastSample<D>& o) {
  const FastAffine& F = a.fast;
  sample_index<D>(a.virt, s, o.idx);
  bool uncertain = false;

  // fixed value
  float fv = 0.0f;
  if (F.fixed_direct) {
    const int32_t fx = o.idx[0] + F.fixed_offset[0] - a.fixed.start[0];
    const int32_t fy = o.idx[1] + F.fixed_offset[1] - a.fixed.start[1];
    const int32_t fz = D == 3 ? o.idx[2] + F.fixed_offset[2] - a.fixed.start[2] : 0;
    fv = a.fixed_px[fx + a.fixed.sy * fy + a.fixed.sz * fz];
  } else {
    uncertain = true;  // general fixed geometry: exact path
  }

  // moving continuous index: one FP64 affine map of the virtual index
  double ci[3];
  const double x = (double)o.idx[0], y = (double)o.idx[1], z = (double)o.idx[2];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double c = F.b[k];
    if (D == 3) c = fma(F.A[k * 3 + 2], z, c);
    c = fma(F.A[k * D + 1], y, c);
    ci[k] = fma(F.A[k * D + 0], x, c);
  }
  bool out = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    out = out || (ci[k] < a.moving.cilo[k] - F.e_ci) || (ci[k] >= a.moving.cihi[k] + F.e_ci);
    uncertain = uncertain || (fabs(ci[k] - a.moving.cilo[k]) <= F.e_ci) || (fabs(ci[k] - a.moving.cihi[k]) <= F.e_ci);
  }
  if (out) return false;  // outside the moving buffer: not counted whatever the fixed side says

  if (!uncertain) {
    o.cell = make_cell<D>(a.moving, ci);
    const float t[3] = {(float)o.cell.d[0], (float)o.cell.d[1], (float)o.cell.d[2]};
    const float v = interp_fast<D>(a.moving_px, o.cell, t);
    if (v < F.mmin - F.ev || v > F.mmax + F.ev) return false;
    uncertain = (v < F.mmin + F.ev) || (v > F.mmax - F.ev);
    const float mterm = fmaf(v, F.inv_mbs, -F.mnm);
    o.mbin = certified_bin(mterm, F.bt_m, a.bins, &uncertain);
    o.arg0 = (float)(o.mbin - 1) - mterm;
    bool fu = false;
    const float fterm = fmaf(fv, F.inv_fbs, -F.fnm);
    o.fbin = certified_bin(fterm, F.bt_f, a.bins, &fu);
    if (fu) {
      // exact FP64 fixed term (ComputeSingleFixedImageParzenWindowIndex)
      const double ft = (double)fv / a.fixed_bin_size - a.fixed_norm_min;
      o.fbin = ft < 2.0 ? 2 : (ft >= (double)(a.bins - 2) ? a.bins - 3 : (int)ft);
    }
    if (!uncertain) return true;
  }
  // exact FP64 path of the reference for this sample
  SampleOut<D> e;
  if (!eval_sample<D, 1>(a, o.idx, e)) return false;
  o.fbin = e.fbin;
  o.mbin = e.mbin;
  o.arg0 = (float)((double)(e.mbin - 1) - e.mterm);
  o.cell = e.cell;
  return true;
}

// ---------------------------------------------------------------------------
// Fast-path kernels: 4 consecutive samples per thread, XCD-local ranges
// ---------------------------------------------------------------------------
//
// Each thread evaluates a run of 4 consecutive virtual samples (x fastest).
// The continuous index is advanced by one FP64 column add per step (re-derived
// from the index at every row wrap), consecutive samples with the same
// (fixed bin, moving bin) are merged in registers before the LDS atomics, and
// the valid count stays in a register until the end of the kernel.  Workgroup
// b runs on XCD b % 8 (round-robin dispatch); the sample range is cut into 8
// contiguous chunks so each XCD's L2 sees one slab of the moving texels.

constexpr int kRun = 4;
constexpr int kXcd = 8;

// Fast evaluation of one sample given its index and FP64 continuous index.
template <int D>
__device__ __forceinline__ bool fast_eval(const EvalArgs& a, const int32_t* idx, const double* ci, int& fbin,
                                          int& mbin, float& arg0, Cell& cell) {
  const FastAffine& F = a.fast;
  bool uncertain = !F.fixed_direct;
  bool out = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    out = out || (ci[k] < a.moving.cilo[k] - F.e_ci) || (ci[k] >= a.moving.cihi[k] + F.e_ci);
    uncertain = uncertain || (fabs(ci[k] - a.moving.cilo[k]) <= F.e_ci) || (fabs(ci[k] - a.moving.cihi[k]) <= F.e_ci);
  }
  if (out) return false;
  if (!uncertain) {
    const int32_t fx = idx[0] + F.fixed_offset[0] - a.fixed.start[0];
    const int32_t fy = idx[1] + F.fixed_offset[1] - a.fixed.start[1];
    const int32_t fz = D == 3 ? idx[2] + F.fixed_offset[2] - a.fixed.start[2] : 0;
    const float fv = a.fixed_px[fx + a.fixed.sy * fy + a.fixed.sz * fz];
    cell = make_cell<D>(a.moving, ci);
    const float t[3] = {(float)cell.d[0], (float)cell.d[1], (float)cell.d[2]};
    const float v = interp_fast<D>(a.moving_px, cell, t);
    if (v < F.mmin - F.ev || v > F.mmax + F.ev) return false;
    uncertain = (v < F.mmin + F.ev) || (v > F.mmax - F.ev);
    const float mterm = fmaf(v, F.inv_mbs, -F.mnm);
    mbin = certified_bin(mterm, F.bt_m, a.bins, &uncertain);
    arg0 = (float)(mbin - 1) - mterm;
    bool fu = false;
    const float fterm = fmaf(fv, F.inv_fbs, -F.fnm);
    fbin = certified_bin(fterm, F.bt_f, a.bins, &fu);
    if (fu) {
      const double ft = (double)fv / a.fixed_bin_size - a.fixed_norm_min;
      fbin = ft < 2.0 ? 2 : (ft >= (double)(a.bins - 2) ? a.bins - 3 : (int)ft);
    }
    if (!uncertain) return true;
  }
  SampleOut<D> e;
  if (!eval_sample<D, 1>(a, idx, e)) return false;
  fbin = e.fbin;
  mbin = e.mbin;
  arg0 = (float)((double)(e.mbin - 1) - e.mterm);
  cell = e.cell;
  return true;
}

// Iterates the runs of this thread: calls f(idx, ci) for every sample.
template <int D, typename F>
__device__ __forceinline__ void for_each_run_sample(const EvalArgs& a, F&& f) {
  const FastAffine& FA = a.fast;
  const uint32_t n = (uint32_t)a.n;
  const uint32_t nruns = (n + kRun - 1) / kRun;
  const uint32_t xcd = blockIdx.x % kXcd;
  const uint32_t lb = blockIdx.x / kXcd;
  const uint32_t bpx = gridDim.x / kXcd;
  const uint32_t chunk = (nruns + kXcd - 1) / kXcd;
  const uint32_t r0 = xcd * chunk;
  const uint32_t r1 = min(nruns, r0 + chunk);
  for (uint32_t r = r0 + lb * blockDim.x + threadIdx.x; r < r1; r += bpx * blockDim.x) {
    uint32_t s = r * kRun;
    int32_t idx[3];
    sample_index<D>(a.virt, s, idx);
    double ci[3];
    const double x = (double)idx[0], y = (double)idx[1], z = (double)idx[2];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double c = FA.b[k];
      if (D == 3) c = fma(FA.A[k * 3 + 2], z, c);
      c = fma(FA.A[k * D + 1], y, c);
      ci[k] = fma(FA.A[k * D + 0], x, c);
    }
#pragma unroll
    for (int j = 0; j < kRun; ++j) {
      if (s >= n) break;
      f(s, idx, ci);
      ++s;
      if (++idx[0] > a.virt.end[0]) {
        sample_index<D>(a.virt, s, idx);
        const double x2 = (double)idx[0], y2 = (double)idx[1], z2 = (double)idx[2];
#pragma unroll
        for (int k = 0; k < D; ++k) {
          double c = FA.b[k];
          if (D == 3) c = fma(FA.A[k * 3 + 2], z2, c);
          c = fma(FA.A[k * D + 1], y2, c);
          ci[k] = fma(FA.A[k * D + 0], x2, c);
        }
      } else {
#pragma unroll
        for (int k = 0; k < D; ++k) ci[k] += FA.A[k * D + 0];
      }
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) pass1_fast_kernel(EvalArgs a) {
  extern __shared__ unsigned long long lds_u64[];
  const int B = a.bins;
  const int BB = B * B;
  unsigned long long* hist = lds_u64;
  unsigned int* fmarg = reinterpret_cast<unsigned int*>(lds_u64 + BB);
  unsigned int* nvalid_lds = fmarg + B;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) hist[i] = 0ull;
  for (int i = threadIdx.x; i < B + 1; i += blockDim.x) fmarg[i] = 0u;
  __syncthreads();
  unsigned int nvalid = 0;
  int cur = -1;  // key of the pending run: fbin * B + mbin
  unsigned long long acc[4] = {0, 0, 0, 0};
  unsigned int cnt = 0;
  auto flush = [&]() {
    if (cur < 0) return;
    const int fb = cur / B, mb = cur - fb * B;
    atomicAdd(&fmarg[fb], cnt);
    unsigned long long* row = hist + fb * B + (mb - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (acc[k]) atomicAdd(&row[k], acc[k]);
  };
  for_each_run_sample<D>(a, [&](uint32_t s, const int32_t* idx, const double* ci) {
    int fbin = 0, mbin = 0;
    float arg0 = 0.0f;
    Cell cell;
    const bool ok = fast_eval<D>(a, idx, ci, fbin, mbin, arg0, cell);
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)mbin : (int16_t)-1;
    }
    if (!ok) return;
    ++nvalid;
    unsigned int q[4];
    parzen_fixed(arg0, q);
    const int key = fbin * B + mbin;
    if (key != cur) {
      flush();
      cur = key;
      cnt = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] = 0ull;
    }
    ++cnt;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += q[k];
  });
  flush();
  // valid count: wavefront sum, one LDS atomic per wavefront
  unsigned int v = nvalid;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(nvalid_lds, v);
  __syncthreads();
  for (int i = threadIdx.x; i < BB; i += blockDim.x) {
    const unsigned long long h = hist[i];
    if (h) atomicAdd(&a.hist[i], h);
  }
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const unsigned int m = fmarg[i];
    if (m) atomicAdd(&a.fmarg[i], (unsigned long long)m);
  }
  if (threadIdx.x == 0 && *nvalid_lds) atomicAdd(a.nvalid, (unsigned long long)*nvalid_lds);
}

template <int D>
__global__ void __launch_bounds__(256) pass2_fast_kernel(EvalArgs a) {
  extern __shared__ float lds_f32[];
  const int B = a.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles
  float* pr = lds_f32 + 128;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = a.pratio_f[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  for_each_run_sample<D>(a, [&](uint32_t, const int32_t* idx, const double* ci) {
    int fbin = 0, mbin = 0;
    float arg0 = 0.0f;
    Cell cell;
    if (!fast_eval<D>(a, idx, ci, fbin, mbin, arg0, cell)) return;
    const float w = parzen_deriv_weight(arg0, pr + fbin * B + (mbin - 1));
    if (w == 0.0f) return;
    const float3 g3 = interp_gradient3<D>(a.texels, cell);
    const double gd[3] = {(double)g3.x, (double)g3.y, (double)g3.z};
    double vp[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double p = a.virt.origin[k];
      p += a.virt.i2p[k * D + 0] * (double)idx[0];
      p += a.virt.i2p[k * D + 1] * (double)idx[1];
      if (D == 3) p += a.virt.i2p[k * D + 2] * (double)idx[2];
      vp[k] = p - a.mtf.center[k];
    }
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const double wg = (double)w * gd[b];
#pragma unroll
      for (int d = 0; d < D; ++d) acc[b * D + d] = fma(wg, vp[d], acc[b * D + d]);
      acc[D * D + b] += wg;
    }
  });
  block_reduce_array<PA>(acc, scratch, a.partials + (int64_t)blockIdx.x * PA);
}

// ---------------------------------------------------------------------------
// pass 1: joint histogram
// ---------------------------------------------------------------------------

template <int D, int MODE>
__global__ void __launch_bounds__(256) pass1_kernel(EvalArgs a) {
  constexpr bool FAST = MODE == kFastAffine;
  constexpr int MK = MODE == kExactBSpline ? 2 : 1;
  extern __shared__ unsigned long long lds_u64[];
  const int B = a.bins;
  const int BB = B * B;
  unsigned long long* hist = a.hist_in_lds ? lds_u64 : a.hist;
  unsigned int* fmarg = reinterpret_cast<unsigned int*>(lds_u64 + (a.hist_in_lds ? BB : 0));
  unsigned int* nvalid = fmarg + B;
  if (a.hist_in_lds) {
    for (int i = threadIdx.x; i < BB; i += blockDim.x) hist[i] = 0ull;
  }
  for (int i = threadIdx.x; i < B + 1; i += blockDim.x) fmarg[i] = 0u;
  __syncthreads();

  const uint32_t n = (uint32_t)a.n;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
    int fbin = 0, mbin = 0;
    float arg0 = 0.0f;
    bool ok;
    if constexpr (FAST) {
      FastSample<D> o;
      ok = fast_sample<D>(a, s, o);
      fbin = o.fbin;
      mbin = o.mbin;
      arg0 = o.arg0;
    } else {
      int32_t idx[3];
      sample_index<D>(a.virt, s, idx);
      SampleOut<D> o;
      ok = eval_sample<D, MK>(a, idx, o);
      fbin = o.fbin;
      mbin = o.mbin;
      arg0 = (float)((double)(o.mbin - 1) - o.mterm);
    }
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)mbin : (int16_t)-1;
    }
    const unsigned long long vm = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && vm) atomicAdd(nvalid, (unsigned int)__popcll(vm));
    if (!ok) continue;
    atomicAdd(&fmarg[fbin], 1u);
    unsigned int q[4];
    parzen_fixed(arg0, q);
    unsigned long long* row = hist + fbin * B + (mbin - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (q[k]) atomicAdd(&row[k], (unsigned long long)q[k]);
  }
  __syncthreads();
  if (a.hist_in_lds) {
    for (int i = threadIdx.x; i < BB; i += blockDim.x) {
      const unsigned long long v = hist[i];
      if (v) atomicAdd(&a.hist[i], v);
    }
  }
  for (int i = threadIdx.x; i < B; i += blockDim.x) {
    const unsigned int v = fmarg[i];
    if (v) atomicAdd(&a.fmarg[i], (unsigned long long)v);
  }
  if (threadIdx.x == 0 && *nvalid) atomicAdd(a.nvalid, (unsigned long long)*nvalid);
}

// ---------------------------------------------------------------------------
// finalize: ComputeResults (Mattes:353-499) in FP64, one workgroup
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f) {
  __shared__ double s_norm;
  __shared__ unsigned long long redu[256];
  __shared__ double reds[256];
  const int B = f.bins;
  const int BB = B * B;
  const int t = threadIdx.x;
  // Joint PDF sum: exact in 64-bit integers.
  unsigned long long part = 0;
  for (int i = t; i < BB; i += blockDim.x) part += f.hist[i];
  redu[t] = part;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) redu[t] += redu[t + w];
    __syncthreads();
  }
  const unsigned long long total = redu[0];
  const unsigned long long nvalid = *f.nvalid;
  const double jointPDFSum = (double)total / kFixedPointScale;
  if (t == 0) {
    f.result->n_valid = nvalid;
    f.result->joint_pdf_sum = jointPDFSum;
    f.result->status = (nvalid == 0) ? 1 : (jointPDFSum < 2.220446049250313e-16 ? 2 : 0);
    s_norm = 1.0 / jointPDFSum;
  }
  __syncthreads();
  if (nvalid == 0 || jointPDFSum < 2.220446049250313e-16) return;
  const double nf = s_norm;
  const double closeToZero = 2.220446049250313e-16;
  for (int i = t; i < BB; i += blockDim.x) f.pdf[i] = ((double)f.hist[i] / kFixedPointScale) * nf;
  for (int i = t; i < B; i += blockDim.x) {
    const double fmv = (double)f.fmarg[i] * nf;
    f.fm[i] = fmv;
    f.rowsum[i] = fmv > closeToZero ? log(fmv) : 0.0;  // log of the fixed marginal, once per row
  }
  __syncthreads();
  // moving marginal: column sums in the reference's row order (Mattes:405-417)
  for (int m = t; m < B; m += blockDim.x) {
    double acc = 0.0;
    for (int r = 0; r < B; ++r) acc += f.pdf[r * B + m];
    f.mm[m] = acc;
  }
  __syncthreads();
  // entropy terms and pRatio over all (f, m) pairs in parallel (Mattes:419-476)
  double part_sum = 0.0;
  for (int i = t; i < BB; i += blockDim.x) {
    const int r = i / B, m = i - (i / B) * B;
    const double fmv = f.fm[r];
    const double mmv = f.mm[m];
    const double jp = f.pdf[i];
    double pr = 0.0;
    if (fmv > closeToZero && mmv > closeToZero && jp > closeToZero) {
      pr = log(jp / mmv);
      part_sum += jp * (pr - f.rowsum[r]);
    }
    f.pratio[i] = pr;
    f.pratio_f[i] = (float)pr;
  }
  reds[t] = part_sum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) reds[t] += reds[t + w];
    __syncthreads();
  }
  if (t == 0) {
    f.result->value = -1.0 * reds[0];
    f.result->n_factor = -1.0 / (f.moving_bin_size * (double)nvalid);
  }
}

// ---------------------------------------------------------------------------
// pass 2: derivative
// ---------------------------------------------------------------------------

// Central differences of the moving image at a mapped point
// (CentralDifferenceImageFunction::EvaluateSpecialized,
// Modules/Core/ImageFunction/include/itkCentralDifferenceImageFunction.hxx:252-318).
template <int D>
__device__ __forceinline__ void central_difference(const EvalArgs& a, const double* p, double* grad) {
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
        const double v2 = interp_exact<D>(a.moving_px, make_cell<D>(a.moving, c2));
        const double v1 = interp_exact<D>(a.moving_px, make_cell<D>(a.moving, c1));
        grad[dim] = (v2 - v1) / delta;
      } else {
        grad[dim] = 0.0;
      }
    }
    n1[dim] = p[dim];
    n2[dim] = p[dim];
  }
}

template <int D, int MODE>
__global__ void __launch_bounds__(256) pass2_kernel(EvalArgs a) {
  constexpr bool FAST = MODE == kFastAffine;
  constexpr int MK = MODE == kExactBSpline ? 2 : 1;
  extern __shared__ float lds_f32[];
  const int B = a.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  // pRatio table: always in LDS on the fast path (the host only selects it
  // when B*B floats fit), else in LDS when it fits and through the caches
  // otherwise.
  float* lds_pr = lds_f32 + 128;
  const float* pr = lds_pr;
  if (FAST || a.pratio_in_lds) {
    for (int i = threadIdx.x; i < BB; i += blockDim.x) lds_pr[i] = a.pratio_f[i];
  } else {
    pr = a.pratio_f;
  }
  __syncthreads();

  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;

  const uint32_t n = (uint32_t)a.n;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
    float g[3] = {0.0f, 0.0f, 0.0f};
    double gd[3] = {0.0, 0.0, 0.0};
    double vp[3];
    float w;
    if constexpr (FAST) {
      FastSample<D> o;
      if (!fast_sample<D>(a, s, o)) continue;
      w = parzen_deriv_weight(o.arg0, pr + o.fbin * B + (o.mbin - 1));
      if (w == 0.0f) continue;
      const float3 g3 = interp_gradient3<D>(a.texels, o.cell);
      gd[0] = (double)g3.x;
      gd[1] = (double)g3.y;
      gd[2] = (double)g3.z;
      const double i0 = (double)o.idx[0], i1 = (double)o.idx[1], i2 = (double)o.idx[2];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double p = a.virt.origin[k];
        p += a.virt.i2p[k * D + 0] * i0;
        p += a.virt.i2p[k * D + 1] * i1;
        if (D == 3) p += a.virt.i2p[k * D + 2] * i2;
        vp[k] = p;
      }
    } else {
      int32_t idx[3];
      sample_index<D>(a.virt, s, idx);
      SampleOut<D> o;
      if (!eval_sample<D, MK>(a, idx, o)) continue;
      w = parzen_deriv_weight((float)((double)(o.mbin - 1) - o.mterm), pr + o.fbin * B + (o.mbin - 1));
      if (w == 0.0f) continue;
      if (a.gradient_source == 1) {
        interp_gradient<D>(a.texels, o.cell, g);
#pragma unroll
        for (int k = 0; k < D; ++k) gd[k] = (double)g[k];
      } else {
        central_difference<D>(a, o.mp, gd);
      }
#pragma unroll
      for (int k = 0; k < D; ++k) vp[k] = o.vp[k];
      if constexpr (MK == 2) {
        // B-spline Jacobian: J[d][node + d*pdim] = weight(node) (BSplTf:594-650)
        if (o.sup.inside) {
          constexpr int NW = D == 3 ? 64 : 16;
          for (int k = 0; k < NW; ++k) {
            const double wk = (double)w * bspline_weight<D>(o.sup, k);
            const int32_t node = bspline_node<D>(a.mtf.grid, o.sup, k);
#pragma unroll
            for (int d = 0; d < D; ++d) atomicAdd(&a.deriv_accum[node + d * a.mtf.pdim], wk * gd[d]);
          }
        }
        continue;
      }
    }
    // affine Jacobian: J[b][b*D+d] = x_d - c_d, J[d][D*D+d] = 1 (MOTB:563-591)
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const double wg = (double)w * gd[b];
#pragma unroll
      for (int d = 0; d < D; ++d) acc[b * D + d] = fma(wg, vp[d] - a.mtf.center[d], acc[b * D + d]);
      acc[D * D + b] += wg;
    }
  }
  if constexpr (MK == 1) block_reduce_array<PA>(acc, scratch, a.partials + (int64_t)blockIdx.x * PA);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------




#define MMI_DISPATCH(KERNEL, a, grid, lds, st, fast)                                                          \
  do {                                                                                                        \
    const int mode_ = (a).mtf.kind == 2 ? kExactBSpline : ((fast) ? kFastAffine : kExactAffine);             \
    if ((a).virt.dim == 2) {                                                                                  \
      if (mode_ == kFastAffine)                                                                               \
        hipLaunchKernelGGL((KERNEL<2, kFastAffine>), dim3(grid), dim3(256), lds, st, a);                      \
      else if (mode_ == kExactBSpline)                                                                        \
        hipLaunchKernelGGL((KERNEL<2, kExactBSpline>), dim3(grid), dim3(256), lds, st, a);                    \
      else                                                                                                    \
        hipLaunchKernelGGL((KERNEL<2, kExactAffine>), dim3(grid), dim3(256), lds, st, a);                     \
    } else {                                                                                                  \
      if (mode_ == kFastAffine)                                                                               \
        hipLaunchKernelGGL((KERNEL<3, kFastAffine>), dim3(grid), dim3(256), lds, st, a);                      \
      else if (mode_ == kExactBSpline)                                                                        \
        hipLaunchKernelGGL((KERNEL<3, kExactBSpline>), dim3(grid), dim3(256), lds, st, a);                    \
      else                                                                                                    \
        hipLaunchKernelGGL((KERNEL<3, kExactAffine>), dim3(grid), dim3(256), lds, st, a);                     \
    }                                                                                                         \
  } while (0)

// grid: a multiple of 8 (one share per XCD) on the fast path.
hipError_t launch_pass1(const EvalArgs& a, int grid, bool fast, hipStream_t st) {
  const size_t lds = (a.hist_in_lds ? (size_t)a.bins * a.bins * 8 : 0) + (size_t)(a.bins + 1) * 4;
  if (fast && a.mtf.kind == 1) {
    if (a.virt.dim == 2)
      hipLaunchKernelGGL(pass1_fast_kernel<2>, dim3(grid), dim3(256), lds, st, a);
    else
      hipLaunchKernelGGL(pass1_fast_kernel<3>, dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
  }
  MMI_DISPATCH(pass1_kernel, a, grid, lds, st, false);
  return hipGetLastError();
}

hipError_t launch_pass2(const EvalArgs& a, int grid, bool fast, hipStream_t st) {
  if (fast && a.mtf.kind == 1) {
    const size_t lds = ((size_t)a.bins * a.bins + 128) * 4;
    if (a.virt.dim == 2)
      hipLaunchKernelGGL(pass2_fast_kernel<2>, dim3(grid), dim3(256), lds, st, a);
    else
      hipLaunchKernelGGL(pass2_fast_kernel<3>, dim3(grid), dim3(256), lds, st, a);
    return hipGetLastError();
  }
  const size_t lds = ((a.pratio_in_lds ? (size_t)a.bins * a.bins : 0) + 128) * 4;
  MMI_DISPATCH(pass2_kernel, a, grid, lds, st, false);
  return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeArgs& f, hipStream_t st) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, f);
  return hipGetLastError();
}

}  // namespace mmi
