// Mattes mutual information v4, dense value + derivative, on gfx950.
//
// One evaluation = three kernels on the context stream:
//   pass 1   every virtual sample -> fixed/moving Parzen bins; 4 cubic B-spline
//            Parzen weights into a per-workgroup joint histogram in LDS held as
//            64-bit fixed point (weight * 2^32), fixed-marginal and valid
//            counts; LDS partials merged into HBM with 64-bit integer atomics.
//            Integer accumulation makes the histogram independent of
//            scheduling order (bit-reproducible).
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

// ProcessVirtualPoint + Mattes ProcessPoint up to the bins
// (ThrBase:174-273, I2IM:289-373, MattesT:215-275).
template <int D>
__device__ __forceinline__ bool eval_sample(const EvalArgs& a, int64_t s, SampleOut<D>& o) {
  int64_t idx[3];
  const int64_t sx = a.virt.size[0];
  idx[0] = a.virt.start[0] + s % sx;
  if (D == 2) {
    idx[1] = a.virt.start[1] + s / sx;
    idx[2] = 0;
  } else {
    const int64_t sy = a.virt.size[1];
    idx[1] = a.virt.start[1] + (s / sx) % sy;
    idx[2] = a.virt.start[2] + s / (sx * sy);
  }
  index_to_point<D>(a.virt, idx, o.vp);

  double fp[3], ci[3];
  BSplineSupport<D> dummy;
  transform_point<D>(a.ftf, o.vp, fp, &dummy);
  point_to_cindex<D>(a.fixed, fp, ci);
  if (!inside_buffer<D>(a.fixed, ci)) return false;
  const Cell fc = make_cell<D>(a.fixed, ci);
  const double fval = (double)(float)interp_exact<D>(a.fixed_px, fc);

  transform_point<D>(a.mtf, o.vp, o.mp, &o.sup);
  point_to_cindex<D>(a.moving, o.mp, ci);
  if (!inside_buffer<D>(a.moving, ci)) return false;
  o.cell = make_cell<D>(a.moving, ci);
  const double mval = (double)(float)interp_exact<D>(a.moving_px, o.cell);

  if (mval < a.moving_min) return false;
  if (mval > a.moving_max) return false;
  const double mterm = mval / a.moving_bin_size - a.moving_norm_min;
  int mb = mterm < 2.0 ? 2 : (mterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)mterm);
  const double fterm = fval / a.fixed_bin_size - a.fixed_norm_min;
  int fb = fterm < 2.0 ? 2 : (fterm >= (double)(a.bins - 2) ? a.bins - 3 : (int)fterm);
  o.fbin = fb;
  o.mbin = mb;
  o.mterm = mterm;
  return true;
}

template <int D>
__global__ void __launch_bounds__(256) pass1_kernel(EvalArgs a) {
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

  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n; s += stride) {
    SampleOut<D> o;
    const bool ok = eval_sample<D>(a, s, o);
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)o.fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)o.mbin : (int16_t)-1;
    }
    if (!ok) continue;
    atomicAdd(&fmarg[o.fbin], 1u);
    atomicAdd(nvalid, 1u);
    double arg = (double)(o.mbin - 1) - o.mterm;
    unsigned long long* row = hist + (int64_t)o.fbin * B + (o.mbin - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double w = bspline3(arg);
      const unsigned long long q = (unsigned long long)__builtin_rint(w * kFixedPointScale);
      if (q) atomicAdd(&row[k], q);
      arg += 1.0;
    }
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

// ComputeResults (Mattes:353-499) in FP64, one workgroup.
__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f) {
  __shared__ double s_norm;
  const int B = f.bins;
  const int BB = B * B;
  const int t = threadIdx.x;
  // Joint PDF sum: exact in 64-bit integers.
  unsigned long long part = 0;
  for (int i = t; i < BB; i += blockDim.x) part += f.hist[i];
  __shared__ unsigned long long redu[256];
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
  }
  __shared__ double reds[256];
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

template <int D>
__device__ __forceinline__ double block_reduce_sum(double v, double* scratch) {
  // wavefront (64 lanes) shuffle reduction, then across wavefronts in LDS
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += scratch[w];
  }
  __syncthreads();
  return r;
}

template <int D>
__global__ void __launch_bounds__(256) pass2_kernel(EvalArgs a) {
  extern __shared__ double lds_f64[];
  const int B = a.bins;
  const int BB = B * B;
  // pRatio table in LDS when it fits (B <= 90), else read through the caches.
  double* scratch = lds_f64;
  const double* pr = a.pratio;
  if (a.hist_in_lds) {
    double* t = lds_f64 + 8;
    for (int i = threadIdx.x; i < BB; i += blockDim.x) t[i] = a.pratio[i];
    pr = t;
  }
  __syncthreads();

  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;

  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n; s += stride) {
    SampleOut<D> o;
    if (!eval_sample<D>(a, s, o)) continue;
    // sum_k B3'(arg_k) * pRatio[f][m-1+k], arg advanced as in MattesT:329-378
    double arg = (double)(o.mbin - 1) - o.mterm;
    const double* prow = pr + o.fbin * B + (o.mbin - 1);
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      w += bspline3_deriv(arg) * prow[k];
      arg += 1.0;
    }
    if (w == 0.0) continue;
    double g[3] = {0.0, 0.0, 0.0};
    if (a.gradient_source == 1) {
      float gf[3];
      interp_gradient<D>(a.texels, o.cell, gf);
#pragma unroll
      for (int k = 0; k < D; ++k) g[k] = (double)gf[k];
    } else {
      central_difference<D>(a, o.mp, g);
    }
    if (a.mtf.kind == 1) {
      // affine Jacobian: J[b][b*D+d] = x_d - c_d, J[d][D*D+d] = 1 (MOTB:563-591)
      double v[3];
#pragma unroll
      for (int d = 0; d < D; ++d) v[d] = o.vp[d] - a.mtf.center[d];
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const double wg = w * g[b];
#pragma unroll
        for (int d = 0; d < D; ++d) acc[b * D + d] += wg * v[d];
        acc[D * D + b] += wg;
      }
    } else if (a.mtf.kind == 2 && o.sup.inside) {
      // B-spline Jacobian: J[d][node + d*pdim] = weight(node) (BSplTf:594-650)
      constexpr int NW = D == 3 ? 64 : 16;
      for (int k = 0; k < NW; ++k) {
        const double wk = w * bspline_weight<D>(o.sup, k);
        const int64_t node = bspline_node<D>(a.mtf.grid, o.sup, k);
#pragma unroll
        for (int d = 0; d < D; ++d) atomicAdd(&a.deriv_accum[node + d * a.mtf.pdim], wk * g[d]);
      }
    }
  }
  if (a.mtf.kind == 1) {
#pragma unroll
    for (int i = 0; i < PA; ++i) {
      const double r = block_reduce_sum<D>(acc[i], scratch);
      if (threadIdx.x == 0) a.partials[(int64_t)blockIdx.x * PA + i] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Certified fast path (affine moving transform)
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

template <int D>
struct FastSample {
  int fbin;
  int mbin;
  float arg0;  // (mbin - 1) - mterm
  Cell cell;   // moving interpolation cell
  int64_t idx[3];
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
__device__ __forceinline__ bool fast_sample(const EvalArgs& a, int64_t s, FastSample<D>& o) {
  const FastAffine& F = a.fast;
  const int64_t sx = a.virt.size[0];
  int64_t lin[3];
  lin[0] = s % sx;
  if (D == 2) {
    lin[1] = s / sx;
    lin[2] = 0;
  } else {
    const int64_t sy = a.virt.size[1];
    lin[1] = (s / sx) % sy;
    lin[2] = s / (sx * sy);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) o.idx[k] = a.virt.start[k] + lin[k];
  bool uncertain = false;

  // fixed value
  float fv = 0.0f;
  if (F.fixed_direct) {
    const int64_t fx = o.idx[0] + F.fixed_offset[0] - a.fixed.start[0];
    const int64_t fy = o.idx[1] + F.fixed_offset[1] - a.fixed.start[1];
    const int64_t fz = D == 3 ? o.idx[2] + F.fixed_offset[2] - a.fixed.start[2] : 0;
    fv = a.fixed_px[fx + a.fixed.sy * fy + a.fixed.sz * fz];
  } else {
    uncertain = true;  // general fixed geometry: exact path
  }

  // moving continuous index, one FP64 affine map
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
  if (!eval_sample<D>(a, s, e)) return false;
  o.fbin = e.fbin;
  o.mbin = e.mbin;
  o.arg0 = (float)((double)(e.mbin - 1) - e.mterm);
  o.cell = e.cell;
  return true;
}

template <int D>
__global__ void __launch_bounds__(256) pass1_fast_kernel(EvalArgs a) {
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
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n; s += stride) {
    FastSample<D> o;
    const bool ok = fast_sample<D>(a, s, o);
    if (a.sample_fbin) {
      a.sample_fbin[s] = ok ? (int16_t)o.fbin : (int16_t)-1;
      a.sample_mbin[s] = ok ? (int16_t)o.mbin : (int16_t)-1;
    }
    const unsigned long long vm = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && vm) atomicAdd(nvalid, (unsigned int)__popcll(vm));
    if (!ok) continue;
    atomicAdd(&fmarg[o.fbin], 1u);
    float arg = o.arg0;
    unsigned long long* row = hist + (int64_t)o.fbin * B + (o.mbin - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float w = bspline3f(arg);
      const unsigned long long q = (unsigned long long)(unsigned int)rintf(w * 4294967296.0f);
      if (q) atomicAdd(&row[k], q);
      arg += 1.0f;
    }
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

template <int D>
__global__ void __launch_bounds__(256) pass2_fast_kernel(EvalArgs a) {
  extern __shared__ float lds_f32[];
  const int B = a.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);
  float* pr = lds_f32 + 16;
  for (int i = threadIdx.x; i < BB; i += blockDim.x) pr[i] = (float)a.pratio[i];
  __syncthreads();
  constexpr int PA = D * D + D;
  double acc[PA];
#pragma unroll
  for (int i = 0; i < PA; ++i) acc[i] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n; s += stride) {
    FastSample<D> o;
    if (!fast_sample<D>(a, s, o)) continue;
    float arg = o.arg0;
    const float* prow = pr + o.fbin * B + (o.mbin - 1);
    float w = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      w = fmaf(bspline3_derivf(arg), prow[k], w);
      arg += 1.0f;
    }
    if (w == 0.0f) continue;
    float g[3];
    interp_gradient<D>(a.texels, o.cell, g);
    double vp[3];
    index_to_point<D>(a.virt, o.idx, vp);
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const double wg = (double)(w * g[b]);
#pragma unroll
      for (int d = 0; d < D; ++d) acc[b * D + d] = fma(wg, vp[d] - a.mtf.center[d], acc[b * D + d]);
      acc[D * D + b] += wg;
    }
  }
#pragma unroll
  for (int i = 0; i < PA; ++i) {
    const double r = block_reduce_sum<D>(acc[i], scratch);
    if (threadIdx.x == 0) a.partials[(int64_t)blockIdx.x * PA + i] = r;
  }
}

// Deterministic sum of per-workgroup partials, then the nFactor scaling.
// One workgroup per parameter; strided partial sums then a fixed-shape tree,
// so the result does not depend on scheduling.
__global__ void __launch_bounds__(256) derivative_reduce_kernel(const double* __restrict__ partials, int nblocks,
                                                                int P, const DevResult* __restrict__ result,
                                                                double* __restrict__ deriv) {
  __shared__ double red[256];
  const int mu = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(int64_t)b * P + mu];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) deriv[mu] = red[0] * result->n_factor;
}

__global__ void derivative_scale_kernel(double* __restrict__ deriv, int64_t P, const DevResult* __restrict__ result) {
  const int64_t mu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mu < P) deriv[mu] = deriv[mu] * result->n_factor;
}

// Intensity range over a whole buffer (MattesMI::Initialize, Mattes:97-229),
// 2 doubles per workgroup, reduced on the host.
__global__ void minmax_kernel(const float* __restrict__ px, int64_t n, float* __restrict__ out) {
  __shared__ float smin[256], smax[256];
  float lo = 3.402823466e38f, hi = -3.402823466e38f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = px[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = smin[0];
    out[2 * blockIdx.x + 1] = smax[0];
  }
}

// Recursive Gaussian along one axis of a double volume, one line per thread
// (RecursiveSeparableImageFilter::FilterDataArray,
// Modules/Filtering/ImageFilterBase/include/itkRecursiveSeparableImageFilter.hxx:62-143):
// the causal pass is written to `out`, the anti-causal pass is added in a
// second sweep; both keep their 4-tap histories in registers.
__global__ void recursive_gaussian_kernel(const double* __restrict__ in, double* __restrict__ out, int64_t ln,
                                          int64_t stride, int64_t nlines, RecCoeffs c) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  const int64_t lo = l % stride, hi = l / stride;
  const double* x = in + lo + hi * stride * ln;
  double* y = out + lo + hi * stride * ln;
  auto X = [&](int64_t i) { return x[i * stride]; };
  // causal
  const double v1 = X(0);
  double o0, o1, o2, o3;  // outs[i-1..i-4]
  {
    double d0 = X(0), d1 = X(1), d2 = X(2), d3 = X(3);
    double q0 = v1 * c.N0 + v1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q1 = d1 * c.N0 + v1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q2 = d2 * c.N0 + d1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q3 = d3 * c.N0 + d2 * c.N1 + d1 * c.N2 + v1 * c.N3;
    (void)d0;
    q0 -= v1 * c.BN1 + v1 * c.BN2 + v1 * c.BN3 + v1 * c.BN4;
    q1 -= q0 * c.D1 + v1 * c.BN2 + v1 * c.BN3 + v1 * c.BN4;
    q2 -= q1 * c.D1 + q0 * c.D2 + v1 * c.BN3 + v1 * c.BN4;
    q3 -= q2 * c.D1 + q1 * c.D2 + q0 * c.D3 + v1 * c.BN4;
    y[0] = q0;
    y[stride] = q1;
    y[2 * stride] = q2;
    y[3 * stride] = q3;
    o0 = q3;
    o1 = q2;
    o2 = q1;
    o3 = q0;
  }
  double xm1 = X(3), xm2 = X(2), xm3 = X(1);
  for (int64_t i = 4; i < ln; ++i) {
    const double xi = X(i);
    double q = xi * c.N0 + xm1 * c.N1 + xm2 * c.N2 + xm3 * c.N3;
    q -= o0 * c.D1 + o1 * c.D2 + o2 * c.D3 + o3 * c.D4;
    y[i * stride] = q;
    o3 = o2;
    o2 = o1;
    o1 = o0;
    o0 = q;
    xm3 = xm2;
    xm2 = xm1;
    xm1 = xi;
  }
  // anti-causal
  const double v2 = X(ln - 1);
  double s0, s1, s2, s3;  // scratch[i+1..i+4]
  {
    const double e1 = X(ln - 1), e2 = X(ln - 2), e3 = X(ln - 3);
    double r1 = v2 * c.M1 + v2 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r2 = e1 * c.M1 + v2 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r3 = e2 * c.M1 + e1 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r4 = e3 * c.M1 + e2 * c.M2 + e1 * c.M3 + v2 * c.M4;
    r1 -= v2 * c.BM1 + v2 * c.BM2 + v2 * c.BM3 + v2 * c.BM4;
    r2 -= r1 * c.D1 + v2 * c.BM2 + v2 * c.BM3 + v2 * c.BM4;
    r3 -= r2 * c.D1 + r1 * c.D2 + v2 * c.BM3 + v2 * c.BM4;
    r4 -= r3 * c.D1 + r2 * c.D2 + r1 * c.D3 + v2 * c.BM4;
    y[(ln - 1) * stride] += r1;
    y[(ln - 2) * stride] += r2;
    y[(ln - 3) * stride] += r3;
    y[(ln - 4) * stride] += r4;
    s0 = r4;
    s1 = r3;
    s2 = r2;
    s3 = r1;
  }
  double xp1 = X(ln - 3), xp2 = X(ln - 2), xp3 = X(ln - 1);  // data[i+1..i+3] for i = ln-4
  for (int64_t i = ln - 4; i > 0; --i) {
    const double xi = X(i);
    // scratch[i-1] = data[i]*M1 + data[i+1]*M2 + data[i+2]*M3 + data[i+3]*M4 - (...)
    double q = xi * c.M1 + xp1 * c.M2 + xp2 * c.M3 + xp3 * c.M4;
    q -= s0 * c.D1 + s1 * c.D2 + s2 * c.D3 + s3 * c.D4;
    y[(i - 1) * stride] += q;
    s3 = s2;
    s2 = s1;
    s1 = s0;
    s0 = q;
    xp3 = xp2;
    xp2 = xp1;
    xp1 = xi;
  }
}

__global__ void float_to_double_kernel(const float* __restrict__ in, double* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

// gradient component `comp` = vol / spacing into the texel's slot comp+1.
__global__ void gradient_component_kernel(const double* __restrict__ vol, int64_t n, double spacing, int comp,
                                          double* __restrict__ grad, int D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grad[i * D + comp] = vol[i] / spacing;
}

// TransformLocalVectorToPhysicalVector (itkImageBase.h:636-653) + packing into
// float4 texels (value, gx, gy, gz).
__global__ void pack_texels_kernel(const float* __restrict__ moving, const double* __restrict__ grad, int64_t n,
                                   int D, Direction dir, float4* __restrict__ tex) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double in[3] = {0, 0, 0}, o[3] = {0, 0, 0};
  for (int j = 0; j < D; ++j) in[j] = grad[i * D + j];
  for (int r = 0; r < D; ++r) {
    double s = 0.0;
    for (int c = 0; c < D; ++c) s += dir.m[r * D + c] * in[c];
    o[r] = s;
  }
  tex[i] = make_float4(moving[i], (float)o[0], (float)o[1], (float)o[2]);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------

hipError_t launch_pass1(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = (a.hist_in_lds ? (size_t)a.bins * a.bins * 8 : 0) + (size_t)(a.bins + 1) * 4;
  if (a.virt.dim == 2)
    hipLaunchKernelGGL(pass1_kernel<2>, dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(pass1_kernel<3>, dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_pass1_fast(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = (a.hist_in_lds ? (size_t)a.bins * a.bins * 8 : 0) + (size_t)(a.bins + 1) * 4;
  if (a.virt.dim == 2)
    hipLaunchKernelGGL(pass1_fast_kernel<2>, dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(pass1_fast_kernel<3>, dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_pass2_fast(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = ((size_t)a.bins * a.bins + 16) * 4;
  if (a.virt.dim == 2)
    hipLaunchKernelGGL(pass2_fast_kernel<2>, dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(pass2_fast_kernel<3>, dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeArgs& f, hipStream_t st) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, f);
  return hipGetLastError();
}

hipError_t launch_pass2(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = (a.hist_in_lds ? (size_t)a.bins * a.bins * 8 : 0) + 8 * sizeof(double);
  if (a.virt.dim == 2)
    hipLaunchKernelGGL(pass2_kernel<2>, dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(pass2_kernel<3>, dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_derivative_reduce(const double* partials, int nblocks, int P, const DevResult* r, double* deriv,
                                    hipStream_t st) {
  hipLaunchKernelGGL(derivative_reduce_kernel, dim3(P), dim3(256), 0, st, partials, nblocks, P, r, deriv);
  return hipGetLastError();
}

hipError_t launch_derivative_scale(double* deriv, int64_t P, const DevResult* r, hipStream_t st) {
  hipLaunchKernelGGL(derivative_scale_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, deriv, P, r);
  return hipGetLastError();
}

hipError_t launch_minmax(const float* px, int64_t n, float* out, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(minmax_kernel, dim3(blocks), dim3(256), 0, st, px, n, out);
  return hipGetLastError();
}

hipError_t launch_recursive_gaussian(const double* in, double* out, int64_t ln, int64_t stride, int64_t nlines,
                                     const RecCoeffs& c, hipStream_t st) {
  hipLaunchKernelGGL(recursive_gaussian_kernel, dim3((unsigned)((nlines + 127) / 128)), dim3(128), 0, st, in, out,
                     ln, stride, nlines, c);
  return hipGetLastError();
}

hipError_t launch_float_to_double(const float* in, double* out, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(float_to_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t launch_gradient_component(const double* vol, int64_t n, double spacing, int comp, double* grad, int D,
                                     hipStream_t st) {
  hipLaunchKernelGGL(gradient_component_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vol, n, spacing,
                     comp, grad, D);
  return hipGetLastError();
}

hipError_t launch_pack_texels(const float* moving, const double* grad, int64_t n, int D, const Direction& dir,
                              float4* tex, hipStream_t st) {
  hipLaunchKernelGGL(pack_texels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, moving, grad, n, D, dir,
                     tex);
  return hipGetLastError();
}

}  // namespace mmi
