// Mattes mutual information v4, dense value + derivative, on gfx950: the exact
// evaluation kernels and the FP64 finalisation.
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
// This file holds the exact path (FP64 in the reference's operation order, any
// transform); mattes_fast.hip holds the certified fast path for affine
// transforms.  Compiled with -ffp-contract=off: the FP64 coordinate path must
// round like the reference's x86-64 build so that bins match bit for bit.
#include "mattes_eval.h"

#include <hip/hip_runtime.h>

namespace mmi {

constexpr int kExactAffine = 0;   // FP64 reference path, affine moving transform
constexpr int kExactBSpline = 2;  // FP64 reference path, cubic B-spline moving transform

// ---------------------------------------------------------------------------
// pass 1: joint histogram
// ---------------------------------------------------------------------------

template <int D, int MODE, typename P>
__global__ void __launch_bounds__(256) pass1_kernel(EvalArgs a) {
  constexpr int MK = MODE == kExactBSpline ? 2 : 1;
  extern __shared__ unsigned long long lds_u64[];
  const int B = a.bins;
  const int BB = B * B;
  unsigned long long* hist = a.hist_in_lds ? lds_u64 : a.hist;
  unsigned int* nvalid = reinterpret_cast<unsigned int*>(lds_u64 + (a.hist_in_lds ? BB : 0));
  if (a.hist_in_lds) {
    for (int i = threadIdx.x; i < BB; i += blockDim.x) hist[i] = 0ull;
  }
  if (threadIdx.x == 0) *nvalid = 0u;
  __syncthreads();

  const uint32_t n = (uint32_t)a.n;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
    double vp[3];
    virtual_point<D>(a, s, vp);
    SampleOut<D> o;
    const bool ok = eval_sample<D, MK, P>(a, vp, o);
    if (a.sample_fbin) {
      const uint32_t so = a.point_ids ? a.point_ids[s] : s;
      a.sample_fbin[so] = ok ? (int16_t)o.fbin : (int16_t)-1;
      a.sample_mbin[so] = ok ? (int16_t)o.mbin : (int16_t)-1;
    }
    const unsigned long long vm = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && vm) atomicAdd(nvalid, (unsigned int)__popcll(vm));
    if (!ok) continue;
    unsigned int q[4];
    parzen_fixed((float)((double)(o.mbin - 1) - o.mterm), q);
    unsigned long long* row = hist + o.fbin * B + (o.mbin - 1);
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
  if (threadIdx.x == 0 && *nvalid) atomicAdd(a.nvalid, (unsigned long long)*nvalid);
}

// ---------------------------------------------------------------------------
// finalize: ComputeResults (Mattes:353-499) in FP64, one workgroup
// ---------------------------------------------------------------------------

// One workgroup of kFinalizeThreads.  The joint PDF, marginals and log rows
// live in LDS when they fit (f.lds != 0), else in the global output buffers.
// Every sum has a fixed shape (per-thread strided sums, wavefront shuffle
// tree, wavefront results in order), so the result does not depend on
// scheduling.
constexpr int kFinalizeThreads = 1024;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  return x;
}

__device__ __forceinline__ void finalize_body(const FinalizeArgs& f) {
  extern __shared__ double lds_d[];
  __shared__ unsigned long long redu[kFinalizeThreads / 64];
  __shared__ double reds[kFinalizeThreads / 64];
  __shared__ double s_norm;
  __shared__ unsigned long long s_nvalid;
  const int B = f.bins;
  const int BB = B * B;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6, nw = blockDim.x >> 6;
  double* pdf = f.lds ? lds_d : f.pdf;
  double* fm = f.lds ? lds_d + BB : f.fm;
  double* mm = f.lds ? lds_d + BB + B : f.mm;
  double* lrow = f.lds ? lds_d + BB + 2 * B : f.rowsum;
  // pass 1's histogram, read once into the PDF's place (as integers) and left
  // zeroed for the next evaluation, with its exact 64-bit total
  unsigned long long* stage = reinterpret_cast<unsigned long long*>(pdf);
  unsigned long long part = 0;
  for (int i = t; i < BB; i += blockDim.x) {
    const unsigned long long h = f.hist[i];
    f.hist[i] = 0ull;
    stage[i] = h;
    part += h;
  }
  part = wave_sum(part);
  if (lane == 0) redu[wid] = part;
  if (t == 0) {
    s_nvalid = *f.nvalid;
    *f.nvalid = 0ull;
    f.result->exc_overflow = *f.overflow ? 1 : 0;
    *f.overflow = 0u;
  }
  __syncthreads();
  if (t == 0) {
    unsigned long long total = 0;
    for (int w = 0; w < nw; ++w) total += redu[w];
    const unsigned long long nvalid = s_nvalid;
    const double jointPDFSum = (double)total / kFixedPointScale;
    f.result->n_valid = nvalid;
    f.result->joint_pdf_sum = jointPDFSum;
    f.result->status = (nvalid == 0) ? 1 : (jointPDFSum < 2.220446049250313e-16 ? 2 : 0);
    s_norm = (nvalid == 0 || jointPDFSum < 2.220446049250313e-16) ? 0.0 : 1.0 / jointPDFSum;
  }
  __syncthreads();
  const double nf = s_norm;
  if (nf == 0.0) return;
  const double closeToZero = 2.220446049250313e-16;
  // fixed marginal (the reference counts each valid sample once in its fixed
  // bin): the Parzen weights of a counted sample sum to 2^32 within a few
  // units, so a row's total rounded to a multiple of 2^32 is the row's count
  // exactly (below 2^30 samples).  Sixteen lanes per row, an exact integer
  // sum.
  for (int r = t >> 4; r < B; r += blockDim.x >> 4) {  // r is uniform over each group of sixteen lanes
    unsigned long long rs = 0ull;
    for (int m = t & 15; m < B; m += 16) rs += stage[r * B + m];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) rs += __shfl_xor(rs, o, 16);
    if ((t & 15) == 0) {
      const double fmv = (double)((rs + (1ull << 31)) >> 32) * nf;
      fm[r] = fmv;
      if (f.lds) f.fm[r] = fmv;
      lrow[r] = fmv > closeToZero ? log(fmv) : 0.0;  // log of the fixed marginal, once per row
    }
  }
  __syncthreads();
  for (int i = t; i < BB; i += blockDim.x) {
    const double v = ((double)stage[i] / kFixedPointScale) * nf;
    pdf[i] = v;
    if (f.lds) f.pdf[i] = v;
  }
  __syncthreads();
  // moving marginal: column sums in the reference's row order (Mattes:405-417),
  // the loads of a few rows issued ahead of their sums
  for (int m = t; m < B; m += blockDim.x) {
    double acc = 0.0;
    int r = 0;
    for (; r + 8 <= B; r += 8) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = pdf[(r + k) * B + m];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += v[k];
    }
    for (; r < B; ++r) acc += pdf[r * B + m];
    mm[m] = acc;
    if (f.lds) f.mm[m] = acc;
  }
  __syncthreads();
  // entropy terms and pRatio over all (f, m) pairs in parallel (Mattes:419-476)
  double part_sum = 0.0;
  for (int i = t; i < BB; i += blockDim.x) {
    const int r = i / B, m = i - r * B;
    const double fmv = fm[r];
    const double mmv = mm[m];
    const double jp = pdf[i];
    double pr = 0.0;
    if (fmv > closeToZero && mmv > closeToZero && jp > closeToZero) {
      pr = log(jp / mmv);
      part_sum += jp * (pr - lrow[r]);
    }
    f.pratio[i] = pr;
    f.pratio_f[i] = (float)pr;
  }
  part_sum = wave_sum(part_sum);
  if (lane == 0) reds[wid] = part_sum;
  __syncthreads();
  if (t == 0) {
    double sum = 0.0;
    for (int w = 0; w < nw; ++w) sum += reds[w];
    f.result->value = -1.0 * sum;
    f.result->n_factor = -1.0 / (f.moving_bin_size * (double)s_nvalid);
  }
}

__global__ void __launch_bounds__(kFinalizeThreads) finalize_kernel(FinalizeArgs f) { finalize_body(f); }

__global__ void __launch_bounds__(kFinalizeThreads) finalize_batch(const FinalizeArgs* __restrict__ fz,
                                                                   const int32_t* __restrict__ halted) {
  if (halted && halted[blockIdx.x]) return;
  finalize_body(fz[blockIdx.x]);
}

// ---------------------------------------------------------------------------
// pass 2: derivative
// ---------------------------------------------------------------------------

template <int D, int MODE, typename P>
__global__ void __launch_bounds__(256) pass2_kernel(EvalArgs a) {
  constexpr int MK = MODE == kExactBSpline ? 2 : 1;
  extern __shared__ float lds_f32[];
  const int B = a.bins;
  const int BB = B * B;
  double* scratch = reinterpret_cast<double*>(lds_f32);  // 4 wavefronts x 12 doubles (128 floats)
  // pRatio table in LDS when it fits, else read through the caches.
  float* lds_pr = lds_f32 + 128;
  const float* pr = lds_pr;
  if (a.pratio_in_lds) {
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
    double vp[3];
    virtual_point<D>(a, s, vp);
    SampleOut<D> o;
    const bool ok = eval_sample<D, MK, P>(a, vp, o);
    float w = 0.0f;
    if (ok) w = parzen_deriv_weight((float)((double)(o.mbin - 1) - o.mterm), pr + o.fbin * B + (o.mbin - 1));
    double gd[3] = {0.0, 0.0, 0.0};
    if (w != 0.0f) {
      if (a.gradient_source == 1) {
        const float3 g3 = interp_gradient3<D, P>(static_cast<const typename P::texel*>(a.texels), o.cell);
        gd[0] = (double)g3.x;
        gd[1] = (double)g3.y;
        gd[2] = (double)g3.z;
      } else {
        central_difference<D, P>(a, o.mp, gd);
      }
      // composite moving transform: the Jacobian is the stages' matrices times
      // the moving transform's (itkCompositeTransform.hxx:467-560), so the
      // gradient meets the moving transform's Jacobian as M0^T .. Mn-1^T g
      for (int st = 0; st < a.n_init; ++st) {
        const double* m = a.init_stages + 12 * st;
        double h[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double sum = 0.0;
#pragma unroll
          for (int r = 0; r < D; ++r) sum += m[r * D + c] * gd[r];
          h[c] = sum;
        }
#pragma unroll
        for (int c = 0; c < D; ++c) gd[c] = h[c];
      }
    }
    if constexpr (MK == 2) {
      // B-spline Jacobian: J[d][node + d*pdim] = weight(node) (BSplTf:594-650)
      const bool use = w != 0.0f && o.sup.inside;
      if (a.bspline_g) {
        // separable geometry: per-sample factors, contracted axis by axis later
#pragma unroll
        for (int d = 0; d < D; ++d) a.bspline_g[(int64_t)d * n + s] = use ? (float)((double)w * gd[d]) : 0.0f;
      } else if (use) {
        constexpr int NW = D == 3 ? 64 : 16;
        for (int k = 0; k < NW; ++k) {
          const double wk = (double)w * bspline_weight<D>(o.sup, k);
          const int32_t node = bspline_node<D>(a.mtf.grid, o.sup, k);
#pragma unroll
          for (int d = 0; d < D; ++d) atomicAdd(&a.deriv_accum[node + d * a.mtf.pdim], wk * gd[d]);
        }
      }
    } else if (w != 0.0f) {
      // affine Jacobian: J[b][b*D+d] = x_d - c_d, J[d][D*D+d] = 1 (MOTB:563-591)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        const double wg = (double)w * gd[b];
#pragma unroll
        for (int d = 0; d < D; ++d) acc[b * D + d] = fma(wg, o.vp[d] - a.mtf.center[d], acc[b * D + d]);
        acc[D * D + b] += wg;
      }
    }
  }
  if constexpr (MK == 1) block_reduce_array<PA>(acc, scratch, a.partials + (int64_t)blockIdx.x * PA);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------

#define MMI_DISPATCH_EXACT_P(KERNEL, P, a, grid, lds, st)                                         \
  do {                                                                                            \
    if ((a).virt.dim == 2) {                                                                      \
      if ((a).mtf.kind == 2)                                                                      \
        hipLaunchKernelGGL((KERNEL<2, kExactBSpline, P>), dim3(grid), dim3(256), lds, st, a);     \
      else                                                                                        \
        hipLaunchKernelGGL((KERNEL<2, kExactAffine, P>), dim3(grid), dim3(256), lds, st, a);      \
    } else {                                                                                      \
      if ((a).mtf.kind == 2)                                                                      \
        hipLaunchKernelGGL((KERNEL<3, kExactBSpline, P>), dim3(grid), dim3(256), lds, st, a);     \
      else                                                                                        \
        hipLaunchKernelGGL((KERNEL<3, kExactAffine, P>), dim3(grid), dim3(256), lds, st, a);      \
    }                                                                                             \
  } while (0)

#define MMI_DISPATCH_EXACT(KERNEL, a, grid, lds, st)                                              \
  do {                                                                                            \
    if ((a).storage == kStorageBF16)                                                              \
      MMI_DISPATCH_EXACT_P(KERNEL, PxBF16, a, grid, lds, st);                                     \
    else if ((a).storage == kStorageF16)                                                          \
      MMI_DISPATCH_EXACT_P(KERNEL, PxF16, a, grid, lds, st);                                      \
    else                                                                                          \
      MMI_DISPATCH_EXACT_P(KERNEL, PxF32, a, grid, lds, st);                                      \
  } while (0)

hipError_t launch_pass1(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = (a.hist_in_lds ? (size_t)a.bins * a.bins * 8 : 0) + 8;
  MMI_DISPATCH_EXACT(pass1_kernel, a, grid, lds, st);
  return hipGetLastError();
}

hipError_t launch_pass2(const EvalArgs& a, int grid, hipStream_t st) {
  const size_t lds = ((a.pratio_in_lds ? (size_t)a.bins * a.bins : 0) + 128) * 4;
  MMI_DISPATCH_EXACT(pass2_kernel, a, grid, lds, st);
  return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeArgs& f, hipStream_t st) {
  FinalizeArgs g = f;
  g.lds = finalize_uses_lds(f.bins) ? 1 : 0;
  const size_t lds = g.lds ? ((size_t)f.bins * f.bins + 3 * (size_t)f.bins) * sizeof(double) : 0;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kFinalizeThreads), lds, st, g);
  return hipGetLastError();
}

hipError_t launch_finalize_batch(const FinalizeArgs* fz, int n, int bins_max, const int32_t* halted, hipStream_t st) {
  const size_t lds =
      finalize_uses_lds(bins_max) ? ((size_t)bins_max * bins_max + 3 * (size_t)bins_max) * sizeof(double) : 0;
  hipLaunchKernelGGL(finalize_batch, dim3(n), dim3(kFinalizeThreads), lds, st, fz, halted);
  return hipGetLastError();
}

}  // namespace mmi
