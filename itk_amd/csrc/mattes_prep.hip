// Preparation kernels (once per level) and small reductions of the Mattes MI
// v4 evaluation: intensity range, recursive-Gaussian gradient image, texel
// packing, derivative reduction.  Compiled with -ffp-contract=off like the
// evaluation kernels, so the gradient image rounds like the reference filter.
#include "mattes_eval.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace mmi {

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
template <typename P>
__global__ void minmax_kernel(const typename P::pixel* __restrict__ px, int64_t n, float* __restrict__ out) {
  __shared__ float smin[256], smax[256];
  float lo = 3.402823466e38f, hi = -3.402823466e38f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = P::ld(px + i, 0);
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

// Recursive Gaussian along one axis of a volume, one line per thread
// (RecursiveSeparableImageFilter::FilterDataArray,
// Modules/Filtering/ImageFilterBase/include/itkRecursiveSeparableImageFilter.hxx:62-143):
// the causal pass is written to `out`, the anti-causal pass is added in a
// second sweep; both keep their 4-tap histories in registers.  The input is
// the float image (first pass of a chain: its values widened exactly) or a
// double volume; the final values go where `sink` says (the double volume,
// value / spacing into one component of D-interleaved gradient vectors, or
// a float volume), so the chain needs no separate conversion kernels.
template <typename TIn>
__device__ __forceinline__ void recursive_gaussian_line(const TIn* __restrict__ in, double* __restrict__ out,
                                                        int64_t ln, int64_t stride, int64_t nlines,
                                                        const RecCoeffs& c, const RecSink& sink, int64_t l) {
  if (l >= nlines) return;
  const int64_t lo = l % stride, hi = l / stride;
  const int64_t base = lo + hi * stride * ln;
  const TIn* x = in + base;
  double* y = out + base;
  auto X = [&](int64_t i) { return (double)x[i * stride]; };
  auto FIN = [&](int64_t i, double v) {
    const int64_t e = base + i * stride;
    if (sink.mode == kSinkGradient)
      sink.grad[e * sink.dim + sink.comp] = v / sink.spacing;
    else if (sink.mode == kSinkFloat)
      sink.outf[e] = (float)v;
    else
      out[e] = v;
  };
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
  // the recursion is sequential along the line; its inputs are loaded a chunk
  // ahead so the loads of one chunk are in flight together
  constexpr int kChunk = 8;
  for (int64_t i0 = 4; i0 < ln; i0 += kChunk) {
    double xs[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) xs[k] = i0 + k < ln ? X(i0 + k) : 0.0;
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      if (i0 + k >= ln) break;
      const double xi = xs[k];
      double q = xi * c.N0 + xm1 * c.N1 + xm2 * c.N2 + xm3 * c.N3;
      q -= o0 * c.D1 + o1 * c.D2 + o2 * c.D3 + o3 * c.D4;
      y[(i0 + k) * stride] = q;
      o3 = o2;
      o2 = o1;
      o1 = o0;
      o0 = q;
      xm3 = xm2;
      xm2 = xm1;
      xm1 = xi;
    }
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
    FIN(ln - 1, y[(ln - 1) * stride] + r1);
    FIN(ln - 2, y[(ln - 2) * stride] + r2);
    FIN(ln - 3, y[(ln - 3) * stride] + r3);
    FIN(ln - 4, y[(ln - 4) * stride] + r4);
    s0 = r4;
    s1 = r3;
    s2 = r2;
    s3 = r1;
  }
  double xp1 = X(ln - 3), xp2 = X(ln - 2), xp3 = X(ln - 1);  // data[i+1..i+3] for i = ln-4
  for (int64_t i0 = ln - 4; i0 > 0; i0 -= kChunk) {
    // i runs i0, i0-1, ... > 0; it reads data[i] and updates out[i-1]
    double xs[kChunk], ys[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t i = i0 - k;
      xs[k] = i > 0 ? X(i) : 0.0;
      ys[k] = i > 0 ? y[(i - 1) * stride] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t i = i0 - k;
      if (i <= 0) break;
      const double xi = xs[k];
      // scratch[i-1] = data[i]*M1 + data[i+1]*M2 + data[i+2]*M3 + data[i+3]*M4 - (...)
      double q = xi * c.M1 + xp1 * c.M2 + xp2 * c.M3 + xp3 * c.M4;
      q -= s0 * c.D1 + s1 * c.D2 + s2 * c.D3 + s3 * c.D4;
      FIN(i - 1, ys[k] + q);
      s3 = s2;
      s2 = s1;
      s1 = s0;
      s0 = q;
      xp3 = xp2;
      xp2 = xp1;
      xp1 = xi;
    }
  }
}

// The same pass over lines along x (stride 1), which lie one after another in
// memory: with a line per thread every load and store of a wavefront would
// touch 64 lines.  The workgroup's kRecLines consecutive lines form one block;
// their elements pass through LDS kRecX at a time, read and written by
// consecutive threads along each line, and each thread runs its line's
// recursion from its LDS row with recursive_gaussian_line's arithmetic (the
// same values in the same order, so the same bits).  The first and last
// elements of a line, which start the two recursions, are read directly.
constexpr int kRecX = 8;              // elements of each line per LDS stage
constexpr int kRecPitch = kRecX + 1;  // LDS row pitch in doubles
constexpr size_t kRecRowsLdsBytes = 2 * kRecLines * kRecPitch * sizeof(double);
static_assert(kRecX == 8 && kRecLines == 128, "the staging index split assumes 8 elements of 128 lines");

// Elements [i0, i0 + n) of the block's nl lines into an LDS tile.
template <typename TIn>
__device__ __forceinline__ void rec_stage(double* tile, const TIn* __restrict__ src, int64_t blk, int64_t ln,
                                          int64_t i0, int n, int nl) {
  for (int r = threadIdx.x; r < kRecLines * kRecX; r += kRecLines) {
    const int j = r >> 3, k = r & 7;
    if (j < nl && k < n) tile[j * kRecPitch + k] = (double)src[blk + (int64_t)j * ln + i0 + k];
  }
}

// A final value of element e where the pass's sink puts it.
__device__ __forceinline__ void rec_sink_store(const RecSink& sink, double* __restrict__ out, int64_t e, double v) {
  if (sink.mode == kSinkGradient)
    sink.grad[e * sink.dim + sink.comp] = v / sink.spacing;
  else if (sink.mode == kSinkFloat)
    sink.outf[e] = (float)v;
  else
    out[e] = v;
}

template <typename TIn>
__device__ __forceinline__ void recursive_gaussian_rows(const TIn* __restrict__ in, double* __restrict__ out,
                                                        int64_t ln, int64_t nlines, const RecCoeffs& cref,
                                                        const RecSink& sref, int64_t l0, double* tx, double* ty) {
  // by value: the coefficients and sink then stay in scalar registers
  const RecCoeffs c = cref;
  const RecSink sink = sref;
  const int t = threadIdx.x;
  const bool live = l0 + t < nlines;
  const int nl = (int)min((int64_t)kRecLines, nlines - l0);
  const int64_t blk = l0 * ln;
  const int64_t base = blk + (int64_t)t * ln;  // this thread's line
  double* rx = tx + t * kRecPitch;
  double* ry = ty + t * kRecPitch;
  // causal
  double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0, xm1 = 0.0, xm2 = 0.0, xm3 = 0.0;
  if (live) {
    const double v1 = (double)in[base];
    const double d1 = (double)in[base + 1], d2 = (double)in[base + 2], d3 = (double)in[base + 3];
    double q0 = v1 * c.N0 + v1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q1 = d1 * c.N0 + v1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q2 = d2 * c.N0 + d1 * c.N1 + v1 * c.N2 + v1 * c.N3;
    double q3 = d3 * c.N0 + d2 * c.N1 + d1 * c.N2 + v1 * c.N3;
    q0 -= v1 * c.BN1 + v1 * c.BN2 + v1 * c.BN3 + v1 * c.BN4;
    q1 -= q0 * c.D1 + v1 * c.BN2 + v1 * c.BN3 + v1 * c.BN4;
    q2 -= q1 * c.D1 + q0 * c.D2 + v1 * c.BN3 + v1 * c.BN4;
    q3 -= q2 * c.D1 + q1 * c.D2 + q0 * c.D3 + v1 * c.BN4;
    out[base] = q0;
    out[base + 1] = q1;
    out[base + 2] = q2;
    out[base + 3] = q3;
    o0 = q3;
    o1 = q2;
    o2 = q1;
    o3 = q0;
    xm1 = d3;
    xm2 = d2;
    xm3 = d1;
  }
  for (int64_t i0 = 4; i0 < ln; i0 += kRecX) {
    const int n = (int)min((int64_t)kRecX, ln - i0);
    __syncthreads();  // the previous stage's tiles are consumed
    rec_stage<TIn>(tx, in, blk, ln, i0, n, nl);
    __syncthreads();
    if (live) {
      for (int k = 0; k < n; ++k) {
        const double xi = rx[k];
        double q = xi * c.N0 + xm1 * c.N1 + xm2 * c.N2 + xm3 * c.N3;
        q -= o0 * c.D1 + o1 * c.D2 + o2 * c.D3 + o3 * c.D4;
        ry[k] = q;
        o3 = o2;
        o2 = o1;
        o1 = o0;
        o0 = q;
        xm3 = xm2;
        xm2 = xm1;
        xm1 = xi;
      }
    }
    __syncthreads();
    for (int r = t; r < kRecLines * kRecX; r += kRecLines) {
      const int j = r >> 3, k = r & 7;
      if (j < nl && k < n) out[blk + (int64_t)j * ln + i0 + k] = ty[j * kRecPitch + k];
    }
  }
  __syncthreads();  // every line's causal values are in place
  // anti-causal
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, xp1 = 0.0, xp2 = 0.0, xp3 = 0.0;
  if (live) {
    const double v2 = (double)in[base + ln - 1];
    const double e1 = (double)in[base + ln - 1], e2 = (double)in[base + ln - 2], e3 = (double)in[base + ln - 3];
    double r1 = v2 * c.M1 + v2 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r2 = e1 * c.M1 + v2 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r3 = e2 * c.M1 + e1 * c.M2 + v2 * c.M3 + v2 * c.M4;
    double r4 = e3 * c.M1 + e2 * c.M2 + e1 * c.M3 + v2 * c.M4;
    r1 -= v2 * c.BM1 + v2 * c.BM2 + v2 * c.BM3 + v2 * c.BM4;
    r2 -= r1 * c.D1 + v2 * c.BM2 + v2 * c.BM3 + v2 * c.BM4;
    r3 -= r2 * c.D1 + r1 * c.D2 + v2 * c.BM3 + v2 * c.BM4;
    r4 -= r3 * c.D1 + r2 * c.D2 + r1 * c.D3 + v2 * c.BM4;
    const double y1 = out[base + ln - 1], y2 = out[base + ln - 2];
    const double y3 = out[base + ln - 3], y4 = out[base + ln - 4];
    rec_sink_store(sink, out, base + ln - 1, y1 + r1);
    rec_sink_store(sink, out, base + ln - 2, y2 + r2);
    rec_sink_store(sink, out, base + ln - 3, y3 + r3);
    rec_sink_store(sink, out, base + ln - 4, y4 + r4);
    s0 = r4;
    s1 = r3;
    s2 = r2;
    s3 = r1;
    xp1 = e3;
    xp2 = e2;
    xp3 = e1;
  }
  for (int64_t i0 = ln - 4; i0 > 0; i0 -= kRecX) {
    // i runs i0, i0 - 1, ..., i0 - n + 1 (> 0): it reads data[i] and the
    // causal value at i - 1 and finishes element i - 1
    const int n = (int)min((int64_t)kRecX, i0);
    const int64_t lo = i0 - n;
    __syncthreads();
    rec_stage<TIn>(tx, in, blk, ln, lo + 1, n, nl);
    rec_stage<double>(ty, out, blk, ln, lo, n, nl);
    __syncthreads();
    if (live) {
      for (int k = 0; k < n; ++k) {
        const double xi = rx[n - 1 - k];
        double q = xi * c.M1 + xp1 * c.M2 + xp2 * c.M3 + xp3 * c.M4;
        q -= s0 * c.D1 + s1 * c.D2 + s2 * c.D3 + s3 * c.D4;
        ry[n - 1 - k] = ry[n - 1 - k] + q;
        s3 = s2;
        s2 = s1;
        s1 = s0;
        s0 = q;
        xp3 = xp2;
        xp2 = xp1;
        xp1 = xi;
      }
    }
    __syncthreads();
    for (int r = t; r < kRecLines * kRecX; r += kRecLines) {
      const int j = r >> 3, k = r & 7;
      if (j < nl && k < n) rec_sink_store(sink, out, blk + (int64_t)j * ln + lo + k, ty[j * kRecPitch + k]);
    }
  }
}


// The same pass of up to kRecJobsMax independent chains in one launch: the
// workgroups are the jobs' in order, so a pass with few lines per chain still
// fills the device.
template <typename TIn>
__global__ void __launch_bounds__(kRecLines) recursive_gaussian_jobs_kernel(RecJobs J) {
  // the workgroup's job, with constant indices into the argument (an index
  // held in a register would put the whole argument block in scratch)
  static_assert(kRecJobsMax == 4, "one test per job boundary below");
  uint32_t b = blockIdx.x;
  int k = 0;
  if (J.n > 1 && b >= (uint32_t)J.j[0].blocks) {
    b -= (uint32_t)J.j[0].blocks;
    k = 1;
    if (J.n > 2 && b >= (uint32_t)J.j[1].blocks) {
      b -= (uint32_t)J.j[1].blocks;
      k = 2;
      if (J.n > 3 && b >= (uint32_t)J.j[2].blocks) {
        b -= (uint32_t)J.j[2].blocks;
        k = 3;
      }
    }
  }
  const int64_t l0 = (int64_t)b * kRecLines;
  extern __shared__ double rec_lds[];  // kRecRowsLdsBytes when the launch holds a pass along x
#define MMI_REC_JOB(K)                                                                                          \
  if (J.j[K].stride == 1)                                                                                        \
    recursive_gaussian_rows<TIn>(static_cast<const TIn*>(J.j[K].in), J.j[K].out, J.j[K].ln, J.j[K].nlines,       \
                                 J.j[K].c, J.j[K].sink, l0, rec_lds, rec_lds + kRecLines * kRecPitch);          \
  else                                                                                                           \
    recursive_gaussian_line<TIn>(static_cast<const TIn*>(J.j[K].in), J.j[K].out, J.j[K].ln, J.j[K].stride,      \
                                 J.j[K].nlines, J.j[K].c, J.j[K].sink, l0 + threadIdx.x)
  // a workgroup-uniform branch per job keeps the arguments in scalar registers
  if (k == 0)
    MMI_REC_JOB(0);
  else if (k == 1)
    MMI_REC_JOB(1);
  else if (k == 2)
    MMI_REC_JOB(2);
  else
    MMI_REC_JOB(3);
#undef MMI_REC_JOB
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

// FP32 -> BF16 / FP16, round to nearest even (NaN stays NaN; FP16 overflows to
// infinity, which the host rejects through the intensity range).
__device__ __forceinline__ uint32_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (u >> 16) | 0x40u;
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ uint32_t to_f16(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

template <bool BF16>
__global__ void to_storage_kernel(const float* __restrict__ in, uint16_t* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint16_t)(BF16 ? to_bf16(in[i]) : to_f16(in[i]));
}

template <bool BF16>
__global__ void texels_to_storage_kernel(const float4* __restrict__ in, uint2* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 t = in[i];
  uint2 u;
  if (BF16) {
    u.x = to_bf16(t.x) | (to_bf16(t.y) << 16);
    u.y = to_bf16(t.z) | (to_bf16(t.w) << 16);
  } else {
    u.x = to_f16(t.x) | (to_f16(t.y) << 16);
    u.y = to_f16(t.z) | (to_f16(t.w) << 16);
  }
  out[i] = u;
}

// Coefficient lines of the B-spline fast path in two contractions, each
// summing in the order the per-sample displacement would (over the y support
// inside, then over the z support):
//   t1[gz][y][d][a]  = sum_b wy[y][b] coef[d][gz][sy(y) + b][a]
//   rows[z][y][d][a] = sum_c wz[z][c] t1[sz(z) + c][y][d][a]
// In 2-D the first one writes the rows themselves (Gz = 1).  Zero where the
// row's support is outside the transform's valid region.
__global__ void __launch_bounds__(256) bspline_rows_y_kernel(const double* __restrict__ coeffs, int32_t pdim, int Gx,
                                                             int Gy, int Ny, int per, uint32_t total, AxisTable ty,
                                                             double* __restrict__ out) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  const uint32_t r = e / (uint32_t)per;
  const int ea = (int)(e - r * (uint32_t)per);
  const int y = (int)(r % (uint32_t)Ny);
  const int gz = (int)(r / (uint32_t)Ny);
  const int d = ea / Gx;
  const int a = ea - d * Gx;
  double inner = 0.0;
  if (ty.inside[y]) {
    const double* c = coeffs + (int64_t)d * pdim + a + (int64_t)Gx * (ty.start[y] + (int64_t)Gy * gz);
    double v[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) v[b] = c[Gx * b];
#pragma unroll
    for (int b = 0; b < 4; ++b) inner += ty.w[y * 4 + b] * v[b];
  }
  out[e] = inner;
}

__global__ void __launch_bounds__(256) bspline_rows_z_kernel(const double* __restrict__ t1, int Ny, int per,
                                                             uint32_t total, AxisTable ty, AxisTable tz,
                                                             double* __restrict__ rows) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  const uint32_t r = e / (uint32_t)per;
  const int ea = (int)(e - r * (uint32_t)per);
  const int y = (int)(r % (uint32_t)Ny);
  const int z = (int)(r / (uint32_t)Ny);
  double acc = 0.0;
  if (ty.inside[y] && tz.inside[z]) {
    const int64_t plane = (int64_t)Ny * per;
    const double* t = t1 + ((int64_t)tz.start[z] * Ny + y) * per + ea;
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = t[c * plane];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc += tz.w[z * 4 + c] * v[c];
  }
  rows[e] = acc;
}

// Per-axis B-spline support of the virtual samples for a separable geometry.
// The grid continuous index along `axis` is computed from the full virtual
// point with the reference's expressions; the other axes' terms are exact
// zeros there, so it equals the per-sample value for every sample with this
// index along the axis.
__global__ void bspline_axis_table_kernel(DevGeom virt, DevGeom grid, int axis, int32_t* __restrict__ start,
                                          double* __restrict__ w, uint8_t* __restrict__ inside) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= virt.size[axis]) return;
  int32_t idx[3] = {virt.start[0], virt.start[1], virt.start[2]};
  idx[axis] += i;
  double p[3], ci[3];
  if (virt.dim == 3) {
    index_to_point<3>(virt, idx, p);
    point_to_cindex<3>(grid, p, ci);
  } else {
    index_to_point<2>(virt, idx, p);
    point_to_cindex<2>(grid, p, ci);
  }
  double c = ci[axis];
  const bool in = bspline_axis_inside(grid, axis, c);
  inside[i] = in ? 1 : 0;
  double w4[4] = {0.0, 0.0, 0.0, 0.0};
  start[i] = in ? bspline_axis_weights(c, w4) : 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) w[i * 4 + k] = w4[k];
}

// One axis of the tensor-product contraction of the B-spline derivative:
// out[o][g][i] = sum over virtual indices l in [lo[g], hi[g]] of
// w[l][g - start[l]] * in[o][l][i].  The support start is monotone in l and the
// valid indices along an axis are one interval, so every l of the range holds
// g in its support.  Every output element is one thread summing in index
// order, so the result does not depend on scheduling.
template <typename T>
__global__ void __launch_bounds__(256) separable_contract_kernel(const T* __restrict__ in, double* __restrict__ out,
                                                                 int64_t total, int L, int64_t inner, int G,
                                                                 AxisTable t, const DevResult* __restrict__ scale) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int64_t i = e % inner;
  const int64_t r = e / inner;
  const int g = (int)(r % G);
  const int64_t o = r / G;
  const T* src = in + o * L * inner + i;
  double acc = 0.0;
  const int hi = t.hi[g];
#pragma unroll 4
  for (int l = t.lo[g]; l <= hi; ++l) acc += t.w[l * 4 + (g - t.start[l])] * (double)src[(int64_t)l * inner];
  if (scale) acc *= scale->n_factor;
  out[e] = acc;
}

// The first (x) axis of the contraction, one thread per row: the row's
// samples stream in index order through a window of the four nodes the
// current support covers; a node is written once the window has moved past
// it.  Every node receives its terms in increasing index order starting from
// zero, as in separable_contract_kernel, so both give the same bits.  Needs
// a nondecreasing support start along the axis.
__global__ void __launch_bounds__(256) separable_rows_kernel(const float* __restrict__ in, double* __restrict__ out,
                                                             int64_t rows, int L, int G, AxisTable t) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const float* x = in + r * L;
  double* o = out + r * G;
  int base = -1, written = 0;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int l = 0; l < L; ++l) {
    if (!t.inside[l]) continue;
    const int s = t.start[l];
    if (base < 0) base = s;
    while (base < s) {
      for (; written < base; ++written) o[written] = 0.0;
      o[base] = a0;
      written = base + 1;
      a0 = a1;
      a1 = a2;
      a2 = a3;
      a3 = 0.0;
      ++base;
    }
    const double v = (double)x[l];
    const double* w = t.w + 4 * l;
    a0 += w[0] * v;
    a1 += w[1] * v;
    a2 += w[2] * v;
    a3 += w[3] * v;
  }
  if (base >= 0) {
    for (; written < base; ++written) o[written] = 0.0;
    o[base] = a0;
    o[base + 1] = a1;
    o[base + 2] = a2;
    o[base + 3] = a3;
    written = base + 4;
  }
  for (; written < G; ++written) o[written] = 0.0;
}

// The first (x) axis when its table and a block of rows fit in LDS: a
// workgroup stages the axis table (support starts, weights, node ranges) once
// and then `nr` rows at a time with coalesced loads (rows padded by one float
// against bank conflicts); each output node of the block is one thread summing
// w[l][g - start[l]] * x[l] over l in [lo[g], hi[g]] in increasing l, the
// order of separable_contract_kernel, so both give the same bits.  The
// outputs of a block are contiguous, so the stores coalesce.
constexpr int kRowsLdsBytes = 64 * 1024;

// LDS of separable_rows_lds_kernel: weights, starts and node ranges, the
// output block, the padded row tile.
__host__ __device__ inline size_t rows_lds_bytes(int L, int G, int nr) {
  return sizeof(double) * ((size_t)4 * L + (L + 2 * G + 1) / 2 + (size_t)nr * G) + sizeof(float) * (size_t)nr * (L + 1);
}

__global__ void __launch_bounds__(256) separable_rows_lds_kernel(const float* __restrict__ in,
                                                                 double* __restrict__ out, int64_t rows, int L, int G,
                                                                 int nr, AxisTable t) {
  extern __shared__ double lds_d[];
  double* wt = lds_d;                                      // [L][4]
  int32_t* st = reinterpret_cast<int32_t*>(wt + 4 * L);   // [L]
  int32_t* lo = st + L;                                    // [G]
  int32_t* hi = lo + G;                                    // [G]
  double* ob = lds_d + 4 * L + (L + 2 * G + 1) / 2;         // [nr][G] outputs of a block
  float* tile = reinterpret_cast<float*>(ob + nr * G);     // [nr][L + 1]
  for (int i = threadIdx.x; i < 4 * L; i += 256) wt[i] = t.w[i];
  for (int i = threadIdx.x; i < L; i += 256) st[i] = t.start[i];
  for (int i = threadIdx.x; i < G; i += 256) {
    lo[i] = t.lo[i];
    hi[i] = t.hi[i];
  }
  const int ld = L + 1;
  for (int64_t r0 = (int64_t)blockIdx.x * nr; r0 < rows; r0 += (int64_t)gridDim.x * nr) {
    const int n = (int)min((int64_t)nr, rows - r0);
    __syncthreads();  // the tables are staged; the previous block is done with the tile
    const float* src = in + r0 * L;
    const int cnt = n * L;
    // eight loads in flight per thread; element i of the block goes to row
    // i / L of the padded tile, i.e. to tile[i + i / L]
    for (int i0 = 0; i0 < cnt; i0 += 8 * 256) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = i0 + k * 256 + (int)threadIdx.x;
        v[k] = i < cnt ? src[i] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = i0 + k * 256 + (int)threadIdx.x;
        if (i < cnt) tile[i + i / L] = v[k];
      }
    }
    __syncthreads();
    // rows fastest: a wavefront shares one or two nodes, so its table reads
    // are broadcasts and its sample reads fall on distinct banks
    const int outs = n * G;
    for (int e = threadIdx.x; e < outs; e += 256) {
      const int g = e / n;
      const int row = e - g * n;
      const float* x = tile + row * ld;
      double acc = 0.0;
      const int h = hi[g];
#pragma unroll 4
      for (int l = lo[g]; l <= h; ++l) acc += wt[l * 4 + (g - st[l])] * (double)x[l];
      ob[row * G + g] = acc;
    }
    __syncthreads();
    double* dst = out + r0 * G;
    for (int e = threadIdx.x; e < outs; e += 256) dst[e] = ob[e];
  }
}

// Intensity range over part of an image (MattesMI::Initialize,
// itkMattesMutualInformationImageToImageMetricv4.hxx:105-220): the voxels
// whose physical points lie inside the mask, or (sparse sampling) the pixels
// the listed virtual points round to after the fixed transform.  Per
// workgroup min and max as minmax_kernel writes them.
template <int D, typename P>
__global__ void __launch_bounds__(256) minmax_select_kernel(const typename P::pixel* __restrict__ px, DevGeom img,
                                                            const DevMask* __restrict__ mask,
                                                            const double* __restrict__ points, DevTransform ftf,
                                                            int64_t n, float* __restrict__ out) {
  __shared__ float smin[256], smax[256];
  float lo = 3.402823466e38f, hi = -3.402823466e38f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t off = 0;
    double p[3] = {0.0, 0.0, 0.0};
    if (points) {
      double vp[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < D; ++k) vp[k] = points[i * D + k];
      fixed_point<D>(ftf, vp, p);
      if (!rounded_index_inside<D>(img, p, &off)) continue;
    } else {
      int32_t idx[3];
      sample_index<D>(img, (uint32_t)i, idx);
      index_to_point<D>(img, idx, p);
      if (!mask_inside<D>(*mask, p)) continue;
      off = (int32_t)i;
    }
    const float v = P::ld(px, off);
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

hipError_t launch_minmax_select(const void* px, int storage, const DevGeom& img, const DevMask* mask,
                                const double* points, const DevTransform& ftf, int64_t n, float* out, int blocks,
                                hipStream_t st) {
#define MMI_MINMAX_SELECT(PX)                                                                                    \
  do {                                                                                                           \
    const auto* p = static_cast<const typename PX::pixel*>(px);                                                  \
    if (img.dim == 2)                                                                                            \
      hipLaunchKernelGGL((minmax_select_kernel<2, PX>), dim3(blocks), dim3(256), 0, st, p, img, mask, points, ftf, \
                         n, out);                                                                                \
    else                                                                                                         \
      hipLaunchKernelGGL((minmax_select_kernel<3, PX>), dim3(blocks), dim3(256), 0, st, p, img, mask, points, ftf, \
                         n, out);                                                                                \
  } while (0)
  if (storage == kStorageBF16)
    MMI_MINMAX_SELECT(PxBF16);
  else if (storage == kStorageF16)
    MMI_MINMAX_SELECT(PxF16);
  else
    MMI_MINMAX_SELECT(PxF32);
#undef MMI_MINMAX_SELECT
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

hipError_t launch_minmax(const void* px, int storage, int64_t n, float* out, int blocks, hipStream_t st) {
  if (storage == kStorageBF16)
    hipLaunchKernelGGL(minmax_kernel<PxBF16>, dim3(blocks), dim3(256), 0, st, static_cast<const uint16_t*>(px), n,
                       out);
  else if (storage == kStorageF16)
    hipLaunchKernelGGL(minmax_kernel<PxF16>, dim3(blocks), dim3(256), 0, st, static_cast<const uint16_t*>(px), n, out);
  else
    hipLaunchKernelGGL(minmax_kernel<PxF32>, dim3(blocks), dim3(256), 0, st, static_cast<const float*>(px), n, out);
  return hipGetLastError();
}

hipError_t launch_to_storage(const float* in, void* out, int storage, int64_t n, hipStream_t st) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (storage == kStorageBF16)
    hipLaunchKernelGGL(to_storage_kernel<true>, grid, dim3(256), 0, st, in, static_cast<uint16_t*>(out), n);
  else if (storage == kStorageF16)
    hipLaunchKernelGGL(to_storage_kernel<false>, grid, dim3(256), 0, st, in, static_cast<uint16_t*>(out), n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_texels_to_storage(const float4* in, void* out, int storage, int64_t n, hipStream_t st) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (storage == kStorageBF16)
    hipLaunchKernelGGL(texels_to_storage_kernel<true>, grid, dim3(256), 0, st, in, static_cast<uint2*>(out), n);
  else if (storage == kStorageF16)
    hipLaunchKernelGGL(texels_to_storage_kernel<false>, grid, dim3(256), 0, st, in, static_cast<uint2*>(out), n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_recursive_gaussian_jobs(const RecJobs& jobs, bool in_float, hipStream_t st) {
  if (jobs.n < 1 || jobs.n > kRecJobsMax) return hipErrorInvalidValue;
  unsigned blocks = 0;
  for (int k = 0; k < jobs.n; ++k) {
    if (jobs.j[k].blocks != (int32_t)((jobs.j[k].nlines + kRecLines - 1) / kRecLines)) return hipErrorInvalidValue;
    blocks += (unsigned)jobs.j[k].blocks;
  }
  size_t lds = 0;
  for (int k = 0; k < jobs.n; ++k)
    if (jobs.j[k].stride == 1) lds = kRecRowsLdsBytes;
  if (in_float)
    hipLaunchKernelGGL(recursive_gaussian_jobs_kernel<float>, dim3(blocks), dim3(kRecLines), lds, st, jobs);
  else
    hipLaunchKernelGGL(recursive_gaussian_jobs_kernel<double>, dim3(blocks), dim3(kRecLines), lds, st, jobs);
  return hipGetLastError();
}

hipError_t launch_pack_texels(const float* moving, const double* grad, int64_t n, int D, const Direction& dir,
                              float4* tex, hipStream_t st) {
  hipLaunchKernelGGL(pack_texels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, moving, grad, n, D, dir,
                     tex);
  return hipGetLastError();
}

// A plain copy, 16 bytes per lane and grid-stride (the copy-rate probe).
__global__ void __launch_bounds__(256) copy_probe_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                         int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

hipError_t launch_copy_probe(const void* in, void* out, int64_t n16, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(copy_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<const uint4*>(in),
                     static_cast<uint4*>(out), n16);
  return hipGetLastError();
}

// A copy of an image buffer (nx x ny x nz elements, x fastest) with a
// one-element border on each side of every axis (z only in 3-D), each border
// element a copy of its nearest buffer element.  A linear interpolation on the
// copy at any cell whose lower corner lies in [start - 1, end] gives the bits
// the reference's branch rules give in the buffer: a corner the reference
// drops is duplicated (lerp(a, a, t) == a), and the border holds exactly that
// duplicate.
template <typename T>
__global__ void pad_replicate_kernel(const T* __restrict__ src, T* __restrict__ dst, int nx, int ny, int nz, int D) {
  const int px = nx + 2, py = ny + 2, pz = D == 3 ? nz + 2 : 1;
  const int64_t n = (int64_t)px * py * pz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % px);
    const int64_t r = i / px;
    const int y = (int)(r % py);
    const int z = (int)(r / py);
    const int sx = min(max(x - 1, 0), nx - 1), sy = min(max(y - 1, 0), ny - 1);
    const int sz = D == 3 ? min(max(z - 1, 0), nz - 1) : 0;
    dst[i] = src[sx + (int64_t)nx * (sy + (int64_t)ny * sz)];
  }
}

hipError_t launch_pad_replicate(const void* src, void* dst, int elem_bytes, int nx, int ny, int nz, int D,
                                hipStream_t st) {
  const int64_t n = (int64_t)(nx + 2) * (ny + 2) * (D == 3 ? nz + 2 : 1);
  const dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 65536));
  switch (elem_bytes) {
    case 2:
      hipLaunchKernelGGL(pad_replicate_kernel<uint16_t>, grid, dim3(256), 0, st, static_cast<const uint16_t*>(src),
                         static_cast<uint16_t*>(dst), nx, ny, nz, D);
      break;
    case 4:
      hipLaunchKernelGGL(pad_replicate_kernel<uint32_t>, grid, dim3(256), 0, st, static_cast<const uint32_t*>(src),
                         static_cast<uint32_t*>(dst), nx, ny, nz, D);
      break;
    case 8:
      hipLaunchKernelGGL(pad_replicate_kernel<uint2>, grid, dim3(256), 0, st, static_cast<const uint2*>(src),
                         static_cast<uint2*>(dst), nx, ny, nz, D);
      break;
    case 16:
      hipLaunchKernelGGL(pad_replicate_kernel<uint4>, grid, dim3(256), 0, st, static_cast<const uint4*>(src),
                         static_cast<uint4*>(dst), nx, ny, nz, D);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The accumulators an evaluation starts from: joint histogram, valid count,
// exception-overflow flag (one launch instead of three memsets).
__global__ void clear_eval_kernel(unsigned long long* __restrict__ hist, int n_hist,
                                  unsigned long long* __restrict__ nvalid, uint32_t* __restrict__ overflow) {
  for (int i = threadIdx.x; i < n_hist; i += blockDim.x) hist[i] = 0ull;
  if (threadIdx.x == 0) {
    *nvalid = 0ull;
    *overflow = 0u;
  }
}

hipError_t launch_clear_eval(unsigned long long* hist, int n_hist, unsigned long long* nvalid, uint32_t* overflow,
                             hipStream_t st) {
  hipLaunchKernelGGL(clear_eval_kernel, dim3(1), dim3(1024), 0, st, hist, n_hist, nvalid, overflow);
  return hipGetLastError();
}

hipError_t launch_bspline_rows(const double* coeffs, int32_t pdim, int Gx, int Gy, int Gz, int D, int Ny, int Nz,
                               const AxisTable* tabs, double* t1, double* rows, hipStream_t st) {
  const int per = D * Gx;
  if (D == 2) {
    const uint32_t total = (uint32_t)((int64_t)Ny * per);
    hipLaunchKernelGGL(bspline_rows_y_kernel, dim3((total + 255) / 256), dim3(256), 0, st, coeffs, pdim, Gx, Gy, Ny,
                       per, total, tabs[1], rows);
    return hipGetLastError();
  }
  const uint32_t t1n = (uint32_t)((int64_t)Gz * Ny * per);
  hipLaunchKernelGGL(bspline_rows_y_kernel, dim3((t1n + 255) / 256), dim3(256), 0, st, coeffs, pdim, Gx, Gy, Ny, per,
                     t1n, tabs[1], t1);
  const uint32_t total = (uint32_t)((int64_t)Nz * Ny * per);
  hipLaunchKernelGGL(bspline_rows_z_kernel, dim3((total + 255) / 256), dim3(256), 0, st, t1, Ny, per, total, tabs[1],
                     tabs[2], rows);
  return hipGetLastError();
}

hipError_t launch_bspline_axis_table(const DevGeom& virt, const DevGeom& grid, int axis, int32_t* start, double* w,
                                     uint8_t* inside, hipStream_t st) {
  const int n = virt.size[axis];
  hipLaunchKernelGGL(bspline_axis_table_kernel, dim3((n + 127) / 128), dim3(128), 0, st, virt, grid, axis, start, w,
                     inside);
  return hipGetLastError();
}

hipError_t launch_separable_contract(const float* in_f, const double* in_d, double* out, int64_t outer, int L,
                                     int64_t inner, int G, const AxisTable& t, const DevResult* scale, hipStream_t st) {
  const int64_t total = outer * G * inner;
  if (total <= 0) return hipSuccess;
  const dim3 grid((unsigned)((total + 255) / 256));
  int nr = (int)std::min<int64_t>(32, outer);
  while (nr > 8 && rows_lds_bytes(L, G, nr) > (size_t)kRowsLdsBytes) nr /= 2;
  if (in_f && inner == 1 && !scale && rows_lds_bytes(L, G, nr) <= (size_t)kRowsLdsBytes) {
    const int64_t blocks = std::min<int64_t>((outer + nr - 1) / nr, 2048);
    hipLaunchKernelGGL(separable_rows_lds_kernel, dim3((unsigned)blocks), dim3(256), rows_lds_bytes(L, G, nr), st,
                       in_f, out, outer, L, G, nr, t);
  } else if (in_f && inner == 1 && !scale && t.increasing) {
    hipLaunchKernelGGL(separable_rows_kernel, dim3((unsigned)((outer + 255) / 256)), dim3(256), 0, st, in_f, out,
                       outer, L, G, t);
  } else if (in_f)
    hipLaunchKernelGGL(separable_contract_kernel<float>, grid, dim3(256), 0, st, in_f, out, total, L, inner, G, t,
                       scale);
  else
    hipLaunchKernelGGL(separable_contract_kernel<double>, grid, dim3(256), 0, st, in_d, out, total, L, inner, G, t,
                       scale);
  return hipGetLastError();
}

}  // namespace mmi
