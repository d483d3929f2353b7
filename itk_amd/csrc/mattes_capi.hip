// C ABI of the MI355X Mattes MI v4 metric: device context, image and
// transform uploads, MattesMutualInformationImageToImageMetricv4::Initialize
// and GetValue / GetValueAndDerivative on the device.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/itk_mattes_hip.h"
#include "mattes_affine.h"
#include "mattes_kernels.h"
#include "mattes_optim.h"

using namespace mmi;

namespace {

thread_local std::string g_last_error = "";

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace

// The last-error string of the calling thread, for the other translation units
// of the library (host-side sampling).
int mmi::set_last_error(int code, const char* msg) { return fail(code, msg); }

namespace {

#define MMI_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      return fail(e_ == hipErrorOutOfMemory ? MMI_ERR_MEMORY : MMI_ERR_DEVICE,                        \
                  std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_));                \
    }                                                                                                  \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

DevGeom to_dev(const mmi_geometry& g) {
  DevGeom d;
  std::memset(&d, 0, sizeof(d));
  d.dim = g.dim;
  for (int i = 0; i < 3; ++i) {
    d.size[i] = (int32_t)(i < g.dim ? g.size[i] : 1);
    d.start[i] = (int32_t)(i < g.dim ? g.start[i] : 0);
    d.end[i] = d.start[i] + d.size[i] - 1;
    d.origin[i] = i < g.dim ? g.origin[i] : 0.0;
    d.spacing[i] = i < g.dim ? g.spacing[i] : 1.0;
    // ImageFunction::SetInputImage bounds (itkImageFunction.hxx:57-64)
    d.cilo[i] = (double)d.start[i] - 0.5;
    d.cihi[i] = (double)d.end[i] + 0.5;
  }
  for (int i = 0; i < 9; ++i) {
    d.i2p[i] = g.index_to_physical[i];
    d.p2i[i] = g.physical_to_index[i];
  }
  d.sy = d.size[0];
  d.sz = d.size[0] * d.size[1];
  d.div_x = make_fastdiv((uint32_t)d.size[0]);
  d.div_y = make_fastdiv((uint32_t)d.size[1]);
  return d;
}

int64_t num_pixels(const mmi_geometry& g) {
  int64_t n = 1;
  for (int i = 0; i < g.dim; ++i) n *= g.size[i];
  return n;
}

int check_geometry(const mmi_geometry* g) {
  if (!g) return fail(MMI_ERR_ARGUMENT, "geometry is NULL");
  if (g->dim != 2 && g->dim != 3) return fail(MMI_ERR_ARGUMENT, "only 2-D and 3-D images are supported");
  int64_t n = 1;
  for (int i = 0; i < g->dim; ++i) {
    if (g->size[i] < 1) return fail(MMI_ERR_ARGUMENT, "image size must be positive");
    if (g->start[i] < -(1ll << 29) || g->start[i] > (1ll << 29))
      return fail(MMI_ERR_ARGUMENT, "image start index out of range");
    n *= g->size[i];
    if (n >= (1ll << 31)) return fail(MMI_ERR_ARGUMENT, "images must hold fewer than 2^31 pixels");
  }
  return MMI_OK;
}

// RecursiveGaussianImageFilter::SetUp for orders 0 and 1
// (Modules/Filtering/Smoothing/include/itkRecursiveGaussianImageFilter.hxx:56-300).
// The exponential-series parameters are the published table of Farneback &
// Westin (2006), identical to the reference's.
bool rec_setup(double sigma, double spacing, int order, bool normalizeAcrossScale, RecCoeffs* c) {
  double direction = 1.0;
  if (spacing < 0.0) {
    direction = -1.0;
    spacing *= -1;
  }
  if (spacing < 1e-8 || !(sigma > 0.0)) return false;
  const double sigmad = sigma / spacing;
  double across = 1.0;
  const double A1[2] = {1.3530, -0.6724};
  const double B1[2] = {1.8151, -3.4327};
  const double W1 = 0.6681, L1 = -1.3932;
  const double A2[2] = {-0.3531, 0.6724};
  const double B2[2] = {0.0902, 0.6100};
  const double W2 = 2.0787, L2 = -1.3732;
  const double Cos1 = std::cos(W1 / sigmad), Cos2 = std::cos(W2 / sigmad);
  const double Sin1 = std::sin(W1 / sigmad), Sin2 = std::sin(W2 / sigmad);
  const double Exp1 = std::exp(L1 / sigmad), Exp2 = std::exp(L2 / sigmad);
  c->D4 = Exp1 * Exp1 * Exp2 * Exp2;
  c->D3 = -2 * Cos1 * Exp1 * Exp2 * Exp2;
  c->D3 += -2 * Cos2 * Exp2 * Exp1 * Exp1;
  c->D2 = 4 * Cos2 * Cos1 * Exp1 * Exp2;
  c->D2 += Exp1 * Exp1 + Exp2 * Exp2;
  c->D1 = -2 * (Exp2 * Cos2 + Exp1 * Cos1);
  const double SD = 1.0 + c->D1 + c->D2 + c->D3 + c->D4;
  const double DD = c->D1 + 2 * c->D2 + 3 * c->D3 + 4 * c->D4;
  const int o = order;
  const double a1 = A1[o], b1 = B1[o], a2 = A2[o], b2 = B2[o];
  c->N0 = a1 + a2;
  c->N1 = Exp2 * (b2 * Sin2 - (a2 + 2 * a1) * Cos2);
  c->N1 += Exp1 * (b1 * Sin1 - (a1 + 2 * a2) * Cos1);
  c->N2 = (a1 + a2) * Cos2 * Cos1;
  c->N2 -= b1 * Cos2 * Sin1 + b2 * Cos1 * Sin2;
  c->N2 *= 2 * Exp1 * Exp2;
  c->N2 += a2 * Exp1 * Exp1 + a1 * Exp2 * Exp2;
  c->N3 = Exp2 * Exp1 * Exp1 * (b2 * Sin2 - a2 * Cos2);
  c->N3 += Exp1 * Exp2 * Exp2 * (b1 * Sin1 - a1 * Cos1);
  const double SN = c->N0 + c->N1 + c->N2 + c->N3;
  const double DN = c->N1 + 2 * c->N2 + 3 * c->N3;
  bool symmetric;
  if (order == 0) {
    const double alpha0 = 2 * SN / SD - c->N0;
    c->N0 *= across / alpha0;
    c->N1 *= across / alpha0;
    c->N2 *= across / alpha0;
    c->N3 *= across / alpha0;
    symmetric = true;
  } else {
    if (normalizeAcrossScale) across = sigma;
    double alpha1 = 2 * (SN * DD - DN * SD) / (SD * SD);
    alpha1 *= direction;
    c->N0 *= across / alpha1;
    c->N1 *= across / alpha1;
    c->N2 *= across / alpha1;
    c->N3 *= across / alpha1;
    symmetric = false;
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
  return true;
}

// Inverse of a small matrix: exact reciprocal for axis-permutation matrices
// (what ITK's SVD pseudo-inverse gives there), Gauss-Jordan otherwise.
void invert(const double* m, int D, double* inv) {
  bool perm = true;
  for (int i = 0; i < D && perm; ++i) {
    int nz = 0;
    for (int j = 0; j < D; ++j) nz += m[i * D + j] != 0.0;
    perm = nz == 1;
  }
  for (int j = 0; j < D && perm; ++j) {
    int nz = 0;
    for (int i = 0; i < D; ++i) nz += m[i * D + j] != 0.0;
    perm = nz == 1;
  }
  if (perm) {
    for (int i = 0; i < D * D; ++i) inv[i] = 0.0;
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j)
        if (m[i * D + j] != 0.0) inv[j * D + i] = 1.0 / m[i * D + j];
    return;
  }
  double a[3][6];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < 2 * D; ++j) a[i][j] = j < D ? m[i * D + j] : (j - D == i ? 1.0 : 0.0);
  for (int c = 0; c < D; ++c) {
    int p = c;
    for (int r = c + 1; r < D; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    for (int j = 0; j < 2 * D; ++j) std::swap(a[c][j], a[p][j]);
    const double piv = a[c][c];
    for (int j = 0; j < 2 * D; ++j) a[c][j] /= piv;
    for (int r = 0; r < D; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int j = 0; j < 2 * D; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) inv[i * D + j] = a[i][j + D];
}

struct HostTransform {
  int32_t kind = MMI_TF_IDENTITY;
  std::vector<double> fixed;
  std::vector<double> params;
};

}  // namespace

struct mmi_context {
  int device = 0;
  uint32_t flags = 0;
  hipStream_t stream = nullptr;
  int cu_count = 256;
  int32_t bins = 50;
  int32_t gradient_source = MMI_GRADIENT_IMAGE;
  int32_t storage = MMI_F32;

  bool have_fixed = false, have_moving = false, have_virtual = false, have_texels = false;
  bool user_texels = false;  // gradient given by mmi_set_moving_gradient (not recomputed per level)
  bool texel_values_level = false;  // the texels' value channel is the level's moving image
  mmi_geometry fixed_g{}, moving_g{}, virtual_g{};
  DevBuf fixed_px, moving_px, texels;  // FP32 uploads and FP32 gradient texels
  DevBuf fixed_lp, moving_lp, texels_lp;  // 16-bit storage copies (MMI_BF16 / MMI_F16)
  int32_t active_storage = MMI_F32;       // storage the evaluation reads, fixed at mmi_initialize
  // per-level smoothing (physical sigma per axis, 0 = none) and the smoothed images
  double fixed_sigma[3] = {0, 0, 0}, moving_sigma[3] = {0, 0, 0};
  double texel_sigma[3] = {0, 0, 0};  // moving smoothing the filter texels were computed from
  DevBuf fixed_work, moving_work;
  bool fixed_smoothed = false, moving_smoothed = false;

  bool initialized = false;
  double fixed_min = 0, fixed_max = 0, moving_min = 0, moving_max = 0;
  double fixed_bin_size = 0, fixed_norm_min = 0, moving_bin_size = 0, moving_norm_min = 0;

  HostTransform ftf, mtf;
  DevTransform dftf{}, dmtf{};
  DevBuf coeffs;
  // composite moving transform: matrix-offset stages before the moving
  // transform in the queue (12 doubles each), on the host and the device
  std::vector<double> init_stages;
  DevBuf init_dev;

  DevBuf hist, nvalid, pdf, fm, mm, pratio, pratio_f, rowsum, partials, sample_fbin, sample_mbin, minmax_out;
  // the evaluation's outputs in one block (one copy to the host): the result,
  // the fast path's exception-overflow flag, the derivative (dev_result,
  // dev_overflow, dev_deriv)
  DevBuf out;
  DevBuf exc_count, exc_list;  // fast-path exception lists
  DevBuf exc_res;              // B-spline: exact results of the listed samples
  DevBuf cls;                  // pass 1's byte per sample for pass 2
  // The exact fixed bin of every dense virtual sample (FastArgs::fbin8):
  // allocated by mmi_initialize, filled by the level's first row-walk
  // evaluation (fixed_bins_kernel), valid until the next mmi_initialize.
  DevBuf fbin8;
  bool fbin8_valid = false;
  // The level's moving values and texels with a one-voxel replicated border
  // (pad_replicate_kernel), for the row walks with the fixed bins; built by
  // mmi_initialize (pad_ok), texels only with a gradient image (pad_tex).
  DevBuf mpad, tpad;
  bool pad_ok = false, pad_tex = false;
  DevBuf cd_defer;             // central differences on the fast pass 2: per-wavefront deferred samples
  DevBuf ftab;                               // fast path: per-axis fixed cells (coarse levels)
  // grow-only work space of the preparation filters (no device frees, which
  // would synchronise the whole device, while other contexts run)
  DevBuf work_a, work_b, work_grad;
  DevBuf work_c, work_d, work_e, work_f;  // with work_a, work_b: the FP64 volumes of up to 3 filter chains
  bool ftab_valid = false;
  mmi_geometry ftab_key_v{}, ftab_key_f{};
  // separable B-spline derivative: per-sample factors, contraction stages, axis tables
  DevBuf bs_g, bs_t1, bs_t2, bs_rows, bs_rows_t1;
  DevBuf bs_start[3], bs_w[3], bs_inside[3], bs_lo[3], bs_hi[3];
  int32_t bs_increasing[3] = {0, 0, 0};
  bool bs_tables = false;
  DevGeom bs_key_virt{}, bs_key_grid{};
  int xcd_chunks = 1;                        // fast path: one contiguous sample share per XCD
  int blocks_per_cu = 4;                     // evaluation grid = CUs x blocks_per_cu workgroups of 256
  int rows_walk = 16;                        // affine fast path: rows per walk (0: sample-strided kernels)
  int default_blocks_per_cu = 4, default_rows_walk = 16;  // the library's choice (environment overrides)
  int hist_copies = 0;                       // pass 1's LDS histogram copies (0: by size)
  // launch settings the caller (mmi_set_launch_config, mmi_set_histogram_copies)
  // or the environment chose; the others come from the tuning table
  bool user_bpc = false, user_walk = false, user_copies = false;
  bool env_bpc = false, env_walk = false, env_copies = false;  // set by the environment (kept by the API's -1 / 0)
  int32_t tuned = 0;                         // 1: the last mmi_initialize took a tuning-table row
  int32_t fixed_kind = kPixelFloat, moving_kind = kPixelFloat;  // reference pixel kinds of the uploads
  int max_lds_per_block = 64 * 1024;         // the device's LDS bytes per workgroup
  int grid = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string last_error;  // the last failed evaluation's message ("" after a success)
  // pinned host staging: B-spline coefficients up, result and derivative down
  void* h_coeffs = nullptr;
  size_t h_coeffs_bytes = 0;
  void* h_out = nullptr;
  size_t h_out_bytes = 0;
  // the fast affine evaluation's result block, written by its last kernel:
  // coherent pinned memory and its device address
  void* h_res = nullptr;
  char* h_res_dev = nullptr;
  size_t h_res_bytes = 0;
  float last_total_ms = 0, last_pass1_ms = 0, last_pass2_ms = 0;
  int32_t last_path = 0;  // 0 exact, 1 fast, 2 fast redone exactly
  bool acc_clean = false;  // pass-1 accumulators known to be zero (finalize zeroes them)
  // sparse sampling: virtual-space points (n_points > 0) instead of the voxels
  int64_t n_points = 0;
  int32_t points_dim = 0;
  DevBuf points;     // n_points x points_dim doubles, in spatial (Morton) order
  DevBuf point_ids;  // the caller's index of each stored point
  // image masks [0] fixed, [1] moving: pixels and the DevMask the kernels read
  bool have_mask[2] = {false, false};
  mmi_geometry mask_g[2] = {};
  DevBuf mask_px[2], mask_desc[2];
};

namespace {
// Morton order of points quantised to 2^10 (3-D) or 2^15 (2-D) cells of their
// bounding box; LSD radix sort, stable (equal keys keep the caller's order).
void spatial_order(const double* p, int dim, int64_t n, std::vector<uint32_t>* order) {
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool any = false;
  for (int64_t k = 0; k < n; ++k) {
    bool fin = true;
    for (int d = 0; d < dim; ++d) fin = fin && std::isfinite(p[k * dim + d]);
    if (!fin) continue;
    for (int d = 0; d < dim; ++d) {
      const double v = p[k * dim + d];
      lo[d] = any ? std::min(lo[d], v) : v;
      hi[d] = any ? std::max(hi[d], v) : v;
    }
    any = true;
  }
  const int bits = dim == 3 ? 10 : 15;
  const double cells = (double)((1 << bits) - 1);
  std::vector<uint32_t> key((size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    uint32_t code = 0;
    for (int d = 0; d < dim; ++d) {
      const double v = p[k * dim + d];
      double q = (hi[d] > lo[d] && std::isfinite(v)) ? (v - lo[d]) / (hi[d] - lo[d]) * cells : 0.0;
      q = std::min(std::max(q, 0.0), cells);
      const uint32_t c = (uint32_t)q;
      for (int b = 0; b < bits; ++b) code |= ((c >> b) & 1u) << (b * dim + d);
    }
    key[(size_t)k] = code;
  }
  std::vector<uint32_t> idx((size_t)n), tmp((size_t)n), ktmp((size_t)n);
  for (int64_t k = 0; k < n; ++k) idx[(size_t)k] = (uint32_t)k;
  for (int shift = 0; shift < 32; shift += 8) {
    size_t count[257] = {0};
    for (int64_t k = 0; k < n; ++k) ++count[((key[(size_t)k] >> shift) & 255u) + 1];
    for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
    for (int64_t k = 0; k < n; ++k) {
      const size_t dst = count[(key[(size_t)k] >> shift) & 255u]++;
      tmp[dst] = idx[(size_t)k];
      ktmp[dst] = key[(size_t)k];
    }
    idx.swap(tmp);
    key.swap(ktmp);
  }
  order->swap(idx);
}

// Samples of one evaluation: the listed points, or the virtual voxels.
int64_t sample_count(const mmi_context* ctx) {
  return ctx->n_points > 0 ? ctx->n_points : num_pixels(ctx->virtual_g);
}
}  // namespace

namespace {

// The output block: DevResult, the exception-overflow flag (16 bytes), then
// P derivative values.
static_assert(sizeof(DevResult) % 8 == 0, "derivative alignment");
size_t out_bytes(int64_t P) { return sizeof(DevResult) + 16 + sizeof(double) * (size_t)P; }
// parameters of the largest transform the fast affine derivative writes
// straight to host memory (a 3-D affine map)
constexpr int64_t kMaxOutParams = 12;
DevResult* dev_result(const mmi_context* ctx) { return ctx->out.as<DevResult>(); }
double* dev_deriv(const mmi_context* ctx) {
  return reinterpret_cast<double*>(static_cast<char*>(ctx->out.p) + sizeof(DevResult) + 16);
}

// Device buffers an evaluation writes: the context's own (grown on first use)
// or the static scratch of mmi_evaluate_static.
struct EvalBufs {
  unsigned long long* hist;  // [B * B]
  unsigned long long* nvalid;
  char* out;  // out_bytes(P): DevResult, exception-overflow flag, derivative
  double* pdf;
  double* fm;
  double* mm;
  double* pratio;
  float* pratio_f;
  double* rowsum;
  double* partials;     // [2 * grid * P]
  uint32_t* exc_count;  // [grid]
  uint32_t* exc_list;   // [grid * kExcCap]
  uint8_t* cls;         // [n] pass 1's byte per sample; NULL: the context's, grown when used
  DevResult* result() const { return reinterpret_cast<DevResult*>(out); }
  uint32_t* overflow() const { return reinterpret_cast<uint32_t*>(out + sizeof(DevResult)); }
  double* deriv() const { return reinterpret_cast<double*>(out + sizeof(DevResult) + 16); }
};

// ---- Device scratch of mmi_evaluate_static ---------------------------------
// One file-scope device array per GPU holds every buffer the evaluation
// writes besides the caller's outputs: the joint histogram and counters, the
// finalize tables, the result block, the derivative partials, the exception
// lists and pass 1's byte per sample.  Sized for the limits below
// (kStaticScratchBytes, about 19 MiB).
constexpr int kStaticMaxBins = 128;
constexpr int64_t kStaticMaxSamples = 1ll << 24;  // 256^3 virtual samples
constexpr int kStaticMaxGrid = 2048;               // workgroups per pass
constexpr int kStaticMaxParams = 12;               // 3-D affine
constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kStaticB2 = (size_t)kStaticMaxBins * kStaticMaxBins;
constexpr size_t kOffHist = 0;
constexpr size_t kOffNvalid = kOffHist + al256(kStaticB2 * 8);
constexpr size_t kOffOut = kOffNvalid + 256;
constexpr size_t kOffPdf = kOffOut + al256(sizeof(DevResult) + 16 + 8 * kStaticMaxParams);
constexpr size_t kOffFm = kOffPdf + al256(kStaticB2 * 8);
constexpr size_t kOffMm = kOffFm + al256(kStaticMaxBins * 8);
constexpr size_t kOffRowsum = kOffMm + al256(kStaticMaxBins * 8);
constexpr size_t kOffPratio = kOffRowsum + al256(kStaticMaxBins * 8);
constexpr size_t kOffPratioF = kOffPratio + al256(kStaticB2 * 8);
constexpr size_t kOffPartials = kOffPratioF + al256(kStaticB2 * 4);
constexpr size_t kOffExcCount = kOffPartials + al256((size_t)2 * kStaticMaxGrid * kStaticMaxParams * 8);
constexpr size_t kOffExcList = kOffExcCount + al256((size_t)kStaticMaxGrid * 4);
constexpr size_t kOffCls = kOffExcList + al256((size_t)kStaticMaxGrid * kExcCap * 4);
constexpr size_t kStaticScratchBytes = kOffCls + al256((size_t)kStaticMaxSamples);
constexpr int kMaxDevices = 64;

}  // namespace

// In host code this name is only a handle for hipGetSymbolAddress: kernels
// and copies take the per-device address resolved below.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

// Resolves `scratch` on `device` (current) once; later calls only read.
int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(MMI_ERR_ARGUMENT, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return MMI_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MMI_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  return MMI_OK;
}

// ---------------------------------------------------------------------------
// Launch-configuration table (written by the tuner, itk_amd/tune.py): per
// workload class the blocks per CU, rows per walk and pass-1 histogram copies
// that timed fastest.
// ---------------------------------------------------------------------------

struct TuneRow {
  int32_t kind;      // MMI_TF_AFFINE / MMI_TF_BSPLINE
  int32_t storage;   // MMI_F32 / MMI_BF16 / MMI_F16
  int32_t dim;
  int32_t gradient;  // MMI_GRADIENT_IMAGE / MMI_GRADIENT_CENTRAL_DIFFERENCE
  int32_t bins;      // the row serves contexts with at most this many bins
  int32_t log2n;     // floor(log2(virtual samples)) of the tuned workload
  int32_t bpc, walk, copies;
};

std::mutex g_tuning_mu;
std::vector<TuneRow> g_tuning;
bool g_tuning_loaded = false;

int hist_copies_default(int bins) { return hist_copies(bins); }

// The table shipped beside the library, or $MMI_TUNING_TABLE.
std::string default_tuning_path() {
  if (const char* e = std::getenv("MMI_TUNING_TABLE")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&hist_copies_default), &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    const size_t slash = so.find_last_of('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning/mi355x.tsv";
  }
  return "tuning/mi355x.tsv";
}

int read_tuning_table(const std::string& path, std::vector<TuneRow>& rows) {
  std::ifstream in(path);
  if (!in) return fail(MMI_ERR_ARGUMENT, "cannot open tuning table " + path);
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ss(line);
    std::string kind, storage, gradient;
    TuneRow r{};
    if (!(ss >> kind)) continue;  // blank or comment
    if (!(ss >> storage >> r.dim >> gradient >> r.bins >> r.log2n >> r.bpc >> r.walk >> r.copies))
      return fail(MMI_ERR_ARGUMENT, path + ":" + std::to_string(lineno) + ": expected 9 fields");
    r.kind = kind == "affine" ? MMI_TF_AFFINE : (kind == "bspline" ? MMI_TF_BSPLINE : -1);
    r.storage = storage == "f32" ? MMI_F32 : (storage == "bf16" ? MMI_BF16 : (storage == "f16" ? MMI_F16 : -1));
    r.gradient = gradient == "image" ? MMI_GRADIENT_IMAGE : (gradient == "cd" ? MMI_GRADIENT_CENTRAL_DIFFERENCE : -1);
    if (r.kind < 0 || r.storage < 0 || r.gradient < 0 || (r.dim != 2 && r.dim != 3) || r.bpc < 1 || r.bpc > 64 ||
        r.walk < 0 || r.walk > 256 || (r.copies != 1 && r.copies != 2 && r.copies != 4))
      return fail(MMI_ERR_ARGUMENT, path + ":" + std::to_string(lineno) + ": field out of range");
    rows.push_back(r);
  }
  return MMI_OK;
}

// The table's settings for this context's workload (the smallest bins bucket
// that holds its bins, then the nearest sample-count bucket), for the
// settings the caller left to the library; the defaults otherwise.  The
// table is read once, at the first initialisation.
void apply_tuning(mmi_context* ctx, int64_t n) {
  {
    std::lock_guard<std::mutex> lock(g_tuning_mu);
    if (!g_tuning_loaded) {
      g_tuning_loaded = true;
      std::vector<TuneRow> rows;
      const std::string saved = g_last_error;
      if (read_tuning_table(default_tuning_path(), rows) == MMI_OK) g_tuning = rows;
      g_last_error = saved;  // no table is not an error
    }
  }
  const TuneRow* best = nullptr;
  int log2n = 0;
  while (log2n < 62 && (int64_t(1) << (log2n + 1)) <= n) ++log2n;
  for (const TuneRow& r : g_tuning) {
    if (r.kind != ctx->mtf.kind || r.storage != ctx->storage || r.dim != ctx->fixed_g.dim ||
        r.gradient != ctx->gradient_source || r.bins < ctx->bins)
      continue;
    if (!best || r.bins < best->bins ||
        (r.bins == best->bins && (std::abs(r.log2n - log2n) < std::abs(best->log2n - log2n) ||
                                  (std::abs(r.log2n - log2n) == std::abs(best->log2n - log2n) &&
                                   r.log2n < best->log2n))))
      best = &r;
  }
  ctx->tuned = best ? 1 : 0;
  if (!ctx->user_bpc) ctx->blocks_per_cu = best ? best->bpc : ctx->default_blocks_per_cu;
  if (!ctx->user_walk) ctx->rows_walk = best ? best->walk : ctx->default_rows_walk;
  if (!ctx->user_copies) ctx->hist_copies = best ? best->copies : 0;
}

EvalBufs scratch_bufs(unsigned char* base) {
  EvalBufs b;
  b.hist = reinterpret_cast<unsigned long long*>(base + kOffHist);
  b.nvalid = reinterpret_cast<unsigned long long*>(base + kOffNvalid);
  b.out = reinterpret_cast<char*>(base + kOffOut);
  b.pdf = reinterpret_cast<double*>(base + kOffPdf);
  b.fm = reinterpret_cast<double*>(base + kOffFm);
  b.mm = reinterpret_cast<double*>(base + kOffMm);
  b.pratio = reinterpret_cast<double*>(base + kOffPratio);
  b.pratio_f = reinterpret_cast<float*>(base + kOffPratioF);
  b.rowsum = reinterpret_cast<double*>(base + kOffRowsum);
  b.partials = reinterpret_cast<double*>(base + kOffPartials);
  b.exc_count = reinterpret_cast<uint32_t*>(base + kOffExcCount);
  b.exc_list = reinterpret_cast<uint32_t*>(base + kOffExcList);
  b.cls = reinterpret_cast<uint8_t*>(base + kOffCls);
  return b;
}

// The context's buffers as they are (no growth).
EvalBufs ctx_bufs(mmi_context* ctx) {
  EvalBufs b;
  b.hist = ctx->hist.as<unsigned long long>();
  b.nvalid = ctx->nvalid.as<unsigned long long>();
  b.out = static_cast<char*>(ctx->out.p);
  b.pdf = ctx->pdf.as<double>();
  b.fm = ctx->fm.as<double>();
  b.mm = ctx->mm.as<double>();
  b.pratio = ctx->pratio.as<double>();
  b.pratio_f = ctx->pratio_f.as<float>();
  b.rowsum = ctx->rowsum.as<double>();
  b.partials = ctx->partials.as<double>();
  b.exc_count = ctx->exc_count.as<uint32_t>();
  b.exc_list = ctx->exc_list.as<uint32_t>();
  b.cls = nullptr;
  return b;
}

int ensure_host_pinned(void** p, size_t* have, size_t need) {
  if (*p && *have >= need) return MMI_OK;
  if (*p) {
    MMI_HIP(hipDeviceSynchronize());  // no copy from the old buffer in flight (rare: growth only)
    MMI_HIP(hipHostFree(*p));
    *p = nullptr;
    *have = 0;
  }
  MMI_HIP(hipHostMalloc(p, need ? need : 16, hipHostMallocDefault));
  *have = need;
  return MMI_OK;
}

int build_transform(mmi_context* ctx, int D, HostTransform& h, DevTransform& d, bool moving) {
  std::memset(&d, 0, sizeof(d));
  d.kind = h.kind;
  if (h.kind == MMI_TF_IDENTITY) return MMI_OK;
  if (h.kind == MMI_TF_AFFINE) {
    if ((int)h.params.size() != D * D + D) return fail(MMI_ERR_ARGUMENT, "affine transform needs D*D+D parameters");
    if ((int)h.fixed.size() != D) return fail(MMI_ERR_ARGUMENT, "affine transform needs D fixed parameters (centre)");
    // MatrixOffsetTransformBase::SetParameters + ComputeOffset (the device
    // optimizer loop uses the same function)
    affine_from_parameters(D, h.params.data(), h.fixed.data(), d);
    return MMI_OK;
  }
  if (h.kind == MMI_TF_BSPLINE) {
    if (!moving) return fail(MMI_ERR_ARGUMENT, "the fixed transform must be identity or affine");
    const int nf = (int)h.fixed.size();
    if (nf != 3 * D + D * D && nf != 3 * D + 2 * D * D)
      return fail(MMI_ERR_ARGUMENT, "B-spline fixed parameters: size, origin, spacing, direction [, physical_to_index]");
    mmi_geometry g{};
    g.dim = D;
    int64_t nodes = 1;
    for (int i = 0; i < D; ++i) {
      g.size[i] = (int64_t)std::llround(h.fixed[i]);
      g.start[i] = 0;
      g.origin[i] = h.fixed[D + i];
      g.spacing[i] = h.fixed[2 * D + i];
      if (g.size[i] < 4) return fail(MMI_ERR_ARGUMENT, "B-spline grid needs at least 4 nodes per axis");
      nodes *= g.size[i];
    }
    for (int i = 0; i < D * D; ++i) g.direction[i] = h.fixed[3 * D + i];
    // ImageBase::ComputeIndexToPhysicalPointMatrices: direction * diag(spacing)
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += g.direction[i * D + k] * (k == j ? g.spacing[j] : 0.0);
        g.index_to_physical[i * D + j] = s;
      }
    if (nf == 3 * D + 2 * D * D) {
      for (int i = 0; i < D * D; ++i) g.physical_to_index[i] = h.fixed[3 * D + D * D + i];
    } else {
      invert(g.index_to_physical, D, g.physical_to_index);
    }
    if ((int64_t)h.params.size() != D * nodes) return fail(MMI_ERR_ARGUMENT, "B-spline needs D * prod(grid) parameters");
    d.grid = to_dev(g);
    d.pdim = (int32_t)nodes;
    MMI_HIP(ctx->coeffs.ensure(sizeof(double) * h.params.size()));
    // through pinned memory: an asynchronous copy, ordered before the next
    // evaluation on the context stream
    if (int rc = ensure_host_pinned(&ctx->h_coeffs, &ctx->h_coeffs_bytes, sizeof(double) * h.params.size()))
      return rc;
    std::memcpy(ctx->h_coeffs, h.params.data(), sizeof(double) * h.params.size());
    MMI_HIP(hipMemcpyAsync(ctx->coeffs.p, ctx->h_coeffs, sizeof(double) * h.params.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    d.coeffs = ctx->coeffs.as<double>();
    return MMI_OK;
  }
  return fail(MMI_ERR_ARGUMENT, "unknown transform kind");
}

// Host pixels of the reference's pixel type as the device's FP32 image;
// *kind receives the pixel kind.  Double and 32-bit integer pixels must be
// exact in float (the FP32 image then holds every pixel exactly).
template <typename T>
int to_float_exact(const void* pixels, int64_t n, std::vector<float>& out, const char* what) {
  const T* src = static_cast<const T*>(pixels);
  out.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const float f = (float)src[i];
    if (!((T)f == src[i]))
      return fail(MMI_ERR_ARGUMENT, std::string(what) + " pixel " + std::to_string(i) +
                                        " is not exact in the device's FP32 storage");
    out[(size_t)i] = f;
  }
  return MMI_OK;
}

int upload_image(mmi_context* ctx, DevBuf& buf, const mmi_geometry* g, int32_t pixel_type, const void* pixels,
                 int32_t* kind) {
  if (int rc = check_geometry(g)) return rc;
  if (!pixels) return fail(MMI_ERR_ARGUMENT, "pixels is NULL");
  const int64_t n = num_pixels(*g);
  std::vector<float> tmp;
  int rc = MMI_OK;
  *kind = kPixelInteger;
  switch (pixel_type) {
    case MMI_F32: *kind = kPixelFloat; break;
    case MMI_F64: *kind = kPixelDouble; rc = to_float_exact<double>(pixels, n, tmp, "double"); break;
    case MMI_U8: rc = to_float_exact<uint8_t>(pixels, n, tmp, "uint8"); break;
    case MMI_S8: rc = to_float_exact<int8_t>(pixels, n, tmp, "int8"); break;
    case MMI_U16: rc = to_float_exact<uint16_t>(pixels, n, tmp, "uint16"); break;
    case MMI_S16: rc = to_float_exact<int16_t>(pixels, n, tmp, "int16"); break;
    case MMI_U32: rc = to_float_exact<uint32_t>(pixels, n, tmp, "uint32"); break;
    case MMI_S32: rc = to_float_exact<int32_t>(pixels, n, tmp, "int32"); break;
    default: return fail(MMI_ERR_ARGUMENT, "pixel_type must be MMI_F32, MMI_F64 or an integer type");
  }
  if (rc) return rc;
  MMI_HIP(buf.ensure(sizeof(float) * n));
  MMI_HIP(hipMemcpyAsync(buf.p, pixel_type == MMI_F32 ? pixels : tmp.data(), sizeof(float) * n,
                         hipMemcpyHostToDevice, ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  return MMI_OK;
}

bool is_diagonal(const double* m, int D) {
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      if (i != j && m[i * D + j] != 0.0) return false;
  return true;
}

double itk_floor_host(double x) { return (double)(((int64_t)std::nearbyint(2.0 * x - 0.5)) >> 1); }

// Fixed side of the certified fast path: with an identity fixed transform and
// axis-aligned virtual and fixed grids, the reference's continuous index of
// every virtual sample along axis a depends on that axis' index only.  If, for
// every index, it lands inside the buffer on a pixel centre (or on an axis the
// interpolator skips), the interpolated fixed value is exactly one pixel, at
// a constant index offset.  Checked here with the reference's expressions.
bool fixed_direct_offsets(const mmi_geometry& v, const mmi_geometry& f, int64_t* offset) {
  const int D = v.dim;
  if (!is_diagonal(v.index_to_physical, D) || !is_diagonal(f.physical_to_index, D)) return false;
  for (int a = 0; a < D; ++a) {
    const double s = v.index_to_physical[a * D + a];
    const double q = f.physical_to_index[a * D + a];
    const int64_t fend = f.start[a] + f.size[a] - 1;
    bool have = false;
    for (int64_t i = v.start[a]; i < v.start[a] + v.size[a]; ++i) {
      const double p = v.origin[a] + s * (double)i;
      const double ci = q * (p - f.origin[a]);
      if (!(ci >= (double)f.start[a] - 0.5 && ci < (double)fend + 0.5)) return false;
      int64_t b = (int64_t)itk_floor_host(ci);
      if (b < f.start[a]) b = f.start[a];
      const double d = ci - (double)b;
      const bool active = d > 0.0 && b + 1 <= fend;
      if (active) return false;
      if (!have) {
        offset[a] = b - i;
        have = true;
      } else if (b - i != offset[a]) {
        return false;
      }
    }
  }
  for (int a = D; a < 3; ++a) offset[a] = 0;
  return true;
}

// Per-axis fixed-image cells of the virtual samples (axis-aligned virtual and
// fixed grids, identity fixed transform): for each virtual index along an axis,
// the reference's continuous index (computed with its expressions: the other
// axes' terms are exact zeros), the inside-buffer test of that axis and the
// interpolation cell of LinearInterpolateImageFunction.  Cached on the device
// per geometry pair.
bool fixed_axis_tables(mmi_context* ctx, const mmi_geometry& v, const mmi_geometry& f, FastArgs* F) {
  const int D = v.dim;
  if (!is_diagonal(v.index_to_physical, D) || !is_diagonal(f.physical_to_index, D)) return false;
  const bool same = ctx->ftab_valid && std::memcmp(&ctx->ftab_key_v, &v, sizeof(v)) == 0 &&
                    std::memcmp(&ctx->ftab_key_f, &f, sizeof(f)) == 0;
  int64_t off[4] = {0, 0, 0, 0};
  for (int a = 0; a < D; ++a) off[a + 1] = off[a] + v.size[a];
  if (!same) {
    ctx->ftab_valid = false;
    std::vector<int4> tab((size_t)off[D]);
    int64_t stride = 1;
    for (int a = 0; a < D; ++a) {
      const double s = v.index_to_physical[a * D + a];
      const double q = f.physical_to_index[a * D + a];
      const int64_t fend = f.start[a] + f.size[a] - 1;
      for (int64_t i = 0; i < v.size[a]; ++i) {
        const double p = v.origin[a] + s * (double)(v.start[a] + i);
        const double ci = q * (p - f.origin[a]);
        int4 e = make_int4(0, 0, 0, 0);
        if (ci >= (double)f.start[a] - 0.5 && ci < (double)fend + 0.5) {
          int64_t b = (int64_t)itk_floor_host(ci);
          if (b < f.start[a]) b = f.start[a];
          const double d = ci - (double)b;
          const bool active = d > 0.0 && b + 1 <= fend;
          const float df = (float)d;
          int32_t bits;
          std::memcpy(&bits, &df, 4);
          e = make_int4((int32_t)((b - f.start[a]) * stride), active ? (int32_t)stride : 0, active ? bits : 0, 1);
        }
        tab[(size_t)(off[a] + i)] = e;
      }
      stride *= f.size[a];
    }
    if (ctx->ftab.ensure(sizeof(int4) * tab.size()) != hipSuccess) return false;
    if (hipMemcpyAsync(ctx->ftab.p, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      return false;
    ctx->ftab_key_v = v;
    ctx->ftab_key_f = f;
    ctx->ftab_valid = true;
  }
  for (int a = 0; a < 3; ++a) F->ftab[a] = a < D ? ctx->ftab.as<int4>() + off[a] : nullptr;
  return true;
}

int bspline_separable(mmi_context* ctx, const DevGeom& virt, int D, AxisTable* tabs, bool* usable);

// B-spline part of the fast path: the per-axis support tables (shared with the
// separable derivative), the moving image's map, and the bound on the moving
// continuous index.  The displacement is evaluated as sum_k wx[k] * row[k]
// over coefficient lines reduced over the y and z supports instead of the
// reference's 64-term sum in node order; both are FP64 sums of non-negative
// weights (summing to 1) times coefficients, so they differ by at most
// ~100 ulp of max|coefficient|.  That, the rounding of vp + U and of the
// index map, is covered with a wide margin by
// e_ci = 1024 eps (rowsum|P2I| (D max|C| + max|mp| + max|o_m|) + max|ci|).
bool setup_fast_bspline(mmi_context* ctx, FastArgs* F) {
  const int D = ctx->fixed_g.dim;
  const mmi_geometry& v = ctx->virtual_g;
  const mmi_geometry& m = ctx->moving_g;
  if (ctx->dmtf.grid.dim != D || !is_diagonal(v.index_to_physical, D)) return false;
  AxisTable tabs[3];
  std::memset(tabs, 0, sizeof(tabs));
  bool usable = false;
  if (bspline_separable(ctx, to_dev(v), D, tabs, &usable) != MMI_OK || !usable) return false;
  for (int a = 0; a < 3; ++a) {
    F->bs_start[a] = a < D ? tabs[a].start : nullptr;
    F->bs_w[a] = a < D ? tabs[a].w : nullptr;
    F->bs_inside[a] = a < D ? tabs[a].inside : nullptr;
  }
  F->bs_gx = ctx->dmtf.grid.size[0];
  for (int k = 0; k < D; ++k) {
    F->vorg[k] = v.origin[k];
    F->vsp[k] = v.index_to_physical[k * D + k];
    F->morg[k] = m.origin[k];
  }
  for (int i = 0; i < D * D; ++i) F->mp2i[i] = m.physical_to_index[i];
  const int64_t Ny = v.size[1], Nz = D == 3 ? v.size[2] : 1;
  if (ctx->bs_rows.ensure(sizeof(double) * Nz * Ny * D * F->bs_gx) != hipSuccess) return false;
  if (ctx->bs_rows_t1.ensure(sizeof(double) * (D == 3 ? ctx->dmtf.grid.size[2] : 1) * Ny * D * F->bs_gx) != hipSuccess)
    return false;
  F->bs_rows = ctx->bs_rows.as<double>();
  double maxC = 0.0;
  for (double c : ctx->mtf.params) maxC = std::max(maxC, std::fabs(c));
  double maxVP = 0.0, maxO = 0.0, rowP2I = 0.0;
  for (int corner = 0; corner < (1 << D); ++corner)
    for (int k = 0; k < D; ++k) {
      const int64_t i = v.start[k] + (((corner >> k) & 1) ? v.size[k] - 1 : 0);
      maxVP = std::max(maxVP, std::fabs(v.origin[k]) + std::fabs(F->vsp[k] * (double)i));
    }
  for (int k = 0; k < D; ++k) {
    maxO = std::max(maxO, std::fabs(m.origin[k]));
    double r = 0.0;
    for (int j = 0; j < D; ++j) r += std::fabs(m.physical_to_index[k * D + j]);
    rowP2I = std::max(rowP2I, r);
  }
  const double maxMP = maxVP + D * maxC;
  const double maxCI = rowP2I * (maxMP + maxO);
  F->e_ci = 1024.0 * std::ldexp(1.0, -52) * (rowP2I * (D * maxC + maxMP + maxO) + maxCI) + 1e-300;
  return true;
}

// Continuous index in geometry g of the virtual index through T(x) = M x +
// off, with its bound (affine_index_map in mattes_affine.h).
void affine_index_map_geom(const mmi_geometry& v, const double* M, const double* off, const mmi_geometry& g,
                           double* A, double* b, double* e) {
  int32_t vs[3] = {0, 0, 0}, vz[3] = {1, 1, 1};
  for (int k = 0; k < v.dim; ++k) {
    vs[k] = (int32_t)v.start[k];
    vz[k] = (int32_t)v.size[k];
  }
  affine_index_map(v.dim, vs, vz, v.origin, v.index_to_physical, M, off, g.origin, g.physical_to_index, A, b, e);
}

// The level-constant inputs of the affine fast path's map and bounds.
AffineMapInputs affine_inputs(const mmi_context* ctx, const FastRows& rows) {
  AffineMapInputs in;
  std::memset(&in, 0, sizeof(in));
  const mmi_geometry& v = ctx->virtual_g;
  const mmi_geometry& m = ctx->moving_g;
  const int D = v.dim;
  in.dim = D;
  in.rows = rows.on;
  in.walk = rows.walk;
  in.bins = ctx->bins;
  for (int k = 0; k < 3; ++k) {
    in.v_start[k] = k < D ? (int32_t)v.start[k] : 0;
    in.v_size[k] = k < D ? (int32_t)v.size[k] : 1;
    in.m_start[k] = k < D ? (int32_t)m.start[k] : 0;
    in.m_size[k] = k < D ? (int32_t)m.size[k] : 1;
    in.v_origin[k] = v.origin[k];
    in.m_origin[k] = m.origin[k];
    in.center[k] = k < D && (int)ctx->mtf.fixed.size() > k ? ctx->mtf.fixed[k] : 0.0;
  }
  for (int i = 0; i < 9; ++i) {
    in.v_i2p[i] = v.index_to_physical[i];
    in.m_p2i[i] = m.physical_to_index[i];
  }
  in.Mm = std::max(std::fabs(ctx->moving_min), std::fabs(ctx->moving_max));
  in.moving_bin_size = ctx->moving_bin_size;
  in.moving_norm_min = ctx->moving_norm_min;
  return in;
}

// Row walks of the affine fast path (FastRows): items of 64 x by `walk` rows;
// the layout here, the fixed-point steps and thresholds with the map
// (rows_thresholds in mattes_affine.h).  R.on stays 0 when they do not apply.
void setup_rows_layout(const mmi_context* ctx, FastRows& R) {
  std::memset(&R, 0, sizeof(R));
  if (ctx->rows_walk <= 0) return;
  const int D = ctx->fixed_g.dim;
  const mmi_geometry& v = ctx->virtual_g;
  const mmi_geometry& m = ctx->moving_g;
  // cell offsets through 24-bit products; index steps far below 2^31
  const int64_t msy = m.size[0], msz = D == 3 ? m.size[0] * m.size[1] : 0;
  if (msy >= (1 << 24) || msz >= (1 << 24)) return;
  for (int k = 0; k < D; ++k)
    if (m.size[k] >= (1 << 24) || std::fabs(m.start[k]) >= (1 << 24)) return;
  const int64_t nx = v.size[0], ny = v.size[1], nz = D == 3 ? v.size[2] : 1;
  const int64_t nxseg = (nx + 63) / 64;
  // walks long enough to amortise the FP64 start, short enough that every
  // wavefront of the grid gets an item
  const int64_t waves = (int64_t)ctx->grid * 4;
  const int64_t walk = std::max<int64_t>(1, std::min<int64_t>({(int64_t)ctx->rows_walk, ny, nxseg * ny * nz / waves}));
  const int64_t nyblk = (ny + walk - 1) / walk;
  const int64_t items = nxseg * nyblk * nz;
  if (items >= (1ll << 31)) return;
  R.walk = (int32_t)walk;
  R.nx = (int32_t)nx;
  R.ny = (int32_t)ny;
  R.items = (uint32_t)items;
  R.dxseg = make_fastdiv((uint32_t)nxseg);
  R.dyblk = make_fastdiv((uint32_t)nyblk);
  R.on = 1;
  // central differences on the fast pass 2: a moving physical-to-index map
  // with exact zeros off the diagonal (a half-spacing step along a physical
  // axis then moves one index coordinate only, by 1/2 to within an ulp)
  if (ctx->gradient_source != MMI_GRADIENT_IMAGE) {
    bool diag = true;
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        const double v = m.physical_to_index[i * D + j];
        if (i != j && v != 0.0) diag = false;
        if (i == j && !(std::fabs(std::fabs(v * m.spacing[i]) - 1.0) <= 4.0 * 2.220446049250313e-16)) diag = false;
      }
    for (int i = 0; i < D; ++i)
      if (!(m.spacing[i] > 1e-12)) diag = false;
    if (diag) {
      R.cd_on = 1;
      for (int i = 0; i < D; ++i)
        R.cd_scale[i] = (float)((m.physical_to_index[i * D + i] > 0.0 ? 1.0 : -1.0) / m.spacing[i]);
    }
  }
}

// Whether the affine row walks can take the fixed bins from the level's byte
// map (FastArgs::fbin8): a dense unmasked domain, bins that fit a byte below
// kClsListed, and texels holding the level's moving image when pass 2 reads
// the gradient image (it takes the moving values from them).  The same for
// GetValue and GetValueAndDerivative, so both see the same histogram.
bool fb_rows_ok(const mmi_context* ctx) {
  return ctx->mtf.kind == MMI_TF_AFFINE && ctx->bins <= 253 && ctx->n_points == 0 && !ctx->have_mask[0] &&
         !ctx->have_mask[1] && ctx->pad_ok && ctx->fbin8.bytes >= (size_t)sample_count(ctx) &&
         (ctx->gradient_source != MMI_GRADIENT_IMAGE || (ctx->texel_values_level && ctx->pad_tex));
}

bool setup_fast(mmi_context* ctx, FastArgs* F) {
  std::memset(F, 0, sizeof(*F));
  // a composite moving transform maps through its earlier stages: the FP64 path
  if (!ctx->init_stages.empty()) return false;
  // integer pixels truncate the interpolated value: the FP64 path
  if (ctx->fixed_kind == kPixelInteger || ctx->moving_kind == kPixelInteger) return false;
  const int D = ctx->fixed_g.dim;
  const int B = ctx->bins;
  if (num_pixels(ctx->virtual_g) >= (1ll << 31)) return false;
  // point sets run on the exact path (masks: affine only, below)
  if (ctx->n_points > 0) return false;
  // the fast kernels address the images with 32-bit byte offsets
  if (num_pixels(ctx->moving_g) * (int64_t)sizeof(float4) >= (1ll << 32) ||
      num_pixels(ctx->fixed_g) * (int64_t)sizeof(float) >= (1ll << 32))
    return false;
  if (ctx->flags & 2u) return false;  // exact path requested
  const bool bs = ctx->mtf.kind == MMI_TF_BSPLINE;
  if ((ctx->mtf.kind != MMI_TF_AFFINE && !bs) || ctx->ftf.kind != MMI_TF_IDENTITY) return false;
  // central differences (affine): the fast pass 1 (the histogram needs no
  // gradient), then the exact pass 2
  if (ctx->gradient_source != MMI_GRADIENT_IMAGE && bs) return false;
  if ((size_t)B * B * 8 + (B + 1) * 4 > 64 * 1024) return false;
  const mmi_geometry& v = ctx->virtual_g;
  const mmi_geometry& m = ctx->moving_g;
  const DevTransform& t = ctx->dmtf;
  F->mk = bs ? 2 : 1;
  if (bs && !setup_fast_bspline(ctx, F)) return false;
  // image masks (affine): the masks' continuous indices are affine maps of the
  // virtual index as well, certified like the moving image's
  const mmi_geometry* mg[2] = {ctx->have_mask[0] ? &ctx->mask_g[0] : nullptr,
                               ctx->have_mask[1] ? &ctx->mask_g[1] : nullptr};
  FastMask* fm[2] = {&F->fmask, &F->mmask};
  for (int k = 0; k < 2; ++k) {
    fm[k]->px = nullptr;
    if (!mg[k]) continue;
    if (bs) return false;
    double ident[9] = {0}, zero[3] = {0, 0, 0};
    for (int i = 0; i < D; ++i) ident[i * D + i] = 1.0;
    // fixed mask at the fixed point (identity fixed transform), moving mask at
    // the mapped moving point
    if (k == 0)
      affine_index_map_geom(v, ident, zero, *mg[k], fm[k]->A, fm[k]->b, &fm[k]->e);
    else
      affine_index_map_geom(v, t.m, t.off, *mg[k], fm[k]->A, fm[k]->b, &fm[k]->e);
    if (fm[k]->e > 1e-6) return false;
    for (int i = 0; i < 3; ++i) {
      fm[k]->start[i] = i < D ? (int32_t)mg[k]->start[i] : 0;
      fm[k]->end[i] = i < D ? (int32_t)(mg[k]->start[i] + mg[k]->size[i] - 1) : 0;
    }
    fm[k]->sy = (int32_t)mg[k]->size[0];
    fm[k]->sz = (int32_t)(mg[k]->size[0] * (D > 1 ? mg[k]->size[1] : 1));
    fm[k]->px = ctx->mask_px[k].as<uint8_t>();
  }
  if (bs && F->e_ci > 1e-6) return false;
  // fixed samples on pixel centres (direct reads), else per-axis cell tables
  // and an FP32 interpolation with the same error bound as the moving value's
  int64_t off[3] = {0, 0, 0};
  const bool direct = fixed_direct_offsets(v, ctx->fixed_g, off);
  if (!direct && !fixed_axis_tables(ctx, v, ctx->fixed_g, F)) return false;
  double ev = 0.0, bt_m = 0.0;
  if (!bs) {
    // the affine map, row walks (their steps widen e_ci) and moving bounds, as
    // the device optimizer loop recomputes them after each step
    // (fixed samples off the fixed pixel centres: only with the fixed bins)
    if ((direct || fb_rows_ok(ctx)) && !mg[0] && !mg[1]) setup_rows_layout(ctx, F->rows);
    DevTransform tt;
    std::memset(&tt, 0, sizeof(tt));
    if (!affine_fast_update(affine_inputs(ctx, F->rows), ctx->mtf.params.data(), tt, *F)) return false;
    ev = F->ev;
    bt_m = F->bt_m;
  } else {
    moving_value_bounds(std::max(std::fabs(ctx->moving_min), std::fabs(ctx->moving_max)), ctx->moving_bin_size,
                        ctx->moving_norm_min, B, F->e_ci, &ev, &bt_m);
  }
  // intensity bounds
  const double Mf = std::max(std::fabs(ctx->fixed_min), std::fabs(ctx->fixed_max));
  const double u = std::ldexp(1.0, -22);
  const double ef = direct ? 0.0 : Mf * std::ldexp(1.0, -19);
  const double bt_f = 2.0 * ((ef + Mf * u) / ctx->fixed_bin_size + u * (Mf / ctx->fixed_bin_size +
                                                                       std::fabs(ctx->fixed_norm_min) + B));
  if (bt_m > 0.05 || bt_f > 0.05) return false;
  F->ev = (float)ev;
  F->bt_m = (float)bt_m;
  F->bt_f = (float)bt_f;
  F->inv_mbs = (float)(1.0 / ctx->moving_bin_size);
  F->mnm = (float)ctx->moving_norm_min;
  F->inv_fbs = (float)(1.0 / ctx->fixed_bin_size);
  F->fnm = (float)ctx->fixed_norm_min;
  F->mmin = (float)ctx->moving_min;
  F->mmax = (float)ctx->moving_max;
  F->fixed_interp = direct ? 0 : 1;
  const mmi_geometry& fg = ctx->fixed_g;
  for (int k = 0; k < 3; ++k) {
    F->foff[k] = k < D ? (int32_t)(off[k] - fg.start[k]) : 0;
    F->vstart[k] = k < D ? (int32_t)v.start[k] : 0;
    F->mstart[k] = k < D ? (int32_t)m.start[k] : 0;
    F->mend[k] = k < D ? (int32_t)(m.start[k] + m.size[k] - 1) : 0;
    F->cilo[k] = k < D ? (float)((double)m.start[k] - 0.5) : 0.0f;
    F->cihi[k] = k < D ? (float)((double)(m.start[k] + m.size[k] - 1) + 0.5) : 0.0f;
    if (k < D && ((double)F->cilo[k] != (double)m.start[k] - 0.5 ||
                  (double)F->cihi[k] != (double)(m.start[k] + m.size[k] - 1) + 0.5))
      return false;  // bounds not exact in FP32 (|index| >= 2^23)
  }
  F->vsize0 = (int32_t)v.size[0];
  F->vsize1 = (int32_t)(D > 1 ? v.size[1] : 1);
  F->vdx = make_fastdiv((uint32_t)F->vsize0);
  F->vdy = make_fastdiv((uint32_t)F->vsize1);
  F->fsy = (int32_t)fg.size[0];
  F->fsz = (int32_t)(fg.size[0] * (D > 1 ? fg.size[1] : 1));
  F->msy = (int32_t)m.size[0];
  F->msz = (int32_t)(m.size[0] * (D > 1 ? m.size[1] : 1));
  F->n = (uint32_t)num_pixels(v);
  F->bins = B;
  return true;
}

// FP32 images of the current level: the uploads, or their smoothed copies.
const float* level_fixed(const mmi_context* ctx) {
  return ctx->fixed_smoothed ? ctx->fixed_work.as<float>() : ctx->fixed_px.as<float>();
}
const float* level_moving(const mmi_context* ctx) {
  return ctx->moving_smoothed ? ctx->moving_work.as<float>() : ctx->moving_px.as<float>();
}

// Device images the evaluation reads: the level's FP32 images or their 16-bit copies.
const void* fixed_pixels(const mmi_context* ctx) {
  return ctx->active_storage == MMI_F32 ? (const void*)level_fixed(ctx) : ctx->fixed_lp.p;
}
const void* moving_pixels(const mmi_context* ctx) {
  return ctx->active_storage == MMI_F32 ? (const void*)level_moving(ctx) : ctx->moving_lp.p;
}
const void* texel_pixels(const mmi_context* ctx) {
  return ctx->active_storage == MMI_F32 ? ctx->texels.p : ctx->texels_lp.p;
}

// SmoothingRecursiveGaussianImageFilter on the device
// (itkSmoothingRecursiveGaussianImageFilter.hxx:30-70, 119-131, 199-260):
// zero-order recursive Gaussian along the last axis, then axes 0 .. D-2, in
// FP64 (normalise across scale off), cast to FP32 at the end.
// A chain of D recursive-Gaussian passes over one float image: pass s along
// axis[s] with coefficients c[s], FP64 between the passes in buf[0] / buf[1],
// the last pass's values into `last`.
struct FilterChain {
  const mmi_geometry* g;
  const float* in;
  int axis[3];
  RecCoeffs c[3];
  double* buf[2];
  RecSink last;
};

// Runs independent chains pass by pass, each pass of every chain in one
// launch (their lines together fill the device; each line's arithmetic is
// that of a chain run alone).
int run_filter_chains(mmi_context* ctx, const FilterChain* ch, int K) {
  if (K <= 0) return MMI_OK;
  const int D = ch[0].g->dim;
  for (int step = 0; step < D; ++step) {
    RecJobs J;
    std::memset(&J, 0, sizeof(J));
    J.n = K;
    for (int k = 0; k < K; ++k) {
      const FilterChain& f = ch[k];
      const mmi_geometry& g = *f.g;
      const int axis = f.axis[step];
      int64_t stride = 1;
      for (int j = 0; j < axis; ++j) stride *= g.size[j];
      RecJob& r = J.j[k];
      r.in = step == 0 ? static_cast<const void*>(f.in) : static_cast<const void*>(f.buf[(step - 1) & 1]);
      r.out = f.buf[step & 1];
      r.ln = g.size[axis];
      r.stride = stride;
      r.nlines = num_pixels(g) / r.ln;
      r.blocks = (int32_t)((r.nlines + kRecLines - 1) / kRecLines);
      r.c = f.c[step];
      r.sink.mode = kSinkDouble;
      if (step == D - 1) r.sink = f.last;
    }
    MMI_HIP(launch_recursive_gaussian_jobs(J, step == 0, ctx->stream));
  }
  return MMI_OK;
}

int check_filter_extent(const mmi_geometry& g, const char* word) {
  for (int i = 0; i < g.dim; ++i)
    if (g.size[i] < 4)
      return fail(MMI_ERR_ARGUMENT, "The number of pixels along " + std::string(word) + " " + std::to_string(i) +
                                        " is less than 4. This filter requires a minimum of four pixels along the "
                                        "dimension to be processed.");
  return MMI_OK;
}

// SmoothingRecursiveGaussianImageFilter on the device
// (itkSmoothingRecursiveGaussianImageFilter.hxx:30-70, 119-131, 199-260) as a
// chain: zero-order recursive Gaussian along the last axis, then axes 0 ..
// D-2, in FP64 (normalise across scale off), cast to FP32 at the end.
int smooth_chain(const mmi_geometry& g, const float* in, const double* sigma, double* b0, double* b1, float* out,
                 FilterChain& f) {
  if (int rc = check_filter_extent(g, "dimension")) return rc;
  const int D = g.dim;
  f.g = &g;
  f.in = in;
  f.buf[0] = b0;
  f.buf[1] = b1;
  for (int step = 0; step < D; ++step) {
    f.axis[step] = step == 0 ? D - 1 : step - 1;
    if (!rec_setup(sigma[f.axis[step]], g.spacing[f.axis[step]], 0, false, &f.c[step]))
      return fail(MMI_ERR_ARGUMENT, "recursive Gaussian: spacing or sigma out of range");
  }
  std::memset(&f.last, 0, sizeof(f.last));
  f.last.mode = kSinkFloat;
  f.last.outf = out;
  return MMI_OK;
}

// A new gradient image after mmi_initialize with 16-bit storage: refresh the
// 16-bit texels the evaluation reads.
int refresh_storage_texels(mmi_context* ctx) {
  if (ctx->active_storage == MMI_F32) return MMI_OK;
  const int64_t nm = num_pixels(ctx->moving_g);
  MMI_HIP(ctx->texels_lp.ensure(sizeof(uint2) * nm));
  MMI_HIP(launch_texels_to_storage(ctx->texels.as<float4>(), ctx->texels_lp.p, ctx->active_storage, nm, ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  return MMI_OK;
}

int device_minmax(mmi_context* ctx, const void* px, int storage, int64_t n, double* lo, double* hi) {
  const int blocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
  MMI_HIP(ctx->minmax_out.ensure(sizeof(float) * 2 * blocks));
  MMI_HIP(launch_minmax(px, storage, n, ctx->minmax_out.as<float>(), blocks, ctx->stream));
  std::vector<float> h(2 * blocks);
  MMI_HIP(hipMemcpyAsync(h.data(), ctx->minmax_out.p, sizeof(float) * 2 * blocks, hipMemcpyDeviceToHost, ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  float a = std::numeric_limits<float>::max(), b = -std::numeric_limits<float>::max();
  for (int i = 0; i < blocks; ++i) {
    a = std::min(a, h[2 * i]);
    b = std::max(b, h[2 * i + 1]);
  }
  *lo = a;
  *hi = b;
  return MMI_OK;
}

int device_minmax_select(mmi_context* ctx, const void* px, const mmi_geometry& g, const DevMask* mask,
                         const double* points, int64_t n, double* lo, double* hi) {
  const int blocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
  MMI_HIP(ctx->minmax_out.ensure(sizeof(float) * 2 * blocks));
  MMI_HIP(launch_minmax_select(px, ctx->active_storage, to_dev(g), mask, points, ctx->dftf, n, ctx->minmax_out.as<float>(),
                               blocks, ctx->stream));
  std::vector<float> h(2 * blocks);
  MMI_HIP(hipMemcpyAsync(h.data(), ctx->minmax_out.p, sizeof(float) * 2 * blocks, hipMemcpyDeviceToHost, ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  float a = std::numeric_limits<float>::max(), b = -std::numeric_limits<float>::max();
  for (int i = 0; i < blocks; ++i) {
    a = std::min(a, h[2 * i]);
    b = std::max(b, h[2 * i + 1]);
  }
  *lo = a;
  *hi = b;
  return MMI_OK;
}

// Axis tables of the separable B-spline derivative for this virtual domain and
// grid, rebuilt when either changes (they do not depend on the coefficients).
int bspline_axis_tables(mmi_context* ctx, const DevGeom& virt, const DevGeom& grid, int D, AxisTable* tabs,
                        bool* usable) {
  *usable = false;
  const bool same = ctx->bs_tables && std::memcmp(&ctx->bs_key_virt, &virt, sizeof(DevGeom)) == 0 &&
                    std::memcmp(&ctx->bs_key_grid, &grid, sizeof(DevGeom)) == 0;
  if (!same) {
    ctx->bs_tables = false;
    for (int j = 0; j < D; ++j) {
      const int len = virt.size[j], nodes = grid.size[j];
      MMI_HIP(ctx->bs_start[j].ensure(sizeof(int32_t) * len));
      MMI_HIP(ctx->bs_w[j].ensure(sizeof(double) * 4 * len));
      MMI_HIP(ctx->bs_inside[j].ensure(len));
      MMI_HIP(ctx->bs_lo[j].ensure(sizeof(int32_t) * nodes));
      MMI_HIP(ctx->bs_hi[j].ensure(sizeof(int32_t) * nodes));
      MMI_HIP(launch_bspline_axis_table(virt, grid, j, ctx->bs_start[j].as<int32_t>(), ctx->bs_w[j].as<double>(),
                                        ctx->bs_inside[j].as<uint8_t>(), ctx->stream));
      std::vector<int32_t> start(len);
      std::vector<uint8_t> inside(len);
      MMI_HIP(hipMemcpyAsync(start.data(), ctx->bs_start[j].p, sizeof(int32_t) * len, hipMemcpyDeviceToHost,
                             ctx->stream));
      MMI_HIP(hipMemcpyAsync(inside.data(), ctx->bs_inside[j].p, len, hipMemcpyDeviceToHost, ctx->stream));
      MMI_HIP(hipStreamSynchronize(ctx->stream));
      std::vector<int32_t> lo(nodes, 0), hi(nodes, -1);
      std::vector<bool> seen(nodes, false);
      for (int i = 0; i < len; ++i) {
        if (!inside[i]) continue;
        for (int k = 0; k < 4; ++k) {
          const int g = start[i] + k;
          if (g < 0 || g >= nodes) continue;
          if (!seen[g]) {
            lo[g] = i;
            seen[g] = true;
          }
          hi[g] = i;
        }
      }
      // The contraction kernels sum every index of [lo, hi]: each must hold g
      // in its support (true when the support start is monotone along the axis
      // and the valid indices form one interval, as for any axis-aligned grid).
      for (int g = 0; g < nodes; ++g)
        for (int i = lo[g]; i <= hi[g]; ++i)
          if (!inside[i] || g < start[i] || g > start[i] + 3) return MMI_OK;  // not usable: per-sample path
      int32_t inc = 1, prev = std::numeric_limits<int32_t>::min();
      for (int i = 0; i < len; ++i) {
        if (!inside[i]) continue;
        if (start[i] < prev) inc = 0;
        prev = start[i];
      }
      ctx->bs_increasing[j] = inc;
      MMI_HIP(hipMemcpyAsync(ctx->bs_lo[j].p, lo.data(), sizeof(int32_t) * nodes, hipMemcpyHostToDevice, ctx->stream));
      MMI_HIP(hipMemcpyAsync(ctx->bs_hi[j].p, hi.data(), sizeof(int32_t) * nodes, hipMemcpyHostToDevice, ctx->stream));
      MMI_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->bs_key_virt = virt;
    ctx->bs_key_grid = grid;
    ctx->bs_tables = true;
  }
  *usable = true;
  for (int j = 0; j < D; ++j) {
    tabs[j].start = ctx->bs_start[j].as<int32_t>();
    tabs[j].w = ctx->bs_w[j].as<double>();
    tabs[j].inside = ctx->bs_inside[j].as<uint8_t>();
    tabs[j].lo = ctx->bs_lo[j].as<int32_t>();
    tabs[j].hi = ctx->bs_hi[j].as<int32_t>();
    tabs[j].increasing = ctx->bs_increasing[j];
  }
  return MMI_OK;
}

// Whether the B-spline derivative of this evaluation can be contracted axis by
// axis: axis-aligned virtual domain and grid, monotone supports.
int bspline_separable(mmi_context* ctx, const DevGeom& virt, int D, AxisTable* tabs, bool* usable) {
  *usable = false;
  if (ctx->mtf.kind != MMI_TF_BSPLINE || ctx->dmtf.grid.dim != D) return MMI_OK;
  if (ctx->n_points > 0) return MMI_OK;  // samples are not the virtual grid's voxels
  if (!is_diagonal(ctx->virtual_g.index_to_physical, D) || !is_diagonal(ctx->dmtf.grid.p2i, D)) return MMI_OK;
  return bspline_axis_tables(ctx, virt, ctx->dmtf.grid, D, tabs, usable);
}

}  // namespace

extern "C" {


const char* mmi_last_error(void) { return g_last_error.c_str(); }

int mmi_context_create(int device, uint32_t flags, mmi_context** out) {
  if (!out) return fail(MMI_ERR_ARGUMENT, "out is NULL");
  int count = 0;
  MMI_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(MMI_ERR_ARGUMENT, "no such HIP device");
  MMI_HIP(hipSetDevice(device));
  // the static scratch's address on this device, once per device (the
  // library's set-up; mmi_evaluate_static only reads it)
  if (int rc = resolve_scratch(device)) return rc;
  auto* ctx = new mmi_context();
  ctx->device = device;
  ctx->flags = flags;
  if (const char* e = std::getenv("MMI_XCD_CHUNKS")) ctx->xcd_chunks = std::atoi(e) != 0;
  if (const char* e = std::getenv("MMI_BLOCKS_PER_CU")) {
    ctx->blocks_per_cu = std::max(1, std::min(64, std::atoi(e)));
    ctx->user_bpc = ctx->env_bpc = true;
  }
  if (const char* e = std::getenv("MMI_ROWS_WALK")) {
    ctx->rows_walk = std::max(0, std::min(256, std::atoi(e)));
    ctx->user_walk = ctx->env_walk = true;
  }
  if (const char* e = std::getenv("MMI_HIST_COPIES")) {
    ctx->hist_copies = std::max(0, std::min(4, std::atoi(e)));
    ctx->user_copies = ctx->env_copies = ctx->hist_copies != 0;
  }
  ctx->default_blocks_per_cu = ctx->blocks_per_cu;
  ctx->default_rows_walk = ctx->rows_walk;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(MMI_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->cu_count = prop.multiProcessorCount;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    ctx->max_lds_per_block = lds;
  for (auto& ev : ctx->ev) {
    if (hipEventCreate(&ev) != hipSuccess) {
      mmi_context_destroy(ctx);
      return fail(MMI_ERR_DEVICE, "hipEventCreate failed");
    }
  }
  *out = ctx;
  return MMI_OK;
}

int mmi_context_destroy(mmi_context* ctx) {
  if (!ctx) return MMI_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DevBuf* bufs[] = {&ctx->fixed_px, &ctx->moving_px, &ctx->texels, &ctx->fixed_lp, &ctx->moving_lp, &ctx->texels_lp,
                    &ctx->fixed_work, &ctx->moving_work,
                    &ctx->coeffs, &ctx->hist,
                    &ctx->nvalid, &ctx->pdf, &ctx->fm, &ctx->mm, &ctx->pratio, &ctx->pratio_f, &ctx->rowsum, &ctx->out, &ctx->points, &ctx->point_ids, &ctx->mask_px[0], &ctx->mask_px[1],
                    &ctx->mask_desc[0], &ctx->mask_desc[1],
                    &ctx->partials, &ctx->sample_fbin, &ctx->sample_mbin, &ctx->minmax_out,
                    &ctx->exc_count, &ctx->exc_list, &ctx->exc_res, &ctx->cls, &ctx->fbin8, &ctx->mpad, &ctx->tpad, &ctx->bs_g, &ctx->bs_t1, &ctx->bs_t2, &ctx->bs_rows, &ctx->bs_rows_t1, &ctx->ftab, &ctx->work_a,
                    &ctx->work_b, &ctx->work_grad, &ctx->work_c, &ctx->work_d, &ctx->work_e, &ctx->work_f};
  for (DevBuf* b : bufs) b->release();
  for (int j = 0; j < 3; ++j) {
    ctx->bs_start[j].release();
    ctx->bs_w[j].release();
    ctx->bs_inside[j].release();
    ctx->bs_lo[j].release();
    ctx->bs_hi[j].release();
  }
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->h_coeffs) (void)hipHostFree(ctx->h_coeffs);
  if (ctx->h_out) (void)hipHostFree(ctx->h_out);
  if (ctx->h_res) (void)hipHostFree(ctx->h_res);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MMI_OK;
}

int mmi_set_storage(mmi_context* ctx, int32_t pixel_type) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (pixel_type != MMI_F32 && pixel_type != MMI_BF16 && pixel_type != MMI_F16)
    return fail(MMI_ERR_ARGUMENT, "storage must be MMI_F32, MMI_BF16 or MMI_F16");
  if (pixel_type != ctx->storage) ctx->initialized = false;
  ctx->storage = pixel_type;
  return MMI_OK;
}

int mmi_set_fixed_image(mmi_context* ctx, const mmi_geometry* g, int32_t pixel_type, const void* pixels) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  MMI_HIP(hipSetDevice(ctx->device));
  if (int rc = upload_image(ctx, ctx->fixed_px, g, pixel_type, pixels, &ctx->fixed_kind)) return rc;
  ctx->fixed_g = *g;
  ctx->have_fixed = true;
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_set_moving_image(mmi_context* ctx, const mmi_geometry* g, int32_t pixel_type, const void* pixels) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  MMI_HIP(hipSetDevice(ctx->device));
  if (int rc = upload_image(ctx, ctx->moving_px, g, pixel_type, pixels, &ctx->moving_kind)) return rc;
  ctx->moving_g = *g;
  ctx->have_moving = true;
  ctx->have_texels = false;
  ctx->user_texels = false;
  ctx->moving_smoothed = false;
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_set_moving_gradient(mmi_context* ctx, const double* gradient) {
  if (!ctx || !gradient) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if (!ctx->have_moving) return fail(MMI_ERR_ARGUMENT, "set the moving image first");
  MMI_HIP(hipSetDevice(ctx->device));
  const int D = ctx->moving_g.dim;
  const int64_t n = num_pixels(ctx->moving_g);
  DevBuf& tmp = ctx->work_grad;
  MMI_HIP(tmp.ensure(sizeof(double) * n * D));
  MMI_HIP(hipMemcpyAsync(tmp.p, gradient, sizeof(double) * n * D, hipMemcpyHostToDevice, ctx->stream));
  MMI_HIP(ctx->texels.ensure(sizeof(float4) * n));
  Direction ident{};
  for (int i = 0; i < D; ++i) ident.m[i * D + i] = 1.0;
  MMI_HIP(launch_pack_texels(ctx->moving_px.as<float>(), tmp.as<double>(), n, D, ident, ctx->texels.as<float4>(),
                             ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_texels = true;
  ctx->user_texels = true;
  ctx->texel_values_level = false;  // packed from the unsmoothed upload; mmi_initialize decides
  return refresh_storage_texels(ctx);
}

int mmi_compute_moving_gradient(mmi_context* ctx) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->have_moving) return fail(MMI_ERR_ARGUMENT, "set the moving image first");
  MMI_HIP(hipSetDevice(ctx->device));
  const mmi_geometry& g = ctx->moving_g;
  const int D = g.dim;
  const int64_t n = num_pixels(g);
  if (int rc = check_filter_extent(g, "direction")) return rc;
  // GradientRecursiveGaussianImageFilter as the metric configures it:
  // sigma = maximum spacing, normalised across scale (I2IM:487-503).
  double sigma = 0.0;
  for (int i = 0; i < D; ++i) sigma = std::max(sigma, g.spacing[i]);
  DevBuf* work[6] = {&ctx->work_a, &ctx->work_b, &ctx->work_c, &ctx->work_d, &ctx->work_e, &ctx->work_f};
  for (int i = 0; i < 2 * D; ++i) MMI_HIP(work[i]->ensure(sizeof(double) * n));
  DevBuf& grad = ctx->work_grad;
  MMI_HIP(grad.ensure(sizeof(double) * n * D));
  // one chain per component, run together: derivative along dim, then zero
  // order along the other axes in increasing order (GradRG:214-238); the
  // float image is widened inside the first pass, FP64 between the passes,
  // the last pass writes component dim (value / spacing)
  FilterChain ch[3];
  for (int dim = 0; dim < D; ++dim) {
    FilterChain& f = ch[dim];
    f.g = &g;
    f.in = level_moving(ctx);
    f.buf[0] = work[2 * dim]->as<double>();
    f.buf[1] = work[2 * dim + 1]->as<double>();
    f.axis[0] = dim;
    int k = 1;
    for (int o = 0; o < D; ++o)
      if (o != dim) f.axis[k++] = o;
    for (int step = 0; step < D; ++step)
      if (!rec_setup(sigma, g.spacing[f.axis[step]], step == 0 ? 1 : 0, true, &f.c[step]))
        return fail(MMI_ERR_ARGUMENT, "recursive Gaussian: spacing or sigma out of range");
    std::memset(&f.last, 0, sizeof(f.last));
    f.last.mode = kSinkGradient;
    f.last.comp = dim;
    f.last.dim = D;
    f.last.spacing = g.spacing[dim];
    f.last.grad = grad.as<double>();
  }
  if (int rc = run_filter_chains(ctx, ch, D)) return rc;
  Direction dir{};
  for (int i = 0; i < D * D; ++i) dir.m[i] = g.direction[i];
  MMI_HIP(ctx->texels.ensure(sizeof(float4) * n));
  MMI_HIP(launch_pack_texels(level_moving(ctx), grad.as<double>(), n, D, dir, ctx->texels.as<float4>(), ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_texels = true;
  ctx->user_texels = false;
  ctx->texel_values_level = true;
  for (int i = 0; i < 3; ++i) ctx->texel_sigma[i] = ctx->moving_smoothed ? ctx->moving_sigma[i] : 0.0;
  return refresh_storage_texels(ctx);
}

int mmi_set_gradient_source(mmi_context* ctx, int32_t source) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (source != MMI_GRADIENT_IMAGE && source != MMI_GRADIENT_CENTRAL_DIFFERENCE)
    return fail(MMI_ERR_ARGUMENT, "unknown gradient source");
  ctx->gradient_source = source;
  return MMI_OK;
}

int mmi_set_virtual_domain(mmi_context* ctx, const mmi_geometry* g) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (int rc = check_geometry(g)) return rc;
  ctx->virtual_g = *g;
  ctx->have_virtual = true;
  return MMI_OK;
}

int mmi_set_virtual_points(mmi_context* ctx, int32_t dim, const double* points, int64_t n) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (n < 0 || (n > 0 && !points)) return fail(MMI_ERR_ARGUMENT, "points is NULL or n < 0");
  if (n > 0 && dim != 2 && dim != 3) return fail(MMI_ERR_ARGUMENT, "points must be 2-D or 3-D");
  if (n >= (1ll << 31)) return fail(MMI_ERR_ARGUMENT, "at most 2^31 - 1 points");
  MMI_HIP(hipSetDevice(ctx->device));
  if (n > 0) {
    // stored in Morton order of their bounding-box cells, so that neighbouring
    // threads read neighbouring image memory; the histogram is an integer sum
    // (order-free) and the bins are reported at the caller's indices
    std::vector<uint32_t> order;
    spatial_order(points, dim, n, &order);
    std::vector<double> sorted((size_t)n * dim);
    for (int64_t k = 0; k < n; ++k)
      for (int d = 0; d < dim; ++d) sorted[(size_t)k * dim + d] = points[(size_t)order[k] * dim + d];
    MMI_HIP(ctx->points.ensure(sizeof(double) * dim * n));
    MMI_HIP(ctx->point_ids.ensure(sizeof(uint32_t) * n));
    MMI_HIP(hipMemcpy(ctx->points.p, sorted.data(), sizeof(double) * dim * n, hipMemcpyHostToDevice));
    MMI_HIP(hipMemcpy(ctx->point_ids.p, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  }
  ctx->n_points = n;
  ctx->points_dim = n > 0 ? dim : 0;
  ctx->initialized = false;
  return MMI_OK;
}

static int set_mask(mmi_context* ctx, int which, const mmi_geometry* g, const uint8_t* pixels) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  ctx->initialized = false;
  if (!g) {  // no geometry: remove the mask
    ctx->have_mask[which] = false;
    return MMI_OK;
  }
  if (!pixels) return fail(MMI_ERR_ARGUMENT, "mask pixels are NULL");
  if (int rc = check_geometry(g)) return rc;
  const int64_t n = num_pixels(*g);
  if (n >= (1ll << 31)) return fail(MMI_ERR_ARGUMENT, "masks of 2^31 or more pixels are not supported");
  MMI_HIP(hipSetDevice(ctx->device));
  MMI_HIP(ctx->mask_px[which].ensure(n));
  MMI_HIP(hipMemcpy(ctx->mask_px[which].p, pixels, n, hipMemcpyHostToDevice));
  DevMask m;
  m.g = to_dev(*g);
  m.px = ctx->mask_px[which].as<uint8_t>();
  MMI_HIP(ctx->mask_desc[which].ensure(sizeof(DevMask)));
  MMI_HIP(hipMemcpy(ctx->mask_desc[which].p, &m, sizeof(DevMask), hipMemcpyHostToDevice));
  ctx->mask_g[which] = *g;
  ctx->have_mask[which] = true;
  return MMI_OK;
}

int mmi_set_fixed_mask(mmi_context* ctx, const mmi_geometry* g, const uint8_t* pixels) {
  return set_mask(ctx, 0, g, pixels);
}

int mmi_set_moving_mask(mmi_context* ctx, const mmi_geometry* g, const uint8_t* pixels) {
  return set_mask(ctx, 1, g, pixels);
}

int mmi_set_smoothing_sigmas(mmi_context* ctx, const double* fixed_sigma, const double* moving_sigma) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->have_fixed || !ctx->have_moving) return fail(MMI_ERR_ARGUMENT, "set the fixed and moving images first");
  const int D = ctx->fixed_g.dim;
  const double* src[2] = {fixed_sigma, moving_sigma};
  double* dst[2] = {ctx->fixed_sigma, ctx->moving_sigma};
  for (int k = 0; k < 2; ++k) {
    int positive = 0;
    for (int i = 0; i < 3; ++i) dst[k][i] = (src[k] && i < D) ? src[k][i] : 0.0;
    for (int i = 0; i < D; ++i) {
      if (!(dst[k][i] >= 0.0)) return fail(MMI_ERR_ARGUMENT, "smoothing sigmas must be >= 0");
      positive += dst[k][i] > 0.0;
    }
    if (positive != 0 && positive != D)
      return fail(MMI_ERR_ARGUMENT, "smoothing sigmas must be all zero or all positive");
  }
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_set_launch_config(mmi_context* ctx, int32_t blocks_per_cu, int32_t rows_walk) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (blocks_per_cu > 64 || rows_walk > 256) return fail(MMI_ERR_ARGUMENT, "launch configuration out of range");
  ctx->blocks_per_cu = blocks_per_cu > 0 ? blocks_per_cu : ctx->default_blocks_per_cu;
  ctx->rows_walk = rows_walk >= 0 ? rows_walk : ctx->default_rows_walk;
  ctx->user_bpc = blocks_per_cu > 0 || ctx->env_bpc;
  ctx->user_walk = rows_walk >= 0 || ctx->env_walk;
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_set_histogram_copies(mmi_context* ctx, int32_t copies) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (copies != 0 && copies != 1 && copies != 2 && copies != 4)
    return fail(MMI_ERR_ARGUMENT, "histogram copies must be 0 (the library's choice), 1, 2 or 4");
  if (copies != 0 || !ctx->env_copies) ctx->hist_copies = copies;
  ctx->user_copies = copies != 0 || ctx->env_copies;
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_get_launch_config(mmi_context* ctx, int32_t* blocks_per_cu, int32_t* rows_walk, int32_t* hist_copies,
                          int32_t* from_table) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (blocks_per_cu) *blocks_per_cu = ctx->blocks_per_cu;
  if (rows_walk) *rows_walk = ctx->rows_walk;
  if (hist_copies) *hist_copies = ctx->hist_copies ? ctx->hist_copies : hist_copies_default(ctx->bins);
  if (from_table) *from_table = ctx->tuned;
  return MMI_OK;
}

int mmi_load_tuning_table(const char* path, int32_t* n_rows) {
  std::vector<TuneRow> rows;
  const std::string p = path ? std::string(path) : default_tuning_path();
  if (int rc = read_tuning_table(p, rows)) return rc;
  std::lock_guard<std::mutex> lock(g_tuning_mu);
  g_tuning = rows;
  g_tuning_loaded = true;
  if (n_rows) *n_rows = (int32_t)rows.size();
  return MMI_OK;
}

int mmi_set_number_of_histogram_bins(mmi_context* ctx, int32_t bins) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  ctx->bins = std::max<int32_t>(5, bins);  // itkSetClampMacro(..., 5, max)
  ctx->initialized = false;
  return MMI_OK;
}

int mmi_initialize(mmi_context* ctx) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->have_fixed) return fail(MMI_ERR_ARGUMENT, "FixedImage is not present");
  if (!ctx->have_moving) return fail(MMI_ERR_ARGUMENT, "MovingImage is not present");
  if (ctx->fixed_g.dim != ctx->moving_g.dim) return fail(MMI_ERR_ARGUMENT, "fixed and moving dimensions differ");
  if (ctx->n_points > 0 && ctx->points_dim != ctx->fixed_g.dim)
    return fail(MMI_ERR_ARGUMENT, "point set and image dimensions differ");
  for (int k = 0; k < 2; ++k)
    if (ctx->have_mask[k] && ctx->mask_g[k].dim != ctx->fixed_g.dim)
      return fail(MMI_ERR_ARGUMENT, "mask and image dimensions differ");
  MMI_HIP(hipSetDevice(ctx->device));
  // Per-level smoothing of the fixed and moving images
  // (ImageRegistrationMethodv4::InitializeRegistrationAtEachLevel,
  // itkImageRegistrationMethodv4.hxx:591-637).
  const int D = ctx->fixed_g.dim;
  // the reference keeps a smoothed image in its pixel type: only float images
  // are smoothed here (their device copy is that type)
  if ((ctx->fixed_sigma[0] > 0.0 && ctx->fixed_kind != kPixelFloat) ||
      (ctx->moving_sigma[0] > 0.0 && ctx->moving_kind != kPixelFloat))
    return fail(MMI_ERR_ARGUMENT, "smoothing takes float images only (other pixel types have no device copy of "
                                  "their smoothed image)");
  ctx->fixed_smoothed = ctx->moving_smoothed = false;
  {
    // the fixed and moving smoothing chains together
    FilterChain ch[2];
    int K = 0;
    if (ctx->fixed_sigma[0] > 0.0) {
      const int64_t nf = num_pixels(ctx->fixed_g);
      MMI_HIP(ctx->work_a.ensure(sizeof(double) * nf));
      MMI_HIP(ctx->work_b.ensure(sizeof(double) * nf));
      MMI_HIP(ctx->fixed_work.ensure(sizeof(float) * nf));
      if (int rc = smooth_chain(ctx->fixed_g, ctx->fixed_px.as<float>(), ctx->fixed_sigma, ctx->work_a.as<double>(),
                                ctx->work_b.as<double>(), ctx->fixed_work.as<float>(), ch[K++]))
        return rc;
    }
    if (ctx->moving_sigma[0] > 0.0) {
      const int64_t nm = num_pixels(ctx->moving_g);
      MMI_HIP(ctx->work_c.ensure(sizeof(double) * nm));
      MMI_HIP(ctx->work_d.ensure(sizeof(double) * nm));
      MMI_HIP(ctx->moving_work.ensure(sizeof(float) * nm));
      if (int rc = smooth_chain(ctx->moving_g, ctx->moving_px.as<float>(), ctx->moving_sigma,
                                ctx->work_c.as<double>(), ctx->work_d.as<double>(), ctx->moving_work.as<float>(),
                                ch[K++]))
        return rc;
    }
    if (int rc = run_filter_chains(ctx, ch, K)) return rc;
    MMI_HIP(hipStreamSynchronize(ctx->stream));
    ctx->fixed_smoothed = ctx->fixed_sigma[0] > 0.0;
    ctx->moving_smoothed = ctx->moving_sigma[0] > 0.0;
  }
  bool stale = !ctx->have_texels;
  if (ctx->have_texels && !ctx->user_texels)
    for (int i = 0; i < D; ++i) stale = stale || ctx->texel_sigma[i] != (ctx->moving_smoothed ? ctx->moving_sigma[i] : 0.0);
  if (ctx->gradient_source == MMI_GRADIENT_IMAGE && stale) {
    if (int rc = mmi_compute_moving_gradient(ctx)) return rc;
  }
  // user texels were packed from the unsmoothed upload
  if (ctx->user_texels) ctx->texel_values_level = !ctx->moving_smoothed;
  // 16-bit storage: every later read of the images (intensity range, bins,
  // interpolation, gradient) sees the rounded copies.
  ctx->active_storage = MMI_F32;
  if (ctx->storage != MMI_F32) {
    const int64_t nf = num_pixels(ctx->fixed_g), nm = num_pixels(ctx->moving_g);
    MMI_HIP(ctx->fixed_lp.ensure(sizeof(uint16_t) * nf));
    MMI_HIP(ctx->moving_lp.ensure(sizeof(uint16_t) * nm));
    MMI_HIP(launch_to_storage(level_fixed(ctx), ctx->fixed_lp.p, ctx->storage, nf, ctx->stream));
    MMI_HIP(launch_to_storage(level_moving(ctx), ctx->moving_lp.p, ctx->storage, nm, ctx->stream));
    if (ctx->have_texels) {
      MMI_HIP(ctx->texels_lp.ensure(sizeof(uint2) * nm));
      MMI_HIP(launch_texels_to_storage(ctx->texels.as<float4>(), ctx->texels_lp.p, ctx->storage, nm, ctx->stream));
    }
    ctx->active_storage = ctx->storage;
  }
  // Mattes:105-220: the fixed range over the sampled points when sparse
  // sampling is on (masks then do not matter), else over the masked voxels;
  // the moving range over its masked voxels
  if (ctx->n_points > 0) {
    if (int rc = device_minmax_select(ctx, fixed_pixels(ctx), ctx->fixed_g, nullptr, ctx->points.as<double>(),
                                      ctx->n_points, &ctx->fixed_min, &ctx->fixed_max))
      return rc;
  } else if (ctx->have_mask[0]) {
    if (int rc = device_minmax_select(ctx, fixed_pixels(ctx), ctx->fixed_g, ctx->mask_desc[0].as<DevMask>(), nullptr,
                                      num_pixels(ctx->fixed_g), &ctx->fixed_min, &ctx->fixed_max))
      return rc;
  } else if (int rc = device_minmax(ctx, fixed_pixels(ctx), ctx->active_storage, num_pixels(ctx->fixed_g),
                                    &ctx->fixed_min, &ctx->fixed_max)) {
    return rc;
  }
  if (ctx->have_mask[1]) {
    if (int rc = device_minmax_select(ctx, moving_pixels(ctx), ctx->moving_g, ctx->mask_desc[1].as<DevMask>(), nullptr,
                                      num_pixels(ctx->moving_g), &ctx->moving_min, &ctx->moving_max))
      return rc;
  } else if (int rc = device_minmax(ctx, moving_pixels(ctx), ctx->active_storage, num_pixels(ctx->moving_g),
                                    &ctx->moving_min, &ctx->moving_max)) {
    return rc;
  }
  if (ctx->fixed_min > ctx->fixed_max || ctx->moving_min > ctx->moving_max)
    return fail(MMI_ERR_DEGENERATE, "no fixed or moving voxel (sampled point) inside the analysis region");
  if (!std::isfinite(ctx->fixed_min) || !std::isfinite(ctx->fixed_max) || !std::isfinite(ctx->moving_min) ||
      !std::isfinite(ctx->moving_max))
    return fail(MMI_ERR_ARGUMENT, "image intensities are not finite in the storage format (FP16 holds |v| <= 65504)");
  // Mattes:273-317
  const double eps = std::numeric_limits<double>::epsilon();
  char buf[512];
  if (ctx->fixed_max - ctx->fixed_min <= eps) {
    std::snprintf(buf, sizeof(buf),
                  "All fixed-image samples in the analysis region have the same intensity value (%g). Mattes mutual "
                  "information is undefined for a constant image: the marginal entropy is zero and no meaningful "
                  "gradient can be computed.",
                  ctx->fixed_min);
    return fail(MMI_ERR_DEGENERATE, buf);
  }
  if (ctx->moving_max - ctx->moving_min <= eps) {
    std::snprintf(buf, sizeof(buf),
                  "All moving-image samples in the analysis region have the same intensity value (%g). Mattes "
                  "mutual information is undefined for a constant image: the marginal entropy is zero and no "
                  "meaningful gradient can be computed.",
                  ctx->moving_min);
    return fail(MMI_ERR_DEGENERATE, buf);
  }
  const int padding = 2;
  ctx->fixed_bin_size = (ctx->fixed_max - ctx->fixed_min) / (double)(ctx->bins - 2 * padding);
  ctx->fixed_norm_min = ctx->fixed_min / ctx->fixed_bin_size - (double)padding;
  ctx->moving_bin_size = (ctx->moving_max - ctx->moving_min) / (double)(ctx->bins - 2 * padding);
  ctx->moving_norm_min = ctx->moving_min / ctx->moving_bin_size - (double)padding;
  if (!ctx->have_virtual) ctx->virtual_g = ctx->fixed_g;
  const int B = ctx->bins;
  MMI_HIP(ctx->hist.ensure(sizeof(unsigned long long) * B * B));
  MMI_HIP(ctx->nvalid.ensure(sizeof(unsigned long long)));
  MMI_HIP(ctx->pdf.ensure(sizeof(double) * B * B));
  MMI_HIP(ctx->pratio.ensure(sizeof(double) * B * B));
  MMI_HIP(ctx->pratio_f.ensure(sizeof(float) * B * B));
  MMI_HIP(ctx->fm.ensure(sizeof(double) * B));
  MMI_HIP(ctx->mm.ensure(sizeof(double) * B));
  MMI_HIP(ctx->rowsum.ensure(sizeof(double) * B));
  MMI_HIP(ctx->out.ensure(out_bytes((int64_t)ctx->mtf.params.size())));
  const int64_t n = sample_count(ctx);
  apply_tuning(ctx, n);
  // a multiple of 8: the fast kernels split the samples into one share per XCD
  const int64_t want = std::min<int64_t>((n + 255) / 256, (int64_t)ctx->cu_count * ctx->blocks_per_cu);
  ctx->grid = (int)std::max<int64_t>(8, (want + 7) / 8 * 8);
  if (ctx->flags & 1u) {
    MMI_HIP(ctx->sample_fbin.ensure(sizeof(int16_t) * n));
    MMI_HIP(ctx->sample_mbin.ensure(sizeof(int16_t) * n));
  }
  // the fixed bins of a dense domain, for the affine row walks (bins fit a
  // byte below kClsListed)
  ctx->fbin8_valid = false;
  ctx->pad_ok = ctx->pad_tex = false;
  if (ctx->n_points == 0 && B <= 253 && n < (1ll << 31) && ctx->mtf.kind == MMI_TF_AFFINE) {
    const mmi_geometry& mg = ctx->moving_g;
    const int64_t px = mg.size[0] + 2, py = mg.size[1] + 2, pz = D == 3 ? mg.size[2] + 2 : 1;
    const int64_t pn = px * py * pz;
    const int vb = ctx->active_storage == MMI_F32 ? 4 : 2, tb = ctx->active_storage == MMI_F32 ? 16 : 8;
    // 32-bit byte offsets and 24-bit index products in the kernels
    if (pn * tb < (1ll << 32) && px * py < (1ll << 24) && pz < (1ll << 24)) {
      MMI_HIP(ctx->fbin8.ensure((size_t)n));
      MMI_HIP(ctx->mpad.ensure((size_t)(pn * vb)));
      MMI_HIP(launch_pad_replicate(moving_pixels(ctx), ctx->mpad.p, vb, (int)mg.size[0], (int)mg.size[1],
                                   D == 3 ? (int)mg.size[2] : 1, D, ctx->stream));
      if (ctx->have_texels) {
        MMI_HIP(ctx->tpad.ensure((size_t)(pn * tb)));
        MMI_HIP(launch_pad_replicate(texel_pixels(ctx), ctx->tpad.p, tb, (int)mg.size[0], (int)mg.size[1],
                                     D == 3 ? (int)mg.size[2] : 1, D, ctx->stream));
        ctx->pad_tex = true;
      }
      ctx->pad_ok = true;
    }
  }
  ctx->initialized = true;
  ctx->acc_clean = false;  // the accumulators may have been reallocated
  return MMI_OK;
}

static int set_transform(mmi_context* ctx, HostTransform& h, DevTransform& d, bool moving, int32_t kind,
                         const double* fixed_params, int32_t n_fixed, const double* params, int64_t n_params) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->have_fixed) return fail(MMI_ERR_ARGUMENT, "set the fixed image before the transforms");
  MMI_HIP(hipSetDevice(ctx->device));
  h.kind = kind;
  h.fixed.assign(fixed_params, fixed_params + std::max(0, n_fixed));
  h.params.assign(params, params + std::max<int64_t>(0, n_params));
  return build_transform(ctx, ctx->fixed_g.dim, h, d, moving);
}

int mmi_set_fixed_transform(mmi_context* ctx, int32_t kind, const double* fixed_params, int32_t n_fixed,
                            const double* params, int64_t n_params) {
  return set_transform(ctx, ctx->ftf, ctx->dftf, false, kind, fixed_params, n_fixed, params, n_params);
}

int mmi_set_moving_transform(mmi_context* ctx, int32_t kind, const double* fixed_params, int32_t n_fixed,
                             const double* params, int64_t n_params) {
  return set_transform(ctx, ctx->mtf, ctx->dmtf, true, kind, fixed_params, n_fixed, params, n_params);
}

int mmi_set_moving_initial_transforms(mmi_context* ctx, int32_t n, const double* matrices, const double* offsets) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (n < 0 || n > kMaxInitialStages)
    return fail(MMI_ERR_ARGUMENT, "a composite moving transform takes 0 to " + std::to_string(kMaxInitialStages) +
                                      " transforms before the moving transform");
  if (n > 0 && (!matrices || !offsets)) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  const int D = ctx->fixed_g.dim ? ctx->fixed_g.dim : ctx->moving_g.dim;
  if (n > 0 && D != 2 && D != 3) return fail(MMI_ERR_ARGUMENT, "set the images before the composite stages");
  std::vector<double> st((size_t)n * 12, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < D * D; ++k) st[12 * i + k] = matrices[(size_t)i * D * D + k];
    for (int k = 0; k < D; ++k) st[12 * i + 9 + k] = offsets[(size_t)i * D + k];
    for (int k = 0; k < 12; ++k)
      if (!std::isfinite(st[12 * i + k])) return fail(MMI_ERR_ARGUMENT, "composite stage is not finite");
  }
  MMI_HIP(hipSetDevice(ctx->device));
  if (n > 0) {
    MMI_HIP(ctx->init_dev.ensure(sizeof(double) * st.size()));
    MMI_HIP(hipStreamSynchronize(ctx->stream));
    MMI_HIP(hipMemcpy(ctx->init_dev.p, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice));
  }
  ctx->init_stages = st;
  return MMI_OK;
}

int mmi_set_moving_parameters(mmi_context* ctx, const double* params, int64_t n_params) {
  if (!ctx || !params) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if ((int64_t)ctx->mtf.params.size() != n_params)
    return fail(MMI_ERR_ARGUMENT, "parameter count differs from the moving transform's");
  ctx->mtf.params.assign(params, params + n_params);
  return build_transform(ctx, ctx->fixed_g.dim, ctx->mtf, ctx->dmtf, true);
}

}  // extern "C"

namespace {

// The separable B-spline derivative from the per-sample factors in bs_g:
// contraction along x (LDS-staged rows), then y, then z with the nFactor scale.
int bspline_contract(mmi_context* ctx, const DevGeom& virt, const AxisTable* tabs, const DevResult* result,
                     double* deriv) {
  const int D = virt.dim;
  hipStream_t st = ctx->stream;
  const int64_t Nx = virt.size[0], Ny = virt.size[1], Nz = D == 3 ? virt.size[2] : 1;
  const int Gx = ctx->dmtf.grid.size[0], Gy = ctx->dmtf.grid.size[1], Gz = D == 3 ? ctx->dmtf.grid.size[2] : 1;
  MMI_HIP(ctx->bs_t1.ensure(sizeof(double) * D * Nz * Ny * Gx));
  MMI_HIP(launch_separable_contract(ctx->bs_g.as<float>(), nullptr, ctx->bs_t1.as<double>(), D * Nz * Ny, (int)Nx, 1,
                                    Gx, tabs[0], nullptr, st));
  if (D == 3) {
    MMI_HIP(ctx->bs_t2.ensure(sizeof(double) * D * Nz * Gy * Gx));
    MMI_HIP(launch_separable_contract(nullptr, ctx->bs_t1.as<double>(), ctx->bs_t2.as<double>(), D * Nz, (int)Ny, Gx,
                                      Gy, tabs[1], nullptr, st));
    MMI_HIP(launch_separable_contract(nullptr, ctx->bs_t2.as<double>(), deriv, D, (int)Nz,
                                      (int64_t)Gy * Gx, Gz, tabs[2], result, st));
  } else {
    MMI_HIP(launch_separable_contract(nullptr, ctx->bs_t1.as<double>(), deriv, D, (int)Ny, Gx, Gy,
                                      tabs[1], result, st));
  }
  return MMI_OK;
}

// Validation, buffers and kernel arguments of one evaluation; `fast` tells
// whether the certified fast path applies (then `fa` is complete, including
// the derivative partials when want_derivative is set).  sb: the buffers to
// write (mmi_evaluate_static's scratch), or NULL for the context's own,
// grown to this evaluation's sizes; `b` receives the ones used.
int prepare_eval(mmi_context* ctx, int32_t want_derivative, EvalArgs& a, FastArgs& fa, bool& fast,
                 const EvalBufs* sb, EvalBufs& b) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->initialized) return fail(MMI_ERR_ARGUMENT, "mmi_initialize must precede mmi_evaluate");
  if (ctx->mtf.kind == MMI_TF_IDENTITY && want_derivative)
    return fail(MMI_ERR_ARGUMENT, "MovingTransform is not present");
  if (ctx->virtual_g.dim != ctx->fixed_g.dim) return fail(MMI_ERR_ARGUMENT, "virtual domain dimension differs");
  MMI_HIP(hipSetDevice(ctx->device));
  const int B = ctx->bins;
  const int64_t n = sample_count(ctx);
  const int64_t P = (int64_t)ctx->mtf.params.size();
  if (!sb) MMI_HIP(ctx->out.ensure(out_bytes(P)));
  b = sb ? *sb : ctx_bufs(ctx);
  std::memset(&a, 0, sizeof(a));
  a.virt = to_dev(ctx->virtual_g);
  a.fixed = to_dev(ctx->fixed_g);
  a.moving = to_dev(ctx->moving_g);
  a.ftf = ctx->dftf;
  a.mtf = ctx->dmtf;
  a.n_init = (int32_t)(ctx->init_stages.size() / 12);
  a.init_stages = a.n_init ? ctx->init_dev.as<double>() : nullptr;
  a.fixed_px = fixed_pixels(ctx);
  a.moving_px = moving_pixels(ctx);
  a.texels = texel_pixels(ctx);
  a.storage = ctx->active_storage;
  a.fixed_kind = ctx->fixed_kind;
  a.moving_kind = ctx->moving_kind;
  a.n = n;
  a.points = ctx->n_points > 0 ? ctx->points.as<double>() : nullptr;
  a.point_ids = ctx->n_points > 0 ? ctx->point_ids.as<uint32_t>() : nullptr;
  a.fixed_mask = ctx->have_mask[0] ? ctx->mask_desc[0].as<DevMask>() : nullptr;
  a.moving_mask = ctx->have_mask[1] ? ctx->mask_desc[1].as<DevMask>() : nullptr;
  a.bins = B;
  a.hist_in_lds = (size_t)B * B * 8 + (B + 1) * 4 <= 64 * 1024 ? 1 : 0;
  a.gradient_source = ctx->gradient_source;
  a.moving_min = ctx->moving_min;
  a.moving_max = ctx->moving_max;
  a.moving_bin_size = ctx->moving_bin_size;
  a.moving_norm_min = ctx->moving_norm_min;
  a.fixed_bin_size = ctx->fixed_bin_size;
  a.fixed_norm_min = ctx->fixed_norm_min;
  a.hist = b.hist;
  a.nvalid = b.nvalid;
  if (ctx->flags & 1u) {
    a.sample_fbin = ctx->sample_fbin.as<int16_t>();
    a.sample_mbin = ctx->sample_mbin.as<int16_t>();
  }
  a.pratio = b.pratio;
  a.pratio_f = b.pratio_f;
  a.pratio_in_lds = (size_t)B * B * 4 + 512 <= 64 * 1024 ? 1 : 0;
  if (want_derivative && ctx->gradient_source == MMI_GRADIENT_IMAGE && !ctx->have_texels)
    return fail(MMI_ERR_ARGUMENT, "moving gradient image missing");

  fast = setup_fast(ctx, &fa);
  ctx->last_path = fast ? 1 : 0;
  if (fast) {
    if (!sb) {
      MMI_HIP(ctx->exc_count.ensure(sizeof(uint32_t) * ctx->grid));
      MMI_HIP(ctx->exc_list.ensure(sizeof(uint32_t) * ctx->grid * kExcCap));
      b.exc_count = ctx->exc_count.as<uint32_t>();
      b.exc_list = ctx->exc_list.as<uint32_t>();
    }
    if (ctx->mtf.kind == MMI_TF_BSPLINE) {
      MMI_HIP(ctx->exc_res.ensure(sizeof(ExcResult) * ctx->grid * kExcCap));
      fa.exc_res = ctx->exc_res.as<ExcResult>();
    }
    fa.xcd_chunks = ctx->xcd_chunks;
    fa.exc_cap = kExcCap;
    fa.hist_copies = (ctx->hist_copies == 1 || ctx->hist_copies == 2 || ctx->hist_copies == 4) &&
                             pass1_lds_bytes_with(B, ctx->hist_copies) <= (size_t)ctx->max_lds_per_block
                         ? ctx->hist_copies
                         : hist_copies(B);
    fa.fixed_px = a.fixed_px;
    fa.moving_px = a.moving_px;
    fa.texels = a.texels;
    fa.storage = a.storage;
    fa.hist = a.hist;
    fa.nvalid = a.nvalid;
    fa.sample_fbin = a.sample_fbin;
    fa.sample_mbin = a.sample_mbin;
    fa.pratio_f = a.pratio_f;
    fa.exc_count = b.exc_count;
    fa.exc_list = b.exc_list;
    fa.exc_overflow = b.overflow();
  }
  if (fast && want_derivative) {
    if (!sb) {
      MMI_HIP(ctx->partials.ensure(sizeof(double) * P * 2 * ctx->grid));
      b.partials = ctx->partials.as<double>();
    }
    fa.partials = b.partials;
  }
  if (fast && want_derivative && fa.rows.on && fa.rows.cd_on) {
    if (sb) {
      fa.rows.cd_on = 0;  // the static entry takes the gradient image only
    } else {
      // a wavefront's share of the walks (rows_range): its items times the
      // samples of an item bound the samples it can defer
      const FastRows& R = fa.rows;
      const uint64_t waves = (uint64_t)ctx->grid * 4;
      uint64_t per_wave = (R.items + waves - 1) / waves;
      if (fa.xcd_chunks) {
        const uint64_t chunk = (R.items + kXcd - 1) / kXcd;
        const uint64_t wx = (uint64_t)(ctx->grid / kXcd) * 4;
        per_wave = std::max<uint64_t>(per_wave, (chunk + wx - 1) / wx);
      }
      const uint64_t cap = per_wave * 64 * (uint64_t)R.walk;
      // the region, then the per-thread counts
      const size_t region = sizeof(uint32_t) * cap * waves;
      MMI_HIP(ctx->cd_defer.ensure(region + sizeof(uint32_t) * waves * 64));
      fa.cd_defer = ctx->cd_defer.as<uint32_t>();
      fa.cd_nd = fa.cd_defer + cap * waves;
      fa.cd_wave_cap = (uint32_t)cap;
    }
  }
  // Row walks over an unmasked dense domain take the fixed bins from the
  // level's byte map (filled here once per level; a kernel launch, no
  // allocation).  Pass 2 then classifies on its own and takes the moving
  // values from the texels, which must hold the level's moving image.
  if (fast && fa.mk == 1 && fa.rows.on && fb_rows_ok(ctx)) {
    if (!ctx->fbin8_valid) {
      MMI_HIP(launch_fixed_bins(a, ctx->fbin8.as<uint8_t>(), (uint32_t)n, std::max(8, ctx->grid), ctx->stream));
      ctx->fbin8_valid = true;
    }
    fa.fbin8 = ctx->fbin8.as<uint8_t>();
    fa.mpad = ctx->mpad.p;
    fa.tpad = ctx->pad_tex ? ctx->tpad.p : nullptr;
    fa.psy = (int32_t)(ctx->moving_g.size[0] + 2);
    fa.psz = fa.psy * (int32_t)(ctx->moving_g.size[1] + 2);
  }
  return MMI_OK;
}

FinalizeArgs finalize_args(const mmi_context* ctx, const EvalBufs& b) {
  FinalizeArgs f;
  std::memset(&f, 0, sizeof(f));
  f.bins = ctx->bins;
  f.hist = b.hist;
  f.nvalid = b.nvalid;
  f.pdf = b.pdf;
  f.fm = b.fm;
  f.mm = b.mm;
  f.pratio = b.pratio;
  f.pratio_f = b.pratio_f;
  f.rowsum = b.rowsum;
  f.result = b.result();
  f.overflow = b.overflow();
  f.moving_bin_size = ctx->moving_bin_size;
  return f;
}

// Index-space derivative moments -> affine parameters.
FastMap fast_map(const mmi_context* ctx) {
  const int D = ctx->virtual_g.dim;
  FastMap map;
  std::memset(&map, 0, sizeof(map));
  for (int i = 0; i < D * D; ++i) map.V[i] = ctx->virtual_g.index_to_physical[i];
  for (int i = 0; i < D; ++i) map.c0[i] = ctx->virtual_g.origin[i] - ctx->dmtf.center[i];
  return map;
}

// GetValueAndDerivative / GetValue into the buffers of `sb` (NULL: the
// context's own).  With sb the context must be on the affine fast path (the
// caller checks): the branches that grow B-spline buffers are not reached,
// the results come back through the caller's host memory instead of the
// context's pinned staging, and the accumulators are cleared first (other
// contexts on the device share the scratch).
int evaluate_impl(mmi_context* ctx, int32_t want_derivative, double* value, double* derivative, int64_t* n_valid,
                  double* joint_pdf, double* fixed_marginal, const EvalBufs* sb) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  EvalArgs a;
  FastArgs fa;
  EvalBufs b;
  bool fast = false;
  if (int rc = prepare_eval(ctx, want_derivative, a, fa, fast, sb, b)) return rc;
  const int D = ctx->fixed_g.dim;
  const int B = ctx->bins;
  const int64_t n = sample_count(ctx);
  const int64_t P = (int64_t)ctx->mtf.params.size();
  hipStream_t st = ctx->stream;
  // An event record costs the queue a few microseconds (a timestamp with a
  // release), so pass 1 and pass 2 are timed apart only on request.
  const bool time_passes = (ctx->flags & 4u) != 0;
  // the fast affine derivative kernel wrote the result block to h_res
  bool direct_out = false;
retry:
  direct_out = false;
  MMI_HIP(hipEventRecord(ctx->ev[0], st));
  // pass 1's accumulators (histogram, count, overflow word) are zero between
  // evaluations of a context: finalize leaves them so; a clear kernel runs
  // only after initialisation, an interrupted evaluation or in the scratch
  if (sb || !ctx->acc_clean)
    MMI_HIP(launch_clear_eval(b.hist, B * B, b.nvalid, b.overflow(), st));
  ctx->acc_clean = false;
  if (fast && fa.mk == 2) {
    AxisTable rt[3];
    for (int k = 0; k < 3; ++k) {
      rt[k].start = fa.bs_start[k];
      rt[k].w = fa.bs_w[k];
      rt[k].inside = fa.bs_inside[k];
      rt[k].lo = rt[k].hi = nullptr;
    }
    MMI_HIP(launch_bspline_rows(ctx->dmtf.coeffs, ctx->dmtf.pdim, ctx->dmtf.grid.size[0], ctx->dmtf.grid.size[1],
                                D == 3 ? ctx->dmtf.grid.size[2] : 1, D, (int)a.virt.size[1],
                                D == 3 ? (int)a.virt.size[2] : 1, rt, ctx->bs_rows_t1.as<double>(),
                                ctx->bs_rows.as<double>(), st));
  }
  // pass 1 leaves a byte per sample (fixed bin or skip) for pass 2, which
  // then also takes the moving values from the texels it gathers for the
  // gradient: their value channel must be the level's moving image
  if (fast && want_derivative && B <= 253 && !(ctx->flags & MMI_FLAG_PASS2_RECLASSIFY) &&
      (ctx->texel_values_level || (fa.rows.on && fa.rows.cd_on)) && !(fa.fbin8 && !fa.rows.cd_on)) {
    if (b.cls) {
      fa.cls = b.cls;
    } else {
      MMI_HIP(ctx->cls.ensure((size_t)n));
      fa.cls = ctx->cls.as<uint8_t>();
    }
  }
  // a row-walk pass 1 needs a pass 2 that follows its byte per sample (or the
  // exact pass 2 of central differences, which classifies on its own)
  if (fast && fa.rows.on && want_derivative && !fa.cls && !fa.fbin8 && ctx->gradient_source == MMI_GRADIENT_IMAGE)
    fa.rows.on = 0;
  if (fast) {
    MMI_HIP(launch_fast_pass1(fa, a, D, ctx->grid, st));
    if (fa.mk == 2) MMI_HIP(launch_fast_exceptions1(a, fa, ctx->grid, st));
  } else {
    MMI_HIP(launch_pass1(a, ctx->grid, st));
  }
  if (time_passes || !want_derivative) MMI_HIP(hipEventRecord(ctx->ev[1], st));

  const FinalizeArgs f = finalize_args(ctx, b);
  MMI_HIP(launch_finalize(f, st));

  AxisTable tabs[3];
  std::memset(tabs, 0, sizeof(tabs));
  bool separable = false;
  if (want_derivative && ctx->mtf.kind == MMI_TF_BSPLINE)
    if (int rc = bspline_separable(ctx, a.virt, D, tabs, &separable)) return rc;
  if (want_derivative) {
    if (time_passes) MMI_HIP(hipEventRecord(ctx->ev[2], st));
    // the fast pass 2 interpolates the gradient image, or takes central
    // differences over row walks after a pass 1 that left its byte per
    // sample; other central differences run the exact pass 2 after the fast
    // pass 1
    const bool fast2 = fast && (ctx->gradient_source == MMI_GRADIENT_IMAGE || (fa.rows.on && fa.rows.cd_on && fa.cls));
    if (fast2 && fa.mk == 2) {
      // B-spline fast path: per-sample factors, then the separable contraction
      MMI_HIP(ctx->bs_g.ensure(sizeof(float) * D * n));
      fa.bspline_g = ctx->bs_g.as<float>();
      MMI_HIP(launch_fast_pass2(fa, a, D, ctx->grid, st));  // its tail writes the listed samples' factors
      if (int rc = bspline_contract(ctx, a.virt, tabs, f.result, b.deriv())) return rc;
      MMI_HIP(hipEventRecord(ctx->ev[3], st));
    } else if (fast2) {
      MMI_HIP(launch_fast_pass2(fa, a, D, ctx->grid, st));
      // the context's evaluation: the derivative kernel also writes the
      // result block into coherent pinned memory (no copy after it)
      if (!sb && !ctx->h_res_dev) {
        if (ctx->h_res) (void)hipHostFree(ctx->h_res);
        ctx->h_res = nullptr;
        MMI_HIP(hipHostMalloc(&ctx->h_res, out_bytes(kMaxOutParams), hipHostMallocCoherent));
        ctx->h_res_bytes = out_bytes(kMaxOutParams);
        void* dp = nullptr;
        MMI_HIP(hipHostGetDevicePointer(&dp, ctx->h_res, 0));
        ctx->h_res_dev = static_cast<char*>(dp);
      }
      direct_out = !sb && (int64_t)P <= kMaxOutParams;
      MMI_HIP(launch_fast_derivative(fa.partials, 2 * ctx->grid, fast_map(ctx), D, f.result, b.deriv(),
                                     direct_out ? ctx->h_res_dev : nullptr, st));
      MMI_HIP(hipEventRecord(ctx->ev[3], st));
    } else if (ctx->mtf.kind == MMI_TF_AFFINE) {
      if (!sb) {
        MMI_HIP(ctx->partials.ensure(sizeof(double) * P * ctx->grid));
        b.partials = ctx->partials.as<double>();
      }
      a.partials = b.partials;
      MMI_HIP(launch_pass2(a, ctx->grid, st));
      MMI_HIP(launch_derivative_reduce(a.partials, ctx->grid, (int)P, f.result, b.deriv(), st));
      MMI_HIP(hipEventRecord(ctx->ev[3], st));
    } else if (separable) {
      // Separable geometry: the B-spline Jacobian weight of node (a, b, c) at
      // virtual index (x, y, z) is wx[x][a] wy[y][b] wz[z][c], so the
      // derivative is a tensor-product contraction of the per-sample factors,
      // done axis by axis without atomics.
      MMI_HIP(ctx->bs_g.ensure(sizeof(float) * D * n));
      a.bspline_g = ctx->bs_g.as<float>();
      MMI_HIP(launch_pass2(a, ctx->grid, st));
      if (int rc = bspline_contract(ctx, a.virt, tabs, f.result, b.deriv())) return rc;
      MMI_HIP(hipEventRecord(ctx->ev[3], st));
    } else {
      MMI_HIP(hipMemsetAsync(b.deriv(), 0, sizeof(double) * P, st));
      a.deriv_accum = b.deriv();
      MMI_HIP(launch_pass2(a, ctx->grid, st));
      MMI_HIP(launch_derivative_scale(b.deriv(), P, f.result, st));
      MMI_HIP(hipEventRecord(ctx->ev[3], st));
    }
  }
  // result, overflow flag and derivative: one copy of the output block
  // (through the context's pinned staging; the static variant copies into a
  // block on its stack and the caller's arrays)
  const size_t ob = out_bytes(want_derivative && derivative ? P : 0);
  DevResult r;
  if (sb) {
    MMI_HIP(hipMemcpyAsync(&r, b.out, sizeof(DevResult), hipMemcpyDeviceToHost, st));
    if (want_derivative && derivative)
      MMI_HIP(hipMemcpyAsync(derivative, b.deriv(), sizeof(double) * P, hipMemcpyDeviceToHost, st));
  } else if (!direct_out) {
    if (int rc = ensure_host_pinned(&ctx->h_out, &ctx->h_out_bytes, out_bytes(P))) return rc;
    MMI_HIP(hipMemcpyAsync(ctx->h_out, b.out, ob, hipMemcpyDeviceToHost, st));
  }
  if (joint_pdf) MMI_HIP(hipMemcpyAsync(joint_pdf, b.pdf, sizeof(double) * B * B, hipMemcpyDeviceToHost, st));
  if (fixed_marginal)
    MMI_HIP(hipMemcpyAsync(fixed_marginal, b.fm, sizeof(double) * B, hipMemcpyDeviceToHost, st));
  MMI_HIP(hipStreamSynchronize(st));
  ctx->acc_clean = !sb;  // finalize ran (it zeroes the accumulators it read)
  if (!sb) {
    const char* hout = static_cast<const char*>(direct_out ? ctx->h_res : ctx->h_out);
    std::memcpy(&r, hout, sizeof(DevResult));
    if (want_derivative && derivative)
      std::memcpy(derivative, hout + sizeof(DevResult) + 16, sizeof(double) * P);
  }
  const bool overflow = r.exc_overflow != 0;
  if (overflow) {
    // more uncertain samples in one workgroup than its list holds: redo exactly
    fast = false;
    ctx->last_path = 2;
    goto retry;
  }
  ctx->last_pass1_ms = ctx->last_pass2_ms = -1.0f;
  if (time_passes) {
    float t1 = 0, t2 = 0;
    MMI_HIP(hipEventElapsedTime(&t1, ctx->ev[0], ctx->ev[1]));
    ctx->last_pass1_ms = t1;
    ctx->last_pass2_ms = 0;
    if (want_derivative) {
      MMI_HIP(hipEventElapsedTime(&t2, ctx->ev[2], ctx->ev[3]));
      ctx->last_pass2_ms = t2;
    }
  }
  float tt = 0;
  MMI_HIP(hipEventElapsedTime(&tt, ctx->ev[0], want_derivative ? ctx->ev[3] : ctx->ev[1]));
  ctx->last_total_ms = tt;
  if (n_valid) *n_valid = (int64_t)r.n_valid;
  if (r.status == 1)
    return fail(MMI_ERR_DEGENERATE,
                "All samples map outside moving image buffer. The images do not sufficiently overlap. They need to "
                "be initialized to have more overlap before this metric will work. For instance, you can align the "
                "image centers by translation.");
  if (r.status == 2) return fail(MMI_ERR_DEGENERATE, "Joint PDF summed to zero");
  if (value) *value = r.value;
  return MMI_OK;
}

}  // namespace

extern "C" {

int mmi_evaluate(mmi_context* ctx, int32_t want_derivative, double* value, double* derivative, int64_t* n_valid,
                 double* joint_pdf, double* fixed_marginal) {
  const int rc = evaluate_impl(ctx, want_derivative, value, derivative, n_valid, joint_pdf, fixed_marginal, nullptr);
  if (ctx) ctx->last_error = rc == MMI_OK ? std::string() : g_last_error;
  return rc;
}

const char* mmi_context_last_error(const mmi_context* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int mmi_evaluate_static(mmi_context* ctx, int32_t want_derivative, double* value, double* derivative,
                        int64_t* n_valid) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->initialized) return fail(MMI_ERR_ARGUMENT, "mmi_initialize must precede mmi_evaluate_static");
  unsigned char* base = g_scratch_ptr[ctx->device];
  if (!base) return fail(MMI_ERR_DEVICE, "static scratch not resolved on this device");
  const int64_t P = (int64_t)ctx->mtf.params.size();
  if (ctx->bins > kStaticMaxBins)
    return fail(MMI_ERR_ARGUMENT, "mmi_evaluate_static: more than 128 histogram bins");
  if (sample_count(ctx) > kStaticMaxSamples)
    return fail(MMI_ERR_ARGUMENT, "mmi_evaluate_static: more than 2^24 virtual samples");
  if (ctx->grid > kStaticMaxGrid)
    return fail(MMI_ERR_ARGUMENT, "mmi_evaluate_static: more than 2048 workgroups per pass");
  if (P > kStaticMaxParams) return fail(MMI_ERR_ARGUMENT, "mmi_evaluate_static: more than 12 parameters");
  // the affine fast path over the virtual voxels: no allocation on any branch
  // it can take (the exact-path redo after a list overflow included)
  const bool ok = ctx->mtf.kind == MMI_TF_AFFINE && ctx->ftf.kind == MMI_TF_IDENTITY && ctx->n_points == 0 &&
                  !ctx->have_mask[0] && !ctx->have_mask[1] && !(ctx->flags & (MMI_FLAG_EXACT_PATH_ONLY |
                                                                              MMI_FLAG_RECORD_SAMPLE_BINS)) &&
                  (!want_derivative || ctx->gradient_source == MMI_GRADIENT_IMAGE);
  if (!ok)
    return fail(MMI_ERR_ARGUMENT,
                "mmi_evaluate_static takes an affine moving transform, identity fixed transform, dense sampling "
                "without masks and the moving gradient image");
  // fixed samples on fixed pixel centres first (other geometries would build
  // per-axis cell tables in the context)
  int64_t off[3];
  FastArgs fa;
  if (!fixed_direct_offsets(ctx->virtual_g, ctx->fixed_g, off) || !setup_fast(ctx, &fa))
    return fail(MMI_ERR_ARGUMENT,
                "mmi_evaluate_static takes virtual samples on fixed pixel centres within the fast path's bounds");
  const EvalBufs sb = scratch_bufs(base);
  return evaluate_impl(ctx, want_derivative, value, derivative, n_valid, nullptr, nullptr, &sb);
}

}  // extern "C"

// Batched evaluation: one launch of each fast-path kernel for every context in
// the batch (optionally replayed as a captured HIP graph), the other contexts
// evaluated one by one while it runs.
struct mmi_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int64_t> sig;  // launch shape and buffer addresses the graph was captured with
  void* h_stage = nullptr;   // pinned: start table and argument arrays
  size_t h_stage_bytes = 0;
  double* h_out = nullptr;   // pinned: gathered results
  size_t h_out_bytes = 0;
  DevBuf d_stage, d_out;
  float last_ms = 0.0f;
  int32_t last_fused = 0;
  // the RegularStep loop on the device: chunks of `chunk` iterations, captured
  // once per launch shape and buffer set
  bool step_on_device = true;
  int32_t chunk = 8;
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  std::vector<int64_t> step_sig;
  int32_t last_device_steps = 0;  // evaluations the last mmi_batch_regular_step ran in the device loop
  int32_t* h_flags = nullptr;     // pinned: the device loop's per-member halted flags
  size_t h_flags_bytes = 0;
};

namespace {

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

int ensure_pinned(void** p, size_t* have, size_t need) {
  if (*p && *have >= need) return MMI_OK;
  if (*p) {
    MMI_HIP(hipHostFree(*p));
    *p = nullptr;
    *have = 0;
  }
  MMI_HIP(hipHostMalloc(p, need ? need : 16, hipHostMallocDefault));
  *have = need;
  return MMI_OK;
}

void drop_eval_graph(mmi_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->sig.clear();
}

void drop_step_graph(mmi_batch* b) {
  if (b->step_exec) (void)hipGraphExecDestroy(b->step_exec);
  if (b->step_graph) (void)hipGraphDestroy(b->step_graph);
  b->step_exec = nullptr;
  b->step_graph = nullptr;
  b->step_sig.clear();
}

// One capture of `work` on the stream, instantiated.  False, with no graph,
// when the capture fails: another host thread may make a call that waits for
// the whole device (a free, a device synchronisation) while the capture is
// open, which invalidates it.  The caller then runs the work without a graph
// and captures again on its next call; the failed capture's error is cleared
// (it is not the work's).
template <typename Work>
bool capture_graph(hipStream_t st, Work&& work, hipGraph_t* graph, hipGraphExec_t* exec) {
  *graph = nullptr;
  *exec = nullptr;
  bool ok = false;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    const hipError_t e = work();
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess && e2 == hipSuccess && g && hipGraphInstantiate(exec, g, nullptr, nullptr, 0) == hipSuccess) {
      *graph = g;
      ok = true;
    } else {
      if (g) (void)hipGraphDestroy(g);
      *exec = nullptr;
    }
  }
  (void)hipGetLastError();
  return ok;
}

void drop_graph(mmi_batch* b) {
  drop_eval_graph(b);
  drop_step_graph(b);
}

const char* degenerate_message(int status) {
  return status == 1 ? "All samples map outside moving image buffer. The images do not sufficiently overlap. They need "
                       "to be initialized to have more overlap before this metric will work. For instance, you can "
                       "align the image centers by translation."
                     : "Joint PDF summed to zero";
}

// Whether a context goes into a batch's fused launch (the certified affine
// fast path, no image masks, the gradient image, at most 12 parameters and,
// with the derivative, pass 1's byte per sample, as it has alone); its
// arguments in a and f.
int fused_candidate(mmi_context* c, int32_t want_derivative, EvalArgs& a, FastArgs& f, bool& ok) {
  bool fast = false;
  EvalBufs eb;
  if (int rc = prepare_eval(c, want_derivative, a, f, fast, nullptr, eb)) return rc;
  const int64_t P = (int64_t)c->mtf.params.size();
  const bool masked = f.fmask.px || f.mmask.px;
  const bool image_gradient = c->gradient_source == MMI_GRADIENT_IMAGE;
  const bool cls_ok = !want_derivative ||
                      (c->bins <= 253 && c->texel_values_level && !(c->flags & MMI_FLAG_PASS2_RECLASSIFY));
  ok = fast && f.mk == 1 && !masked && image_gradient && P <= 12 && cls_ok;
  return MMI_OK;
}

// Offsets of the fused launch's tables in the batch's staging buffers.
struct StageLayout {
  size_t o_start, o_fa, o_ea, o_fz, o_dv, o_it, o_extra, bytes;
};

// Stages the fused launch of the members ctxs[fused[k]] (arguments ea[k],
// fa[k]) in the batch's pinned buffer, sized for `extra` more bytes at
// S.o_extra, and fills L (no gather: L.out NULL, L.halted NULL).  The members'
// own streams are drained first (their uploads precede the batch's reads).
int stage_fused(mmi_batch* b, mmi_context* const* ctxs, const std::vector<int>& fused, std::vector<EvalArgs>& ea,
                std::vector<FastArgs>& fa, int32_t want_derivative, int dim, int storage, size_t extra,
                StageLayout& S, BatchLaunch& L) {
  const int nf = (int)fused.size();
  for (int k = 0; k < nf; ++k) {
    MMI_HIP(hipStreamSynchronize(ctxs[fused[k]]->stream));
    ctxs[fused[k]]->acc_clean = false;  // the batch clears and uses them on its own stream
  }
  S.o_start = 0;
  S.o_fa = align16(sizeof(int32_t) * (nf + 1));
  S.o_ea = S.o_fa + align16(sizeof(FastArgs) * nf);
  S.o_fz = S.o_ea + align16(sizeof(EvalArgs) * nf);
  S.o_dv = S.o_fz + align16(sizeof(FinalizeArgs) * nf);
  S.o_it = S.o_dv + align16(sizeof(BatchDeriv) * nf);
  S.o_extra = S.o_it + align16(sizeof(BatchOut) * nf);
  S.bytes = S.o_extra + align16(extra);
  if (int rc = ensure_pinned(&b->h_stage, &b->h_stage_bytes, S.bytes)) return rc;
  MMI_HIP(b->d_stage.ensure(S.bytes));
  char* hs = static_cast<char*>(b->h_stage);
  char* ds = static_cast<char*>(b->d_stage.p);
  int32_t* start = reinterpret_cast<int32_t*>(hs + S.o_start);
  int bins_max = 0;
  int total = 0;
  size_t lds1 = 0;
  for (int k = 0; k < nf; ++k) {
    start[k] = total;
    total += ctxs[fused[k]]->grid;  // multiples of 8
    bins_max = std::max(bins_max, ctxs[fused[k]]->bins);
    lds1 = std::max(lds1, pass1_lds_bytes_with(ctxs[fused[k]]->bins, fa[k].hist_copies));
  }
  start[nf] = total;
  // pass 1's byte per sample: every fused member can use it (fused_candidate)
  const bool cls = want_derivative != 0;
  for (int k = 0; k < nf; ++k) {
    fa[k].cls = nullptr;
    if (cls && !fa[k].fbin8) {
      mmi_context* c = ctxs[fused[k]];
      MMI_HIP(c->cls.ensure((size_t)sample_count(c)));
      fa[k].cls = c->cls.as<uint8_t>();
    }
  }
  const int32_t lds = finalize_uses_lds(bins_max) ? 1 : 0;
  for (int k = 0; k < nf; ++k) {
    mmi_context* c = ctxs[fused[k]];
    c->last_path = 1;
    std::memcpy(hs + S.o_fa + sizeof(FastArgs) * k, &fa[k], sizeof(FastArgs));
    std::memcpy(hs + S.o_ea + sizeof(EvalArgs) * k, &ea[k], sizeof(EvalArgs));
    FinalizeArgs z = finalize_args(c, ctx_bufs(c));
    z.lds = lds;
    std::memcpy(hs + S.o_fz + sizeof(FinalizeArgs) * k, &z, sizeof(z));
    BatchDeriv d;
    std::memset(&d, 0, sizeof(d));
    d.partials = fa[k].partials;
    d.rows = 2 * c->grid;
    d.map = fast_map(c);
    d.result = dev_result(c);
    d.deriv = dev_deriv(c);
    std::memcpy(hs + S.o_dv + sizeof(BatchDeriv) * k, &d, sizeof(d));
  }
  std::memset(&L, 0, sizeof(L));
  L.n = nf;
  L.dim = dim;
  L.storage = storage;
  L.want_derivative = want_derivative ? 1 : 0;
  L.total_blocks = total;
  L.bins_max = bins_max;
  L.lds1 = (int32_t)lds1;
  L.cls = cls ? 1 : 0;
  L.start = reinterpret_cast<const int32_t*>(ds + S.o_start);
  L.fa = reinterpret_cast<const FastArgs*>(ds + S.o_fa);
  L.ea = reinterpret_cast<const EvalArgs*>(ds + S.o_ea);
  L.fz = reinterpret_cast<const FinalizeArgs*>(ds + S.o_fz);
  L.dv = reinterpret_cast<const BatchDeriv*>(ds + S.o_dv);
  L.items = reinterpret_cast<const BatchOut*>(ds + S.o_it);
  return MMI_OK;
}

}  // namespace

extern "C" {

int mmi_batch_create(int device, mmi_batch** out) {
  if (!out) return fail(MMI_ERR_ARGUMENT, "out is NULL");
  int count = 0;
  MMI_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(MMI_ERR_ARGUMENT, "no such HIP device");
  MMI_HIP(hipSetDevice(device));
  auto* b = new mmi_batch();
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    mmi_batch_destroy(b);
    return fail(MMI_ERR_DEVICE, "HIP stream or event creation failed");
  }
  *out = b;
  return MMI_OK;
}

int mmi_batch_destroy(mmi_batch* b) {
  if (!b) return MMI_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  b->d_stage.release();
  b->d_out.release();
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MMI_OK;
}

int mmi_batch_set_graphs(mmi_batch* b, int32_t enabled) {
  if (!b) return fail(MMI_ERR_ARGUMENT, "batch is NULL");
  b->use_graph = enabled != 0;
  if (!b->use_graph) drop_graph(b);
  return MMI_OK;
}

int mmi_batch_evaluate(mmi_batch* b, mmi_context* const* ctxs, int32_t n, int32_t want_derivative, double* values,
                       double* const* derivatives, int64_t* n_valid, int32_t* status) {
  if (!b) return fail(MMI_ERR_ARGUMENT, "batch is NULL");
  if (n < 0 || (n > 0 && (!ctxs || !values || !n_valid || !status)))
    return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if (want_derivative && !derivatives) return fail(MMI_ERR_ARGUMENT, "derivatives is NULL");
  MMI_HIP(hipSetDevice(b->device));
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(MMI_ERR_ARGUMENT, "context is NULL");
    if (ctxs[i]->device != b->device) return fail(MMI_ERR_ARGUMENT, "contexts of a batch must share its device");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(MMI_ERR_ARGUMENT, "a context appears twice in one batch");
  }
  // classify: fast-path contexts of one dimension and storage go into the fused launch
  std::vector<int> fused, single;
  std::vector<EvalArgs> ea;
  std::vector<FastArgs> fa;
  int dim = 0, storage = -1;
  for (int i = 0; i < n; ++i) {
    EvalArgs a;
    FastArgs f;
    bool ok = false;
    if (int rc = fused_candidate(ctxs[i], want_derivative, a, f, ok)) return rc;
    if (ok && (storage < 0 || (a.virt.dim == dim && a.storage == storage))) {
      dim = a.virt.dim;
      storage = a.storage;
      fused.push_back(i);
      ea.push_back(a);
      fa.push_back(f);
    } else {
      single.push_back(i);
    }
  }
  const int nf = (int)fused.size();
  b->last_fused = nf;
  const int stride = 16;  // value, count, status, up to 12 derivative entries (+ pad)
  if (nf > 0) {
    BatchLaunch L;
    StageLayout S;
    if (int rc = stage_fused(b, ctxs, fused, ea, fa, want_derivative, dim, storage, 0, S, L)) return rc;
    const size_t bytes = S.bytes;
    if (int rc = ensure_pinned(reinterpret_cast<void**>(&b->h_out), &b->h_out_bytes, sizeof(double) * stride * nf))
      return rc;
    MMI_HIP(b->d_out.ensure(sizeof(double) * stride * nf));
    for (int k = 0; k < nf; ++k) {
      BatchOut it;
      std::memset(&it, 0, sizeof(it));
      it.result = dev_result(ctxs[fused[k]]);
      it.deriv = want_derivative ? dev_deriv(ctxs[fused[k]]) : nullptr;
      it.P = (int32_t)ctxs[fused[k]]->mtf.params.size();
      std::memcpy(static_cast<char*>(b->h_stage) + S.o_it + sizeof(BatchOut) * k, &it, sizeof(it));
    }
    const int total = L.total_blocks, bins_max = L.bins_max;
    L.out_stride = stride;
    L.out = b->d_out.as<double>();
    MMI_HIP(hipEventRecord(b->ev0, b->stream));
    auto work = [&]() -> hipError_t {
      hipError_t e = hipMemcpyAsync(b->d_stage.p, b->h_stage, bytes, hipMemcpyHostToDevice, b->stream);
      if (e == hipSuccess) e = launch_fast_batch(L, b->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_out, b->d_out.p, sizeof(double) * stride * nf, hipMemcpyDeviceToHost, b->stream);
      return e;
    };
    bool replayed = false;
    if (b->use_graph) {
      const std::vector<int64_t> sig = {nf, dim, storage, L.want_derivative, total, stride, bins_max, L.lds1, L.cls,
                                        (int64_t)bytes,
                                        (int64_t)(intptr_t)b->d_stage.p, (int64_t)(intptr_t)b->d_out.p,
                                        (int64_t)(intptr_t)b->h_stage, (int64_t)(intptr_t)b->h_out};
      if (!b->exec || sig != b->sig) {
        drop_eval_graph(b);
        if (capture_graph(b->stream, work, &b->graph, &b->exec)) b->sig = sig;
      }
      if (b->exec) {
        MMI_HIP(hipGraphLaunch(b->exec, b->stream));
        replayed = true;
      }
    }
    if (!replayed) MMI_HIP(work());
    MMI_HIP(hipEventRecord(b->ev1, b->stream));
  }
  // the other contexts, each on its own stream from its own host thread, while
  // the fused launch runs
  auto run_single = [&](int i) {
    double v = 0.0;
    int64_t nv = 0;
    status[i] = mmi_evaluate(ctxs[i], want_derivative, &v, want_derivative ? derivatives[i] : nullptr, &nv, nullptr,
                             nullptr);
    values[i] = v;
    n_valid[i] = nv;
  };
  if (single.size() == 1) {
    run_single(single[0]);
  } else if (!single.empty()) {
    std::vector<std::thread> workers;
    workers.reserve(single.size());
    for (int i : single) workers.emplace_back(run_single, i);
    for (auto& w : workers) w.join();
  }
  if (nf > 0) {
    MMI_HIP(hipStreamSynchronize(b->stream));
    MMI_HIP(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
    for (int k = 0; k < nf; ++k) {
      const int i = fused[k];
      mmi_context* c = ctxs[i];
      const double* o = b->h_out + (size_t)stride * k;
      int st = (int)o[2];
      c->last_path = 1;
      if (st & 16) {  // exception-list overflow: redo this one on its own (exact)
        double v = 0.0;
        int64_t nv = 0;
        status[i] = mmi_evaluate(c, want_derivative, &v, want_derivative ? derivatives[i] : nullptr, &nv, nullptr,
                                 nullptr);
        values[i] = v;
        n_valid[i] = nv;
        continue;
      }
      n_valid[i] = (int64_t)o[1];
      if (st != 0) {
        status[i] = fail(MMI_ERR_DEGENERATE, degenerate_message(st));
        c->last_error = degenerate_message(st);
        continue;
      }
      status[i] = MMI_OK;
      c->last_error.clear();
      values[i] = o[0];
      if (want_derivative && derivatives[i])
        std::memcpy(derivatives[i], o + 3, sizeof(double) * c->mtf.params.size());
    }
  }
  return MMI_OK;
}

int mmi_regular_step_start(mmi_regular_step_state* st, const double* scales, int32_t n_scales) {
  if (!st) return fail(MMI_ERR_ARGUMENT, "state is NULL");
  if (n_scales < 0 || n_scales > 12 || (n_scales > 0 && !scales)) return fail(MMI_ERR_ARGUMENT, "bad scales");
  std::memset(st, 0, sizeof(*st));
  st->relaxation = 1.0;
  st->n_scales = n_scales;
  st->scales_identity = 1;
  for (int i = 0; i < n_scales; ++i) {
    if (!(scales[i] > std::numeric_limits<double>::epsilon()))
      return fail(MMI_ERR_ARGUMENT, "m_Scales values must be > epsilon.");
    st->scales[i] = scales[i];
    // identity within 0.01 (itkObjectToObjectOptimizerBase.cxx:113-133)
    if (std::fabs(1.0 - scales[i]) > 0.01) st->scales_identity = 0;
  }
  return MMI_OK;
}

}  // extern "C"

namespace {

// The host loop: up to budget[i] iterations of member i, the running members
// evaluated together in each (mmi_batch_evaluate), then stepped on the host.
int regular_step_host(mmi_batch* b, mmi_context* const* ctxs, int32_t n, const mmi_regular_step_options* o,
                      mmi_regular_step_state* states, const std::vector<int32_t>& budget) {
  std::vector<mmi_context*> act;
  std::vector<int> idx;
  std::vector<double> values;
  std::vector<double*> derivs;
  std::vector<int64_t> nvalid;
  std::vector<int32_t> status;
  for (int step = 0;; ++step) {
    act.clear();
    idx.clear();
    for (int i = 0; i < n; ++i) {
      mmi_regular_step_state& s = states[i];
      if (s.stop != MMI_STOP_RUNNING || step >= budget[i]) continue;
      if (s.iteration >= o->number_of_iterations) {  // checked before evaluating (GDv4:82-89)
        s.stop = MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS;
        continue;
      }
      act.push_back(ctxs[i]);
      idx.push_back(i);
    }
    if (act.empty()) break;
    const int m = (int)act.size();
    values.assign(m, 0.0);
    nvalid.assign(m, 0);
    status.assign(m, 0);
    derivs.assign(m, nullptr);
    for (int k = 0; k < m; ++k) {
      mmi_regular_step_state& s = states[idx[k]];
      std::memcpy(s.previous_gradient, s.gradient, sizeof(s.gradient));  // swap; the evaluation overwrites
      derivs[k] = s.gradient;
    }
    if (int rc = mmi_batch_evaluate(b, act.data(), m, 1, values.data(), derivs.data(), nvalid.data(), status.data()))
      return rc;
    for (int k = 0; k < m; ++k) {
      mmi_regular_step_state& s = states[idx[k]];
      mmi_context* c = act[k];
      if (status[k] != MMI_OK) {
        s.stop = MMI_STOP_METRIC_ERROR;
        s.status = status[k];
        continue;
      }
      s.value = values[k];
      std::vector<double> params = c->mtf.params;
      const int stop = regular_step_advance(s, *o, params.data(), (int)params.size());
      s.iteration += 1;
      if (stop != MMI_STOP_RUNNING) {
        s.stop = stop;
        continue;
      }
      if (int rc = mmi_set_moving_parameters(c, params.data(), (int64_t)params.size())) return rc;
    }
  }
  return MMI_OK;
}

// The device loop, for the members that go into the fused launch with the
// derivative (fused_candidate; one dimension and storage): chunks of
// b->chunk iterations (fused evaluation, then regular_step_kernel's step and
// new map) replayed until every member has stopped, used up max_steps or
// handed back to the host.  budget[i] receives the iterations member i has
// left for the host loop (members not taken keep theirs).
int regular_step_device(mmi_batch* b, mmi_context* const* ctxs, int32_t n, const mmi_regular_step_options* o,
                        mmi_regular_step_state* states, int32_t max_steps, std::vector<int32_t>& budget) {
  b->last_device_steps = 0;
  if (max_steps <= 0) return MMI_OK;
  std::vector<int> fused;
  std::vector<EvalArgs> ea;
  std::vector<FastArgs> fa;
  int dim = 0, storage = -1;
  for (int i = 0; i < n; ++i) {
    // stopped members, and those the host loop's first check stops, stay there
    if (states[i].stop != MMI_STOP_RUNNING || states[i].iteration >= o->number_of_iterations) continue;
    EvalArgs a;
    FastArgs f;
    bool ok = false;
    if (int rc = fused_candidate(ctxs[i], 1, a, f, ok)) return rc;
    if (!ok || (storage >= 0 && (a.virt.dim != dim || a.storage != storage))) continue;
    dim = a.virt.dim;
    storage = a.storage;
    fused.push_back(i);
    ea.push_back(a);
    fa.push_back(f);
  }
  const int nf = (int)fused.size();
  if (nf == 0) return MMI_OK;
  const size_t o_ctl = align16(sizeof(StepItem) * nf);
  const size_t o_halt = o_ctl + align16(sizeof(StepCtl));
  const size_t extra = o_halt + align16(sizeof(int32_t) * nf);
  BatchLaunch L;
  StageLayout S;
  if (int rc = stage_fused(b, ctxs, fused, ea, fa, 1, dim, storage, extra, S, L)) return rc;
  char* hs = static_cast<char*>(b->h_stage) + S.o_extra;
  char* ds = static_cast<char*>(b->d_stage.p) + S.o_extra;
  StepItem* h_items = reinterpret_cast<StepItem*>(hs);
  for (int k = 0; k < nf; ++k) {
    mmi_context* c = ctxs[fused[k]];
    StepItem it;
    std::memset(&it, 0, sizeof(it));
    it.s = states[fused[k]];
    it.P = (int32_t)c->mtf.params.size();
    for (int j = 0; j < it.P; ++j) it.params[j] = c->mtf.params[j];
    it.in = affine_inputs(c, fa[k].rows);
    it.fa = reinterpret_cast<FastArgs*>(static_cast<char*>(b->d_stage.p) + S.o_fa + sizeof(FastArgs) * k);
    it.ea = reinterpret_cast<EvalArgs*>(static_cast<char*>(b->d_stage.p) + S.o_ea + sizeof(EvalArgs) * k);
    it.result = dev_result(c);
    it.deriv = dev_deriv(c);
    std::memcpy(&h_items[k], &it, sizeof(it));
  }
  StepCtl ctl;
  std::memset(&ctl, 0, sizeof(ctl));
  ctl.o = *o;
  ctl.max_steps = max_steps;
  ctl.n = nf;
  std::memcpy(hs + o_ctl, &ctl, sizeof(ctl));
  std::memset(hs + o_halt, 0, sizeof(int32_t) * nf);
  StepItem* d_items = reinterpret_cast<StepItem*>(ds);
  const StepCtl* d_ctl = reinterpret_cast<const StepCtl*>(ds + o_ctl);
  int32_t* d_halt = reinterpret_cast<int32_t*>(ds + o_halt);
  L.halted = d_halt;
  if (int rc = ensure_pinned(reinterpret_cast<void**>(&b->h_flags), &b->h_flags_bytes, sizeof(int32_t) * nf))
    return rc;
  // a chunk of at most the call's own budget (no empty iterations when the
  // caller steps in short calls)
  const int K = std::min<int>(b->chunk, max_steps);
  auto chunk = [&]() -> hipError_t {
    for (int k = 0; k < K; ++k) {
      hipError_t e = launch_fast_batch(L, b->stream);
      if (e == hipSuccess) e = launch_regular_step(d_items, d_ctl, d_halt, nf, dim, b->stream);
      if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(b->h_flags, d_halt, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, b->stream);
  };
  MMI_HIP(hipMemcpyAsync(b->d_stage.p, b->h_stage, S.bytes, hipMemcpyHostToDevice, b->stream));
  if (b->use_graph) {
    const std::vector<int64_t> sig = {nf, dim, storage, L.total_blocks, L.bins_max, L.lds1, K, (int64_t)S.bytes,
                                      (int64_t)(intptr_t)b->d_stage.p, (int64_t)(intptr_t)b->h_flags};
    if (!b->step_exec || sig != b->step_sig) {
      drop_step_graph(b);
      if (capture_graph(b->stream, chunk, &b->step_graph, &b->step_exec)) b->step_sig = sig;
    }
  }
  MMI_HIP(hipEventRecord(b->ev0, b->stream));
  const int chunks = (max_steps + K - 1) / K;
  for (int c = 0; c < chunks; ++c) {
    if (b->use_graph && b->step_exec) {
      MMI_HIP(hipGraphLaunch(b->step_exec, b->stream));
    } else {
      MMI_HIP(chunk());
    }
    MMI_HIP(hipStreamSynchronize(b->stream));
    bool all = true;
    for (int k = 0; k < nf; ++k) all = all && b->h_flags[k] != 0;
    if (all) break;
  }
  MMI_HIP(hipEventRecord(b->ev1, b->stream));
  MMI_HIP(hipMemcpyAsync(h_items, d_items, sizeof(StepItem) * nf, hipMemcpyDeviceToHost, b->stream));
  MMI_HIP(hipStreamSynchronize(b->stream));
  MMI_HIP(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  b->last_fused = nf;
  for (int k = 0; k < nf; ++k) {
    const int i = fused[k];
    const StepItem& it = h_items[k];
    states[i] = it.s;
    b->last_device_steps += it.steps;
    // the context takes the parameters the device reached (its host set-up
    // computes the map the device computed)
    if (int rc = mmi_set_moving_parameters(ctxs[i], it.params, it.P)) return rc;
    if (it.dstatus) {
      (void)fail(MMI_ERR_DEGENERATE, degenerate_message(it.dstatus));
      ctxs[i]->last_error = degenerate_message(it.dstatus);
    }
    budget[i] = it.handback != kStepOnDevice ? max_steps - it.steps : 0;
  }
  return MMI_OK;
}

}  // namespace

extern "C" {

int mmi_batch_regular_step(mmi_batch* b, mmi_context* const* ctxs, int32_t n, const mmi_regular_step_options* o,
                           mmi_regular_step_state* states, int32_t max_steps) {
  if (!b || !o || (n > 0 && (!ctxs || !states))) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if (o->relaxation_factor < 0.0)
    return fail(MMI_ERR_ARGUMENT, "Relaxation factor must be positive. Current value is " +
                                      std::to_string(o->relaxation_factor));
  if (o->relaxation_factor >= 1.0)
    return fail(MMI_ERR_ARGUMENT, "Relaxation factor must be less than 1.0. Current value is " +
                                      std::to_string(o->relaxation_factor));
  if (o->gradient_magnitude_tolerance < 0.0)
    return fail(MMI_ERR_ARGUMENT, "Gradient magnitude tolerance must be greater or equal 0.0.");
  MMI_HIP(hipSetDevice(b->device));
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(MMI_ERR_ARGUMENT, "context is NULL");
    if (ctxs[i]->device != b->device) return fail(MMI_ERR_ARGUMENT, "contexts of a batch must share its device");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(MMI_ERR_ARGUMENT, "a context appears twice in one batch");
    const size_t P = ctxs[i]->mtf.params.size();
    if (P > 12) return fail(MMI_ERR_ARGUMENT, "mmi_batch_regular_step handles transforms of at most 12 parameters");
    if (states[i].n_scales != 0 && (size_t)states[i].n_scales != P)
      return fail(MMI_ERR_ARGUMENT, "The size of Scales differs from the number of parameters");
  }
  std::vector<int32_t> budget(n, std::max<int32_t>(max_steps, 0));
  b->last_device_steps = 0;
  if (b->step_on_device)
    if (int rc = regular_step_device(b, ctxs, n, o, states, max_steps, budget)) return rc;
  return regular_step_host(b, ctxs, n, o, states, budget);
}

int mmi_batch_set_step_loop(mmi_batch* b, int32_t on_device, int32_t chunk) {
  if (!b) return fail(MMI_ERR_ARGUMENT, "batch is NULL");
  if (chunk < 1 || chunk > 64) return fail(MMI_ERR_ARGUMENT, "chunk must be in [1, 64] iterations");
  b->step_on_device = on_device != 0;
  if (chunk != b->chunk) drop_step_graph(b);
  b->chunk = chunk;
  return MMI_OK;
}

int mmi_batch_get_step_count(mmi_batch* b, int32_t* device_steps) {
  if (!b || !device_steps) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  *device_steps = b->last_device_steps;
  return MMI_OK;
}

int mmi_get_moving_parameters(mmi_context* ctx, double* params, int64_t n_params) {
  if (!ctx || !params) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if ((int64_t)ctx->mtf.params.size() != n_params)
    return fail(MMI_ERR_ARGUMENT, "parameter count differs from the moving transform's");
  std::memcpy(params, ctx->mtf.params.data(), sizeof(double) * n_params);
  return MMI_OK;
}

int mmi_batch_get_last_timing(mmi_batch* b, float* device_ms, int32_t* n_fused) {
  if (!b) return fail(MMI_ERR_ARGUMENT, "batch is NULL");
  if (device_ms) *device_ms = b->last_ms;
  if (n_fused) *n_fused = b->last_fused;
  return MMI_OK;
}

int mmi_get_sample_bins(mmi_context* ctx, int16_t* fixed_bin, int16_t* moving_bin) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!(ctx->flags & 1u)) return fail(MMI_ERR_ARGUMENT, "create the context with flag 1 to record sample bins");
  if (!ctx->initialized) return fail(MMI_ERR_ARGUMENT, "not initialized");
  const int64_t n = sample_count(ctx);
  MMI_HIP(hipSetDevice(ctx->device));
  if (fixed_bin)
    MMI_HIP(hipMemcpyAsync(fixed_bin, ctx->sample_fbin.p, sizeof(int16_t) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (moving_bin)
    MMI_HIP(hipMemcpyAsync(moving_bin, ctx->sample_mbin.p, sizeof(int16_t) * n, hipMemcpyDeviceToHost, ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  return MMI_OK;
}

int mmi_get_moving_texels(mmi_context* ctx, float* texels4) {
  if (!ctx || !texels4) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  if (!ctx->initialized) return fail(MMI_ERR_ARGUMENT, "not initialized");
  if (!ctx->have_texels) return fail(MMI_ERR_ARGUMENT, "no moving gradient image (central differences)");
  MMI_HIP(hipSetDevice(ctx->device));
  MMI_HIP(hipMemcpyAsync(texels4, ctx->texels.p, sizeof(float4) * num_pixels(ctx->moving_g), hipMemcpyDeviceToHost,
                         ctx->stream));
  MMI_HIP(hipStreamSynchronize(ctx->stream));
  return MMI_OK;
}

int mmi_get_histogram_parameters(mmi_context* ctx, double* fixed_min, double* fixed_max, double* moving_min,
                                 double* moving_max, double* fixed_bin_size, double* moving_bin_size) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (!ctx->initialized) return fail(MMI_ERR_ARGUMENT, "not initialized");
  if (fixed_min) *fixed_min = ctx->fixed_min;
  if (fixed_max) *fixed_max = ctx->fixed_max;
  if (moving_min) *moving_min = ctx->moving_min;
  if (moving_max) *moving_max = ctx->moving_max;
  if (fixed_bin_size) *fixed_bin_size = ctx->fixed_bin_size;
  if (moving_bin_size) *moving_bin_size = ctx->moving_bin_size;
  return MMI_OK;
}

int mmi_get_exception_count(mmi_context* ctx, int64_t* n_exact, int64_t* n_samples) {
  if (!ctx || !n_exact || !n_samples) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  *n_samples = ctx->initialized ? sample_count(ctx) : 0;
  *n_exact = -1;
  if (ctx->last_path != 1 || !ctx->exc_count.p) return MMI_OK;
  MMI_HIP(hipSetDevice(ctx->device));
  std::vector<uint32_t> h((size_t)ctx->grid);
  MMI_HIP(hipMemcpy(h.data(), ctx->exc_count.p, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost));
  int64_t total = 0;
  for (uint32_t c : h) total += c;
  *n_exact = total;
  return MMI_OK;
}

int mmi_measure_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs) {
  if (!gbs || bytes < 16 || reps < 1) return fail(MMI_ERR_ARGUMENT, "copy probe: bytes >= 16, reps >= 1, gbs set");
  MMI_HIP(hipSetDevice(device));
  int cus = 0;
  MMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const int64_t n16 = bytes / 16;
  DevBuf a, b;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MMI_OK;
  double best = 0.0;
  auto run = [&]() -> hipError_t {
    hipError_t e = a.ensure((size_t)n16 * 16);
    if (e == hipSuccess) e = b.ensure((size_t)n16 * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a.p, 0, (size_t)n16 * 16, st);
    const int blocks = std::max(1, cus) * 8;
    if (e == hipSuccess) e = launch_copy_probe(a.p, b.p, n16, blocks, st);  // warm-up
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      e = hipEventRecord(e0, st);
      if (e == hipSuccess) e = launch_copy_probe(a.p, b.p, n16, blocks, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && ms > 0.0f) best = std::max(best, 2.0 * (double)n16 * 16.0 / (ms * 1e-3) / 1e9);
    }
    return e;
  };
  const hipError_t e = run();
  if (e != hipSuccess) rc = fail(MMI_ERR_DEVICE, std::string("copy probe: ") + hipGetErrorString(e));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  a.release();
  b.release();
  *gbs = best;
  return rc;
}

int mmi_get_last_path(mmi_context* ctx, int32_t* path) {
  if (!ctx || !path) return fail(MMI_ERR_ARGUMENT, "NULL argument");
  *path = ctx->last_path;
  return MMI_OK;
}

int mmi_get_last_timing(mmi_context* ctx, float* total_ms, float* pass1_ms, float* pass2_ms) {
  if (!ctx) return fail(MMI_ERR_ARGUMENT, "ctx is NULL");
  if (total_ms) *total_ms = ctx->last_total_ms;
  if (pass1_ms) *pass1_ms = ctx->last_pass1_ms;
  if (pass2_ms) *pass2_ms = ctx->last_pass2_ms;
  return MMI_OK;
}

}  // extern "C"
