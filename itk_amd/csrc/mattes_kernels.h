// Kernel argument blocks and launch wrappers shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mattes_device.h"

namespace mmi {

// Image storage codes (the C ABI's MMI_F32 / MMI_BF16 / MMI_F16).
constexpr int kStorageF32 = 0;
constexpr int kStorageBF16 = 2;
constexpr int kStorageF16 = 3;

// Pixel kinds of the reference's images: the interpolated value assigned to a
// float (rounded), a double (kept) or an integer pixel (truncated).
constexpr int32_t kPixelFloat = 0;
constexpr int32_t kPixelDouble = 1;
constexpr int32_t kPixelInteger = 2;

struct DevResult {
  unsigned long long n_valid;
  int32_t status;  // 0 ok, 1 no valid sample, 2 joint PDF summed to zero
  int32_t exc_overflow;  // fast path: a pass-1 workgroup listed more samples than it holds
  double value;
  double joint_pdf_sum;
  double n_factor;  // -1 / (moving bin size * NValid)
};

// Compact arguments of the certified fast path (mattes_fast.hip): an affine
// moving transform, identity fixed transform and fixed samples on fixed pixel
// centres.  The moving continuous index is one FP64 affine map of the virtual
// index (|error| <= e_ci against the reference's chain); interpolation and
// Parzen terms run in FP32 with the error bounds below, and any sample whose
// inside test, range test or bin index could differ from the reference's
// within those bounds is listed and redone on the exact FP64 path.
// An image mask on the fast path: its continuous index is A idx + b of the
// virtual index with |error| <= e against the reference's chain; px NULL: no
// mask.
struct FastMask {
  const uint8_t* px;
  double A[9];
  double b[3];
  double e;
  int32_t start[3], end[3];
  int32_t sy, sz;
};

// Row walks of the affine fast path (dense sampling, no masks, fixed samples on
// fixed pixel centres): a wavefront takes 64 consecutive x of `walk`
// consecutive rows (one item); each lane maps its first sample with the FP64
// affine map and steps the moving continuous index down its column in 32.32
// fixed point (|error| <= e_ci including the steps).  Items are numbered x
// segment fastest, then walk, then z.
struct FastRows {
  int32_t on;
  int32_t walk;   // rows per walk
  int32_t nx, ny; // virtual size along x and y
  uint32_t items; // nxseg * nyblk * nz
  int32_t pad_;
  FastDiv dxseg;  // division by the x segments per row
  FastDiv dyblk;  // division by the walks per column
  long long qy[3];                 // A[k][1] * 2^32, rounded
  long long lo_out[3];             // below: certainly outside the buffer
  unsigned long long span_out[3];  // lo_out + span_out and above: certainly outside
  long long lo_in[3];              // [lo_in, lo_in + span_in): certainly inside
  unsigned long long span_in[3];
  // Central differences of the moving image on the fast pass 2 (cd_on; a
  // moving geometry with a diagonal physical-to-index map): along axis b the
  // reference's two points at -/+ half a spacing are the continuous index
  // -/+ 1/2 along index axis b.  Both are inside the buffer iff the index lies
  // in [start, end - 1); in fixed point [cd_lo_in, +cd_span_in) certainly
  // holds it, outside [cd_lo_out, +cd_span_out) it certainly fails, between
  // the two the reference's own FP64 points decide.
  int32_t cd_on;
  float cd_scale[3];  // sign of the index step of a physical + step, over the spacing
  long long cd_lo_in[3];
  unsigned long long cd_span_in[3];
  long long cd_lo_out[3];
  unsigned long long cd_span_out[3];
};

struct FastArgs {
  uint32_t n;  // virtual samples
  int32_t bins;
  int32_t vstart[3];
  int32_t vsize0, vsize1;
  FastDiv vdx, vdy;
  const void* fixed_px;  // storage type given by `storage` (MMI_F32 / MMI_BF16 / MMI_F16)
  int32_t fsy, fsz;
  int32_t foff[3];  // fixed_interp == 0: fixed pixel index = virtual index + foff (buffer-relative)
  int32_t fixed_interp;
  // fixed_interp == 1: per-axis fixed cells of the virtual indices, one int4 per
  // index: (lower-corner element offset, upper-corner delta or 0, FP32 bits of
  // the interpolation weight or 0, inside-buffer flag)
  const int4* ftab[3];
  int32_t storage;
  int32_t mk;  // moving transform: 1 affine (ci = A idx + b), 2 cubic B-spline
  const void* moving_px;
  const void* texels;
  int32_t mstart[3], mend[3];
  int32_t msy, msz;
  double A[9];  // virtual index -> moving continuous index
  double b[3];
  double e_ci;  // bound on |ci_fast - ci_reference|
  float cilo[3], cihi[3];
  float inv_mbs, mnm;  // 1 / moving bin size, moving normalised minimum
  float inv_fbs, fnm;
  float mmin, mmax;  // moving TrueMin / TrueMax
  float ev;          // bound on |v_fast - v_reference| for moving values
  float bt_m;        // bound on |mterm_fast - mterm_reference|
  float bt_f;        // bound on |fterm_fast - fterm_reference|
  int32_t xcd_chunks;
  int32_t exc_cap;
  int32_t hist_copies;  // pass 1's LDS histogram copies (1 or 2; see hist_copies)
  int32_t pad_hc_;
  unsigned long long* hist;
  unsigned long long* nvalid;
  int16_t* sample_fbin;
  int16_t* sample_mbin;
  const float* pratio_f;
  double* partials;  // [2 * grid * P]: fast rows then exception rows
  // rows.cd_on: per thread of pass 2 (workgroup * 256 + thread) a region of
  // cd_wave_cap / 64 samples whose central-difference points need the exact
  // test (cd_wave_cap: the bound of one wavefront's samples)
  uint32_t* cd_defer;
  uint32_t* cd_nd;  // [grid * 256] deferred samples per thread
  uint32_t cd_wave_cap;
  int32_t pad_cd_;
  uint32_t* exc_count;
  uint32_t* exc_list;
  uint32_t* exc_overflow;
  // Cubic B-spline moving transform on a separable geometry (kernels built
  // with MK == 2): per-axis support tables of the virtual indices, the
  // coefficient lines reduced over y and z for every virtual row, and the
  // moving image's physical-to-index map.  ci = P2I_m (vp + U - o_m) with
  // U_d = sum_k wx[x][k] * rows[(z, y, d)][sx + k].
  const int32_t* bs_start[3];
  const double* bs_w[3];
  const uint8_t* bs_inside[3];
  const double* bs_rows;  // [vsize2][vsize1][D][Gx]
  int32_t bs_gx;
  int32_t pad_bs_;
  double vorg[3], vsp[3];     // virtual origin and (diagonal) spacing
  double mp2i[9], morg[3];    // moving physical_to_index and origin
  float* bspline_g;           // [D][n] per-sample factors of the separable derivative
  // B-spline: the exact results of the listed samples, one per list slot,
  // written by the pass-1 exception kernel for pass 2's tail
  struct ExcResult* exc_res;
  // One byte per sample from pass 1 for pass 2 (NULL: pass 2 classifies
  // again): the fixed bin of a counted certain sample, kClsListed for a
  // sample on the exact path, kClsSkip for one not counted.
  uint8_t* cls;
  // The exact fixed Parzen bin of every virtual sample, or kClsSkip where the
  // fixed side does not count it (fixed_bins_kernel, once per level; NULL:
  // the kernels read the fixed image).  Set only for row walks over an
  // unmasked dense domain, where the moving range is the min/max of every
  // moving value the interpolation reads, so the reference's range test
  // (itkMattesMutualInformationImageToImageMetricv4GetValueAndDerivativeThreader.hxx:238-245)
  // passes for every sample inside the moving buffer: an FP64 multilinear
  // value of corners in [min, max] stays in [min, max] (each lerp
  // a + (b - a) t with exact b - a and t in [0, 1] rounds between a and b).
  const uint8_t* fbin8;
  // With fbin8: the moving values (mpad) and texels (tpad; NULL without a
  // gradient image) with a one-voxel replicated border (pad_replicate_kernel),
  // element strides psy, psz: a certainly inside sample's cell, lower corner
  // in [start - 1, end], needs no clamp there.
  const void* mpad;
  const void* tpad;
  int32_t psy, psz;
  FastMask fmask, mmask;  // fixed mask at the fixed point, moving mask at the mapped moving point
  FastRows rows;          // rows.on: the row-walk kernels (pass 2 then needs cls)
};

// XCDs of the device: the fast path gives each a contiguous share of the work.
constexpr int kXcd = 8;

constexpr uint32_t kClsListed = 0xFEu;
constexpr uint32_t kClsSkip = 0xFFu;

// Exact result of one listed B-spline sample: bins, the FP32 Parzen
// argument, the gradient at its cell; fbin < 0 when the sample is not counted
// or outside the transform's valid region.
struct ExcResult {
  int32_t fbin, mbin;
  float arg;
  float g[3];
  int32_t pad_[2];
};

// Map from index-space derivative moments to the affine parameters.
struct FastMap {
  double V[9];   // virtual index_to_physical
  double c0[3];  // virtual origin - transform centre
};

struct EvalArgs {
  DevGeom virt;
  DevGeom fixed;
  DevGeom moving;
  DevTransform ftf;
  DevTransform mtf;
  const void* fixed_px;   // storage type given by `storage`
  const void* moving_px;
  const void* texels;
  int64_t n;  // virtual samples
  // sparse sampling: n virtual-space points (D doubles each) instead of the
  // virtual voxels; NULL for dense sampling
  const double* points;
  const uint32_t* point_ids;   // the caller's index of stored point k (points are kept in spatial order)
  const DevMask* fixed_mask;   // NULL: no mask
  const DevMask* moving_mask;  // NULL: no mask
  int32_t storage;        // MMI_F32, MMI_BF16 or MMI_F16
  int32_t pad2_;
  // the reference images' pixel kinds (kPixelFloat / Double / Integer): how a
  // sample's interpolated value is assigned to the pixel type
  int32_t fixed_kind, moving_kind;
  int32_t bins;
  int32_t hist_in_lds;
  int32_t gradient_source;
  int32_t pad_;
  double moving_min, moving_max;
  double moving_bin_size, moving_norm_min;
  double fixed_bin_size, fixed_norm_min;
  unsigned long long* hist;   // [B*B] fixed point 2^-32
  unsigned long long* nvalid;
  int16_t* sample_fbin;  // optional [n]
  int16_t* sample_mbin;  // optional [n]
  const double* pratio;  // [B*B]
  const float* pratio_f; // [B*B] FP32 copy for pass 2 when it does not fit in LDS
  int32_t pratio_in_lds;
  int32_t pad1_;
  double* partials;      // [grid * P] (affine)
  double* deriv_accum;   // [P] (B-spline, non-separable geometry)
  float* bspline_g;      // [D][n] (B-spline, separable geometry): per-sample w * gradient, 0 if not counted
  // composite moving transform: n_init matrix-offset stages (12 doubles each:
  // row-major matrix, offset) before the moving transform in the
  // CompositeTransform queue; the mapped point goes through stages n_init-1 .. 0
  const double* init_stages;
  int32_t n_init;
  int32_t pad3_;
};

constexpr int kMaxInitialStages = 4;

// Support of the cubic B-spline along one axis of a separable geometry (virtual
// index_to_physical and grid physical_to_index both diagonal): for virtual
// index i along the axis, its support start, the four 1-D weights and the
// axis' part of the valid-region test; for grid node g, the virtual index range
// [lo[g], hi[g]] whose support can hold g (lo > hi when empty).
struct AxisTable {
  const int32_t* start;  // [len]
  const double* w;       // [len][4]
  const uint8_t* inside; // [len]
  const int32_t* lo;     // [nodes]
  const int32_t* hi;     // [nodes]
  int32_t increasing;    // support start nondecreasing along the axis
  int32_t pad_;
};

struct FinalizeArgs {
  int32_t bins;
  int32_t lds;  // set by launch_finalize: PDF and marginals staged in LDS
  // pass 1's accumulators: read, then left zeroed for the next evaluation
  unsigned long long* hist;
  unsigned long long* nvalid;
  uint32_t* overflow;  // exception-list overflow word: moved into the result, zeroed
  double* pdf;     // [B*B] normalised joint PDF
  double* fm;      // [B]
  double* mm;      // [B]
  double* pratio;  // [B*B]
  float* pratio_f; // [B*B]
  double* rowsum;  // [B]
  DevResult* result;
  double moving_bin_size;
};

// Batched evaluation (several contexts per launch).
struct BatchDeriv {
  const double* partials;
  int32_t rows;
  int32_t pad_;
  FastMap map;
  const DevResult* result;
  double* deriv;
};

struct BatchOut {
  const DevResult* result;
  const double* deriv;  // NULL for value-only
  int32_t P;
  int32_t pad_;
};

struct BatchLaunch {
  int32_t n;  // evaluations
  int32_t dim, storage, want_derivative;
  int32_t total_blocks;  // start[n]
  int32_t out_stride;
  int32_t bins_max;
  int32_t lds1;  // pass-1 LDS bytes: the largest pass1_lds_bytes of the members
  int32_t cls;   // 1: every member has pass 1's byte per sample (pass 2 reads it)
  int32_t pad_cls_;
  const int32_t* start;  // [n + 1] first workgroup of each evaluation
  const FastArgs* fa;    // [n]
  const EvalArgs* ea;    // [n]
  const FinalizeArgs* fz;
  const BatchDeriv* dv;
  const BatchOut* items;
  double* out;  // [n][out_stride]: value, n_valid, status, derivative (NULL: not gathered)
  const int32_t* halted;  // NULL, or [n]: nonzero skips the evaluation (every kernel of the sequence)
};

// Uncertain samples a fast-path workgroup can list (more: the evaluation is
// redone on the exact path).  The B-spline exception kernel sums the fixed
// marginal and the valid count of its listed samples in (B + 1) * 4 bytes of
// LDS and adds their histogram weights with global 64-bit atomics.
constexpr int kExcCap = 256;

// Copies of the pass-1 joint histogram in LDS (plus two counters): two when
// they fit in 64 KiB, updated by even and odd lanes.  Neighbouring samples
// (adjacent lanes) usually fall in the same bins, so this halves the
// same-address serialisation of the 64-bit LDS atomics.
__host__ __device__ inline size_t pass1_lds_bytes_with(int bins, int copies) {
  return (size_t)copies * bins * bins * 8 + 2 * 4;
}
__host__ __device__ inline int hist_copies(int bins) {
  return pass1_lds_bytes_with(bins, 2) <= 64 * 1024 ? 2 : 1;
}
__host__ __device__ inline size_t pass1_lds_bytes(int bins) { return pass1_lds_bytes_with(bins, hist_copies(bins)); }

// The finalize workgroup stages the PDF and marginals in LDS when they fit.
inline bool finalize_uses_lds(int bins) {
  return ((size_t)bins * bins + 3 * (size_t)bins) * sizeof(double) <= 48 * 1024;
}

struct RecCoeffs {
  double N0, N1, N2, N3, D1, D2, D3, D4, M1, M2, M3, M4, BN1, BN2, BN3, BN4, BM1, BM2, BM3, BM4;
};

// Destination of a recursive-Gaussian pass's final values.
constexpr int32_t kSinkDouble = 0;    // the pass's double output volume
constexpr int32_t kSinkGradient = 1;  // grad[e * dim + comp] = value / spacing
constexpr int32_t kSinkFloat = 2;     // outf[e] = (float)value
struct RecSink {
  int32_t mode;
  int32_t comp, dim;
  int32_t pad_;
  double spacing;
  double* grad;
  float* outf;
};

// One line set of a batched recursive-Gaussian launch: the lines of one pass
// of one filter chain (a volume, an axis, the pass's coefficients and sink).
struct RecJob {
  const void* in;  // float volume (the first pass of a chain) or double volume
  double* out;
  int64_t ln, stride, nlines;
  int32_t blocks;  // workgroups of this job (kRecLines lines each)
  int32_t pad_;
  RecCoeffs c;
  RecSink sink;
};
constexpr int kRecJobsMax = 4;
constexpr int kRecLines = 128;
// The jobs of one launch (the same pass of independent chains: all take float
// input, or all double).
struct RecJobs {
  RecJob j[kRecJobsMax];
  int32_t n;
};

struct Direction {
  double m[9];
};

hipError_t launch_pass1(const EvalArgs& a, int grid, hipStream_t st);
hipError_t launch_finalize(const FinalizeArgs& f, hipStream_t st);
// n finalize workgroups, item lds flags already set with finalize_uses_lds.
// halted: NULL, or one flag per evaluation (nonzero: its workgroups return
// at once; the device optimizer loop's finished members)
hipError_t launch_finalize_batch(const FinalizeArgs* fz, int n, int bins_max, const int32_t* halted, hipStream_t st);
// The whole fast-path sequence for a batch (clear, pass 1, exceptions,
// finalize, pass 2, exceptions, derivative, gather), in launch order.
hipError_t launch_fast_batch(const BatchLaunch& b, hipStream_t st);
hipError_t launch_pass2(const EvalArgs& a, int grid, hipStream_t st);
// Pass 1 of the fast path; each workgroup ends by evaluating its own uncertain
// samples on the exact path (the EvalArgs) into its LDS histogram.
hipError_t launch_fast_pass1(const FastArgs& f, const EvalArgs& a, int dim, int grid, hipStream_t st);
// The exact fixed bin of each of the n dense virtual samples (FastArgs::fbin8).
hipError_t launch_fixed_bins(const EvalArgs& a, uint8_t* out, uint32_t n, int grid, hipStream_t st);
// B-spline fast path: pass 1's listed samples on the exact path.
hipError_t launch_fast_exceptions1(const EvalArgs& a, const FastArgs& f, int grid, hipStream_t st);
// Pass 2 of the fast path; affine workgroups end with their listed samples on
// the exact path (their own partial row).
hipError_t launch_fast_pass2(const FastArgs& f, const EvalArgs& a, int dim, int grid, hipStream_t st);
// The affine derivative from the fast pass 2's partial rows (fixed-order
// reduction, then the index-space moments mapped to the parameters); with
// host_out, the result block and the derivative are also written into that
// pinned host block (out_bytes layout).
hipError_t launch_fast_derivative(const double* partials, int rows, const FastMap& map, int dim,
                                  const DevResult* r, double* deriv, char* host_out, hipStream_t st);
hipError_t launch_derivative_reduce(const double* partials, int nblocks, int P, const DevResult* r, double* deriv,
                                    hipStream_t st);
hipError_t launch_derivative_scale(double* deriv, int64_t P, const DevResult* r, hipStream_t st);
// Coefficient lines reduced over the y and z supports, per virtual row:
// rows[((z * Ny + y) * D + d) * Gx + a] = sum_c wz[z][c] sum_b wy[y][b] C_d[sz+c][sy+b][a]
// (0 for rows outside the valid region); t1: D * Gz * Ny * Gx doubles of
// scratch (3-D).
hipError_t launch_bspline_rows(const double* coeffs, int32_t pdim, int Gx, int Gy, int Gz, int D, int Ny, int Nz,
                               const AxisTable* tabs, double* t1, double* rows, hipStream_t st);
hipError_t launch_bspline_axis_table(const DevGeom& virt, const DevGeom& grid, int axis, int32_t* start, double* w,
                                     uint8_t* inside, hipStream_t st);
// out[o][g][i] = sum_l w[l][g - start[l]] * in[o][l][i] over the virtual
// indices l whose support holds node g, times scale->n_factor when scale is set.
hipError_t launch_separable_contract(const float* in_f, const double* in_d, double* out, int64_t outer, int L,
                                     int64_t inner, int G, const AxisTable& t, const DevResult* scale, hipStream_t st);
hipError_t launch_minmax(const void* px, int storage, int64_t n, float* out, int blocks, hipStream_t st);
// FP32 image -> 16-bit storage (round to nearest even), and FP32 texels -> 16-bit texels.
hipError_t launch_to_storage(const float* in, void* out, int storage, int64_t n, hipStream_t st);
// A copy of n16 16-byte elements by `blocks` workgroups of 256 (copy-rate probe).
hipError_t launch_copy_probe(const void* in, void* out, int64_t n16, int blocks, hipStream_t st);
// The buffer with a one-element replicated border on each axis (pad_replicate_kernel).
hipError_t launch_pad_replicate(const void* src, void* dst, int elem_bytes, int nx, int ny, int nz, int D,
                                hipStream_t st);
hipError_t launch_texels_to_storage(const float4* in, void* out, int storage, int64_t n, hipStream_t st);
// Independent recursive-Gaussian passes in one launch (one job per chain).
hipError_t launch_recursive_gaussian_jobs(const RecJobs& jobs, bool in_float, hipStream_t st);
hipError_t launch_clear_eval(unsigned long long* hist, int n_hist, unsigned long long* nvalid, uint32_t* overflow,
                             hipStream_t st);
// Per-workgroup min/max over the masked voxels of an image (mask) or over the
// pixels the listed virtual points round to after the fixed transform (points).
hipError_t launch_minmax_select(const void* px, int storage, const DevGeom& img, const DevMask* mask,
                                const double* points, const DevTransform& ftf, int64_t n, float* out, int blocks,
                                hipStream_t st);
hipError_t launch_pack_texels(const float* moving, const double* grad, int64_t n, int D, const Direction& dir,
                              float4* tex, hipStream_t st);

// Sets the calling thread's mmi_last_error text and returns code.
int set_last_error(int code, const char* msg);

}  // namespace mmi
