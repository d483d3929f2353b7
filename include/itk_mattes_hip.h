/*
 * C ABI of the MI355X Mattes mutual-information v4 metric.
 *
 * Plain C types only; host pointers are borrowed for the duration of a call.
 * Every function returns MMI_OK (0) or a negative status; the message of the
 * last failure on the calling thread is returned by mmi_last_error().  The
 * C++ ITK subclass converts non-zero statuses into itk::ExceptionObject with
 * the same text the reference throws.
 *
 * The entry points replace, inside ITK, everything from
 * ImageToImageMetricv4::GetValueAndDerivativeExecute down
 * (Modules/Registration/Metricsv4/include/itkImageToImageMetricv4.hxx:239-259):
 * the dense threader, ProcessVirtualPoint, Mattes ProcessPoint and the
 * post-processing in MattesMutualInformationImageToImageMetricv4::ComputeResults.
 * A context is not re-entrant: one metric instance drives it from one thread.
 */
#ifndef ITK_MATTES_HIP_H
#define ITK_MATTES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMI_OK 0
#define MMI_ERR_ARGUMENT -1   /* bad argument or call order */
#define MMI_ERR_DEVICE -2     /* HIP runtime failure */
#define MMI_ERR_DEGENERATE -3 /* the reference would throw (constant image, no overlap, ...) */
#define MMI_ERR_MEMORY -4     /* device memory exhausted */

enum mmi_pixel_type {
  MMI_F32 = 0,
  MMI_F64 = 1,
  MMI_BF16 = 2, /* device storage only (mmi_set_storage) */
  MMI_F16 = 3,  /* device storage only (mmi_set_storage) */
  MMI_U8 = 4,
  MMI_S8 = 5,
  MMI_U16 = 6,
  MMI_S16 = 7,
  MMI_U32 = 8,
  MMI_S32 = 9
};
enum mmi_transform_kind { MMI_TF_IDENTITY = 0, MMI_TF_AFFINE = 1, MMI_TF_BSPLINE = 2 };
enum mmi_gradient_source { MMI_GRADIENT_CENTRAL_DIFFERENCE = 0, MMI_GRADIENT_IMAGE = 1 };

/* Image geometry in ITK conventions (x fastest in memory).
 * physical = origin + index_to_physical * index;
 * continuous index = physical_to_index * (physical - origin).
 * index_to_physical and physical_to_index are passed as ITK computed them
 * (ImageBase::GetIndexToPhysicalPoint / GetPhysicalPointToIndex,
 * Modules/Core/Common/include/itkImageBase.h:517-531,594-604) so that the
 * device reproduces ITK's continuous indices bit for bit. */
typedef struct mmi_geometry {
  int32_t dim; /* 2 or 3 */
  int32_t pad_;
  int64_t size[3];
  int64_t start[3];
  double origin[3];
  double spacing[3];
  double direction[9];
  double index_to_physical[9];
  double physical_to_index[9];
} mmi_geometry;

typedef struct mmi_context mmi_context;

/* Last error message of the calling thread (never NULL). */
const char* mmi_last_error(void);

/* Context flags. */
#define MMI_FLAG_RECORD_SAMPLE_BINS 1u /* keep per-sample Parzen bins for mmi_get_sample_bins */
#define MMI_FLAG_EXACT_PATH_ONLY 2u    /* evaluate every sample on the FP64 path (no certified FP32 fast path) */
#define MMI_FLAG_TIME_PASSES 4u        /* time pass 1 and pass 2 apart (two more events per evaluation) */
#define MMI_FLAG_PASS2_RECLASSIFY 8u   /* fast path: pass 2 classifies every sample again instead of reading
                                          the byte per sample pass 1 leaves (same results; comparison runs) */

/* Create a context on a HIP device with the flags above (0 for defaults). */
int mmi_context_create(int device, uint32_t flags, mmi_context** out);
int mmi_context_destroy(mmi_context* ctx);

/* Images: ImageToImageMetricv4::SetFixedImage / SetMovingImage
 * (itkImageToImageMetricv4.h:362-390).  pixels are host memory in buffer
 * order; pixel_type is the reference image's pixel type (MMI_F32, MMI_F64 or
 * one of the integer types MMI_U8 ... MMI_S32).  A sample's interpolated value
 * takes the reference's assignment to that pixel type
 * (itkImageToImageMetricv4.hxx:318,369): rounded to float for MMI_F32, kept
 * for MMI_F64, truncated toward zero for the integer types.  The device stores
 * FP32 (or the storage of mmi_set_storage), so MMI_F64 and 32-bit integer
 * pixels must be exact in float (else MMI_ERR_ARGUMENT: no silent rounding);
 * images of other pixel types than MMI_F32 take no
 * smoothing (mmi_set_smoothing_sigmas: the reference keeps the smoothed image
 * in the pixel type), and integer pixels evaluate on the FP64 path. */
int mmi_set_fixed_image(mmi_context* ctx, const mmi_geometry* g, int32_t pixel_type, const void* pixels);
int mmi_set_moving_image(mmi_context* ctx, const mmi_geometry* g, int32_t pixel_type, const void* pixels);

/* Moving-image gradient: dim doubles per voxel in buffer order, as
 * GradientRecursiveGaussianImageFilter produces it
 * (ImageToImageMetricv4::ComputeMovingImageGradientFilterImage,
 * itkImageToImageMetricv4.hxx:448-455). */
int mmi_set_moving_gradient(mmi_context* ctx, const double* gradient);

/* Compute the moving-image gradient on the device with the recursive Gaussian
 * derivative filter, sigma = maximum spacing, normalised across scale
 * (ImageToImageMetricv4::InitializeDefaultMovingImageGradientFilter,
 * itkImageToImageMetricv4.hxx:487-503). */
int mmi_compute_moving_gradient(mmi_context* ctx);

/* MMI_GRADIENT_IMAGE (UseMovingImageGradientFilter on, the metric default) or
 * MMI_GRADIENT_CENTRAL_DIFFERENCE (filter off: CentralDifferenceImageFunction). */
int mmi_set_gradient_source(mmi_context* ctx, int32_t source);

/* Device storage of fixed/moving/gradient images: MMI_F32 (default),
 * MMI_BF16 or MMI_F16 (round to nearest even at mmi_initialize; FP16 holds
 * |v| <= 65504).  With 16-bit storage every read of the images (intensity
 * range, bins, interpolation, gradient) sees the rounded values; Parzen
 * weights stay FP32 accumulated in fixed point, coordinates, normalisation and
 * the entropy sum FP64.  Takes effect at the next mmi_initialize.  No
 * counterpart in the reference (its images are the pixel type itself). */
int mmi_set_storage(mmi_context* ctx, int32_t pixel_type);

/* Per-level smoothing of the fixed and moving images, as the registration
 * method prepares each level (itkImageRegistrationMethodv4.hxx:591-637 with
 * SmoothingRecursiveGaussianImageFilter): physical sigma per axis, NULL or all
 * zero for none.  The uploaded images stay resident; each mmi_initialize
 * smooths them on the device, and the gradient filter, intensity range and
 * 16-bit storage use the smoothed images. */
int mmi_set_smoothing_sigmas(mmi_context* ctx, const double* fixed_sigma, const double* moving_sigma);

/* Virtual domain (ObjectToObjectMetric::SetVirtualDomain,
 * Modules/Numerics/Optimizersv4/include/itkObjectToObjectMetric.hxx:210-244).
 * If never called, the fixed image's geometry is used. */
int mmi_set_virtual_domain(mmi_context* ctx, const mmi_geometry* g);

/* Sparse sampling over a virtual-space point set
 * (ImageToImageMetricv4::SetVirtualSampledPointSet with
 * SetUseSampledPointSet(true) and SetUseVirtualSampledPointSet(true),
 * itkImageToImageMetricv4.h:100-101,484; the sparse threader iterates the
 * points in order, itkImageToImageMetricv4GetValueAndDerivativeThreader.hxx:44-63).
 * points: n x dim doubles (x, y[, z]) in point order; n = 0 returns to dense
 * sampling over the virtual domain.  The fixed intensity range is then taken
 * over the fixed pixels the points round to after the fixed transform, and the
 * fixed mask does not enter it (itkMattesMutualInformationImageToImageMetricv4.hxx:102-152).
 * Sample i of mmi_get_sample_bins is point i.  Takes effect at the next
 * mmi_initialize. */
int mmi_set_virtual_points(mmi_context* ctx, int32_t dim, const double* points, int64_t n);

/* Fixed and moving image masks (ImageToImageMetricv4::SetFixedImageMask /
 * SetMovingImageMask, itkImageToImageMetricv4.h:420-438) as
 * ImageMaskSpatialObject with its default settings: a point is inside when its
 * rounded index (RoundHalfIntegerUp) lies in the mask's buffer and the mask
 * pixel there is non-zero (itkImageMaskSpatialObject.hxx:36-48).  The fixed
 * mask is tested at the fixed-space point, the moving mask at the mapped
 * moving point (itkImageToImageMetricv4.hxx:298-309,351-362); without a point
 * set, the fixed (moving) intensity range is taken over the voxels whose
 * physical points lie inside the fixed (moving) mask
 * (itkMattesMutualInformationImageToImageMetricv4.hxx:155-220).  g: the mask
 * image's geometry, pixels: one byte per voxel in x-fastest order; g == NULL
 * removes the mask.  Takes effect at the next mmi_initialize. */
int mmi_set_fixed_mask(mmi_context* ctx, const mmi_geometry* g, const uint8_t* pixels);
int mmi_set_moving_mask(mmi_context* ctx, const mmi_geometry* g, const uint8_t* pixels);

/* Metric sampling strategies (ImageRegistrationMethodv4Enums::MetricSamplingStrategy,
 * itkImageRegistrationMethodv4.h:51-56). */
#define MMI_SAMPLING_NONE 0
#define MMI_SAMPLING_REGULAR 1
#define MMI_SAMPLING_RANDOM 2

/* ImageRegistrationMethodv4::SetMetricSamplePoints for one level
 * (itkImageRegistrationMethodv4.hxx:867-1024), on the host, no device work:
 * the virtual-space sample points for mmi_set_virtual_points.  REGULAR keeps
 * the first voxel of the virtual region and then every ceil(1/percentage)-th
 * in buffer order; RANDOM draws (unsigned long)((float)N * percentage) voxels
 * with replacement (ImageRandomConstIteratorWithIndex).  Each point is moved
 * by GetNormalVariate() * spacing / 3 per axis and kept only when the fixed
 * mask (mask_geom / mask as in mmi_set_fixed_mask, or NULL) holds it.
 * seed: in/out, the method's m_CurrentRandomSeed; the perturbation stream is
 * seeded with *seed and, for RANDOM, the voxel stream with *seed + 1, and
 * *seed advances by the seeds used (1 or 2).  points: capacity x dim doubles;
 * with points == NULL, *n_points receives the capacity needed (the count
 * before the mask) and nothing else happens. */
int mmi_metric_sample_points(const mmi_geometry* virt, int32_t strategy, double percentage, uint32_t* seed,
                             const mmi_geometry* mask_geom, const uint8_t* mask, double* points, int64_t capacity,
                             int64_t* n_points);

/* MattesMutualInformationImageToImageMetricv4::SetNumberOfHistogramBins
 * (clamped to >= 5, itkMattesMutualInformationImageToImageMetricv4.h:146). */
int mmi_set_number_of_histogram_bins(mmi_context* ctx, int32_t bins);

/* MattesMutualInformationImageToImageMetricv4::Initialize
 * (itkMattesMutualInformationImageToImageMetricv4.hxx:70-327): intensity
 * ranges (whole buffers, masked voxels or sampled points, as set above),
 * degenerate-range errors with the reference's messages, bin size and
 * normalised minimum.  MMI_ERR_DEGENERATE when the point set or a mask leaves
 * no pixel to take a range over (the reference goes on with an inverted
 * range). */
int mmi_initialize(mmi_context* ctx);

/* Fixed transform (identity by default) and moving transform.
 * MMI_TF_AFFINE:  params = matrix row-major then translation (P = D*D + D),
 *                 fixed_params = centre (D values).
 * MMI_TF_BSPLINE: params = coefficients, dimension-major (P = D * prod(grid)),
 *                 fixed_params = ITK's BSplineTransform fixed parameters:
 *                 grid size (D), grid origin (D), grid spacing (D), grid
 *                 direction (D*D); optionally followed by the grid's
 *                 physical-to-index matrix (D*D) as ITK computed it. */
int mmi_set_fixed_transform(mmi_context* ctx, int32_t kind, const double* fixed_params, int32_t n_fixed,
                            const double* params, int64_t n_params);
int mmi_set_moving_transform(mmi_context* ctx, int32_t kind, const double* fixed_params, int32_t n_fixed,
                             const double* params, int64_t n_params);
/* Replace only the moving transform's parameters (optimizer step). */
int mmi_set_moving_parameters(mmi_context* ctx, const double* params, int64_t n_params);
/* A composite moving transform (CompositeTransform with only its most recent
 * transform optimized, as ImageRegistrationMethodv4 builds it from a moving
 * initial transform: itkImageRegistrationMethodv4.hxx:344-355): n
 * matrix-offset transforms (MatrixOffsetTransformBase, y = matrix * x +
 * offset) before the moving transform in the composite's queue, in queue
 * order.  A point goes through the moving transform first, then through
 * stages n-1 .. 0 (itkCompositeTransform.hxx:60-71); the derivative takes the
 * composite's Jacobian, the stages' matrices times the moving transform's
 * parameter Jacobian (itkCompositeTransform.hxx:467-560).  matrices: n
 * row-major D x D, offsets: n x D; n = 0 (the default) removes them, at most
 * 4.  Set after the images.  Such contexts evaluate on the exact FP64 path. */
int mmi_set_moving_initial_transforms(mmi_context* ctx, int32_t n, const double* matrices, const double* offsets);

/* GetValueAndDerivative (want_derivative = 1) or GetValue (0):
 * ImageToImageMetricv4::GetValueAndDerivative (itkImageToImageMetricv4.hxx:217-231).
 * value = -MI, derivative[P] = -d value / d p (the reference's sign),
 * n_valid = NumberOfValidPoints.  joint_pdf (B*B, normalised) and
 * fixed_marginal (B) are optional (NULL to skip). */
int mmi_evaluate(mmi_context* ctx, int32_t want_derivative, double* value, double* derivative, int64_t* n_valid,
                 double* joint_pdf, double* fixed_marginal);

/* Launch configuration of the evaluation kernels (no counterpart in the
 * reference; tuning and comparison runs): workgroups of 256 per compute unit
 * (the grid holds at most one workgroup per 256 samples) and the rows per walk
 * of the affine row-walk kernels (0: the sample-strided kernels).  A negative
 * value restores the library's choice.  Takes effect at the next
 * mmi_initialize. */
int mmi_set_launch_config(mmi_context* ctx, int32_t blocks_per_cu, int32_t rows_walk);
/* Pass 1's LDS histogram copies (lane i adds to copy i mod copies): 1, 2, 4,
 * or 0 for the library's choice.  No counterpart in the reference (tuning). */
int mmi_set_histogram_copies(mmi_context* ctx, int32_t copies);
/* The launch configuration in effect after mmi_initialize; from_table: 1 when
 * the tuning table chose the settings the caller left to the library. */
int mmi_get_launch_config(mmi_context* ctx, int32_t* blocks_per_cu, int32_t* rows_walk, int32_t* hist_copies,
                          int32_t* from_table);
/* Launch-configuration table (no counterpart in the reference): text lines
 *   kind storage dim gradient bins log2_samples blocks_per_cu rows_walk hist_copies
 * (kind affine|bspline, storage f32|bf16|f16, gradient image|cd; '#' starts
 * a comment), as the tuner itk_amd/tune.py writes them from timed sweeps.
 * mmi_initialize takes, for the settings neither mmi_set_launch_config,
 * mmi_set_histogram_copies nor the environment chose, the row of the
 * context's transform kind, storage, dimension and gradient source with the
 * smallest bins >= the context's and the log2 sample count nearest to its
 * own; no row: the library's defaults.  The table is read at the first
 * mmi_initialize from $MMI_TUNING_TABLE or tuning/mi355x.tsv beside the
 * library; this call replaces it (path NULL: that default).  n_rows: rows
 * read (may be NULL). */
int mmi_load_tuning_table(const char* path, int32_t* n_rows);

/* mmi_evaluate without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * buffer the evaluation writes besides value, derivative and n_valid lives in
 * one static device array per GPU (about 19 MiB, resolved at
 * mmi_context_create), so the call makes no hipMalloc, hipFree or pinned-host
 * allocation.  Limits: an initialised context on the certified affine fast
 * path (affine moving transform, identity fixed transform, dense sampling of
 * virtual voxels on fixed pixel centres, no masks or point set, the moving
 * gradient image for the derivative, no MMI_FLAG_RECORD_SAMPLE_BINS or
 * MMI_FLAG_EXACT_PATH_ONLY), at most 2^24 virtual samples, 128 histogram
 * bins, 12 parameters and 2048 workgroups per pass; anything else returns
 * MMI_ERR_ARGUMENT.  The array is shared by every context on the GPU and
 * keeps the last call's contents: calls that overlap on one GPU (several host
 * threads, or contexts whose streams run concurrently) race.  Results equal
 * mmi_evaluate's on the same context. */
int mmi_evaluate_static(mmi_context* ctx, int32_t want_derivative, double* value, double* derivative,
                        int64_t* n_valid);

/* Per-sample Parzen bins of the last evaluation (int16, -1 = sample not
 * counted), in virtual buffer order; for parity tests.  Either may be NULL. */
int mmi_get_sample_bins(mmi_context* ctx, int16_t* fixed_bin, int16_t* moving_bin);

/* The moving texels the evaluation interpolates, FP32 (value, gradient x, y,
 * z; 2-D: z = 0) per moving pixel in buffer order: the value channel is the
 * level's moving image (smoothed when the level smooths it), the gradient the
 * output of GradientRecursiveGaussianImageFilter
 * (itkGradientRecursiveGaussianImageFilter.hxx:214-238) or the caller's
 * gradient, in physical space.  Diagnostic (parity tests of the device
 * filters); texels4 holds 4 floats per pixel. */
int mmi_get_moving_texels(mmi_context* ctx, float* texels4);

/* Bin geometry chosen by mmi_initialize. */
int mmi_get_histogram_parameters(mmi_context* ctx, double* fixed_min, double* fixed_max, double* moving_min,
                                 double* moving_max, double* fixed_bin_size, double* moving_bin_size);

/* Timing support: device time (ms) of the last mmi_evaluate, from HIP events
 * on the context stream around its kernels; pass1_ms (clear, pass 1 with its
 * exceptions) and pass2_ms (finalize excluded: pass 2 through the derivative)
 * only with MMI_FLAG_TIME_PASSES, else -1. */
int mmi_get_last_timing(mmi_context* ctx, float* total_ms, float* pass1_ms, float* pass2_ms);

/* Diagnostics: the samples of the last fast-path evaluation that its
 * certification sent to the exact FP64 path (-1 when the last evaluation was
 * not on the fast path), and the samples per evaluation. */
int mmi_get_exception_count(mmi_context* ctx, int64_t* n_exact, int64_t* n_samples);

/* Diagnostics (no counterpart in the reference): the device-to-device copy
 * rate of the device, in GB/s (bytes read + written over the kernel time of a
 * plain 16-byte-per-lane copy of `bytes` bytes, the best of `reps` launches
 * timed with HIP events): the practical HBM ceiling beside the published
 * peak.  Allocates and frees 2 x bytes of device memory. */
int mmi_measure_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs);

/* Diagnostics (no counterpart in the reference): which kernels the last
 * mmi_evaluate ran: 0 the exact FP64 path, 1 the certified fast path, 2 the
 * fast path redone on the exact path after an exception-list overflow. */
int mmi_get_last_path(mmi_context* ctx, int32_t* path);

/* ---- Batched evaluation (serving many registrations at once) ----------------
 * No counterpart in the reference (its metric evaluates one registration per
 * call).  A batch evaluates several initialised contexts in one pass: those on
 * the certified fast path (affine, identity fixed transform, gradient image;
 * one dimension and storage per batch) share one launch of each kernel, which
 * the batch captures as a HIP graph and replays while the set of contexts and
 * their shapes stay the same; the others are evaluated one by one meanwhile.
 * A context may appear once per batch; the moving parameters are the ones
 * last set on each context.  values[i], n_valid[i], derivatives[i] (the
 * context's parameter count, when want_derivative) and status[i] (MMI_OK or
 * the code mmi_evaluate would return) receive context i's results; the call
 * itself fails only on argument or device errors. */
typedef struct mmi_batch mmi_batch;

int mmi_batch_create(int device, mmi_batch** out);
int mmi_batch_destroy(mmi_batch* batch);
/* HIP graph capture and replay of the fused launch (default on). */
int mmi_batch_set_graphs(mmi_batch* batch, int32_t enabled);
int mmi_batch_evaluate(mmi_batch* batch, mmi_context* const* ctxs, int32_t n, int32_t want_derivative,
                       double* values, double* const* derivatives, int64_t* n_valid, int32_t* status);
/* The message of the context's last failed evaluation ("" when it succeeded;
 * never NULL), also when a batch evaluated it on another thread: the text of
 * the itk::ExceptionObject the reference's GetValueAndDerivative would throw
 * (no separate counterpart in the reference, whose exceptions carry it). */
const char* mmi_context_last_error(const mmi_context* ctx);
/* Device time of the last fused launch and how many contexts it held. */
int mmi_batch_get_last_timing(mmi_batch* batch, float* device_ms, int32_t* n_fused);

/* RegularStepGradientDescentOptimizerv4 over a batch
 * (Modules/Numerics/Optimizersv4/include/itkRegularStepGradientDescentOptimizerv4.hxx:25-171,
 * itkGradientDescentOptimizerv4.hxx:72-165): every iteration evaluates all
 * still-running contexts in one batched pass, then takes each one's
 * AdvanceOneStep in the reference's arithmetic (Kahan sums, the relaxation on
 * a direction change, the previous gradient re-multiplied by 1/scales each
 * step), so each context's iterates equal a one-at-a-time run bit for bit.
 * Contexts of the fused launch (see mmi_batch_evaluate) iterate on the
 * device: the step and the fast path's new map run in a kernel after each
 * evaluation, in chunks of iterations replayed as one HIP graph, with no host
 * round trip per iteration; the others, and a context whose next evaluation
 * leaves the fast path, continue in the host loop with the same arithmetic.
 * For transforms of at most 12 parameters (affine).  The
 * caller owns one state per context, started with mmi_regular_step_start;
 * each call runs at most max_steps iterations, so contexts can join and
 * leave between calls.  The final parameters stay in each context
 * (mmi_get_moving_parameters). */
typedef struct mmi_regular_step_options {
  double learning_rate;
  double minimum_step_length;
  double relaxation_factor;
  double gradient_magnitude_tolerance;
  int32_t number_of_iterations;
  int32_t pad_;
} mmi_regular_step_options;

enum mmi_stop_condition {
  MMI_STOP_RUNNING = 0,
  MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS = 1,
  MMI_STOP_GRADIENT_MAGNITUDE_TOLERANCE = 2,
  MMI_STOP_STEP_TOO_SMALL = 3,
  MMI_STOP_METRIC_ERROR = 4
};

typedef struct mmi_regular_step_state {
  double gradient[12];
  double previous_gradient[12];
  double scales[12];
  double relaxation;  /* CurrentLearningRateRelaxation */
  double value;       /* last metric value */
  double gradient_magnitude;
  int32_t n_scales;   /* 0: identity scales */
  int32_t scales_identity;
  int32_t iteration;  /* CurrentIteration */
  int32_t stop;       /* enum mmi_stop_condition */
  int32_t status;     /* metric status when stop == MMI_STOP_METRIC_ERROR */
  int32_t pad_;
} mmi_regular_step_state;

/* StartOptimization: zero gradients, relaxation 1, iteration 0; scales (one
 * per parameter, > eps) or NULL for identity. */
int mmi_regular_step_start(mmi_regular_step_state* state, const double* scales, int32_t n_scales);
int mmi_batch_regular_step(mmi_batch* batch, mmi_context* const* ctxs, int32_t n,
                           const mmi_regular_step_options* options, mmi_regular_step_state* states, int32_t max_steps);
/* The moving transform parameters a context holds (P doubles). */
int mmi_get_moving_parameters(mmi_context* ctx, double* params, int64_t n_params);
/* Where mmi_batch_regular_step iterates fused contexts: on the device
 * (on_device != 0, the default) in chunks of `chunk` iterations (1..64,
 * default 8; a chunk's iterations after every context stopped are empty
 * launches), or in the host loop.  No counterpart in the reference (its
 * optimizer loop is host code); results are the same either way. */
int mmi_batch_set_step_loop(mmi_batch* batch, int32_t on_device, int32_t chunk);
/* Iterations the last mmi_batch_regular_step ran in the device loop (summed
 * over contexts). */
int mmi_batch_get_step_count(mmi_batch* batch, int32_t* device_steps);

#ifdef __cplusplus
}
#endif

#endif
