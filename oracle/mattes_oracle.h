/*
 * CPU FP64 restatement of ITK's dense Mattes mutual-information v4 metric.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (pytest, the smoke check and the CPU-baseline leg of
 * bench.py).  Product code never links, loads or calls it.
 *
 * Every function follows a reference file named in mattes_oracle.cpp, in the
 * same floating-point operation order, compiled without FMA contraction.
 */
#ifndef ORC_MATTES_ORACLE_H
#define ORC_MATTES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Image (or B-spline coefficient grid) geometry, ITK conventions: x fastest,
 * physical = origin + index_to_physical * index,
 * continuous index = physical_to_index * (physical - origin). */
typedef struct {
  int32_t dim;
  int32_t pad_;
  int64_t size[3];
  int64_t start[3];
  double origin[3];
  double spacing[3];
  double direction[9];         /* row-major dim x dim */
  double index_to_physical[9]; /* row-major dim x dim = direction * diag(spacing) */
  double physical_to_index[9]; /* row-major dim x dim */
} orc_geometry;

enum { ORC_TF_IDENTITY = 0, ORC_TF_AFFINE = 1, ORC_TF_BSPLINE = 2 };

typedef struct orc_transform_s {
  int32_t kind;
  int32_t pad_;
  double matrix[9];  /* affine: row-major */
  double offset[3];  /* affine: y = matrix * x + offset */
  double center[3];  /* affine: Jacobian is taken about this point */
  orc_geometry grid; /* B-spline coefficient grid, size = mesh + 3 */
  const double* coeffs; /* B-spline parameters, dim * prod(grid.size), dimension-major */
  /* moving transform of a CompositeTransform whose only optimized transform
   * is this one (its queue's last entry): the n_initial affine transforms
   * before it, in queue order (a point goes through this one, then
   * initial[n_initial - 1] .. initial[0]); NULL / 0 for none */
  const struct orc_transform_s* initial;
  int32_t n_initial;
  int32_t pad2_;
} orc_transform;

typedef struct {
  int32_t bins;
  int32_t pixel_f64;        /* 1: pixels are double (interpolated values not rounded to float) */
  int32_t gradient_source;  /* 0: central differences of the moving image, 1: gradient image */
  int32_t compute_derivative;
  int32_t threads;          /* slabs along the slowest virtual axis */
  int32_t pixel_integer;    /* 1: the images' pixel type is integral (interpolated values truncated) */
  double fixed_min, fixed_max, moving_min, moving_max;
} orc_options;

typedef struct {
  double value;
  double joint_pdf_sum;
  int64_t n_valid;
  double* derivative;      /* [P] or NULL */
  double* joint_pdf;       /* [B*B] normalised, row = fixed bin; or NULL */
  double* fixed_marginal;  /* [B] normalised; or NULL */
  double* moving_marginal; /* [B]; or NULL */
  int16_t* fixed_bin;      /* [N virtual], -1 when the sample is not counted; or NULL */
  int16_t* moving_bin;     /* [N virtual], Parzen window index, -1 when not counted; or NULL */
  double* moving_value;    /* [N virtual], interpolated moving value (NaN when outside); or NULL */
} orc_result;

/* Sample selection beyond the dense virtual domain.  points: n_points x dim
 * virtual-space points (the sparse threader's point set), or NULL for every
 * virtual voxel.  Masks: ImageMaskSpatialObject with default settings over a
 * uint8 image of the given geometry, or NULL geometry for none. */
typedef struct {
  const double* points;
  int64_t n_points;
  const orc_geometry* fixed_mask_geom;
  const uint8_t* fixed_mask;
  const orc_geometry* moving_mask_geom;
  const uint8_t* moving_mask;
} orc_sampling;

/* Mattes MI value (and derivative) over every voxel of the virtual domain.
 * Returns 0 on success, 1 with the reference's exception text in err. */
int orc_mattes_evaluate(const orc_geometry* virt, const orc_geometry* fixed_geom, const void* fixed_pixels,
                        const orc_geometry* moving_geom, const void* moving_pixels,
                        const double* moving_gradient, const orc_transform* fixed_tf,
                        const orc_transform* moving_tf, const orc_options* opt, orc_result* res, char* err,
                        int errlen);

/* orc_mattes_evaluate over the samples and masks of smp (NULL: dense, no
 * masks).  Result arrays indexed by sample have n_points entries when points
 * are given. */
int orc_mattes_evaluate_ex(const orc_geometry* virt, const orc_geometry* fixed_geom, const void* fixed_pixels,
                           const orc_geometry* moving_geom, const void* moving_pixels,
                           const double* moving_gradient, const orc_transform* fixed_tf,
                           const orc_transform* moving_tf, const orc_options* opt, const orc_sampling* smp,
                           orc_result* res, char* err, int errlen);

/* Intensity range of MattesMutualInformationImageToImageMetricv4::Initialize
 * over one image: with points, the pixels the points (mapped by fixed_tf)
 * round to; else with a mask, the voxels whose physical points the mask holds;
 * else every voxel.  Returns the number of pixels taken; lo/hi keep the
 * reference's start values (max, -max of the pixel type) when it is 0. */
int64_t orc_intensity_range(const orc_geometry* g, const void* pixels, int32_t pixel_f64,
                            const orc_geometry* mask_geom, const uint8_t* mask, const double* points,
                            int64_t n_points, const orc_transform* fixed_tf, double* lo, double* hi);

/* Math::RoundHalfIntegerUp<int64>(double) and the rounded-index test of
 * ImageBase::TransformPhysicalPointToIndex: 1 with idx set when the rounded
 * index is inside the image. */
int64_t orc_itk_round(double x);
int orc_physical_point_to_index(const orc_geometry* g, const double* point, int64_t* idx);

/* MersenneTwisterRandomVariateGenerator seeded with seed: n integer variates,
 * or n normal variates (mean 0, variance 1). */
void orc_mt_integers(uint32_t seed, uint32_t* out, int64_t n);
void orc_mt_normals(uint32_t seed, double* out, int64_t n);

/* ImageRegistrationMethodv4::SetMetricSamplePoints for one level: strategy 1
 * REGULAR, 2 RANDOM; seed in/out (m_CurrentRandomSeed); fixed mask optional;
 * out has room for the bound (REGULAR ceil(N / ceil(1/p)), RANDOM
 * (unsigned long)((float)N * p)) points of dim doubles.  Returns the number of
 * points kept, -1 on a bad argument. */
int64_t orc_metric_sample_points(const orc_geometry* virt, int32_t strategy, double percentage, uint32_t* seed,
                                 const orc_geometry* mask_geom, const uint8_t* mask, double* out);

/* Bin geometry of MattesMutualInformationImageToImageMetricv4::Initialize. */
int orc_bin_parameters(double true_min, double true_max, int32_t bins, double* bin_size, double* normalized_min,
                       char* err, int errlen);

/* Parzen bin index of one intensity (ComputeSingleFixedImageParzenWindowIndex). */
int64_t orc_parzen_index(double value, double bin_size, double normalized_min, int32_t bins);

double orc_bspline3(double u);
double orc_bspline3_derivative(double u);
int64_t orc_itk_floor(double x);

/* Linear interpolation of a scalar image at a physical point; returns 0 when
 * the point is outside the buffer. */
int orc_interpolate(const orc_geometry* g, const void* pixels, int32_t pixel_f64, const double* point, double* out);

/* Transform a point (affine or B-spline), as the moving transform does. */
void orc_transform_point(const orc_transform* tf, int32_t dim, const double* in, double* out);

/* GradientRecursiveGaussianImageFilter (sigma, normalise across scale, image
 * direction applied): out has dim doubles per voxel. */
int orc_gradient_recursive_gaussian(const orc_geometry* g, const void* pixels, int32_t pixel_f64, double sigma,
                                    double* out, int32_t threads);

/* SmoothingRecursiveGaussianImageFilter: zero-order recursive Gaussian with
 * sigma[axis] (physical units) along the last axis, then axes 0 .. dim-2;
 * normalise across scale off; FP64 internally, out is float or double like
 * the input. */
int orc_smooth_recursive_gaussian(const orc_geometry* g, const void* pixels, int32_t pixel_f64, const double* sigma,
                                  void* out, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif
