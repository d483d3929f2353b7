/*
 * C++ layer over the C ABI of itk_mattes_hip.h, with no ITK dependency: an
 * owning context, statuses turned into exceptions that carry the library's
 * message (the reference's exception text for degenerate inputs), the pixel
 * type codes of the images the reference takes, and image geometry in ITK's
 * conventions.  The ITK metric subclass
 * (itkMattesMutualInformationImageToImageMetricv4HIP.h) is written on top of
 * it; a host program without ITK can use it directly.
 */
#ifndef ITK_MATTES_HIP_HPP
#define ITK_MATTES_HIP_HPP

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "itk_mattes_hip.h"

namespace itk_amd {

// A failed C ABI call: the status (MMI_ERR_*) and mmi_last_error()'s text.
class MetricError : public std::runtime_error {
 public:
  MetricError(int code, const std::string& message) : std::runtime_error(message), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int status) {
  if (status != MMI_OK) throw MetricError(status, mmi_last_error());
}

// The C ABI's code of a host pixel type (mmi_set_fixed_image /
// mmi_set_moving_image); other pixel types do not compile.
template <typename TPixel>
struct pixel_code;
template <>
struct pixel_code<float> {
  static constexpr int32_t value = MMI_F32;
};
template <>
struct pixel_code<double> {
  static constexpr int32_t value = MMI_F64;
};
template <>
struct pixel_code<uint8_t> {
  static constexpr int32_t value = MMI_U8;
};
template <>
struct pixel_code<int8_t> {
  static constexpr int32_t value = MMI_S8;
};
template <>
struct pixel_code<uint16_t> {
  static constexpr int32_t value = MMI_U16;
};
template <>
struct pixel_code<int16_t> {
  static constexpr int32_t value = MMI_S16;
};
template <>
struct pixel_code<uint32_t> {
  static constexpr int32_t value = MMI_U32;
};
template <>
struct pixel_code<int32_t> {
  static constexpr int32_t value = MMI_S32;
};

// Geometry of an image in ITK's conventions from row-major D x D matrices
// (direction, ITK's index-to-physical and physical-to-index matrices as
// ImageBase computed them; itkImageBase.h:517-531).
inline mmi_geometry make_geometry(int dim, const int64_t* size, const int64_t* start, const double* origin,
                                  const double* spacing, const double* direction, const double* index_to_physical,
                                  const double* physical_to_index) {
  mmi_geometry g;
  std::memset(&g, 0, sizeof(g));
  g.dim = dim;
  for (int i = 0; i < dim; ++i) {
    g.size[i] = size[i];
    g.start[i] = start[i];
    g.origin[i] = origin[i];
    g.spacing[i] = spacing[i];
  }
  for (int i = 0; i < dim * dim; ++i) {
    g.direction[i] = direction[i];
    g.index_to_physical[i] = index_to_physical[i];
    g.physical_to_index[i] = physical_to_index[i];
  }
  return g;
}

// One metric context on a HIP device (not copyable; one thread drives it).
class Context {
 public:
  explicit Context(int device = 0, uint32_t flags = 0) { check(mmi_context_create(device, flags, &ctx_)); }
  ~Context() {
    if (ctx_) mmi_context_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  mmi_context* get() const { return ctx_; }

  template <typename TPixel>
  void set_fixed_image(const mmi_geometry& g, const TPixel* pixels) {
    check(mmi_set_fixed_image(ctx_, &g, pixel_code<TPixel>::value, pixels));
  }
  template <typename TPixel>
  void set_moving_image(const mmi_geometry& g, const TPixel* pixels) {
    check(mmi_set_moving_image(ctx_, &g, pixel_code<TPixel>::value, pixels));
  }
  void set_virtual_domain(const mmi_geometry& g) { check(mmi_set_virtual_domain(ctx_, &g)); }
  void set_number_of_histogram_bins(int32_t bins) { check(mmi_set_number_of_histogram_bins(ctx_, bins)); }
  // true: the recursive-Gaussian gradient image (computed on the device unless
  // `gradient` gives it: dim doubles per moving pixel); false: central differences
  void set_moving_gradient(bool use_filter, const double* gradient = nullptr) {
    check(mmi_set_gradient_source(ctx_, use_filter ? MMI_GRADIENT_IMAGE : MMI_GRADIENT_CENTRAL_DIFFERENCE));
    if (use_filter && gradient) check(mmi_set_moving_gradient(ctx_, gradient));
  }
  void set_fixed_mask(const mmi_geometry* g, const uint8_t* inside) { check(mmi_set_fixed_mask(ctx_, g, inside)); }
  void set_moving_mask(const mmi_geometry* g, const uint8_t* inside) { check(mmi_set_moving_mask(ctx_, g, inside)); }
  // virtual-space points, dim doubles each; empty: dense sampling
  void set_virtual_points(int32_t dim, const std::vector<double>& xyz) {
    check(mmi_set_virtual_points(ctx_, dim, xyz.empty() ? nullptr : xyz.data(), (int64_t)xyz.size() / dim));
  }
  void set_fixed_transform(int32_t kind, const std::vector<double>& fixed, const std::vector<double>& params) {
    check(mmi_set_fixed_transform(ctx_, kind, fixed.data(), (int32_t)fixed.size(), params.data(),
                                  (int64_t)params.size()));
  }
  void set_moving_transform(int32_t kind, const std::vector<double>& fixed, const std::vector<double>& params) {
    check(mmi_set_moving_transform(ctx_, kind, fixed.data(), (int32_t)fixed.size(), params.data(),
                                   (int64_t)params.size()));
  }
  void set_moving_parameters(const double* params, int64_t n) { check(mmi_set_moving_parameters(ctx_, params, n)); }
  // a composite moving transform's matrix-offset entries before the moving
  // transform, in queue order (n = 0: none)
  void set_moving_initial_transforms(int32_t n, const double* matrices, const double* offsets) {
    check(mmi_set_moving_initial_transforms(ctx_, n, matrices, offsets));
  }
  void initialize() { check(mmi_initialize(ctx_)); }

  // GetValueAndDerivative (derivative != NULL) or GetValue: the reference's
  // signs (value = -MI, derivative = -d value / d p); returns the valid count.
  int64_t evaluate(double* value, double* derivative, double* joint_pdf = nullptr) {
    int64_t n_valid = 0;
    check(mmi_evaluate(ctx_, derivative ? 1 : 0, value, derivative, &n_valid, joint_pdf, nullptr));
    return n_valid;
  }

 private:
  mmi_context* ctx_ = nullptr;
};

}  // namespace itk_amd

#endif
