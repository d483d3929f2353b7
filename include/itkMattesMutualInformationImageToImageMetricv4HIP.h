/*
 * itk::MattesMutualInformationImageToImageMetricv4HIP: the Mattes mutual
 * information v4 metric evaluated on an MI355X through the C ABI
 * (itk_mattes_hip.h via the C++ layer itk_mattes_hip.hpp).  A drop-in for
 * itk::MattesMutualInformationImageToImageMetricv4
 * (Modules/Registration/Metricsv4/include/itkMattesMutualInformationImageToImageMetricv4.h:85-108):
 * the same template parameters, an IMAGE_METRIC for ImageRegistrationMethodv4
 * and the v4 optimizers, GetValue / GetDerivative / GetValueAndDerivative with
 * the reference's signs, SetNumberOfHistogramBins clamped to >= 5, the
 * valid-point count and the reference's exception messages (library statuses
 * become itk::ExceptionObject).
 *
 * What the device takes: fixed and moving images of float, double or integer
 * pixels (double and 32-bit integer pixels exact in float: the device stores
 * FP32; interpolated values assigned to the pixel type as the reference does:
 * rounded to float, kept as double, truncated for integers), dense
 * sampling of the virtual domain or a virtual point set, ImageMaskSpatialObject
 * masks, an identity or affine fixed transform, and a moving transform that is
 * an AffineTransform or a cubic BSplineTransform, alone or as the last entry
 * of a CompositeTransform whose earlier entries (at most 4) are
 * MatrixOffsetTransformBase transforms and whose last entry alone is
 * optimized (ImageRegistrationMethodv4's moving initial transform).  Other
 * settings throw at Initialize.
 *
 * The base class's Initialize sets up the virtual domain, the transforms and a
 * mapped fixed point set; the moving-image gradient filter runs on the device
 * instead of on the host (unless the caller installed a gradient filter of its
 * own, whose host output is uploaded).
 */
#ifndef itkMattesMutualInformationImageToImageMetricv4HIP_h
#define itkMattesMutualInformationImageToImageMetricv4HIP_h

#include <cstdint>
#include <memory>
#include <type_traits>
#include <vector>

#include "itkAffineTransform.h"
#include "itkBSplineTransform.h"
#include "itkCompositeTransform.h"
#include "itkIdentityTransform.h"
#include "itkImageMaskSpatialObject.h"
#include "itkImageRegionConstIterator.h"
#include "itkImageToImageMetricv4.h"
#include "itkMatrixOffsetTransformBase.h"
#include "itk_mattes_hip.hpp"

namespace itk {

template <typename TFixedImage, typename TMovingImage, typename TVirtualImage = TFixedImage,
          typename TInternalComputationValueType = double,
          typename TMetricTraits =
              DefaultImageToImageMetricTraitsv4<TFixedImage, TMovingImage, TVirtualImage, TInternalComputationValueType>>
class MattesMutualInformationImageToImageMetricv4HIP
    : public ImageToImageMetricv4<TFixedImage, TMovingImage, TVirtualImage, TInternalComputationValueType,
                                  TMetricTraits> {
 public:
  ITK_DISALLOW_COPY_AND_MOVE(MattesMutualInformationImageToImageMetricv4HIP);

  using Self = MattesMutualInformationImageToImageMetricv4HIP;
  using Superclass =
      ImageToImageMetricv4<TFixedImage, TMovingImage, TVirtualImage, TInternalComputationValueType, TMetricTraits>;
  using Pointer = SmartPointer<Self>;
  using ConstPointer = SmartPointer<const Self>;

  itkNewMacro(Self);
  itkOverrideGetNameOfClassMacro(MattesMutualInformationImageToImageMetricv4HIP);

  using typename Superclass::DerivativeType;
  using typename Superclass::MeasureType;
  using typename Superclass::FixedImageType;
  using typename Superclass::MovingImageType;
  using typename Superclass::VirtualImageType;
  static constexpr unsigned int Dimension = TVirtualImage::ImageDimension;

  itkSetClampMacro(NumberOfHistogramBins, SizeValueType, 5, NumericTraits<SizeValueType>::max());
  itkGetConstReferenceMacro(NumberOfHistogramBins, SizeValueType);

  /** HIP device of the metric's context (set before the first Initialize). */
  itkSetMacro(Device, int);
  itkGetConstMacro(Device, int);

  void Initialize() override {
    // the base class without its host gradient filter: the device computes it
    const bool use_filter = this->m_UseMovingImageGradientFilter;
    const bool user_filter = use_filter && this->m_MovingImageGradientFilter.IsNotNull() &&
                             this->m_MovingImageGradientFilter != this->m_DefaultMovingImageGradientFilter;
    if (!user_filter) this->m_UseMovingImageGradientFilter = false;
    Superclass::Initialize();
    this->m_UseMovingImageGradientFilter = use_filter;
    try {
      if (!m_Context) m_Context.reset(new itk_amd::Context(m_Device, 0));
      itk_amd::Context& c = *m_Context;
      c.set_fixed_image(Geometry(this->m_FixedImage.GetPointer()), this->m_FixedImage->GetBufferPointer());
      c.set_moving_image(Geometry(this->m_MovingImage.GetPointer()), this->m_MovingImage->GetBufferPointer());
      c.set_virtual_domain(VirtualGeometry());
      c.set_number_of_histogram_bins(static_cast<int32_t>(m_NumberOfHistogramBins));
      std::vector<double> gradient;
      if (user_filter) gradient = HostGradient();
      c.set_moving_gradient(use_filter, gradient.empty() ? nullptr : gradient.data());
      PushMask(true, this->m_FixedImageMask.GetPointer());
      PushMask(false, this->m_MovingImageMask.GetPointer());
      std::vector<double> xyz;
      if (this->m_UseSampledPointSet) {  // mapped to virtual space by the base class
        const auto* points = this->m_VirtualSampledPointSet->GetPoints();
        for (auto it = points->Begin(); it != points->End(); ++it)
          for (unsigned int d = 0; d < Dimension; ++d) xyz.push_back(it.Value()[d]);
      }
      c.set_virtual_points(Dimension, xyz);
      PushTransform(true, this->m_FixedTransform.GetPointer());
      PushTransform(false, this->m_MovingTransform.GetPointer());
      c.initialize();
    } catch (const itk_amd::MetricError& e) {
      itkExceptionMacro(<< e.what());
    }
  }

  MeasureType GetValue() const override {
    MeasureType value{};
    Evaluate(value, nullptr);
    return value;
  }

  void GetDerivative(DerivativeType& derivative) const override {
    MeasureType value{};
    GetValueAndDerivative(value, derivative);
  }

  void GetValueAndDerivative(MeasureType& value, DerivativeType& derivative) const override {
    // sized and written whole, as InitializeForIteration sizes and zeroes it
    // (itkImageToImageMetricv4.hxx:267-281)
    derivative.SetSize(this->GetNumberOfParameters());
    Evaluate(value, derivative.data_block());
  }

 protected:
  MattesMutualInformationImageToImageMetricv4HIP() = default;
  ~MattesMutualInformationImageToImageMetricv4HIP() override = default;

 private:
  void Evaluate(MeasureType& value, double* derivative) const {
    if (!m_Context) itkExceptionMacro(<< "Initialize must precede GetValue and GetValueAndDerivative");
    try {
      const auto& p = this->m_MovingTransform->GetParameters();
      m_Context->set_moving_parameters(p.data_block(), static_cast<int64_t>(p.Size()));
      double v = 0.0;
      const int64_t valid = m_Context->evaluate(&v, derivative);
      this->m_NumberOfValidPoints = static_cast<SizeValueType>(valid);
      value = static_cast<MeasureType>(v);
      this->m_Value = value;
    } catch (const itk_amd::MetricError& e) {
      itkExceptionMacro(<< e.what());
    }
  }

  // ITK's index-to-physical matrices, computed as
  // ImageBase::ComputeIndexToPhysicalPointMatrices does (direction * spacing,
  // then its inverse), so the device sees ITK's continuous indices bit for bit.
  template <typename TImage>
  static mmi_geometry Geometry(const TImage* image) {
    constexpr unsigned int D = TImage::ImageDimension;
    const auto& region = image->GetBufferedRegion();
    typename TImage::DirectionType scale;
    scale.Fill(0.0);
    for (unsigned int i = 0; i < D; ++i) scale[i][i] = image->GetSpacing()[i];
    const typename TImage::DirectionType i2p = image->GetDirection() * scale;
    const typename TImage::DirectionType p2i(i2p.GetInverse());
    int64_t size[3] = {1, 1, 1}, start[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0}, spacing[3] = {1, 1, 1}, dir[9], m1[9], m2[9];
    for (unsigned int i = 0; i < D; ++i) {
      size[i] = static_cast<int64_t>(region.GetSize()[i]);
      start[i] = static_cast<int64_t>(region.GetIndex()[i]);
      origin[i] = image->GetOrigin()[i];
      spacing[i] = image->GetSpacing()[i];
      for (unsigned int j = 0; j < D; ++j) {
        dir[i * D + j] = image->GetDirection()[i][j];
        m1[i * D + j] = i2p[i][j];
        m2[i * D + j] = p2i[i][j];
      }
    }
    return itk_amd::make_geometry(static_cast<int>(D), size, start, origin, spacing, dir, m1, m2);
  }

  mmi_geometry VirtualGeometry() const {
    // the virtual image's geometry over the virtual region the base class uses
    mmi_geometry g = Geometry(this->m_VirtualImage.GetPointer());
    const auto region = this->GetVirtualRegion();
    for (unsigned int i = 0; i < Dimension; ++i) {
      g.size[i] = static_cast<int64_t>(region.GetSize()[i]);
      g.start[i] = static_cast<int64_t>(region.GetIndex()[i]);
    }
    return g;
  }

  // The caller's gradient filter output (its own choice of filter), D
  // doubles per moving pixel in buffer order.
  std::vector<double> HostGradient() const {
    const auto* grad = this->m_MovingImageGradientImage.GetPointer();
    std::vector<double> out;
    out.reserve(grad->GetBufferedRegion().GetNumberOfPixels() * Dimension);
    ImageRegionConstIterator<std::remove_const_t<std::remove_pointer_t<decltype(grad)>>> it(grad,
                                                                                            grad->GetBufferedRegion());
    for (; !it.IsAtEnd(); ++it)
      for (unsigned int d = 0; d < Dimension; ++d) out.push_back(static_cast<double>(it.Get()[d]));
    return out;
  }

  // An ImageMaskSpatialObject as one byte per mask voxel (1 inside, under its
  // UseMaskValue rule); no mask: NULL.  Other spatial objects throw.
  template <typename TMask>
  void PushMask(bool fixed, const TMask* mask) {
    itk_amd::Context& c = *m_Context;
    if (!mask) {
      fixed ? c.set_fixed_mask(nullptr, nullptr) : c.set_moving_mask(nullptr, nullptr);
      return;
    }
    using MaskSO = ImageMaskSpatialObject<Dimension>;
    const auto* so = dynamic_cast<const MaskSO*>(mask);
    if (!so) itkExceptionMacro(<< "the HIP metric takes ImageMaskSpatialObject masks only");
    const auto* image = so->GetImage();
    std::vector<uint8_t> inside;
    inside.reserve(image->GetBufferedRegion().GetNumberOfPixels());
    ImageRegionConstIterator<typename MaskSO::ImageType> it(image, image->GetBufferedRegion());
    for (; !it.IsAtEnd(); ++it)
      inside.push_back(so->GetUseMaskValue() ? (it.Get() == so->GetMaskValue()) : (it.Get() != 0));
    const mmi_geometry g = Geometry(image);
    fixed ? c.set_fixed_mask(&g, inside.data()) : c.set_moving_mask(&g, inside.data());
  }

  // Affine (fixed parameters: the centre), cubic B-spline (grid size, origin,
  // spacing, direction) or identity; a CompositeTransform with one entry is
  // the entry (the driver's output transform, itkCompositeTransform.hxx:28-54).
  template <typename TTransform>
  void PushTransform(bool fixed, const TTransform* t) {
    using Scalar = typename TTransform::ParametersValueType;
    const TransformBaseTemplate<Scalar>* base = t;
    std::vector<double> stage_m, stage_o;  // moving composite: the entries before the last
    int32_t n_stages = 0;
    if (const auto* comp = dynamic_cast<const CompositeTransform<Scalar, Dimension>*>(base)) {
      const SizeValueType n = comp->GetNumberOfTransforms();
      if (n == 0) itkExceptionMacro(<< "the moving CompositeTransform is empty");
      if (fixed && n != 1) itkExceptionMacro(<< "the HIP metric takes a fixed CompositeTransform with one transform");
      if (n - 1 > 4) itkExceptionMacro(<< "the HIP metric takes at most 4 transforms before the optimized one");
      using MatrixOffset = MatrixOffsetTransformBase<Scalar, Dimension, Dimension>;
      for (SizeValueType i = 0; i + 1 < n; ++i) {
        if (comp->GetNthTransformToOptimize(i))
          itkExceptionMacro(<< "the HIP metric optimizes only the CompositeTransform's most recent transform");
        const auto* s = dynamic_cast<const MatrixOffset*>(comp->GetNthTransformConstPointer(i));
        if (!s) itkExceptionMacro(<< "the transforms before the optimized one must be matrix-offset transforms");
        for (unsigned int r = 0; r < Dimension; ++r) {
          for (unsigned int c = 0; c < Dimension; ++c) stage_m.push_back(static_cast<double>(s->GetMatrix()[r][c]));
          stage_o.push_back(static_cast<double>(s->GetOffset()[r]));
        }
        ++n_stages;
      }
      base = comp->GetNthTransformConstPointer(n - 1);
    }
    if (!fixed) m_Context->set_moving_initial_transforms(n_stages, stage_m.data(), stage_o.data());
    int32_t kind = MMI_TF_IDENTITY;
    if (dynamic_cast<const AffineTransform<Scalar, Dimension>*>(base))
      kind = MMI_TF_AFFINE;
    else if (!fixed && dynamic_cast<const BSplineTransform<Scalar, Dimension, 3>*>(base))
      kind = MMI_TF_BSPLINE;
    else if (!dynamic_cast<const IdentityTransform<Scalar, Dimension>*>(base))
      itkExceptionMacro(<< "the HIP metric takes affine, cubic B-spline or identity transforms");
    std::vector<double> fp, p;
    if (kind != MMI_TF_IDENTITY) {
      const auto& f = base->GetFixedParameters();
      const auto& q = base->GetParameters();
      fp.assign(f.begin(), f.end());
      p.assign(q.begin(), q.end());
    }
    fixed ? m_Context->set_fixed_transform(kind, fp, p) : m_Context->set_moving_transform(kind, fp, p);
  }

  SizeValueType m_NumberOfHistogramBins{ 50 };
  int m_Device{ 0 };
  mutable std::unique_ptr<itk_amd::Context> m_Context;
};

}  // namespace itk

#endif
