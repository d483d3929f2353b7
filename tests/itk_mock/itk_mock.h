// Test-only stand-in for the subset of ITK's interface that
// include/itkMattesMutualInformationImageToImageMetricv4HIP.h uses: the class,
// member and method names the ITK headers declare for them
// (itkImageToImageMetricv4.h, itkObjectToObjectMetric.h, itkImageBase.h,
// itkMatrixOffsetTransformBase.h, itkBSplineTransform.h,
// itkCompositeTransform.h, itkImageMaskSpatialObject.h, itkMacro.h), with
// minimal behaviour of its own, so the subclass compiles and runs in this
// repository's tests without ITK's CMake-generated configuration.  It is not
// ITK and ships in no product path.
#pragma once

#include <cmath>
#include <cstddef>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace itk {

using SizeValueType = unsigned long;
using IndexValueType = long;

class ExceptionObject : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

#define itkExceptionMacro(x)                            \
  do {                                                  \
    std::ostringstream itk_mock_message_;               \
    itk_mock_message_ x;                                \
    throw ::itk::ExceptionObject(itk_mock_message_.str()); \
  } while (0)
#define ITK_DISALLOW_COPY_AND_MOVE(T) \
  T(const T&) = delete;               \
  T& operator=(const T&) = delete
#define itkNewMacro(T) \
  static Pointer New() { return Pointer(new T()); }
#define itkOverrideGetNameOfClassMacro(T) \
  const char* GetNameOfClass() const override { return #T; }
#define itkSetClampMacro(name, type, lo, hi) \
  virtual void Set##name(type v) { this->m_##name = v < (lo) ? (lo) : (v > (hi) ? (hi) : v); }
#define itkGetConstReferenceMacro(name, type) \
  virtual const type& Get##name() const { return this->m_##name; }
#define itkSetMacro(name, type) \
  virtual void Set##name(type v) { this->m_##name = v; }
#define itkGetConstMacro(name, type) \
  virtual type Get##name() const { return this->m_##name; }

template <typename T>
struct NumericTraits {
  static T max() { return std::numeric_limits<T>::max(); }
};

// Intrusive reference counting as ITK's objects do it.
class LightObject {
 public:
  virtual ~LightObject() = default;
  virtual const char* GetNameOfClass() const { return "LightObject"; }
  void Register() const { ++m_References; }
  void UnRegister() const {
    if (--m_References == 0) delete this;
  }

 private:
  mutable int m_References = 0;
};

template <typename T>
class SmartPointer {
 public:
  SmartPointer() = default;
  SmartPointer(std::nullptr_t) {}
  SmartPointer(T* p) : m_Pointer(p) { Take(); }
  SmartPointer(const SmartPointer& o) : m_Pointer(o.m_Pointer) { Take(); }
  template <typename U, typename = std::enable_if_t<std::is_convertible<U*, T*>::value>>
  SmartPointer(const SmartPointer<U>& o) : m_Pointer(o.GetPointer()) {
    Take();
  }
  ~SmartPointer() { Drop(); }
  SmartPointer& operator=(const SmartPointer& o) {
    SmartPointer tmp(o);
    std::swap(m_Pointer, tmp.m_Pointer);
    return *this;
  }
  T* GetPointer() const { return m_Pointer; }
  T* operator->() const { return m_Pointer; }
  T& operator*() const { return *m_Pointer; }
  operator T*() const { return m_Pointer; }
  bool IsNull() const { return m_Pointer == nullptr; }
  bool IsNotNull() const { return m_Pointer != nullptr; }

 private:
  void Take() {
    if (m_Pointer) m_Pointer->Register();
  }
  void Drop() {
    if (m_Pointer) m_Pointer->UnRegister();
    m_Pointer = nullptr;
  }
  T* m_Pointer = nullptr;
};

template <typename T, unsigned int R, unsigned int C>
class Matrix {
 public:
  T* operator[](unsigned int i) { return m_Data[i]; }
  const T* operator[](unsigned int i) const { return m_Data[i]; }
  void Fill(T v) {
    for (auto& row : m_Data)
      for (auto& x : row) x = v;
  }
  static Matrix Identity() {
    Matrix m;
    m.Fill(0);
    for (unsigned int i = 0; i < R && i < C; ++i) m[i][i] = 1;
    return m;
  }
  Matrix operator*(const Matrix& b) const {
    Matrix r;
    for (unsigned int i = 0; i < R; ++i)
      for (unsigned int j = 0; j < C; ++j) {
        T s = 0;
        for (unsigned int k = 0; k < C; ++k) s += m_Data[i][k] * b[k][j];
        r[i][j] = s;
      }
    return r;
  }
  // Gauss-Jordan with partial pivoting (exact for the diagonal matrices the
  // tests use)
  Matrix GetInverse() const {
    double a[R][2 * R];
    for (unsigned int i = 0; i < R; ++i)
      for (unsigned int j = 0; j < 2 * R; ++j) a[i][j] = j < R ? m_Data[i][j] : (j - R == i ? 1.0 : 0.0);
    for (unsigned int c = 0; c < R; ++c) {
      unsigned int p = c;
      for (unsigned int r = c + 1; r < R; ++r)
        if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
      for (unsigned int j = 0; j < 2 * R; ++j) std::swap(a[c][j], a[p][j]);
      const double d = a[c][c];
      for (unsigned int j = 0; j < 2 * R; ++j) a[c][j] /= d;
      for (unsigned int r = 0; r < R; ++r)
        if (r != c) {
          const double f = a[r][c];
          if (f != 0.0)
            for (unsigned int j = 0; j < 2 * R; ++j) a[r][j] -= f * a[c][j];
        }
    }
    Matrix inv;
    for (unsigned int i = 0; i < R; ++i)
      for (unsigned int j = 0; j < R; ++j) inv[i][j] = (T)a[i][j + R];
    return inv;
  }

 private:
  T m_Data[R][C] = {};
};

template <typename T, unsigned int D>
struct FixedArray {
  T v[D] = {};
  T& operator[](unsigned int i) { return v[i]; }
  const T& operator[](unsigned int i) const { return v[i]; }
};

template <typename T, unsigned int D>
using Point = FixedArray<T, D>;
template <typename T, unsigned int D>
using CovariantVector = FixedArray<T, D>;

template <unsigned int D>
struct ImageRegion {
  FixedArray<IndexValueType, D> m_Index;
  FixedArray<SizeValueType, D> m_Size;
  const FixedArray<IndexValueType, D>& GetIndex() const { return m_Index; }
  const FixedArray<SizeValueType, D>& GetSize() const { return m_Size; }
  SizeValueType GetNumberOfPixels() const {
    SizeValueType n = 1;
    for (unsigned int i = 0; i < D; ++i) n *= m_Size[i];
    return n;
  }
};

template <typename TPixel, unsigned int D>
class Image : public LightObject {
 public:
  using Self = Image;
  using Pointer = SmartPointer<Self>;
  using ConstPointer = SmartPointer<const Self>;
  using PixelType = TPixel;
  static constexpr unsigned int ImageDimension = D;
  using DirectionType = Matrix<double, D, D>;
  using SpacingType = FixedArray<double, D>;
  using PointType = Point<double, D>;
  using RegionType = ImageRegion<D>;
  static Pointer New() { return Pointer(new Self()); }
  void SetRegions(const RegionType& r) { m_Region = r; }
  void Allocate() { m_Buffer.assign(m_Region.GetNumberOfPixels(), TPixel{}); }
  void SetOrigin(const PointType& o) { m_Origin = o; }
  void SetSpacing(const SpacingType& s) { m_Spacing = s; }
  void SetDirection(const DirectionType& d) { m_Direction = d; }
  const RegionType& GetBufferedRegion() const { return m_Region; }
  const PointType& GetOrigin() const { return m_Origin; }
  const SpacingType& GetSpacing() const { return m_Spacing; }
  const DirectionType& GetDirection() const { return m_Direction; }
  const TPixel* GetBufferPointer() const { return m_Buffer.data(); }
  TPixel* GetBufferPointer() { return m_Buffer.data(); }
  const std::vector<TPixel>& Buffer() const { return m_Buffer; }

 private:
  RegionType m_Region;
  PointType m_Origin;
  SpacingType m_Spacing;
  DirectionType m_Direction = DirectionType::Identity();
  std::vector<TPixel> m_Buffer;
};

template <typename TImage>
class ImageRegionConstIterator {
 public:
  ImageRegionConstIterator(const TImage* image, const typename TImage::RegionType&) : m_Image(image) {}
  bool IsAtEnd() const { return m_Offset >= m_Image->Buffer().size(); }
  ImageRegionConstIterator& operator++() {
    ++m_Offset;
    return *this;
  }
  const typename TImage::PixelType& Get() const { return m_Image->Buffer()[m_Offset]; }

 private:
  const TImage* m_Image;
  size_t m_Offset = 0;
};

template <typename T>
class Array {
 public:
  using ValueType = T;
  Array() = default;
  explicit Array(SizeValueType n) : m_Data(n) {}
  void SetSize(SizeValueType n) { m_Data.assign(n, T{}); }
  SizeValueType Size() const { return m_Data.size(); }
  T* data_block() { return m_Data.data(); }
  const T* data_block() const { return m_Data.data(); }
  T& operator[](SizeValueType i) { return m_Data[i]; }
  const T& operator[](SizeValueType i) const { return m_Data[i]; }
  typename std::vector<T>::const_iterator begin() const { return m_Data.begin(); }
  typename std::vector<T>::const_iterator end() const { return m_Data.end(); }

 private:
  std::vector<T> m_Data;
};

template <typename T>
class TransformBaseTemplate : public LightObject {
 public:
  using ParametersValueType = T;
  using ParametersType = Array<T>;
  using FixedParametersType = Array<double>;
  virtual const ParametersType& GetParameters() const { return m_Parameters; }
  virtual const FixedParametersType& GetFixedParameters() const { return m_FixedParameters; }
  virtual void SetParameters(const ParametersType& p) { m_Parameters = p; }
  virtual void SetFixedParameters(const FixedParametersType& p) { m_FixedParameters = p; }
  virtual SizeValueType GetNumberOfParameters() const { return m_Parameters.Size(); }

 protected:
  ParametersType m_Parameters;
  FixedParametersType m_FixedParameters;
};

template <typename T, unsigned int NIn, unsigned int NOut>
class Transform : public TransformBaseTemplate<T> {
 public:
  using Pointer = SmartPointer<Transform>;
};

// MatrixOffsetTransformBase: y = matrix * x + offset; parameters = matrix
// row-major then translation, fixed parameters = centre, offset from
// ComputeOffset (itkMatrixOffsetTransformBase.hxx:420-500, 615-631).
template <typename T = double, unsigned int NIn = 3, unsigned int NOut = 3>
class MatrixOffsetTransformBase : public Transform<T, NIn, NOut> {
 public:
  using MatrixType = Matrix<T, NOut, NIn>;
  using OutputVectorType = FixedArray<T, NOut>;
  const MatrixType& GetMatrix() const { return m_Matrix; }
  const OutputVectorType& GetOffset() const { return m_Offset; }
  void SetParameters(const typename TransformBaseTemplate<T>::ParametersType& p) override {
    TransformBaseTemplate<T>::SetParameters(p);
    Update();
  }
  void SetFixedParameters(const typename TransformBaseTemplate<T>::FixedParametersType& p) override {
    TransformBaseTemplate<T>::SetFixedParameters(p);
    Update();
  }

 private:
  void Update() {
    const auto& p = this->m_Parameters;
    const auto& c = this->m_FixedParameters;
    if (p.Size() != NOut * NIn + NOut || c.Size() != NIn) return;
    for (unsigned int i = 0; i < NOut; ++i)
      for (unsigned int j = 0; j < NIn; ++j) m_Matrix[i][j] = p[i * NIn + j];
    for (unsigned int i = 0; i < NOut; ++i) {
      T off = p[NOut * NIn + i] + c[i];
      for (unsigned int j = 0; j < NIn; ++j) off -= m_Matrix[i][j] * c[j];
      m_Offset[i] = off;
    }
  }
  MatrixType m_Matrix;
  OutputVectorType m_Offset;
};

template <typename T = double, unsigned int D = 3>
class AffineTransform : public MatrixOffsetTransformBase<T, D, D> {
 public:
  using Pointer = SmartPointer<AffineTransform>;
  static Pointer New() {
    Pointer p(new AffineTransform());
    typename TransformBaseTemplate<T>::ParametersType q(D * D + D);
    for (unsigned int i = 0; i < D; ++i) q[i * D + i] = 1;
    p->SetFixedParameters(Array<double>(D));
    p->SetParameters(q);
    return p;
  }
};

template <typename T = double, unsigned int D = 3, unsigned int VSplineOrder = 3>
class BSplineTransform : public Transform<T, D, D> {
 public:
  using Pointer = SmartPointer<BSplineTransform>;
  static Pointer New() { return Pointer(new BSplineTransform()); }
};

template <typename T = double, unsigned int D = 3>
class IdentityTransform : public Transform<T, D, D> {
 public:
  using Pointer = SmartPointer<IdentityTransform>;
  static Pointer New() { return Pointer(new IdentityTransform()); }
};

// Parameters of the transforms to optimise; here every entry is optimised and
// the parameters are those of the last entry when there is one.
template <typename T = double, unsigned int D = 3>
class CompositeTransform : public Transform<T, D, D> {
 public:
  using Pointer = SmartPointer<CompositeTransform>;
  using TransformType = Transform<T, D, D>;
  static Pointer New() { return Pointer(new CompositeTransform()); }
  // a new entry is optimized (as ITK's PushBackTransform sets it)
  void AddTransform(TransformType* t) {
    m_Queue.push_back(typename TransformType::Pointer(t));
    m_Optimize.push_back(true);
  }
  void SetOnlyMostRecentTransformToOptimizeOn() {
    for (size_t i = 0; i < m_Optimize.size(); ++i) m_Optimize[i] = i + 1 == m_Optimize.size();
  }
  bool GetNthTransformToOptimize(SizeValueType n) const { return m_Optimize[n]; }
  SizeValueType GetNumberOfTransforms() const { return m_Queue.size(); }
  const TransformType* GetNthTransformConstPointer(SizeValueType n) const { return m_Queue[n].GetPointer(); }
  const typename TransformBaseTemplate<T>::ParametersType& GetParameters() const override {
    return m_Queue.back()->GetParameters();
  }
  void SetParameters(const typename TransformBaseTemplate<T>::ParametersType& p) override {
    const_cast<TransformType*>(m_Queue.back().GetPointer())->SetParameters(p);
  }
  SizeValueType GetNumberOfParameters() const override { return m_Queue.back()->GetNumberOfParameters(); }

 private:
  std::vector<typename TransformType::Pointer> m_Queue;
  std::vector<bool> m_Optimize;
};

template <unsigned int D>
class SpatialObject : public LightObject {
 public:
  using Pointer = SmartPointer<SpatialObject>;
  using ConstPointer = SmartPointer<const SpatialObject>;
};

template <unsigned int D = 3, typename TPixel = unsigned char>
class ImageMaskSpatialObject : public SpatialObject<D> {
 public:
  using Pointer = SmartPointer<ImageMaskSpatialObject>;
  using ImageType = Image<TPixel, D>;
  using PixelType = TPixel;
  static Pointer New() { return Pointer(new ImageMaskSpatialObject()); }
  void SetImage(const ImageType* image) { m_Image = image; }
  const ImageType* GetImage() const { return m_Image.GetPointer(); }
  itkSetMacro(UseMaskValue, bool);
  itkGetConstReferenceMacro(UseMaskValue, bool);
  itkSetMacro(MaskValue, PixelType);
  itkGetConstReferenceMacro(MaskValue, PixelType);

 private:
  typename ImageType::ConstPointer m_Image;
  bool m_UseMaskValue = false;
  PixelType m_MaskValue = 1;
};

template <typename TCoordinate, unsigned int D>
class PointSet : public LightObject {
 public:
  using Pointer = SmartPointer<PointSet>;
  using ConstPointer = SmartPointer<const PointSet>;
  using PointType = Point<TCoordinate, D>;
  struct PointsContainer {
    std::vector<PointType> m_Points;
    struct ConstIterator {
      const PointType* p;
      const PointType& Value() const { return *p; }
      ConstIterator& operator++() {
        ++p;
        return *this;
      }
      bool operator!=(const ConstIterator& o) const { return p != o.p; }
    };
    ConstIterator Begin() const { return {m_Points.data()}; }
    ConstIterator End() const { return {m_Points.data() + m_Points.size()}; }
    SizeValueType Size() const { return m_Points.size(); }
  };
  static Pointer New() { return Pointer(new PointSet()); }
  const PointsContainer* GetPoints() const { return &m_Container; }
  void AddPoint(const PointType& p) { m_Container.m_Points.push_back(p); }

 private:
  PointsContainer m_Container;
};

template <typename TFixedImage, typename TMovingImage, typename TVirtualImage, typename TInternal>
struct DefaultImageToImageMetricTraitsv4 {};

class GradientFilterBase : public LightObject {};

template <typename TFixedImage, typename TMovingImage, typename TVirtualImage = TFixedImage,
          typename TInternal = double,
          typename TTraits = DefaultImageToImageMetricTraitsv4<TFixedImage, TMovingImage, TVirtualImage, TInternal>>
class ImageToImageMetricv4 : public LightObject {
 public:
  using MeasureType = TInternal;
  using DerivativeType = Array<TInternal>;
  using FixedImageType = TFixedImage;
  using MovingImageType = TMovingImage;
  using VirtualImageType = TVirtualImage;
  static constexpr unsigned int VirtualImageDimension = TVirtualImage::ImageDimension;
  using FixedTransformType = Transform<TInternal, VirtualImageDimension, TFixedImage::ImageDimension>;
  using MovingTransformType = Transform<TInternal, VirtualImageDimension, TMovingImage::ImageDimension>;
  using MovingImageGradientImageType =
      Image<CovariantVector<double, TMovingImage::ImageDimension>, TMovingImage::ImageDimension>;
  using VirtualPointSetType = PointSet<double, VirtualImageDimension>;
  using FixedImageMaskType = SpatialObject<TFixedImage::ImageDimension>;
  using MovingImageMaskType = SpatialObject<TMovingImage::ImageDimension>;

  void SetFixedImage(const FixedImageType* i) { m_FixedImage = i; }
  void SetMovingImage(const MovingImageType* i) { m_MovingImage = i; }
  void SetFixedTransform(FixedTransformType* t) { m_FixedTransform = t; }
  void SetMovingTransform(MovingTransformType* t) { m_MovingTransform = t; }
  void SetVirtualDomainFromImage(const VirtualImageType* i) { m_VirtualImage = const_cast<VirtualImageType*>(i); }
  void SetFixedImageMask(const FixedImageMaskType* m) { m_FixedImageMask = m; }
  void SetMovingImageMask(const MovingImageMaskType* m) { m_MovingImageMask = m; }
  void SetUseMovingImageGradientFilter(bool on) { m_UseMovingImageGradientFilter = on; }
  void SetVirtualSampledPointSet(const VirtualPointSetType* p) { m_VirtualSampledPointSet = p; }
  void SetUseSampledPointSet(bool on) { m_UseSampledPointSet = on; }
  SizeValueType GetNumberOfValidPoints() const { return m_NumberOfValidPoints; }
  MeasureType GetCurrentValue() const { return m_Value; }
  SizeValueType GetNumberOfParameters() const { return m_MovingTransform->GetNumberOfParameters(); }
  typename VirtualImageType::RegionType GetVirtualRegion() const { return m_VirtualImage->GetBufferedRegion(); }

  // the base class's checks and defaults: identity fixed transform, the fixed
  // image's domain as the virtual domain
  virtual void Initialize() {
    if (m_FixedImage.IsNull()) itkExceptionMacro(<< "FixedImage is not present");
    if (m_MovingImage.IsNull()) itkExceptionMacro(<< "MovingImage is not present");
    if (m_MovingTransform.IsNull()) itkExceptionMacro(<< "MovingTransform is not present");
    if (m_FixedTransform.IsNull()) {
      auto id = IdentityTransform<TInternal, VirtualImageDimension>::New();
      m_FixedTransform = id.GetPointer();
    }
    if (m_VirtualImage.IsNull()) {
      auto v = VirtualImageType::New();
      v->SetRegions(m_FixedImage->GetBufferedRegion());
      v->SetOrigin(m_FixedImage->GetOrigin());
      v->SetSpacing(m_FixedImage->GetSpacing());
      v->SetDirection(m_FixedImage->GetDirection());
      m_VirtualImage = v;
    }
  }
  virtual MeasureType GetValue() const = 0;
  virtual void GetDerivative(DerivativeType&) const = 0;
  virtual void GetValueAndDerivative(MeasureType&, DerivativeType&) const = 0;

 protected:
  ImageToImageMetricv4() : m_DefaultMovingImageGradientFilter(new GradientFilterBase()) {
    m_MovingImageGradientFilter = m_DefaultMovingImageGradientFilter;
  }

  typename TFixedImage::ConstPointer m_FixedImage;
  typename TMovingImage::ConstPointer m_MovingImage;
  typename FixedTransformType::Pointer m_FixedTransform;
  typename MovingTransformType::Pointer m_MovingTransform;
  typename TVirtualImage::Pointer m_VirtualImage;
  bool m_UseMovingImageGradientFilter = true;
  SmartPointer<GradientFilterBase> m_MovingImageGradientFilter;
  SmartPointer<GradientFilterBase> m_DefaultMovingImageGradientFilter;
  mutable typename MovingImageGradientImageType::Pointer m_MovingImageGradientImage;
  typename FixedImageMaskType::ConstPointer m_FixedImageMask;
  typename MovingImageMaskType::ConstPointer m_MovingImageMask;
  bool m_UseSampledPointSet = false;
  typename VirtualPointSetType::ConstPointer m_VirtualSampledPointSet;
  mutable SizeValueType m_NumberOfValidPoints = 0;
  mutable MeasureType m_Value = 0;
};

}  // namespace itk
