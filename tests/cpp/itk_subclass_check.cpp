// Drives itk::MattesMutualInformationImageToImageMetricv4HIP (compiled against
// the test-only ITK interface mock in tests/itk_mock) on images and a
// transform read from a directory, and prints one JSON line: value, valid
// count and derivative, or the exception text.  Test infrastructure for
// tests/test_itk_subclass.py and tests/test_gpu_itk_subclass.py.
//
//   itk_subclass_check <dir>
//
// <dir>/spec.txt:  dim pixel(f32|f64) size[dim] spacing[dim] origin[dim]
//                  kind(affine|bspline) n_fixed fixed[n_fixed] n_params params[n_params]
//                  bins use_filter(0|1) evaluations
//                  then evaluations - 1 parameter vectors of n_params values
//                  then optionally n_initial and, per initial affine transform,
//                  centre[dim] params[dim*dim+dim] (the composite's entries
//                  before the moving transform, in queue order)
// <dir>/fixed.raw, <dir>/moving.raw: pixels in buffer order.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "itkMattesMutualInformationImageToImageMetricv4HIP.h"

namespace {

struct Spec {
  unsigned int dim = 0;
  std::string pixel, kind;
  std::vector<double> size, spacing, origin, fixed, params;
  std::vector<std::vector<double>> more;
  std::vector<std::vector<double>> initial_center, initial_params;
  int bins = 50, use_filter = 1, evaluations = 1;
};

Spec read_spec(const std::string& dir) {
  std::ifstream in(dir + "/spec.txt");
  Spec s;
  in >> s.dim >> s.pixel;
  auto vec = [&](std::vector<double>& v, size_t n) {
    v.resize(n);
    for (auto& x : v) in >> x;
  };
  vec(s.size, s.dim);
  vec(s.spacing, s.dim);
  vec(s.origin, s.dim);
  size_t n = 0;
  in >> s.kind >> n;
  vec(s.fixed, n);
  in >> n;
  vec(s.params, n);
  in >> s.bins >> s.use_filter >> s.evaluations;
  for (int e = 1; e < s.evaluations; ++e) {
    std::vector<double> p;
    vec(p, n);
    s.more.push_back(p);
  }
  if (!in) throw std::runtime_error("bad spec.txt in " + dir);
  size_t n_init = 0;
  if (in >> n_init) {
    for (size_t i = 0; i < n_init; ++i) {
      std::vector<double> c, p;
      vec(c, s.dim);
      vec(p, s.dim * s.dim + s.dim);
      s.initial_center.push_back(c);
      s.initial_params.push_back(p);
    }
    if (!in) throw std::runtime_error("bad initial transforms in " + dir + "/spec.txt");
  }
  return s;
}

template <typename TImage>
typename TImage::Pointer read_image(const std::string& path, const Spec& s) {
  auto image = TImage::New();
  typename TImage::RegionType region;
  typename TImage::PointType origin;
  typename TImage::SpacingType spacing;
  for (unsigned int i = 0; i < s.dim; ++i) {
    region.m_Size[i] = (itk::SizeValueType)s.size[i];
    origin[i] = s.origin[i];
    spacing[i] = s.spacing[i];
  }
  image->SetRegions(region);
  image->SetOrigin(origin);
  image->SetSpacing(spacing);
  image->Allocate();
  std::ifstream in(path, std::ios::binary);
  in.read(reinterpret_cast<char*>(image->GetBufferPointer()),
          (std::streamsize)(region.GetNumberOfPixels() * sizeof(typename TImage::PixelType)));
  if (!in) throw std::runtime_error("short read of " + path);
  return image;
}

void print_result(double value, itk::SizeValueType valid, const itk::Array<double>& d) {
  std::printf("{\"value\": %.17g, \"n_valid\": %lu, \"derivative\": [", value, (unsigned long)valid);
  for (itk::SizeValueType i = 0; i < d.Size(); ++i) std::printf("%s%.17g", i ? ", " : "", d[i]);
  std::printf("]}\n");
}

template <typename TPixel, unsigned int D>
int run(const std::string& dir, const Spec& s) {
  using ImageType = itk::Image<TPixel, D>;
  using MetricType = itk::MattesMutualInformationImageToImageMetricv4HIP<ImageType, ImageType>;
  auto fixed = read_image<ImageType>(dir + "/fixed.raw", s);
  auto moving = read_image<ImageType>(dir + "/moving.raw", s);
  typename MetricType::MovingTransformType::Pointer tf;
  if (s.kind == "affine") {
    auto a = itk::AffineTransform<double, D>::New();
    tf = a.GetPointer();
  } else {
    auto b = itk::BSplineTransform<double, D, 3>::New();
    tf = b.GetPointer();
  }
  itk::Array<double> fp(s.fixed.size()), p(s.params.size());
  for (size_t i = 0; i < s.fixed.size(); ++i) fp[i] = s.fixed[i];
  for (size_t i = 0; i < s.params.size(); ++i) p[i] = s.params[i];
  tf->SetFixedParameters(fp);
  tf->SetParameters(p);
  // the registration method's output transform: a CompositeTransform holding it
  auto composite = itk::CompositeTransform<double, D>::New();
  for (size_t i = 0; i < s.initial_center.size(); ++i) {
    auto a = itk::AffineTransform<double, D>::New();
    itk::Array<double> c(D), q(D * D + D);
    for (unsigned int k = 0; k < D; ++k) c[k] = s.initial_center[i][k];
    for (unsigned int k = 0; k < D * D + D; ++k) q[k] = s.initial_params[i][k];
    a->SetFixedParameters(c);
    a->SetParameters(q);
    composite->AddTransform(a.GetPointer());
  }
  composite->AddTransform(tf.GetPointer());
  composite->SetOnlyMostRecentTransformToOptimizeOn();
  auto metric = MetricType::New();
  metric->SetFixedImage(fixed.GetPointer());
  metric->SetMovingImage(moving.GetPointer());
  metric->SetMovingTransform(composite.GetPointer());
  metric->SetNumberOfHistogramBins((itk::SizeValueType)s.bins);
  metric->SetUseMovingImageGradientFilter(s.use_filter != 0);
  try {
    metric->Initialize();
    typename MetricType::DerivativeType d;
    double value = 0;
    metric->GetValueAndDerivative(value, d);
    print_result(value, metric->GetNumberOfValidPoints(), d);
    for (const auto& q : s.more) {
      for (size_t i = 0; i < q.size(); ++i) p[i] = q[i];
      composite->SetParameters(p);
      metric->GetValueAndDerivative(value, d);
      print_result(value, metric->GetNumberOfValidPoints(), d);
    }
  } catch (const itk::ExceptionObject& e) {
    std::string msg = e.what();
    for (auto& ch : msg)
      if (ch == '"' || ch == '\\' || ch == '\n') ch = ' ';
    std::printf("{\"error\": \"%s\"}\n", msg.c_str());
    return 3;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const Spec s = read_spec(dir);
  if (s.dim == 2 && s.pixel == "f32") return run<float, 2>(dir, s);
  if (s.dim == 2 && s.pixel == "f64") return run<double, 2>(dir, s);
  if (s.dim == 3 && s.pixel == "f32") return run<float, 3>(dir, s);
  if (s.dim == 3 && s.pixel == "f64") return run<double, 3>(dir, s);
  std::fprintf(stderr, "unsupported dim/pixel\n");
  return 2;
}
