// Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer: the
// oracle (oracle/mattes_oracle.cpp) and the library's host-side sample
// selection (itk_amd/csrc/mattes_sampling.cpp), compiled with g++ from their
// sources into this one program (no GPU code).  Exercises every oracle entry
// point on small 2-D and 3-D problems: dense and sparse sampling, masks,
// affine, B-spline and composite moving transforms, both gradient sources,
// the recursive Gaussian filters, the samplers and the degenerate inputs;
// prints "ok" when it finishes.  Test infrastructure for tests/test_sanitize.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/itk_mattes_hip.h"
#include "../../oracle/mattes_oracle.h"

namespace mmi {
// the library's error channel (mattes_capi.hip), for mattes_sampling.cpp
int set_last_error(int code, const char*) { return code; }
}  // namespace mmi

namespace {

orc_geometry geometry(int dim, const int64_t* size, const double* spacing, const double* origin) {
  orc_geometry g;
  std::memset(&g, 0, sizeof(g));
  g.dim = dim;
  for (int i = 0; i < dim; ++i) {
    g.size[i] = size[i];
    g.spacing[i] = spacing[i];
    g.origin[i] = origin[i];
    g.direction[i * dim + i] = 1.0;
    g.index_to_physical[i * dim + i] = spacing[i];
    g.physical_to_index[i * dim + i] = 1.0 / spacing[i];
  }
  return g;
}

int64_t pixels(const orc_geometry& g) {
  int64_t n = 1;
  for (int i = 0; i < g.dim; ++i) n *= g.size[i];
  return n;
}

std::vector<float> blobs(const orc_geometry& g, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  const int64_t n = pixels(g);
  std::vector<float> px((size_t)n);
  double c[3][4];
  for (auto& b : c)
    for (int k = 0; k < 4; ++k) b[k] = u(rng);
  for (int64_t i = 0; i < n; ++i) {
    int64_t r = i;
    double x[3] = {0, 0, 0};
    for (int d = 0; d < g.dim; ++d) {
      x[d] = (double)(r % g.size[d]) / (double)g.size[d];
      r /= g.size[d];
    }
    double v = 0.1 * u(rng);
    for (auto& b : c) {
      double s = 0.0;
      for (int d = 0; d < g.dim; ++d) s += (x[d] - b[d]) * (x[d] - b[d]);
      v += std::exp(-s / (0.02 + 0.05 * b[3]));
    }
    px[(size_t)i] = (float)v;
  }
  return px;
}

orc_transform affine(int dim, double scale, unsigned seed, const orc_geometry& g) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  orc_transform t;
  std::memset(&t, 0, sizeof(t));
  t.kind = ORC_TF_AFFINE;
  for (int i = 0; i < dim; ++i) {
    t.center[i] = g.origin[i] + 0.5 * g.spacing[i] * (double)(g.size[i] - 1);
    for (int j = 0; j < dim; ++j) t.matrix[i * dim + j] = (i == j ? 1.0 : 0.0) + 0.04 * scale * u(rng);
  }
  for (int i = 0; i < dim; ++i) {
    double off = 2.0 * scale * u(rng) + t.center[i];
    for (int j = 0; j < dim; ++j) off -= t.matrix[i * dim + j] * t.center[j];
    t.offset[i] = off;
  }
  return t;
}

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

void run_dim(int dim) {
  const int64_t size[3] = {22, 20, dim == 3 ? 18 : 1};
  const double spacing[3] = {1.1, 0.9, 1.3}, origin[3] = {-3.0, 2.0, 5.0};
  const orc_geometry g = geometry(dim, size, spacing, origin);
  const std::vector<float> fixed = blobs(g, 1 + dim), moving = blobs(g, 11 + dim);
  const int64_t n = pixels(g);
  char err[256];

  std::vector<double> grad((size_t)(n * dim));
  expect(orc_gradient_recursive_gaussian(&g, moving.data(), 0, 1.0, grad.data(), 2) == 0, "gradient filter");
  std::vector<float> smooth((size_t)n);
  const double sig[3] = {1.0, 1.5, 0.8};
  expect(orc_smooth_recursive_gaussian(&g, moving.data(), 0, sig, smooth.data(), 2) == 0, "smoothing filter");

  orc_options opt;
  std::memset(&opt, 0, sizeof(opt));
  opt.bins = 32;
  opt.compute_derivative = 1;
  opt.threads = 3;
  double lo = 0, hi = 0;
  orc_intensity_range(&g, fixed.data(), 0, nullptr, nullptr, nullptr, 0, nullptr, &lo, &hi);
  opt.fixed_min = lo;
  opt.fixed_max = hi;
  orc_intensity_range(&g, moving.data(), 0, nullptr, nullptr, nullptr, 0, nullptr, &lo, &hi);
  opt.moving_min = lo;
  opt.moving_max = hi;

  orc_transform ident;
  std::memset(&ident, 0, sizeof(ident));
  const orc_transform aff = affine(dim, 1.0, 21 + dim, g);
  std::vector<double> deriv(4096), pdf(32 * 32), fm(32), mm(32), mval((size_t)n);
  std::vector<int16_t> fb((size_t)n), mb((size_t)n);
  orc_result res;
  std::memset(&res, 0, sizeof(res));
  res.derivative = deriv.data();
  res.joint_pdf = pdf.data();
  res.fixed_marginal = fm.data();
  res.moving_marginal = mm.data();
  res.fixed_bin = fb.data();
  res.moving_bin = mb.data();
  res.moving_value = mval.data();
  for (int gs = 0; gs < 2; ++gs) {
    opt.gradient_source = gs;
    expect(orc_mattes_evaluate(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &aff, &opt, &res, err,
                               sizeof(err)) == 0,
           "affine evaluation");
  }

  // B-spline on a 5-node grid per axis, then as a composite after an affine
  const int64_t gsz[3] = {5, 5, dim == 3 ? 5 : 1};
  const double gsp[3] = {6.0, 5.0, 6.5}, gor[3] = {-9.0, -4.0, -2.0};
  orc_transform bs;
  std::memset(&bs, 0, sizeof(bs));
  bs.kind = ORC_TF_BSPLINE;
  bs.grid = geometry(dim, gsz, gsp, gor);
  std::vector<double> coeffs((size_t)(dim * pixels(bs.grid)));
  std::mt19937 rng(31 + dim);
  std::normal_distribution<double> nd(0.0, 0.5);
  for (auto& c : coeffs) c = nd(rng);
  bs.coeffs = coeffs.data();
  opt.gradient_source = 1;
  expect(orc_mattes_evaluate(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &bs, &opt, &res, err,
                             sizeof(err)) == 0,
         "B-spline evaluation");
  orc_transform stages[2] = {affine(dim, 0.5, 41, g), affine(dim, 0.5, 42, g)};
  bs.initial = stages;
  bs.n_initial = 2;
  expect(orc_mattes_evaluate(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &bs, &opt, &res, err,
                             sizeof(err)) == 0,
         "composite B-spline evaluation");
  orc_transform aff2 = aff;
  aff2.initial = stages;
  aff2.n_initial = 1;
  opt.gradient_source = 0;
  expect(orc_mattes_evaluate(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &aff2, &opt, &res, err,
                             sizeof(err)) == 0,
         "composite affine evaluation");

  // sparse points and masks
  std::vector<uint8_t> mask((size_t)n);
  for (int64_t i = 0; i < n; ++i) mask[(size_t)i] = (uint8_t)((i * 7) % 5 != 0);
  uint32_t seed = 121;
  std::vector<double> pts((size_t)(n * dim));
  const int64_t np = orc_metric_sample_points(&g, 2, 0.3, &seed, &g, mask.data(), pts.data());
  expect(np > 0, "random sample points");
  seed = 121;
  const int64_t npr = orc_metric_sample_points(&g, 1, 0.25, &seed, nullptr, nullptr, pts.data());
  expect(npr > 0, "regular sample points");
  orc_sampling smp;
  std::memset(&smp, 0, sizeof(smp));
  smp.points = pts.data();
  smp.n_points = npr;
  smp.fixed_mask_geom = &g;
  smp.fixed_mask = mask.data();
  smp.moving_mask_geom = &g;
  smp.moving_mask = mask.data();
  expect(orc_mattes_evaluate_ex(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &aff, &opt, &smp,
                                &res, err, sizeof(err)) == 0,
         "sparse masked evaluation");

  // degenerate: a constant moving image, and a map that misses the image
  std::vector<float> flat((size_t)n, 2.0f);
  expect(orc_bin_parameters(2.0, 2.0, 32, &lo, &hi, err, sizeof(err)) == 1, "constant image refused");
  orc_transform far = aff;
  for (int i = 0; i < dim; ++i) far.offset[i] += 1e4;
  expect(orc_mattes_evaluate(&g, &g, fixed.data(), &g, moving.data(), grad.data(), &ident, &far, &opt, &res, err,
                             sizeof(err)) == 1,
         "no overlap refused");

  // the library's host sampler (mattes_sampling.cpp)
  mmi_geometry mg;
  std::memset(&mg, 0, sizeof(mg));
  mg.dim = dim;
  for (int i = 0; i < dim; ++i) {
    mg.size[i] = g.size[i];
    mg.origin[i] = g.origin[i];
    mg.spacing[i] = g.spacing[i];
    mg.direction[i * dim + i] = 1.0;
    mg.index_to_physical[i * dim + i] = g.spacing[i];
    mg.physical_to_index[i * dim + i] = 1.0 / g.spacing[i];
  }
  for (int strategy : {MMI_SAMPLING_REGULAR, MMI_SAMPLING_RANDOM}) {
    uint32_t s = 7;
    int64_t bound = 0;
    expect(mmi_metric_sample_points(&mg, strategy, 0.2, &s, nullptr, nullptr, nullptr, 0, &bound) == MMI_OK,
           "sampler size query");
    std::vector<double> out((size_t)(bound * dim) + 1);
    int64_t kept = 0;
    s = 7;
    expect(mmi_metric_sample_points(&mg, strategy, 0.2, &s, &mg, mask.data(), out.data(), bound, &kept) == MMI_OK,
           "sampler");
    expect(kept <= bound, "sampler bound");
    expect(mmi_metric_sample_points(&mg, strategy, 0.2, &s, nullptr, nullptr, out.data(), bound - 1, &kept) ==
               MMI_ERR_ARGUMENT,
           "sampler capacity check");
  }
}

}  // namespace

int main() {
  run_dim(2);
  run_dim(3);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
