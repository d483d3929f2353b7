// Host-side metric sample points of ImageRegistrationMethodv4
// (SetMetricSamplePoints, Modules/Registration/RegistrationMethodsv4/include/
// itkImageRegistrationMethodv4.hxx:867-1024): REGULAR keeps every k-th voxel
// of the virtual region, RANDOM draws voxels with replacement; each kept
// voxel's physical point is moved by a normal variate times a third of the
// virtual spacing per axis and dropped when the fixed mask does not hold it.
//
// The random streams are MT19937 (Matsumoto & Nishimura 1998) as ITK's
// MersenneTwisterRandomVariateGenerator produces them
// (Modules/Core/Common/src/itkMersenneTwisterRandomVariateGenerator.cxx:175-240;
// seeding with the 1812433253 recurrence, the standard tempering), with ITK's
// variate mappings (itkMersenneTwisterRandomVariateGenerator.h:189-226,364-375)
// and ImageRandomConstIteratorWithIndex::RandomJump
// (itkImageRandomConstIteratorWithIndex.hxx:47-68).  The point arithmetic
// rounds in the reference's order (built with -ffp-contract=off).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/itk_mattes_hip.h"

namespace mmi {
int set_last_error(int code, const char* msg);
}

namespace {

constexpr uint32_t kMtN = 624;
constexpr uint32_t kMtM = 397;

class Mt19937 {
 public:
  explicit Mt19937(uint32_t seed) {
    state_[0] = seed;
    for (uint32_t i = 1; i < kMtN; ++i) state_[i] = 1812433253u * (state_[i - 1] ^ (state_[i - 1] >> 30)) + i;
    pos_ = kMtN;  // the first draw regenerates the block
  }

  uint32_t next() {
    if (pos_ == kMtN) regenerate();
    uint32_t y = state_[pos_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
  }

  // (u + 0.5) / 2^32: in (0, 1)
  double open() { return ((double)next() + 0.5) / 4294967296.0; }
  // u / 2^32: in [0, 1)
  double open_upper() { return (double)next() / 4294967296.0; }

  // Box-Muller with the cosine branch only, mean 0 and variance 1, as
  // GetNormalVariate() evaluates it: the radius draw first, then the angle.
  double normal() {
    const double r = std::sqrt(-2.0 * std::log(1.0 - open()) * 1.0);
    const double phi = 2.0 * 3.14159265358979323846 * open_upper();
    return 0.0 + r * std::cos(phi);
  }

 private:
  void regenerate() {
    for (uint32_t i = 0; i < kMtN; ++i) {
      const uint32_t y = (state_[i] & 0x80000000u) | (state_[(i + 1) % kMtN] & 0x7fffffffu);
      state_[i] = state_[(i + kMtM) % kMtN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    pos_ = 0;
  }

  uint32_t state_[kMtN];
  uint32_t pos_;
};

int64_t pixels_of(const mmi_geometry& g) {
  int64_t n = 1;
  for (int i = 0; i < g.dim; ++i) n *= g.size[i];
  return n;
}

// ImageBase::TransformIndexToPhysicalPoint: origin, then += M[i][j] * idx[j].
void index_to_point(const mmi_geometry& g, const int64_t* idx, double* p) {
  const int D = g.dim;
  for (int i = 0; i < D; ++i) {
    p[i] = g.origin[i];
    for (int j = 0; j < D; ++j) p[i] += g.index_to_physical[i * D + j] * (double)idx[j];
  }
}

// ImageMaskSpatialObject::IsInsideInObjectSpace, UseMaskValue off: the
// rounded index (RoundHalfIntegerUp of the continuous index, summed from zero
// in j order) in the buffer and a non-zero pixel there.
bool mask_holds(const mmi_geometry& g, const uint8_t* px, const double* p) {
  const int D = g.dim;
  int64_t off = 0, stride = 1;
  for (int i = 0; i < D; ++i) {
    double s = 0.0;
    for (int j = 0; j < D; ++j) s += g.physical_to_index[i * D + j] * (p[j] - g.origin[j]);
    if (!(std::fabs(s) < 1.0e18)) return false;
    const int64_t k = (int64_t)std::nearbyint(2.0 * s + 0.5) >> 1;
    if (k < g.start[i] || k >= g.start[i] + g.size[i]) return false;
    off += (k - g.start[i]) * stride;
    stride *= g.size[i];
  }
  return px[off] != 0;
}

int64_t sample_bound(const mmi_geometry& v, int32_t strategy, double percentage, uint64_t* per) {
  const int64_t n = pixels_of(v);
  if (strategy == MMI_SAMPLING_REGULAR) {
    const uint64_t every = (uint64_t)std::ceil(1.0 / percentage);
    *per = every;
    return (int64_t)(((uint64_t)n + every - 1) / every);
  }
  const uint64_t count = (uint64_t)((double)(float)n * percentage);
  *per = count;
  return (int64_t)count;
}

}  // namespace

extern "C" {

int mmi_metric_sample_points(const mmi_geometry* virt, int32_t strategy, double percentage, uint32_t* seed,
                             const mmi_geometry* mask_geom, const uint8_t* mask, double* points, int64_t capacity,
                             int64_t* n_points) {
  if (!virt || !seed || !n_points) return mmi::set_last_error(MMI_ERR_ARGUMENT, "virt, seed or n_points is NULL");
  if (virt->dim != 2 && virt->dim != 3) return mmi::set_last_error(MMI_ERR_ARGUMENT, "dimension must be 2 or 3");
  if (strategy != MMI_SAMPLING_REGULAR && strategy != MMI_SAMPLING_RANDOM)
    return mmi::set_last_error(MMI_ERR_ARGUMENT, "Invalid sampling strategy requested.");
  if (!(percentage > 0.0 && percentage <= 1.0))
    return mmi::set_last_error(MMI_ERR_ARGUMENT, "sampling percentage must be in (0, 1]");
  if (mask_geom && (!mask || mask_geom->dim != virt->dim))
    return mmi::set_last_error(MMI_ERR_ARGUMENT, "mask pixels missing or mask dimension differs");
  const mmi_geometry& v = *virt;
  const int D = v.dim;
  uint64_t per = 0;
  const int64_t bound = sample_bound(v, strategy, percentage, &per);
  if (!points) {  // size query
    *n_points = bound;
    return MMI_OK;
  }
  if (capacity < bound) return mmi::set_last_error(MMI_ERR_ARGUMENT, "points capacity below the sample bound");

  double third[3];
  for (int d = 0; d < D; ++d) third[d] = v.spacing[d] / 3.0;
  Mt19937 perturb(*seed);
  *seed += 1;
  int64_t kept = 0;
  auto emit = [&](const int64_t* idx) {
    double p[3];
    index_to_point(v, idx, p);
    for (int d = 0; d < D; ++d) p[d] += perturb.normal() * third[d];
    if (mask_geom && !mask_holds(*mask_geom, mask, p)) return;
    for (int d = 0; d < D; ++d) points[kept * D + d] = p[d];
    ++kept;
  };

  if (strategy == MMI_SAMPLING_REGULAR) {
    // the first voxel is kept, then every per-th in buffer order
    uint64_t count = per;
    int64_t idx[3] = {0, 0, 0};
    const int64_t sz = D == 3 ? v.size[2] : 1;
    for (int64_t z = 0; z < sz; ++z)
      for (int64_t y = 0; y < v.size[1]; ++y)
        for (int64_t x = 0; x < v.size[0]; ++x) {
          if (count == per) {
            count = 0;
            idx[0] = v.start[0] + x;
            idx[1] = v.start[1] + y;
            idx[2] = D == 3 ? v.start[2] + z : 0;
            emit(idx);
          }
          ++count;
        }
  } else {
    Mt19937 jump(*seed);
    *seed += 1;
    const int64_t n = pixels_of(v);
    const double span = (double)n - 0.5;
    for (uint64_t k = 0; k < per; ++k) {
      int64_t pos = (int64_t)(jump.open() * span);
      int64_t idx[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) {
        const int64_t r = pos % v.size[d];
        idx[d] = r + v.start[d];
        pos = (pos - r) / v.size[d];
      }
      emit(idx);
    }
  }
  *n_points = kept;
  return MMI_OK;
}

}  // extern "C"
