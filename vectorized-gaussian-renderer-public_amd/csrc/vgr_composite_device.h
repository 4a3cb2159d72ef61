// Device-side compositing helpers shared by the forward compositor and the
// backward (gradient) compositor: record gathering, the staged splat form with
// its exact log-threshold, the conservative ellipse box, and the per-wave
// ordered sub-lists.  Both passes therefore visit the same (pixel, splat) pairs
// and evaluate them with bit-identical arithmetic.
#ifndef VGR_COMPOSITE_DEVICE_H
#define VGR_COMPOSITE_DEVICE_H

#include "vgr_internal.h"
#include "vgr_project_device.h"

namespace vgr {
namespace dev {

// Native 4-wide vector (first-class in LLVM, so per-thread arrays of it stay in
// registers; HIP's float4 wrapper struct defeats scalar replacement).
typedef float f32x4 __attribute__((ext_vector_type(4)));


__device__ __forceinline__ int find_active_camera(const CameraParams* table, int active_count, int64_t block) {
  int lo = 0, hi = active_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (table[mid].active_tile_begin <= block) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// Loads the records of one batch of the sorted list into registers.  With
// kSlots the sorted values are intersection slots and slot_records maps each
// slot to its record (training forwards, whose backward accumulates per slot).
template <int kThreads, int kBatch, int kPerThread, bool kSlots = false>
__device__ __forceinline__ void gather_batch(const int32_t* __restrict__ sorted_values,
                                             const vgr_record* __restrict__ records, int64_t visible_capacity,
                                             int64_t batch_start, int64_t range_end, int lane,
                                             f32x4 (&a)[kPerThread], f32x4 (&b)[kPerThread],
                                             f32x4 (&c)[kPerThread],
                                             const int32_t* __restrict__ slot_records = nullptr) {
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t index = batch_start + k * kThreads + lane;
    if (k * kThreads + lane < kBatch && index < range_end) {
      int32_t record = sorted_values[index];
      if constexpr (kSlots) record = record >= 0 ? slot_records[record] : -1;
      if (record >= 0 && record < visible_capacity) {
        const f32x4* src = reinterpret_cast<const f32x4*>(records + record);
        a[k] = src[0];
        b[k] = src[1];
        c[k] = src[2];
      } else {
        // Only reachable with a corrupted list: contributes nothing (opacity 0).
        a[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        b[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        c[k] = f32x4{0.0f, 0.0f, __int_as_float(-1), 0.0f};
      }
    }
  }
}

// Margin subtracted from the per-Gaussian log-threshold: far larger than the
// error of vgr_logf/vgr_expf and of the alpha product, so "power below the
// threshold" implies the exact alpha test fails.
constexpr float kThresholdMargin = 1.0e-4f;

// Staged form of one record.  The power of the reference,
//   -0.5 * (c0*dx*dx + 2*c1*dx*dy + c2*dy*dy),
// is evaluated as (A*dx*dx + B*dx*dy) + C*dy*dy with A = -0.5*c0, B = -c1,
// C = -0.5*c2: scaling by powers of two commutes with rounding, so the value is
// bit-identical and two multiplies cheaper.
struct StagedSplat {
  f32x4 geometry;  // mean_x, mean_y, A, B
  f32x4 shading;   // C, opacity, power threshold, depth
  f32x4 colour;    // red, green, blue, gaussian id bits
};

__device__ __forceinline__ StagedSplat stage_record(const f32x4& a, const f32x4& b, const f32x4& c) {
  StagedSplat s;
  const float opacity = b.y;
  // power < threshold  =>  min(0.99, opacity * exp(power)) < 1/255 exactly.
  const float threshold = opacity > 0.0f ? vgr_logf(VGR_ALPHA_THRESHOLD / opacity) - kThresholdMargin
                                         : __int_as_float(0x7f800000);
  s.geometry = f32x4{a.x, a.y, -0.5f * a.z, -a.w};
  s.shading = f32x4{-0.5f * b.x, opacity, threshold, b.z};
  s.colour = f32x4{b.w, c.x, c.y, c.z};
  return s;
}

// Half extents of a box containing every pixel offset whose float-evaluated
// power can reach the threshold: the ellipse d^T C d <= R^2 (R^2 = -2 threshold)
// inflated by a bound on the rounding error of the power expression, which
// grows with the conic's condition.  Returns false when no finite box is
// certain (non-positive-definite or extremely anisotropic conic).
__device__ __forceinline__ bool splat_extent(const f32x4& a, const f32x4& b, float threshold, float& half_x,
                                             float& half_y) {
  const double c0 = a.z, c1 = a.w, c2 = b.x;
  const double det = c0 * c2 - c1 * c1;
  if (!(det > 0.0) || !(threshold < 0.0f)) return false;
  const double sxx = c2 / det;
  const double syy = c0 / det;
  const double condition = (fabs(c0) + 2.0 * fabs(c1) + fabs(c2)) * (sxx + syy);
  if (!(condition < 1.0e5)) return false;
  const double r2 = -2.0 * static_cast<double>(threshold) * (1.0 + 1.0e-6 * condition) * 1.001 + 1.0e-6;
  half_x = static_cast<float>(sqrt(r2 * sxx)) + 1.0e-3f;
  half_y = static_cast<float>(sqrt(r2 * syy)) + 1.0e-3f;
  return true;
}

// Stages one batch of gathered records into LDS and builds, for every wave of
// the tile, the ordered list of batch slots whose conservative ellipse box
// reaches that wave's rows of pixel centres.  A slot left out of a wave's list
// has alpha < 1/255 at every pixel of the wave, so skipping it is exact.
// Slots are ordered by (slice k, source wave, lane) = batch-slot order = depth
// order.  Ends with a __syncthreads(); the caller synchronizes before it (LDS
// of the previous batch fully consumed).
template <int kTileSize, int kBatch>
struct WaveLists {
  static constexpr int kThreads = kTileSize * kTileSize;
  static constexpr int kWaves = kThreads / kWave;
  static constexpr int kPerThread = (kBatch + kThreads - 1) / kThreads;
  // Each wave owns an 8x8 block of the tile (16x16 tiles: four quadrants), the
  // squarest shape of 64 pixels, so a splat's box reaches as few waves as possible.
  static constexpr int kBlock = 8;
  static constexpr int kBlocksX = kTileSize / kBlock;
  __host__ __device__ static constexpr int pixel_dx(int lane) {
    return ((lane / kWave) % kBlocksX) * kBlock + (lane & (kBlock - 1));
  }
  __host__ __device__ static constexpr int pixel_dy(int lane) {
    return ((lane / kWave) / kBlocksX) * kBlock + ((lane & (kWave - 1)) / kBlock);
  }
  f32x4 geometry[kBatch];
  f32x4 shading[kBatch];
  f32x4 colour[kBatch];
  uint16_t list[kWaves][kBatch];
  uint32_t totals[kPerThread][kWaves][kWaves];  // [slice][source wave][list]

  __device__ __forceinline__ uint32_t count(int wave) const {
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < kPerThread; ++k)
#pragma unroll
      for (int q = 0; q < kWaves; ++q) n += totals[k][q][wave];
    return n;
  }
};

template <int kTileSize, int kBatch>
__device__ __forceinline__ void build_wave_lists(WaveLists<kTileSize, kBatch>& s,
                                                 const f32x4 (&na)[WaveLists<kTileSize, kBatch>::kPerThread],
                                                 const f32x4 (&nb)[WaveLists<kTileSize, kBatch>::kPerThread],
                                                 const f32x4 (&nc)[WaveLists<kTileSize, kBatch>::kPerThread],
                                                 int count, int lane, float rect_x0, float rect_x1,
                                                 float rect_y_base) {
  (void)rect_x1;
  using L = WaveLists<kTileSize, kBatch>;
  constexpr int kThreads = L::kThreads;
  constexpr int kWaves = L::kWaves;
  constexpr int kPerThread = L::kPerThread;
  const int wave = lane / kWave;
  const int wave_lane = lane & (kWave - 1);
  uint32_t wave_bits[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int slot = k * kThreads + lane;
    wave_bits[k] = 0u;
    if (slot < count) {
      const StagedSplat staged = stage_record(na[k], nb[k], nc[k]);
      s.geometry[slot] = staged.geometry;
      s.shading[slot] = staged.shading;
      s.colour[slot] = staged.colour;
      float half_x, half_y;
      if (splat_extent(na[k], nb[k], staged.shading.z, half_x, half_y)) {
        const float mx = na[k].x, my = na[k].y;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
          const float x0 = rect_x0 + static_cast<float>((w % L::kBlocksX) * L::kBlock);
          const float y0 = rect_y_base + static_cast<float>((w / L::kBlocksX) * L::kBlock);
          const float gap_x = fabsf(mx - fminf(fmaxf(mx, x0), x0 + static_cast<float>(L::kBlock - 1)));
          const float gap_y = fabsf(my - fminf(fmaxf(my, y0), y0 + static_cast<float>(L::kBlock - 1)));
          if (gap_x <= half_x && gap_y <= half_y) wave_bits[k] |= 1u << w;
        }
      } else {
        wave_bits[k] = (1u << kWaves) - 1u;  // no certain box: every wave evaluates it
      }
    }
  }
  const uint64_t below = wave_lane == 0 ? 0ull : (~0ull >> (kWave - wave_lane));
  uint32_t prefix[kPerThread][kWaves];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint64_t vote = __ballot((wave_bits[k] >> w) & 1u);
      prefix[k][w] = static_cast<uint32_t>(__popcll(vote & below));
      if (wave_lane == 0) s.totals[k][wave][w] = static_cast<uint32_t>(__popcll(vote));
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      if ((wave_bits[k] >> w) & 1u) {
        uint32_t position = prefix[k][w];
        for (int kk = 0; kk < k; ++kk)
          for (int q = 0; q < kWaves; ++q) position += s.totals[kk][q][w];
        for (int q = 0; q < wave; ++q) position += s.totals[k][q][w];
        s.list[w][position] = static_cast<uint16_t>(k * kThreads + lane);
      }
    }
  }
  __syncthreads();
}

// The 48-byte record of (camera, packed Gaussian g), recomputed with the
// projection write kernel's arithmetic (same functions, same operation order),
// so it is bit-identical to the record that kernel would store.
template <int kDegree>
__device__ __forceinline__ void record_from_gaussian(const ProjectionArgs& pa, const CameraParams& cam, int32_t g,
                                                     f32x4& a, f32x4& b, f32x4& c) {
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(pa, g, loaded, m0, m1, m2, opacity, cov);
  Projected p;
  if (!loaded || !project_pair(cam, m0, m1, m2, opacity, cov, pa.near_plane, pa.far_plane, pa.support_sigma,
                               pa.covariance_epsilon, pa.antialiased != 0, p)) {
    a = f32x4{0.0f, 0.0f, 0.0f, 0.0f};  // unreachable for listed pairs: contributes nothing
    b = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    c = f32x4{0.0f, 0.0f, __int_as_float(-1), 0.0f};
    return;
  }
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  float f[kChannels];
  load_features<kDegree>(pa.features, g, pa.feature_width, f);
  float rgb[3];
  sh_colour<kDegree>(f, m0 - cam.origin_x, m1 - cam.origin_y, m2 - cam.origin_z, rgb);
  const int32_t id = pa.gaussian_ids != nullptr ? pa.gaussian_ids[g] : g;
  a = f32x4{p.mean_x, p.mean_y, p.conic_0, p.conic_1};
  b = f32x4{p.conic_2, p.opacity, p.depth, rgb[0]};
  c = f32x4{rgb[1], rgb[2], __int_as_float(id), __int_as_float(cam.camera)};
}

}  // namespace dev
}  // namespace vgr

#endif  // VGR_COMPOSITE_DEVICE_H
