// Device-side projection helpers shared by the forward projection/emission
// kernels and the backward (gradient) kernels, so both passes make the same
// visibility, radius and colour decisions from the same arithmetic.
#ifndef VGR_PROJECT_DEVICE_H
#define VGR_PROJECT_DEVICE_H

#include "vgr_internal.h"

namespace vgr {
namespace dev {

constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f;
constexpr float kShC2_1 = -1.0925484305920792f;
constexpr float kShC2_2 = 0.31539156525252005f;
constexpr float kShC2_3 = -1.0925484305920792f;
constexpr float kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f;
constexpr float kShC3_1 = 2.890611442640554f;
constexpr float kShC3_2 = -0.4570457994644658f;
constexpr float kShC3_3 = 0.3731763325901154f;
constexpr float kShC3_4 = -0.4570457994644658f;
constexpr float kShC3_5 = 1.445305721320277f;
constexpr float kShC3_6 = -0.5900435899266435f;


__device__ __forceinline__ int float_to_int_clamped(float value) {
  // ceilf/floorf results of finite splat extents; clamp keeps the int
  // conversion defined for pathological covariances.
  if (!(value < 1073741824.0f)) return 1073741824;
  if (!(value > -1073741824.0f)) return -1073741824;
  return static_cast<int>(value);
}

struct Projected {
  float mean_x, mean_y;
  float conic_0, conic_1, conic_2;
  float depth, opacity;
  int radius_x, radius_y;
};

// Screen-space projection of one (camera, Gaussian) pair.  Operation order
// follows renderer_cuda.cu:614-1055 term by term (left-to-right sums, no
// contraction), with vgr_logf in place of __logf.
// kRay: exact-ray renders cull on the support quadric's screen ellipse instead
// (ray_extent), so the screen-space cull of the splat is skipped here.
template <bool kRay = false>
__device__ __forceinline__ bool project_pair(
    const CameraParams& cam, float m0, float m1, float m2, float opacity,
    const float* cov, float near_plane, float far_plane, float support_sigma,
    float covariance_epsilon, bool antialiased, Projected& out) {
  if (!(opacity >= VGR_ALPHA_THRESHOLD)) return false;
  const float* e = cam.env;
  const float env_x = e[0] * m0 + e[1] * m1 + e[2] * m2 + e[3];
  const float env_y = e[4] * m0 + e[5] * m1 + e[6] * m2 + e[7];
  const float env_z = e[8] * m0 + e[9] * m1 + e[10] * m2 + e[11];
  const float* v = cam.view;
  const float camera_x = v[0] * env_x + v[1] * env_y + v[2] * env_z + v[3];
  const float camera_y = v[4] * env_x + v[5] * env_y + v[6] * env_z + v[7];
  const float camera_z = v[8] * env_x + v[9] * env_y + v[10] * env_z + v[11];
  if (camera_z < near_plane || camera_z > far_plane) return false;

  const float inverse_z = 1.0f / camera_z;
  const float mean_x = cam.focal_x * camera_x * inverse_z + cam.center_x;
  const float mean_y = cam.focal_y * camera_y * inverse_z + cam.center_y;

  const float* r = cam.rotation;
  float rc[9];
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    const float r0 = r[row * 3 + 0];
    const float r1 = r[row * 3 + 1];
    const float r2 = r[row * 3 + 2];
    rc[row * 3 + 0] = r0 * cov[0] + r1 * cov[1] + r2 * cov[2];
    rc[row * 3 + 1] = r0 * cov[1] + r1 * cov[3] + r2 * cov[4];
    rc[row * 3 + 2] = r0 * cov[2] + r1 * cov[4] + r2 * cov[5];
  }
  const float c00 = rc[0] * r[0] + rc[1] * r[1] + rc[2] * r[2];
  const float c01 = rc[0] * r[3] + rc[1] * r[4] + rc[2] * r[5];
  const float c02 = rc[0] * r[6] + rc[1] * r[7] + rc[2] * r[8];
  const float c11 = rc[3] * r[3] + rc[4] * r[4] + rc[5] * r[5];
  const float c12 = rc[3] * r[6] + rc[4] * r[7] + rc[5] * r[8];
  const float c22 = rc[6] * r[6] + rc[7] * r[7] + rc[8] * r[8];

  const float clamped_x = camera_z * fminf(cam.limit_x_pos, fmaxf(-cam.limit_x_neg, camera_x * inverse_z));
  const float clamped_y = camera_z * fminf(cam.limit_y_pos, fmaxf(-cam.limit_y_neg, camera_y * inverse_z));
  const float inverse_z_squared = inverse_z * inverse_z;
  const float j00 = cam.focal_x * inverse_z;
  const float j02 = -cam.focal_x * clamped_x * inverse_z_squared;
  const float j11 = cam.focal_y * inverse_z;
  const float j12 = -cam.focal_y * clamped_y * inverse_z_squared;
  const float raw00 = j00 * j00 * c00 + 2.0f * j00 * j02 * c02 + j02 * j02 * c22;
  const float cov01 = j00 * j11 * c01 + j00 * j12 * c02 + j02 * j11 * c12 + j02 * j12 * c22;
  const float raw11 = j11 * j11 * c11 + 2.0f * j11 * j12 * c12 + j12 * j12 * c22;
  const float cov00 = raw00 + covariance_epsilon;
  const float cov11 = raw11 + covariance_epsilon;
  const float determinant = cov00 * cov11 - cov01 * cov01;
  if (!(determinant > 0.0f)) return false;
  if (antialiased) {
    const float original = raw00 * raw11 - cov01 * cov01;
    const float compensation = sqrtf(fmaxf(
        VGR_MIN_ANTIALIAS_COMPENSATION * VGR_MIN_ANTIALIAS_COMPENSATION,
        original / determinant));
    opacity *= compensation;
    if (!(opacity >= VGR_ALPHA_THRESHOLD)) return false;
  }
  const float opacity_extend = sqrtf(2.0f * vgr_logf(opacity / VGR_ALPHA_THRESHOLD));
  const float extend = fminf(support_sigma, opacity_extend);
  const int radius_x = float_to_int_clamped(ceilf(extend * sqrtf(cov00)));
  const int radius_y = float_to_int_clamped(ceilf(extend * sqrtf(cov11)));
  if (radius_x <= 0 && radius_y <= 0) return false;
  if (!kRay && (mean_x + static_cast<float>(radius_x) <= 0.0f ||
      mean_y + static_cast<float>(radius_y) <= 0.0f ||
      mean_x - static_cast<float>(radius_x) >= static_cast<float>(cam.width) ||
      mean_y - static_cast<float>(radius_y) >= static_cast<float>(cam.height))) {
    return false;
  }
  const float inverse_determinant = 1.0f / determinant;
  out.mean_x = mean_x;
  out.mean_y = mean_y;
  out.conic_0 = cov11 * inverse_determinant;
  out.conic_1 = -cov01 * inverse_determinant;
  out.conic_2 = cov00 * inverse_determinant;
  out.depth = camera_z;
  out.opacity = fminf(fmaxf(opacity, 0.0f), 1.0f);
  out.radius_x = radius_x;
  out.radius_y = radius_y;
  return true;
}

// ---- exact-ray evaluation ------------------------------------------------------

// Camera-space precision of a splat for the ray test: P = Sigma^-1 (6), P mu (3),
// mu^T P mu (1), as stored in workspace.ray_records.
struct RayData {
  float precision[6];  // 00, 01, 02, 11, 12, 22
  float precision_mean[3];
  float mean_precision_mean;
  // Half extents (pixels) of a box around the binning centre holding every pixel
  // whose float-evaluated m^2 can reach the support: the screen ellipse of a
  // slightly larger support, sigma_c^2 = sigma^2 (1 + 1e-4) + 1e-5 mu^T P mu, plus
  // one pixel.  The compositor skips rows of pixels outside it.
  float cull_half[2];
};

// Screen ellipse {r = (x, y, 1): r^T (q P - P mu mu^T P) r >= 0} of the rays whose
// closest approach to the mean lies within the support (q = mu^T P mu - sigma^2):
// centre and half extents in normalized ray coordinates; false when unbounded.
__device__ __forceinline__ bool support_ellipse(double p00, double p01, double p02, double p11, double p12,
                                                double p22, double pm0, double pm1, double pm2, double q,
                                                double& centre_x, double& centre_y, double& extent_x,
                                                double& extent_y) {
  const double k00 = q * p00 - pm0 * pm0;
  const double k01 = q * p01 - pm0 * pm1;
  const double k02 = q * p02 - pm0 * pm2;
  const double k11 = q * p11 - pm1 * pm1;
  const double k12 = q * p12 - pm1 * pm2;
  const double k22 = q * p22 - pm2 * pm2;
  const double k_det = k00 * k11 - k01 * k01;
  if (!(k00 > 0.0 && k_det > 0.0)) return false;
  const double inverse_k = 1.0 / k_det;
  centre_x = (k01 * k12 - k11 * k02) * inverse_k;
  centre_y = (k01 * k02 - k00 * k12) * inverse_k;
  const double scale = -(k22 + k02 * centre_x + k12 * centre_y);
  if (!(scale > 0.0)) return false;
  extent_x = sqrt(scale * k11 * inverse_k);
  extent_y = sqrt(scale * k00 * inverse_k);
  return true;
}

// Exact-ray extent of a projected pair (renderer_cuda.cu:795-956): the camera-
// space precision in double, the screen ellipse of the rays whose closest point
// lies within support_sigma (the quadric r^T (q P - P mu mu^T P) r with
// q = mu^T P mu - sigma^2, centre and half extents), which replaces the splat's
// mean and radii for binning and culling; a degenerate ellipse covers the image.
// Returns false when the covariance is singular or the ellipse misses the image.
__device__ __forceinline__ bool ray_extent(const CameraParams& cam, float m0, float m1, float m2, const float* cov,
                                           float support_sigma, Projected& p, RayData& ray) {
  // Camera-space mean and covariance with project_pair's operations.
  const float* e = cam.env;
  const float env_x = e[0] * m0 + e[1] * m1 + e[2] * m2 + e[3];
  const float env_y = e[4] * m0 + e[5] * m1 + e[6] * m2 + e[7];
  const float env_z = e[8] * m0 + e[9] * m1 + e[10] * m2 + e[11];
  const float* v = cam.view;
  const float camera_x = v[0] * env_x + v[1] * env_y + v[2] * env_z + v[3];
  const float camera_y = v[4] * env_x + v[5] * env_y + v[6] * env_z + v[7];
  const float camera_z = v[8] * env_x + v[9] * env_y + v[10] * env_z + v[11];
  const float* r = cam.rotation;
  float rc[9];
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    const float r0 = r[row * 3 + 0];
    const float r1 = r[row * 3 + 1];
    const float r2 = r[row * 3 + 2];
    rc[row * 3 + 0] = r0 * cov[0] + r1 * cov[1] + r2 * cov[2];
    rc[row * 3 + 1] = r0 * cov[1] + r1 * cov[3] + r2 * cov[4];
    rc[row * 3 + 2] = r0 * cov[2] + r1 * cov[4] + r2 * cov[5];
  }
  const double c00 = rc[0] * r[0] + rc[1] * r[1] + rc[2] * r[2];
  const double c01 = rc[0] * r[3] + rc[1] * r[4] + rc[2] * r[5];
  const double c02 = rc[0] * r[6] + rc[1] * r[7] + rc[2] * r[8];
  const double c11 = rc[3] * r[3] + rc[4] * r[4] + rc[5] * r[5];
  const double c12 = rc[3] * r[6] + rc[4] * r[7] + rc[5] * r[8];
  const double c22 = rc[6] * r[6] + rc[7] * r[7] + rc[8] * r[8];
  const double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c02 * c12) + c02 * (c01 * c12 - c02 * c11);
  if (!(det > 1.0e-48)) return false;
  const double inverse_det = 1.0 / det;
  const double p00 = (c11 * c22 - c12 * c12) * inverse_det;
  const double p01 = (c02 * c12 - c01 * c22) * inverse_det;
  const double p02 = (c01 * c12 - c02 * c11) * inverse_det;
  const double p11 = (c00 * c22 - c02 * c02) * inverse_det;
  const double p12 = (c01 * c02 - c00 * c12) * inverse_det;
  const double p22 = (c00 * c11 - c01 * c01) * inverse_det;
  const double x = camera_x, y = camera_y, z = camera_z;
  const double pm0 = p00 * x + p01 * y + p02 * z;
  const double pm1 = p01 * x + p11 * y + p12 * z;
  const double pm2 = p02 * x + p12 * y + p22 * z;
  const double mpm = x * pm0 + y * pm1 + z * pm2;
  ray.precision[0] = static_cast<float>(p00);
  ray.precision[1] = static_cast<float>(p01);
  ray.precision[2] = static_cast<float>(p02);
  ray.precision[3] = static_cast<float>(p11);
  ray.precision[4] = static_cast<float>(p12);
  ray.precision[5] = static_cast<float>(p22);
  ray.precision_mean[0] = static_cast<float>(pm0);
  ray.precision_mean[1] = static_cast<float>(pm1);
  ray.precision_mean[2] = static_cast<float>(pm2);
  ray.mean_precision_mean = static_cast<float>(mpm);

  const double support2 = static_cast<double>(support_sigma) * static_cast<double>(support_sigma);
  const double fx = fabs(static_cast<double>(cam.focal_x)), fy = fabs(static_cast<double>(cam.focal_y));
  double centre_x, centre_y, extent_x, extent_y;
  bool finite = support_ellipse(p00, p01, p02, p11, p12, p22, pm0, pm1, pm2, mpm - support2, centre_x, centre_y,
                                extent_x, extent_y);
  if (finite) {
    const double mean_x = static_cast<double>(cam.focal_x) * centre_x + static_cast<double>(cam.center_x);
    const double mean_y = static_cast<double>(cam.focal_y) * centre_y + static_cast<double>(cam.center_y);
    p.mean_x = static_cast<float>(mean_x);
    p.mean_y = static_cast<float>(mean_y);
    p.radius_x = max(1, float_to_int_clamped(ceilf(static_cast<float>(fx * extent_x))));
    p.radius_y = max(1, float_to_int_clamped(ceilf(static_cast<float>(fy * extent_y))));
    finite = isfinite(mean_x) && isfinite(mean_y) && isfinite(extent_x) && isfinite(extent_y);
  }
  if (!finite) {
    p.mean_x = 0.5f * static_cast<float>(cam.width);
    p.mean_y = 0.5f * static_cast<float>(cam.height);
    p.radius_x = p.radius_y = max(cam.width, cam.height);
  }
  // Cull box of the compositor (never used for binning): the ellipse of the
  // enlarged support, measured from the binning centre; unbounded covers all.
  ray.cull_half[0] = ray.cull_half[1] = 3.0e38f;
  double cc_x, cc_y, ce_x, ce_y;
  if (finite && support_ellipse(p00, p01, p02, p11, p12, p22, pm0, pm1, pm2,
                                mpm - (support2 * (1.0 + 1.0e-4) + 1.0e-5 * fabs(mpm)), cc_x, cc_y, ce_x, ce_y)) {
    const double half_x = fabs(static_cast<double>(cam.focal_x) * cc_x + static_cast<double>(cam.center_x) -
                               static_cast<double>(p.mean_x)) + fx * ce_x + 1.0;
    const double half_y = fabs(static_cast<double>(cam.focal_y) * cc_y + static_cast<double>(cam.center_y) -
                               static_cast<double>(p.mean_y)) + fy * ce_y + 1.0;
    if (half_x < 1.0e30 && half_y < 1.0e30) {
      ray.cull_half[0] = static_cast<float>(half_x) * (1.0f + 1.0e-6f);
      ray.cull_half[1] = static_cast<float>(half_y) * (1.0f + 1.0e-6f);
    }
  }
  return !(p.mean_x + static_cast<float>(p.radius_x) <= 0.0f || p.mean_y + static_cast<float>(p.radius_y) <= 0.0f ||
           p.mean_x - static_cast<float>(p.radius_x) >= static_cast<float>(cam.width) ||
           p.mean_y - static_cast<float>(p.radius_y) >= static_cast<float>(cam.height));
}

struct TileRect {
  int min_x, min_y, max_x, max_y;
};

// Tile rectangle of a projected splat (renderer_cuda.cu:1729-1748).
__device__ __forceinline__ TileRect tile_rect(const Projected& p, int tile_size, int tiles_x, int tiles_y) {
  const float ts = static_cast<float>(tile_size);
  TileRect rect;
  rect.min_x = max(0, static_cast<int>(floorf(p.mean_x / ts - static_cast<float>(p.radius_x) / ts)));
  rect.min_y = max(0, static_cast<int>(floorf(p.mean_y / ts - static_cast<float>(p.radius_y) / ts)));
  rect.max_x = min(tiles_x, static_cast<int>(ceilf(p.mean_x / ts + static_cast<float>(p.radius_x) / ts)));
  rect.max_y = min(tiles_y, static_cast<int>(ceilf(p.mean_y / ts + static_cast<float>(p.radius_y) / ts)));
  return rect;
}

__device__ __forceinline__ int rect_tiles(const TileRect& rect) {
  if (rect.min_x >= rect.max_x || rect.min_y >= rect.max_y) return 0;
  return (rect.max_x - rect.min_x) * (rect.max_y - rect.min_y);
}

// View-dependent colour: base RGB plus real-SH bands 1..degree evaluated at the
// unit direction from the camera centre to the Gaussian (3DGS basis/constants),
// clamped at zero.  f[3k + c] holds band k, channel c (held in registers).
template <int kDegree>
__device__ __forceinline__ void sh_colour(const float (&f)[3 * (kDegree + 1) * (kDegree + 1)], float dx, float dy,
                                          float dz, float* rgb) {
  constexpr int kBands = (kDegree + 1) * (kDegree + 1);
  rgb[0] = f[0];
  rgb[1] = f[1];
  rgb[2] = f[2];
  if (kDegree <= 0) return;
  const float norm2 = dx * dx + dy * dy + dz * dz;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (norm2 > 0.0f) {
    const float inverse_norm = 1.0f / sqrtf(norm2);
    x = dx * inverse_norm;
    y = dy * inverse_norm;
    z = dz * inverse_norm;
  }
  float basis[16];
  basis[1] = -kShC1 * y;
  basis[2] = kShC1 * z;
  basis[3] = -kShC1 * x;
  if (kDegree >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    basis[4] = kShC2_0 * xy;
    basis[5] = kShC2_1 * yz;
    basis[6] = kShC2_2 * (2.0f * zz - xx - yy);
    basis[7] = kShC2_3 * xz;
    basis[8] = kShC2_4 * (xx - yy);
    if (kDegree >= 3) {
      basis[9] = kShC3_0 * y * (3.0f * xx - yy);
      basis[10] = kShC3_1 * xy * z;
      basis[11] = kShC3_2 * y * (4.0f * zz - xx - yy);
      basis[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      basis[13] = kShC3_4 * x * (4.0f * zz - xx - yy);
      basis[14] = kShC3_5 * z * (xx - yy);
      basis[15] = kShC3_6 * x * (xx - 3.0f * yy);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 1; k < kBands; ++k) acc = fmaf(basis[k], f[3 * k + c], acc);
    rgb[c] = fmaxf(f[c] + acc, 0.0f);
  }
}

// Loads the 3*(degree+1)^2 colour channels of Gaussian g into registers, with
// 16-byte loads when the row stride allows it.
template <int kDegree>
__device__ __forceinline__ void load_features(const float* __restrict__ features, int64_t g, int width,
                                              float (&f)[3 * (kDegree + 1) * (kDegree + 1)]) {
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  const float* row = features + g * width;
  if (kChannels % 4 == 0 && (width & 3) == 0 && (reinterpret_cast<uintptr_t>(features) & 15u) == 0) {
    const float4* row4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int i = 0; i < kChannels / 4; ++i) {
      const float4 v = row4[i];
      f[4 * i + 0] = v.x;
      f[4 * i + 1] = v.y;
      f[4 * i + 2] = v.z;
      f[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kChannels; ++i) f[i] = row[i];
  }
}

struct ProjectionArgs {
  const float* means;
  const float* covariances;
  const float* opacities;
  const float* features;
  const CameraParams* table;
  int64_t gaussian_count;
  int64_t chunks;
  int active_count;
  int cameras_per_block;
  int feature_width;
  int sh_degree;
  int tile_size;
  int antialiased;
  float near_plane, far_plane, support_sigma, covariance_epsilon;
  const int32_t* gaussian_ids;   // registration-order id per packed entry, or null (identity)
  const int32_t* item_blocks;    // kept (group, chunk) blocks in order, or null (every block)
  const uint64_t* item_masks;    // cameras of the group that may see the chunk
  const int64_t* item_count;     // device count of kept blocks
  // Dense (tile size 1) passes: 0 records and pixel intersections together, 1
  // records only (the cutoff pass needs them first), 2 pixel intersections only.
  int dense_mode;
  int depth_buckets;             // depth buckets of the cutoff (depth_cutoff non-null)
  const int32_t* depth_cutoff;   // per global pixel: last depth bucket emitted, or null (no cutoff)
  const int32_t* dense_block_cutoff;  // per 8x8 pixel block: largest cutoff of its pixels (with depth_cutoff)
};

// ---- dense (tile size 1) per-pixel decisions --------------------------------------

constexpr float kDenseMinPower = -20.0f;
// Fixed-point scale of the optical-thickness sums and the cutoff threshold
// ceil(9.210340371976184 * 65536) (transmittance 1e-4, renderer_cuda.cu:1672).
constexpr float kTauScale = 65536.0f;
constexpr uint64_t kTauThresholdFixed = 603610ull;

// Depth bucket of a splat: uniform buckets over [near, far] (renderer_cuda.cu:168-177).
__device__ __forceinline__ int depth_bucket(float depth, float near_plane, float far_plane, int buckets) {
  const float normalized = (depth - near_plane) / (far_plane - near_plane);
  const int index = static_cast<int>(normalized * static_cast<float>(buckets));
  return min(buckets - 1, max(0, index));
}

// Alpha of a splat at the centre of pixel (x, y), 0 below 1/255
// (renderer_cuda.cu:179-200, with vgr_expf for the exponential).
__device__ __forceinline__ float pixel_alpha(const Projected& p, int x, int y) {
  const float dx = (static_cast<float>(x) + 0.5f) - p.mean_x;
  const float dy = (static_cast<float>(y) + 0.5f) - p.mean_y;
  const float power = -0.5f * (p.conic_0 * dx * dx + 2.0f * p.conic_1 * dx * dy + p.conic_2 * dy * dy);
  if (power > 0.0f || power < kDenseMinPower) return 0.0f;
  const float alpha = fminf(VGR_MAX_ALPHA, p.opacity * vgr_expf(power));
  return alpha >= VGR_ALPHA_THRESHOLD ? alpha : 0.0f;
}

// Conservative optical thickness of one splat: 0.999 x the first eight terms of
// -log(1 - alpha) = sum alpha^k / k (renderer_cuda.cu:247-257), in 1/65536 units
// rounded down.
__device__ __forceinline__ uint32_t optical_thickness_fixed(float alpha) {
  float term = alpha;
  float result = term;
#pragma unroll
  for (int order = 2; order <= 8; ++order) {
    term *= alpha;
    result += term / static_cast<float>(order);
  }
  return static_cast<uint32_t>((result * 0.999f) * kTauScale);
}

// Resolves the (group, chunk) block and camera mask processed by this workgroup.
// Returns false for workgroups beyond the kept item count.
__device__ __forceinline__ bool resolve_block(const ProjectionArgs& args, int64_t& block, uint64_t& mask) {
  block = blockIdx.x;
  mask = ~0ull;
  if (args.item_blocks == nullptr) return true;
  if (block >= *args.item_count) return false;
  mask = args.item_masks[block];
  block = args.item_blocks[block];
  return true;
}

// Pair results for one camera of the group; identical in both passes.
__device__ __forceinline__ bool evaluate_pair(const ProjectionArgs& args, const CameraParams& cam, int64_t g,
                                              bool loaded, float m0, float m1, float m2, float opacity,
                                              const float* cov, Projected& p, TileRect& rect, int& tiles) {
  tiles = 0;
  if (!loaded || cam.scene_slot < 0 || g < cam.gaussian_begin || g >= cam.gaussian_end) return false;
  if (!project_pair(cam, m0, m1, m2, opacity, cov, args.near_plane, args.far_plane, args.support_sigma,
                    args.covariance_epsilon, args.antialiased != 0, p)) {
    return false;
  }
  rect = tile_rect(p, args.tile_size, cam.tiles_x, cam.tiles_y);
  tiles = rect_tiles(rect);
  return true;
}


// Per-pixel intersections of a wave's splats at tile size 1 (renderer_cuda.cu:1764-1805:
// the pixels of a splat's rectangle where its alpha reaches 1/255), found cooperatively:
// the wave walks the union of its splats' rectangles 64 pixels at a time in
// (lane, row-major) order, pixel slot j going to the lane whose scanned
// rectangle range holds it (binary search over the lanes' inclusive areas), so
// large and small splats share the work.  visit(keep, index, owner, pixel, alpha)
// is called by every lane for every 64-slot step (uniform control flow, so it
// may shuffle): keep marks a pixel where the owner's alpha reaches 1/255 and that
// lies in front of the pixel's depth cutoff (when cutoff is not null), index is
// its position among the wave's kept pixels in that order — the order a per-lane
// row-major loop would emit them in.  Per-lane inputs: the splat (seen), its
// pixel rectangle, its camera's tile base and width in tiles, its depth bucket.
// Returns the wave's kept total.
template <typename Visit>
__device__ __forceinline__ int dense_wave_pixels(bool seen, const Projected& p, const TileRect& rect,
                                                 uint32_t tile_base, int tiles_x, int bucket,
                                                 const int32_t* __restrict__ cutoff, Visit&& visit) {
  const int lane = threadIdx.x & (kWave - 1);
  const int width = seen ? rect.max_x - rect.min_x : 0;
  const int area = seen ? width * (rect.max_y - rect.min_y) : 0;
  int inclusive = area;
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const int t = __shfl_up(inclusive, offset, kWave);
    if (lane >= offset) inclusive += t;
  }
  const int total = __shfl(inclusive, kWave - 1, kWave);
  const int exclusive = inclusive - area;
  const int min_x = seen ? rect.min_x : 0;
  const int min_y = seen ? rect.min_y : 0;
  const float mean_x = seen ? p.mean_x : 0.0f, mean_y = seen ? p.mean_y : 0.0f;
  const float conic_0 = seen ? p.conic_0 : 0.0f, conic_1 = seen ? p.conic_1 : 0.0f;
  const float conic_2 = seen ? p.conic_2 : 0.0f, opacity = seen ? p.opacity : 0.0f;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  int kept = 0;
  for (int run = 0; run < total; run += kWave) {  // uniform over the wave
    const int j = run + lane;
    int owner = 0;
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
      if (__shfl(inclusive, owner + step - 1, kWave) <= j) owner += step;
    }
    const bool in = j < total;
    const int local = j - __shfl(exclusive, owner, kWave);
    const int row_width = __shfl(width, owner, kWave);
    const int row = row_width > 0 ? local / row_width : 0;
    const int x = __shfl(min_x, owner, kWave) + (local - row * row_width);
    const int y = __shfl(min_y, owner, kWave) + row;
    Projected q;
    q.mean_x = __shfl(mean_x, owner, kWave);
    q.mean_y = __shfl(mean_y, owner, kWave);
    q.conic_0 = __shfl(conic_0, owner, kWave);
    q.conic_1 = __shfl(conic_1, owner, kWave);
    q.conic_2 = __shfl(conic_2, owner, kWave);
    q.opacity = __shfl(opacity, owner, kWave);
    const uint32_t owner_base = static_cast<uint32_t>(__shfl(static_cast<int>(tile_base), owner, kWave));
    const int owner_tiles_x = __shfl(tiles_x, owner, kWave);
    const int owner_bucket = __shfl(bucket, owner, kWave);
    const float alpha = in ? pixel_alpha(q, x, y) : 0.0f;
    const uint32_t pixel = in ? owner_base + static_cast<uint32_t>(y * owner_tiles_x + x) : 0u;
    bool keep = alpha != 0.0f;
    if (keep && cutoff != nullptr && owner_bucket > cutoff[pixel]) keep = false;
    const uint64_t vote = __ballot(keep);
    visit(keep, kept + static_cast<int>(__popcll(vote & below)), owner, pixel, alpha);
    kept += static_cast<int>(__popcll(vote));
  }
  return kept;
}

// Per-block bounds of the depth cutoff (tile size 1).  A camera's 8x8 pixel
// blocks are stored from its first pixel on (index tile_base + block_y *
// blocks_x + block_x; a camera has at least as many pixels as blocks).
__device__ __forceinline__ int dense_blocks_x(int tiles_x) { return (tiles_x + 7) >> 3; }

// False only when every pixel of the rectangle has its cutoff in front of
// `bucket` (the block maxima say so), so a pixel walk would keep none of it.
__device__ __forceinline__ bool dense_rect_reaches(const TileRect& rect, uint32_t tile_base, int tiles_x, int bucket,
                                                   const int32_t* __restrict__ block_cutoff) {
  if (block_cutoff == nullptr) return true;
  const int blocks_x = dense_blocks_x(tiles_x);
  for (int by = rect.min_y >> 3; by <= (rect.max_y - 1) >> 3; ++by) {
    for (int bx = rect.min_x >> 3; bx <= (rect.max_x - 1) >> 3; ++bx) {
      if (block_cutoff[tile_base + static_cast<uint32_t>(by * blocks_x + bx)] >= bucket) return true;
    }
  }
  return false;
}

// evaluate_pair for exact-ray renders: the ray extent replaces the splat's mean
// and radii (binning and culling); `ray` receives the record's precision data.
__device__ __forceinline__ bool evaluate_ray_pair(const ProjectionArgs& args, const CameraParams& cam, int64_t g,
                                                  bool loaded, float m0, float m1, float m2, float opacity,
                                                  const float* cov, Projected& p, TileRect& rect, int& tiles,
                                                  RayData& ray) {
  tiles = 0;
  if (!loaded || cam.scene_slot < 0 || g < cam.gaussian_begin || g >= cam.gaussian_end) return false;
  if (!project_pair<true>(cam, m0, m1, m2, opacity, cov, args.near_plane, args.far_plane, args.support_sigma,
                          args.covariance_epsilon, args.antialiased != 0, p)) {
    return false;
  }
  if (!ray_extent(cam, m0, m1, m2, cov, args.support_sigma, p, ray)) return false;
  rect = tile_rect(p, args.tile_size, cam.tiles_x, cam.tiles_y);
  tiles = rect_tiles(rect);
  return true;
}


__device__ __forceinline__ void load_gaussian(const ProjectionArgs& args, int64_t g, bool& loaded, float& m0,
                                              float& m1, float& m2, float& opacity, float* cov) {
  loaded = g < args.gaussian_count;
  m0 = m1 = m2 = opacity = 0.0f;
  for (int i = 0; i < 6; ++i) cov[i] = 0.0f;
  if (!loaded) return;
  m0 = args.means[g * 3 + 0];
  m1 = args.means[g * 3 + 1];
  m2 = args.means[g * 3 + 2];
  opacity = args.opacities[g];
  for (int i = 0; i < 6; ++i) cov[i] = args.covariances[g * 6 + i];
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t value) {
  for (int offset = 32; offset > 0; offset >>= 1) value += __shfl_xor(value, offset, kWave);
  return value;
}

// Inclusive wave scan of two int32 values.
__device__ __forceinline__ void wave_inclusive_scan2(int& a, int& b) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int offset = 1; offset < kWave; offset <<= 1) {
    const int ta = __shfl_up(a, offset, kWave);
    const int tb = __shfl_up(b, offset, kWave);
    if (lane >= offset) {
      a += ta;
      b += tb;
    }
  }
}

}  // namespace dev

// Fills the projection arguments of a render (also the kept-item list left in
// scratch by chunk culling, when culling is on).  Shared by the forward
// projection and the backward pass.
void make_projection_args(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_options& options,
                          const ScratchLayout& layout, const char* scratch, dev::ProjectionArgs& args);

// Chunk culling of the (camera group x chunk) candidates: leaves the kept
// blocks, their camera masks and their count in scratch (the projection args
// given must have no item list yet).
int launch_cull(const dev::ProjectionArgs& args, const vgr_scene& scene, const ScratchLayout& layout, char* scratch,
                int64_t* statistics, hipStream_t stream);

}  // namespace vgr

#endif  // VGR_PROJECT_DEVICE_H
