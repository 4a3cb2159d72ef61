// Backward pass of the batched camera path for gfx950: compositing,
// projection/EWA, SH colour, covariance and camera-pose gradients.
//
// The reference renderer is forward-only (KERNEL_ARCHITECTURE.md:454-460);
// this pass differentiates the forward implemented in vgr_project.hip and
// vgr_composite.hip, term by term:
//
//   compositing  alpha_i = min(0.99, o_i exp(power_i)),
//                power_i = -0.5 (c0 dx^2 + 2 c1 dx dy + c2 dy^2), dx = px - mx,
//                C = sum_i T_i alpha_i c_i,  D = sum_i T_i alpha_i z_i,
//                A = 1 - T_final,  depth = D / A.
//   projection   mean2d = f * xyz_cam.xy / z + c, conic = (J W Sigma W^T J^T + eps I)^-1,
//                optional antialias compensation, record opacity = clamp(o comp, 0, 1).
//   colour       rgb = max(0, f_0 + sum_k Y_k(dir) f_k), dir = mean - camera centre.
//
// Compositing backward (one workgroup per (camera, tile), one lane per pixel)
// walks the tile list back to front from the checkpointed final
// transmittance: T_i = T_{i+1} / (1 - alpha_i) is recovered by division and
// the colour / depth / alpha composited "behind" each splat is rebuilt on the
// way (as one scalar: its dot product with the pixel's upstream gradient), so
// no per-intersection state is stored.  The same per-wave ordered sub-lists as
// the forward are rebuilt from the same records, and every skip test is the
// forward's own, so exactly the forward's blended (pixel, splat) pairs
// contribute.  Per-splat gradients are summed over the wave's 64 pixels with
// DPP row operations and stored, with plain stores, in the row of the
// (record, tile) pair's intersection slot: a training forward sorts slots
// instead of record indices (slot_records maps them back), and every slot of
// every tile list receives exactly one row, so no float atomics and no zero
// fill are needed.
//
// Projection backward mirrors projection_write_kernel: the same (camera group
// x Gaussian chunk) grid, the same visibility decisions and block scan, so the
// record of every visible pair and its consecutive intersection slots are found
// without any index buffer; the pair sums its slots' rows.  Each
// thread owns one Gaussian and accumulates its gradients over the group's
// cameras in registers (one atomic per value per group); camera-pose partials
// are reduced over the workgroup and added per camera.
#include <type_traits>

#include "vgr_composite_device.h"
#include "vgr_project_device.h"

namespace vgr {
namespace {

using namespace dev;

constexpr int kRecordGrads = 10;  // mean_x, mean_y, conic_0..2, opacity, depth, r, g, b
constexpr int kCameraGrads = 24;  // point outer product 9, translation 3, rotation 9, centre 3

// ---- wave64 sum with DPP row operations -------------------------------------------
// quad xor 1, quad xor 2, half-row mirror, row mirror give every lane of a row
// the row sum; row_bcast:15 and row_bcast:31 then carry the rows into lane 63.
// The wave total is valid in lane 63 only.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_term(float value) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, value), kCtrl, kRowMask, 0xF, false));
}

__device__ __forceinline__ float wave_sum_to_last_lane(float v) {
  v += dpp_term<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_term<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_term<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_term<0x140, 0xF>(v);  // row_mirror
  v += dpp_term<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_term<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Sum of a value and its copy from the neighbouring row (16 lanes) or half (32
// lanes).  The swap instructions exchange rows between two registers.  Empty asm
// statements keep the optimizer from relating the two operands or the two
// results: it otherwise rewrites the second result as the first (treating the
// exchange as a plain register swap), which drops the neighbour's rows.
__device__ __forceinline__ float permlane_pair_sum(float v, bool across_32) {
  unsigned first = __builtin_bit_cast(unsigned, v);
  unsigned second = first;
  asm volatile("" : "+v"(second));
  unsigned out0, out1;
  if (across_32) {
    const auto r = __builtin_amdgcn_permlane32_swap(first, second, false, false);
    out0 = r[0];
    out1 = r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(first, second, false, false);
    out0 = r[0];
    out1 = r[1];
  }
  asm volatile("" : "+v"(out0), "+v"(out1));
  return __builtin_bit_cast(float, out0) + __builtin_bit_cast(float, out1);
}

// Sums ten per-lane values over the wave, leaving value i's total in the lanes
// whose in-row position L (lane % 16) maps to i, in every row.  Inside a row,
// three exchange steps each halve the values a lane carries (it keeps one half
// and sends the other to a partner that holds the same value indices): xor 1
// and xor 2 (quad permutes), then rotations by 4 and by 8 lanes, which pair
// the four quads so that every quad's partial reaches every kept value once.
// The rows are then summed with the permlane swaps (rows 0+1 and 2+3, then the
// two halves).  37 instructions against 60 for ten separate 6-step sums.
// L = b + 2c + 4d + 8e (bits) holds value 5b + 3c + 2d + e when d == 0 or
// (c == 0 and e == 0); its *index is -1 otherwise.
__device__ __forceinline__ float wave_transpose_sum10(const float (&v)[10], int lane, int* index) {
  const int l = lane & 15;
  const bool b = (l & 1) != 0, c = (l & 2) != 0, d = (l & 4) != 0, e = (l & 8) != 0;
  float w[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) w[j] = (b ? v[j + 5] : v[j]) + dpp_term<0xB1, 0xF>(b ? v[j] : v[j + 5]);
  float x[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float high = j < 2 ? w[3 + j] : 0.0f;
    x[j] = (c ? high : w[j]) + dpp_term<0x4E, 0xF>(c ? w[j] : high);
  }
  float y[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const float high = j == 0 ? x[2] : 0.0f;
    y[j] = (d ? high : x[j]) + dpp_term<0x124, 0xF>(d ? x[j] : high);  // row_ror:4
  }
  float z = (e ? y[1] : y[0]) + dpp_term<0x128, 0xF>(e ? y[0] : y[1]);  // row_ror:8
  z = permlane_pair_sum(z, false);
  z = permlane_pair_sum(z, true);
  *index = (!d || (!c && !e)) ? 5 * static_cast<int>(b) + 3 * static_cast<int>(c) + 2 * static_cast<int>(d) +
                                    static_cast<int>(e)
                              : -1;
  return z;
}

__device__ __forceinline__ void atomic_add_f32(float* address, float value) {
  unsafeAtomicAdd(address, value);
}

// ---- projection backward ---------------------------------------------------------

// dL/d f (accumulated) and dL/d direction (returned) of the SH colour.
template <int kDegree>
__device__ __forceinline__ void sh_colour_backward(const float (&f)[3 * (kDegree + 1) * (kDegree + 1)], float dx,
                                                   float dy, float dz, const float* grad_rgb,
                                                   float (&grad_f)[3 * (kDegree + 1) * (kDegree + 1)],
                                                   float* grad_dir) {
  constexpr int kBands = (kDegree + 1) * (kDegree + 1);
  grad_dir[0] = grad_dir[1] = grad_dir[2] = 0.0f;
  if (kDegree <= 0) {
    // Degree 0: rgb = f[0..2] with no clamp.
    grad_f[0] += grad_rgb[0];
    grad_f[1] += grad_rgb[1];
    grad_f[2] += grad_rgb[2];
    return;
  }
  const float norm2 = dx * dx + dy * dy + dz * dz;
  float x = 0.0f, y = 0.0f, z = 0.0f, inverse_norm = 0.0f;
  if (norm2 > 0.0f) {
    inverse_norm = 1.0f / sqrtf(norm2);
    x = dx * inverse_norm;
    y = dy * inverse_norm;
    z = dz * inverse_norm;
  }
  float basis[16];
  basis[0] = 1.0f;
  basis[1] = -kShC1 * y;
  basis[2] = kShC1 * z;
  basis[3] = -kShC1 * x;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  if (kDegree >= 2) {
    basis[4] = kShC2_0 * xy;
    basis[5] = kShC2_1 * yz;
    basis[6] = kShC2_2 * (2.0f * zz - xx - yy);
    basis[7] = kShC2_3 * xz;
    basis[8] = kShC2_4 * (xx - yy);
  }
  if (kDegree >= 3) {
    basis[9] = kShC3_0 * y * (3.0f * xx - yy);
    basis[10] = kShC3_1 * xy * z;
    basis[11] = kShC3_2 * y * (4.0f * zz - xx - yy);
    basis[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    basis[13] = kShC3_4 * x * (4.0f * zz - xx - yy);
    basis[14] = kShC3_5 * z * (xx - yy);
    basis[15] = kShC3_6 * x * (xx - 3.0f * yy);
  }
  // Clamp mask of max(0, .): recompute the forward sum.
  float g[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 1; k < kBands; ++k) acc = fmaf(basis[k], f[3 * k + c], acc);
    g[c] = f[c] + acc > 0.0f ? grad_rgb[c] : 0.0f;
  }
  // dL/d basis_k = sum_c f[3k+c] g_c; dL/d f[3k+c] = basis_k g_c.
  float gb[16];
#pragma unroll
  for (int k = 0; k < kBands; ++k) {
    gb[k] = f[3 * k + 0] * g[0] + f[3 * k + 1] * g[1] + f[3 * k + 2] * g[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) grad_f[3 * k + c] += basis[k] * g[c];
  }
  if (norm2 <= 0.0f) return;
  float gx = -kShC1 * gb[3];
  float gy = -kShC1 * gb[1];
  float gz = kShC1 * gb[2];
  if (kDegree >= 2) {
    gx += kShC2_0 * y * gb[4] - 2.0f * kShC2_2 * x * gb[6] + kShC2_3 * z * gb[7] + 2.0f * kShC2_4 * x * gb[8];
    gy += kShC2_0 * x * gb[4] + kShC2_1 * z * gb[5] - 2.0f * kShC2_2 * y * gb[6] - 2.0f * kShC2_4 * y * gb[8];
    gz += kShC2_1 * y * gb[5] + 4.0f * kShC2_2 * z * gb[6] + kShC2_3 * x * gb[7];
  }
  if (kDegree >= 3) {
    gx += kShC3_0 * 6.0f * xy * gb[9] + kShC3_1 * yz * gb[10] - kShC3_2 * 2.0f * xy * gb[11] -
          kShC3_3 * 6.0f * xz * gb[12] + kShC3_4 * (4.0f * zz - 3.0f * xx - yy) * gb[13] +
          kShC3_5 * 2.0f * xz * gb[14] + kShC3_6 * 3.0f * (xx - yy) * gb[15];
    gy += kShC3_0 * 3.0f * (xx - yy) * gb[9] + kShC3_1 * xz * gb[10] +
          kShC3_2 * (4.0f * zz - xx - 3.0f * yy) * gb[11] - kShC3_3 * 6.0f * yz * gb[12] -
          kShC3_4 * 2.0f * xy * gb[13] - kShC3_5 * 2.0f * yz * gb[14] - kShC3_6 * 6.0f * xy * gb[15];
    gz += kShC3_1 * xy * gb[10] + kShC3_2 * 8.0f * yz * gb[11] +
          kShC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy) * gb[12] + kShC3_4 * 8.0f * xz * gb[13] +
          kShC3_5 * (xx - yy) * gb[14];
  }
  // Through the normalization u = d / |d|: dL/dd = (g - u (u . g)) / |d|.
  const float dot = x * gx + y * gy + z * gz;
  grad_dir[0] = (gx - x * dot) * inverse_norm;
  grad_dir[1] = (gy - y * dot) * inverse_norm;
  grad_dir[2] = (gz - z * dot) * inverse_norm;
}

// Gradient of one visible pair w.r.t. the Gaussian (mean, covariance, opacity)
// and the camera (point outer product, translation, rotation, centre), given
// the record gradient r[10].  Recomputes the forward quantities of project_pair.
struct PairGradients {
  float mean[3];
  float covariance[6];
  float opacity;
  float camera_point[3];  // dL/d xyz_cam
  float rotation[9];      // dL/d W, W = R_view R_env (covariance transform)
};

__device__ __forceinline__ void project_pair_backward(const CameraParams& cam, float m0, float m1, float m2,
                                                      float opacity, const float* cov, float covariance_epsilon,
                                                      bool antialiased, const float* r, PairGradients& out) {
  const float* e = cam.env;
  const float env_x = e[0] * m0 + e[1] * m1 + e[2] * m2 + e[3];
  const float env_y = e[4] * m0 + e[5] * m1 + e[6] * m2 + e[7];
  const float env_z = e[8] * m0 + e[9] * m1 + e[10] * m2 + e[11];
  const float* v = cam.view;
  const float camera_x = v[0] * env_x + v[1] * env_y + v[2] * env_z + v[3];
  const float camera_y = v[4] * env_x + v[5] * env_y + v[6] * env_z + v[7];
  const float camera_z = v[8] * env_x + v[9] * env_y + v[10] * env_z + v[11];
  const float inverse_z = 1.0f / camera_z;
  const float* w = cam.rotation;
  // Sigma_cam = W Sigma W^T.
  float ws[9];  // W Sigma
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    ws[row * 3 + 0] = w[row * 3 + 0] * cov[0] + w[row * 3 + 1] * cov[1] + w[row * 3 + 2] * cov[2];
    ws[row * 3 + 1] = w[row * 3 + 0] * cov[1] + w[row * 3 + 1] * cov[3] + w[row * 3 + 2] * cov[4];
    ws[row * 3 + 2] = w[row * 3 + 0] * cov[2] + w[row * 3 + 1] * cov[4] + w[row * 3 + 2] * cov[5];
  }
  const float c00 = ws[0] * w[0] + ws[1] * w[1] + ws[2] * w[2];
  const float c01 = ws[0] * w[3] + ws[1] * w[4] + ws[2] * w[5];
  const float c02 = ws[0] * w[6] + ws[1] * w[7] + ws[2] * w[8];
  const float c11 = ws[3] * w[3] + ws[4] * w[4] + ws[5] * w[5];
  const float c12 = ws[3] * w[6] + ws[4] * w[7] + ws[5] * w[8];
  const float c22 = ws[6] * w[6] + ws[7] * w[7] + ws[8] * w[8];
  const float tx = camera_x * inverse_z;
  const float ty = camera_y * inverse_z;
  const float clamp_x = fminf(cam.limit_x_pos, fmaxf(-cam.limit_x_neg, tx));
  const float clamp_y = fminf(cam.limit_y_pos, fmaxf(-cam.limit_y_neg, ty));
  const bool inside_x = clamp_x == tx;
  const bool inside_y = clamp_y == ty;
  const float clamped_x = camera_z * clamp_x;
  const float clamped_y = camera_z * clamp_y;
  const float inverse_z_squared = inverse_z * inverse_z;
  const float fx = cam.focal_x, fy = cam.focal_y;
  const float j00 = fx * inverse_z;
  const float j02 = -fx * clamped_x * inverse_z_squared;
  const float j11 = fy * inverse_z;
  const float j12 = -fy * clamped_y * inverse_z_squared;
  const float raw00 = j00 * j00 * c00 + 2.0f * j00 * j02 * c02 + j02 * j02 * c22;
  const float cov01 = j00 * j11 * c01 + j00 * j12 * c02 + j02 * j11 * c12 + j02 * j12 * c22;
  const float raw11 = j11 * j11 * c11 + 2.0f * j11 * j12 * c12 + j12 * j12 * c22;
  const float cov00 = raw00 + covariance_epsilon;
  const float cov11 = raw11 + covariance_epsilon;
  const float determinant = cov00 * cov11 - cov01 * cov01;
  const float inverse_determinant = 1.0f / determinant;

  // Record opacity = clamp(opacity * compensation, 0, 1).
  float compensation = 1.0f, ratio = 0.0f;
  bool compensation_live = false;
  float original = 0.0f;
  if (antialiased) {
    original = raw00 * raw11 - cov01 * cov01;
    ratio = original / determinant;
    const float floor2 = VGR_MIN_ANTIALIAS_COMPENSATION * VGR_MIN_ANTIALIAS_COMPENSATION;
    compensation = sqrtf(fmaxf(floor2, ratio));
    compensation_live = ratio > floor2;
  }
  const float effective = opacity * compensation;
  const float g_effective = (effective >= 0.0f && effective <= 1.0f) ? r[5] : 0.0f;
  out.opacity = g_effective * compensation;

  // conic = (cov')^-1: dL/dcov' = -M G M with G = [[g0, g1/2], [g1/2, g2]].
  const float a = cov11 * inverse_determinant;
  const float b = -cov01 * inverse_determinant;
  const float c = cov00 * inverse_determinant;
  const float g0 = r[2], g1h = 0.5f * r[3], g2 = r[4];
  const float mg00 = a * g0 + b * g1h, mg01 = a * g1h + b * g2;
  const float mg10 = b * g0 + c * g1h, mg11 = b * g1h + c * g2;
  float g_raw00 = -(mg00 * a + mg01 * b);
  float g_cov01 = -2.0f * (mg00 * b + mg01 * c);
  float g_raw11 = -(mg10 * b + mg11 * c);
  if (compensation_live) {
    const float g_ratio = g_effective * opacity * 0.5f / compensation;
    const float inverse_det2 = inverse_determinant * inverse_determinant;
    g_raw00 += g_ratio * (raw11 * determinant - original * cov11) * inverse_det2;
    g_raw11 += g_ratio * (raw00 * determinant - original * cov00) * inverse_det2;
    g_cov01 += g_ratio * 2.0f * cov01 * (original - determinant) * inverse_det2;
  }

  // 2D covariance = J Sigma_cam J^T: gradients of the six Sigma_cam scalars and of J.
  const float gc00 = g_raw00 * j00 * j00;
  const float gc01 = g_cov01 * j00 * j11;
  const float gc02 = g_raw00 * 2.0f * j00 * j02 + g_cov01 * j00 * j12;
  const float gc11 = g_raw11 * j11 * j11;
  const float gc12 = g_cov01 * j02 * j11 + g_raw11 * 2.0f * j11 * j12;
  const float gc22 = g_raw00 * j02 * j02 + g_cov01 * j02 * j12 + g_raw11 * j12 * j12;
  const float g_j00 = g_raw00 * (2.0f * j00 * c00 + 2.0f * j02 * c02) + g_cov01 * (j11 * c01 + j12 * c02);
  const float g_j02 = g_raw00 * (2.0f * j00 * c02 + 2.0f * j02 * c22) + g_cov01 * (j11 * c12 + j12 * c22);
  const float g_j11 = g_raw11 * (2.0f * j11 * c11 + 2.0f * j12 * c12) + g_cov01 * (j00 * c01 + j02 * c12);
  const float g_j12 = g_raw11 * (2.0f * j11 * c12 + 2.0f * j12 * c22) + g_cov01 * (j00 * c02 + j02 * c22);

  // Camera-space point: mean2d, depth and J.
  float g_x = r[0] * fx * inverse_z;
  float g_y = r[1] * fy * inverse_z;
  float g_z = -(r[0] * fx * camera_x + r[1] * fy * camera_y) * inverse_z_squared + r[6];
  g_z -= (g_j00 * fx + g_j11 * fy) * inverse_z_squared;
  const float inverse_z_cubed = inverse_z_squared * inverse_z;
  g_z += 2.0f * (g_j02 * fx * clamped_x + g_j12 * fy * clamped_y) * inverse_z_cubed;
  const float g_clamped_x = -g_j02 * fx * inverse_z_squared;
  const float g_clamped_y = -g_j12 * fy * inverse_z_squared;
  if (inside_x) {
    g_x += g_clamped_x;
  } else {
    g_z += g_clamped_x * clamp_x;
  }
  if (inside_y) {
    g_y += g_clamped_y;
  } else {
    g_z += g_clamped_y * clamp_y;
  }
  out.camera_point[0] = g_x;
  out.camera_point[1] = g_y;
  out.camera_point[2] = g_z;
  // Mean: xyz_cam = V_R (E_R m + E_t) + V_t.
  const float ge0 = v[0] * g_x + v[4] * g_y + v[8] * g_z;
  const float ge1 = v[1] * g_x + v[5] * g_y + v[9] * g_z;
  const float ge2 = v[2] * g_x + v[6] * g_y + v[10] * g_z;
  out.mean[0] = e[0] * ge0 + e[4] * ge1 + e[8] * ge2;
  out.mean[1] = e[1] * ge0 + e[5] * ge1 + e[9] * ge2;
  out.mean[2] = e[2] * ge0 + e[6] * ge1 + e[10] * ge2;

  // Sigma_cam = W Sigma W^T: dL/dSigma = W^T G W, dL/dW = 2 G W Sigma (G full symmetric).
  const float G[9] = {gc00, 0.5f * gc01, 0.5f * gc02, 0.5f * gc01, gc11, 0.5f * gc12,
                      0.5f * gc02, 0.5f * gc12, gc22};
  float gw[9];  // G W
#pragma unroll
  for (int row = 0; row < 3; ++row)
#pragma unroll
    for (int col = 0; col < 3; ++col)
      gw[row * 3 + col] = G[row * 3 + 0] * w[0 * 3 + col] + G[row * 3 + 1] * w[1 * 3 + col] +
                          G[row * 3 + 2] * w[2 * 3 + col];
  // W^T (G W)
  float s[9];
#pragma unroll
  for (int row = 0; row < 3; ++row)
#pragma unroll
    for (int col = 0; col < 3; ++col)
      s[row * 3 + col] = w[0 * 3 + row] * gw[0 * 3 + col] + w[1 * 3 + row] * gw[1 * 3 + col] +
                         w[2 * 3 + row] * gw[2 * 3 + col];
  out.covariance[0] = s[0];
  out.covariance[1] = s[1] + s[3];
  out.covariance[2] = s[2] + s[6];
  out.covariance[3] = s[4];
  out.covariance[4] = s[5] + s[7];
  out.covariance[5] = s[8];
  const float sigma[9] = {cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]};
#pragma unroll
  for (int row = 0; row < 3; ++row)
#pragma unroll
    for (int col = 0; col < 3; ++col)
      out.rotation[row * 3 + col] = 2.0f * (gw[row * 3 + 0] * sigma[0 * 3 + col] + gw[row * 3 + 1] * sigma[1 * 3 + col] +
                                            gw[row * 3 + 2] * sigma[2 * 3 + col]);
}

// Pose partials of one visible pair from its ten record terms r (mean_x, mean_y,
// conic_0..2, opacity, depth, r, g, b), as the projection backward forms them:
// dL/d xyz_cam outer the environment-space point (9), dL/d xyz_cam (3),
// dL/d W (9) and minus the SH direction gradient (the camera centre, 3).
__device__ __forceinline__ void fill_pose_terms(const CameraParams& cam, float m0, float m1, float m2,
                                                const PairGradients& pg, const float* grad_dir,
                                                float (&pose)[kCameraGrads]) {
  const float* e = cam.env;
  const float pe[3] = {e[0] * m0 + e[1] * m1 + e[2] * m2 + e[3], e[4] * m0 + e[5] * m1 + e[6] * m2 + e[7],
                       e[8] * m0 + e[9] * m1 + e[10] * m2 + e[11]};
#pragma unroll
  for (int row = 0; row < 3; ++row) {
#pragma unroll
    for (int col = 0; col < 3; ++col) pose[row * 3 + col] = pg.camera_point[row] * pe[col];
    pose[9 + row] = pg.camera_point[row];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) pose[12 + k] = pg.rotation[k];
  pose[21] = -grad_dir[0];
  pose[22] = -grad_dir[1];
  pose[23] = -grad_dir[2];
}

// Pose partials of (camera, packed Gaussian g) from the pair's record terms.
template <int kDegree>
__device__ __forceinline__ void gaussian_pose_terms(const ProjectionArgs& pa, const CameraParams& cam, int32_t g,
                                                    const float (&r)[kRecordGrads], float (&pose)[kCameraGrads]) {
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(pa, g, loaded, m0, m1, m2, opacity, cov);
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  float f[kChannels];
  load_features<kDegree>(pa.features, g, pa.feature_width, f);
  PairGradients pg;
  project_pair_backward(cam, m0, m1, m2, opacity, cov, pa.covariance_epsilon, pa.antialiased != 0, r, pg);
  float grad_f[kChannels];  // scene gradient: not needed here
#pragma unroll
  for (int k = 0; k < kChannels; ++k) grad_f[k] = 0.0f;
  float grad_dir[3];
  sh_colour_backward<kDegree>(f, m0 - cam.origin_x, m1 - cam.origin_y, m2 - cam.origin_z, r + 7, grad_f, grad_dir);
  fill_pose_terms(cam, m0, m1, m2, pg, grad_dir, pose);
}

// ---- compositing backward --------------------------------------------------------

struct CompositeBackwardArgs {
  const CameraParams* table;
  const int32_t* tile_starts;
  const int32_t* tile_ends;
  const int32_t* sorted_values;
  const vgr_record* records;
  const float* final_transmittance;
  const int32_t* last_blended;
  const float* depth;
  const float* grad_rgb;
  const float* grad_depth;
  const float* grad_alpha;
  // Gradient terms of each (record, tile) pair, stored at the pair's intersection
  // slot (the sorted value): plain stores, one row per slot, no atomics.  The
  // projection backward sums the rows of a record's slots.
  float* partials;               // [intersection_capacity, 10]
  const int32_t* slot_records;   // intersection slot -> visible record
  const int32_t* tile_order;  // optional launch order of the (camera, tile) blocks
  int64_t visible_capacity;
  int64_t intersection_capacity;
  int active_count;
  int uniform_tiles;
};

__device__ __forceinline__ void store_partial(float* __restrict__ partials, int32_t slot, float v0, float v1,
                                              float v2, float v3, float v4, float v5, float v6, float v7, float v8,
                                              float v9) {
  float2* row = reinterpret_cast<float2*>(partials + static_cast<int64_t>(slot) * kRecordGrads);
  row[0] = make_float2(v0, v1);
  row[1] = make_float2(v2, v3);
  row[2] = make_float2(v4, v5);
  row[3] = make_float2(v6, v7);
  row[4] = make_float2(v8, v9);
}

// Zero rows for the list entries [from, to) of a tile that the walk does not
// reach (past every pixel's last blended splat), so every slot of every tile
// list holds its pair's terms after the pass.
__device__ __forceinline__ void zero_partials(const CompositeBackwardArgs& args, int64_t from, int64_t to,
                                              int lane, int threads) {
  for (int64_t i = from + lane; i < to; i += threads) {
    const int32_t slot = args.sorted_values[i];
    if (slot >= 0 && slot < args.intersection_capacity) {
      store_partial(args.partials, slot, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
}

// (active camera, image tile) of a compositing block; false for a block past the
// last tile of its camera (uneven resolutions).
__device__ __forceinline__ bool backward_block_tile(const CompositeBackwardArgs& args, int64_t block, int& active,
                                                    int64_t& image_tile) {
  if (args.uniform_tiles > 0) {
    active = static_cast<int>(block / args.uniform_tiles);
    if (active >= args.active_count) return false;
    image_tile = block - static_cast<int64_t>(active) * args.uniform_tiles;
    return true;
  }
  active = find_active_camera(args.table, args.active_count, block);
  image_tile = block - args.table[active].active_tile_begin;
  return image_tile < static_cast<int64_t>(args.table[active].tiles_x) * args.table[active].tiles_y;
}

// Intersection slots of one batch (the sorted values), -1 outside the capacity.
template <int kThreads, int kBatch, int kPerThread>
__device__ __forceinline__ void gather_batch_slots(const int32_t* __restrict__ sorted_values,
                                                   int64_t intersection_capacity, int64_t batch_start,
                                                   int64_t range_end, int lane, int32_t (&ids)[kPerThread]) {
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t index = batch_start + k * kThreads + lane;
    ids[k] = -1;
    if (k * kThreads + lane < kBatch && index < range_end) {
      const int32_t slot = sorted_values[index];
      ids[k] = (slot >= 0 && slot < intersection_capacity) ? slot : -1;
    }
  }
}

template <int kTileSize, int kBatch>
__global__ __launch_bounds__(kTileSize* kTileSize) void composite_backward_kernel(CompositeBackwardArgs args) {
  using Lists = WaveLists<kTileSize, kBatch>;
  constexpr int kThreads = Lists::kThreads;
  constexpr int kWaves = Lists::kWaves;
  constexpr int kPerThread = Lists::kPerThread;
  __shared__ Lists s;
  __shared__ int32_t s_record[kBatch];  // intersection slot of each batch entry
  __shared__ __align__(16) float s_partial[kBatch][kRecordGrads];
  __shared__ uint32_t s_blended[kBatch];
  __shared__ int s_last[kWaves];

  const int64_t block = blockIdx.x;
  int active;
  int64_t image_tile;
  if (args.uniform_tiles > 0) {
    active = static_cast<int>(block / args.uniform_tiles);
    if (active >= args.active_count) return;
    image_tile = block - static_cast<int64_t>(active) * args.uniform_tiles;
  } else {
    active = find_active_camera(args.table, args.active_count, block);
    image_tile = block - args.table[active].active_tile_begin;
    if (image_tile >= static_cast<int64_t>(args.table[active].tiles_x) * args.table[active].tiles_y) return;
  }
  const CameraParams& cam = args.table[active];
  if (cam.camera < 0) return;
  const int tile_x = static_cast<int>(image_tile % cam.tiles_x);
  const int tile_y = static_cast<int>(image_tile / cam.tiles_x);
  const uint32_t global_tile = cam.tile_base + static_cast<uint32_t>(image_tile);
  const int lane = threadIdx.x;
  const int wave = lane / kWave;
  const int wave_lane = lane & (kWave - 1);
  const int pixel_x = tile_x * kTileSize + Lists::pixel_dx(lane);
  const int pixel_y = tile_y * kTileSize + Lists::pixel_dy(lane);
  const bool valid_pixel = pixel_x < cam.width && pixel_y < cam.height;
  const float px = static_cast<float>(pixel_x) + 0.5f;
  const float py = static_cast<float>(pixel_y) + 0.5f;
  const float rect_x0 = static_cast<float>(tile_x * kTileSize) + 0.5f;
  const float rect_x1 = rect_x0 + static_cast<float>(kTileSize - 1);
  const float rect_y_base = static_cast<float>(tile_y * kTileSize) + 0.5f;

  // Per-pixel upstream gradients; depth = D / A is folded into dL/dD and dL/dA.
  int last = 0;
  float transmittance = 1.0f;
  float grad_r = 0.0f, grad_g = 0.0f, grad_b = 0.0f, grad_d = 0.0f, grad_a = 0.0f;
  if (valid_pixel) {
    const int64_t pixel = cam.pixel_offset + static_cast<int64_t>(pixel_y) * cam.width + pixel_x;
    last = args.last_blended[pixel];
    transmittance = args.final_transmittance[pixel];
    grad_r = args.grad_rgb[pixel * 3 + 0];
    grad_g = args.grad_rgb[pixel * 3 + 1];
    grad_b = args.grad_rgb[pixel * 3 + 2];
    grad_a = args.grad_alpha != nullptr ? args.grad_alpha[pixel] : 0.0f;
    const float accumulated = 1.0f - transmittance;
    if (args.grad_depth != nullptr && accumulated > 1.0e-8f) {
      const float upstream = args.grad_depth[pixel];
      if (upstream != 0.0f) {
        grad_d = upstream / accumulated;
        grad_a -= upstream * args.depth[pixel] / accumulated;
      }
    }
  }
  int wave_last = last;
  for (int offset = 32; offset > 0; offset >>= 1) wave_last = max(wave_last, __shfl_xor(wave_last, offset, kWave));
  if (wave_lane == 0) s_last[wave] = wave_last;
  __syncthreads();
  int tile_last = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) tile_last = max(tile_last, s_last[w]);
  const int64_t range_start = args.tile_starts[global_tile];
  const int64_t range_end_full = args.tile_ends[global_tile];
  if (range_start < 0 || range_end_full <= range_start) return;
  const int64_t range_end = min(range_end_full, range_start + static_cast<int64_t>(tile_last > 0 ? tile_last : 0));
  zero_partials(args, range_end, range_end_full, lane, kThreads);
  if (tile_last <= 0) return;

  // behind = (colour, depth, alpha) composited behind the current splat, dotted with
  // the pixel's upstream gradient (one scalar: the recursion is linear).
  float behind = 0.0f;
  const int64_t batches = (range_end - range_start + kBatch - 1) / kBatch;
  f32x4 next_a[kPerThread] = {}, next_b[kPerThread] = {}, next_c[kPerThread] = {};
  int32_t next_ids[kPerThread];
  {
    const int64_t start = range_start + (batches - 1) * kBatch;
    gather_batch<kThreads, kBatch, kPerThread, true>(args.sorted_values, args.records, args.visible_capacity, start,
                                                     range_end, lane, next_a, next_b, next_c, args.slot_records);
    gather_batch_slots<kThreads, kBatch, kPerThread>(args.sorted_values, args.intersection_capacity, start,
                                                     range_end, lane, next_ids);
  }
  for (int64_t b = batches - 1; b >= 0; --b) {
    const int64_t batch_start = range_start + b * kBatch;
    const int count = static_cast<int>(range_end - batch_start < kBatch ? range_end - batch_start : kBatch);
    __syncthreads();  // previous batch fully consumed
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const int slot = k * kThreads + lane;
      if (slot < count) {
        s_record[slot] = next_ids[k];
        s_blended[slot] = 0u;
      }
    }
    build_wave_lists<kTileSize, kBatch>(s, next_a, next_b, next_c, count, lane, rect_x0, rect_x1, rect_y_base);
    if (b > 0) {
      const int64_t start = batch_start - kBatch;
      gather_batch<kThreads, kBatch, kPerThread, true>(args.sorted_values, args.records, args.visible_capacity,
                                                       start, range_end, lane, next_a, next_b, next_c,
                                                       args.slot_records);
      gather_batch_slots<kThreads, kBatch, kPerThread>(args.sorted_values, args.intersection_capacity, start,
                                                       range_end, lane, next_ids);
    }
    const int position_base = static_cast<int>(batch_start - range_start);
    const uint32_t list_count = s.count(wave);
    for (uint32_t i = list_count; i-- > 0;) {
      const int item = s.list[wave][i];
      const f32x4 g = s.geometry[item];
      const f32x4 sh = s.shading[item];
      const float delta_x = px - g.x;
      const float delta_y = py - g.y;
      // Forward arithmetic and skip tests, unchanged.
      const float power = g.z * delta_x * delta_x + g.w * delta_x * delta_y + sh.x * delta_y * delta_y;
      bool blended = position_base + item < last && !(power > 0.0f || power < sh.z);
      float gaussian = 0.0f, splat_alpha = 0.0f;
      if (blended) {
        gaussian = vgr_expf(power);
        splat_alpha = fminf(VGR_MAX_ALPHA, sh.y * gaussian);
        blended = splat_alpha >= VGR_ALPHA_THRESHOLD;
      }
      float d_mx = 0.0f, d_my = 0.0f, d_c0 = 0.0f, d_c1 = 0.0f, d_c2 = 0.0f, d_o = 0.0f;
      float d_z = 0.0f, d_r = 0.0f, d_g = 0.0f, d_b = 0.0f;
      if (blended) {
        // Gradient arithmetic only (the replayed forward decisions above stay uncontracted).
#pragma clang fp contract(fast)
        const f32x4 col = s.colour[item];
        const float keep = 1.0f - splat_alpha;
        // keep >= 0.01; the hardware reciprocal (1 ulp) is ample for a gradient.
        const float before = transmittance * __builtin_amdgcn_rcpf(keep);
        const float weight = splat_alpha * before;
        d_r = weight * grad_r;
        d_g = weight * grad_g;
        d_b = weight * grad_b;
        d_z = weight * grad_d;
        const float own = col.x * grad_r + col.y * grad_g + col.z * grad_b + sh.w * grad_d + grad_a;
        const float d_alpha = (own - behind) * before;
        behind = splat_alpha * own + keep * behind;
        transmittance = before;
        if (sh.y * gaussian < VGR_MAX_ALPHA) {
          d_o = d_alpha * gaussian;
          const float d_power = d_alpha * splat_alpha;
          // power = A dx^2 + B dx dy + C dy^2 with A = -c0/2, B = -c1, C = -c2/2.
          d_mx = -d_power * (2.0f * g.z * delta_x + g.w * delta_y);
          d_my = -d_power * (g.w * delta_x + 2.0f * sh.x * delta_y);
          d_c0 = -0.5f * d_power * delta_x * delta_x;
          d_c1 = -d_power * delta_x * delta_y;
          d_c2 = -0.5f * d_power * delta_y * delta_y;
        }
      }
      if (__ballot(blended) == 0ull) continue;
      d_mx = wave_sum_to_last_lane(d_mx);
      d_my = wave_sum_to_last_lane(d_my);
      d_c0 = wave_sum_to_last_lane(d_c0);
      d_c1 = wave_sum_to_last_lane(d_c1);
      d_c2 = wave_sum_to_last_lane(d_c2);
      d_o = wave_sum_to_last_lane(d_o);
      d_z = wave_sum_to_last_lane(d_z);
      d_r = wave_sum_to_last_lane(d_r);
      d_g = wave_sum_to_last_lane(d_g);
      d_b = wave_sum_to_last_lane(d_b);
      if (wave_lane == kWave - 1) {
        // One wave per 8x8 tile: the wave sums are the tile's terms of this splat.
        float* row = s_partial[item];
        row[0] = d_mx;
        row[1] = d_my;
        row[2] = d_c0;
        row[3] = d_c1;
        row[4] = d_c2;
        row[5] = d_o;
        row[6] = d_z;
        row[7] = d_r;
        row[8] = d_g;
        row[9] = d_b;
        s_blended[item] = 1u;
      }
    }
    __syncthreads();
    for (int slot = lane; slot < count; slot += kThreads) {
      const int32_t id = s_record[slot];
      if (id < 0) continue;
      if (s_blended[slot] != 0u) {
        const float* row = s_partial[slot];
        store_partial(args.partials, id, row[0], row[1], row[2], row[3], row[4], row[5], row[6], row[7], row[8],
                      row[9]);
      } else {
        store_partial(args.partials, id, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
  }
}

// Tile-front lists (VGR_ORDERING_TILE_FRONT) hold packed Gaussian indices: the
// records are recomputed from the scene for the tile's camera (as the forward
// compositor did), and the pass computes camera-pose gradients only.  Each
// (Gaussian, tile) pair's ten terms go straight through the pair's projection
// backward (linear in them, so summing over tiles afterwards equals the
// per-pair sum of the global-ordering pass up to rounding); the 24 pose
// partials are summed per lane in LDS and added once per tile.
struct FrontBackwardArgs : CompositeBackwardArgs {
  ProjectionArgs projection;
  float* camera_grads;      // [VGR_POSE_REPLICAS * active_count, 24]
  const int64_t* counters;  // the forward's counter block
};

// Device-side guard: a forward whose intersections overflowed the buffer left
// incomplete lists, so every pose term becomes NaN instead of silently wrong.
__device__ __forceinline__ bool front_backward_invalid(const FrontBackwardArgs& args) {
  return args.counters[VGR_COUNTER_INTERSECTIONS] > args.intersection_capacity ||
         args.counters[VGR_COUNTER_SORT_FAILURES] != 0;
}

template <int kDegree>
using BackwardArgsFor = typename std::conditional<(kDegree < 0), CompositeBackwardArgs, FrontBackwardArgs>::type;

// Wave-per-tile form for 16x16 tiles: one wave64 owns the whole tile and every
// lane four pixels (column lane % 16, rows lane / 16 + 4k).  Pixel k of every
// lane lies in the tile's k-th band of four rows, so the per-band reach mask of
// a splat (the same conservative ellipse box as the forward's per-wave lists)
// is a uniform branch.  Each lane sums its four pixels' gradient terms of a
// splat; a transposing butterfly (DPP and permlane swaps) sums them over the
// wave into ten lanes, which park them in LDS, and after the batch every lane
// stores the row of one batch entry's intersection slot (zeros for entries no
// pixel blended).  The float atomics this replaced took 3 ms of a 7.9 ms
// training step on 16 cameras of 512x512 (the step ran in 4.9 ms without them).
// Five waves per SIMD: at most 102 VGPRs (the 6.6 KB of LDS per wave would allow 24 per CU);
// the batch is gathered at its start rather than prefetched, to stay within them.
template <int kBatch, int kDegree = -1>
__global__ __launch_bounds__(kWave, (kDegree < 0 ? 5 : (kDegree >= 2 ? 2 : 3))) void composite_backward_wave_kernel(
    BackwardArgsFor<kDegree> args) {
  constexpr bool kFront = kDegree >= 0;
  constexpr int kTileSize = 16;
  constexpr int kBands = 4;
  constexpr int kPerLane = kBatch / kWave;
  static_assert(kBatch % kWave == 0 && kBatch <= 512, "batch must be a multiple of 64 and at most 512");
  __shared__ f32x4 s_geometry[kBatch];
  __shared__ f32x4 s_shading[kBatch];
  __shared__ f32x4 s_colour[kBatch];
  __shared__ int32_t s_record[kBatch];
  __shared__ uint16_t s_list[kBatch];
  // Per batch slot: the wave sums of its ten gradient terms (m_x, m_y, m_xx, m_xy,
  // m_yy, opacity, depth, r, g, b; padded to 12 floats for 16-byte reads) and
  // whether any pixel blended it.
  __shared__ __align__(16) float s_partial[kBatch][12];
  __shared__ uint32_t s_blended[kBatch];
  // Tile-front: per-lane sums of the 24 pose partials (column = lane).
  __shared__ float s_pose[kFront ? kCameraGrads : 1][kWave + 1];  // padded: row reads hit distinct banks

  const int64_t block = args.tile_order != nullptr ? args.tile_order[blockIdx.x] : blockIdx.x;
  int active;
  int64_t image_tile;
  if (!backward_block_tile(args, block, active, image_tile)) return;
  const CameraParams& cam = args.table[active];
  if (cam.camera < 0) return;
  const int tile_x = static_cast<int>(image_tile % cam.tiles_x);
  const int tile_y = static_cast<int>(image_tile / cam.tiles_x);
  const uint32_t global_tile = cam.tile_base + static_cast<uint32_t>(image_tile);
  const int lane = threadIdx.x;
  const int pixel_x = tile_x * kTileSize + (lane % kTileSize);
  const float px = static_cast<float>(pixel_x) + 0.5f;
  const float rect_x0 = static_cast<float>(tile_x * kTileSize) + 0.5f;
  const float rect_x1 = rect_x0 + static_cast<float>(kTileSize - 1);
  const float rect_y_base = static_cast<float>(tile_y * kTileSize) + 0.5f;

  float py[kBands], transmittance[kBands], grad_r[kBands], grad_g[kBands], grad_b[kBands], grad_d[kBands],
      grad_a[kBands], behind[kBands];
  int last[kBands];
  int lane_last = 0;
#pragma unroll
  for (int k = 0; k < kBands; ++k) {
    const int pixel_y = tile_y * kTileSize + lane / kTileSize + 4 * k;
    py[k] = static_cast<float>(pixel_y) + 0.5f;
    last[k] = 0;
    transmittance[k] = 1.0f;
    grad_r[k] = grad_g[k] = grad_b[k] = grad_d[k] = grad_a[k] = 0.0f;
    behind[k] = 0.0f;  // what lies behind, dotted with the upstream gradient
    if (pixel_x < cam.width && pixel_y < cam.height) {
      const int64_t pixel = cam.pixel_offset + static_cast<int64_t>(pixel_y) * cam.width + pixel_x;
      last[k] = args.last_blended[pixel];
      transmittance[k] = args.final_transmittance[pixel];
      grad_r[k] = args.grad_rgb[pixel * 3 + 0];
      grad_g[k] = args.grad_rgb[pixel * 3 + 1];
      grad_b[k] = args.grad_rgb[pixel * 3 + 2];
      grad_a[k] = args.grad_alpha != nullptr ? args.grad_alpha[pixel] : 0.0f;
      const float accumulated = 1.0f - transmittance[k];
      if (args.grad_depth != nullptr && accumulated > 1.0e-8f) {
        const float upstream = args.grad_depth[pixel];
        if (upstream != 0.0f) {
          grad_d[k] = upstream / accumulated;
          grad_a[k] -= upstream * args.depth[pixel] / accumulated;
        }
      }
      lane_last = max(lane_last, last[k]);
    }
  }
  int tile_last = lane_last;
  for (int offset = 32; offset > 0; offset >>= 1) tile_last = max(tile_last, __shfl_xor(tile_last, offset, kWave));
  // Per band, the largest last-blended position of its 64 pixels: list entries at or
  // beyond it are skipped for the whole band.
  tile_last = __builtin_amdgcn_readfirstlane(tile_last);
  int band_last[kBands];
#pragma unroll
  for (int k = 0; k < kBands; ++k) {
    int value = last[k];
    for (int offset = 32; offset > 0; offset >>= 1) value = max(value, __shfl_xor(value, offset, kWave));
    band_last[k] = __builtin_amdgcn_readfirstlane(value);
  }
  const int64_t range_start = args.tile_starts[global_tile];
  const int64_t range_end_full = args.tile_ends[global_tile];
  if (range_start < 0 || range_end_full <= range_start) return;
  const int64_t range_end = min(range_end_full, range_start + static_cast<int64_t>(tile_last > 0 ? tile_last : 0));
  if constexpr (!kFront) zero_partials(args, range_end, range_end_full, lane, kWave);
  if (tile_last <= 0) return;
  if constexpr (kFront) {
#pragma unroll
    for (int k = 0; k < kCameraGrads; ++k) s_pose[k][lane] = 0.0f;
  }
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));

  const int64_t batches = (range_end - range_start + kBatch - 1) / kBatch;
  for (int64_t b = batches - 1; b >= 0; --b) {
    const int64_t batch_start = range_start + b * kBatch;
    const int count = static_cast<int>(range_end - batch_start < kBatch ? range_end - batch_start : kBatch);
    f32x4 next_a[kPerLane] = {}, next_b[kPerLane] = {}, next_c[kPerLane] = {};
    int32_t next_ids[kPerLane];
    if constexpr (kFront) {
      // Records recomputed from the listed Gaussians; the entry's id is the Gaussian.
#pragma unroll
      for (int k = 0; k < kPerLane; ++k) {
        const int64_t index = batch_start + k * kWave + lane;
        next_ids[k] = -1;
        if (index < range_end) {
          const int32_t g = args.sorted_values[index];
          if (g >= 0 && g < args.projection.gaussian_count) {
            record_from_gaussian<kDegree>(args.projection, cam, g, next_a[k], next_b[k], next_c[k]);
            next_ids[k] = g;
          } else {
            next_a[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            next_b[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            next_c[k] = f32x4{0.0f, 0.0f, __int_as_float(-1), 0.0f};
          }
        }
      }
    } else {
      gather_batch<kWave, kBatch, kPerLane, true>(args.sorted_values, args.records, args.visible_capacity,
                                                  batch_start, range_end, lane, next_a, next_b, next_c,
                                                  args.slot_records);
      gather_batch_slots<kWave, kBatch, kPerLane>(args.sorted_values, args.intersection_capacity, batch_start,
                                                  range_end, lane, next_ids);
    }
    __syncthreads();  // previous batch fully consumed
    uint32_t list_count = 0;
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int slot = k * kWave + lane;
      uint32_t bands = 0u;
      if (slot < count) {
        const StagedSplat staged = stage_record(next_a[k], next_b[k], next_c[k]);
        s_geometry[slot] = staged.geometry;
        s_shading[slot] = staged.shading;
        s_colour[slot] = staged.colour;
        s_record[slot] = next_ids[k];
        s_blended[slot] = 0u;
        float half_x, half_y;
        if (splat_extent(next_a[k], next_b[k], staged.shading.z, half_x, half_y)) {
          const float mx = next_a[k].x, my = next_a[k].y;
          if (fabsf(mx - fminf(fmaxf(mx, rect_x0), rect_x1)) <= half_x) {
#pragma unroll
            for (int band = 0; band < kBands; ++band) {
              const float y0 = rect_y_base + static_cast<float>(4 * band);
              const float y1 = y0 + 3.0f;
              if (fabsf(my - fminf(fmaxf(my, y0), y1)) <= half_y) bands |= 1u << band;
            }
          }
        } else {
          bands = (1u << kBands) - 1u;
        }
        // Drop bands whose pixels all stopped blending before this list position.
        const int position = static_cast<int>(batch_start - range_start) + slot;
#pragma unroll
        for (int band = 0; band < kBands; ++band) {
          if (position >= band_last[band]) bands &= ~(1u << band);
        }
      }
      const uint64_t vote = __ballot(bands != 0u);
      if (bands != 0u) {
        s_list[list_count + static_cast<uint32_t>(__popcll(vote & below))] =
            static_cast<uint16_t>(slot | (bands << 12));
      }
      list_count += static_cast<uint32_t>(__popcll(vote));
    }
    __syncthreads();
    const int position_base = static_cast<int>(batch_start - range_start);
    for (uint32_t i = list_count; i-- > 0;) {
      const uint32_t entry = s_list[i];
      const int item = static_cast<int>(entry & 0x1ffu);
      const uint32_t bands = entry >> 12;
      const f32x4 g = s_geometry[item];
      const f32x4 sh = s_shading[item];
      const f32x4 col = s_colour[item];
      const int position = position_base + item;
      const float delta_x = px - g.x;
      // Per-lane moments of q = dL/d alpha * G over the lane's pixels; the mean and
      // conic gradients are linear in them and are formed once per splat below.
      float m_x = 0.0f, m_y = 0.0f, m_xx = 0.0f, m_xy = 0.0f, m_yy = 0.0f, d_o = 0.0f;
      float d_z = 0.0f, d_r = 0.0f, d_g = 0.0f, d_b = 0.0f;
      bool any = false;
#pragma unroll
      for (int k = 0; k < kBands; ++k) {
        // Uniform skips: the band is out of the splat's reach, or every pixel of the
        // band stopped blending before this list position.
        if (((bands >> k) & 1u) == 0u || position >= band_last[k]) continue;
        const float delta_y = py[k] - g.y;
        const float power = g.z * delta_x * delta_x + g.w * delta_x * delta_y + sh.x * delta_y * delta_y;
        const bool reachable = position < last[k] && !(power > 0.0f || power < sh.z);
        if (__ballot(reachable) == 0ull) continue;
        // From here the band runs without divergent branches: lanes that do not
        // blend the splat take alpha = 0 (keep = 1, nothing changes) and zero
        // gradients, so every lane executes the same straight-line code.
        const float exact = vgr_expf_nonpositive(power);  // unused where power > 0 (not reachable)
        const float splat_alpha = fminf(VGR_MAX_ALPHA, sh.y * exact);
        const bool blended = reachable && splat_alpha >= VGR_ALPHA_THRESHOLD;
        if (__ballot(blended) == 0ull) continue;
        any = true;
        {
        // Gradient arithmetic only (the replayed forward decisions above stay uncontracted).
#pragma clang fp contract(fast)
        const float alpha = blended ? splat_alpha : 0.0f;
        const float gaussian = blended && sh.y * exact < VGR_MAX_ALPHA ? exact : 0.0f;  // 0 where alpha is capped
        const float keep = 1.0f - alpha;
        const float before = transmittance[k] * __builtin_amdgcn_rcpf(keep);
        const float weight = alpha * before;
        d_r += weight * grad_r[k];
        d_g += weight * grad_g[k];
        d_b += weight * grad_b[k];
        d_z += weight * grad_d[k];
        const float own = col.x * grad_r[k] + col.y * grad_g[k] + col.z * grad_b[k] + sh.w * grad_d[k] + grad_a[k];
        const float d_alpha = (own - behind[k]) * before;
        behind[k] = alpha * own + keep * behind[k];
        transmittance[k] = before;
        const float q = d_alpha * gaussian;  // dL/d power = q * opacity (= d_alpha * alpha below the cap)
        d_o += q;
        const float q_x = q * delta_x;
        const float q_y = q * delta_y;
        m_x += q_x;
        m_y += q_y;
        m_xx += q_x * delta_x;
        m_xy += q_x * delta_y;
        m_yy += q_y * delta_y;
        }
      }
      if (__ballot(any) == 0ull) continue;
      // The wave sums land in ten lanes of row 0, which park them in LDS; after the
      // batch every lane stores one entry's row.
      const float terms[10] = {m_x, m_y, m_xx, m_xy, m_yy, d_o, d_z, d_r, d_g, d_b};
      int index;
      const float total = wave_transpose_sum10(terms, lane, &index);
      if (lane < 16 && index >= 0) s_partial[item][index] = total;
      if (lane == 0) s_blended[item] = 1u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int slot = k * kWave + lane;
      if (slot >= count) continue;
      const int32_t id = s_record[slot];  // intersection slot (tile-front: Gaussian) of the entry
      if (id < 0) continue;
      if (s_blended[slot] == 0u) {
        if constexpr (!kFront) {
          store_partial(args.partials, id, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f);
        }
        continue;
      }
      const f32x4* partial = reinterpret_cast<const f32x4*>(s_partial[slot]);
      const f32x4 p0 = partial[0];  // m_x, m_y, m_xx, m_xy
      const f32x4 p1 = partial[1];  // m_yy, opacity, depth, r
      const f32x4 p2 = partial[2];  // g, b
      const f32x4 g = s_geometry[slot];
      const f32x4 sh = s_shading[slot];
      float d_mx, d_my, d_c0, d_c1, d_c2;
      {
#pragma clang fp contract(fast)
        // power = g.z dx^2 + g.w dx dy + sh.x dy^2 with dx = px - mean_x, dy = py - mean_y;
        // dL/d power = q * opacity, and the moments of q were summed above.
        const float o = sh.y;
        d_mx = -o * (2.0f * g.z * p0.x + g.w * p0.y);
        d_my = -o * (g.w * p0.x + 2.0f * sh.x * p0.y);
        d_c0 = -0.5f * o * p0.z;
        d_c1 = -o * p0.w;
        d_c2 = -0.5f * o * p1.x;
      }
      if constexpr (kFront) {
        float r[kRecordGrads] = {d_mx, d_my, d_c0, d_c1, d_c2, p1.y, p1.z, p1.w, p2.x, p2.y};
        if (front_backward_invalid(args)) {
#pragma unroll
          for (int t = 0; t < kRecordGrads; ++t) r[t] = __int_as_float(0x7fc00000);
        }
        float pose[kCameraGrads];
        gaussian_pose_terms<kDegree>(args.projection, cam, id, r, pose);
#pragma unroll
        for (int t = 0; t < kCameraGrads; ++t) s_pose[t][lane] += pose[t];
      } else {
        store_partial(args.partials, id, d_mx, d_my, d_c0, d_c1, d_c2, p1.y, p1.z, p1.w, p2.x, p2.y);
      }
    }
  }
  if constexpr (kFront) {
    // Lane t < 24 sums pose partial t over the 64 lanes' columns and adds it to
    // the camera's row in replica (tile % VGR_POSE_REPLICAS).
    __syncthreads();
    if (lane < kCameraGrads) {
      float sum = 0.0f;
      for (int column = 0; column < kWave; ++column) sum += s_pose[lane][column];
      const int64_t replica = global_tile % VGR_POSE_REPLICAS;
      if (sum != 0.0f) atomic_add_f32(args.camera_grads + (replica * args.active_count + active) * kCameraGrads + lane,
                                      sum);
    }
  }
}

struct ProjectionBackwardArgs {
  ProjectionArgs forward;
  const int64_t* block_bases;
  const float* partials;           // [intersection_capacity, 10] per-slot terms of the compositing backward
  const int64_t* counters;         // the forward's counter block
  int64_t visible_capacity;
  int64_t intersection_capacity;
  float* grad_means;
  float* grad_covariances;
  float* grad_opacities;
  float* grad_features;
  float* camera_grads;
  int exclusive;  // one camera group: every Gaussian's gradients are owned by one thread
};

// Adds a per-Gaussian gradient: plain read-modify-write when one thread owns the
// Gaussian (a single camera group), a float atomic otherwise.
__device__ __forceinline__ void add_gradient(float* address, float value, bool exclusive) {
  if (exclusive) {
    *address += value;
  } else {
    atomic_add_f32(address, value);
  }
}

template <int kDegree, bool kScene, bool kPose>
__global__ __launch_bounds__(kBlockThreads) void projection_backward_kernel(ProjectionBackwardArgs args) {
  const ProjectionArgs& fw = args.forward;
  int64_t block;
  uint64_t mask;
  if (!resolve_block(fw, block, mask)) return;
  const int group = static_cast<int>(block / fw.chunks);
  const int64_t chunk = block - static_cast<int64_t>(group) * fw.chunks;
  const int64_t g = chunk * kBlockThreads + threadIdx.x;
  bool loaded;
  float m0, m1, m2, opacity, cov[6];
  load_gaussian(fw, g, loaded, m0, m1, m2, opacity, cov);
  constexpr int kChannels = 3 * (kDegree + 1) * (kDegree + 1);
  float f[kChannels];
  float grad_f[kChannels];
#pragma unroll
  for (int i = 0; i < kChannels; ++i) {
    f[i] = 0.0f;
    grad_f[i] = 0.0f;
  }
  if (loaded && opacity >= VGR_ALPHA_THRESHOLD) load_features<kDegree>(fw.features, g, fw.feature_width, f);
  float grad_mean[3] = {0.0f, 0.0f, 0.0f};
  float grad_cov[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float grad_opacity = 0.0f;
  bool touched = false;
  const int64_t wave_slot = static_cast<int64_t>(blockIdx.x) * kProjectionWaves + threadIdx.x / kWave;
  int64_t visible_base = args.block_bases[wave_slot * 2 + 0];
  int64_t intersection_base = args.block_bases[wave_slot * 2 + 1];
  // Device-side guard: when the forward overflowed a buffer (its lists are then
  // incomplete) or its sort gave up a look-back wait (its order is then wrong),
  // every pair's terms are NaN, so the gradients come out NaN, never silently wrong.
  const bool invalid = args.counters[VGR_COUNTER_VISIBLE] > args.visible_capacity ||
                       args.counters[VGR_COUNTER_INTERSECTIONS] > args.intersection_capacity ||
                       args.counters[VGR_COUNTER_SORT_FAILURES] != 0;
  const int first = group * fw.cameras_per_block;
  const int last = min(fw.active_count, first + fw.cameras_per_block);
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  // Records are found with the forward's per-wave bases and wave scans (no block barrier).
  __shared__ float s_pose[kBlockThreads / kWave][kCameraGrads];
  for (int a = first; a < last; ++a) {
    if (((mask >> (a - first)) & 1ull) == 0) continue;
    const CameraParams& cam = fw.table[a];
    Projected p;
    TileRect rect;
    int tiles;
    const bool visible = evaluate_pair(fw, cam, g, loaded, m0, m1, m2, opacity, cov, p, rect, tiles);
    int inclusive_visible = visible ? 1 : 0;
    int inclusive_tiles = tiles;
    wave_inclusive_scan2(inclusive_visible, inclusive_tiles);
    const int total_visible = __shfl(inclusive_visible, kWave - 1, kWave);
    const int total_tiles = __shfl(inclusive_tiles, kWave - 1, kWave);
    float pose[kCameraGrads];
    if (kPose) {
#pragma unroll
      for (int k = 0; k < kCameraGrads; ++k) pose[k] = 0.0f;
    }
    if (visible) {
      const int64_t record = visible_base + inclusive_visible - 1;
      if (record < args.visible_capacity) {
        // The pair's intersections occupy consecutive slots (the forward wrote them
        // at the same scanned positions); their rows sum to the record's terms.
        float r[kRecordGrads];
#pragma unroll
        for (int k = 0; k < kRecordGrads; ++k) r[k] = 0.0f;
        const int64_t first_slot = intersection_base + inclusive_tiles - tiles;
        for (int t = 0; t < tiles; ++t) {
          const int64_t slot = first_slot + t;
          if (slot >= args.intersection_capacity) break;
          const float2* row = reinterpret_cast<const float2*>(args.partials + slot * kRecordGrads);
#pragma unroll
          for (int k = 0; k < kRecordGrads / 2; ++k) {
            const float2 v = row[k];
            r[2 * k] += v.x;
            r[2 * k + 1] += v.y;
          }
        }
        if (invalid) {
#pragma unroll
          for (int k = 0; k < kRecordGrads; ++k) r[k] = __int_as_float(0x7fc00000);
        }
        PairGradients pg;
        project_pair_backward(cam, m0, m1, m2, opacity, cov, fw.covariance_epsilon, fw.antialiased != 0, r, pg);
        float grad_dir[3];
        sh_colour_backward<kDegree>(f, m0 - cam.origin_x, m1 - cam.origin_y, m2 - cam.origin_z, r + 7, grad_f,
                                    grad_dir);
        touched = true;
        if (kScene) {
          grad_mean[0] += pg.mean[0] + grad_dir[0];
          grad_mean[1] += pg.mean[1] + grad_dir[1];
          grad_mean[2] += pg.mean[2] + grad_dir[2];
#pragma unroll
          for (int k = 0; k < 6; ++k) grad_cov[k] += pg.covariance[k];
          grad_opacity += pg.opacity;
        }
        if (kPose) fill_pose_terms(cam, m0, m1, m2, pg, grad_dir, pose);
      }
    }
    // Block-uniform test (the barrier below needs every wave).
    if (kPose && __syncthreads_or(total_visible > 0)) {
#pragma unroll
      for (int k = 0; k < kCameraGrads; ++k) {
        const float sum = wave_sum_to_last_lane(pose[k]);
        if (lane == kWave - 1) s_pose[wave][k] = sum;
      }
      __syncthreads();
      if (threadIdx.x < kCameraGrads) {
        float sum = 0.0f;
#pragma unroll
        for (int w = 0; w < kBlockThreads / kWave; ++w) sum += s_pose[w][threadIdx.x];
        // One of VGR_POSE_REPLICAS copies of the camera's row, picked by workgroup:
        // the workgroups of a camera spread their adds over that many addresses.
        const int64_t replica = blockIdx.x % VGR_POSE_REPLICAS;
        if (sum != 0.0f) {
          atomic_add_f32(args.camera_grads + (replica * fw.active_count + a) * kCameraGrads + threadIdx.x, sum);
        }
      }
    }
    visible_base += total_visible;
    intersection_base += total_tiles;
  }
  if (!kScene || !touched) return;
  const bool exclusive = args.exclusive != 0;
  add_gradient(args.grad_means + g * 3 + 0, grad_mean[0], exclusive);
  add_gradient(args.grad_means + g * 3 + 1, grad_mean[1], exclusive);
  add_gradient(args.grad_means + g * 3 + 2, grad_mean[2], exclusive);
#pragma unroll
  for (int k = 0; k < 6; ++k) add_gradient(args.grad_covariances + g * 6 + k, grad_cov[k], exclusive);
  add_gradient(args.grad_opacities + g, grad_opacity, exclusive);
  float* row = args.grad_features + g * fw.feature_width;
#pragma unroll
  for (int k = 0; k < kChannels; ++k) add_gradient(row + k, grad_f[k], exclusive);
}

// dL/d viewmat of every active camera from the accumulated partials:
//   xyz_cam = V_R p + V_t, p = E_R m + E_t          -> dV_R += sum g p^T, dV_t += sum g
//   W = V_R E_R                                      -> dV_R += (dL/dW) E_R^T
//   centre o = -W^-1 (V_R E_t + V_t) (SH direction)  -> u = -W^-T dL/do,
//                                                      dV_R += u o^T E_R^T + u E_t^T, dV_t += u
__global__ void camera_grads_finalize_kernel(const CameraParams* __restrict__ table, int active_count,
                                             const float* __restrict__ camera_grads,
                                             float* __restrict__ grad_viewmats) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= active_count) return;
  const CameraParams& cam = table[a];
  if (cam.camera < 0) return;
  // Sum of the replicas of the camera's row (fixed order).
  double acc[kCameraGrads];
  for (int k = 0; k < kCameraGrads; ++k) acc[k] = 0.0;
  for (int r = 0; r < VGR_POSE_REPLICAS; ++r) {
    const float* row = camera_grads + (static_cast<int64_t>(r) * active_count + a) * kCameraGrads;
    for (int k = 0; k < kCameraGrads; ++k) acc[k] += row[k];
  }
  const float* e = cam.env;
  const float* v = cam.view;
  double m[9];
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col)
      m[row * 3 + col] = static_cast<double>(v[row * 4 + 0]) * e[0 * 4 + col] +
                         static_cast<double>(v[row * 4 + 1]) * e[1 * 4 + col] +
                         static_cast<double>(v[row * 4 + 2]) * e[2 * 4 + col];
  // Adjugate inverse of W (double).
  const double a00 = m[4] * m[8] - m[5] * m[7];
  const double a01 = m[2] * m[7] - m[1] * m[8];
  const double a02 = m[1] * m[5] - m[2] * m[4];
  const double a10 = m[5] * m[6] - m[3] * m[8];
  const double a11 = m[0] * m[8] - m[2] * m[6];
  const double a12 = m[2] * m[3] - m[0] * m[5];
  const double a20 = m[3] * m[7] - m[4] * m[6];
  const double a21 = m[1] * m[6] - m[0] * m[7];
  const double a22 = m[0] * m[4] - m[1] * m[3];
  const double det = m[0] * a00 + m[1] * a10 + m[2] * a20;
  double u[3] = {0.0, 0.0, 0.0};
  if (det != 0.0) {
    // W^-1 = adj / det; u = -W^-T go.
    const double inv[9] = {a00 / det, a01 / det, a02 / det, a10 / det, a11 / det,
                           a12 / det, a20 / det, a21 / det, a22 / det};
    const double go[3] = {acc[21], acc[22], acc[23]};
    for (int i = 0; i < 3; ++i) u[i] = -(inv[0 * 3 + i] * go[0] + inv[1 * 3 + i] * go[1] + inv[2 * 3 + i] * go[2]);
  }
  const double origin[3] = {cam.origin_x, cam.origin_y, cam.origin_z};
  float* out = grad_viewmats + static_cast<int64_t>(cam.camera) * 16;
  for (int row = 0; row < 3; ++row) {
    for (int col = 0; col < 3; ++col) {
      // (dL/dW) E_R^T and (u o^T) E_R^T: sum_k X[row][k] E_R[col][k].
      double value = acc[row * 3 + col];
      for (int k = 0; k < 3; ++k) {
        value += (static_cast<double>(acc[12 + row * 3 + k]) + u[row] * origin[k]) * e[col * 4 + k];
      }
      value += u[row] * e[col * 4 + 3];
      atomic_add_f32(out + row * 4 + col, static_cast<float>(value));
    }
    atomic_add_f32(out + row * 4 + 3, static_cast<float>(acc[9 + row] + u[row]));
  }
}

// Chain rule of precompute_covariances_kernel (Sigma = R(q) diag(s^2) R(q)^T).
__global__ void covariance_backward_kernel(const float* __restrict__ scales, const float* __restrict__ rotations,
                                           const float* __restrict__ grad_covariances, int64_t count,
                                           float* __restrict__ grad_scales, float* __restrict__ grad_rotations) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= count) return;
  const float qw = rotations[g * 4 + 0], qx = rotations[g * 4 + 1];
  const float qy = rotations[g * 4 + 2], qz = rotations[g * 4 + 3];
  const float norm2 = qw * qw + qx * qx + qy * qy + qz * qz;
  const float inverse_norm = 1.0f / sqrtf(fmaxf(norm2, 1.0e-20f));
  const float w = qw * inverse_norm, x = qx * inverse_norm, y = qy * inverse_norm, z = qz * inverse_norm;
  float R[9];
  R[0] = 1.0f - 2.0f * (y * y + z * z);
  R[1] = 2.0f * (x * y - w * z);
  R[2] = 2.0f * (x * z + w * y);
  R[3] = 2.0f * (x * y + w * z);
  R[4] = 1.0f - 2.0f * (x * x + z * z);
  R[5] = 2.0f * (y * z - w * x);
  R[6] = 2.0f * (x * z - w * y);
  R[7] = 2.0f * (y * z + w * x);
  R[8] = 1.0f - 2.0f * (x * x + y * y);
  const float* gc = grad_covariances + g * 6;
  // Full symmetric gradient of Sigma.
  const float G[9] = {gc[0], 0.5f * gc[1], 0.5f * gc[2], 0.5f * gc[1], gc[3], 0.5f * gc[4],
                      0.5f * gc[2], 0.5f * gc[4], gc[5]};
  const float s[3] = {scales[g * 3 + 0], scales[g * 3 + 1], scales[g * 3 + 2]};
  // GR = G R; dL/dR = 2 G R diag(s^2); dL/ds_i = 2 s_i (R^T G R)_ii.
  float GR[9];
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col)
      GR[row * 3 + col] = G[row * 3 + 0] * R[0 * 3 + col] + G[row * 3 + 1] * R[1 * 3 + col] +
                          G[row * 3 + 2] * R[2 * 3 + col];
  for (int i = 0; i < 3; ++i) {
    const float rgr = R[0 * 3 + i] * GR[0 * 3 + i] + R[1 * 3 + i] * GR[1 * 3 + i] + R[2 * 3 + i] * GR[2 * 3 + i];
    grad_scales[g * 3 + i] = 2.0f * s[i] * rgr;
  }
  float dR[9];
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) dR[row * 3 + col] = 2.0f * GR[row * 3 + col] * s[col] * s[col];
  // R(w, x, y, z) of the normalized quaternion.
  const float dw = 2.0f * (-z * dR[1] + y * dR[2] + z * dR[3] - x * dR[5] - y * dR[6] + x * dR[7]);
  const float dx = 2.0f * (y * dR[1] + z * dR[2] + y * dR[3] - 2.0f * x * dR[4] - w * dR[5] + z * dR[6] +
                           w * dR[7] - 2.0f * x * dR[8]);
  const float dy = 2.0f * (-2.0f * y * dR[0] + x * dR[1] + w * dR[2] + x * dR[3] + z * dR[5] - w * dR[6] +
                           z * dR[7] - 2.0f * y * dR[8]);
  const float dz = 2.0f * (-2.0f * z * dR[0] - w * dR[1] + x * dR[2] + w * dR[3] - 2.0f * z * dR[4] +
                           y * dR[5] + x * dR[6] + y * dR[7]);
  // Normalization q / |q|: dL/dq = (dn - n (n . dn)) / |q|.
  const float dot = w * dw + x * dx + y * dy + z * dz;
  const float scale = norm2 > 1.0e-20f ? inverse_norm : 0.0f;
  grad_rotations[g * 4 + 0] = (dw - w * dot) * scale;
  grad_rotations[g * 4 + 1] = (dx - x * dot) * scale;
  grad_rotations[g * 4 + 2] = (dy - y * dot) * scale;
  grad_rotations[g * 4 + 3] = (dz - z * dot) * scale;
}

template <int kTileSize, int kBatch>
void launch_composite_backward(const CompositeBackwardArgs& args, int64_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL((composite_backward_kernel<kTileSize, kBatch>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kTileSize * kTileSize), 0, stream, args);
}

template <int kBatch>
void launch_composite_backward_wave(const CompositeBackwardArgs& args, int64_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL((composite_backward_wave_kernel<kBatch>), dim3(static_cast<unsigned>(blocks)), dim3(kWave), 0,
                     stream, args);
}

template <int kTileSize>
void launch_composite_backward_batch(int batch, const CompositeBackwardArgs& args, int64_t blocks,
                                     hipStream_t stream) {
  switch (batch) {
    case 64:
      launch_composite_backward<kTileSize, 64>(args, blocks, stream);
      break;
    case 128:
      launch_composite_backward<kTileSize, 128>(args, blocks, stream);
      break;
    case 512:
      launch_composite_backward<kTileSize, 512>(args, blocks, stream);
      break;
    default:
      launch_composite_backward<kTileSize, 256>(args, blocks, stream);
      break;
  }
}

template <int kDegree>
void launch_projection_backward_degree(bool scene, bool pose, const ProjectionBackwardArgs& args, int64_t blocks,
                                       hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(blocks)), threads(kBlockThreads);
  if (scene && pose) {
    hipLaunchKernelGGL((projection_backward_kernel<kDegree, true, true>), grid, threads, 0, stream, args);
  } else if (scene) {
    hipLaunchKernelGGL((projection_backward_kernel<kDegree, true, false>), grid, threads, 0, stream, args);
  } else if (pose) {
    hipLaunchKernelGGL((projection_backward_kernel<kDegree, false, true>), grid, threads, 0, stream, args);
  }
}

}  // namespace

int launch_backward(const vgr_scene& scene, const vgr_cameras& cameras, const vgr_outputs& outputs,
                    const vgr_options& options, const vgr_workspace& ws, const ScratchLayout& layout,
                    const char* scratch, const vgr_gradients& grads, hipStream_t stream) {
  if (cameras.active_count <= 0) return 0;
  const CameraParams* table = reinterpret_cast<const CameraParams*>(scratch + layout.camera_table);
  const bool pose = grads.grad_viewmats != nullptr;
  if (pose) {
    launch_fill_f32(grads.camera_grads,
                    static_cast<int64_t>(VGR_POSE_REPLICAS) * cameras.active_count * kCameraGrads, 0.0f, stream);
  }

  CompositeBackwardArgs cargs;
  cargs.table = table;
  cargs.tile_starts = ws.tile_starts;
  cargs.tile_ends = ws.tile_ends;
  cargs.sorted_values = grads.sorted_selector != 0 ? ws.values_out : ws.values_in;
  cargs.records = ws.records;
  cargs.final_transmittance = outputs.final_transmittance;
  cargs.last_blended = outputs.last_blended;
  cargs.depth = outputs.depth;
  cargs.grad_rgb = grads.grad_rgb;
  cargs.grad_depth = grads.grad_depth;
  cargs.grad_alpha = grads.grad_alpha;
  cargs.partials = grads.record_grads;
  cargs.slot_records = ws.slot_records;
  cargs.tile_order = nullptr;
  cargs.visible_capacity = ws.visible_capacity;
  cargs.intersection_capacity = ws.intersection_capacity;
  cargs.active_count = cameras.active_count;
  int64_t tile_blocks;
  if (cameras.widths == nullptr) {
    const int tiles_x = (options.width + options.tile_size - 1) / options.tile_size;
    const int tiles_y = (options.height + options.tile_size - 1) / options.tile_size;
    cargs.uniform_tiles = tiles_x * tiles_y;
    tile_blocks = static_cast<int64_t>(cameras.active_count) * cargs.uniform_tiles;
  } else {
    cargs.uniform_tiles = 0;
    tile_blocks = cameras.max_global_tiles;
  }
  if (tile_blocks > 0x7fffffffLL) {
    set_error("backward composite grid of %lld tiles exceeds the launch limit", static_cast<long long>(tile_blocks));
    return VGR_E_INVALID_ARGUMENT;
  }
  if (options.ordering == VGR_ORDERING_TILE_FRONT) {
    // Pose gradients only, fused into the compositing backward (no record rows,
    // no projection pass).  Grids the forward ordered longest list first reuse
    // its order; larger ones run in natural order.
    FrontBackwardArgs fargs;
    static_cast<CompositeBackwardArgs&>(fargs) = cargs;
    make_projection_args(scene, cameras, options, layout, scratch, fargs.projection);
    fargs.camera_grads = grads.camera_grads;
    fargs.counters = ws.counters;
    fargs.tile_order = tile_blocks <= kOrderedCompositeBlocks
                           ? reinterpret_cast<const int32_t*>(scratch + layout.tile_order)
                           : nullptr;
    if (tile_blocks > 0) {
      const dim3 grid(static_cast<unsigned>(tile_blocks)), threads(kWave);
      switch (scene.sh_degree) {
        case 1:
          hipLaunchKernelGGL((composite_backward_wave_kernel<64, 1>), grid, threads, 0, stream, fargs);
          break;
        case 2:
          hipLaunchKernelGGL((composite_backward_wave_kernel<64, 2>), grid, threads, 0, stream, fargs);
          break;
        case 3:
          hipLaunchKernelGGL((composite_backward_wave_kernel<64, 3>), grid, threads, 0, stream, fargs);
          break;
        default:
          hipLaunchKernelGGL((composite_backward_wave_kernel<64, 0>), grid, threads, 0, stream, fargs);
          break;
      }
    }
    const int finalize_blocks = (cameras.active_count + 63) / 64;
    hipLaunchKernelGGL(camera_grads_finalize_kernel, dim3(finalize_blocks), dim3(64), 0, stream, table,
                       cameras.active_count, grads.camera_grads, grads.grad_viewmats);
    const hipError_t status = hipGetLastError();
    if (status != hipSuccess) {
      set_error("backward launch failed: %s", hipGetErrorString(status));
      return VGR_E_LAUNCH;
    }
    return 0;
  }
  if (tile_blocks > 0) {
    if (options.tile_size == 8) {
      launch_composite_backward_batch<8>(options.raster_batch, cargs, tile_blocks, stream);
    } else {
      // The backward's batch is independent of the forward's (decisions are per pair);
      // 64 keeps the prefetch registers and LDS small enough for five waves per SIMD.
      // One wave walks a whole tile list, so a long tile that starts late runs
      // alone at the end of the grid: tiles are launched longest first.
      // The checkpointed forward ordered grids of up to kOrderedCompositeBlocks tiles
      // into this scratch from the same lists; its order is reused.
      int32_t* order = reinterpret_cast<int32_t*>(const_cast<char*>(scratch) + layout.tile_order);
      if (tile_blocks > kOrderedCompositeBlocks) {
        const int status = launch_tile_order(table, ws.tile_starts, ws.tile_ends, cameras.active_count,
                                             cargs.uniform_tiles, tile_blocks, order, stream);
        if (status != 0) return status;
      }
      cargs.tile_order = order;
      launch_composite_backward_wave<64>(cargs, tile_blocks, stream);
    }
  }

  ProjectionBackwardArgs pargs;
  make_projection_args(scene, cameras, options, layout, scratch, pargs.forward);
  pargs.block_bases = reinterpret_cast<const int64_t*>(scratch + layout.block_bases);
  pargs.partials = grads.record_grads;
  pargs.counters = ws.counters;
  pargs.visible_capacity = ws.visible_capacity;
  pargs.intersection_capacity = ws.intersection_capacity;
  pargs.grad_means = grads.grad_means;
  pargs.grad_covariances = grads.grad_covariances;
  pargs.grad_opacities = grads.grad_opacities;
  pargs.grad_features = grads.grad_features;
  pargs.camera_grads = grads.camera_grads;
  pargs.exclusive = layout.projection_groups == 1 ? 1 : 0;
  const bool scene_grads = grads.grad_means != nullptr;
  const int64_t blocks = layout.projection_blocks;
  if (blocks > 0 && (scene_grads || pose)) {
    switch (scene.sh_degree) {
      case 1:
        launch_projection_backward_degree<1>(scene_grads, pose, pargs, blocks, stream);
        break;
      case 2:
        launch_projection_backward_degree<2>(scene_grads, pose, pargs, blocks, stream);
        break;
      case 3:
        launch_projection_backward_degree<3>(scene_grads, pose, pargs, blocks, stream);
        break;
      default:
        launch_projection_backward_degree<0>(scene_grads, pose, pargs, blocks, stream);
        break;
    }
  }
  if (pose) {
    const int finalize_blocks = (cameras.active_count + 63) / 64;
    hipLaunchKernelGGL(camera_grads_finalize_kernel, dim3(finalize_blocks), dim3(64), 0, stream, table,
                       cameras.active_count, grads.camera_grads, grads.grad_viewmats);
  }
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    set_error("backward launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_covariance_backward(const float* scales, const float* rotations, const float* grad_covariances,
                               int64_t count, float* grad_scales, float* grad_rotations, hipStream_t stream) {
  if (count <= 0) return 0;
  const int64_t blocks = (count + kBlockThreads - 1) / kBlockThreads;
  hipLaunchKernelGGL(covariance_backward_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlockThreads), 0, stream,
                     scales, rotations, grad_covariances, count, grad_scales, grad_rotations);
  const hipError_t status = hipGetLastError();
  if (status != hipSuccess) {
    set_error("covariance backward launch failed: %s", hipGetErrorString(status));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
