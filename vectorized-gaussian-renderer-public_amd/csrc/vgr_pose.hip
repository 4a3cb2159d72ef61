// Camera-pose composition for training: views[b] = exp(xi_b^) @ base[b], with
// xi = (rho, phi) an se(3) twist, and its analytic backward.  One thread per
// camera, all arithmetic in double (a batch holds at most a few thousand
// cameras, so the cost is one short launch either way; double keeps the
// derivative coefficients free of cancellation at small angles).
//
// exp(xi^) = [[R, J rho], [0, 1]] with K = [phi]_x, theta = |phi| and
//   R = I + a K + b K^2,  J = I + b K + c K^2,
//   a = sin(t)/t,  b = (1 - cos t)/t^2,  c = (t - sin t)/t^3.
// For the backward, dK/dphi_i = E_i (the generator of axis i) and
//   d a/d phi_i = a1 phi_i with a1 = (da/dt)/t (likewise b1, c1), so
//   dR/dphi_i = a1 phi_i K + a E_i + b1 phi_i K^2 + b (E_i K + K E_i)
//   dJ/dphi_i = b1 phi_i K + b E_i + c1 phi_i K^2 + c (E_i K + K E_i).
#include <hip/hip_runtime.h>

#include "vgr_internal.h"

namespace vgr {
namespace {

struct PoseCoefficients {
  double a, b, c, a1, b1, c1;
};

// Below theta = 0.05 the Taylor series in t = theta^2 (truncated after t^3;
// the first omitted term is below 1e-12) replace the closed forms, whose
// numerators cancel.
__device__ PoseCoefficients pose_coefficients(double theta2) {
  PoseCoefficients k;
  if (theta2 < 2.5e-3) {
    const double t = theta2;
    k.a = 1.0 - t / 6.0 + t * t / 120.0 - t * t * t / 5040.0;
    k.b = 0.5 - t / 24.0 + t * t / 720.0 - t * t * t / 40320.0;
    k.c = 1.0 / 6.0 - t / 120.0 + t * t / 5040.0 - t * t * t / 362880.0;
    k.a1 = -1.0 / 3.0 + t / 30.0 - t * t / 840.0 + t * t * t / 45360.0;
    k.b1 = -1.0 / 12.0 + t / 180.0 - t * t / 6720.0 + t * t * t / 453600.0;
    k.c1 = -1.0 / 60.0 + t / 1260.0 - t * t / 60480.0 + t * t * t / 4989600.0;
  } else {
    const double theta = sqrt(theta2);
    const double s = sin(theta), co = cos(theta);
    const double t3 = theta2 * theta, t4 = theta2 * theta2, t5 = t4 * theta;
    k.a = s / theta;
    k.b = (1.0 - co) / theta2;
    k.c = (theta - s) / t3;
    k.a1 = (theta * co - s) / t3;
    k.b1 = (theta * s - 2.0 * (1.0 - co)) / t4;
    k.c1 = (1.0 - co) / t4 - 3.0 * (theta - s) / t5;
  }
  return k;
}

__device__ void skew(const double* phi, double* k) {
  k[0] = 0.0;
  k[1] = -phi[2];
  k[2] = phi[1];
  k[3] = phi[2];
  k[4] = 0.0;
  k[5] = -phi[0];
  k[6] = -phi[1];
  k[7] = phi[0];
  k[8] = 0.0;
}

__device__ void mul3(const double* x, const double* y, double* out) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) out[r * 3 + c] = x[r * 3] * y[c] + x[r * 3 + 1] * y[3 + c] + x[r * 3 + 2] * y[6 + c];
  }
}

// Top 3x4 block of exp(xi^): [R | J rho].
__device__ void pose_transform(const float* xi, double* top, double* rotation, double* jacobian, double* kmat,
                               double* kk, PoseCoefficients& coeff) {
  const double rho[3] = {xi[0], xi[1], xi[2]};
  const double phi[3] = {xi[3], xi[4], xi[5]};
  coeff = pose_coefficients(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
  skew(phi, kmat);
  mul3(kmat, kmat, kk);
  for (int i = 0; i < 9; ++i) {
    const double eye = (i % 4 == 0) ? 1.0 : 0.0;
    rotation[i] = eye + coeff.a * kmat[i] + coeff.b * kk[i];
    jacobian[i] = eye + coeff.b * kmat[i] + coeff.c * kk[i];
  }
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) top[r * 4 + c] = rotation[r * 3 + c];
    top[r * 4 + 3] = jacobian[r * 3] * rho[0] + jacobian[r * 3 + 1] * rho[1] + jacobian[r * 3 + 2] * rho[2];
  }
}

__global__ void pose_compose_kernel(const float* __restrict__ twists, const float* __restrict__ base,
                                    int64_t count, float* __restrict__ views) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= count) return;
  double top[12], rotation[9], jacobian[9], kmat[9], kk[9];
  PoseCoefficients coeff;
  pose_transform(twists + b * 6, top, rotation, jacobian, kmat, kk, coeff);
  const float* v = base + b * 16;
  float* out = views + b * 16;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 4; ++c) {
      const double value = top[r * 4] * v[c] + top[r * 4 + 1] * v[4 + c] + top[r * 4 + 2] * v[8 + c] +
                           top[r * 4 + 3] * v[12 + c];
      out[r * 4 + c] = static_cast<float>(value);
    }
  }
  for (int c = 0; c < 4; ++c) out[12 + c] = v[12 + c];
}

__global__ void pose_compose_backward_kernel(const float* __restrict__ twists, const float* __restrict__ base,
                                             const float* __restrict__ grad_views, int64_t count,
                                             float* __restrict__ grad_twists, float* __restrict__ grad_base) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= count) return;
  const float* xi = twists + b * 6;
  double top[12], rotation[9], jacobian[9], kmat[9], kk[9];
  PoseCoefficients coeff;
  pose_transform(xi, top, rotation, jacobian, kmat, kk, coeff);
  const float* v = base + b * 16;
  const float* g = grad_views + b * 16;
  // dL/dT = G V^T (rows 0..2 matter: the bottom row of T is constant).
  double grad_t[12];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 4; ++c) {
      grad_t[r * 4 + c] = static_cast<double>(g[r * 4]) * v[c * 4] + static_cast<double>(g[r * 4 + 1]) * v[c * 4 + 1] +
                          static_cast<double>(g[r * 4 + 2]) * v[c * 4 + 2] +
                          static_cast<double>(g[r * 4 + 3]) * v[c * 4 + 3];
    }
  }
  const double rho[3] = {xi[0], xi[1], xi[2]};
  const double phi[3] = {xi[3], xi[4], xi[5]};
  const double gt[3] = {grad_t[3], grad_t[7], grad_t[11]};
  double out[6];
  // dL/drho = J^T gt.
  for (int i = 0; i < 3; ++i) out[i] = jacobian[i] * gt[0] + jacobian[3 + i] * gt[1] + jacobian[6 + i] * gt[2];
  for (int i = 0; i < 3; ++i) {
    double e[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double unit[3] = {0.0, 0.0, 0.0};
    unit[i] = 1.0;
    skew(unit, e);
    double ek[9], ke[9];
    mul3(e, kmat, ek);
    mul3(kmat, e, ke);
    double total = 0.0;
    for (int r = 0; r < 3; ++r) {
      double dj_rho = 0.0;
      for (int c = 0; c < 3; ++c) {
        const int m = r * 3 + c;
        const double sym = ek[m] + ke[m];
        const double d_rotation = coeff.a1 * phi[i] * kmat[m] + coeff.a * e[m] + coeff.b1 * phi[i] * kk[m] +
                                  coeff.b * sym;
        const double d_jacobian = coeff.b1 * phi[i] * kmat[m] + coeff.b * e[m] + coeff.c1 * phi[i] * kk[m] +
                                  coeff.c * sym;
        total += grad_t[r * 4 + c] * d_rotation;
        dj_rho += d_jacobian * rho[c];
      }
      total += gt[r] * dj_rho;
    }
    out[3 + i] = total;
  }
  for (int i = 0; i < 6; ++i) grad_twists[b * 6 + i] = static_cast<float>(out[i]);
  if (grad_base != nullptr) {
    // dL/dV = T^T G with T's bottom row (0, 0, 0, 1).
    for (int r = 0; r < 4; ++r) {
      for (int c = 0; c < 4; ++c) {
        double value = r == 3 ? static_cast<double>(g[12 + c]) : 0.0;
        for (int k = 0; k < 3; ++k) value += top[k * 4 + r] * g[k * 4 + c];
        grad_base[b * 16 + r * 4 + c] = static_cast<float>(value);
      }
    }
  }
}

}  // namespace

int launch_pose_compose(const float* twists, const float* base, int64_t count, float* views, hipStream_t stream) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(pose_compose_kernel, dim3(static_cast<unsigned>((count + 63) / 64)), dim3(64), 0, stream,
                     twists, base, count, views);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("pose compose launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

int launch_pose_compose_backward(const float* twists, const float* base, const float* grad_views, int64_t count,
                                 float* grad_twists, float* grad_base, hipStream_t stream) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(pose_compose_backward_kernel, dim3(static_cast<unsigned>((count + 63) / 64)), dim3(64), 0,
                     stream, twists, base, grad_views, count, grad_twists, grad_base);
  const hipError_t error = hipGetLastError();
  if (error != hipSuccess) {
    set_error("pose compose backward launch failed: %s", hipGetErrorString(error));
    return VGR_E_LAUNCH;
  }
  return 0;
}

}  // namespace vgr
