// Conduction velocity from activation times at scattered points (utility/conduction_velocity.py:
// 23-70) on the device: per point, the plane lat = a x + b y + c z + d fitted by least squares
// to the points within `radius`, v = g / |g|^2 with g = (a, b, c).
//
// One wavefront per point.  Lanes stride over all points and accumulate, in FP64, the normal
// equations of the fit on the raw coordinates (the reference's column order x, y, z, 1): the
// 4x4 matrix M = A^T A and the vector A^T lat.  M is diagonalised by cyclic Jacobi rotations
// and the minimum-norm solution x = sum_k (v_k . A^T lat / lambda_k) v_k is formed over the
// eigenvalues above a relative threshold, which is what np.linalg.lstsq returns for full-rank
// and exactly rank-deficient neighbourhoods (coplanar or collinear points, e.g. a flat sheet).
// Neighbour test: (dx^2 + dy^2) + dz^2 <= radius^2 in FP64, the rule of the reference's
// cKDTree.query_ball_point (inclusive, same summation order).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "phs_internal.h"

namespace {

constexpr int kCvWave = 64;
constexpr int kMeanThreads = 256;
constexpr int kSums = 15;  // count, then the upper triangle of M (10), then A^T lat (4)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Eigen-decomposition of a symmetric 4x4 matrix by cyclic Jacobi rotations (each rotation
// zeroes one off-diagonal pair; the rotation angle follows the classical t = sgn(theta) /
// (|theta| + sqrt(theta^2 + 1)) form).  On return a holds the eigenvalues on its diagonal and
// the columns of v the eigenvectors.
__device__ void jacobi4(double (&a)[4][4], double (&v)[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < 4; ++p) {
      diag += a[p][p] * a[p][p];
      for (int q = p + 1; q < 4; ++q) off += a[p][q] * a[p][q];
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < 3; ++p) {
      for (int q = p + 1; q < 4; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {  // columns p, q
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {  // rows p, q
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        a[p][q] = a[q][p] = 0.0;
        for (int k = 0; k < 4; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kCvWave) cv_fit_k(const double* __restrict__ xyz,
                                                     const double* __restrict__ lat, int32_t n,
                                                     double r2, double* __restrict__ vec,
                                                     double* __restrict__ mag) {
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  double s[kSums];
#pragma unroll
  for (int k = 0; k < kSums; ++k) s[k] = 0.0;
  for (int j = lane; j < n; j += kCvWave) {
    const double x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
    const double dx = x - xi, dy = y - yi, dz = z - zi;
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    if (d2 <= r2) {
      const double t = lat[j];
      s[0] += 1.0;
      s[1] += x * x;
      s[2] += x * y;
      s[3] += x * z;
      s[4] += x;
      s[5] += y * y;
      s[6] += y * z;
      s[7] += y;
      s[8] += z * z;
      s[9] += z;
      s[10] += 1.0;
      s[11] += x * t;
      s[12] += y * t;
      s[13] += z * t;
      s[14] += t;
    }
  }
#pragma unroll
  for (int k = 0; k < kSums; ++k) s[k] = wave_sum(s[k]);
  if (lane != 0) return;
  const double nan = __builtin_nan("");
  const double cnt = s[0];
  if (cnt < 4.0) {
    vec[3 * i] = vec[3 * i + 1] = vec[3 * i + 2] = nan;
    mag[i] = nan;
    return;
  }
  double a[4][4] = {{s[1], s[2], s[3], s[4]},
                    {s[2], s[5], s[6], s[7]},
                    {s[3], s[6], s[8], s[9]},
                    {s[4], s[7], s[9], s[10]}};
  const double g[4] = {s[11], s[12], s[13], s[14]};
  double v[4][4];
  jacobi4(a, v);
  double lmax = 0.0;
  for (int k = 0; k < 4; ++k) lmax = fmax(lmax, a[k][k]);
  // eigenvalues of M are the squared singular values of A; those within the rounding of
  // forming M (a few count * eps * lambda_max) belong to an exactly rank-deficient A
  const double cut = 8.0 * fmax(cnt, 4.0) * 2.220446049250313e-16 * lmax;
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < 4; ++k) {
    const double lk = a[k][k];
    if (!(lk > cut)) continue;
    double p = 0.0;
    for (int m = 0; m < 4; ++m) p += v[m][k] * g[m];
    p /= lk;
    for (int m = 0; m < 4; ++m) x[m] += p * v[m][k];
  }
  const double gn2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  if (gn2 > 1e-8) {
    vec[3 * i] = x[0] / gn2;
    vec[3 * i + 1] = x[1] / gn2;
    vec[3 * i + 2] = x[2] / gn2;
    mag[i] = 1.0 / sqrt(gn2);
  } else {
    vec[3 * i] = vec[3 * i + 1] = vec[3 * i + 2] = nan;
    mag[i] = nan;
  }
}

// Mean over the finite magnitudes (NaN if there is none), in a fixed summation order.
__global__ void __launch_bounds__(kMeanThreads) cv_mean_k(const double* __restrict__ mag, int32_t n,
                                                          double* __restrict__ mean) {
  __shared__ double ss[kMeanThreads];
  __shared__ double sc[kMeanThreads];
  double sum = 0.0, cnt = 0.0;
  for (int i = threadIdx.x; i < n; i += kMeanThreads) {
    const double m = mag[i];
    if (!isnan(m)) {
      sum += m;
      cnt += 1.0;
    }
  }
  ss[threadIdx.x] = sum;
  sc[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = kMeanThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *mean = sc[0] > 0.0 ? ss[0] / sc[0] : __builtin_nan("");
}

}  // namespace

extern "C" {

int phs_conduction_velocity(const double* xyz, const double* lat, int32_t n_points, double radius,
                            double* velocity, double* magnitude, double* mean, void* stream) {
  if (!xyz || !lat || !velocity || !magnitude || n_points <= 0 || !(radius >= 0.0))
    return phs::fail(PHS_ERR_ARG, "invalid conduction velocity arguments");
  hipStream_t s = (hipStream_t)stream;
  cv_fit_k<<<n_points, kCvWave, 0, s>>>(xyz, lat, n_points, radius * radius, velocity, magnitude);
  if (mean) cv_mean_k<<<1, kMeanThreads, 0, s>>>(magnitude, n_points, mean);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

}  // extern "C"
