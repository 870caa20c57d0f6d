// Error reporting, version, Chebyshev coefficients.
#include <cmath>
#include <string>

#include "phs_internal.h"

namespace phs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

void chebyshev_coefficients(double lmin, double lmax, int n, ChebCoef* out) {
  const double theta = 0.5 * (lmax + lmin);
  const double delta = 0.5 * (lmax - lmin);
  double rho_prev = delta / theta;
  if (n > 0) out[0] = ChebCoef{0.0, 0.0, 1.0 / theta};
  for (int k = 1; k < n; ++k) {
    const double rho = 1.0 / (2.0 * theta / delta - rho_prev);
    out[k].a = 1.0 + rho * rho_prev;
    out[k].b = -rho * rho_prev;
    out[k].g = 2.0 * rho / delta;
    rho_prev = rho;
  }
}

int chebyshev_sweeps(double lmin, double lmax, double tol) {
  const double kappa = lmax / lmin;
  const double q = (std::sqrt(kappa) - 1.0) / (std::sqrt(kappa) + 1.0);
  if (q <= 0.0) return 1;
  for (int m = 1; m < 100000; ++m) {
    const double qm = std::pow(q, m);
    if (2.0 * qm / (1.0 + qm * qm) <= tol) return m;
  }
  return 100000;
}

}  // namespace phs

extern "C" {

int phs_version(void) { return 10000; }

const char* phs_last_error(void) { return phs::g_last_error.c_str(); }

int phs_device_count(int32_t* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return phs::fail(PHS_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return PHS_OK;
}

}  // extern "C"
