// Error reporting, version, Chebyshev coefficients.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <string>
#include <vector>

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

// Largest root modulus of a monic polynomial of degree n <= 3 (coefficients c[0..n-1] of
// mu^{n-1} .. mu^0), by Durand-Kerner iteration.
static double max_root_modulus(int n, const double* c) {
  using cd = std::complex<double>;
  cd z[3];
  const cd seed(0.4, 0.9);
  z[0] = 1.0;
  for (int i = 1; i < n; ++i) z[i] = z[i - 1] * seed;
  auto eval = [&](cd x) {
    cd v = 1.0;
    for (int i = 0; i < n; ++i) v = v * x + c[i];
    return v;
  };
  for (int it = 0; it < 200; ++it) {
    double change = 0.0;
    for (int i = 0; i < n; ++i) {
      cd den = 1.0;
      for (int j = 0; j < n; ++j)
        if (j != i) den *= z[i] - z[j];
      if (std::abs(den) < 1e-300) den = 1e-300;
      const cd step = eval(z[i]) / den;
      z[i] -= step;
      change = std::max(change, std::abs(step));
    }
    if (change < 1e-15) break;
  }
  double r = 0.0;
  for (int i = 0; i < n; ++i) r = std::max(r, std::abs(z[i]));
  return r;
}

// The largest root modulus over n + 1 grid points of [lam_lo, lmax] (n divides 4096, so the
// grid is a subset of the public function's and its maximum a lower bound of that one).
static double amplification_grid(double lmax, int guess, int sweeps, int refine,
                                 int refine_sweeps, double lam_lo, double cheb_lo, int n);

double step_amplification(double lmax, int guess, int sweeps, int refine, int refine_sweeps,
                          double lam_lo, double cheb_lo) {
  return amplification_grid(lmax, guess, sweeps, refine, refine_sweeps, lam_lo, cheb_lo, 4096);
}

static double amplification_grid(double lmax, int guess, int sweeps, int refine,
                                 int refine_sweeps, double lam_lo, double cheb_lo, int n) {
  if (!(lmax > 1.0)) return 1.0;
  if (!(lam_lo < lmax)) return 0.0;
  const int kmax = std::max(sweeps, refine_sweeps);
  std::vector<ChebCoef> cf(std::max(kmax, 1));
  chebyshev_coefficients(cheb_lo, lmax, kmax, cf.data());
  // residual polynomial of m sweeps from the recurrence, evaluated at lam
  auto poly = [&](double lam, int m) {
    double ep = 1.0, e = 1.0 - cf[0].g * lam;
    for (int k = 1; k < m; ++k) {
      const double en = cf[k].a * e + cf[k].b * ep - cf[k].g * lam * e;
      ep = e;
      e = en;
    }
    return e;
  };
  double worst = 0.0;
  for (int i = 0; i <= n; ++i) {
    const double lam = lam_lo + (lmax - lam_lo) * i / n;
    double P = poly(lam, sweeps);
    for (int r = 0; r < refine; ++r) P *= poly(lam, refine_sweeps);
    const double a = 1.0 + (1.0 - P) * 2.0 * (1.0 - lam) / lam;
    double rad;
    if (guess == 0) {
      rad = std::fabs(a);
    } else if (guess == 1) {
      const double c[2] = {-(a + P), P};
      rad = max_root_modulus(2, c);
    } else {
      const double c[3] = {-(a + 2.0 * P), 3.0 * P, -P};
      rad = max_root_modulus(3, c);
    }
    worst = std::max(worst, rad);
  }
  return worst;
}

int stable_sweeps(double lmax, int guess, int refine, int refine_sweeps, int kmax) {
  for (int k = 2; k <= kmax; ++k)
    if (step_amplification(lmax, guess, k, refine, refine_sweeps) <= 1.0 + 1e-9) return k;
  return -1;
}

// Chebyshev bound on max |p_m| over [1, lmax]: 2 q^m / (1 + q^2m).
static double cheb_bound(double lmax, int m) {
  if (m <= 0) return 1.0;
  const double s = std::sqrt(lmax);
  const double q = (s - 1.0) / (s + 1.0);
  const double qm = std::pow(q, m);
  return 2.0 * qm / (1.0 + qm * qm);
}

// Chebyshev polynomial T_m at x (any real x).
static double cheb_t(int m, double x) {
  if (std::fabs(x) <= 1.0) return std::cos(m * std::acos(x));
  const double c = std::cosh(m * std::acosh(std::fabs(x)));
  return (x < 0.0 && (m & 1)) ? -c : c;
}

static int auto_sweeps_uncached(double lmax, int guess, int refine, int refine_sweeps,
                                int kmax, double resolution, double* cheb_lo);

// Simulations of one voxel size and time step share a spectral bound, so the (count,
// interval) of an input is computed once per process.
int auto_sweeps(double lmax, int guess, int refine, int refine_sweeps, int kmax,
                double resolution, double* cheb_lo) {
  static std::mutex mu;
  static std::map<std::array<double, 6>, std::pair<int, double>> memo;
  const std::array<double, 6> key = {lmax, (double)guess, (double)refine,
                                     (double)refine_sweeps, (double)kmax, resolution};
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = memo.find(key);
    if (it != memo.end()) {
      if (cheb_lo) *cheb_lo = it->second.second;
      return it->second.first;
    }
  }
  double lo = 1.0;
  const int m = auto_sweeps_uncached(lmax, guess, refine, refine_sweeps, kmax, resolution, &lo);
  {
    std::lock_guard<std::mutex> g(mu);
    memo[key] = {m, lo};
  }
  if (cheb_lo) *cheb_lo = lo;
  return m;
}

static int auto_sweeps_uncached(double lmax, int guess, int refine, int refine_sweeps,
                                int kmax, double resolution, double* cheb_lo) {
  if (cheb_lo) *cheb_lo = 1.0;
  const int k0 = stable_sweeps(lmax, guess, refine, refine_sweeps, kmax);
  if (k0 < 0) return -1;
  const double tail = std::pow(cheb_bound(lmax, refine_sweeps), refine);
  const double stiff_max = 1.0 - 1.0 / kStiffSteps;
  const double r = resolution > kFineResolution ? kFineResolution / resolution : 1.0;
  const double accuracy = kAccuracy * r * r;
  // Raised interval [lo, lmax] for the first solve: fine time steps without refinement.
  if (refine == 0 && resolution <= kFineResolution && lmax > 2.0 && lmax <= kMaxRaisedLambda) {
    // The fewest sweeps m for which some lo on a 0.02 grid of [1, kMaxChebLo] meets every
    // condition, with the smallest such lo.  Over lo, max|P| on [lo, lmax] falls and |P(1)|
    // rises, so the closed-form bounds admit one run of grid points; the stiff-mode
    // amplification falls as lo grows, so its threshold is bisected within that run.
    const int nlo = (int)std::lround((kMaxChebLo - 1.0) / 0.02);
    for (int m = 2; m <= kmax; ++m) {
      auto bounds = [&](double lo) {
        const double z0 = (lmax + lo) / (lmax - lo);
        const double bint = 1.0 / cheb_t(m, z0);  // max |P| over [lo, lmax]
        const double p1 = std::fabs(cheb_t(m, (lmax + lo - 2.0) / (lmax - lo))) * bint;
        return lmax * bint <= 0.25 && bint <= accuracy && p1 <= kSmoothAccuracy;
      };
      // a coarse grid first: its maximum bounds the full grid's from below, so a coarse
      // failure is a failure
      auto damped = [&](double lo) {
        return amplification_grid(lmax, guess, m, 0, 0, 2.0, lo, 256) <= stiff_max &&
               amplification_grid(lmax, guess, m, 0, 0, 1.0, lo, 256) <= 1.0 + 1e-9 &&
               step_amplification(lmax, guess, m, 0, 0, 2.0, lo) <= stiff_max &&
               step_amplification(lmax, guess, m, 0, 0, 1.0, lo) <= 1.0 + 1e-9;
      };
      int ja = -1, jb = -1;
      for (int j = 0; j <= nlo; ++j)
        if (bounds(1.0 + 0.02 * j)) {
          if (ja < 0) ja = j;
          jb = j;
        }
      if (ja < 0) continue;
      int a = ja, b = jb;  // damped at b; not at a
      if (damped(1.0 + 0.02 * ja)) b = ja;
      else if (jb == ja || !damped(1.0 + 0.02 * jb)) continue;
      while (b - a > 1) {
        const int mid = (a + b) / 2;
        if (damped(1.0 + 0.02 * mid)) b = mid;
        else a = mid;
      }
      if (cheb_lo) *cheb_lo = 1.0 + 0.02 * b;
      return m;
    }
  }
  for (int k = k0; k <= kmax; ++k)
    if (lmax * cheb_bound(lmax, k) * tail <= 0.25 &&
        (lmax <= 2.0 || cheb_bound(lmax, k) * tail <= accuracy) &&
        (lmax <= 2.0 ||
         step_amplification(lmax, guess, k, refine, refine_sweeps, 2.0) <= stiff_max))
      return k;
  return kmax;
}

double step_resolution(int model, const double* params, double dt) {
  if (model == 0) return dt * (4.0 / 27.0) / params[2];
  // AP: f(u) = k u (u - a) (1 - u) peaks where f'(u) = k (-3 u^2 + 2 (1 + a) u - a) = 0
  const double k = params[0], a = params[1];
  const double disc = (1.0 + a) * (1.0 + a) - 3.0 * a;
  double best = 0.0;
  for (double u : {0.0, 1.0, ((1.0 + a) + std::sqrt(std::max(disc, 0.0))) / 3.0,
                   ((1.0 + a) - std::sqrt(std::max(disc, 0.0))) / 3.0})
    if (u >= 0.0 && u <= 1.0) best = std::max(best, std::fabs(k * u * (u - a) * (1.0 - u)));
  return dt * best;
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

int phs_stable_sweeps(double lambda_max, int32_t guess, int32_t refine, int32_t refine_sweeps,
                      int32_t* stable, int32_t* automatic) {
  if (!stable) return phs::fail(PHS_ERR_ARG, "NULL argument");
  if (guess < 0 || guess > 2 || refine < 0 || refine > 4 ||
      (refine > 0 && (refine_sweeps < 2 || refine_sweeps > 64)) || !(lambda_max >= 1.0))
    return phs::fail(PHS_ERR_ARG, "invalid stability query");
  const int rs = refine > 0 ? refine_sweeps : 0;
  *stable = phs::stable_sweeps(lambda_max, guess, refine, rs, 64);
  if (automatic) *automatic = phs::auto_sweeps(lambda_max, guess, refine, rs, 64);
  return *stable < 0 ? phs::fail(PHS_ERR_UNSUPPORTED, "no stable sweep count <= 64") : PHS_OK;
}

int phs_auto_sweeps(double lambda_max, int32_t guess, int32_t refine, int32_t refine_sweeps,
                    int32_t model, const double* params, double dt, int32_t* sweeps,
                    double* resolution, double* lambda_lo) {
  if (!sweeps || !params || (model != 0 && model != 1) || !(dt > 0))
    return phs::fail(PHS_ERR_ARG, "invalid arguments");
  if (guess < 0 || guess > 2 || refine < 0 || refine > 4 ||
      (refine > 0 && (refine_sweeps < 2 || refine_sweeps > 64)) || !(lambda_max >= 1.0))
    return phs::fail(PHS_ERR_ARG, "invalid stability query");
  const double s = phs::step_resolution(model, params, dt);
  if (resolution) *resolution = s;
  *sweeps = phs::auto_sweeps(lambda_max, guess, refine, refine > 0 ? refine_sweeps : 0, 64, s,
                             lambda_lo);
  return *sweeps < 0 ? phs::fail(PHS_ERR_UNSUPPORTED, "no stable sweep count <= 64") : PHS_OK;
}

int phs_step_amplification(double lambda_max, int32_t guess, int32_t sweeps, int32_t refine,
                           int32_t refine_sweeps, double* amplification) {
  if (!amplification) return phs::fail(PHS_ERR_ARG, "NULL argument");
  if (guess < 0 || guess > 2 || sweeps < 1 || sweeps > 64 || refine < 0 || refine > 4 ||
      (refine > 0 && (refine_sweeps < 1 || refine_sweeps > 64)) || !(lambda_max >= 1.0))
    return phs::fail(PHS_ERR_ARG, "invalid stability query");
  *amplification = phs::step_amplification(lambda_max, guess, sweeps, refine,
                                            refine > 0 ? refine_sweeps : 0);
  return PHS_OK;
}

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
