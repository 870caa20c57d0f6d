// Internal helpers shared by the libphs translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/phs.h"

namespace phs {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Chebyshev semi-iteration for A d = r with spec(A) in [lmin, lmax]:
//   d_1 = r / theta,  d_{k+1} = a_k d_k + b_k d_{k-1} + g_k (r - A d_k)   (k >= 1)
// with rho_0 = delta/theta, rho_k = 1 / (2 theta/delta - rho_{k-1}),
//   a_k = 1 + rho_k rho_{k-1},  b_k = -rho_k rho_{k-1},  g_k = 2 rho_k / delta.
struct ChebCoef {
  double a, b, g;
};
void chebyshev_coefficients(double lmin, double lmax, int n, ChebCoef* out /* [n], out[0].g = 1/theta */);
// Sweeps needed so that the Chebyshev error bound 2 q^m / (1 + q^2m) <= tol.
int chebyshev_sweeps(double lmin, double lmax, double tol);

// Time-step stability of the inexact CN step.  For an eigenmode of A with eigenvalue lam in
// [1, lmax] (CN right-hand side b = 2(1 - lam) u / lam per unit mode), a solve whose error
// polynomial is P(lam) = p_sweeps(lam) p_refine(lam)^refine maps
//   guess 0: u' = a u                               a = 1 + (1 - P) 2 (1 - lam) / lam
//   guess 1: (u, d) -> companion of mu^2 - (a + P) mu + P
//   guess 2: (u, d, d_prev) -> companion of mu^3 - (a + 2P) mu^2 + 3P mu - P
// (exact solve: P = 0, u' = (2/lam - 1) u, the CN factor).  Returns the largest root
// modulus over a 4097-point grid of [lam_lo, lmax] (default the whole spectrum [1, lmax]);
// > 1 means the step grows that mode.
double step_amplification(double lmax, int guess, int sweeps, int refine, int refine_sweeps,
                          double lam_lo = 1.0);
// Fewest sweeps (2..kmax) whose step amplification is <= 1, or -1.
int stable_sweeps(double lmax, int guess, int refine, int refine_sweeps, int kmax);
// Automatic count: the fewest stable sweeps m for which
//  (1) the solve error at the stiffest mode stays under a quarter of CN's own damping there,
//      lmax * max|P| <= 1/4 with the Chebyshev bound max|p_m| = 2 q^m / (1 + q^2m),
//      q = (sqrt(lmax) - 1) / (sqrt(lmax) + 1), and
//  (2) every stiff mode (lam >= 2, the modes CN damps with a change of sign) shrinks by a
//      factor e within kStiffSteps steps: amplification over [2, lmax] <= 1 - 1/kStiffSteps.
//      The smooth modes (lam < 2) keep CN's physical factor close to 1.
// (clamped to kmax; -1 when no count is stable).
//  (3) the solve is accurate enough for the step's time resolution: with s = dt times the cell
//      model's steepest upstroke rate (step_resolution), the Chebyshev bound (times the
//      refinement passes' bounds) is <= kAccuracy * min(1, (kFineResolution / s)^2).  A
//      warm-started solve corrects only what the previous steps did not predict, which grows
//      like s^2 for the extrapolated warm start; the constants are those of the 0.4 mm shell at
//      dt = 0.05 ms (s = 0.025, 6 sweeps), which meets the activation-time and electrogram bars
//      over 40 000 steps, so steps that coarse or finer keep the first two conditions only.
constexpr double kStiffSteps = 40.0;
constexpr double kAccuracy = 0.03;
constexpr double kFineResolution = 0.025;
int auto_sweeps(double lmax, int guess, int refine, int refine_sweeps, int kmax,
                double resolution = 0.0);
// dt times the steepest upstroke rate of the cell model at rest-to-peak (h = 1 for
// Mitchell-Schaeffer, h = 0 for Aliev-Panfilov): MS (4/27) / tau_in, AP max_u k u (u - a) (1 - u).
// params as in phs_sim_desc.
double step_resolution(int model, const double* params, double dt);

}  // namespace phs

#define PHS_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return phs::fail(PHS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PHS_TRY_ALLOC(ptr, bytes)                                                       \
  do {                                                                                  \
    hipError_t _e = hipMalloc((void**)&(ptr), (bytes) > 0 ? (bytes) : 1);               \
    if (_e != hipSuccess)                                                               \
      return phs::fail(PHS_ERR_OOM, std::string("hipMalloc ") + std::to_string(bytes) + \
                                        " bytes: " + hipGetErrorString(_e));            \
  } while (0)

// LAT bookkeeping in the per-voxel flag byte: bits 4-6 beat count, bit 7 above threshold.
#define PHS_VF_BEAT_SHIFT 4
#define PHS_VF_BEAT_MASK 0x70
#define PHS_VF_ABOVE 0x80
