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
// cheb_lo: lower end of the Chebyshev interval of the solves, [cheb_lo, lmax].
double step_amplification(double lmax, int guess, int sweeps, int refine, int refine_sweeps,
                          double lam_lo = 1.0, double cheb_lo = 1.0);
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
// At fine time steps without refinement the first solve may run its Chebyshev semi-iteration
// on a raised interval [lo, lmax], lo in [1, kMaxChebLo] (*cheb_lo, else 1; lo = 1 + dt c k^2 / 2
// for a mode of wavenumber k whatever the voxel size, so the cap keeps the modes that the
// validation below solved to the interval's accuracy solved so everywhere): conditions (1)-(3)
// then hold with max|P| taken over [lo, lmax] (a closed form), (2) for the polynomial of that
// interval, and the smoothest mode's solve error |P(1)| (the largest on [1, lo]) must stay
// <= kSmoothAccuracy.  Raising lo damps the stiff modes with fewer sweeps and leaves more of
// the smooth modes' warm-start residual, which the extrapolated warm start keeps small: on
// the 0.4 mm shell (lambda_max 8.5) 5 sweeps on [1.4, 8.5] (|P(1)| 0.14) and [1.5, 8.5]
// (0.16) kept every activation within 1 step of the plain FP64 form over 20 000 steps and
// the chain electrogram within rel-L2 2e-4 (6 sweeps on [1, 8.5]: 3e-5).  The fewest such m
// with the smallest lo on a 0.02 grid is taken; it is never more sweeps than on [1, lmax].
// The raised interval is limited to lmax <= kMaxRaisedLambda: on the 0.2 mm shell with scar
// (lmax 31) the rule's 14 sweeps on [1.38, 31] left 212 of 1.3 M activations unmatched and
// one 2 steps late over 6000 steps, where 17 on [1, 31] match them all.
constexpr double kStiffSteps = 40.0;
constexpr double kAccuracy = 0.03;
constexpr double kFineResolution = 0.025;
constexpr double kSmoothAccuracy = 0.16;
constexpr double kMaxChebLo = 1.5;
constexpr double kMaxRaisedLambda = 10.0;
int auto_sweeps(double lmax, int guess, int refine, int refine_sweeps, int kmax,
                double resolution = 0.0, double* cheb_lo = nullptr);
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
