"""Time-step stability of the inexact CN step (host-only C ABI: phs_step_amplification,
phs_stable_sweeps, phs_auto_sweeps).

Pins:
  * the per-mode growth factor equals the spectral radius of the mode map built here
    directly from its definition (u' = u + d, d = d* + P (d0 - d*), d* = b / lam) with
    numpy.linalg.eigvals -- not from the characteristic polynomials the library uses;
  * an exact solve (many sweeps) reduces to the CN factor |2/lam - 1| <= 1;
  * a 1D chain with zero-flux ends, stepped on the CPU with the same Chebyshev recurrence,
    grows without bound two sweep counts below the stability limit and stays bounded at
    the limit (a plain-matrix restatement, no library code);
  * the stability boundaries on [1, lambda_max] measured on the shells: 14 sweeps at
    lambda_max 31 (0.2 mm) and 5 at 8.5 (0.4 mm) with the previous-step warm start;
  * the automatic rule recomputed here from the mode matrices and closed-form Chebyshev
    bounds: at fine time steps the fewest sweeps m for which some lower end lo of the
    interval [lo, lambda_max] (a 0.02 grid of [1, 1.5]) keeps the step stable, shrinks every
    stiff mode (lam >= 2) by e within 40 steps, bounds the error on the interval by
    1/(4 lambda_max) and 0.03, and the smoothest mode's error |P(1)| by 0.16; the smallest
    such lo, for lambda_max <= 10.  With the extrapolated warm start that is 5 sweeps on
    [1.42, 8.5] for the 0.4 mm shell (validated over 20 000 steps against the plain FP64
    form: LAT within 1 step, no mismatched activation, chain electrogram rel-L2 <= 2e-4).
    Larger spectra (0.2 mm: 31, where 14 sweeps on [1.38, 31] left activations unmatched)
    and coarse time steps keep the interval [1, lambda_max]: 17 sweeps at 31, and the
    accuracy bound 0.03 (0.025 / s)^2 at coarse steps.
"""
import numpy as np
import pytest

from pyheartsim_amd import runtime


def cheb_coefs(lmin, lmax, n):
    th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    rp = de / th
    out = [(0.0, 0.0, 1.0 / th)]
    for _ in range(1, n):
        r = 1.0 / (2.0 * th / de - rp)
        out.append((1.0 + r * rp, -r * rp, 2.0 * r / de))
        rp = r
    return out


def cheb_solve(apply_a, b, x0, lmax, m):
    """m sweeps of the Chebyshev semi-iteration on A x = b from x0, spec(A) in [1, lmax]."""
    c = cheb_coefs(1.0, lmax, m)
    xp = x0
    x = x0 + c[0][2] * (b - apply_a(x0))
    for k in range(1, m):
        a, bb, g = c[k]
        x, xp = a * x + bb * xp + g * (b - apply_a(x)), x
    return x


def err_poly(lam, lmax, m, lo=1.0):
    """Error polynomial of m Chebyshev sweeps on [lo, lmax] at lam (the recurrence applied to
    the scalar problem lam x = 0 from x = 1)."""
    c = cheb_coefs(lo, lmax, m)
    xp = np.ones_like(lam)
    x = xp - c[0][2] * lam * xp
    for k in range(1, m):
        a, bb, g = c[k]
        x, xp = a * x + bb * xp - g * lam * x, x
    return x


def mode_radius(lam, lmax, m, guess, refine=0, rs=8, lo=1.0):
    P = err_poly(np.array([lam]), lmax, m, lo)[0]
    for _ in range(refine):
        P *= err_poly(np.array([lam]), lmax, rs)[0]
    return radius_of(lam, P, guess)


def radius_of(lam, P, guess):
    c = 2.0 * (1.0 - lam) / lam          # d* = c u
    if guess == 0:
        M = np.array([[1.0 + (1.0 - P) * c]])
    elif guess == 1:                      # state (u, d_prev), d0 = d_prev
        d = np.array([(1.0 - P) * c, P])
        M = np.vstack([d, d])
        M[0, 0] += 1.0
    else:                                 # state (u, d_prev, d_prev2), d0 = 2 d_prev - d_prev2
        d = np.array([(1.0 - P) * c, 2.0 * P, -P])
        M = np.zeros((3, 3))
        M[0] = d
        M[0, 0] += 1.0
        M[1] = d
        M[2, 1] = 1.0
    return np.max(np.abs(np.linalg.eigvals(M)))


@pytest.mark.parametrize("lmax,m,guess,refine", [
    (8.5, 6, 1, 0), (8.5, 4, 1, 0), (31.0, 6, 1, 0), (31.0, 12, 1, 0), (31.0, 16, 0, 0),
    (31.0, 9, 2, 0), (31.0, 6, 1, 1), (61.0, 20, 1, 0), (16.0, 7, 0, 0)])
def test_amplification_matches_mode_matrix(lmax, m, guess, refine):
    lams = np.linspace(1.0, lmax, 4097)
    want = max(mode_radius(l, lmax, m, guess, refine) for l in lams)
    got = runtime.step_amplification(lmax, m, guess, refine, 8)
    assert abs(got - want) <= 1e-9 * max(1.0, want)


def test_exact_solve_reduces_to_cn_factor():
    # 64 sweeps at lambda_max 8.5: |P| < 1e-30, the map is the CN factor, max 1 at lam = 1
    assert abs(runtime.step_amplification(8.5, 64, 1) - 1.0) <= 1e-12
    lam = np.linspace(1.0, 8.5, 101)
    assert np.all(np.abs(2.0 / lam - 1.0) <= 1.0)


def test_counts_measured_on_shells():
    assert runtime.stable_sweeps(31.0, 1)[0] == 14
    assert runtime.stable_sweeps(8.5, 1)[0] == 5
    assert runtime.stable_sweeps(31.0, 2)[0] == 16
    assert runtime.stable_sweeps(8.5, 2) == (6, 5)
    assert runtime.stable_sweeps(31.0, 2)[1] == 17
    assert runtime.step_amplification(31.0, 13, 1) > 1.005
    assert runtime.step_amplification(31.0, 14, 1) <= 1.0 + 1e-9


def _chain_run(m, lmax_bound, n=48, steps=400, alpha=2.5):
    # 1D chain with zero-flux ends: (L u)_i = sum over present neighbours j of (u_j - u_i),
    # a missing neighbour contributing nothing (the reference's operator,
    # legacy/compute_simulation_cpu.py:21-111); A = I - alpha L
    L = np.zeros((n, n))
    for i in range(n):
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                L[i, j] += 1.0
                L[i, i] -= 1.0
    A = np.eye(n) - alpha * L
    rng = np.random.default_rng(0)
    u = rng.standard_normal(n)
    d_prev = np.zeros(n)
    peak = 0.0
    for _ in range(steps):
        b = 2.0 * alpha * (L @ u)
        d = cheb_solve(lambda v: A @ v, b, d_prev, lmax_bound, m)
        u = u + d
        d_prev = d
        peak = max(peak, float(np.max(np.abs(u))))
    return peak


def test_chain_diverges_below_limit_and_not_at_it():
    lmax = 1.0 + 2.0 * 2.5 * 2          # Gershgorin bound, alpha 2.5, one axis: 11
    stable, _ = runtime.stable_sweeps(lmax, 1)
    start = _chain_run(64, lmax, steps=1)
    assert _chain_run(stable, lmax) <= 10.0 * start
    assert _chain_run(stable - 2, lmax) > 1e3 * start


def _stiff_radius(lmax, m, guess, lo=1.0, start=2.0, n=4097):
    lams = np.linspace(start, lmax, n)
    P = err_poly(lams, lmax, m, lo)
    return max(radius_of(l, p, guess) for l, p in zip(lams, P))


def _bound(lmax, m, lo=1.0):
    """max |P| over [lo, lmax] = 1 / T_m((lmax + lo) / (lmax - lo))."""
    return 1.0 / np.cosh(m * np.arccosh((lmax + lo) / (lmax - lo)))


def _p1(lmax, m, lo):
    """|P(1)|: T_m at the image of lam = 1 (outside the interval for lo > 1)."""
    z = (lmax + lo - 2.0) / (lmax - lo)
    return np.cosh(m * np.arccosh(max(z, 1.0))) * _bound(lmax, m, lo)


def _accurate(lmax, m, resolution=0.0):
    target = 0.03 * min(1.0, (0.025 / resolution) ** 2) if resolution > 0 else 0.03
    return lmax * _bound(lmax, m) <= 0.25 and (lmax <= 2.0 or _bound(lmax, m) <= target)


def _fine_ok(lmax, m, lo, guess):
    """Every condition of the rule at fine time steps for m sweeps on [lo, lmax]."""
    b = _bound(lmax, m, lo)
    if not (lmax * b <= 0.25 and b <= 0.03 and _p1(lmax, m, lo) <= 0.16):
        return False
    # a coarse grid (a subset of the fine one) first: failing there is failing
    for n in (257, 4097):
        if _stiff_radius(lmax, m, guess, lo, 2.0, n) > 1.0 - 1.0 / 40.0:
            return False
        if _stiff_radius(lmax, m, guess, lo, 1.0, n) > 1.0 + 1e-9:
            return False
    return True


LO_GRID = [1.0 + 0.02 * j for j in range(26)]


@pytest.mark.parametrize("lmax,guess", [(3.0, 2), (8.5, 1), (8.5, 2), (16.0, 2), (31.0, 1),
                                        (31.0, 2), (60.0, 1)])
def test_automatic_count_rule(lmax, guess):
    """The library's (count, interval) at fine time steps meets the rule, no smaller count
    meets it with any interval of the grid, and no lower interval end meets it at that count."""
    from tests.helpers import small_shell
    w = small_shell(dt=0.05, t_final=0.1)
    P = w["P_2d"]
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    m, s, lo = runtime.auto_sweeps(lmax, prep, guess=guess)
    assert s <= 0.025
    assert runtime.stable_sweeps(lmax, guess)[1] == m
    stable, _ = runtime.stable_sweeps(lmax, guess)
    old = next(k for k in range(stable, 65) if _accurate(lmax, k)
               and _stiff_radius(lmax, k, guess) <= 1.0 - 1.0 / 40.0)
    if lmax > 10.0:           # the interval stays [1, lambda_max]: the rule on it alone
        assert lo == 1.0 and m == old
        return
    j = round((lo - 1.0) / 0.02)
    assert abs(lo - LO_GRID[j]) <= 1e-12
    assert _fine_ok(lmax, m, lo, guess)
    assert not any(_fine_ok(lmax, m, x, guess) for x in LO_GRID[:j])
    assert not any(_fine_ok(lmax, k, x, guess) for k in range(2, m) for x in LO_GRID)
    assert m <= old           # never more sweeps than the rule on [1, lambda_max] alone


@pytest.mark.parametrize("dt,expect,expect_lo", [(0.05, 5, 1.42), (0.5, 16, 1.0)])
def test_automatic_count_follows_time_resolution(dt, expect, expect_lo):
    """The 0.4 mm shell's spectrum at dt = 0.05 ms (s = 0.025) takes the fine-step rule: 5
    sweeps on [1.42, 8.5].  The reference's patient-mesh step dt = 0.5 ms (s = 0.25, ten
    times coarser) keeps [1, lambda_max] and needs the Chebyshev bound below 0.03 / 100:
    16 sweeps at lambda_max = 13 (1 mm voxels)."""
    from tests.helpers import small_shell
    w = small_shell(dt=dt, t_final=2 * dt)
    P = w["P_2d"]
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    lmax = 8.5 if dt == 0.05 else 13.0
    m, s, lo = runtime.auto_sweeps(lmax, prep, guess=2)
    # Mitchell-Schaeffer: s = dt (4/27) / tau_in
    assert s == pytest.approx(dt * (4.0 / 27.0) / 0.3, rel=1e-12)
    assert m == expect and lo == pytest.approx(expect_lo, abs=1e-12)
    if dt > 0.05:
        stable, _ = runtime.stable_sweeps(lmax, 2)

        def ok(k):
            return (k >= stable and _accurate(lmax, k, s)
                    and _stiff_radius(lmax, k, 2) <= 1.0 - 1.0 / 40.0)

        assert ok(m) and not any(ok(k) for k in range(stable, m))


def test_refinement_keeps_the_whole_interval():
    """With FP64-residual refinement the interval stays [1, lambda_max]."""
    from tests.helpers import small_shell
    w = small_shell(dt=0.05, t_final=0.1)
    P = w["P_2d"]
    prep = runtime.prepare(P.shape[0], P, w["geometry"], w["sim"], w["arr"])
    m, s, lo = runtime.auto_sweeps(24.5, prep, guess=2, refine=3, refine_sweeps=12)
    assert lo == 1.0 and m >= 2
