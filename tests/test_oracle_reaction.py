"""Pins of the oracle's reaction step (oracle/sim_oracle.py: reaction_rhs, rk4_step).

Each check is fixed by mathematics, not by re-typing the formula:
  * rest states are exact fixed points (MS (0, 1), AP (0, 0));
  * closed-form single-cell solutions on the invariant sub-cases of
    Mitchell-Schaeffer (u = 0 gate opening; h = 0 with u above v_gate);
  * the Aliev-Panfilov cubic vanishes exactly at u in {0, a, 1} when h = 0;
  * RK4 converges at fourth order to scipy's high-accuracy ODE solution.
"""
import numpy as np
from scipy.integrate import solve_ivp

from oracle import sim_oracle as so

MS = (120.0, 150.0, 0.3, 6.0, 0.13)    # tau_open, tau_close, tau_in, tau_out, v_gate
AP = (8.0, 0.15, 0.002, 0.2, 0.3)      # k, a, eps0, mu1, mu2


def _p(params, n=1):
    return tuple(np.full(n, v) for v in params)


def _run(model, u0, h0, params, J, dt, n):
    u, h = np.array([u0], float), np.array([h0], float)
    Jv = np.array([J], float)
    for _ in range(n):
        u, h = so.rk4_step(model, u, h, _p(params), Jv, dt)
    return u[0], h[0]


def test_rest_states_are_fixed_points():
    for model, (u0, h0), params in ((0, (0.0, 1.0), MS), (1, (0.0, 0.0), AP)):
        u, h = _run(model, u0, h0, params, 0.0, 0.05, 200)
        assert u == u0 and h == h0


def test_ms_gate_opens_exponentially_at_u_zero():
    # u = 0 is invariant (du = 0); dh = (1 - h) / tau_open  ->  h = 1 - (1 - h0) e^{-t/tau_open}
    dt, n, h0 = 0.05, 2000, 0.3
    u, h = _run(0, 0.0, h0, MS, 0.0, dt, n)
    t = dt * n
    assert u == 0.0
    assert abs(h - (1 - (1 - h0) * np.exp(-t / MS[0]))) < 1e-12


def test_ms_decay_above_gate_with_closed_gate():
    # h = 0 stays 0 while u >= v_gate (dh = -h / tau_close); du = -u / tau_out + J.
    dt, n = 0.01, 1000
    for u0, J in ((0.9, 0.0), (0.5, 1.0)):
        u, h = _run(0, u0, 0.0, MS, J, dt, n)
        t = dt * n
        exact = J * MS[3] + (u0 - J * MS[3]) * np.exp(-t / MS[3])
        assert exact > MS[4]
        assert h == 0.0
        assert abs(u - exact) < 1e-12


def test_ms_gate_closes_exponentially_above_threshold():
    # while u >= v_gate, h evolves alone: h = h0 e^{-t/tau_close}
    dt, n, h0 = 0.01, 200, 0.5
    u = np.array([0.95]); h = np.array([h0])
    for _ in range(n):
        u, h = so.rk4_step(0, u, h, _p(MS), np.zeros(1), dt)
        assert u[0] >= MS[4]
    assert abs(h[0] - h0 * np.exp(-dt * n / MS[1])) < 1e-13


def test_ap_cubic_roots_and_known_value():
    k, a = AP[0], AP[1]
    for u in (0.0, a, 1.0):
        du, _ = so.reaction_rhs(1, np.array([u]), np.zeros(1), *_p(AP), np.zeros(1))
        assert du[0] == 0.0
    # hand value: u = 0.5, h = 0.2, J = 0: du = -8*0.5*0.35*(-0.5) - 0.1 = 0.6
    #             dh = (0.002 + 0.2*0.2/0.8) * (-0.2 - 8*0.5*(0.5-1.15)) = 0.052 * 2.4
    du, dh = so.reaction_rhs(1, np.array([0.5]), np.array([0.2]), *_p(AP), np.zeros(1))
    assert abs(du[0] - 0.6) < 1e-15
    assert abs(dh[0] - 0.052 * 2.4) < 1e-15
    # MS hand value: u = 0.5, h = 0.8, J = 0.25 below/above the gate
    du, dh = so.reaction_rhs(0, np.array([0.5]), np.array([0.8]), *_p(MS), np.array([0.25]))
    assert abs(du[0] - (0.8 * 0.25 * 0.5 / 0.3 - 0.5 / 6 + 0.25)) < 1e-15
    assert abs(dh[0] - (-0.8 / 150)) < 1e-18


def test_rk4_fourth_order_against_scipy():
    # Aliev-Panfilov is smooth; compare RK4 at dt, dt/2, dt/4 to a tight DOP853 solution.
    def f(t, y):
        du, dh = so.reaction_rhs(1, np.array([y[0]]), np.array([y[1]]), *_p(AP), np.zeros(1))
        return [du[0], dh[0]]
    y0, T = (0.4, 0.05), 2.0
    ref = solve_ivp(f, (0, T), y0, method="DOP853", rtol=1e-13, atol=1e-15).y[:, -1]
    errs = []
    for dt in (0.1, 0.05, 0.025):
        u, h = _run(1, y0[0], y0[1], AP, 0.0, dt, int(round(T / dt)))
        errs.append(np.hypot(u - ref[0], h - ref[1]))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12 < r1 < 20 and 12 < r2 < 20, (errs, r1, r2)
