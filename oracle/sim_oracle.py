"""FP64 CPU oracle of the simulation hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product path never routes
through it.

A NumPy/SciPy restatement of the reference's FP64 CPU path:
  * reaction RHS        legacy/compute_simulation_cpu.py:144-160
                        (= simulation/simulation_gpu.py:27-43)
  * RK4 step            legacy/compute_simulation_cpu.py:162-209
  * diffusion operator  legacy/compute_simulation_cpu.py:21-111
                        (= simulation/simulation_gpu.py:138-225)
  * Crank-Nicolson      legacy/compute_simulation_cpu.py:113-142
  * time loop, blocks,  simulation/simulation_gpu.py:261-409
    sampling, init
  * pacing rule         simulation/pacing.py:102-170

Documented deviation: the CN system is solved tightly (sparse LU, or
BiCGSTAB to 1e-13 relative residual) instead of the reference's loose
BiCGSTAB stop rule, so that solver iterate paths never need matching.  The
reference's stop rule is available as solver="bicgstab_ref".

Additions not in the reference (the build's metric definitions):
  * per-voxel LAT: the step index s whose post-step u first reaches
    ``lat_threshold`` (v_gate = 0.13) from below, per beat, up to
    ``max_beats`` beats;
  * probe traces: u after every step at a list of voxels.

Pinned by tests/test_oracle_*.py: rest-state fixed points, closed-form
single-cell decays, RK4 convergence order against scipy's solve_ivp, row
sums / symmetry of L, the Neumann cosine-mode known answer of one CN step,
block semantics, and hand-derived pacing step tables.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


# ---------------------------------------------------------------- reaction
def reaction_rhs(model, u, h, p15, p16, p17, p18, p19, J):
    """legacy/compute_simulation_cpu.py:144-160, same operation order."""
    if model == 0:
        du = (h * (u ** 2) * (1 - u)) / p17 - u / p18 + J
        dh = np.where(u < p19, (1 - h) / p15, -h / p16)
    else:
        du = -p15 * u * (u - p16) * (u - 1) - u * h + J
        dh = (p17 + p18 * h / (u + p19)) * (-h - p15 * u * (u - p16 - 1))
    return du, dh


def rk4_step(model, u, h, params, J, dt):
    """legacy/compute_simulation_cpu.py:162-209 (classic RK4, J constant)."""
    k1u, k1h = reaction_rhs(model, u, h, *params, J)
    k2u, k2h = reaction_rhs(model, u + 0.5 * dt * k1u, h + 0.5 * dt * k1h, *params, J)
    k3u, k3h = reaction_rhs(model, u + 0.5 * dt * k2u, h + 0.5 * dt * k2h, *params, J)
    k4u, k4h = reaction_rhs(model, u + dt * k3u, h + dt * k3h, *params, J)
    u_next = u + (dt / 6.0) * (k1u + 2 * k2u + 2 * k3u + k4u)
    h_next = h + (dt / 6.0) * (k1h + 2 * k2h + 2 * k3h + k4h)
    return u_next, h_next


# ---------------------------------------------------------------- operator
_CROSS = ((6, 8, 9), (9, 7, 10), (14, 16, 11), (17, 15, 12), (10, 12, 13), (13, 11, 14))


def build_diffusion_matrix(P_2d, neighbor_id_2d, Delta):
    """legacy/compute_simulation_cpu.py:21-111 as a vectorised COO build.

    ``neighbor_id_2d`` may hold -1; it is mapped to 0 as the caller does at
    :255-256 (the matching P_2d weight is 0, so those entries are zeros).
    """
    P = np.asarray(P_2d, dtype=np.float64)
    n = P.shape[0]
    nb = np.asarray(neighbor_id_2d).copy()
    nb[nb == -1] = 0
    Dc = P[:, 20] / (4 * Delta ** 2)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [Dc * -(P[:, 0] + P[:, 1] + P[:, 2] +
                                                             P[:, 3] + P[:, 4] + P[:, 5])]
    ar = np.arange(n)
    for k in range(6):
        rows.append(ar); cols.append(nb[:, k]); vals.append(Dc * P[:, k])
    for k in range(3):
        rows.append(ar); cols.append(nb[:, 2 * k]); vals.append(Dc * P[:, 6 + k])
        rows.append(ar); cols.append(nb[:, 2 * k + 1]); vals.append(-Dc * P[:, 6 + k])
    for a, b, c in _CROSS:
        rows.append(ar); cols.append(nb[:, a]); vals.append(Dc * P[:, c])
        rows.append(ar); cols.append(nb[:, b]); vals.append(-Dc * P[:, c])
    L = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return L.tocsr()


class CNSolver:
    """Crank-Nicolson step of legacy/compute_simulation_cpu.py:113-142:
    solve (I - dt/2 L) x = u* + dt/2 L u*, x0 = u*."""

    def __init__(self, L, dt, method="direct"):
        self.L = L
        self.dt = dt
        n = L.shape[0]
        self.A = (sp.identity(n, format="csr") - (dt / 2.0) * L).tocsr()
        self.method = method
        self.lu = spla.splu(self.A.tocsc()) if method == "direct" else None
        self.iterations = []

    def __call__(self, u_star):
        b = u_star + (self.dt / 2.0) * (self.L @ u_star)
        if self.method == "direct":
            return self.lu.solve(b)
        if self.method == "bicgstab_ref":
            # reference stop rule: atol 1e-6, scipy default rtol 1e-5, maxiter 1000
            x, info = spla.bicgstab(self.A, b, x0=u_star, atol=1e-6, maxiter=1000)
            return x
        x, info = spla.bicgstab(self.A, b, x0=u_star, rtol=1e-13, atol=0.0, maxiter=10000)
        return x


# ---------------------------------------------------------------- pacing
def stim_on(arr, sim, flag, model_time, s1_t, duration, s2_t):
    """simulation/pacing.py:136-170 decision rule -> (S1 on, S2 on)."""
    s1 = s2 = False
    if flag == 0:
        t = model_time
        while t - arr["pacing_start_time"] > arr["pacing_cycle_length"]:
            t = t - arr["pacing_cycle_length"]
        if t >= arr["pacing_start_time"] and t <= arr["pacing_start_time"] + duration:
            s1 = True
    if flag == 5:
        f1, f2 = 0 / sim["time_scale"], 300 / sim["time_scale"]
        if model_time >= f1 and model_time <= f1 + duration:
            s1 = True
        if model_time >= f2 and model_time <= f2 + duration:
            s2 = True
    if flag == 6:
        f1, f2 = 0 / sim["time_scale"], 50 / sim["time_scale"]
        if model_time >= f1 and model_time <= f1 + duration:
            s1 = True
        if model_time >= f2 and model_time <= f2 + duration:
            s2 = True
    elif flag in (1, 2, 3, 4):
        if model_time >= s1_t and model_time <= s1_t + duration:
            s1 = True
    if flag in (1, 2, 3) and model_time >= s2_t and model_time <= s2_t + duration:
        s2 = True
    return s1, s2


def pacing_constants(arr, sim):
    """simulation/pacing.py:102-134 (magnitude 1; duration 5 or 5/time_scale)."""
    s1_t = arr["s1_t"]
    s2_t = s1_t + arr["s1_s2_delta_t"]
    duration = 5 if sim["heart_model_flag"] == 0 else 5 / sim["time_scale"]
    duration = sim.get("pacing_duration_override", duration)
    return s1_t, duration, s2_t


# ---------------------------------------------------------------- time loop
def simulate(n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters, *,
             solver="direct", n_steps=None, lat_threshold=0.13, max_beats=8, probes=None,
             record_samples=True, sample_dtype=np.float64, u0=None, h0=None, step0=0):
    """simulation/simulation_gpu.py:261-409 in FP64.

    Returns a dict with ``u``, ``h`` (samples x N), ``physical_time``,
    ``lat`` (max_beats x N int64, -1 = none), ``beats`` (N), ``probe_u``
    (steps x len(probes)) and the final state ``u_final``, ``h_final``.
    """
    sim, arr = simulation_parameters, arrhythmia_parameters
    P = np.asarray(P_2d, dtype=np.float64)
    nb = np.asarray(geometry_data["neighbor_id_2d"])
    Delta = geometry_data["Delta"]
    node_flag = np.asarray(arr["node_flag"])
    temp_ids = np.nonzero(node_flag == 3)[0]
    perm_ids = np.nonzero(node_flag == 4)[0]
    t_final = float(sim["t_final"])
    dt = sim["dt"]
    flag = sim["arrhythmia_flag"]
    model = sim["heart_model_flag"]
    ts = sim["time_scale"]

    s1_t, duration, s2_t = pacing_constants(arr, sim)
    s1 = arr["s1_pacing_voxel_id"]
    s2 = arr.get("s2_pacing_voxel_id", [])

    u = np.zeros(n_voxel) if u0 is None else np.array(u0, dtype=np.float64)
    if h0 is not None:
        h = np.array(h0, dtype=np.float64)
    else:
        h = np.ones(n_voxel) if model == 0 else np.zeros(n_voxel)
    n_samples = int(np.round(t_final)) if model == 0 else int(np.round(t_final * ts))

    P_perm = P.copy()
    P_perm[perm_ids, 20] = 0.0
    L = build_diffusion_matrix(P_perm, nb, Delta)
    cn = CNSolver(L, dt, solver)
    if len(temp_ids):
        P_tb = P_perm.copy()
        P_tb[temp_ids, 20] = 0.0
        cn_tb = CNSolver(build_diffusion_matrix(P_tb, nb, Delta), dt, solver)
    else:
        cn_tb = cn

    params = tuple(P[:, c] for c in range(15, 20))
    total = int(np.round(t_final / dt)) if n_steps is None else int(n_steps)
    nspm = int(1 / (dt * ts))

    su = np.zeros((n_samples, n_voxel), dtype=sample_dtype) if record_samples else None
    sh = np.zeros((n_samples, n_voxel), dtype=sample_dtype) if record_samples else None
    ptime = np.zeros(n_samples, dtype=np.float64)
    lat = np.full((max_beats, n_voxel), -1, dtype=np.int64)
    beats = np.zeros(n_voxel, dtype=np.int64)
    above = u >= lat_threshold
    probes = np.asarray(probes if probes is not None else [], dtype=np.int64)
    probe_u = np.zeros((total, len(probes)))
    J = np.zeros(n_voxel)
    id_save = 0
    for step in range(step0, step0 + total):
        model_time = step * dt
        active = cn_tb if (model_time >= 0 and model_time < 350 * dt) else cn
        J.fill(0.0)
        on1, on2 = stim_on(arr, sim, flag, model_time, s1_t, duration, s2_t)
        if on1:
            J[s1] = 1
        if on2:
            J[s2] = 1
        u_star, h = rk4_step(model, u, h, params, J, dt)
        u = active(u_star)
        now = u >= lat_threshold
        cross = np.nonzero(now & ~above)[0]
        if len(cross):
            b = beats[cross]
            ok = b < max_beats
            lat[b[ok], cross[ok]] = step
            beats[cross] += 1
        above = now
        if len(probes):
            probe_u[step - step0] = u[probes]
        save = (step % nspm == 0) if nspm > 0 else True
        if save and id_save < n_samples:
            if record_samples:
                su[id_save] = u
                sh[id_save] = h
            ptime[id_save] = model_time * ts
            id_save += 1
    out = dict(physical_time=ptime[:id_save].astype(np.float32), lat=lat, beats=beats,
               probe_u=probe_u, u_final=u, h_final=h, n_steps=total)
    if record_samples:
        out["u"] = su[:id_save]
        out["h"] = sh[:id_save]
    return out


def compute(n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters):
    """Reference-shaped return (sim_u_voxel, sim_h_voxel, physical_time)."""
    r = simulate(n_voxel, P_2d, geometry_data, simulation_parameters, arrhythmia_parameters)
    return r["u"], r["h"], r["physical_time"]
