// Plain FP64 form of the simulation hot path.
//
// The reference algorithm on one GPU, voxels in the caller's order:
//   reaction  simulation/simulation_gpu.py:71-130 (rk4_reaction_kernel) with the FP64
//             arithmetic and operation order of legacy/compute_simulation_cpu.py:144-209
//   diffusion simulation/simulation_gpu.py:227-259 (CN step).  The reference builds a CSR
//             L (simulation_gpu.py:138-225) and runs GMRES; here L is applied matrix-free
//             from the 6 face-neighbour ids (a missing neighbour contributes u_m - u_n = 0,
//             the zero-flux mirror), and the CN system
//                 (I - dt/2 L) x = u* + dt/2 L u*
//             is solved in correction form x = u* + d, A d = dt L u*, by Chebyshev
//             semi-iteration with Gershgorin bounds spec(A) in [1, 1 + 2 max(alpha deg)],
//             iterated until the error bound reaches `tol`.
// One thread per voxel; one kernel launch per reaction, right-hand side and sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "phs_internal.h"

namespace {

constexpr int kBlock = 256;

struct PlainParams {
  int n;
  int model;
  double p0, p1, p2, p3, p4;
  double dt;
  const double* pv;  // [5][n] per-voxel parameters, or nullptr (uniform p0..p4)
};

// Reaction right-hand side, operation order of legacy/compute_simulation_cpu.py:144-160;
// p[k] is the voxel's P_2d[n, 15 + k].
__device__ __forceinline__ void rhs64(int model, const double (&p)[5], double u, double h,
                                      double J, double& du, double& dh) {
#pragma clang fp contract(off)
  if (model == 0) {
    du = (h * (u * u) * (1.0 - u)) / p[2] - u / p[3] + J;
    dh = (u < p[4]) ? (1.0 - h) / p[0] : -h / p[1];
  } else {
    du = -p[0] * u * (u - p[1]) * (u - 1.0) - u * h + J;
    dh = (p[2] + p[3] * h / (u + p[4])) * (-h - p[0] * u * (u - p[1] - 1.0));
  }
}

// RK4 of legacy/compute_simulation_cpu.py:162-209 (J constant over the step).
__global__ void __launch_bounds__(kBlock) plain_reaction(PlainParams P, double* __restrict__ U,
                                                          double* __restrict__ H,
                                                          const uint8_t* __restrict__ vf,
                                                          uint8_t sched) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= P.n) return;
  const uint8_t f = vf[n];
  const double J = (((sched & PHS_SCHED_S1) && (f & PHS_VF_S1)) ||
                    ((sched & PHS_SCHED_S2) && (f & PHS_VF_S2)))
                       ? 1.0
                       : 0.0;
  const double dt = P.dt;
  const double u = U[n], h = H[n];
  double p[5] = {P.p0, P.p1, P.p2, P.p3, P.p4};
  if (P.pv)
#pragma unroll
    for (int k = 0; k < 5; ++k) p[k] = P.pv[(size_t)k * P.n + n];
  double k1u, k1h, k2u, k2h, k3u, k3h, k4u, k4h;
  rhs64(P.model, p, u, h, J, k1u, k1h);
  rhs64(P.model, p, u + 0.5 * dt * k1u, h + 0.5 * dt * k1h, J, k2u, k2h);
  rhs64(P.model, p, u + 0.5 * dt * k2u, h + 0.5 * dt * k2h, J, k3u, k3h);
  rhs64(P.model, p, u + dt * k3u, h + dt * k3h, J, k4u, k4h);
  U[n] = u + (dt / 6.0) * (k1u + 2.0 * k2u + 2.0 * k3u + k4u);
  H[n] = h + (dt / 6.0) * (k1h + 2.0 * k2h + 2.0 * k3h + k4h);
}

struct Nb6 {
  const int* __restrict__ c[6];  // SoA face neighbours, self id where missing
};

__device__ __forceinline__ double alpha_eff(const double* __restrict__ alpha,
                                            const uint8_t* __restrict__ vf, int n,
                                            uint8_t sched) {
  const double a = alpha[n];
  return ((sched & PHS_SCHED_TEMP_BLOCK) && (vf[n] & PHS_VF_TEMP_BLOCK)) ? 0.0 : a;
}

// r0 = dt L u* = 2 alpha_n sum_m (u*_m - u*_n)
__global__ void __launch_bounds__(kBlock) plain_rhs(int N, Nb6 nb, const double* __restrict__ U,
                                                     const double* __restrict__ alpha,
                                                     const uint8_t* __restrict__ vf, uint8_t sched,
                                                     double* __restrict__ R0) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double un = U[n];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) s += U[nb.c[k][n]] - un;
  R0[n] = 2.0 * alpha_eff(alpha, vf, n, sched) * s;
}

struct SweepOut {
  double* U;          // final: x = u* + d written here (in place over u*)
  const double* H;
  uint8_t* vf;
  int32_t* lat;
  int max_beats;
  double thr;
  long step;
  void* su;           // sample rows (float or double), row index `row` or -1
  void* sh;
  int fp64;
  long row;
  int32_t* nonfinite;  // first step whose new u is not finite (INT32_MAX = none)
};

// One Chebyshev sweep d_{k+1} = a d_k + b d_{k-1} + g (r0 - A d_k).
// cur == nullptr means d_k = r0 * cur_scale (k = 1); prev == nullptr means d_{k-1} = 0,
// prev_from_r0 means d_{k-1} = r0 * cur_scale (k = 2).
template <bool FINAL>
__global__ void __launch_bounds__(kBlock) plain_sweep(int N, Nb6 nb, const double* __restrict__ R0,
                                                       const double* __restrict__ cur,
                                                       const double* prev,
                                                       int prev_from_r0, double r0_scale,
                                                       const double* __restrict__ alpha,
                                                       uint8_t sched, double ca, double cb,
                                                       double cg, double* out,
                                                       SweepOut fo) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double r0 = R0[n];
  const double* src = cur ? cur : R0;
  const double sc = cur ? 1.0 : r0_scale;
  const double dn = src[n] * sc;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) s += src[nb.c[k][n]] * sc - dn;
  const double a = alpha_eff(alpha, fo.vf, n, sched);
  const double Ad = dn - a * s;
  double dp = 0.0;
  if (prev) dp = prev[n];
  else if (prev_from_r0) dp = r0 * r0_scale;
  const double dnew = ca * dn + cb * dp + cg * (r0 - Ad);
  if (!FINAL) {
    out[n] = dnew;
    return;
  }
  const double x = fo.U[n] + dnew;
  fo.U[n] = x;
  if (!__builtin_isfinite(x)) atomicMin(fo.nonfinite, (int32_t)fo.step);
  uint8_t f = fo.vf[n];
  const bool now = x >= fo.thr;
  const bool above = (f & PHS_VF_ABOVE) != 0;
  if (now && !above) {
    int b = (f & PHS_VF_BEAT_MASK) >> PHS_VF_BEAT_SHIFT;
    if (b < fo.max_beats && fo.lat) fo.lat[(long)b * N + n] = (int32_t)fo.step;
    b = b < 7 ? b + 1 : 7;
    fo.vf[n] = (uint8_t)((f & 0x0F) | (b << PHS_VF_BEAT_SHIFT) | PHS_VF_ABOVE);
  } else if (!now && above) {
    fo.vf[n] = (uint8_t)(f & ~PHS_VF_ABOVE);
  }
  if (fo.row >= 0) {
    const long o = fo.row * (long)N + n;
    if (fo.fp64) {
      if (fo.su) ((double*)fo.su)[o] = x;
      if (fo.sh) ((double*)fo.sh)[o] = fo.H[n];
    } else {
      if (fo.su) ((float*)fo.su)[o] = (float)x;
      if (fo.sh) ((float*)fo.sh)[o] = (float)fo.H[n];
    }
  }
}

template <typename T>
__global__ void plain_probe(const double* __restrict__ U, const int32_t* __restrict__ ids, int np,
                            T* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < np) out[j] = (T)U[ids[j]];
}

}  // namespace

struct phs_plain {
  PlainParams P{};
  int N = 0;
  int sweeps = 0;
  double lmin = 1.0, lmax = 1.0;
  std::vector<phs::ChebCoef> coef;
  double *U = nullptr, *H = nullptr, *R0 = nullptr, *DA = nullptr, *DB = nullptr,
         *alpha = nullptr, *pv = nullptr;
  int* nbd = nullptr;
  uint8_t* vf = nullptr;
  int32_t* probe_ids = nullptr;
  int32_t* nonfinite = nullptr;  // device: first step with a non-finite u (INT32_MAX = none)
  std::vector<uint8_t> sched;
  phs_outputs out{};
  int64_t n_steps = 0;
  int64_t step = 0;
  int32_t steps_per_sample = 1;
  int32_t n_samples = 0;
  int64_t id_save = 0;
  double thr = 0.13;
};

static int validate_desc(const phs_sim_desc* d) {
  if (!d) return phs::fail(PHS_ERR_ARG, "desc is NULL");
  if (d->n_voxel <= 0) return phs::fail(PHS_ERR_ARG, "n_voxel must be > 0");
  if (d->model != 0 && d->model != 1) return phs::fail(PHS_ERR_ARG, "model must be 0 or 1");
  if (!d->nb || d->nb_stride < 6) return phs::fail(PHS_ERR_ARG, "nb table with stride >= 6 required");
  if (!(d->dt > 0) || !(d->delta > 0)) return phs::fail(PHS_ERR_ARG, "dt and delta must be > 0");
  if (d->n_steps < 0) return phs::fail(PHS_ERR_ARG, "n_steps must be >= 0");
  if (d->n_steps > 0 && !d->schedule) return phs::fail(PHS_ERR_ARG, "schedule required");
  if (d->steps_per_sample < 0 || d->n_samples < 0) return phs::fail(PHS_ERR_ARG, "bad sampling");
  for (int32_t i = 0; i < d->n_voxel; ++i)
    for (int k = 0; k < 6; ++k) {
      const int32_t m = d->nb[(int64_t)i * d->nb_stride + k];
      if (m < -1 || m >= d->n_voxel) return phs::fail(PHS_ERR_ARG, "neighbour id out of range");
    }
  return PHS_OK;
}

extern "C" {

int phs_plain_create(const phs_sim_desc* d, const phs_outputs* out, double tol, phs_plain** handle) {
  int rc = validate_desc(d);
  if (rc) return rc;
  if (!handle) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (!(tol > 0)) return phs::fail(PHS_ERR_ARG, "tol must be > 0");
  if (out && out->max_beats > 7) return phs::fail(PHS_ERR_ARG, "max_beats must be <= 7");
  phs_plain* p = new phs_plain();
  const int N = d->n_voxel;
  p->N = N;
  p->P = PlainParams{N, d->model, d->params[0], d->params[1], d->params[2], d->params[3],
                     d->params[4], d->dt, nullptr};
  p->n_steps = d->n_steps;
  p->steps_per_sample = d->steps_per_sample;
  p->n_samples = d->n_samples;
  p->thr = d->lat_threshold;
  if (out) p->out = *out;
  p->sched.assign(d->schedule, d->schedule + d->n_steps);

  // Host-side operator metadata: SoA neighbours (self where missing), alpha with scar zeroed.
  std::vector<int> nbh((size_t)6 * N);
  std::vector<double> alpha(N);
  std::vector<uint8_t> vf(N);
  double amax_deg = 0.0;
  for (int i = 0; i < N; ++i) {
    int deg = 0;
    for (int k = 0; k < 6; ++k) {
      const int m = d->nb[(int64_t)i * d->nb_stride + k];
      nbh[(size_t)k * N + i] = m >= 0 ? m : i;
      deg += m >= 0;
    }
    const uint8_t f = d->vflags ? d->vflags[i] : 0;
    vf[i] = f & 0x0F;
    const double c = d->c_voxel ? d->c_voxel[i] : d->c;
    alpha[i] = (f & PHS_VF_SCAR) ? 0.0 : d->dt * c / (2.0 * d->delta * d->delta);
    if (alpha[i] < 0) {
      delete p;
      return phs::fail(PHS_ERR_ARG, "negative conductivity");
    }
    amax_deg = std::max(amax_deg, alpha[i] * deg);
  }
  p->lmin = 1.0;
  p->lmax = 1.0 + 2.0 * amax_deg;
  p->sweeps = std::max(2, phs::chebyshev_sweeps(p->lmin, p->lmax, tol));
  if (p->lmax <= 1.0) p->sweeps = 2;
  p->coef.resize(p->sweeps);
  if (p->lmax > 1.0) phs::chebyshev_coefficients(p->lmin, p->lmax, p->sweeps, p->coef.data());

  auto cleanup = [&](int code) {
    phs_plain_destroy(p);
    return code;
  };
  hipError_t e = hipSuccess;
  const size_t nd = sizeof(double) * (size_t)N;
  if ((e = hipMalloc(&p->U, nd)) || (e = hipMalloc(&p->H, nd)) || (e = hipMalloc(&p->R0, nd)) ||
      (e = hipMalloc(&p->DA, nd)) || (e = hipMalloc(&p->DB, nd)) ||
      (e = hipMalloc(&p->alpha, nd)) || (e = hipMalloc(&p->nbd, sizeof(int) * 6 * (size_t)N)) ||
      (e = hipMalloc(&p->vf, (size_t)N)) ||
      (d->params_voxel && (e = hipMalloc(&p->pv, 5 * nd))) ||
      (e = hipMalloc(&p->nonfinite, sizeof(int32_t))))
    return cleanup(phs::fail(PHS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e)));
  const int32_t none = INT32_MAX;
  if ((e = hipMemcpy(p->nonfinite, &none, sizeof none, hipMemcpyHostToDevice)))
    return cleanup(phs::fail(PHS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
  if (d->params_voxel) {
    if ((e = hipMemcpy(p->pv, d->params_voxel, 5 * nd, hipMemcpyHostToDevice)))
      return cleanup(phs::fail(PHS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
    p->P.pv = p->pv;
  }
  std::vector<double> u0(N), h0(N);
  for (int i = 0; i < N; ++i) {
    u0[i] = d->u0 ? (double)d->u0[i] : 0.0;
    h0[i] = d->h0 ? (double)d->h0[i] : (d->model == 0 ? 1.0 : 0.0);
    if (u0[i] >= d->lat_threshold) vf[i] |= PHS_VF_ABOVE;
  }
  if ((e = hipMemcpy(p->U, u0.data(), nd, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(p->H, h0.data(), nd, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(p->alpha, alpha.data(), nd, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(p->nbd, nbh.data(), sizeof(int) * 6 * (size_t)N, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(p->vf, vf.data(), (size_t)N, hipMemcpyHostToDevice)))
    return cleanup(phs::fail(PHS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
  if (p->out.lat && p->out.max_beats > 0) {
    if ((e = hipMemset(p->out.lat, 0xFF, sizeof(int32_t) * (size_t)p->out.max_beats * N)))
      return cleanup(phs::fail(PHS_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e)));
  }
  if (p->out.n_probes > 0) {
    if (!p->out.probe_ids || !p->out.probe_u) return cleanup(phs::fail(PHS_ERR_ARG, "probe arrays"));
    for (int j = 0; j < p->out.n_probes; ++j)
      if (p->out.probe_ids[j] < 0 || p->out.probe_ids[j] >= N)
        return cleanup(phs::fail(PHS_ERR_ARG, "probe id out of range"));
    if ((e = hipMalloc(&p->probe_ids, sizeof(int32_t) * p->out.n_probes)) ||
        (e = hipMemcpy(p->probe_ids, p->out.probe_ids, sizeof(int32_t) * p->out.n_probes,
                       hipMemcpyHostToDevice)))
      return cleanup(phs::fail(PHS_ERR_HIP, std::string("probe ids: ") + hipGetErrorString(e)));
  }
  *handle = p;
  return PHS_OK;
}

int phs_plain_run(phs_plain* p, int64_t n_steps, void* stream) {
  if (!p) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (n_steps < 0) return phs::fail(PHS_ERR_ARG, "n_steps < 0");
  hipStream_t s = (hipStream_t)stream;
  const int N = p->N;
  const int grid = (N + kBlock - 1) / kBlock;
  Nb6 nb;
  for (int k = 0; k < 6; ++k) nb.c[k] = p->nbd + (size_t)k * N;
  const double inv_theta = p->coef[0].g;
  const int64_t end = std::min<int64_t>(p->n_steps, p->step + n_steps);
  for (; p->step < end; ++p->step) {
    const uint8_t sc = p->sched[p->step];
    plain_reaction<<<grid, kBlock, 0, s>>>(p->P, p->U, p->H, p->vf, sc);
    plain_rhs<<<grid, kBlock, 0, s>>>(N, nb, p->U, p->alpha, p->vf, sc, p->R0);
    SweepOut fo{};
    fo.U = p->U;
    fo.H = p->H;
    fo.vf = p->vf;
    fo.lat = p->out.lat;
    fo.max_beats = p->out.max_beats;
    fo.thr = p->thr;
    fo.step = (long)p->step;
    fo.su = p->out.samples_u;
    fo.sh = p->out.samples_h;
    fo.fp64 = p->out.sample_fp64;
    fo.nonfinite = p->nonfinite;
    const bool save = p->steps_per_sample > 0 ? (p->step % p->steps_per_sample == 0) : true;
    fo.row = (save && p->id_save < p->n_samples) ? (long)p->id_save : -1;
    if (fo.row >= 0) p->id_save++;
    if (p->lmax <= 1.0) {
      // No conduction anywhere: x = u*.  A single final pass with zero coefficients.
      plain_sweep<true><<<grid, kBlock, 0, s>>>(N, nb, p->R0, nullptr, nullptr, 0, 0.0, p->alpha,
                                                sc, 0.0, 0.0, 0.0, nullptr, fo);
    } else {
      double* bufs[2] = {p->DA, p->DB};
      for (int k = 1; k < p->sweeps; ++k) {
        const double* cur = k == 1 ? nullptr : bufs[k % 2];
        const double* prev = k <= 2 ? nullptr : bufs[(k - 1) % 2];
        const int prev_r0 = k == 2;
        double* outb = bufs[(k + 1) % 2];
        const phs::ChebCoef& c = p->coef[k];
        if (k == p->sweeps - 1)
          plain_sweep<true><<<grid, kBlock, 0, s>>>(N, nb, p->R0, cur, prev, prev_r0, inv_theta,
                                                    p->alpha, sc, c.a, c.b, c.g, outb, fo);
        else
          plain_sweep<false><<<grid, kBlock, 0, s>>>(N, nb, p->R0, cur, prev, prev_r0, inv_theta,
                                                     p->alpha, sc, c.a, c.b, c.g, outb, fo);
      }
    }
    if (p->out.n_probes > 0) {
      const size_t off = (size_t)p->step * p->out.n_probes;
      if (p->out.sample_fp64)
        plain_probe<double><<<(p->out.n_probes + 63) / 64, 64, 0, s>>>(
            p->U, p->probe_ids, p->out.n_probes, (double*)p->out.probe_u + off);
      else
        plain_probe<float><<<(p->out.n_probes + 63) / 64, 64, 0, s>>>(
            p->U, p->probe_ids, p->out.n_probes, (float*)p->out.probe_u + off);
    }
  }
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

int phs_plain_state(phs_plain* p, double* u, double* h, void* stream) {
  if (!p) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  hipStream_t s = (hipStream_t)stream;
  if (u) PHS_HIP_CHECK(hipMemcpyAsync(u, p->U, sizeof(double) * p->N, hipMemcpyDeviceToDevice, s));
  if (h) PHS_HIP_CHECK(hipMemcpyAsync(h, p->H, sizeof(double) * p->N, hipMemcpyDeviceToDevice, s));
  return PHS_OK;
}

int phs_plain_info(phs_plain* p, int32_t* sweeps, double* lmin, double* lmax) {
  if (!p) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (sweeps) *sweeps = p->sweeps;
  if (lmin) *lmin = p->lmin;
  if (lmax) *lmax = p->lmax;
  return PHS_OK;
}

int phs_plain_status(phs_plain* p, int64_t* first_nonfinite_step, void* stream) {
  if (!p || !first_nonfinite_step) return phs::fail(PHS_ERR_ARG, "NULL argument");
  int32_t v = INT32_MAX;
  hipStream_t s = (hipStream_t)stream;
  PHS_HIP_CHECK(hipMemcpyAsync(&v, p->nonfinite, sizeof v, hipMemcpyDeviceToHost, s));
  PHS_HIP_CHECK(hipStreamSynchronize(s));
  *first_nonfinite_step = v == INT32_MAX ? -1 : (int64_t)v;
  return PHS_OK;
}

int phs_plain_destroy(phs_plain* p) {
  if (!p) return PHS_OK;
  hipFree(p->U);
  hipFree(p->H);
  hipFree(p->R0);
  hipFree(p->DA);
  hipFree(p->DB);
  hipFree(p->alpha);
  hipFree(p->pv);
  hipFree(p->nonfinite);
  hipFree(p->nbd);
  hipFree(p->vf);
  hipFree(p->probe_ids);
  delete p;
  return PHS_OK;
}

}  // extern "C"
