/*
 * FP64 CPU restatement of the reference's CPU simulation path, OpenMP over host cores.
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may call it, and only as the checker / the reported CPU baseline.
 *
 * Follows legacy/compute_simulation_cpu.py:
 *   reaction   :144-209  RK4, same operation order (numba prange passes -> omp for)
 *   operator   :21-111   CSR L with L[n,n] = c_n/(4 Delta^2) * -(sum_k P[n,k]),
 *                        L[n,m_k] = c_n/(4 Delta^2) * P[n,k] for present face neighbours
 *                        (P = 4 delta_k for D = I; the anisotropy terms are zero)
 *   CN step    :113-142  A = I - dt/2 L formed as its own CSR matrix every step (:131-133),
 *                        b = u* + dt/2 L u* (:134), BiCGSTAB on A from x0 = u* (:137)
 *                        stopping at ||r|| < max(rtol ||b||, atol) (scipy's rule; the
 *                        reference passes atol = 1e-6 and keeps rtol = 1e-5), maxiter 1000
 * with the pacing schedule, scar / temporary block and sampling of
 * simulation/simulation_gpu.py:261-409 supplied as per-step bytes by the caller.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t n;
  int32_t model;
  double p[5];
  const double* c;        /* [n] conductivity, scar already zeroed */
  const uint8_t* tb;      /* [n] 1 = temporary block voxel, or NULL */
  const int32_t* nb;      /* [n][6] face neighbours, -1 = none */
  const uint8_t* s1;      /* [n] S1 site */
  const uint8_t* s2;      /* [n] S2 site */
  double dt, delta;
  double rtol, atol;
  int32_t maxiter;
} cpu_desc;

typedef struct {
  int32_t n;
  int32_t* rowptr;
  int32_t* col;
  double* val;
} csr;

static void rhs(int model, const double* p, double u, double h, double J, double* du, double* dh) {
  if (model == 0) {
    *du = (h * (u * u) * (1 - u)) / p[2] - u / p[3] + J;
    *dh = (u < p[4]) ? (1 - h) / p[0] : -h / p[1];
  } else {
    *du = -p[0] * u * (u - p[1]) * (u - 1) - u * h + J;
    *dh = (p[2] + p[3] * h / (u + p[4])) * (-h - p[0] * u * (u - p[1] - 1));
  }
}

static void build_L(const cpu_desc* d, int temp_block_on, csr* L) {
  const int n = d->n;
  L->n = n;
  L->rowptr = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  L->col = (int32_t*)malloc(sizeof(int32_t) * 7 * (size_t)n);
  L->val = (double*)malloc(sizeof(double) * 7 * (size_t)n);
  int32_t k = 0;
  for (int i = 0; i < n; ++i) {
    L->rowptr[i] = k;
    double c = d->c[i];
    if (temp_block_on && d->tb && d->tb[i]) c = 0.0;
    const double Dc = c / (4 * d->delta * d->delta);
    double psum = 0.0;
    for (int q = 0; q < 6; ++q) psum += d->nb[6 * (size_t)i + q] >= 0 ? 4.0 : 0.0;
    L->col[k] = i;
    L->val[k++] = Dc * -psum;
    for (int q = 0; q < 6; ++q) {
      const int m = d->nb[6 * (size_t)i + q];
      if (m >= 0) {
        L->col[k] = m;
        L->val[k++] = Dc * 4.0;
      }
    }
  }
  L->rowptr[n] = k;
}

static void free_L(csr* L) {
  free(L->rowptr);
  free(L->col);
  free(L->val);
}

/* y = x - s * (L x)  (s = -dt/2 gives b = u* + dt/2 L u*) */
static void apply(const csr* L, double s, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < L->n; ++i) {
    double acc = 0.0;
    for (int k = L->rowptr[i]; k < L->rowptr[i + 1]; ++k) acc += L->val[k] * x[L->col[k]];
    y[i] = x[i] - s * acc;
  }
}

/* y = M x */
static void spmv(const csr* M, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M->n; ++i) {
    double acc = 0.0;
    for (int k = M->rowptr[i]; k < M->rowptr[i + 1]; ++k) acc += M->val[k] * x[M->col[k]];
    y[i] = acc;
  }
}

/* A = I - hdt L as a new CSR matrix with L's pattern (build_L stores the diagonal first in
 * every row), in freshly allocated arrays, as scipy forms I - dt/2 L every step. */
static void form_A(const csr* L, double hdt, csr* A) {
  const int n = L->n;
  const int32_t nnz = L->rowptr[n];
  A->n = n;
  A->rowptr = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  A->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
  A->val = (double*)malloc(sizeof(double) * (size_t)nnz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    A->rowptr[i] = L->rowptr[i];
    for (int k = L->rowptr[i]; k < L->rowptr[i + 1]; ++k) {
      A->col[k] = L->col[k];
      A->val[k] = (L->col[k] == i ? 1.0 : 0.0) - hdt * L->val[k];
    }
  }
  A->rowptr[n] = nnz;
}

static double dot(const double* a, const double* b, int n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Unpreconditioned BiCGSTAB on A (the algorithm of scipy.sparse.linalg.bicgstab with M = I). */
static int bicgstab(const csr* A, const double* b, double* x, double rtol, double atol,
                    int maxiter, double* work) {
  const int n = A->n;
  double *r = work, *rt = work + n, *p = work + 2 * (size_t)n, *v = work + 3 * (size_t)n,
         *s = work + 4 * (size_t)n, *t = work + 5 * (size_t)n;
  const double bn = sqrt(dot(b, b, n));
  if (bn == 0.0) {  /* scipy returns the exact solution x = 0 for b = 0 */
    memset(x, 0, sizeof(double) * n);
    return 0;
  }
  const double tol = fmax(atol, rtol * bn);
  spmv(A, x, v);
#pragma omp parallel for
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - v[i];
    rt[i] = r[i];
  }
  if (sqrt(dot(r, r, n)) < tol) return 0;
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  for (int it = 0; it < maxiter; ++it) {
    const double rho = dot(rt, r, n);
    if (rho == 0.0) return -1;
    if (it == 0) {
      memcpy(p, r, sizeof(double) * n);
    } else {
      const double beta = (rho / rho_prev) * (alpha / omega);
#pragma omp parallel for
      for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    spmv(A, p, v);
    alpha = rho / dot(rt, v, n);
#pragma omp parallel for
    for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    if (sqrt(dot(s, s, n)) < tol) {
#pragma omp parallel for
      for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
      return 0;
    }
    spmv(A, s, t);
    omega = dot(t, s, n) / dot(t, t, n);
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i] + omega * s[i];
      r[i] = s[i] - omega * t[i];
    }
    if (sqrt(dot(r, r, n)) < tol) return 0;
    rho_prev = rho;
  }
  return 1;
}

/* Advance (u, h) by n_steps steps; sched[i] bits: 1 S1 on, 2 S2 on, 4 temporary block.
 * rebuild != 0 forms A = I - dt/2 L anew every step (the reference's cost); otherwise once
 * per call for each of the two operators.  Returns the number of CN solves that hit maxiter. */
int cpu_simulate(const cpu_desc* d, const uint8_t* sched, int64_t n_steps, double* u, double* h,
                 int threads, int rebuild) {
  if (threads > 0) omp_set_num_threads(threads);
  const int n = d->n;
  csr L0, L1, A0, A1;
  build_L(d, 0, &L0);
  build_L(d, 1, &L1);
  if (!rebuild) {
    form_A(&L0, d->dt / 2.0, &A0);
    form_A(&L1, d->dt / 2.0, &A1);
  }
  double* us = (double*)malloc(sizeof(double) * n);
  double* b = (double*)malloc(sizeof(double) * n);
  double* work = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  int fails = 0;
  const double dt = d->dt;
  for (int64_t st = 0; st < n_steps; ++st) {
    const uint8_t sc = sched[st];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const double J = (((sc & 1) && d->s1[i]) || ((sc & 2) && d->s2[i])) ? 1.0 : 0.0;
      double k1u, k1h, k2u, k2h, k3u, k3h, k4u, k4h;
      rhs(d->model, d->p, u[i], h[i], J, &k1u, &k1h);
      rhs(d->model, d->p, u[i] + 0.5 * dt * k1u, h[i] + 0.5 * dt * k1h, J, &k2u, &k2h);
      rhs(d->model, d->p, u[i] + 0.5 * dt * k2u, h[i] + 0.5 * dt * k2h, J, &k3u, &k3h);
      rhs(d->model, d->p, u[i] + dt * k3u, h[i] + dt * k3h, J, &k4u, &k4h);
      us[i] = u[i] + (dt / 6.0) * (k1u + 2 * k2u + 2 * k3u + k4u);
      h[i] = h[i] + (dt / 6.0) * (k1h + 2 * k2h + 2 * k3h + k4h);
    }
    const csr* L = (sc & 4) ? &L1 : &L0;
    csr Astep;
    const csr* A = (sc & 4) ? &A1 : &A0;
    if (rebuild) {
      form_A(L, dt / 2.0, &Astep);
      A = &Astep;
    }
    apply(L, -dt / 2.0, us, b);
    memcpy(u, us, sizeof(double) * n);
    if (bicgstab(A, b, u, d->rtol, d->atol, d->maxiter, work)) ++fails;
    if (rebuild) free_L(&Astep);
  }
  free(us);
  free(b);
  free(work);
  free_L(&L0);
  free_L(&L1);
  if (!rebuild) {
    free_L(&A0);
    free_L(&A1);
  }
  return fails;
}
