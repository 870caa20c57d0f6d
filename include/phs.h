/*
 * phs.h -- C ABI of libphs.so, the MI355X (gfx950) simulation hot path of
 * PyHeartSim: RK4 Mitchell-Schaeffer / Aliev-Panfilov reaction,
 * Crank-Nicolson diffusion on the voxel lattice, per-voxel activation times,
 * 1 kHz sampling, and the unipolar electrogram.
 *
 * The reference has no FFI: its boundary is two Python functions,
 *   simulation/simulation_gpu.py:261          compute(n_voxel, P_2d, geometry_data,
 *                                              simulation_parameters, arrhythmia_parameters)
 *   simulation/unipolar_electrogram_gpu.py:71 compute(electrode_xyz, voxel, D0, c_voxel,
 *                                              action_potential, Delta, neighbor_id_2d, batch_size)
 * The Python package pyheartsim_amd.simulation keeps those two signatures and binds this
 * header through ctypes.  Each entry point below names the reference code it replaces.
 *
 * Conventions
 *   - Return 0 on success, a negative PHS_ERR_* code otherwise; phs_last_error() gives the
 *     message (thread-local).  No C++ exception crosses this boundary.
 *   - Pointers documented "host" are read during the call only.  Pointers documented
 *     "device" are caller-owned HBM buffers that must stay valid while the library uses them.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).  Every call that
 *     takes a stream is ordered on it.
 *   - Handles are not thread-safe: use one handle per host thread.
 */
#ifndef PHS_H
#define PHS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHS_OK 0
#define PHS_ERR_ARG (-1)         /* invalid argument */
#define PHS_ERR_HIP (-2)         /* HIP runtime error */
#define PHS_ERR_OOM (-3)         /* device allocation failed */
#define PHS_ERR_UNSUPPORTED (-4) /* input outside the supported physics (e.g. anisotropic D) */
#define PHS_ERR_CAPACITY (-5)    /* arena has no room for the simulation */
#define PHS_ERR_STATE (-6)       /* call not valid in the handle's current state */

/* Voxel flag bits (vflags, one byte per voxel). */
#define PHS_VF_SCAR 0x01       /* node_flag 4: non-conductive for the whole run */
#define PHS_VF_TEMP_BLOCK 0x02 /* node_flag 3: non-conductive while schedule bit 2 is set */
#define PHS_VF_S1 0x04         /* S1 pacing site */
#define PHS_VF_S2 0x08         /* S2 pacing site */

/* Schedule bits (one byte per step). */
#define PHS_SCHED_S1 0x01
#define PHS_SCHED_S2 0x02
#define PHS_SCHED_TEMP_BLOCK 0x04

/* One simulation, all host arrays in the caller's (original) voxel order.
 * Mirrors what simulation/simulation_gpu.py:261-357 reads from its five arguments:
 *   model, params[5]  <- heart_model_flag, P_2d[:, 15:20] (uniform over voxels; params_voxel
 *                        when they vary)
 *   c, c_voxel        <- P_2d[:, 20] (c_voxel NULL when uniform; scar handled by vflags)
 *   dt, delta         <- simulation_parameters['dt'], geometry_data['Delta']
 *   nb                <- geometry_data['neighbor_id_2d'] columns 0-5 (row stride nb_stride)
 *   vflags            <- arrhythmia_parameters node_flag (3, 4) and the S1/S2 site lists
 *   schedule          <- simulation/pacing.py:136-170 and the 350-step temporary block,
 *                        evaluated in FP64 on the host, one byte per step
 *   steps_per_sample, n_samples  <- :286-290, :356, :395-407 (1 kHz sampling) */
typedef struct phs_sim_desc {
  int32_t n_voxel;
  int32_t model;           /* 0 = Mitchell-Schaeffer, 1 = Aliev-Panfilov */
  double params[5];        /* MS: tau_open, tau_close, tau_in, tau_out, v_gate;
                              AP: k, a, epsilon_0, mu1, mu2 */
  double c;                /* conductivity of conducting voxels */
  const double* c_voxel;   /* host [n_voxel] or NULL (uniform c) */
  double dt;
  double delta;
  int64_t n_steps;         /* length of `schedule` */
  int32_t steps_per_sample;/* save after step s when s % steps_per_sample == 0 (0 = every step) */
  int32_t n_samples;       /* capacity of the sample outputs */
  const int32_t* ijk;      /* host [n_voxel][3] integer lattice coordinates (bricked form) */
  const int32_t* nb;       /* host [n_voxel][nb_stride] neighbour ids, columns 0-5 used, -1 = none */
  int32_t nb_stride;
  const uint8_t* vflags;   /* host [n_voxel] PHS_VF_* */
  const uint8_t* schedule; /* host [n_steps] PHS_SCHED_* */
  const float* u0;         /* host [n_voxel] initial u or NULL (rest: 0) */
  const float* h0;         /* host [n_voxel] initial h or NULL (rest: MS 1, AP 0) */
  double lat_threshold;    /* activation threshold for the per-voxel LAT (0.13) */
  const double* params_voxel; /* host [5][n_voxel] per-voxel cell parameters (parameter k of
                              voxel i at params_voxel[k * n_voxel + i], in the order of
                              params) or NULL (params holds the uniform values); replaces the
                              per-voxel reads of P_2d[n, 15..19] at simulation_gpu.py:29-41 */
} phs_sim_desc;

/* Caller-owned device outputs of one simulation. */
typedef struct phs_outputs {
  void* samples_u;         /* device [n_samples][n_voxel] or NULL; float, or double for the
                              plain FP64 form when sample_fp64 = 1 */
  void* samples_h;         /* device, same layout, or NULL */
  int32_t sample_fp64;     /* plain FP64 form only: samples are double */
  int32_t* lat;            /* device [max_beats][n_voxel] step index of each upward crossing,
                              -1 = none (the library initialises it) */
  int32_t max_beats;       /* <= 7 */
  void* probe_u;           /* device [n_steps][n_probes] u after every step, or NULL (float;
                              double for the plain FP64 form when sample_fp64 = 1) */
  const int32_t* probe_ids;/* host [n_probes] voxel ids */
  int32_t n_probes;
  int32_t sample_ring;     /* bricked arena: > 0 = the sample buffers hold sample_ring rows used
                              cyclically (sample s goes to row s % sample_ring) */
} phs_outputs;

int phs_version(void);
const char* phs_last_error(void);
int phs_device_count(int32_t* count);

/* Host only.  Time-step stability of the inexact CN step of the bricked arena: with
 * spec(I - dt/2 L) in [1, lambda_max], a Chebyshev solve of `sweeps` sweeps on [1, lambda_max]
 * (plus `refine` FP64-residual passes of `refine_sweeps`) started from warm start `guess`
 * (0, 1, 2) grows the stiffest modes by *amplification per step (> 1 = unstable; the exact
 * solve gives the CN factor, <= 1).  phs_stable_sweeps gives the fewest such sweeps with
 * amplification <= 1 (*stable, the least fixed count the arena accepts) and the arena's
 * automatic count at fine time steps (*automatic, may be NULL; see phs_auto_sweeps).
 * No counterpart in the reference, whose BiCGSTAB solve runs to a tolerance. */
int phs_stable_sweeps(double lambda_max, int32_t guess, int32_t refine, int32_t refine_sweeps,
                      int32_t* stable, int32_t* automatic);
int phs_step_amplification(double lambda_max, int32_t guess, int32_t sweeps, int32_t refine,
                           int32_t refine_sweeps, double* amplification);
/* Host only.  The arena's automatic sweep count for one simulation (model and params as in
 * phs_sim_desc, dt its step), with s = dt times the cell model's steepest upstroke rate
 * (*resolution, may be NULL): the fewest sweeps m, on a Chebyshev interval [lo, lambda_max]
 * (*lambda_lo, may be NULL), such that the step is stable, every stiff mode (lambda >= 2)
 * shrinks by e within 40 steps (amplification over [2, lambda_max] <= 1 - 1/40), the error
 * bound max|P| over the interval (times the refinement passes' bounds) is <= 1/(4 lambda_max)
 * and <= 0.03 min(1, (0.025 / s)^2) (coarse time steps leave the warm start less to predict),
 * and, for lo > 1, the smoothest mode's error |P(1)| is <= 0.16.  lo > 1 is considered only
 * at fine steps (s <= 0.025) without refinement and for lambda_max <= 10 (the spectra where
 * it was validated); it damps the stiff modes with fewer sweeps.
 * No counterpart in the reference, whose solver runs to a tolerance. */
int phs_auto_sweeps(double lambda_max, int32_t guess, int32_t refine, int32_t refine_sweeps,
                    int32_t model, const double* params, double dt, int32_t* sweeps,
                    double* resolution, double* lambda_lo);

/* Host only (no device work): the Morton brick plan of a voxel set for brick shape
 * brick[3] (bx*by*bz = 64).  perm [n_voxel] receives the compact -> caller order,
 * *n_bricks the brick count, masks (NULL or capacity n_voxel) the per-brick occupancy.
 * When nb is not NULL the face-neighbour table is checked against the lattice: each entry
 * must be the adjacent lattice voxel or -1 (a -1 where that voxel exists cuts the link from
 * that row of the operator, as the reference's table-built operator does); any other entry
 * gives PHS_ERR_UNSUPPORTED (the plain FP64 form takes any table).  Replaces nothing in the
 * reference: it is the voxel reordering that the device forms add. */
int phs_plan_build(int32_t n_voxel, const int32_t* ijk, const int32_t* nb, int32_t nb_stride,
                   const int32_t* brick, int32_t* perm, int32_t* n_bricks, uint64_t* masks);

/* ------------------------------------------------------------------------------------------
 * Plain FP64 form: the reference algorithm (simulation/simulation_gpu.py:358-407 with the
 * FP64 arithmetic of legacy/compute_simulation_cpu.py:144-228) on one GPU, voxels in the
 * caller's order, neighbour-list operator, Chebyshev-accelerated CN solve iterated to the
 * FP64 tolerance `tol`.  Outputs in the caller's voxel order.
 * ------------------------------------------------------------------------------------------ */
typedef struct phs_plain phs_plain;

/* Replaces simulation_gpu.py:296-353 (operator build, A = I - dt/2 L, buffers). */
int phs_plain_create(const phs_sim_desc* desc, const phs_outputs* out, double tol,
                     phs_plain** handle);
/* Replaces one pass of the loop simulation_gpu.py:358-407 for n_steps steps. */
int phs_plain_run(phs_plain* handle, int64_t n_steps, void* stream);
/* Device-to-device copy of the current state (double[n_voxel] each; either may be NULL). */
int phs_plain_state(phs_plain* handle, double* u, double* h, void* stream);
/* Divergence report: the first step whose new transmembrane potential was not finite (NaN or
 * inf) at any voxel, or -1.  Recorded on the device by the step itself; reading it waits for
 * the work queued on `stream`.  The reference has no such report (simulation_gpu.py:250-255
 * only prints non-convergence of the solver). */
int phs_plain_status(phs_plain* handle, int64_t* first_nonfinite_step, void* stream);
/* Chebyshev sweeps per step and the spectral bounds used. */
int phs_plain_info(phs_plain* handle, int32_t* sweeps, double* lambda_min, double* lambda_max);
int phs_plain_destroy(phs_plain* handle);

/* ------------------------------------------------------------------------------------------
 * Bricked FP32 arena: many simulations of uneven size packed into one set of launches.
 * Voxels of each simulation are reordered into Morton-ordered bricks of 64 lattice cells
 * (one wavefront per brick).  Reaction and diffusion run in FP32; the CN right-hand side
 * and the refinement residuals are formed in FP64.  Simulations are added and removed
 * while others run (continuous batching); each advances by its own step counter.
 * ------------------------------------------------------------------------------------------ */
typedef struct phs_arena phs_arena;

typedef struct phs_arena_config {
  int64_t max_voxels;      /* voxel capacity over all resident simulations */
  int64_t max_bricks;      /* brick capacity */
  int32_t max_sims;        /* slot capacity */
  int32_t brick[3];        /* brick shape (bx, by, bz), bx*by*bz = 64 */
  int32_t sweeps;          /* Chebyshev sweeps of the first CN solve: 2..64, or 0 = automatic
                              (per slot the count of phs_stable_sweeps, the arena runs the
                              largest over resident slots); a fixed count that is unstable
                              for a slot's spectrum makes phs_arena_add fail */
  int32_t refine;          /* FP64-residual refinement passes (0..4) */
  int32_t refine_sweeps;   /* Chebyshev sweeps per refinement (>= 2) */
  int32_t xcd_remap;       /* 1 = map consecutive bricks to one XCD */
  int32_t diag;            /* 1 = keep u* and the correction of the last step so that
                              phs_arena_last_residual can recompute the CN residual */
  int32_t bricks_per_wave; /* bricks handled by one wavefront: 1 or 2 */
  int32_t guess;           /* warm start of the CN correction: 0 none, 1 previous step's
                              correction, 2 linear extrapolation of the last two */
} phs_arena_config;

int phs_arena_create(const phs_arena_config* cfg, phs_arena** handle);
/* Replaces simulation_gpu.py:296-353 for one simulation; the plan (Morton bricks) is
 * built on the host.  Fails with PHS_ERR_CAPACITY when the arena is full. */
int phs_arena_add(phs_arena* a, const phs_sim_desc* desc, const phs_outputs* out,
                  int32_t* slot, void* stream);
/* Stream-ordered: steps queued earlier on `stream` still advance the simulation. */
int phs_arena_remove(phs_arena* a, int32_t slot, void* stream);
/* Launches cover the bricks up to the highest resident simulation's, so removals leave holes
 * that are launched for nothing.  When the launched bricks exceed (1 + max_waste) times the
 * resident ones, slides the resident simulations' brick records down in brick order until
 * they are contiguous from brick 0.  With `voxels` != 0 it first slides their voxel data down
 * in voxel order the same way (the state arrays, rebasing the records' voxel indices and the
 * slots' offsets), so that the free voxels form one range: an add that failed with
 * PHS_ERR_CAPACITY although enough voxels were free in total then fits.  Every simulation
 * computes bitwise what it computed before; the copies are ordered on `stream` (a move onto
 * part of its old range goes through a device scratch range, allocated on first need).
 * *moved (may be NULL) receives the number of range moves.  Replaces nothing in the
 * reference, whose simulations do not share a device. */
int phs_arena_compact(phs_arena* a, double max_waste, int32_t voxels, int32_t* moved,
                      void* stream);
/* Advance every resident simulation by n_steps steps (each stops at its own n_steps).
 * Replaces the loop simulation_gpu.py:358-407. */
int phs_arena_step(phs_arena* a, int32_t n_steps, void* stream);
/* Record n_steps steps into a HIP graph once and replay it on later calls with the
 * same n_steps (the launch pattern does not depend on which slots are resident). */
int phs_arena_step_graph(phs_arena* a, int32_t n_steps, void* stream);
/* Capture (without launching) the graphs that phs_arena_step_graph replays for n_steps
 * steps, for both launch parities, so that a later replay of that length does no capture.
 * Host only; the graphs stay valid until the set of resident simulations changes. */
int phs_arena_graph_prepare(phs_arena* a, int32_t n_steps);
/* One eager step (as phs_arena_step with n_steps = 1) whose launches are each bracketed by
 * HIP events recorded on `stream`; phs_arena_step_timing then gives the duration of each
 * launch of that step measured in place, in launch order: *n = launches, and for i <
 * min(*n, cap) ms[i] the duration (ms) and kinds[i] the kernel: 0 right-hand side, 1 first
 * sweep, 2 intermediate sweep, 3 last sweep (activation times, samples, next reaction),
 * 4 refinement residual, 5 refinement sweep.  phs_arena_step_timing waits for the marked step
 * to complete.  No counterpart in the reference (measurement only). */
int phs_arena_step_marked(phs_arena* a, void* stream);
int phs_arena_step_timing(phs_arena* a, int32_t cap, double* ms, int32_t* kinds, int32_t* n);
/* Host-side view of a slot: steps done, voxels, bricks, first compact index. */
int phs_arena_slot_info(phs_arena* a, int32_t slot, int64_t* steps_done, int32_t* n_voxel,
                        int32_t* n_bricks, int64_t* voxel_offset);
/* Compact (brick) order -> caller order map of a slot: perm[i] = caller id of voxel i. */
int phs_arena_perm(phs_arena* a, int32_t slot, int32_t* perm_host);
/* Device-to-device copy of a slot's current state in caller order (float[n_voxel]).  The
 * next step's reaction runs at the end of each step, so between steps u is that reaction's
 * u* (and h its h); after a simulation's last step it is the final state. */
int phs_arena_state(phs_arena* a, int32_t slot, float* u, float* h, void* stream);
/* Divergence report of one slot, as phs_plain_status: the first step whose new u was not
 * finite at any of its voxels, or -1. */
int phs_arena_slot_status(phs_arena* a, int32_t slot, int64_t* first_nonfinite_step,
                          void* stream);
/* Chebyshev sweeps and spectral bounds used for a slot. */
int phs_arena_slot_bounds(phs_arena* a, int32_t slot, double* lambda_min, double* lambda_max);
/* FP64 relative residual ||b - A x|| / ||b|| of the last CN solve of a slot, recomputed on
 * the device from the stored correction (diagnostic; refine > 0 keeps the correction). */
int phs_arena_last_residual(phs_arena* a, int32_t slot, double* rel_residual, void* stream);
/* Sweeps of the first CN solve that the next step runs (fixed, or automatic over slots). */
int phs_arena_sweeps(phs_arena* a, int32_t* sweeps);
/* Average duration (ms, HIP events on `stream`) of the reaction, right-hand-side and
 * Chebyshev-sweep kernels over `reps` launches each: ms[0..2].  Simulation state is not
 * changed (the reaction writes to scratch, the sweep runs in its intermediate mode). */
int phs_arena_profile(phs_arena* a, int32_t reps, double* ms, void* stream);
/* Resident voxels, bricks, voxels of simulations that read the flag byte in the sweeps
 * (scar / temporary block), and bricks covered by one launch. */
int phs_arena_counts(phs_arena* a, int64_t* voxels, int64_t* bricks, int64_t* vf_voxels,
                     int64_t* launch_bricks);
int phs_arena_destroy(phs_arena* a);

/* Row permutation: dst[r][perm[i]] = src[r][i] for r < n_rows, i < n (float). */
int phs_scatter_rows_f32(float* dst, const float* src, const int32_t* perm_dev, int64_t n_rows,
                         int64_t n, void* stream);
int phs_scatter_rows_i32(int32_t* dst, const int32_t* src, const int32_t* perm_dev,
                         int64_t n_rows, int64_t n, void* stream);

/* ------------------------------------------------------------------------------------------
 * Unipolar electrogram (replaces simulation/unipolar_electrogram_gpu.py:71-152 and the
 * kernel :21-69).  The per-voxel dipole sum
 *     egm[t, e] = sum_v c_v / l_ve^3 * ((D_v grad u(v, t)) . l_vec_ve)
 * is linear in u, so the central-difference gradient is folded into a lead field
 * Wt (E x N) once per geometry and the electrogram becomes the contraction
 *     egm^T (E x T) = Wt (E x N) . U (N x T)
 * run on MFMA as BF16 hi/lo split products with FP32 accumulation.
 * ------------------------------------------------------------------------------------------ */
typedef struct phs_egm phs_egm;

/* Builds the lead field on the device.  voxel [n][3], electrodes [n_electrode][3],
 * c_voxel [n], D [9][n] (row-major D11..D33, NULL = identity), nb [n][nb_stride] (host).
 * `perm` (host, may be NULL) maps the sample column order to caller voxel ids: column i of
 * the samples holds voxel perm[i]. */
int phs_egm_create(int32_t n_voxel, int32_t n_electrode, const double* voxel,
                   const double* electrodes, const double* c_voxel, const double* D,
                   const int32_t* nb, int32_t nb_stride, double delta, const int32_t* perm,
                   phs_egm** handle, void* stream);
/* samples: device float [n_time][sample_stride] (first n_voxel columns used).
 * egm: device float [n_time][n_electrode] (reference layout), un-normalised.
 * precision: 0 = BF16 hi/lo split MFMA, 1 = FP32 MFMA (16x16x4), 2 = single BF16 MFMA. */
int phs_egm_run(phs_egm* handle, const float* samples, int64_t n_time, int64_t sample_stride,
                float* egm, int32_t precision, void* stream);
int phs_egm_destroy(phs_egm* handle);
/* phs_egm_run without allocation: no counterpart in the reference; for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  Same arguments, kernels and results as phs_egm_run (bitwise), but the split-K
 * workspace is a library-owned static device array of phs_egm_scratch_bytes bytes (64 MiB)
 * per device, held until the process ends, instead of a hipMalloc'd buffer; a request that
 * needs more (splits x n_time x padded electrodes x 4 bytes) fails with PHS_ERR_ARG before
 * any launch (contract fewer samples per call).  Nothing in the array needs to start at zero:
 * the contraction writes every element the reduction then reads.  Every call on one device
 * shares that array: calls that overlap on one GPU (other streams, other host threads) race. */
int phs_egm_run_noalloc(phs_egm* handle, const float* samples, int64_t n_time,
                        int64_t sample_stride, float* egm, int32_t precision, void* stream);
int phs_egm_scratch_bytes(int64_t* bytes);
/* One job of a streamed contraction: the n_time (1..16) newest sample rows of one simulation,
 * samples device float [n_time][sample_stride] (first n_voxel columns used, in the order the
 * lead field was built for), egm device float [n_time][n_electrode] receiving those rows of
 * the (un-normalised) electrogram. */
typedef struct phs_egm_job {
  phs_egm* lead_field;
  const float* samples;
  int64_t n_time;
  int64_t sample_stride;
  float* egm;
} phs_egm_job;
/* Replaces the per-batch calls of simulation/unipolar_electrogram_gpu.py:124-146 for the rows
 * that a batch of simulations produced since the last call: every job's rows in one launch
 * (per 64 jobs) of BF16 hi/lo split MFMA products with FP32 accumulation, K split into
 * 1024-voxel chunks summed in chunk order, so a job's result does not depend on the other
 * jobs.  workspace: caller-owned device memory of at least phs_egm_stream_workspace bytes,
 * reused by every call (calls on one workspace must be ordered on one stream).  No
 * allocation; stream-ordered. */
int phs_egm_stream_workspace(const phs_egm_job* jobs, int32_t n_jobs, int64_t* bytes);
int phs_egm_stream(const phs_egm_job* jobs, int32_t n_jobs, void* workspace,
                   int64_t workspace_bytes, void* stream);
/* Scale egm [n_time][n_electrode] in place so the mean peak-to-peak is `typical`
 * (simulation/unipolar_electrogram_equation_parts.py:118-124). */
int phs_egm_normalize(float* egm, int64_t n_time, int32_t n_electrode, double typical,
                      void* stream);
/* Electrode activation times (replaces utility/lat_map.py:20-27 compute_electrode_lat):
 * lat[e] = argmax_t -(egm[t+1][e] - egm[t][e]) (first index on ties) minus its minimum over
 * electrodes.  egm device float [n_time][n_electrode] (n_time >= 2); lat device int32
 * [n_electrode].  No allocation; stream-ordered. */
int phs_egm_electrode_lat(const float* egm, int64_t n_time, int32_t n_electrode, int32_t* lat,
                          void* stream);
/* Activation markers and local activation times of sampled action potentials (replaces
 * legacy/calculate_local_activation_time.py:18-66 execute_on_action_potential with the
 * markers of :67-136 find_action_potential_marker).  ap: device float, sample t of signal i at
 * ap[t * stride + i] (stride = n_signal for a row-major [n_time][n_signal] array).  Markers
 * of a signal: for each pair of the sorted unique boundary points
 * of its runs above `threshold`, the first index of the largest forward difference, kept when
 * that difference is > 0 (or NaN).  lat[i] = the first marker in [t_start, t_start +
 * round(CL * cycle_length_fraction)] (CL = mean cycle length over all signals, 180 if none),
 * then lat - min(lat) + 1; NaN where there is none.  stats (device double[5], may be NULL):
 * CL, window start, window end, total markers, total cycle lengths.  any_above (device uint8
 * [n_signal], may be NULL): whether the signal exceeds the threshold.  workspace: device
 * memory of phs_ap_lat_workspace bytes; it keeps the offsets that phs_ap_markers uses.  No
 * allocation; stream-ordered. */
int phs_ap_lat_workspace(int64_t n_signal, int64_t* bytes);
int phs_ap_lat(const float* ap, int64_t n_time, int64_t n_signal, int64_t stride, double threshold,
               double t_start, double cycle_length_fraction, void* workspace, double* lat,
               double* stats, uint8_t* any_above, void* stream);
/* After phs_ap_lat on the same workspace and samples: every signal's markers (device int32
 * [total markers], signal-major, in time order) and cycle lengths (device double [total cycle
 * lengths], the differences of consecutive markers of each signal), either may be NULL. */
int phs_ap_markers(const float* ap, int64_t n_time, int64_t n_signal, int64_t stride,
                   double threshold, void* workspace, int32_t* markers, double* cycle_lengths,
                   void* stream);
/* Conduction velocity at scattered points (replaces the fit loop of
 * utility/conduction_velocity.py:23-70): for each point i, the plane lat = a x + b y + c z + d
 * fitted by least squares (minimum-norm, as np.linalg.lstsq) to the points j with
 * (dx^2 + dy^2) + dz^2 <= radius^2 (i included); g = (a, b, c); velocity[i] = g / |g|^2 and
 * magnitude[i] = 1 / |g| if |g|^2 > 1e-8, else NaN, also NaN with fewer than 4 points.
 * mean (optional, may be NULL) = mean of the non-NaN magnitudes (NaN if none).
 * xyz device double [n_points][3] (mm), lat device double [n_points]; velocity device double
 * [n_points][3], magnitude device double [n_points], mean device double [1].  FP64; no
 * allocation; stream-ordered. */
int phs_conduction_velocity(const double* xyz, const double* lat, int32_t n_points, double radius,
                            double* velocity, double* magnitude, double* mean, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PHS_H */
