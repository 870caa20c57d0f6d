// Bricked FP32 arena: the chosen form of the simulation hot path.
//
// Many simulations of uneven voxel counts share one set of device arrays; every launch
// covers all resident bricks and each wavefront finds its simulation through the brick's
// slot id.  Per step (simulation/simulation_gpu.py:358-407):
//   reaction   RK4 in FP32, parameters as wave-uniform scalars, J from the per-step
//              schedule byte and the voxel's site bits (replaces rk4_reaction_kernel
//              simulation_gpu.py:71-130 and the per-step J_stim upload :373-377)
//   rhs        r0 = dt L u* = 2 alpha sum_m (u*_m - u*_n), formed in FP64 from FP32 u*
//   sweeps     Chebyshev semi-iteration on A d = r0, A = I - dt/2 L, FP32; the last sweep
//              writes x = u* + d, the activation-time bookkeeping and the 1 kHz samples
//              (replaces build_diffusion_matrix_gpu :138-225 and the GMRES CN solve
//              :227-259, :388, and the sampling copies :391-407)
//   refine     optional FP64-residual refinement: r = r0 - A d in FP64, Chebyshev on the
//              correction in FP32, accumulation in FP64
//   advance    per-simulation step counters (device side, so steps can be graph-replayed)
// A brick is 64 lattice cells handled by one wavefront; in-brick neighbour values move
// between lanes with ds_bpermute (__shfl), cross-brick face neighbours are gathered from
// the neighbour brick's compact range.  A missing neighbour contributes u_m - u_n = 0
// (zero-flux mirror).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "phs_internal.h"
#include "plan.h"

#include "arena_kernels.h"

using namespace arena;

namespace {

// ------------------------------------------------------------------ host side
class RangeAlloc {
 public:
  void init(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int64_t alloc(int64_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t at = it->first;
        const int64_t rem = it->second - n;
        free_.erase(it);
        if (rem > 0) free_[at + n] = rem;
        return at;
      }
    }
    return -1;
  }
  void release(int64_t at, int64_t n) {
    auto it = free_.emplace(at, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

struct SlotHost {
  bool used = false;
  int64_t voxel_off = 0, brick_off = 0;
  int32_t n_voxel = 0, n_bricks = 0;
  int64_t n_steps = 0, steps_done = 0;
  std::vector<int32_t> perm;
  int32_t* perm_dev = nullptr;
  uint8_t* sched_dev = nullptr;
  double lmin = 1.0, lmax = 1.0;
  int32_t need_sweeps = 2;  // sweeps this slot asks of the arena (automatic mode)
  bool needs_vf = false;    // scar or temporary-block voxels
  bool needs_tb = false;    // temporary-block voxels (the lean kernels' flag test)
  bool hetero = false;      // per-voxel alpha
  bool hparams = false;     // per-voxel cell parameters
  void* staging = nullptr;  // pinned upload buffer, freed after its copies complete
  hipEvent_t upload_done = nullptr;
};

}  // namespace

struct DeferredFree {
  hipEvent_t done = nullptr;
  void* host = nullptr;
  void* dev[2] = {nullptr, nullptr};
};

struct phs_arena {
  phs_arena_config cfg{};
  std::vector<DeferredFree> graveyard;
  BrickRec* bricks = nullptr;
  SlotRec* slots = nullptr;
  float *U = nullptr, *H = nullptr, *R0 = nullptr, *DA = nullptr, *DB = nullptr,
        *USTAR = nullptr, *ALPHA = nullptr, *D1 = nullptr, *D2 = nullptr;
  double* DACC = nullptr;
  uint8_t* VF = nullptr;
  double* diag_acc = nullptr;
  float* PV[5] = {};  // per-voxel cell parameters, allocated by the first add that has them
  int32_t* NONFINITE = nullptr;  // [max_sims] first step with a non-finite u per slot
  std::vector<SlotHost> sh;
  RangeAlloc vox, brk;
  int64_t brick_high = 0;
  void* move_tmp = nullptr;  // scratch for overlapping brick-record moves (phs_arena_compact)
  size_t move_bytes = 0;
  hipStream_t cap_stream = nullptr;
  // One captured sequence of steps per launch parity at its first step (the kernels' slot
  // copy alternates with the step), so that odd step counts replay without recapture too.
  struct Graph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int32_t steps = -1;
    int64_t grid = -1;
    int64_t high = -1;  // brick_high captured in the kernel arguments
    int32_t sweeps = -1;
    int32_t flav = -1;
    const void* pv = nullptr;  // PV captured in the kernel arguments
  } g[2];
  int32_t par = 0;         // launch parity of the next step (flips with every launched step)
  uint32_t walk = 0;       // launches so far: odd ones walk the bricks backwards
  int32_t sweeps_eff = 2;  // sweeps of the first CN solve in the launched sequence
  int32_t flav = 0;        // kernel flavour (FLAV_*) the resident simulations allow
  // Launch timing of one eager step: while `marking`, every launch of the step is followed
  // by an event record (event 0 before the first launch), and kinds[i] labels launch i.
  static constexpr int kMaxMarks = 2 + kMaxSweeps + 4 * (kMaxSweeps + 1);
  hipEvent_t marks[kMaxMarks] = {};
  int32_t kinds[kMaxMarks] = {};
  int32_t n_marks = 0;     // events recorded by the marked step (launches + 1)
  bool marking = false;
  bool marks_valid = false;  // a marked step has been launched
};

namespace {

const int32_t kNoStep = INT32_MAX;  // "no non-finite step yet", uploaded at each add

enum { KIND_RHS = 0, KIND_FIRST = 1, KIND_MID = 2, KIND_LAST = 3, KIND_RESIDUAL = 4,
       KIND_REFINE = 5 };

// After a launch of kind `kind` in a marked step: record the event that closes it (event 0
// is recorded before the step's first launch).
void mark(phs_arena* a, hipStream_t s, int kind) {
  if (!a->marking || a->n_marks < 1 || a->n_marks >= phs_arena::kMaxMarks) return;
  a->kinds[a->n_marks - 1] = kind;
  hipEventRecord(a->marks[a->n_marks], s);
  ++a->n_marks;
}

int nb_per_wave(const phs_arena* a);

// Record prefetch distance (bricks ahead in a wavefront's walk): one round of resident
// wavefronts per XCD (32 CUs x 32 waves), measured flat from 512 to 2048 bricks.  Arenas of
// fewer than 16 rounds' bricks stay in L2 and the memory-side cache anyway, so they skip it.
int prefetch_distance(const phs_arena* a) {
  const int dist = 1024 * nb_per_wave(a);
  return a->brick_high >= 16 * (int64_t)dist ? dist : 0;
}

ArenaDev dev_view(const phs_arena* a) {
  ArenaDev A;
  A.bricks = a->bricks;
  A.slots = a->slots;
  A.U = a->U;
  A.H = a->H;
  A.UO = a->U;
  A.HO = a->H;
  A.R0 = a->R0;
  A.DACC = a->DACC;
  A.USTAR = a->USTAR;
  A.VF = a->VF;
  A.ALPHA = a->ALPHA;
  A.D1 = a->D1;
  A.D2 = a->D2;
  for (int k = 0; k < 5; ++k) A.PV[k] = a->PV[k];
  A.NONFINITE = a->NONFINITE;
  A.n_bricks = (int32_t)a->brick_high;
  A.flags = (a->cfg.xcd_remap ? 1 : 0) | (a->par << 1) | (prefetch_distance(a) << 8);
  A.diag = a->cfg.diag ? 1 : 0;
  A.max_sims = a->cfg.max_sims;
  return A;
}

int nb_per_wave(const phs_arena* a) { return a->cfg.bricks_per_wave == 2 ? 2 : 1; }

int grid_for_nb1(const phs_arena* a) {
  int g = (int)((a->brick_high + kWaves - 1) / kWaves);
  if (g < 1) g = 1;
  if (a->cfg.xcd_remap) g = (g + 7) / 8 * 8;
  return g;
}

int grid_for(const phs_arena* a) {
  const int per_block = kWaves * nb_per_wave(a);
  int g = (int)((a->brick_high + per_block - 1) / per_block);
  if (g < 1) g = 1;
  if (a->cfg.xcd_remap) g = (g + 7) / 8 * 8;
  return g;
}

// Launches of one time step for brick shape (BX, BY, BZ) and NB bricks per wavefront.
// The lean kernel flavours (one brick per wavefront only) serve arenas whose simulations need
// neither per-voxel alpha nor per-voxel cell parameters, with or without the temporary-block
// test and the identity sweeps of simulations whose own sweep count is below the arena's.
template <int BX, int BY, int BZ, int NB, int GUESS>
struct Step {
  template <int MODE, int FLAV>
  static void sweep_m(int grid, hipStream_t s, const ArenaDev& A, int k, int K, const float* cur,
                      const float* prev, int prev_kind, float* out) {
    if (k == 1)
      k_sweep<BX, BY, BZ, MODE, NB, GUESS, true, FLAV><<<grid, kBlock, 0, s>>>(A, k, K, cur, prev,
                                                                               prev_kind, out);
    else
      k_sweep<BX, BY, BZ, MODE, NB, GUESS, false, FLAV><<<grid, kBlock, 0, s>>>(A, k, K, cur, prev,
                                                                                prev_kind, out);
  }
  template <int FLAV>
  static void sweep(int mode, int grid, hipStream_t s, const ArenaDev& A, int k, int K,
                    const float* cur, const float* prev, int prev_kind, float* out) {
    switch (mode) {
      case MODE_MID: sweep_m<MODE_MID, FLAV>(grid, s, A, k, K, cur, prev, prev_kind, out); break;
      case MODE_FINAL_X: sweep_m<MODE_FINAL_X, FLAV>(grid, s, A, k, K, cur, prev, prev_kind, out); break;
      case MODE_DACC_SET: sweep_m<MODE_DACC_SET, FLAV>(grid, s, A, k, K, cur, prev, prev_kind, out); break;
      case MODE_DACC_ADD: sweep_m<MODE_DACC_ADD, FLAV>(grid, s, A, k, K, cur, prev, prev_kind, out); break;
      default: sweep_m<MODE_DACC_ADD_X, FLAV>(grid, s, A, k, K, cur, prev, prev_kind, out);
    }
  }

  static void run(phs_arena* a, hipStream_t s) {
    if constexpr (NB == 1) {
      switch (a->flav) {
        case FLAV_LEAN: return run_t<FLAV_LEAN>(a, s);
        case FLAV_LEAN_FLAGS: return run_t<FLAV_LEAN_FLAGS>(a, s);
        case FLAV_LEAN_IDENT: return run_t<FLAV_LEAN_IDENT>(a, s);
        case FLAV_LEAN_FLAGS_IDENT: return run_t<FLAV_LEAN_FLAGS_IDENT>(a, s);
        default: break;
      }
    }
    run_t<FLAV_GENERAL>(a, s);
  }

  template <int FLAV>
  static void run_t(phs_arena* a, hipStream_t s) {
    const ArenaDev A0 = dev_view(a);
    const int grid = grid_for(a);
    const int grid1 = grid_for_nb1(a);
    const phs_arena_config& c = a->cfg;
    // Launches alternate the brick walk (forward, backward, ...) across steps, so that each
    // starts on the bricks the previous one finished with.  The walk changes no result: every
    // launch computes each brick from the previous launch's buffers alone.
    ArenaDev A = A0;
    auto next = [&]() {
      A = A0;
      if (a->walk++ & 1) A.flags |= 4;
    };
    next();
    // the step's reaction already ran in the previous step's last sweep (or at add)
    k_rhs<BX, BY, BZ, NB, GUESS, FLAV><<<grid, kBlock, 0, s>>>(A);
    mark(a, s, KIND_RHS);
    float* bufs[2] = {a->DA, a->DB};
    // One Chebyshev solve of m sweeps on the right-hand side held in R0; per_slot: each
    // simulation runs its own count (<= m) of the first solve.
    auto solve = [&](int m, int last_mode, bool per_slot) {
      const int K = per_slot ? m : 0;
      for (int k = 1; k < m; ++k) {
        const float* cur = k == 1 ? nullptr : bufs[k % 2];
        const float* prev = k <= 2 ? nullptr : bufs[(k - 1) % 2];
        const int prev_kind = k == 1 ? 0 : (k == 2 ? 1 : 2);
        next();
        sweep<FLAV>(k < m - 1 ? MODE_MID : last_mode, grid, s, A, k, K, cur, prev, prev_kind,
                    bufs[(k + 1) % 2]);
        mark(a, s, !per_slot ? KIND_REFINE : k == 1 ? KIND_FIRST : k < m - 1 ? KIND_MID
                                                                             : KIND_LAST);
      }
    };
    if (c.refine <= 0) {
      solve(a->sweeps_eff, MODE_FINAL_X, true);
    } else {
      solve(a->sweeps_eff, MODE_DACC_SET, true);
      for (int r = 0; r < c.refine; ++r) {
        next();
        k_residual<BX, BY, BZ><<<grid1, kBlock, 0, s>>>(A);
        mark(a, s, KIND_RESIDUAL);
        solve(c.refine_sweeps, r == c.refine - 1 ? MODE_DACC_ADD_X : MODE_DACC_ADD, false);
      }
    }
    // the next step's per-step values were staged by this step's k_rhs
    a->par ^= 1;
  }

  // Average duration (ms) of the reaction, rhs and sweep kernels over `reps` launches,
  // timed with HIP events on `s`.  The reaction writes into scratch (DA, DB) and the sweep
  // runs in MID mode, so no simulation state changes.  The kernels are the ones a step of
  // the arena launches (of its flavour).
  static hipError_t profile(phs_arena* a, int reps, double* ms, hipStream_t s) {
    if constexpr (NB == 1) {
      switch (a->flav) {
        case FLAV_LEAN: return profile_t<FLAV_LEAN>(a, reps, ms, s);
        case FLAV_LEAN_FLAGS: return profile_t<FLAV_LEAN_FLAGS>(a, reps, ms, s);
        case FLAV_LEAN_IDENT: return profile_t<FLAV_LEAN_IDENT>(a, reps, ms, s);
        case FLAV_LEAN_FLAGS_IDENT: return profile_t<FLAV_LEAN_FLAGS_IDENT>(a, reps, ms, s);
        default: break;
      }
    }
    return profile_t<FLAV_GENERAL>(a, reps, ms, s);
  }
  template <int FLAV>
  static hipError_t profile_t(phs_arena* a, int reps, double* ms, hipStream_t s) {
    const ArenaDev A = dev_view(a);
    const int grid = grid_for(a);
    hipEvent_t ev[4];
    for (auto& e : ev)
      if (hipError_t r = hipEventCreate(&e)) return r;
    ArenaDev R = A;
    R.UO = a->DA;
    R.HO = a->DB;
    hipEventRecord(ev[0], s);
    for (int i = 0; i < reps; ++i)
      k_reaction<NB><<<grid, kBlock, 0, s>>>(R, 0, (int)a->brick_high);
    hipEventRecord(ev[1], s);
    for (int i = 0; i < reps; ++i) k_rhs<BX, BY, BZ, NB, GUESS, FLAV><<<grid, kBlock, 0, s>>>(A);
    hipEventRecord(ev[2], s);
    for (int i = 0; i < reps; ++i)
      k_sweep<BX, BY, BZ, MODE_MID, NB, GUESS, false, FLAV><<<grid, kBlock, 0, s>>>(
          A, 3, 0, a->DA, a->DB, 2, a->DB);
    hipEventRecord(ev[3], s);
    hipError_t r = hipEventSynchronize(ev[3]);
    for (int i = 0; i < 3; ++i) {
      float t = 0.0f;
      hipEventElapsedTime(&t, ev[i], ev[i + 1]);
      ms[i] = (double)t / reps;
    }
    for (auto& e : ev) hipEventDestroy(e);
    return r;
  }

  // Reaction of step 0 for the bricks [b_lo, b_hi) of a newly added simulation.
  static void react_range(phs_arena* a, int b_lo, int b_hi, hipStream_t s) {
    ArenaDev A = dev_view(a);
    A.flags &= ~1;  // no XCD remap: the grid covers a sub-range
    const int per = kWaves * NB;
    const int grid = (b_hi - b_lo + per - 1) / per;
    if (grid > 0) k_reaction<NB><<<grid, kBlock, 0, s>>>(A, b_lo, b_hi);
  }

  static void diag(phs_arena* a, int slot, hipStream_t s) {
    k_diag<BX, BY, BZ><<<grid_for_nb1(a), kBlock, 0, s>>>(dev_view(a), slot, a->diag_acc);
  }
};

bool shape_supported(const int* b) {
  const int s[][3] = {{4, 4, 4}, {8, 8, 1}, {8, 4, 2}, {4, 8, 2}, {16, 4, 1}};
  for (auto& q : s)
    if (q[0] == b[0] && q[1] == b[1] && q[2] == b[2]) return true;
  return false;
}

template <int BX, int BY, int BZ, typename F>
void with_nb(const phs_arena* a, F&& f) {
  switch (nb_per_wave(a)) {
    case 2:
      if (a->cfg.guess == 2) f(Step<BX, BY, BZ, 2, 2>());
      else if (a->cfg.guess == 1) f(Step<BX, BY, BZ, 2, 1>());
      else f(Step<BX, BY, BZ, 2, 0>());
      break;
    default:
      if (a->cfg.guess == 2) f(Step<BX, BY, BZ, 1, 2>());
      else if (a->cfg.guess == 1) f(Step<BX, BY, BZ, 1, 1>());
      else f(Step<BX, BY, BZ, 1, 0>());
  }
}

// Calls f(Step<shape, NB>()) for the arena's brick shape and bricks per wavefront.
template <typename F>
void dispatch(const phs_arena* a, F&& f) {
  const int* b = a->cfg.brick;
  if (b[0] == 8 && b[1] == 8) with_nb<8, 8, 1>(a, f);
  else if (b[0] == 8 && b[1] == 4) with_nb<8, 4, 2>(a, f);
  else if (b[0] == 4 && b[1] == 8) with_nb<4, 8, 2>(a, f);
  else if (b[0] == 16) with_nb<16, 4, 1>(a, f);
  else with_nb<4, 4, 4>(a, f);
}

void launch_step(phs_arena* a, hipStream_t s) {
  dispatch(a, [&](auto st) { decltype(st)::run(a, s); });
}

hipError_t profile_kernels(phs_arena* a, int reps, double* ms, hipStream_t s) {
  hipError_t r = hipSuccess;
  dispatch(a, [&](auto st) { r = decltype(st)::profile(a, reps, ms, s); });
  return r;
}

hipError_t launch_react_range(phs_arena* a, int b_lo, int b_hi, hipStream_t s) {
  dispatch(a, [&](auto st) { decltype(st)::react_range(a, b_lo, b_hi, s); });
  return hipGetLastError();
}

void launch_diag(phs_arena* a, int slot, hipStream_t s) {
  dispatch(a, [&](auto st) { decltype(st)::diag(a, slot, s); });
}

void free_slot_host(SlotHost& h) {
  if (h.upload_done) {
    hipEventSynchronize(h.upload_done);
    hipEventDestroy(h.upload_done);
    h.upload_done = nullptr;
  }
  if (h.staging) hipHostFree(h.staging);
  h.staging = nullptr;
  hipFree(h.perm_dev);
  hipFree(h.sched_dev);
  h.perm_dev = nullptr;
  h.sched_dev = nullptr;
}

// Free slot resources whose last use (recorded by `done`) has completed; `wait` blocks.
void reap(phs_arena* a, bool wait) {
  std::vector<DeferredFree> keep;
  for (auto& g : a->graveyard) {
    if (g.done && !wait && hipEventQuery(g.done) == hipErrorNotReady) {
      keep.push_back(g);
      continue;
    }
    if (g.done) {
      hipEventSynchronize(g.done);
      hipEventDestroy(g.done);
    }
    if (g.host) hipHostFree(g.host);
    hipFree(g.dev[0]);
    hipFree(g.dev[1]);
  }
  a->graveyard.swap(keep);
}

void recompute_high(phs_arena* a) {
  int64_t hi = 0;
  int32_t sw = 2;
  for (auto& h : a->sh)
    if (h.used) {
      hi = std::max(hi, h.brick_off + h.n_bricks);
      sw = std::max(sw, h.need_sweeps);
    }
  a->brick_high = hi;
  a->sweeps_eff = a->cfg.sweeps > 0 ? a->cfg.sweeps : sw;
  bool lean = true, flags = false, ident = false;
  for (auto& h : a->sh)
    if (h.used) {
      if (h.hetero || h.hparams) lean = false;
      ident = ident || h.need_sweeps != a->sweeps_eff;
      flags = flags || h.needs_tb;
    }
  a->flav = !lean  ? FLAV_GENERAL
            : flags ? (ident ? FLAV_LEAN_FLAGS_IDENT : FLAV_LEAN_FLAGS)
                    : (ident ? FLAV_LEAN_IDENT : FLAV_LEAN);
}

}  // namespace

extern "C" {

int phs_arena_create(const phs_arena_config* cfg, phs_arena** handle) {
  if (!cfg || !handle) return phs::fail(PHS_ERR_ARG, "NULL argument");
  if (cfg->max_voxels <= 0 || cfg->max_bricks <= 0 || cfg->max_sims <= 0)
    return phs::fail(PHS_ERR_ARG, "capacities must be > 0");
  if (cfg->max_voxels >= (1LL << 29) || cfg->max_bricks >= (1LL << 28))
    return phs::fail(PHS_ERR_ARG, "capacity exceeds 32-bit byte offsets (max 2^29 voxels)");
  if (!shape_supported(cfg->brick))
    return phs::fail(PHS_ERR_ARG, "brick shape must be one of 4x4x4, 8x8x1, 8x4x2, 4x8x2, 16x4x1");
  if (cfg->sweeps != 0 && (cfg->sweeps < 2 || cfg->sweeps > kMaxSweeps))
    return phs::fail(PHS_ERR_ARG, "sweeps must be 0 (automatic) or in [2, 64]");
  if (cfg->refine < 0 || cfg->refine > 4) return phs::fail(PHS_ERR_ARG, "refine must be in [0, 4]");
  if (cfg->bricks_per_wave != 0 && cfg->bricks_per_wave != 1 && cfg->bricks_per_wave != 2)
    return phs::fail(PHS_ERR_ARG, "bricks_per_wave must be 1 or 2");
  if (cfg->guess < 0 || cfg->guess > 2) return phs::fail(PHS_ERR_ARG, "guess must be 0, 1 or 2");
  if (cfg->refine > 0 && (cfg->refine_sweeps < 2 || cfg->refine_sweeps > kMaxSweeps))
    return phs::fail(PHS_ERR_ARG, "refine_sweeps must be in [2, 64]");
  phs_arena* a = new phs_arena();
  a->cfg = *cfg;
  a->sweeps_eff = cfg->sweeps > 0 ? cfg->sweeps : 2;
  const size_t nv = (size_t)cfg->max_voxels;
  hipError_t e = hipSuccess;
  const bool need_dacc = cfg->refine > 0 || cfg->diag;
  if ((e = hipMalloc(&a->bricks, sizeof(BrickRec) * (size_t)cfg->max_bricks)) ||
      (e = hipMalloc(&a->slots, sizeof(SlotRec) * (size_t)cfg->max_sims)) ||
      (e = hipMalloc(&a->U, 4 * nv)) || (e = hipMalloc(&a->H, 4 * nv)) ||
      (e = hipMalloc(&a->R0, 4 * nv)) || (e = hipMalloc(&a->DA, 4 * nv)) ||
      (e = hipMalloc(&a->DB, 4 * nv)) || (e = hipMalloc(&a->VF, nv)) ||
      (e = hipMalloc(&a->ALPHA, 4 * nv)) || (e = hipMalloc(&a->diag_acc, 16)) ||
      (e = hipMalloc(&a->NONFINITE, 4 * (size_t)cfg->max_sims)) ||
      (need_dacc && (e = hipMalloc(&a->DACC, 8 * nv))) ||
      (cfg->guess >= 1 && (e = hipMalloc(&a->D1, 4 * nv))) ||
      (cfg->guess == 2 && (e = hipMalloc(&a->D2, 4 * nv))) ||
      (cfg->diag && (e = hipMalloc(&a->USTAR, 4 * nv)))) {
    phs_arena_destroy(a);
    return phs::fail(PHS_ERR_OOM, std::string("arena hipMalloc: ") + hipGetErrorString(e));
  }
  std::vector<SlotRec> zero(cfg->max_sims);
  std::memset(zero.data(), 0, sizeof(SlotRec) * zero.size());
  if ((e = hipMemcpy(a->slots, zero.data(), sizeof(SlotRec) * zero.size(), hipMemcpyHostToDevice)) ||
      (e = hipMemset(a->bricks, 0xFF, sizeof(BrickRec) * (size_t)cfg->max_bricks))) {
    phs_arena_destroy(a);
    return phs::fail(PHS_ERR_HIP, std::string("arena init: ") + hipGetErrorString(e));
  }
  a->sh.resize(cfg->max_sims);
  a->vox.init(cfg->max_voxels);
  a->brk.init(cfg->max_bricks);
  *handle = a;
  return PHS_OK;
}

int phs_arena_add(phs_arena* a, const phs_sim_desc* d, const phs_outputs* out, int32_t* slot_out,
                  void* stream) {
  if (!a || !d || !slot_out) return phs::fail(PHS_ERR_ARG, "NULL argument");
  if (d->n_voxel <= 0 || (d->model != 0 && d->model != 1) || !(d->dt > 0) || !(d->delta > 0))
    return phs::fail(PHS_ERR_ARG, "invalid simulation description");
  if (!d->ijk) return phs::fail(PHS_ERR_ARG, "lattice coordinates (ijk) required");
  if (d->n_steps > 0 && !d->schedule) return phs::fail(PHS_ERR_ARG, "schedule required");
  if (out && out->max_beats > 7) return phs::fail(PHS_ERR_ARG, "max_beats must be <= 7");
  if (out && out->n_probes > 0) return phs::fail(PHS_ERR_ARG, "probes are a plain-form output");
  if (d->nb && d->nb_stride < 6) return phs::fail(PHS_ERR_ARG, "nb_stride must be >= 6");
  hipStream_t s = (hipStream_t)stream;
  reap(a, false);
  int slot = -1;
  for (int i = 0; i < a->cfg.max_sims; ++i)
    if (!a->sh[i].used) {
      slot = i;
      break;
    }
  if (slot < 0) return phs::fail(PHS_ERR_CAPACITY, "no free simulation slot");
  phs::BrickPlan P;
  std::string err;
  const int rc = phs::build_plan(d->n_voxel, d->ijk, d->nb, d->nb_stride, a->cfg.brick[0],
                                 a->cfg.brick[1], a->cfg.brick[2], P, err);
  if (rc == -2) return phs::fail(PHS_ERR_UNSUPPORTED, err);
  if (rc) return phs::fail(PHS_ERR_ARG, err);
  const int64_t vo = a->vox.alloc(d->n_voxel);
  if (vo < 0) return phs::fail(PHS_ERR_CAPACITY, "arena voxel capacity exhausted");
  const int64_t bo = a->brk.alloc(P.n_bricks);
  if (bo < 0) {
    a->vox.release(vo, d->n_voxel);
    return phs::fail(PHS_ERR_CAPACITY, "arena brick capacity exhausted");
  }
  SlotHost& h = a->sh[slot];
  h = SlotHost();
  h.voxel_off = vo;
  h.brick_off = bo;
  h.n_voxel = d->n_voxel;
  h.n_bricks = P.n_bricks;
  h.n_steps = d->n_steps;
  h.perm = P.perm;
  const int N = d->n_voxel;
  auto undo = [&](int code, const std::string& msg) {
    free_slot_host(h);
    a->vox.release(vo, N);
    a->brk.release(bo, P.n_bricks);
    h = SlotHost();
    return phs::fail(code, msg);
  };

  // Host staging (pinned): bricks | slot | U | H | VF | ALPHA | perm | schedule
  const size_t sz_br = sizeof(BrickRec) * (size_t)P.n_bricks;
  const size_t sz_sl = sizeof(SlotRec);
  const size_t sz_f = 4 * (size_t)N;
  const bool hparams = d->params_voxel != nullptr;
  const size_t total = sz_br + sz_sl + 4 * sz_f + (size_t)N + (size_t)d->n_steps + 64 +
                       (hparams ? 5 * sz_f + 64 : 0);
  hipError_t e = hipHostMalloc(&h.staging, total, hipHostMallocDefault);
  if (e != hipSuccess) return undo(PHS_ERR_OOM, "hipHostMalloc staging");
  char* st = (char*)h.staging;
  BrickRec* hb = (BrickRec*)st;
  SlotRec* hs = (SlotRec*)(st + sz_br);
  float* hu = (float*)(st + sz_br + sz_sl);
  float* hh = hu + N;
  float* hal = hh + N;
  int32_t* hperm = (int32_t*)(hal + N);
  uint8_t* hvf = (uint8_t*)(hperm + N);
  uint8_t* hsched = hvf + N;
  // per-voxel parameters (compact order), 16-byte aligned after the schedule
  float* hpv = hparams ? (float*)(((uintptr_t)(hsched + d->n_steps) + 15) & ~(uintptr_t)15)
                       : nullptr;

  bool hetero = false;
  bool needs_vf = false, needs_tb = false;
  double amax_deg = 0.0;
  // time resolution of the step (the automatic sweep count's accuracy term): the steepest
  // upstroke over the voxels' cell parameters
  double resolution = phs::step_resolution(d->model, d->params, d->dt);
  const double a0 = d->dt * d->c / (2.0 * d->delta * d->delta);
  const int bd[3] = {a->cfg.brick[0], a->cfg.brick[1], a->cfg.brick[2]};
  const int stride[3] = {1, bd[0], bd[0] * bd[1]};
  uint64_t inb[6];  // lanes whose d-neighbour cell lies in the same brick
  for (int k = 0; k < 6; ++k) {
    inb[k] = 0;
    for (int l = 0; l < 64; ++l) {
      const int c[3] = {l % bd[0], (l / bd[0]) % bd[1], l / (bd[0] * bd[1])};
      const int ax = k >> 1;
      if ((k & 1) ? c[ax] > 0 : c[ax] < bd[ax] - 1) inb[k] |= 1ULL << l;
    }
  }
  auto shl = [](uint64_t m, int n) { return n >= 64 ? 0ULL : m << n; };
  auto shr = [](uint64_t m, int n) { return n >= 64 ? 0ULL : m >> n; };
  for (int32_t b = 0; b < P.n_bricks; ++b) {
    BrickRec r;
    const uint64_t mask = P.mask[b];
    r.mask = mask;
    r.start = (int32_t)(vo + P.start[b]);
    r.slot = slot;
    for (int k = 0; k < 6; ++k) {
      const int ax = k >> 1;
      const int st = stride[ax];
      const int wrap = (bd[ax] - 1) * stride[ax];
      const int32_t nb = P.nbb[(size_t)b * 6 + k];
      const uint64_t nmask = nb >= 0 ? P.mask[nb] : 0;
      // + direction: the neighbour of lane l is lane l + st (in the brick) or the neighbour
      // brick's cell l - wrap; - direction: l - st, or the cell l + wrap
      // presence from the lattice, less the links the caller's table cuts
      const uint64_t keep = ~P.cut[(size_t)b * 6 + k];
      r.pin[k] = ((k & 1) ? shl(mask, st) : shr(mask, st)) & inb[k] & mask & keep;
      r.sh[k] = (k & 1) ? shr(nmask, wrap) : shl(nmask, wrap);
      r.pcr[k] = r.sh[k] & ~inb[k] & mask & keep;
      int32_t below = 0;
      if (nb >= 0 && (k & 1)) {
        // - face: the gather rank is counted from the face-adjacent offset (B-1)*stride
        below = wrap > 0 ? __builtin_popcountll(nmask & ((1ULL << wrap) - 1ULL)) : 0;
      }
      r.nstart[k] = nb >= 0 ? (int32_t)(vo + P.start[nb] + below) : 0;
    }
    r.tblock = 0;  // set below, from the voxel flags, with the scar rows' presence cut
    hb[b] = r;
  }
  for (int32_t c = 0; c < N; ++c) {
    const int32_t o = P.perm[c];
    hperm[c] = o;
    hu[c] = d->u0 ? d->u0[o] : 0.0f;
    hh[c] = d->h0 ? d->h0[o] : (d->model == 0 ? 1.0f : 0.0f);
    uint8_t f = d->vflags ? (uint8_t)(d->vflags[o] & 0x0F) : 0;
    if (f & (PHS_VF_SCAR | PHS_VF_TEMP_BLOCK)) needs_vf = true;
    if (f & PHS_VF_TEMP_BLOCK) needs_tb = true;
    if (hu[c] >= (float)d->lat_threshold) f |= PHS_VF_ABOVE;
    hvf[c] = f;
    const double cv = d->c_voxel ? d->c_voxel[o] : d->c;
    if (cv < 0) return undo(PHS_ERR_ARG, "negative conductivity");
    if (d->c_voxel && cv != d->c) hetero = true;
    const double al = d->dt * cv / (2.0 * d->delta * d->delta);
    hal[c] = (float)al;
    if (hparams) {
      double q[5];
      for (int k = 0; k < 5; ++k) q[k] = d->params_voxel[(int64_t)k * N + o];
      if (d->model == 0 && !(q[0] > 0 && q[1] > 0 && q[2] > 0 && q[3] > 0))
        return undo(PHS_ERR_ARG, "Mitchell-Schaeffer time constants must be > 0");
      for (int k = 0; k < 5; ++k)
        hpv[(size_t)k * N + c] = d->model == 0 && k < 4 ? (float)(1.0 / q[k]) : (float)q[k];
      resolution = c == 0 ? phs::step_resolution(d->model, q, d->dt)
                          : std::max(resolution, phs::step_resolution(d->model, q, d->dt));
    }
  }
  // Spectral bound: 1 + 2 max(alpha_n deg_n) over conducting voxels; the scar and
  // temporary-block lane masks of the brick records.
  for (int32_t b = 0; b < P.n_bricks; ++b) {
    const uint64_t m = P.mask[b];
    for (int lane = 0; lane < 64; ++lane) {
      if (!((m >> lane) & 1)) continue;
      const int32_t c = P.start[b] + __builtin_popcountll(m & ((1ULL << lane) - 1));
      if (hvf[c] & PHS_VF_SCAR)  // the row of a non-conductive voxel: no neighbour terms
        for (int k = 0; k < 6; ++k) {
          hb[b].pin[k] &= ~(1ULL << lane);
          hb[b].pcr[k] &= ~(1ULL << lane);
        }
      if (hvf[c] & PHS_VF_TEMP_BLOCK) hb[b].tblock |= 1ULL << lane;
      if (hvf[c] & PHS_VF_SCAR) continue;
      int deg = 0;
      const int32_t o = P.perm[c];
      if (d->nb) {
        for (int k = 0; k < 6; ++k) deg += d->nb[(int64_t)o * d->nb_stride + k] >= 0;
      } else {
        deg = 6;
      }
      amax_deg = std::max(amax_deg, (double)hal[c] * deg);
    }
  }
  for (int64_t i = 0; i < d->n_steps; ++i) hsched[i] = d->schedule[i];

  h.lmin = 1.0;
  h.lmax = 1.0 + 2.0 * amax_deg;
  h.needs_vf = needs_vf;
  h.needs_tb = needs_tb;
  h.hetero = hetero;
  {
    // Inexact CN steps amplify the stiffest modes unless the sweep polynomial is small
    // enough there; reject a fixed count that would, or pick one (automatic mode).
    const phs_arena_config& c = a->cfg;
    const int rs = c.refine > 0 ? c.refine_sweeps : 0;
    const int need = phs::stable_sweeps(h.lmax, c.guess, c.refine, rs, kMaxSweeps);
    char msg[256];
    if (need < 0) {
      std::snprintf(msg, sizeof msg, "no sweep count <= %d keeps the CN step stable for "
                    "lambda_max = %.3f (guess %d)", kMaxSweeps, h.lmax, c.guess);
      return undo(PHS_ERR_UNSUPPORTED, msg);
    }
    if (c.sweeps > 0 && c.sweeps < need) {
      std::snprintf(msg, sizeof msg, "sweeps = %d makes the CN step unstable for lambda_max = "
                    "%.3f (guess %d, refine %d): needs >= %d, or sweeps = 0 (automatic)",
                    c.sweeps, h.lmax, c.guess, c.refine, need);
      return undo(PHS_ERR_ARG, msg);
    }
    // a fixed count runs on [1, lmax]; the automatic one may raise the interval's lower end
    h.need_sweeps = c.sweeps > 0 ? c.sweeps
                                 : phs::auto_sweeps(h.lmax, c.guess, c.refine, rs, kMaxSweeps,
                                                    resolution, &h.lmin);
  }
  std::vector<phs::ChebCoef> coef(kMaxSweeps);
  SlotRec sr;
  std::memset(&sr, 0, sizeof(sr));
  if (h.lmax > 1.0) {
    phs::chebyshev_coefficients(h.lmin, h.lmax, kMaxSweeps, coef.data());
  } else {
    for (auto& q : coef) q = phs::ChebCoef{0.0, 0.0, 0.0};
  }
  for (int k = 0; k < kMaxSweeps; ++k) {
    sr.coef[k][0] = (float)coef[k].a;
    sr.coef[k][1] = (float)coef[k].b;
    sr.coef[k][2] = (float)coef[k].g;
    sr.coef[k][3] = 0.0f;
  }
  SlotDyn d0;
  std::memset(&d0, 0, sizeof(d0));
  d0.step = 0;
  sr.n_steps = d->n_steps;
  d0.cur_sched = d->n_steps > 0 ? d->schedule[0] : 0;
  d0.next_sched = d->n_steps > 1 ? d->schedule[1] : 0;
  d0.cur_off = d->n_samples > 0 ? 0 : -1;  // step 0 writes sample row 0
  sr.voxel_off = vo;
  sr.su = out ? (float*)out->samples_u : nullptr;
  sr.sh = out ? (float*)out->samples_h : nullptr;
  sr.lat = out ? out->lat : nullptr;
  sr.n_voxel = N;
  d0.live = d->n_steps > 0 ? 1 : 0;
  sr.dyn[0] = d0;  // both copies: the slot's first step may run at either launch parity
  sr.dyn[1] = d0;
  sr.model = d->model;
  sr.nspm = d->steps_per_sample;
  sr.n_samples = d->n_samples;
  sr.max_beats = out ? out->max_beats : 0;
  sr.needs_vf = needs_vf ? 1 : 0;
  sr.hetero = hetero ? 1 : 0;
  sr.ring = out ? std::max(0, out->sample_ring) : 0;
  sr.nsw = h.need_sweeps;
  sr.hparams = hparams ? 1 : 0;
  h.hparams = hparams;
  if (d->model == 0) {
    sr.p[0] = (float)(1.0 / d->params[0]);
    sr.p[1] = (float)(1.0 / d->params[1]);
    sr.p[2] = (float)(1.0 / d->params[2]);
    sr.p[3] = (float)(1.0 / d->params[3]);
    sr.p[4] = (float)d->params[4];
  } else {
    for (int i = 0; i < 5; ++i) sr.p[i] = (float)d->params[i];
  }
  sr.dt = (float)d->dt;
  sr.alpha = (float)a0;
  sr.alpha64 = a0;
  sr.lat_thr = (float)d->lat_threshold;

  if ((e = hipMalloc(&h.perm_dev, 4 * (size_t)N)) ||
      (e = hipMalloc(&h.sched_dev, (size_t)std::max<int64_t>(d->n_steps, 1))))
    return undo(PHS_ERR_OOM, std::string("slot hipMalloc: ") + hipGetErrorString(e));
  if (hparams && !a->PV[0]) {
    // the arena's per-voxel parameter arrays, kept until it is destroyed
    for (int k = 0; k < 5; ++k)
      if ((e = hipMalloc(&a->PV[k], 4 * (size_t)a->cfg.max_voxels))) {
        for (auto& q : a->PV) {
          hipFree(q);
          q = nullptr;
        }
        return undo(PHS_ERR_OOM, std::string("per-voxel parameters: ") + hipGetErrorString(e));
      }
  }
  for (int k = 0; hparams && k < 5; ++k)
    if ((e = hipMemcpyAsync(a->PV[k] + vo, hpv + (size_t)k * N, sz_f, hipMemcpyHostToDevice, s)))
      return undo(PHS_ERR_HIP, std::string("slot upload: ") + hipGetErrorString(e));
  sr.sched = h.sched_dev;
  *hs = sr;
  if ((e = hipMemcpyAsync(a->bricks + bo, hb, sz_br, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(a->U + vo, hu, sz_f, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(a->H + vo, hh, sz_f, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(a->ALPHA + vo, hal, sz_f, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(a->VF + vo, hvf, (size_t)N, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(h.perm_dev, hperm, sz_f, hipMemcpyHostToDevice, s)) ||
      (d->n_steps > 0 &&
       (e = hipMemcpyAsync(h.sched_dev, hsched, (size_t)d->n_steps, hipMemcpyHostToDevice, s))) ||
      (out && out->lat && out->max_beats > 0 &&
       (e = hipMemsetAsync(out->lat, 0xFF, 4 * (size_t)out->max_beats * N, s))) ||
      (a->D1 && (e = hipMemsetAsync(a->D1 + vo, 0, sz_f, s))) ||
      (a->D2 && (e = hipMemsetAsync(a->D2 + vo, 0, sz_f, s))) ||
      (e = hipMemcpyAsync(a->slots + slot, hs, sz_sl, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(a->NONFINITE + slot, &kNoStep, 4, hipMemcpyHostToDevice, s)) ||
      (e = launch_react_range(a, (int)bo, (int)(bo + P.n_bricks), s)) ||
      (e = hipEventCreateWithFlags(&h.upload_done, hipEventDisableTiming)) ||
      (e = hipEventRecord(h.upload_done, s)))
    return undo(PHS_ERR_HIP, std::string("slot upload: ") + hipGetErrorString(e));
  h.used = true;
  recompute_high(a);
  *slot_out = slot;
  return PHS_OK;
}

int phs_arena_remove(phs_arena* a, int32_t slot, void* stream) {
  if (!a || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  SlotHost& h = a->sh[slot];
  hipStream_t s = (hipStream_t)stream;
  // Stream-ordered: steps already queued on `stream` still see the simulation; later work
  // (including an add that reuses its ranges) does not.
  k_invalidate<<<(h.n_bricks + 255) / 256, 256, 0, s>>>(a->bricks, h.brick_off, h.n_bricks);
  k_deactivate<<<1, 1, 0, s>>>(a->slots, slot);
  PHS_HIP_CHECK(hipGetLastError());
  DeferredFree g;
  PHS_HIP_CHECK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  PHS_HIP_CHECK(hipEventRecord(g.done, s));
  g.host = h.staging;
  g.dev[0] = h.perm_dev;
  g.dev[1] = h.sched_dev;
  if (h.upload_done) hipEventDestroy(h.upload_done);
  a->graveyard.push_back(g);
  a->vox.release(h.voxel_off, h.n_voxel);
  a->brk.release(h.brick_off, h.n_bricks);
  h = SlotHost();
  recompute_high(a);
  reap(a, false);
  return PHS_OK;
}

// Launches cover bricks [0, brick_high), so the ranges that finished simulations leave below
// the highest resident one are launched for nothing.  A brick record holds voxel indices only
// (its own start and its neighbours' gather bases, all in its simulation's voxel range), never
// another brick's index, so a simulation's records can move to any brick range as they are.
// Compaction slides the simulations down in brick order until they are contiguous from brick
// 0; with `voxels` it also slides their voxel data down in voxel order (so that the free
// voxels form one range an admission can use), rebasing the records' voxel indices and the
// slot's offset.  Per-step scratch arrays (R0, DA, DB) hold nothing between steps and stay.
}  // extern "C"

namespace {

// `bytes` of the arena's move scratch; copies queued on `s` may still read the old range.
hipError_t move_scratch(phs_arena* a, size_t bytes, hipStream_t s) {
  if (a->move_bytes >= bytes) return hipSuccess;
  const hipError_t w = a->move_tmp ? hipStreamSynchronize(s) : hipSuccess;
  hipFree(a->move_tmp);
  a->move_tmp = nullptr;
  a->move_bytes = 0;
  if (w != hipSuccess) return w;
  const hipError_t e = hipMalloc(&a->move_tmp, bytes);
  if (e == hipSuccess) a->move_bytes = bytes;
  return e;
}

// Elements [from, from + n) of `base` to [to, to + n), to < from, through the scratch when
// the ranges overlap.
hipError_t move_down(phs_arena* a, void* base, size_t elem, int64_t from, int64_t to, int64_t n,
                     hipStream_t s) {
  char* b = (char*)base;
  const size_t bytes = elem * (size_t)n;
  if (to + n <= from)
    return hipMemcpyAsync(b + elem * to, b + elem * from, bytes, hipMemcpyDeviceToDevice, s);
  hipError_t e = move_scratch(a, bytes, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(a->move_tmp, b + elem * from, bytes, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b + elem * to, a->move_tmp, bytes, hipMemcpyDeviceToDevice, s);
  return e;
}

// Slides the resident simulations' ranges (brick or voxel) down until they are contiguous
// from 0; `off`/`len` pick the range of a slot, `move` moves one slot to a lower offset.
template <typename Off, typename Len, typename Move>
int slide(phs_arena* a, RangeAlloc& alloc, Off off, Len len, Move move, int32_t& n_moved) {
  std::vector<int> order;
  for (int i = 0; i < a->cfg.max_sims; ++i)
    if (a->sh[i].used) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return off(x) < off(y); });
  int64_t next = 0;  // [0, next) holds the simulations already placed
  for (int i : order) {
    const int64_t n = len(i), old = off(i);
    if (old > next) {
      if (int rc = move(i, next)) return rc;
      // [0, next) is full, so once the old range is free the lowest range that fits n
      // starts at next
      alloc.release(old, n);
      if (alloc.alloc(n) != next) return phs::fail(PHS_ERR_STATE, "ranges out of order");
      ++n_moved;
    }
    next += n;
  }
  return PHS_OK;
}

}  // namespace

extern "C" {

int phs_arena_compact(phs_arena* a, double max_waste, int32_t voxels, int32_t* moved,
                      void* stream) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (!(max_waste >= 0.0)) return phs::fail(PHS_ERR_ARG, "max_waste must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  if (moved) *moved = 0;
  int32_t n_moved = 0;
  int rc = PHS_OK;
  if (voxels) {
    rc = slide(
        a, a->vox, [&](int i) { return a->sh[i].voxel_off; },
        [&](int i) { return (int64_t)a->sh[i].n_voxel; },
        [&](int i, int64_t to) -> int {
          SlotHost& h = a->sh[i];
          const int64_t from = h.voxel_off, n = h.n_voxel;
          // the arrays that carry a simulation from one step to the next
          const struct {
            void* p;
            size_t elem;
          } arr[] = {{a->U, 4},     {a->H, 4},     {a->D1, 4},    {a->D2, 4},
                     {a->ALPHA, 4}, {a->VF, 1},    {a->DACC, 8},  {a->USTAR, 4},
                     {h.hparams ? a->PV[0] : nullptr, 4}, {h.hparams ? a->PV[1] : nullptr, 4},
                     {h.hparams ? a->PV[2] : nullptr, 4}, {h.hparams ? a->PV[3] : nullptr, 4},
                     {h.hparams ? a->PV[4] : nullptr, 4}};
          for (const auto& x : arr)
            if (x.p) {
              const hipError_t e = move_down(a, x.p, x.elem, from, to, n, s);
              if (e != hipSuccess)
                return phs::fail(PHS_ERR_HIP, std::string("voxel move: ") + hipGetErrorString(e));
            }
          k_rebase<<<(unsigned)((h.n_bricks + 255) / 256), 256, 0, s>>>(
              a->bricks, h.brick_off, h.n_bricks, (int32_t)(to - from));
          k_set_voxel_off<<<1, 64, 0, s>>>(a->slots, i, to);
          PHS_HIP_CHECK(hipGetLastError());
          h.voxel_off = to;
          return PHS_OK;
        },
        n_moved);
  }
  int64_t live = 0;
  for (const auto& h : a->sh)
    if (h.used) live += h.n_bricks;
  const int64_t launched = (int64_t)grid_for(a) * kWaves * nb_per_wave(a);
  if (rc == PHS_OK && live > 0 && (double)launched > (1.0 + max_waste) * (double)live) {
    rc = slide(
        a, a->brk, [&](int i) { return a->sh[i].brick_off; },
        [&](int i) { return (int64_t)a->sh[i].n_bricks; },
        [&](int i, int64_t to) -> int {
          SlotHost& h = a->sh[i];
          const int64_t from = h.brick_off, n = h.n_bricks;
          const hipError_t e = move_down(a, a->bricks, sizeof(BrickRec), from, to, n, s);
          if (e != hipSuccess)
            return phs::fail(PHS_ERR_HIP, std::string("brick move: ") + hipGetErrorString(e));
          // the old records not overwritten by the new range
          const int64_t lo = std::max(from, to + n);
          k_invalidate<<<(unsigned)((from + n - lo + 255) / 256), 256, 0, s>>>(a->bricks, lo,
                                                                                from + n - lo);
          PHS_HIP_CHECK(hipGetLastError());
          h.brick_off = to;
          return PHS_OK;
        },
        n_moved);
  }
  recompute_high(a);  // the moves made stand, also after a failure
  if (moved) *moved = n_moved;
  return rc;
}

int phs_arena_step(phs_arena* a, int32_t n_steps, void* stream) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (n_steps < 0) return phs::fail(PHS_ERR_ARG, "n_steps < 0");
  hipStream_t s = (hipStream_t)stream;
  if (a->brick_high == 0) return PHS_OK;
  for (int i = 0; i < n_steps; ++i) launch_step(a, s);
  PHS_HIP_CHECK(hipGetLastError());
  for (auto& h : a->sh)
    if (h.used) h.steps_done = std::min<int64_t>(h.n_steps, h.steps_done + n_steps);
  return PHS_OK;
}

}  // extern "C"

namespace {

// The cached graph of n_steps steps starting at launch parity par0, captured (not launched)
// when the cache holds none for the arena's current launch pattern.
int ensure_graph(phs_arena* a, int32_t n_steps, int32_t par0) {
  const int64_t grid = grid_for(a);
  phs_arena::Graph& G = a->g[par0];
  if (G.exec && G.steps == n_steps && G.grid == grid && G.high == a->brick_high &&
      G.sweeps == a->sweeps_eff && G.flav == a->flav && G.pv == a->PV[0])
    return PHS_OK;
  if (G.exec) hipGraphExecDestroy(G.exec);
  if (G.graph) hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
  if (!a->cap_stream) PHS_HIP_CHECK(hipStreamCreateWithFlags(&a->cap_stream, hipStreamNonBlocking));
  PHS_HIP_CHECK(hipStreamBeginCapture(a->cap_stream, hipStreamCaptureModeThreadLocal));
  const int32_t par_saved = a->par;
  a->par = par0;
  for (int i = 0; i < n_steps; ++i) launch_step(a, a->cap_stream);  // flips a->par
  a->par = par_saved;  // nothing ran: only a replay advances the parity
  hipGraph_t g = nullptr;
  PHS_HIP_CHECK(hipStreamEndCapture(a->cap_stream, &g));
  G.graph = g;
  PHS_HIP_CHECK(hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0));
  G.steps = n_steps;
  G.grid = grid;
  G.high = a->brick_high;
  G.sweeps = a->sweeps_eff;
  G.flav = a->flav;
  G.pv = a->PV[0];
  return PHS_OK;
}

}  // namespace

extern "C" {

int phs_arena_graph_prepare(phs_arena* a, int32_t n_steps) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (n_steps <= 0) return phs::fail(PHS_ERR_ARG, "n_steps must be > 0");
  if (a->brick_high == 0) return PHS_OK;
  for (int32_t p = 0; p < 2; ++p)
    if (int rc = ensure_graph(a, n_steps, p)) return rc;
  return PHS_OK;
}

int phs_arena_step_graph(phs_arena* a, int32_t n_steps, void* stream) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (n_steps <= 0) return phs::fail(PHS_ERR_ARG, "n_steps must be > 0");
  if (a->brick_high == 0) return PHS_OK;
  hipStream_t s = (hipStream_t)stream;
  const int32_t par0 = a->par;
  if (int rc = ensure_graph(a, n_steps, par0)) return rc;
  PHS_HIP_CHECK(hipGraphLaunch(a->g[par0].exec, s));
  a->par ^= n_steps & 1;
  for (auto& h : a->sh)
    if (h.used) h.steps_done = std::min<int64_t>(h.n_steps, h.steps_done + n_steps);
  return PHS_OK;
}

int phs_arena_step_marked(phs_arena* a, void* stream) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  if (a->brick_high == 0) return phs::fail(PHS_ERR_STATE, "arena is empty");
  hipStream_t s = (hipStream_t)stream;
  if (!a->marks[0])
    for (auto& e : a->marks) PHS_HIP_CHECK(hipEventCreate(&e));
  a->marking = true;
  PHS_HIP_CHECK(hipEventRecord(a->marks[0], s));
  a->n_marks = 1;
  launch_step(a, s);
  a->marking = false;
  PHS_HIP_CHECK(hipGetLastError());
  a->marks_valid = true;
  for (auto& h : a->sh)
    if (h.used) h.steps_done = std::min<int64_t>(h.n_steps, h.steps_done + 1);
  return PHS_OK;
}

int phs_arena_step_timing(phs_arena* a, int32_t cap, double* ms, int32_t* kinds, int32_t* n) {
  if (!a || !n || cap < 0 || (cap > 0 && (!ms || !kinds)))
    return phs::fail(PHS_ERR_ARG, "invalid arguments");
  if (!a->marks_valid || a->n_marks < 2)
    return phs::fail(PHS_ERR_STATE, "no marked step has been launched");
  const int launches = a->n_marks - 1;
  PHS_HIP_CHECK(hipEventSynchronize(a->marks[launches]));
  for (int i = 0; i < launches && i < cap; ++i) {
    float t = 0.0f;
    PHS_HIP_CHECK(hipEventElapsedTime(&t, a->marks[i], a->marks[i + 1]));
    ms[i] = (double)t;
    kinds[i] = a->kinds[i];
  }
  *n = launches;
  return PHS_OK;
}

int phs_arena_slot_info(phs_arena* a, int32_t slot, int64_t* steps_done, int32_t* n_voxel,
                        int32_t* n_bricks, int64_t* voxel_offset) {
  if (!a || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  const SlotHost& h = a->sh[slot];
  if (steps_done) *steps_done = h.steps_done;
  if (n_voxel) *n_voxel = h.n_voxel;
  if (n_bricks) *n_bricks = h.n_bricks;
  if (voxel_offset) *voxel_offset = h.voxel_off;
  return PHS_OK;
}

int phs_arena_perm(phs_arena* a, int32_t slot, int32_t* perm_host) {
  if (!a || !perm_host || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  const SlotHost& h = a->sh[slot];
  std::memcpy(perm_host, h.perm.data(), 4 * h.perm.size());
  return PHS_OK;
}

int phs_arena_state(phs_arena* a, int32_t slot, float* u, float* h, void* stream) {
  if (!a || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  const SlotHost& sl = a->sh[slot];
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((sl.n_voxel + 255) / 256, 1);
  if (u) k_scatter_f32<<<grid, 256, 0, s>>>(u, a->U + sl.voxel_off, sl.perm_dev, 1, sl.n_voxel);
  if (h) k_scatter_f32<<<grid, 256, 0, s>>>(h, a->H + sl.voxel_off, sl.perm_dev, 1, sl.n_voxel);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

int phs_arena_slot_status(phs_arena* a, int32_t slot, int64_t* first_nonfinite_step,
                          void* stream) {
  if (!a || !first_nonfinite_step || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  int32_t v = INT32_MAX;
  hipStream_t s = (hipStream_t)stream;
  PHS_HIP_CHECK(hipMemcpyAsync(&v, a->NONFINITE + slot, sizeof v, hipMemcpyDeviceToHost, s));
  PHS_HIP_CHECK(hipStreamSynchronize(s));
  *first_nonfinite_step = v == INT32_MAX ? -1 : (int64_t)v;
  return PHS_OK;
}

int phs_arena_slot_bounds(phs_arena* a, int32_t slot, double* lmin, double* lmax) {
  if (!a || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  if (lmin) *lmin = a->sh[slot].lmin;
  if (lmax) *lmax = a->sh[slot].lmax;
  return PHS_OK;
}

int phs_arena_last_residual(phs_arena* a, int32_t slot, double* rel, void* stream) {
  if (!a || !rel || slot < 0 || slot >= a->cfg.max_sims || !a->sh[slot].used)
    return phs::fail(PHS_ERR_ARG, "invalid slot");
  if (!a->cfg.diag) return phs::fail(PHS_ERR_STATE, "arena created without diag");
  hipStream_t s = (hipStream_t)stream;
  double zero[2] = {0.0, 0.0};
  PHS_HIP_CHECK(hipMemcpyAsync(a->diag_acc, zero, 16, hipMemcpyHostToDevice, s));
  launch_diag(a, slot, s);
  double acc[2];
  PHS_HIP_CHECK(hipMemcpyAsync(acc, a->diag_acc, 16, hipMemcpyDeviceToHost, s));
  PHS_HIP_CHECK(hipStreamSynchronize(s));
  if (acc[1] > 0)
    *rel = std::sqrt(acc[0] / acc[1]);
  else  // b = 0 with x = 0 is an exact solve; anything else (NaN included) is not
    *rel = (acc[0] == 0.0 && acc[1] == 0.0) ? 0.0 : NAN;
  return PHS_OK;
}

int phs_arena_sweeps(phs_arena* a, int32_t* sweeps) {
  if (!a || !sweeps) return phs::fail(PHS_ERR_ARG, "NULL argument");
  *sweeps = a->sweeps_eff;
  return PHS_OK;
}

int phs_arena_profile(phs_arena* a, int32_t reps, double* ms, void* stream) {
  if (!a || !ms || reps <= 0) return phs::fail(PHS_ERR_ARG, "invalid profile arguments");
  if (a->brick_high == 0) return phs::fail(PHS_ERR_STATE, "arena is empty");
  PHS_HIP_CHECK(profile_kernels(a, reps, ms, (hipStream_t)stream));
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

int phs_arena_counts(phs_arena* a, int64_t* voxels, int64_t* bricks, int64_t* vf_voxels,
                     int64_t* launch_bricks) {
  if (!a) return phs::fail(PHS_ERR_ARG, "handle is NULL");
  int64_t v = 0, b = 0;
  for (auto& h : a->sh)
    if (h.used) {
      v += h.n_voxel;
      b += h.n_bricks;
    }
  if (voxels) *voxels = v;
  if (bricks) *bricks = b;
  if (vf_voxels) *vf_voxels = 0;
  for (auto& h : a->sh)
    if (h.used && h.needs_vf && vf_voxels) *vf_voxels += h.n_voxel;
  if (launch_bricks) *launch_bricks = (int64_t)grid_for(a) * kWaves * nb_per_wave(a);
  return PHS_OK;
}

int phs_arena_destroy(phs_arena* a) {
  if (!a) return PHS_OK;
  hipDeviceSynchronize();
  for (auto& h : a->sh) free_slot_host(h);
  reap(a, true);
  for (auto& G : a->g) {
    if (G.exec) hipGraphExecDestroy(G.exec);
    if (G.graph) hipGraphDestroy(G.graph);
  }
  if (a->cap_stream) hipStreamDestroy(a->cap_stream);
  for (auto& e : a->marks)
    if (e) hipEventDestroy(e);
  hipFree(a->bricks);
  hipFree(a->slots);
  hipFree(a->U);
  hipFree(a->H);
  hipFree(a->R0);
  hipFree(a->DA);
  hipFree(a->DB);
  hipFree(a->USTAR);
  hipFree(a->ALPHA);
  hipFree(a->D1);
  hipFree(a->D2);
  hipFree(a->DACC);
  hipFree(a->VF);
  hipFree(a->diag_acc);
  hipFree(a->NONFINITE);
  hipFree(a->move_tmp);
  for (auto& q : a->PV) hipFree(q);
  delete a;
  return PHS_OK;
}

int phs_scatter_rows_f32(float* dst, const float* src, const int32_t* perm, int64_t rows, int64_t n,
                         void* stream) {
  if (!dst || !src || !perm || rows < 0 || n < 0) return phs::fail(PHS_ERR_ARG, "bad arguments");
  if (rows == 0 || n == 0) return PHS_OK;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)std::min<int64_t>(rows, 65535));
  k_scatter_f32<<<grid, 256, 0, (hipStream_t)stream>>>(dst, src, perm, rows, n);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

int phs_scatter_rows_i32(int32_t* dst, const int32_t* src, const int32_t* perm, int64_t rows,
                         int64_t n, void* stream) {
  if (!dst || !src || !perm || rows < 0 || n < 0) return phs::fail(PHS_ERR_ARG, "bad arguments");
  if (rows == 0 || n == 0) return PHS_OK;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)std::min<int64_t>(rows, 65535));
  k_scatter_i32<<<grid, 256, 0, (hipStream_t)stream>>>(dst, src, perm, rows, n);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

}  // extern "C"
