// Device code of the bricked FP32 arena (included by sim_arena.hip only).
//
// A brick is 64 lattice cells handled by one wavefront (lane l <-> cell
// (l % BX, (l / BX) % BY, l / (BX*BY))).  The brick record (192 B, scalar loads) holds the
// brick's occupancy mask and compact start, its per-direction neighbour-presence lane masks
// and the shifted masks and gather bases of its six face-neighbour bricks, all precomputed on
// the host, so every lane's loads depend on that one record only, and all of them are issued
// before any value is used.  Presence masks are used as lane masks (inverse ballot);
// compact indices come from mbcnt.
// In-brick neighbour values move between lanes with ds_bpermute (__shfl); cross-brick
// values are gathered (lanes without a cross neighbour re-load their own value, which hits
// the same cache line and keeps the code branch-free).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "phs_internal.h"

namespace arena {

constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;
constexpr int kMaxSweeps = 64;

// Brick record, built on the host when a simulation is added; everything a wavefront needs
// to address its brick's stencil is here, so the kernels spend no scalar work on deriving it.
// Directions d: 0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z.
struct BrickRec {
  uint64_t mask;       // occupied cells
  int32_t start;       // compact index of the first occupied cell
  int32_t slot;        // simulation slot, -1 = unused brick
  uint64_t pin[6];     // lanes whose d-neighbour is an occupied cell of this brick
  uint64_t pcr[6];     // occupied lanes whose d-neighbour is an occupied cell of the
                       // d-neighbour brick
  uint64_t sh[6];      // the d-neighbour brick's mask moved so that bit l is the cell
                       // adjacent to lane l across the face (its mbcnt is the gather rank)
  int32_t nstart[6];   // gather base of each face-neighbour brick: its start, plus, for a
                       // - face, the number of its cells below the face-adjacent offset
  uint64_t tblock;     // lanes of temporary-block voxels (non-conductive while the schedule
                       // byte has the temporary-block bit).  Scar voxels need no mask: their
                       // presence bits are all clear, so their rows are the identity, as
                       // alpha = 0 makes them.
};
static_assert(sizeof(BrickRec) == 192, "brick record layout");

// The values of a simulation that change from step to step.  A slot keeps two copies: the
// kernels of a step read copy `par` (the parity of the arena's launched-step count), and the
// step's first kernel writes the next step's values into copy par ^ 1, which no kernel of the
// step reads, so no separate advance launch is needed.
struct SlotDyn {
  int64_t step;       // steps done
  int64_t cur_off;    // element offset of the sample row written by this step in the sample
                      // buffers (row modulo the ring length, times n_voxel), -1 = none
  int32_t live;       // resident and step < n_steps
  int32_t cur_sched;  // schedule byte of this step
  int32_t next_sched; // schedule byte of the next step (its reaction runs in this step's
                      // last sweep), 0 after the last step
  int32_t pad;
};

struct SlotRec {
  SlotDyn dyn[2];
  int64_t n_steps;
  int64_t voxel_off;
  const uint8_t* sched;
  float* su;
  float* sh;
  int32_t* lat;
  int32_t n_voxel;
  int32_t model;
  int32_t nspm;
  int32_t n_samples;
  int32_t max_beats;
  int32_t needs_vf;   // scar or temporary-block voxels present
  int32_t hetero;     // per-voxel alpha in the ALPHA array
  int32_t ring;       // > 0: sample s is written to row s % ring
  int32_t nsw;        // sweeps of this simulation's first CN solve (<= the launched count)
  int32_t hparams;    // per-voxel cell parameters in the PV arrays (general kernels only)
  float p[5];       // MS: 1/tau_open, 1/tau_close, 1/tau_in, 1/tau_out, v_gate; AP: k,a,eps0,mu1,mu2
  float dt;
  float alpha;
  float lat_thr;
  double alpha64;
  float coef[kMaxSweeps][4];  // a, b, g, -; coef[0][2] = 1/theta
};

struct ArenaDev {
  // The brick table, its length and the order bits first: one 16-byte kernel-argument load
  // gives every kernel what it needs to address its brick record.
  const BrickRec* __restrict__ bricks;
  int32_t n_bricks;
  int32_t flags;  // bit 0: XCD remap; bit 1: launch parity of this step (the slots' copy of
                  // the per-step values); bit 2: walk the bricks backwards
  SlotRec* slots;
  float* U;
  float* H;
  float* UO;  // reaction outputs (= U, H except when the kernel is timed in isolation)
  float* HO;
  float* R0;
  double* DACC;
  float* USTAR;
  uint8_t* VF;
  const float* ALPHA;
  float* D1;  // CN correction of the previous step (the warm start with guess 1)
  float* D2;  // guess 2: the next step's warm start 2 D1 - (correction of the step before),
              // formed in FP32 when D1 is written
  const float* PV[5];  // per-voxel cell parameters (same encoding as SlotRec::p) of the
                       // simulations with hparams, or null
  int32_t* NONFINITE;  // per slot: first step whose new u is not finite (INT32_MAX = none)
  int32_t diag;
  int32_t max_sims;  // slots to advance
};

// Record prefetch (flags bits 8-23: a distance in bricks, 0 = none).  Each wavefront loads
// the record of the brick that distance ahead in its walk, which a wavefront of a later round
// of resident wavefronts processes, so that wavefront's record load hits L2 instead of waiting
// on HBM.  The loaded value is kept alive to the end of the kernel (keep_alive), whose
// s_endpgm waits for every outstanding memory operation anyway.
__device__ __forceinline__ void keep_alive(uint32_t v) { asm volatile("" ::"v"(v)); }


// The brick-order bits of flags (XCD remap, backwards walk) for first_brick.
__device__ __forceinline__ int xcd_of(const ArenaDev& A) { return A.flags & 5; }
__device__ __forceinline__ int par_of(const ArenaDev& A) { return (A.flags >> 1) & 1; }

enum { MODE_MID = 0, MODE_FINAL_X = 1, MODE_DACC_SET = 2, MODE_DACC_ADD = 3, MODE_DACC_ADD_X = 4 };

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Element access through a 32-bit byte offset from a wave-uniform base, so that the load or
// store uses the scalar-base addressing form (arena arrays stay below 2^32 bytes).
template <typename T>
__device__ __forceinline__ T ld(const T* base, int i) {
  return *(const T*)((const char*)base + (uint32_t)i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void stv(T* base, int i, T v) {
  *(T*)((char*)base + (uint32_t)i * (uint32_t)sizeof(T)) = v;
}

template <int NB>
__device__ __forceinline__ uint32_t prefetch_record(const ArenaDev& A, int b0) {
  const int dist = (A.flags >> 8) & 0xffff;
  const int t = (A.flags & 4) ? b0 - dist : b0 + dist;
  const int l = (int)(threadIdx.x & 63);
  uint32_t v = 0;
  if (dist != 0 && t >= 0 && t + NB <= A.n_bricks && l < 48 * NB)
    v = ((const uint32_t*)(A.bricks + t))[l];
  return v;
}

// Warm start of the CN correction d (x = u* + d): 0, the previous step's correction d_{n-1}
// (D1), or the linear extrapolation 2 d_{n-1} - d_{n-2}, kept ready in D2 so that the
// right-hand side gathers one array for it.
template <int GUESS>
__device__ __forceinline__ const float* guess_array(const ArenaDev& A) {
  return GUESS == 2 ? A.D2 : A.D1;
}

// Stores the step's total correction dt (own index i; o1 = the previous one, read from D1).
template <int GUESS>
__device__ __forceinline__ void keep_correction(const ArenaDev& A, int i, float o1, float dt) {
  if (GUESS == 2) stv(A.D2, i, 2.0f * dt - o1);
  stv(A.D1, i, dt);
}

__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// number of bits of m below this lane
__device__ __forceinline__ int mbcnt64(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// First brick of this wavefront.  With bit 0 of `order`, the round-robin dispatch of
// workgroups over the 8 XCDs is undone so that each XCD works on a contiguous run of bricks
// and the halo gathers hit that XCD's L2 (the grid is then a multiple of 8).  With bit 2 the
// bricks are walked backwards (each XCD its own run), so that a launch after a forward one
// starts on the bricks whose data that launch touched last, still in L2 and the memory-side
// cache.  The mappings are computed either way and selected, so that the kernel-argument
// loads issue together, not behind a branch.
template <int NB>
__device__ __forceinline__ int first_brick(int order) {
  const int blk = blockIdx.x;
  const int g = gridDim.x;
  const int per = g >> 3;
  const int fwd = (order & 1) ? (blk & 7) * per + (blk >> 3) : blk;
  const int bwd = (order & 1) ? (blk & 7) * per + (per - 1 - (blk >> 3)) : g - 1 - blk;
  const int b = (order & 4) ? bwd : fwd;
  return __builtin_amdgcn_readfirstlane((b * kWaves + (int)(threadIdx.x >> 6)) * NB);
}

// Lane patterns of a brick shape: INB[d] = lanes whose neighbour in direction d
// (0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z) lies in the same brick.
template <int BX, int BY, int BZ>
struct Shape {
  static_assert(BX * BY * BZ == 64, "64 cells per brick");
  static constexpr int B[3] = {BX, BY, BZ};
  static constexpr int St[3] = {1, BX, BX * BY};
  static constexpr uint64_t inb(int d) {
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l) {
      const int c[3] = {l % BX, (l / BX) % BY, l / (BX * BY)};
      const int ax = d >> 1;
      const bool in = (d & 1) ? c[ax] > 0 : c[ax] < B[ax] - 1;
      if (in) m |= 1ULL << l;
    }
    return m;
  }
  static constexpr int step(int d) { return St[d >> 1]; }
  static constexpr int wrap(int d) { return (B[d >> 1] - 1) * St[d >> 1]; }
};

// Per-brick, per-lane stencil addressing (everything derived from the brick record).
// Cross-brick values are gathered once per axis: a lane lies on at most one face of an axis
// whose brick extent is >= 2, so one load brings its + or - cross neighbour (or, on lanes with
// neither, its own value).  Axes of extent 1 (both faces on every lane) load both.
template <int BX, int BY, int BZ>
struct Stencil {
  using SH = Shape<BX, BY, BZ>;
  static constexpr bool both(int ax) { return SH::B[ax] == 1; }
  uint64_t pin[6];  // lanes whose d-neighbour is an occupied cell of this brick
  uint64_t pcr[6];  // occupied lanes whose d-neighbour is an occupied cell across the face
  int gaddr[6];     // gather addresses: [2 ax] per axis (+ face, - face or own index), and
                    // [2 ax + 1] the - face of an extent-1 axis
  int idx;          // compact index (the brick's first voxel on empty lanes)
  bool valid;

  // `br` is a register copy of the brick record (all of it loaded before any use).
  __device__ __forceinline__ void init(const BrickRec& br) {
    valid = lane_bit(br.mask);
    const int own = mbcnt64(br.mask);
    idx = br.start + (valid ? own : 0);
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      pin[d] = br.pin[d];
      pcr[d] = br.pcr[d];
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const int cp = br.nstart[2 * ax] + mbcnt64(br.sh[2 * ax]);
      const int cm = br.nstart[2 * ax + 1] + mbcnt64(br.sh[2 * ax + 1]);
      if (both(ax)) {
        gaddr[2 * ax] = lane_bit(pcr[2 * ax]) ? cp : idx;
        gaddr[2 * ax + 1] = lane_bit(pcr[2 * ax + 1]) ? cm : idx;
      } else {
        gaddr[2 * ax] = lane_bit(pcr[2 * ax]) ? cp : (lane_bit(pcr[2 * ax + 1]) ? cm : idx);
      }
    }
  }

  // g[d]: the gathered value for direction d (shared by the two directions of an axis).
  template <typename V>
  __device__ __forceinline__ void gather(const V* __restrict__ src, V (&g)[6]) const {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      g[2 * ax] = ld(src, gaddr[2 * ax]);
      g[2 * ax + 1] = both(ax) ? ld(src, gaddr[2 * ax + 1]) : g[2 * ax];
    }
  }

  // Value of the in-brick neighbour lane in direction D: a DPP row shift when the axis
  // stays inside one 16-lane row, ds_bpermute otherwise.
  template <int D, typename V>
  static __device__ __forceinline__ V exch(V v, int lane) {
    constexpr int ax = D >> 1;
    constexpr bool plus = !(D & 1);
    constexpr int stp = SH::St[ax];
    if constexpr (SH::B[ax] == 1) {
      return v;
    } else {
      static_assert(sizeof(V) % 4 == 0, "32-bit words");
      constexpr int W = sizeof(V) / 4;
      struct Words {
        int w[W];
      };
      Words in = __builtin_bit_cast(Words, v), out;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        if constexpr (stp * SH::B[ax] <= 16) {
          out.w[i] = __builtin_amdgcn_mov_dpp(in.w[i], plus ? 0x100 + stp : 0x110 + stp, 0xf, 0xf,
                                              false);
        } else {
          out.w[i] = __builtin_amdgcn_ds_bpermute(((lane + (plus ? stp : -stp)) & 63) << 2, in.w[i]);
        }
      }
      return __builtin_bit_cast(V, out);
    }
  }

  // sum over present face neighbours of (v_m - v_n) in type T; own value v, gathered values g
  // scaled by sc.  Absent neighbours give m = v.  All six lane exchanges are issued before
  // any is consumed.
  template <typename T, typename V>
  __device__ __forceinline__ T sum(V v, const V (&g)[6], V sc, int lane) const {
    const V pv[6] = {exch<0>(v, lane), exch<1>(v, lane), exch<2>(v, lane),
                     exch<3>(v, lane), exch<4>(v, lane), exch<5>(v, lane)};
    T s;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      V m = lane_bit(pcr[d]) ? g[d] * sc : v;
      m = lane_bit(pin[d]) ? pv[d] : m;
      const T t = (T)m - (T)v;
      s = d == 0 ? t : s + t;
    }
    return s;
  }

  // 2 sum(u_m - u_n) + sum(g_m - g_n) over present neighbours in one pass: per direction
  // fma(2, u_m - u_n, g_m - g_n) of FP32 differences, each exact or rounded relative to
  // itself.  An axis gather reads the cross neighbour on a face lane that has one and the
  // lane's own value otherwise (a lane lies on at most one face of an axis of extent > 1), so
  // its differences enter with no select (they are 0 on a lane without one); the in-brick
  // terms are selected by presence.
  template <typename V>
  __device__ __forceinline__ float sum_pair(V u, const V (&gu)[6], V g, const V (&gg)[6],
                                            int lane) const {
    const V pu[6] = {exch<0>(u, lane), exch<1>(u, lane), exch<2>(u, lane),
                     exch<3>(u, lane), exch<4>(u, lane), exch<5>(u, lane)};
    const V pg[6] = {exch<0>(g, lane), exch<1>(g, lane), exch<2>(g, lane),
                     exch<3>(g, lane), exch<4>(g, lane), exch<5>(g, lane)};
    float s = __builtin_fmaf(2.0f, gu[0] - u, gg[0] - g);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (ax > 0) s += __builtin_fmaf(2.0f, gu[2 * ax] - u, gg[2 * ax] - g);
      if (both(ax)) s += __builtin_fmaf(2.0f, gu[2 * ax + 1] - u, gg[2 * ax + 1] - g);
    }
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if (SH::B[d >> 1] > 1)
        s += lane_bit(pin[d]) ? __builtin_fmaf(2.0f, pu[d] - u, pg[d] - g) : 0.0f;
    return s;
  }

  // The same sum formed as (sum of terms) - n v in T, with one select per in-brick direction
  // only.  An axis gather reads the cross neighbour on a face lane that has one and the lane's
  // own value otherwise (a lane lies on at most one face of an axis of extent > 1), so
  // (g sc - v) is the axis' cross term as it stands: g sc and v are the same product on a lane
  // without one.  The in-brick terms select the neighbour lane or v.  n counts the terms
  // (9 for 4 x 4 x 4 bricks).  For T = double and V = float the rounding is ~1e-16 of |v|,
  // far below the FP32 rounding of the right-hand side it feeds.
  template <typename T, typename V>
  __device__ __forceinline__ T sum_total(V v, const V (&g)[6], V sc, int lane) const {
    const V pv[6] = {exch<0>(v, lane), exch<1>(v, lane), exch<2>(v, lane),
                     exch<3>(v, lane), exch<4>(v, lane), exch<5>(v, lane)};
    T s = (T)(g[0] * sc);
    constexpr int n = 3 + both(0) + both(1) + both(2) + 2 * ((SH::B[0] > 1) + (SH::B[1] > 1) +
                                                          (SH::B[2] > 1));
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (ax > 0) s += (T)(g[2 * ax] * sc);
      if (both(ax)) s += (T)(g[2 * ax + 1] * sc);
    }
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if (SH::B[d >> 1] > 1) s += (T)(lane_bit(pin[d]) ? pv[d] : v);
    if constexpr (sizeof(T) == 4)
      return __builtin_fmaf(-(float)n, (float)v, s);
    else
      return __builtin_fma(-(double)n, (double)v, s);
  }
};

// Kernel flavours.  The lean ones run when no simulation in the arena has heterogeneous
// alpha or per-voxel cell parameters: they drop the per-voxel loads, and each keeps only the
// tests its arena needs (the scalar and vector work they cost is a measurable share of a
// sweep, which is bound by instruction issue as much as by memory): the temporary-block test
// (FLAGS; scar needs no test) and the identity-sweep logic of simulations whose own sweep
// count is below the launched one (IDENT).
enum {
  FLAV_GENERAL = 0,
  FLAV_LEAN_FLAGS = 1,
  FLAV_LEAN = 2,
  FLAV_LEAN_IDENT = 3,
  FLAV_LEAN_FLAGS_IDENT = 4
};
constexpr bool flav_hetero(int f) { return f == FLAV_GENERAL; }
constexpr bool flav_tb(int f) {
  return f == FLAV_GENERAL || f == FLAV_LEAN_FLAGS || f == FLAV_LEAN_FLAGS_IDENT;
}
constexpr bool flav_ident(int f) {
  return f == FLAV_GENERAL || f == FLAV_LEAN_IDENT || f == FLAV_LEAN_FLAGS_IDENT;
}

// Wave-uniform slot values a stencil kernel needs, loaded together.
// Per-simulation sweep counts: a solve of K launched sweeps carries a simulation whose own
// count is nsw < K by identity sweeps (x_{k+1} = x_k) from launch nsw on, so it computes
// exactly what it computes alone.  Identity launches other than the last one only move the
// value between the two ping-pong buffers, and only when the parity of the remaining
// launches needs it: the simulation's waves skip every launch in [nsw, K-2] except launch
// nsw when K - 1 - nsw is odd.  K = 0 (refinement solves, timing) disables all of this.
struct SlotView {
  int live, needs_vf, hetero, cur_sched;
  bool ident, skip;
  float alpha, inv_theta, ca, cb, cg;
  double alpha64;
  template <int FLAV = FLAV_GENERAL>
  __device__ __forceinline__ void load(const SlotRec& S, int par, int k, int K = 0,
                                       bool last = true) {
    if (!flav_ident(FLAV)) {
      ident = skip = false;
    } else {
      const int nsw = S.nsw;
      ident = K > 0 && k >= nsw;
      skip = ident && !last && !(k == nsw && ((K - 1 - nsw) & 1));
    }
    live = S.dyn[par].live;
    needs_vf = S.needs_vf;
    hetero = S.hetero;
    cur_sched = S.dyn[par].cur_sched;
    alpha = S.alpha;
    alpha64 = S.alpha64;
    inv_theta = S.coef[0][2];
    ca = S.coef[k][0];
    cb = S.coef[k][1];
    cg = S.coef[k][2];
  }
};

// The per-voxel inputs of alpha: a heterogeneous value (general flavour, loaded in the
// kernel's load phase with the stencil values) and the temporary-block membership, which the
// brick record carries as a lane mask.  Scar rows need neither: their presence bits are clear,
// so their neighbour sums are 0 and their rows the identity, exactly as alpha = 0 gives.
struct AlphaIn {
  float al;
  bool tblock;
  template <int FLAV>
  __device__ __forceinline__ void load(const ArenaDev& A, const SlotView& V, int idx,
                                       const BrickRec& br) {
    al = 0.0f;
    tblock = false;
    if (flav_hetero(FLAV) && V.hetero) al = ld(A.ALPHA, idx);
    if (flav_tb(FLAV)) tblock = lane_bit(br.tblock);
  }
};

template <typename T, int FLAV = FLAV_GENERAL>
__device__ __forceinline__ T alpha_at(const SlotView& V, const AlphaIn& in) {
  T a = sizeof(T) == 8 ? (T)V.alpha64 : (T)V.alpha;
  if (flav_hetero(FLAV) && V.hetero) a = (T)in.al;
  if (flav_tb(FLAV) && (V.cur_sched & PHS_SCHED_TEMP_BLOCK) && in.tblock) a = (T)0;
  return a;
}

// Cell-model right-hand side (simulation/simulation_gpu.py:27-43) in FP32, branch-free:
// MODEL 0 Mitchell-Schaeffer (p = 1/tau_open, 1/tau_close, 1/tau_in, 1/tau_out, v_gate),
// MODEL 1 Aliev-Panfilov (p = k, a, eps0, mu1, mu2).
template <int MODEL>
__device__ __forceinline__ void rhs32(const float (&p)[5], float u, float h, float J, float& du,
                                      float& dh) {
  if (MODEL == 0) {
    du = h * u * u * (1.0f - u) * p[2] - u * p[3] + J;
    const float open = (1.0f - h) * p[0], close = -h * p[1];
    dh = u < p[4] ? open : close;
  } else {
    const float k = p[0], a = p[1];
    du = -k * u * (u - a) * (u - 1.0f) - u * h + J;
    dh = (p[2] + p[3] * h * __builtin_amdgcn_rcpf(u + p[4])) * (-h - k * u * (u - a - 1.0f));
  }
}

template <int MODEL>
__device__ __forceinline__ void rk4_model(const float (&p)[5], float dt, float J, float u, float h,
                                          float& uo, float& ho) {
  float k1u, k1h, k2u, k2h, k3u, k3h, k4u, k4h;
  rhs32<MODEL>(p, u, h, J, k1u, k1h);
  rhs32<MODEL>(p, u + 0.5f * dt * k1u, h + 0.5f * dt * k1h, J, k2u, k2h);
  rhs32<MODEL>(p, u + 0.5f * dt * k2u, h + 0.5f * dt * k2h, J, k3u, k3h);
  rhs32<MODEL>(p, u + dt * k3u, h + dt * k3h, J, k4u, k4h);
  const float s6 = dt * (1.0f / 6.0f);
  uo = u + s6 * (k1u + 2.0f * k2u + 2.0f * k3u + k4u);
  ho = h + s6 * (k1h + 2.0f * k2h + 2.0f * k3h + k4h);
}

// One RK4 step of the cell model (simulation/simulation_gpu.py:71-130) in FP32; the stimulus
// J is 1 on the S1 (S2) sites while the schedule byte sc has the S1 (S2) bit.  The model is
// wave-uniform: one scalar branch per call.
__device__ __forceinline__ void rk4(const float (&p)[5], float dt, int model, int sc, uint8_t f,
                                    float u, float h, float& uo, float& ho) {
  const float J = (((sc & PHS_SCHED_S1) && (f & PHS_VF_S1)) ||
                   ((sc & PHS_SCHED_S2) && (f & PHS_VF_S2)))
                      ? 1.0f
                      : 0.0f;
  if (model == 0)
    rk4_model<0>(p, dt, J, u, h, uo, ho);
  else
    rk4_model<1>(p, dt, J, u, h, uo, ho);
}

// The slot values the last sweep needs to finish a voxel.  They are read with the other
// wave-uniform values, before the kernel's first store: the stores could alias the slot
// table as far as the compiler knows, so reads after them become vector loads, each a
// memory round trip of its own.
struct SlotFinish {
  int64_t step, n_steps, voxel_off, cur_off;
  float* su;
  float* sh;
  int32_t* lat;
  int32_t n_voxel, max_beats, next_sched, model, hparams, slot;
  float p[5];
  float dt, lat_thr;
  __device__ __forceinline__ void load(const SlotRec& S, int par, int slot_id) {
    hparams = S.hparams;
    slot = slot_id;
    step = S.dyn[par].step;
    n_steps = S.n_steps;
    voxel_off = S.voxel_off;
    cur_off = S.dyn[par].cur_off;
    su = S.su;
    sh = S.sh;
    lat = S.lat;
    n_voxel = S.n_voxel;
    max_beats = S.max_beats;
    next_sched = S.dyn[par].next_sched;
    model = S.model;
#pragma unroll
    for (int i = 0; i < 5; ++i) p[i] = S.p[i];
    dt = S.dt;
    lat_thr = S.lat_thr;
  }
};

// The reaction of a step on its own: the first step of a newly added simulation (bricks
// [b_lo, b_hi) only), and the timing of the reaction in isolation.  Every later step's
// reaction runs inside the previous step's last sweep.  Loads are issued on the brick
// record alone; the slot's liveness is checked before anything is stored.
template <int NB>
__global__ void __launch_bounds__(kBlock) k_reaction(ArenaDev A, int b_lo, int b_hi) {
  const int b0 = first_brick<NB>(xcd_of(A)) + b_lo;
  float u[NB], h[NB];
  uint8_t f[NB];
  int idx[NB];
  bool on[NB], valid[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int b = b0 + j;
    on[j] = b < b_hi && A.bricks[b].slot >= 0;
    if (on[j]) {
      const BrickRec& br = A.bricks[b];
      valid[j] = lane_bit(br.mask);
      idx[j] = br.start + (valid[j] ? mbcnt64(br.mask) : 0);
      u[j] = ld(A.U, idx[j]);
      h[j] = ld(A.H, idx[j]);
      f[j] = ld(A.VF, idx[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (!on[j]) continue;
    const SlotRec& S = A.slots[A.bricks[b0 + j].slot];
    if (!S.dyn[par_of(A)].live || !valid[j]) continue;
    float p[5] = {S.p[0], S.p[1], S.p[2], S.p[3], S.p[4]};
    if (S.hparams)
#pragma unroll
      for (int k = 0; k < 5; ++k) p[k] = ld(A.PV[k], idx[j]);
    float uo, ho;
    rk4(p, S.dt, S.model, S.dyn[par_of(A)].cur_sched, f[j], u[j], h[j], uo, ho);
    stv(A.UO, idx[j], uo);
    stv(A.HO, idx[j], ho);
  }
}

// The next step's per-step values of slot S (copy par ^ 1) from this step's (copy par): the
// step counter, the schedule bytes and the sample row.  A finished or unused slot's values are
// carried over unchanged, so that its copies never disagree.  Every read comes before the
// first store.
__device__ __forceinline__ void advance_slot(SlotRec& S, int par) {
  const SlotDyn cur = S.dyn[par];
  const int64_t t = cur.step + 1;
  const int64_t n_steps = S.n_steps;
  const uint8_t* sched = S.sched;
  const int nspm = S.nspm, n_samples = S.n_samples, ring = S.ring, n_voxel = S.n_voxel;
  SlotDyn nx = cur;
  if (cur.live) {
    const bool more = t < n_steps;
    nx.step = t;
    nx.live = more ? 1 : 0;
    if (more) {
      nx.cur_sched = sched[t];
      nx.next_sched = t + 1 < n_steps ? sched[t + 1] : 0;
      int64_t row = -1;
      if (nspm > 0) {
        if (t % nspm == 0) row = t / nspm;
      } else {
        row = t;
      }
      nx.cur_off = row >= 0 && row < n_samples ? (ring > 0 ? row % ring : row) * n_voxel : -1;
    }
  }
  S.dyn[par ^ 1] = nx;
}

// Right-hand side of the correction solve A e = r, x = u* + g + e, with the warm start g:
//   r = dt L u* - A g = 2 alpha sum (u*_m - u*_n) - (g_n - alpha sum (g_m - g_n)),
// formed in FP32 from the neighbour differences: each u*_m - u*_n is exact, or rounded
// relative to the difference itself, so r carries FP32 rounding of its own size, as the
// sweeps that follow do (no cancellation against |u*|).
template <int BX, int BY, int BZ, int NB, int GUESS, int FLAV>
__global__ void __launch_bounds__(kBlock) k_rhs(ArenaDev A) {
  const int b0 = first_brick<NB>(xcd_of(A));
  const uint32_t pfr = prefetch_record<NB>(A, b0);
  const int lane = lane_id();
  Stencil<BX, BY, BZ> st[NB];
  SlotView sv[NB];
  bool on[NB];
  float un[NB], gn0[NB];
  float gu[NB][6], gg[NB][6];
  AlphaIn ain[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int b = b0 + j;
    BrickRec br;
    on[j] = false;
    if (b < A.n_bricks) {
      br = A.bricks[b];
      on[j] = br.slot >= 0;
    }
    if (on[j]) {
      st[j].init(br);
      sv[j].template load<FLAV>(A.slots[br.slot], par_of(A), 0);
      un[j] = ld(A.U, st[j].idx);
      st[j].gather(A.U, gu[j]);
      if (GUESS >= 1) {
        const float* G = guess_array<GUESS>(A);
        gn0[j] = ld(G, st[j].idx);
        st[j].gather(G, gg[j]);
      }
      ain[j].template load<FLAV>(A, sv[j], st[j].idx, br);
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (!on[j]) continue;
    const float uv = st[j].valid ? un[j] : 0.0f;
    // r0 = a (2 sum(u*) + sum(d0)) - d0: the two neighbour sums in one pass
    float s2, gn = 0.0f;
    if (GUESS >= 1) {
      gn = st[j].valid ? gn0[j] : 0.0f;
      s2 = st[j].sum_pair(uv, gu[j], gn, gg[j], lane);
    } else {
      s2 = 2.0f * st[j].template sum<float>(uv, gu[j], 1.0f, lane);
    }
    if (!st[j].valid || !sv[j].live) continue;
    const float a = alpha_at<float, FLAV>(sv[j], ain[j]);
    stv(A.R0, st[j].idx, a * s2 - gn);
  }
  keep_alive(pfr);
  // The first workgroup's first wavefront also stages every slot's next-step values.  This
  // comes after all of the kernel's loads: a store to the slot table before them would leave
  // the compiler unable to keep the brick records on the scalar path.
  if (blockIdx.x == 0 && threadIdx.x < 64)
    for (int s = (int)threadIdx.x; s < A.max_sims; s += 64) advance_slot(A.slots[s], par_of(A));
}

// Last pass of a step: activation time and samples of the post-step state x, then the next
// step's reaction, so that U holds u* of the next step between steps (x itself after a
// simulation's last step).  f and hv are the voxel's flags and gate, loaded by the caller
// together with its stencil values.
__device__ __forceinline__ void finish_voxel(const ArenaDev& A, const SlotFinish& S, int idx,
                                             float x, uint8_t f, float hv,
                                             const float (&p)[5]) {
  if (!__builtin_isfinite(x)) atomicMin(A.NONFINITE + S.slot, (int32_t)S.step);
  const bool now = x >= S.lat_thr;
  const bool above = (f & PHS_VF_ABOVE) != 0;
  const int64_t local = (int64_t)idx - S.voxel_off;
  if (now && !above) {
    int bt = (f & PHS_VF_BEAT_MASK) >> PHS_VF_BEAT_SHIFT;
    if (bt < S.max_beats && S.lat) S.lat[(int64_t)bt * S.n_voxel + local] = (int32_t)S.step;
    bt = bt < 7 ? bt + 1 : 7;
    A.VF[idx] = (uint8_t)((f & 0x0F) | (bt << PHS_VF_BEAT_SHIFT) | PHS_VF_ABOVE);
  } else if (!now && above) {
    A.VF[idx] = (uint8_t)(f & ~PHS_VF_ABOVE);
  }
  if (S.cur_off >= 0) {
    const int64_t o = S.cur_off + local;
    if (S.su) S.su[o] = x;
    if (S.sh) S.sh[o] = hv;
  }
  if (S.step + 1 < S.n_steps) {
    float uo, ho;
    rk4(p, S.dt, S.model, S.next_sched, f, x, hv, uo, ho);
    A.U[idx] = uo;
    A.H[idx] = ho;
  } else {
    A.U[idx] = x;
  }
}

// One Chebyshev sweep d_{k+1} = a_k d_k + b_k d_{k-1} + g_k (r - A d_k), A d = d - alpha sum.
// FIRST (k = 1): d_k = r / theta, read from R0.  prev_kind 0: d_{k-1} = 0; 1: r / theta;
// 2: prev[].  GUESS: warm start of the correction (see guess_array); the final modes add it
// back and keep the step's total correction in D1 (and, with guess 2, the next step's
// extrapolated warm start in D2).
template <int BX, int BY, int BZ, int MODE, int NB, int GUESS, bool FIRST, int FLAV>
__global__ void __launch_bounds__(kBlock) k_sweep(ArenaDev A, int k, int K, const float* cur,
                                                  const float* prev, int prev_kind, float* out) {
  const int b0 = first_brick<NB>(xcd_of(A));
  const uint32_t pfr = prefetch_record<NB>(A, b0);
  const int lane = lane_id();
  const float* src = FIRST ? A.R0 : cur;
  Stencil<BX, BY, BZ> st[NB];
  SlotView sv[NB];
  bool on[NB];
  float dn[NB], r0[NB], dp[NB];
  float g6[NB][6];
  // the last sweep's own-voxel state (u*, h, flags, previous corrections), loaded with the
  // stencil values so that finishing the step waits on no second round trip
  constexpr bool kFinish = MODE == MODE_FINAL_X || MODE == MODE_DACC_ADD_X;
  float pu[NB], ph[NB], p1[NB], p2[NB];
  uint8_t pf[NB];
  SlotFinish sf[NB];
  float pp[NB][5];  // the voxel's cell parameters (the slot's, or per voxel)
  AlphaIn ain[NB];
  // the refinement modes' accumulator (or warm start), likewise loaded up front
  double pd[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int b = b0 + j;
    BrickRec br;
    on[j] = false;
    if (b < A.n_bricks) {
      br = A.bricks[b];
      on[j] = br.slot >= 0;
    }
    if (on[j]) {
      // the loads are issued before the skip test so that they depend on the brick record
      // only; a skipped wave drops them
      st[j].init(br);
      sv[j].template load<FLAV>(A.slots[br.slot], par_of(A), k, K, MODE != MODE_MID);
      const int i = st[j].idx;
      dn[j] = ld(src, i);
      r0[j] = ld(A.R0, i);
      dp[j] = !FIRST && prev_kind == 2 ? ld(prev, i) : 0.0f;
      st[j].gather(src, g6[j]);
      ain[j].template load<FLAV>(A, sv[j], i, br);
      if (MODE == MODE_DACC_ADD || MODE == MODE_DACC_ADD_X) pd[j] = A.DACC[i];
      if (MODE == MODE_DACC_SET) pd[j] = GUESS >= 1 ? (double)guess_array<GUESS>(A)[i] : 0.0;
      if (kFinish) {
        sf[j].load(A.slots[br.slot], par_of(A), br.slot);
        pu[j] = ld(A.U, i);
        ph[j] = ld(A.H, i);
        pf[j] = ld(A.VF, i);
        p1[j] = GUESS >= 1 ? ld(A.D1, i) : 0.0f;
        p2[j] = GUESS == 2 ? ld(A.D2, i) : 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) pp[j][k] = sf[j].p[k];
        if (flav_hetero(FLAV) && sf[j].hparams)
#pragma unroll
          for (int k = 0; k < 5; ++k) pp[j][k] = ld(A.PV[k], i);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (!on[j] || sv[j].skip) continue;
    const float inv_theta = sv[j].inv_theta;
    const float sc = FIRST ? inv_theta : 1.0f;
    const float d = st[j].valid ? dn[j] * sc : 0.0f;
    // d is the CN correction, so the rounding of (sum of terms) - n d (a few ulp of |d|)
    // stays far below the sweep's own truncation error; this form takes fewer VALU
    // instructions than the sum of differences
    const float s = st[j].template sum_total<float>(d, g6[j], sc, lane);
    if (!st[j].valid || !sv[j].live) continue;
    const int i = st[j].idx;
    const float a = alpha_at<float, FLAV>(sv[j], ain[j]);
    const float Ad = d - a * s;
    const float dpv = FIRST ? 0.0f : prev_kind == 1 ? r0[j] * inv_theta : dp[j];
    const float dnew = sv[j].ident ? d : sv[j].ca * d + sv[j].cb * dpv + sv[j].cg * (r0[j] - Ad);
    if (MODE == MODE_MID) {
      stv(out, i, dnew);
    } else if (MODE == MODE_FINAL_X) {
      const float us = pu[j];
      float dt = dnew;
      if (GUESS >= 1) {
        const float o1 = p1[j];
        dt += GUESS == 2 ? p2[j] : o1;
        keep_correction<GUESS>(A, i, o1, dt);
      }
      if (A.diag) {
        A.USTAR[i] = us;
        A.DACC[i] = (double)dt;
      }
      finish_voxel(A, sf[j], i, us + dt, pf[j], ph[j], pp[j]);
    } else if (MODE == MODE_DACC_SET) {
      A.DACC[i] = pd[j] + (double)dnew;
    } else if (MODE == MODE_DACC_ADD) {
      A.DACC[i] = pd[j] + (double)dnew;
    } else {
      const double acc = pd[j] + (double)dnew;
      const float us = pu[j];
      A.DACC[i] = acc;
      if (GUESS >= 1) keep_correction<GUESS>(A, i, p1[j], (float)acc);
      if (A.diag) A.USTAR[i] = us;
      finish_voxel(A, sf[j], i, (float)((double)us + acc), pf[j], ph[j], pp[j]);
    }
  }
  // s_endpgm waits for every outstanding memory operation, so this wait costs nothing extra
  keep_alive(pfr);
}

// FP64 residual of the accumulated correction: r = 2 alpha sum(u*_m - u*_n) - A dacc.
template <int BX, int BY, int BZ>
__global__ void __launch_bounds__(kBlock) k_residual(ArenaDev A) {
  const int b = first_brick<1>(xcd_of(A));
  if (b >= A.n_bricks || A.bricks[b].slot < 0) return;
  const int lane = lane_id();
  Stencil<BX, BY, BZ> st;
  const BrickRec br = A.bricks[b];
  st.init(br);
  SlotView sv;
  sv.load(A.slots[br.slot], par_of(A), 0);
  AlphaIn ain;
  ain.load<FLAV_GENERAL>(A, sv, st.idx, br);
  const float un = A.U[st.idx];
  const double dn = A.DACC[st.idx];
  float gu[6];
  double gd[6];
  st.gather((const float*)A.U, gu);
  st.gather((const double*)A.DACC, gd);
  const double su = st.template sum<double>(st.valid ? un : 0.0f, gu, 1.0f, lane);
  const double sd = st.template sum<double>(st.valid ? dn : 0.0, gd, 1.0, lane);
  if (!st.valid || !sv.live) return;
  const double a = alpha_at<double>(sv, ain);
  A.R0[st.idx] = (float)(2.0 * a * su - (dn - a * sd));
}

__global__ void k_deactivate(SlotRec* slots, int slot) {
  slots[slot].dyn[0].live = 0;
  slots[slot].dyn[1].live = 0;
}

__global__ void k_invalidate(BrickRec* bricks, int64_t b0, int64_t nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) bricks[b0 + i].slot = -1;
}

// A simulation's voxel range moved by delta: its brick records' voxel indices (own start and
// gather bases) move with it.  A gather base of a face without a neighbour brick is never
// read (no lane of that face gathers), so it may move with the others.
__global__ void k_rebase(BrickRec* bricks, int64_t b0, int64_t nb, int32_t delta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  BrickRec& r = bricks[b0 + i];
  r.start += delta;
#pragma unroll
  for (int k = 0; k < 6; ++k) r.nstart[k] += delta;
}

__global__ void k_set_voxel_off(SlotRec* slots, int slot, int64_t off) {
  if (threadIdx.x == 0) slots[slot].voxel_off = off;
}

// Diagnostic: sum of r^2 and b^2 with b = u* + alpha sum(u*_m - u*_n), x = u* + dacc,
// r = b - (x - alpha sum(x_m - x_n)), for the last completed step of one slot.
template <int BX, int BY, int BZ>
__global__ void __launch_bounds__(kBlock) k_diag(ArenaDev A, int slot_sel, double* acc) {
  const int b = first_brick<1>(xcd_of(A));
  if (b >= A.n_bricks || A.bricks[b].slot != slot_sel) return;
  const SlotRec& S = A.slots[slot_sel];
  const int lane = lane_id();
  Stencil<BX, BY, BZ> st;
  st.init(A.bricks[b]);
  const float us = st.valid ? A.USTAR[st.idx] : 0.0f;
  const double dn = st.valid ? A.DACC[st.idx] : 0.0;
  float gu[6];
  double gd[6];
  st.gather((const float*)A.USTAR, gu);
  st.gather((const double*)A.DACC, gd);
  const double su = st.template sum<double>(us, gu, 1.0f, lane);
  const double sd = st.template sum<double>(dn, gd, 1.0, lane);
  double r2 = 0.0, b2 = 0.0;
  if (st.valid) {
    const int64_t done = S.dyn[par_of(A)].step;
    const uint8_t sc = done > 0 ? S.sched[done - 1] : 0;
    double a = S.hetero ? (double)A.ALPHA[st.idx] : S.alpha64;
    if (S.needs_vf) {
      const uint8_t f = A.VF[st.idx];
      if ((f & PHS_VF_SCAR) || ((sc & PHS_SCHED_TEMP_BLOCK) && (f & PHS_VF_TEMP_BLOCK))) a = 0.0;
    }
    const double bb = (double)us + a * su;
    const double r = bb - ((double)us + dn - a * (su + sd));
    r2 = r * r;
    b2 = bb * bb;
  }
  for (int o = 32; o > 0; o >>= 1) {
    r2 += __shfl_xor(r2, o);
    b2 += __shfl_xor(b2, o);
  }
  if (lane == 0) {
    atomicAdd(acc, r2);
    atomicAdd(acc + 1, b2);
  }
}

__global__ void k_scatter_f32(float* __restrict__ dst, const float* __restrict__ src,
                              const int32_t* __restrict__ perm, int64_t rows, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y)
    if (i < n) dst[r * n + perm[i]] = src[r * n + i];
}

__global__ void k_scatter_i32(int32_t* __restrict__ dst, const int32_t* __restrict__ src,
                              const int32_t* __restrict__ perm, int64_t rows, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y)
    if (i < n) dst[r * n + perm[i]] = src[r * n + i];
}

}  // namespace arena
