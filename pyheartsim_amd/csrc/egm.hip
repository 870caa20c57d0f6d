// Unipolar electrogram on MFMA.
//
// Reference: simulation/unipolar_electrogram_gpu.py:21-69 (one thread per (voxel, electrode),
// FP32 atomics per sample) fed by the host gradients of
// simulation/unipolar_electrogram_equation_parts.py:60-103 and the distance arrays of :41-58.
//
// The per-voxel dipole sum
//     egm[t, e] = sum_v c_v / l^3 * sum_k (D_v^T l)_k * g_k(v, t),
//     g_k(v, t) = (u(p_k(v), t) - u(m_k(v), t)) / (2 Delta)  when both axis neighbours exist
// is linear in u.  Moving the central-difference stencil onto the electrode side gives a lead
// field Wt (E x N),
//     Wt[e, w] = 1/(2 Delta) sum_k ( ok_k(m_k(w)) W_k[e, m_k(w)] - ok_k(p_k(w)) W_k[e, p_k(w)] ),
//     W_k[e, v] = c_v / l_ve^3 * (D_v^T l_ve)_k,  l_ve = r_v - x_e,  l = max(|l_ve|, 1 mm),
// built once per geometry in FP64, and the electrogram becomes one contraction
//     egm (T x E) = U (T x N) . Wt^T (N x E)
// over the 1 kHz samples U.  It runs on MFMA (v_mfma_f32_16x16x32_bf16) as BF16 hi/lo split
// products (hi*hi + hi*lo + lo*hi) with FP32 accumulation, split over K = N voxels with a
// fixed-order reduction (deterministic, unlike the reference's atomics).  Samples and lead
// field are staged through LDS per 64-deep K stage; the lead field is stored K-stage-major
// so that a stage's tile is one contiguous block, and workgroups sharing a K range are placed
// on one XCD so that its L2 serves the lead-field reuse across time tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "phs_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kTM = 128;  // time rows per workgroup
constexpr int kTE = 64;   // electrodes per workgroup
constexpr int kKS = 32;   // K per MFMA step of the FP32 contraction
constexpr int kK2 = 64;   // K of one LDS stage of the BF16 contraction (two 32-deep MFMA steps)

__device__ __forceinline__ unsigned short bf16_rn(float x) {
  unsigned int b = __float_as_uint(x);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}

__device__ __forceinline__ double w_axis(const double* __restrict__ vox, const double* __restrict__ el,
                                         const double* __restrict__ c, const double* __restrict__ D,
                                         int N, int v, int e, int k) {
  const double lx = vox[3 * (size_t)v] - el[3 * e];
  const double ly = vox[3 * (size_t)v + 1] - el[3 * e + 1];
  const double lz = vox[3 * (size_t)v + 2] - el[3 * e + 2];
  double l = sqrt(lx * lx + ly * ly + lz * lz);
  if (l < 1.0) l = 1.0;
  double proj;
  if (D) {
    // (D^T l)_k = sum_j l_j D_jk, D stored [9][N] row-major (D11, D12, D13, D21, ...)
    proj = lx * D[(size_t)(0 + k) * N + v] + ly * D[(size_t)(3 + k) * N + v] +
           lz * D[(size_t)(6 + k) * N + v];
  } else {
    proj = k == 0 ? lx : (k == 1 ? ly : lz);
  }
  return c[v] / (l * l * l) * proj;
}

// The lead-field column of voxel w is the transpose of the central-difference stencil:
//   Wt[e, w] = 1/(2 Delta) sum over the stencil terms that read u(w):
//     + W_k[e, v] for every v whose axis-k pair is complete and p_k(v) = w,
//     - W_k[e, v] for every v whose axis-k pair is complete and m_k(v) = w.
// The terms of w are listed on the host (inv_ptr / inv_v / inv_code = 2 k + (minus)), in the
// order axis k, then + before -, then v, so the sum is formed in a fixed order; for a
// symmetric table they are + W_k[m_k(w)] and - W_k[p_k(w)].
__global__ void build_leadfield(int N, int E, int Kpad, const double* __restrict__ vox,
                                const double* __restrict__ el, const double* __restrict__ c,
                                const double* __restrict__ D, const int* __restrict__ inv_ptr,
                                const int* __restrict__ inv_v,
                                const unsigned char* __restrict__ inv_code,
                                const int* __restrict__ perm, double inv2d, int Epad,
                                float* __restrict__ Wf, unsigned short* __restrict__ Wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (i >= Kpad) return;
  const size_t o = (size_t)e * Kpad + i;
  double acc = 0.0;
  if (i < N && e < E) {
    const int w = perm ? perm[i] : i;
    for (int q = inv_ptr[w]; q < inv_ptr[w + 1]; ++q) {
      const int code = inv_code[q];
      const double t = w_axis(vox, el, c, D, N, inv_v[q], e, code >> 1);
      acc = (code & 1) ? acc - t : acc + t;
    }
    acc *= inv2d;
  }
  const float f = (float)acc;
  const unsigned short hi = bf16_rn(f);
  Wf[o] = f;
  // BF16 planes in K-stage-major tiles [Kpad / 64][plane][Epad][64]: one K stage of every
  // electrode and both planes is one contiguous block
  const size_t tile = (size_t)(i / kK2) * 2 * Epad;
  Wt[(tile + e) * kK2 + i % kK2] = hi;
  Wt[(tile + Epad + e) * kK2 + i % kK2] = bf16_rn(f - bf16_to_f(hi));
}

// FP32 contraction on MFMA (v_mfma_f32_16x16x4f32), the accuracy reference of the BF16 form:
// workgroup 4 waves = 128 time rows x 64 electrodes, samples read straight from memory.
__global__ void __launch_bounds__(256) egm_gemm_f32(const float* __restrict__ S, int64_t T,
                                                    int64_t stride, int N, int Kpad, int k_chunk,
                                                    const float* __restrict__ Wf, int Epad,
                                                    float* __restrict__ ws) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wt = wave & 1, we = wave >> 1;
  const int64_t t0 = (int64_t)blockIdx.x * kTM + 64 * wt;
  const int e0 = blockIdx.y * kTE + 32 * we;
  const int split = blockIdx.z;
  const int kb = split * k_chunk;
  const int ke = min(N, kb + k_chunk);
  const int q = lane >> 4;   // k sub-block of 8
  const int r16 = lane & 15;

  f32x4 acc[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) acc[r][cc] = f32x4{0.f, 0.f, 0.f, 0.f};

  const float* arow[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int64_t t = t0 + 16 * r + r16;
    if (t > T - 1) t = T - 1;
    arow[r] = S + t * stride;
  }
  for (int kk = kb; kk < ke; kk += kKS) {
    const int kl = kk + 8 * q;
    float a[4][8];
    const bool full = kk + kKS <= N;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int j = 0; j < 8; ++j) a[r][j] = (full || kl + j < N) ? arow[r][kl + j] : 0.0f;
    }
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const size_t wo = (size_t)(e0 + 16 * cc + r16) * Kpad + kl;
      const float4 b0 = *(const float4*)(Wf + wo);
      const float4 b1 = *(const float4*)(Wf + wo + 4);
      const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          acc[r][cc] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][j], bv[j], acc[r][cc], 0, 0, 0);
    }
  }
  // D tile layout: lane holds column (electrode) r16, rows (time) 4*q + j.
  float* wsp = ws + (size_t)split * T * Epad;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int cc = 0; cc < 2; ++cc)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t t = t0 + 16 * r + 4 * q + j;
        if (t < T) wsp[t * Epad + e0 + 16 * cc + r16] = acc[r][cc][j];
      }
}

// BF16 split of a pair of samples for the MFMA A operand: hi = x rounded to BF16, lo = x - hi
// (exact in FP32) rounded to BF16, both with the hardware round-to-nearest-even conversion.
// hi + lo carries 17 significant bits of x.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned& hi, unsigned& lo) {
  const bf16x2 h = {(__bf16)x0, (__bf16)x1};
  hi = __builtin_bit_cast(unsigned, h);
  const float d0 = x0 - __uint_as_float(hi << 16);
  const float d1 = x1 - __uint_as_float(hi & 0xFFFF0000u);
  const bf16x2 l = {(__bf16)d0, (__bf16)d1};
  lo = __builtin_bit_cast(unsigned, l);
}

// Shape of the BF16 contraction per electrode-tile width CE (64 CE electrodes per workgroup):
//   RW   16-row time blocks per wave, TG groups of 4 waves across time per workgroup
//        (workgroup: 16 RW TG time rows), KS the K depth of one LDS stage;
//   BLDS the lead-field tile is staged through LDS (shared by the TG time groups), else each
//        wave reads its own lead-field fragments straight from the L2-resident tile.
// Up to 128 electrodes: RW 4, TG 2, 64-deep stages, lead field through LDS (the sample stream
// dominates).  256 electrodes: RW 8 (halves the lead-field reads per flop), 32-deep stages and
// one time group whose waves read their lead-field fragments directly, which keeps the 128
// accumulator registers, the staged samples and the operands within the 256 registers of a
// wave at 2 waves per SIMD (no spills) and the LDS at 20 KB, so that two workgroups share a
// CU and one's stage barrier and loads overlap the other's MFMAs.
template <int CE>
struct GemmCfg {
  static constexpr int RW = CE == 4 ? 8 : 4;
  static constexpr int TG = CE == 4 ? 1 : 2;
  static constexpr int KS = CE == 4 ? 32 : 64;
  static constexpr bool BLDS = CE != 4;
  static constexpr int DEPTH = CE == 4 ? 2 : 1;  // stages of samples in flight per workgroup
  static constexpr int THREADS = 256 * TG;
  static constexpr int TW = 16 * RW * TG;  // time rows per workgroup
};

// Staging registers of one K stage: KS-deep sample rows (FP32, split into BF16 hi/lo when
// written to LDS) and lead-field rows (BF16 planes, copied as they are).
template <int KTW, int EW, int NP, int KS, int THREADS, bool BLDS>
struct GemmStage {
  static constexpr int LPR = KS / 4;                           // float4 chunks per sample row
  static constexpr int RPI = 64 / LPR;                         // sample rows per wave load
  static constexpr int ACH = KTW * LPR / THREADS;              // sample chunks per thread
  static constexpr int BRC = KS / 8;                           // 16-byte chunks per lead-field row
  static constexpr int BCH = BLDS ? NP * EW * BRC / THREADS : 1;  // lead-field chunks per thread
  static constexpr int PA = KS + 8;  // LDS row pitch in bf16 (16 B pad against bank conflicts)
  static_assert(ACH >= 1 && ACH * THREADS == KTW * LPR, "sample staging");
  static_assert(!BLDS || (BCH >= 1 && BCH * THREADS == NP * EW * BRC), "lead-field staging");
  // Sample chunk i of a lane: 4 floats of row (wave ACH + i) RPI + lane / LPR at k offset
  // 4 (lane % LPR), so that one wave load covers RPI whole KS-deep row segments.
  // the staged values of one stage (a workgroup keeps DEPTH of them in flight)
  struct Data {
    float st[ACH][4];
    uint4 bst[BCH];
  };
  int srw[ACH];            // sample chunk: row in the tile
  int sk;                  // k offset of the lane's sample chunks in the stage
  const float* srow[ACH];  // their sample rows (rows past T repeat row T - 1, never written)
  int boff[BCH];           // lead-field chunk: element offset within a 64-deep Wt tile
  int bl[BCH];             // its LDS word offset (plane, row, k)

  __device__ __forceinline__ void init(int tid, const float* S, int64_t T, int64_t stride,
                                       int64_t tb, int Epad, int eb) {
    const int wave = tid >> 6, lane = tid & 63;
    sk = 4 * (lane % LPR);
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      srw[i] = (wave * ACH + i) * RPI + lane / LPR;
      int64_t t = tb + srw[i];
      if (t > T - 1) t = T - 1;
      srow[i] = S + t * stride;
    }
#pragma unroll
    for (int i = 0; i < BCH; ++i) {
      const int j = tid + THREADS * i;
      const int pl = j / (EW * BRC), rw = (j / BRC) % EW, k = 8 * (j % BRC);
      boff[i] = (pl * Epad + eb + rw) * kK2 + k;
      bl[i] = pl * (EW * PA / 2) + (rw * PA + k) / 2;
    }
  }

  // Loads the stage at k0 (a multiple of KS) into the staging registers; ke bounds the samples.
  // The lead field is zero-padded to whole 64-deep tiles, so its loads are always in bounds.
  // FULL: the whole stage is in range (k0 + KS <= ke), known at the call site, so that the
  // loop of full stages has no branch between these loads and their use (a branch merge makes
  // the compiler's wait for earlier loads also wait for these).
  template <bool FULL = false>
  __device__ __forceinline__ void load(Data& d, int k0, int ke,
                                       const unsigned short* __restrict__ Wt, int Epad) const {
    auto& st = d.st;
    auto& bst = d.bst;
    const int k = k0 + sk;
    if (FULL || k0 + KS <= ke) {  // whole stage in range (wave-uniform): plain 16-byte loads
#pragma unroll
      for (int i = 0; i < ACH; ++i) {
        float4 x;
        __builtin_memcpy(&x, srow[i] + k, 16);
        st[i][0] = x.x;
        st[i][1] = x.y;
        st[i][2] = x.z;
        st[i][3] = x.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < ACH; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) st[i][j] = k + j < ke ? srow[i][k + j] : 0.0f;
    }
    if (BLDS) {
      const unsigned short* tile = Wt + (size_t)(k0 / kK2) * 2 * Epad * kK2 + (k0 % kK2);
#pragma unroll
      for (int i = 0; i < BCH; ++i) bst[i] = *(const uint4*)(tile + boff[i]);
    }
  }

  // Writes the staged stage to LDS: sample hi (and lo) planes, lead-field planes.
  __device__ __forceinline__ void store(const Data& d, unsigned* __restrict__ la,
                                        unsigned* __restrict__ lb) const {
    const auto& st = d.st;
    const auto& bst = d.bst;
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      unsigned h0, l0, h1, l1;
      split_pair(st[i][0], st[i][1], h0, l0);
      split_pair(st[i][2], st[i][3], h1, l1);
      const int o = (srw[i] * PA + sk) / 2;
      *(uint2*)&la[o] = uint2{h0, h1};
      if (NP == 2) *(uint2*)&la[KTW * PA / 2 + o] = uint2{l0, l1};
    }
    if (BLDS)
#pragma unroll
      for (int i = 0; i < BCH; ++i) *(uint4*)&lb[bl[i]] = bst[i];
  }
};

// egm partial (T x Epad, one K split), samples (and lead field) staged through LDS.
// Workgroup = 4 TG waves: wave (tg, w) = (wave / 4, wave % 4) owns time rows
// [tb + 16 RW tg, tb + 16 RW (tg + 1)) and electrodes [e0 + 16 CE w, e0 + 16 CE (w + 1)) as
// RW x CE tiles of 16 x 16 with FP32 accumulators.  Per KS-deep K stage the workgroup
//   * reads its (16 RW TG) x KS FP32 sample tile from HBM once (4 KS B per time row, one wave
//     load covering whole row segments) and splits it into BF16 hi/lo planes in LDS;
//   * copies its (64 CE) x KS lead-field tile (BF16 hi/lo planes, L2/MALL-resident across
//     time tiles) into LDS, once for all time groups (BLDS), or each wave reads its CE
//     fragments per plane from that tile directly;
//   * runs every wave's MFMAs with the sample operand from LDS.
// The next stage's samples (and lead field) are loaded into registers while the current
// stage's MFMAs run.  LO: add the hi*lo and lo*hi products.
// Operand layout of v_mfma_f32_16x16x32_bf16: lane (r16 = lane % 16, q = lane / 16) holds
// A row r16 and B column r16, k = 8q .. 8q+7; the result lane holds column r16, rows 4q .. 4q+3.
//
// Tile order (1D grid, mt time tiles x ncols (electrode tile, K split) columns, the column count
// padded to a multiple of 8): workgroups are dispatched to the 8 XCDs round-robin, so workgroup
// id runs on XCD id % 8.  Slot id / 8 of XCD x walks the mt time tiles of column
// 8 (slot / mt) + x, i.e. all time tiles of a column run on one XCD at about the same time
// and read each lead-field stage once from HBM, then from that XCD's L2.
template <int CE, bool LO>
__global__ void __launch_bounds__(GemmCfg<CE>::THREADS, 2) egm_gemm_lds(
    const float* __restrict__ S, int64_t T, int64_t stride, int N, int Kpad, int k_chunk,
    const unsigned short* __restrict__ Wt, int Epad, float* __restrict__ ws, int mt, int nt,
    int ncols) {
  using Cfg = GemmCfg<CE>;
  constexpr int RW = Cfg::RW, KS = Cfg::KS;
  constexpr bool BLDS = Cfg::BLDS;
  constexpr int NP = LO ? 2 : 1;  // planes
  constexpr int kTW = Cfg::TW;    // time rows per workgroup
  constexpr int EW = 64 * CE;     // electrodes per workgroup
  using Stage = GemmStage<kTW, EW, NP, KS, Cfg::THREADS, BLDS>;
  constexpr int PA = Stage::PA;
  __shared__ unsigned lds_a[NP * kTW * PA / 2];
  __shared__ unsigned lds_b[BLDS ? NP * EW * PA / 2 : 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int tg = wave >> 2, wv = wave & 3;
  const int slot = (int)(blockIdx.x >> 3);
  const int col = (slot / mt) * 8 + (int)(blockIdx.x & 7);
  if (col >= ncols) return;  // padding column (whole workgroup)
  const int split = col / nt;
  const int64_t tb = (int64_t)(slot % mt) * kTW;
  const int eb = (col % nt) * EW;
  const int kb = split * k_chunk;
  const int ke = min(N, kb + k_chunk);
  const int r16 = lane & 15, q = lane >> 4;
  // direct lead-field fragments: row e = eb + 16 CE wv + 16 c + r16 of each plane's tile
  const int bdir = (eb + 16 * CE * wv + r16) * kK2 + 8 * q;
  const size_t bplane = (size_t)Epad * kK2;

  Stage stg;  // staging addresses
  stg.init(tid, S, T, stride, tb, Epad, eb);
  f32x4 acc[RW][CE];
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int c = 0; c < CE; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Direct lead-field fragments (one 32-deep MFMA step per stage): those of a stage are
  // loaded once the previous stage's MFMAs have consumed theirs, into the same registers, so
  // that they arrive during the stage barrier and the sample store.  (Two register sets, each
  // loaded a whole stage ahead, measured slower: the registers they take cost the samples
  // their second stage in flight.)
  static_assert(BLDS || KS == 32, "direct lead-field fragments: one MFMA step per stage");
  struct Frags {
    bf16x8 h[CE], l[CE];
  };
  auto load_b = [&](Frags& b, int k) {
    const unsigned short* f = Wt + (size_t)(k / kK2) * 2 * bplane + (k % kK2) + bdir;
#pragma unroll
    for (int c = 0; c < CE; ++c) {
      b.h[c] = __builtin_bit_cast(bf16x8, *(const uint4*)(f + 16 * c * kK2));
      if (LO) b.l[c] = __builtin_bit_cast(bf16x8, *(const uint4*)(f + 16 * c * kK2 + bplane));
    }
  };
  using Data = typename Stage::Data;
  constexpr int DEPTH = Cfg::DEPTH;
  // One stage at k0 from the staged values d: LDS store, the prefetch of stage k0 + DEPTH KS
  // into d, the MFMAs, then the next stage's direct fragments into b.
  // FULL: the prefetched stage and the next one are whole stages in range.
  auto run_stage = [&](Data& d, Frags& b, int k0, auto full) {
    constexpr bool FULL = decltype(full)::value;
    __syncthreads();  // the previous stage's LDS reads are done
    stg.store(d, lds_a, lds_b);
    __syncthreads();
    if (FULL)
      stg.template load<true>(d, k0 + DEPTH * KS, ke, Wt, Epad);
    else if (k0 + DEPTH * KS < ke)
      stg.load(d, k0 + DEPTH * KS, ke, Wt, Epad);
#pragma unroll
    for (int s2 = 0; s2 < KS / 32; ++s2) {
      if (BLDS) {
#pragma unroll
        for (int c = 0; c < CE; ++c) {
          const int o = ((16 * CE * wv + 16 * c + r16) * PA + 32 * s2 + 8 * q) / 2;
          b.h[c] = __builtin_bit_cast(bf16x8, *(const uint4*)&lds_b[o]);
          if (LO) b.l[c] = __builtin_bit_cast(bf16x8, *(const uint4*)&lds_b[EW * PA / 2 + o]);
        }
      }
      // per time block: the products of all CE electrode tiles, pass by pass, so that
      // consecutive MFMAs on one accumulator are CE apart; the next block's sample operands
      // are read from LDS before this block's MFMAs issue
      auto read_a = [&](int r, bf16x8& h, bf16x8& l) {
        const int o = ((16 * RW * tg + 16 * r + r16) * PA + 32 * s2 + 8 * q) / 2;
        h = __builtin_bit_cast(bf16x8, *(const uint4*)&lds_a[o]);
        if (LO) l = __builtin_bit_cast(bf16x8, *(const uint4*)&lds_a[kTW * PA / 2 + o]);
      };
      // keep the stage's sample-operand reads after the LDS store and the lead-field reads
      // (register pressure; 0.50 -> 0.45 ms at 128 electrodes)
      __builtin_amdgcn_sched_barrier(0);
      bf16x8 ah, al, ahn, aln;
      read_a(0, ah, al);
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        if (r + 1 < RW) read_a(r + 1, ahn, aln);
#pragma unroll
        for (int c = 0; c < CE; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.h[c], acc[r][c], 0, 0, 0);
        if (LO) {
#pragma unroll
          for (int c = 0; c < CE; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.l[c], acc[r][c], 0, 0, 0);
#pragma unroll
          for (int c = 0; c < CE; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b.h[c], acc[r][c], 0, 0, 0);
        }
        ah = ahn;
        al = aln;
      }
    }
    if (!BLDS && (FULL || k0 + KS < ke)) load_b(b, k0 + KS);
  };
  using Full = std::true_type;
  using Part = std::false_type;
  Data d0, d1;
  Frags b;
  Data& d_odd = DEPTH == 2 ? d1 : d0;
  if (kb < ke) {
    stg.load(d0, kb, ke, Wt, Epad);
    if (DEPTH == 2 && kb + KS < ke) stg.load(d1, kb + KS, ke, Wt, Epad);
    if (!BLDS) load_b(b, kb);
  }
  // stages in pairs: even ones from d0, odd ones from d_odd; pairs whose prefetches are all
  // whole stages first, then the last few stages
  int k0 = kb;
  for (; k0 + (DEPTH + 2) * KS <= ke; k0 += 2 * KS) {
    run_stage(d0, b, k0, Full());
    run_stage(d_odd, b, k0 + KS, Full());
  }
  for (; k0 < ke; k0 += 2 * KS) {
    run_stage(d0, b, k0, Part());
    if (k0 + KS < ke) run_stage(d_odd, b, k0 + KS, Part());
  }
  float* wsp = ws + (size_t)split * T * Epad;
  const int e0 = eb + 16 * CE * wv;
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int c = 0; c < CE; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t t = tb + 16 * RW * tg + 16 * r + 4 * q + j;
        if (t < T) wsp[t * Epad + e0 + 16 * c + r16] = acc[r][c][j];
      }
}

__global__ void egm_reduce(const float* __restrict__ ws, int splits, int64_t T, int Epad, int E,
                           float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T * E) return;
  const int64_t t = i / E;
  const int e = (int)(i % E);
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += ws[((size_t)k * T + t) * Epad + e];
  out[i] = s;
}

// ---------------------------------------------------------------- streamed contraction
// A few new sample rows (T <= 16) of many simulations at once, e.g. the rows a batch of
// simulations produced during the last few hundred steps.  One launch covers up to
// kStreamJobs jobs (a job: one lead field, its T new rows and their output rows); work items are
// (job, 1024-voxel chunk, 64-electrode tile), one wavefront each.  Per 64-deep K stage the
// wavefront runs the BF16 hi/lo split products of one 16-row time block against four
// 16-electrode blocks on v_mfma_f32_16x16x32_bf16, with the lead-field operands loaded
// straight from the K-stage-major tiles (each lane's 8 k-values of one electrode are 16
// contiguous bytes) and the samples split into hi/lo in registers: no LDS and no barriers.
// At T = 1 the work is bound by the lead-field stream (4 bytes per electrode and voxel), not
// by the MFMAs.  Each item writes its partial (T x 64) block to the workspace; a second kernel
// sums the chunks of each output element in a fixed order (egm_stream_reduce_k), so the result
// depends only on the job itself, not on the other jobs of the launch.
constexpr int kStreamJobs = 64;
constexpr int kStreamRows = 16;    // time rows of one job (one MFMA row block)
constexpr int kStreamStages = 16;  // 64-deep K stages per work item (1024 voxels)
constexpr int kStreamWaves = 4;    // wavefronts (independent items) per workgroup

struct StreamJob {
  const unsigned short* Wt;
  const float* S;
  float* out;
  int64_t stride;
  int32_t T, N, Epad, E;
  int32_t item0;  // first work item of the job
  int32_t tile0;  // first (time row, 64-electrode tile) of the job (reduction)
};
struct StreamTable {
  int32_t n;
  StreamJob j[kStreamJobs];
};

__device__ __forceinline__ int stream_chunks(int N) {
  return ((N + kK2 - 1) / kK2 + kStreamStages - 1) / kStreamStages;
}

__global__ void __launch_bounds__(64 * kStreamWaves) egm_stream_k(const StreamTable tab,
                                                                   float* __restrict__ ws,
                                                                   int n_items) {
  const int item = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kStreamWaves +
                                                        (threadIdx.x >> 6)));
  if (item >= n_items) return;
  int j = 0;
  for (int q = 1; q < tab.n; ++q)
    if (tab.j[q].item0 <= item) j = q;
  const StreamJob& J = tab.j[j];
  const int lane = threadIdx.x & 63, r16 = lane & 15, q = lane >> 4;
  const int etiles = J.Epad >> 6;
  const int li = item - J.item0;
  const int et = li % etiles, ch = li / etiles;
  const int nst = (J.N + kK2 - 1) / kK2;
  const int s0 = ch * kStreamStages, s1 = min(nst, s0 + kStreamStages);
  const int t = r16 < J.T ? r16 : J.T - 1;  // rows >= T repeat row T - 1 (never written)
  const float* arow = J.S + (int64_t)t * J.stride;
  const size_t plane = (size_t)J.Epad * kK2;
  f32x4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int s = s0; s < s1; ++s) {
    const unsigned short* tile = J.Wt + (size_t)s * 2 * plane + (size_t)(et * 64 + r16) * kK2;
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      const int k = s * kK2 + 32 * s2 + 8 * q;
      float a[8];
      if (k + 8 <= J.N) {
        float4 x0, x1;
        __builtin_memcpy(&x0, arow + k, 16);
        __builtin_memcpy(&x1, arow + k + 4, 16);
        a[0] = x0.x; a[1] = x0.y; a[2] = x0.z; a[3] = x0.w;
        a[4] = x1.x; a[5] = x1.y; a[6] = x1.z; a[7] = x1.w;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = k + i < J.N ? arow[k + i] : 0.0f;
      }
      unsigned h[4], l[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) split_pair(a[2 * i], a[2 * i + 1], h[i], l[i]);
      const bf16x8 ah = __builtin_bit_cast(bf16x8, uint4{h[0], h[1], h[2], h[3]});
      const bf16x8 al = __builtin_bit_cast(bf16x8, uint4{l[0], l[1], l[2], l[3]});
      uint4 bh[4], bl[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const unsigned short* p = tile + (size_t)(16 * c) * kK2 + 32 * s2 + 8 * q;
        bh[c] = *(const uint4*)p;
        bl[c] = *(const uint4*)(p + plane);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bf16x8 h8 = __builtin_bit_cast(bf16x8, bh[c]);
        const bf16x8 l8 = __builtin_bit_cast(bf16x8, bl[c]);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, h8, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, l8, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, h8, acc[c], 0, 0, 0);
      }
    }
  }
  // result lane: electrode 16 c + r16 of the tile, time rows 4 q .. 4 q + 3
  float* w = ws + (size_t)item * kStreamRows * 64;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int row = 4 * q + jj;
      if (row < J.T) w[row * 64 + 16 * c + r16] = acc[c][jj];
    }
}

// Sum of the chunk partials of one (job, time row, 64-electrode tile) per workgroup: wave w
// adds chunks w, w + 16, ... in order, then the 16 partial sums are added in wave order, so the
// result depends only on the job itself.
constexpr int kReduceWaves = 16;
__global__ void __launch_bounds__(64 * kReduceWaves) egm_stream_reduce_k(const StreamTable tab,
                                                                         const float* __restrict__ ws) {
  __shared__ float part[kReduceWaves][64];
  const int tile = blockIdx.x;
  int j = 0;
  for (int q = 1; q < tab.n; ++q)
    if (tab.j[q].tile0 <= tile) j = q;
  const StreamJob& J = tab.j[j];
  const int etiles = J.Epad >> 6;
  const int li = tile - J.tile0;
  const int t = li / etiles, et = li % etiles;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int chunks = stream_chunks(J.N);
  const float* p = ws + ((size_t)(J.item0 + et) * kStreamRows + t) * 64 + lane;
  const size_t cstride = (size_t)etiles * kStreamRows * 64;
  float sum = 0.0f;
  for (int c = w; c < chunks; c += kReduceWaves) sum += p[c * cstride];
  part[w][lane] = sum;
  __syncthreads();
  if (w == 0) {
    float tot = part[0][lane];
#pragma unroll
    for (int k = 1; k < kReduceWaves; ++k) tot += part[k][lane];
    const int e = et * 64 + lane;
    if (e < J.E) J.out[(size_t)t * J.E + e] = tot;
  }
}

constexpr int kNormThreads = 1024;
constexpr int kNormMaxE = 4096;

// One workgroup: peak-to-peak of every electrode over time, their mean, and the scaling of
// the whole (T x E) electrogram (simulation/unipolar_electrogram_equation_parts.py:118-124).
// No allocation and no host synchronisation, so it can follow the contraction on a stream.
__global__ void __launch_bounds__(kNormThreads) egm_normalize_k(float* __restrict__ egm, int64_t T,
                                                                int E, double typical) {
  __shared__ float slo[kNormThreads], shi[kNormThreads];
  __shared__ float ptp[kNormMaxE];
  __shared__ double scale;
  const int tid = threadIdx.x;
  const int ec = E < kNormThreads ? E : kNormThreads;
  const int G = kNormThreads / ec;
  for (int e0 = 0; e0 < E; e0 += ec) {
    const int el = tid % ec, g = tid / ec;
    const int e = e0 + el;
    float lo = INFINITY, hi = -INFINITY;
    if (g < G && e < E)
      for (int64_t t = g; t < T; t += G) {
        const float v = egm[t * E + e];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
      }
    slo[tid] = lo;
    shi[tid] = hi;
    __syncthreads();
    if (tid < ec && e0 + tid < E) {
      float l = slo[tid], h = shi[tid];
      for (int q = 1; q < G; ++q) {
        l = fminf(l, slo[tid + q * ec]);
        h = fmaxf(h, shi[tid + q * ec]);
      }
      ptp[e0 + tid] = h - l;
    }
    __syncthreads();
  }
  if (tid == 0) {
    double m = 0.0;
    for (int e = 0; e < E; ++e) m += ptp[e];
    scale = typical / (m / E);
  }
  __syncthreads();
  const int64_t n = T * E;
  for (int64_t i = tid; i < n; i += kNormThreads) egm[i] = (float)(egm[i] * scale);
}

// Electrode activation time from the electrogram (utility/lat_map.py:20-27): per electrode
// the first time index of the steepest downstroke, argmax_t -(egm[t+1] - egm[t]), then
// shifted so that the earliest electrode is 0.  One workgroup per electrode; the order is
// NumPy argmax's: a NaN is above every number (the first NaN wins), and ties go to the lower
// index.
constexpr int kLatThreads = 256;

// (v, i) comes before (w, j) in argmax order.
__device__ __forceinline__ bool argmax_before(float v, int64_t i, float w, int64_t j) {
  const bool nv = __builtin_isnan(v), nw = __builtin_isnan(w);
  if (nv || nw) return nv && (!nw || i < j);
  return v > w || (v == w && i < j);
}

__global__ void __launch_bounds__(kLatThreads) egm_downstroke_k(const float* __restrict__ egm,
                                                                int64_t T, int E,
                                                                int32_t* __restrict__ idx) {
  __shared__ float sv[kLatThreads];
  __shared__ int64_t si[kLatThreads];
  const int e = blockIdx.x, tid = threadIdx.x;
  float best = -INFINITY;
  int64_t bi = T;  // sentinel above every index
  for (int64_t t = tid; t + 1 < T; t += kLatThreads) {
    const float v = -(egm[(t + 1) * E + e] - egm[t * E + e]);
    if (argmax_before(v, t, best, bi)) {
      best = v;
      bi = t;
    }
  }
  sv[tid] = best;
  si[tid] = bi;
  __syncthreads();
  for (int o = kLatThreads / 2; o > 0; o >>= 1) {
    if (tid < o) {
      const float v = sv[tid + o];
      const int64_t i = si[tid + o];
      if (argmax_before(v, i, sv[tid], si[tid])) {
        sv[tid] = v;
        si[tid] = i;
      }
    }
    __syncthreads();
  }
  if (tid == 0) idx[e] = si[0] < T ? (int32_t)si[0] : 0;
}

__global__ void __launch_bounds__(kLatThreads) lat_shift_k(int32_t* __restrict__ lat, int E) {
  __shared__ int32_t sm[kLatThreads];
  const int tid = threadIdx.x;
  int32_t m = INT32_MAX;
  for (int e = tid; e < E; e += kLatThreads) m = min(m, lat[e]);
  sm[tid] = m;
  __syncthreads();
  for (int o = kLatThreads / 2; o > 0; o >>= 1) {
    if (tid < o) sm[tid] = min(sm[tid], sm[tid + o]);
    __syncthreads();
  }
  const int32_t lo = sm[0];
  for (int e = tid; e < E; e += kLatThreads) lat[e] -= lo;
}

}  // namespace

struct phs_egm {
  int N = 0, E = 0, Kpad = 0, Epad = 0;
  int ce = 4;  // 16-electrode tiles per wave in egm_gemm_lds (workgroup: 64 CE electrodes)
  float* Wf = nullptr;
  unsigned short* Wt = nullptr;  // BF16 hi/lo planes, tiles [Kpad / 64][2][Epad][64]
  float* ws = nullptr;
  size_t ws_bytes = 0;
  int resident[2] = {0, 0};  // workgroups of egm_gemm_lds (without / with lo) the device
                             // holds at once, 0 = not yet queried
};

extern "C" {

int phs_egm_create(int32_t N, int32_t E, const double* voxel, const double* electrodes,
                   const double* c_voxel, const double* D, const int32_t* nb, int32_t nb_stride,
                   double delta, const int32_t* perm, phs_egm** handle, void* stream) {
  if (N <= 0 || E <= 0 || !voxel || !electrodes || !c_voxel || !nb || nb_stride < 6 || !handle ||
      !(delta > 0))
    return phs::fail(PHS_ERR_ARG, "invalid electrogram arguments");
  // Transpose of the central-difference stencil (simulation/unipolar_electrogram_equation_parts
  // .py:60-103: the axis-k gradient at v is (u(p_k(v)) - u(m_k(v))) / (2 Delta) when both
  // exist, else 0), for any face table: the terms reading u(w), grouped by w.
  for (int32_t w = 0; w < N; ++w)
    for (int k = 0; k < 6; ++k) {
      const int32_t m = nb[(int64_t)w * nb_stride + k];
      if (m < -1 || m >= N) return phs::fail(PHS_ERR_ARG, "neighbour id out of range");
    }
  std::vector<int32_t> inv_ptr((size_t)N + 1, 0);
  for (int32_t v = 0; v < N; ++v)
    for (int k = 0; k < 3; ++k) {
      const int32_t p = nb[(int64_t)v * nb_stride + 2 * k], m = nb[(int64_t)v * nb_stride + 2 * k + 1];
      if (p >= 0 && m >= 0) {
        ++inv_ptr[(size_t)p + 1];
        ++inv_ptr[(size_t)m + 1];
      }
    }
  for (int32_t w = 0; w < N; ++w) inv_ptr[(size_t)w + 1] += inv_ptr[w];
  std::vector<int32_t> inv_v(inv_ptr[N]);
  std::vector<unsigned char> inv_code(inv_ptr[N]);
  {
    std::vector<int32_t> fill(inv_ptr.begin(), inv_ptr.end() - 1);
    for (int32_t v = 0; v < N; ++v)
      for (int k = 0; k < 3; ++k) {
        const int32_t p = nb[(int64_t)v * nb_stride + 2 * k], m = nb[(int64_t)v * nb_stride + 2 * k + 1];
        if (p >= 0 && m >= 0) {
          inv_v[fill[p]] = v;
          inv_code[fill[p]++] = (unsigned char)(2 * k);
          inv_v[fill[m]] = v;
          inv_code[fill[m]++] = (unsigned char)(2 * k + 1);
        }
      }
    // order within each voxel: axis, then + before -, then v (v ascending already)
    std::vector<std::pair<unsigned char, int32_t>> tmp;
    for (int32_t w = 0; w < N; ++w) {
      const int32_t a = inv_ptr[w], b = inv_ptr[(size_t)w + 1];
      if (b - a <= 1) continue;
      tmp.clear();
      for (int32_t q = a; q < b; ++q) tmp.emplace_back(inv_code[q], inv_v[q]);
      std::sort(tmp.begin(), tmp.end());
      for (int32_t q = a; q < b; ++q) {
        inv_code[q] = tmp[q - a].first;
        inv_v[q] = tmp[q - a].second;
      }
    }
  }
  if (perm) {
    std::vector<char> seen(N, 0);
    for (int32_t i = 0; i < N; ++i) {
      if (perm[i] < 0 || perm[i] >= N || seen[perm[i]])
        return phs::fail(PHS_ERR_ARG, "perm is not a permutation");
      seen[perm[i]] = 1;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  phs_egm* g = new phs_egm();
  g->N = N;
  g->E = E;
  g->Kpad = (N + 63) / 64 * 64;
  // workgroup width of the LDS-staged contraction: 4 waves x CE tiles of 16 electrodes;
  // Epad is a whole number of workgroup columns
  g->ce = E <= 64 ? 1 : (E <= 128 ? 2 : 4);
  g->Epad = (E + 64 * g->ce - 1) / (64 * g->ce) * (64 * g->ce);
  const size_t nw = (size_t)g->Epad * g->Kpad;
  double *dv = nullptr, *de = nullptr, *dc = nullptr, *dD = nullptr;
  int *dptr = nullptr, *dinv = nullptr, *dperm = nullptr;
  unsigned char* dcode = nullptr;
  const size_t n_inv = std::max<size_t>(inv_v.size(), 1);
  hipError_t e = hipSuccess;
  auto cleanup = [&](int code, const std::string& msg) {
    hipFree(dv);
    hipFree(de);
    hipFree(dc);
    hipFree(dD);
    hipFree(dptr);
    hipFree(dinv);
    hipFree(dcode);
    hipFree(dperm);
    if (code) phs_egm_destroy(g);
    return code ? phs::fail(code, msg) : PHS_OK;
  };
  if ((e = hipMalloc(&g->Wf, 4 * nw)) || (e = hipMalloc(&g->Wt, 4 * nw)) ||
      (e = hipMalloc(&dv, 24 * (size_t)N)) ||
      (e = hipMalloc(&de, 24 * (size_t)E)) || (e = hipMalloc(&dc, 8 * (size_t)N)) ||
      (D && (e = hipMalloc(&dD, 72 * (size_t)N))) ||
      (e = hipMalloc(&dptr, 4 * ((size_t)N + 1))) || (e = hipMalloc(&dinv, 4 * n_inv)) ||
      (e = hipMalloc(&dcode, n_inv)) ||
      (perm && (e = hipMalloc(&dperm, 4 * (size_t)N))))
    return cleanup(PHS_ERR_OOM, std::string("egm hipMalloc: ") + hipGetErrorString(e));
  if ((e = hipMemcpyAsync(dv, voxel, 24 * (size_t)N, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(de, electrodes, 24 * (size_t)E, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(dc, c_voxel, 8 * (size_t)N, hipMemcpyHostToDevice, s)) ||
      (D && (e = hipMemcpyAsync(dD, D, 72 * (size_t)N, hipMemcpyHostToDevice, s))) ||
      (e = hipMemcpyAsync(dptr, inv_ptr.data(), 4 * ((size_t)N + 1), hipMemcpyHostToDevice, s)) ||
      (!inv_v.empty() &&
       ((e = hipMemcpyAsync(dinv, inv_v.data(), 4 * inv_v.size(), hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(dcode, inv_code.data(), inv_code.size(), hipMemcpyHostToDevice, s)))) ||
      (perm && (e = hipMemcpyAsync(dperm, perm, 4 * (size_t)N, hipMemcpyHostToDevice, s))))
    return cleanup(PHS_ERR_HIP, std::string("egm upload: ") + hipGetErrorString(e));
  const dim3 grid((g->Kpad + 255) / 256, g->Epad);
  build_leadfield<<<grid, 256, 0, s>>>(N, E, g->Kpad, dv, de, dc, dD, dptr, dinv, dcode, dperm,
                                       1.0 / (2.0 * delta), g->Epad, g->Wf, g->Wt);
  if ((e = hipGetLastError()) || (e = hipStreamSynchronize(s)))
    return cleanup(PHS_ERR_HIP, std::string("build_leadfield: ") + hipGetErrorString(e));
  *handle = g;
  return cleanup(PHS_OK, "");
}

}  // extern "C"

namespace {

// The split-K contraction and its fixed-order reduction.  Workspace: with `fixed` (caller-
// provided, `fixed_bytes` long) nothing is allocated and a request that needs more fails;
// otherwise the handle's own workspace grows on demand.
int egm_run(phs_egm* g, const float* samples, int64_t T, int64_t stride, float* egm,
            int32_t precision, hipStream_t s, float* fixed, size_t fixed_bytes) {
  if (!g || !samples || !egm || T <= 0 || stride < g->N)
    return phs::fail(PHS_ERR_ARG, "invalid electrogram run arguments");
  if (precision < 0 || precision > 2) return phs::fail(PHS_ERR_ARG, "precision must be 0, 1 or 2");
  const bool lds = precision != 1;
  const int tm = !lds ? kTM : g->ce == 4 ? GemmCfg<4>::TW : g->ce == 2 ? GemmCfg<2>::TW
                                                                       : GemmCfg<1>::TW;
  const int threads = g->ce == 4 ? GemmCfg<4>::THREADS : g->ce == 2 ? GemmCfg<2>::THREADS
                                                                    : GemmCfg<1>::THREADS;
  const int te = lds ? 64 * g->ce : kTE;
  const int kstep = lds ? kK2 : kKS;
  const int mt = (int)((T + tm - 1) / tm);
  const int nt = g->Epad / te;
  const int ksteps = (g->N + kstep - 1) / kstep;
  // Split K so that the workgroups fill the device once: as many as it holds at once (the
  // occupancy of the kernel times the CU count).  At 64 electrodes (2 per CU) this is 5 %
  // faster than 4 per CU, whose second round is partly a tail; at 256 electrodes (1 per CU)
  // the count makes no measurable difference.  Chunks are whole K stages.
  int target = 1024;
  if (lds) {
    const bool lo = precision == 0;
    int& res = g->resident[lo ? 1 : 0];
    if (res == 0) {
      const void* kern = g->ce == 4 ? (lo ? (const void*)egm_gemm_lds<4, true>
                                          : (const void*)egm_gemm_lds<4, false>)
                         : g->ce == 2 ? (lo ? (const void*)egm_gemm_lds<2, true>
                                            : (const void*)egm_gemm_lds<2, false>)
                                      : (lo ? (const void*)egm_gemm_lds<1, true>
                                            : (const void*)egm_gemm_lds<1, false>);
      int per_cu = 0, dev = 0, cus = 0;
      PHS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0));
      PHS_HIP_CHECK(hipGetDevice(&dev));
      PHS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      res = std::max(1, per_cu) * std::max(1, cus);
    }
    target = res;
  }
  // (at most the resident count: a partial second round of workgroups would double the time)
  int splits = std::max(1, std::min(ksteps, target / (mt * nt)));
  const int k_chunk = (ksteps + splits - 1) / splits * kstep;
  splits = (g->N + k_chunk - 1) / k_chunk;
  const size_t need = sizeof(float) * (size_t)splits * T * g->Epad;
  float* ws = fixed;
  if (fixed) {
    if (need > fixed_bytes)
      return phs::fail(PHS_ERR_ARG, "electrogram needs " + std::to_string(need) +
                                        " bytes of split-K workspace, more than the " +
                                        std::to_string(fixed_bytes) +
                                        " of the static scratch: contract fewer samples per call");
  } else {
    if (need > g->ws_bytes) {
      hipFree(g->ws);
      g->ws = nullptr;
      g->ws_bytes = 0;
      PHS_TRY_ALLOC(g->ws, need);
      g->ws_bytes = need;
    }
    ws = g->ws;
  }
  const int ncols = nt * splits;
  const int ncols_pad = (ncols + 7) / 8 * 8;
  const dim3 grid(mt, nt, splits);
  if (precision == 1) {
    egm_gemm_f32<<<grid, 256, 0, s>>>(samples, T, stride, g->N, g->Kpad, k_chunk, g->Wf,
                                      g->Epad, ws);
  } else {
    auto launch = [&](auto kern) {
      kern<<<dim3(mt * ncols_pad), threads, 0, s>>>(samples, T, stride, g->N, g->Kpad,
                                                         k_chunk, g->Wt, g->Epad, ws, mt, nt,
                                                         ncols);
    };
    const bool lo = precision == 0;
    if (g->ce == 4)
      lo ? launch(egm_gemm_lds<4, true>) : launch(egm_gemm_lds<4, false>);
    else if (g->ce == 2)
      lo ? launch(egm_gemm_lds<2, true>) : launch(egm_gemm_lds<2, false>);
    else
      lo ? launch(egm_gemm_lds<1, true>) : launch(egm_gemm_lds<1, false>);
  }
  const int64_t n = T * g->E;
  egm_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ws, splits, T, g->Epad, g->E, egm);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

// Split-K workspace of phs_egm_run_noalloc: 64 MiB per device, for callers that cannot
// allocate.  In host code `scratch` names a host-side stand-in, never the device array: its
// device address comes from hipGetSymbolAddress, once per device (scratch_on_device).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first call there.
int scratch_on_device(float** ptr) {
  int dev = 0;
  PHS_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return phs::fail(PHS_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return phs::fail(PHS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
    g_scratch_dev[dev] = (float*)p;
  }
  *ptr = g_scratch_dev[dev];
  return PHS_OK;
}

}  // namespace

extern "C" {

int phs_egm_run(phs_egm* g, const float* samples, int64_t T, int64_t stride, float* egm,
                int32_t precision, void* stream) {
  return egm_run(g, samples, T, stride, egm, precision, (hipStream_t)stream, nullptr, 0);
}

int phs_egm_run_noalloc(phs_egm* g, const float* samples, int64_t T, int64_t stride, float* egm,
                        int32_t precision, void* stream) {
  float* ws = nullptr;
  if (int rc = scratch_on_device(&ws)) return rc;
  return egm_run(g, samples, T, stride, egm, precision, (hipStream_t)stream, ws, kScratchBytes);
}

int phs_egm_scratch_bytes(int64_t* bytes) {
  if (!bytes) return phs::fail(PHS_ERR_ARG, "NULL argument");
  *bytes = (int64_t)kScratchBytes;
  return PHS_OK;
}

}  // extern "C"

namespace {

int host_stream_chunks(int N) { return ((N + kK2 - 1) / kK2 + kStreamStages - 1) / kStreamStages; }

// Items of the launch groups (kStreamJobs consecutive jobs each); -1 with the error set.
int64_t stream_check(const phs_egm_job* jobs, int32_t n_jobs, int64_t* max_items) {
  *max_items = 0;
  for (int32_t g0 = 0; g0 < n_jobs; g0 += kStreamJobs) {
    int64_t items = 0;
    for (int32_t i = g0; i < std::min(n_jobs, g0 + kStreamJobs); ++i) {
      const phs_egm_job& J = jobs[i];
      if (!J.lead_field || !J.samples || !J.egm || J.n_time < 1 || J.n_time > kStreamRows ||
          J.sample_stride < J.lead_field->N)
        return phs::fail(PHS_ERR_ARG, "invalid streamed electrogram job (n_time must be 1..16, "
                                      "sample_stride >= n_voxel)");
      items += (int64_t)host_stream_chunks(J.lead_field->N) * (J.lead_field->Epad / 64);
    }
    *max_items = std::max(*max_items, items);
  }
  if (*max_items >= (1LL << 30)) return phs::fail(PHS_ERR_ARG, "too many streamed work items");
  return PHS_OK;
}

}  // namespace

extern "C" {

int phs_egm_stream_workspace(const phs_egm_job* jobs, int32_t n_jobs, int64_t* bytes) {
  if (!jobs || n_jobs <= 0 || !bytes) return phs::fail(PHS_ERR_ARG, "invalid arguments");
  int64_t items = 0;
  if (int rc = stream_check(jobs, n_jobs, &items)) return rc;
  *bytes = items * kStreamRows * 64 * (int64_t)sizeof(float);
  return PHS_OK;
}

int phs_egm_stream(const phs_egm_job* jobs, int32_t n_jobs, void* workspace,
                   int64_t workspace_bytes, void* stream) {
  if (!jobs || n_jobs <= 0 || !workspace) return phs::fail(PHS_ERR_ARG, "invalid arguments");
  int64_t items = 0;
  if (int rc = stream_check(jobs, n_jobs, &items)) return rc;
  if (workspace_bytes < items * kStreamRows * 64 * (int64_t)sizeof(float))
    return phs::fail(PHS_ERR_ARG, "workspace smaller than phs_egm_stream_workspace");
  hipStream_t s = (hipStream_t)stream;
  float* ws = (float*)workspace;
  for (int32_t g0 = 0; g0 < n_jobs; g0 += kStreamJobs) {
    StreamTable tab;
    std::memset(&tab, 0, sizeof tab);
    int32_t n_items = 0, n_tiles = 0;
    for (int32_t i = g0; i < std::min(n_jobs, g0 + kStreamJobs); ++i) {
      const phs_egm_job& J = jobs[i];
      const phs_egm* g = J.lead_field;
      StreamJob& d = tab.j[tab.n++];
      d.Wt = g->Wt;
      d.S = J.samples;
      d.out = J.egm;
      d.stride = J.sample_stride;
      d.T = (int32_t)J.n_time;
      d.N = g->N;
      d.Epad = g->Epad;
      d.E = g->E;
      d.item0 = n_items;
      d.tile0 = n_tiles;
      n_items += host_stream_chunks(g->N) * (g->Epad / 64);
      n_tiles += d.T * (g->Epad / 64);
    }
    egm_stream_k<<<(n_items + kStreamWaves - 1) / kStreamWaves, 64 * kStreamWaves, 0, s>>>(
        tab, ws, n_items);
    egm_stream_reduce_k<<<n_tiles, 64 * kReduceWaves, 0, s>>>(tab, ws);
    PHS_HIP_CHECK(hipGetLastError());
  }
  return PHS_OK;
}

int phs_egm_destroy(phs_egm* g) {
  if (!g) return PHS_OK;
  hipFree(g->Wf);
  hipFree(g->Wt);
  hipFree(g->ws);
  delete g;
  return PHS_OK;
}

int phs_egm_normalize(float* egm, int64_t T, int32_t E, double typical, void* stream) {
  if (!egm || T <= 0 || E <= 0) return phs::fail(PHS_ERR_ARG, "invalid normalize arguments");
  if (E > kNormMaxE) return phs::fail(PHS_ERR_ARG, "normalize supports at most 4096 electrodes");
  egm_normalize_k<<<1, kNormThreads, 0, (hipStream_t)stream>>>(egm, T, E, typical);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

int phs_egm_electrode_lat(const float* egm, int64_t T, int32_t E, int32_t* lat, void* stream) {
  if (!egm || !lat || T < 2 || E <= 0) return phs::fail(PHS_ERR_ARG, "invalid electrode LAT arguments");
  hipStream_t s = (hipStream_t)stream;
  egm_downstroke_k<<<E, kLatThreads, 0, s>>>(egm, T, E, lat);
  lat_shift_k<<<1, kLatThreads, 0, s>>>(lat, E);
  PHS_HIP_CHECK(hipGetLastError());
  return PHS_OK;
}

}  // extern "C"
