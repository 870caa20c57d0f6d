// Host-side brick planner: Morton-ordered bricks of 64 lattice cells.
//
// Voxels (integer lattice coordinates) are grouped into bricks of bx*by*bz = 64 cells; a
// brick is processed by one 64-lane wavefront, lane l <-> cell
//   (l % bx, (l / bx) % by, l / (bx*by)).
// Bricks are sorted by the Morton code of their brick coordinates so that bricks close in
// space are close in memory; voxels are stored compactly brick by brick (occupied cells
// only, in lane order).  A voxel's compact index is
//   start[b] + popcount(mask[b] & ((1 << lane) - 1)).
// Each brick records its 6 face-neighbour bricks (-1 = none).
//
// The plan is checked against the caller's neighbour table: face neighbour k of every
// voxel must be the voxel at the adjacent lattice point, or -1.  A -1 where the lattice point
// is occupied is a cut link, recorded per brick and lane so that that row of the operator
// leaves the neighbour out, as the reference's operator built from the table does (its rows
// may then differ from their transposes).  The reference builds the table from coordinates
// by KD-tree lookup at unit offsets (utility/voxelization.py:102-220).
#include <algorithm>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "plan.h"

namespace phs {

static inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

static inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

static const int kDir[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

int build_plan(int32_t n, const int32_t* ijk, const int32_t* nb, int32_t nb_stride, int bx, int by,
               int bz, BrickPlan& P, std::string& err) {
  if (bx <= 0 || by <= 0 || bz <= 0 || bx * by * bz != 64) {
    err = "brick shape must have bx*by*bz == 64";
    return -1;
  }
  if (n <= 0 || !ijk) {
    err = "lattice coordinates required";
    return -1;
  }
  P = BrickPlan();
  P.bx = bx;
  P.by = by;
  P.bz = bz;
  int32_t lo[3] = {ijk[0], ijk[1], ijk[2]};
  int32_t hi[3] = {ijk[0], ijk[1], ijk[2]};
  for (int32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], ijk[3 * (int64_t)i + a]);
      hi[a] = std::max(hi[a], ijk[3 * (int64_t)i + a]);
    }
  const int bd[3] = {bx, by, bz};
  for (int a = 0; a < 3; ++a) {
    if ((int64_t)hi[a] - lo[a] + 1 > (int64_t)bd[a] * (1 << 20)) {
      err = "lattice extent too large";
      return -1;
    }
  }
  // key = morton(brick) << 6 | cell
  std::vector<std::pair<uint64_t, int32_t>> keys(n);
  for (int32_t i = 0; i < n; ++i) {
    const uint32_t x = ijk[3 * (int64_t)i] - lo[0], y = ijk[3 * (int64_t)i + 1] - lo[1],
                   z = ijk[3 * (int64_t)i + 2] - lo[2];
    const uint32_t cell = (x % bx) + bx * ((y % by) + by * (z % bz));
    keys[i] = {(morton3(x / bx, y / by, z / bz) << 6) | cell, i};
  }
  std::sort(keys.begin(), keys.end());
  P.perm.resize(n);
  P.inv.resize(n);
  std::unordered_map<uint64_t, int32_t> bmap;
  bmap.reserve((size_t)n / 16 + 16);
  for (int32_t c = 0; c < n; ++c) {
    const uint64_t bkey = keys[c].first >> 6;
    const uint32_t cell = keys[c].first & 63;
    if (c > 0 && keys[c].first == keys[c - 1].first) {
      err = "duplicate lattice coordinates";
      return -1;
    }
    if (c == 0 || bkey != (keys[c - 1].first >> 6)) {
      bmap[bkey] = (int32_t)P.mask.size();
      P.mask.push_back(0);
      P.start.push_back(c);
      const int32_t o = keys[c].second;
      P.bcoord.push_back((ijk[3 * (int64_t)o] - lo[0]) / bx);
      P.bcoord.push_back((ijk[3 * (int64_t)o + 1] - lo[1]) / by);
      P.bcoord.push_back((ijk[3 * (int64_t)o + 2] - lo[2]) / bz);
    }
    P.mask.back() |= 1ULL << cell;
    P.perm[c] = keys[c].second;
    P.inv[keys[c].second] = c;
  }
  const int32_t nbk = (int32_t)P.mask.size();
  P.n_bricks = nbk;
  P.nbb.assign((size_t)nbk * 6, -1);
  for (int32_t b = 0; b < nbk; ++b)
    for (int d = 0; d < 6; ++d) {
      const int64_t x = (int64_t)P.bcoord[3 * b] + kDir[d][0];
      const int64_t y = (int64_t)P.bcoord[3 * b + 1] + kDir[d][1];
      const int64_t z = (int64_t)P.bcoord[3 * b + 2] + kDir[d][2];
      if (x < 0 || y < 0 || z < 0) continue;
      auto it = bmap.find(morton3((uint32_t)x, (uint32_t)y, (uint32_t)z));
      if (it != bmap.end()) P.nbb[(size_t)b * 6 + d] = it->second;
    }
  P.cut.assign((size_t)nbk * 6, 0);
  if (!nb) return 0;
  // Verify the caller's face-neighbour table against the lattice; record its cut links.
  for (int32_t b = 0; b < nbk; ++b) {
    const uint64_t m = P.mask[b];
    for (int lane = 0; lane < 64; ++lane) {
      if (!((m >> lane) & 1)) continue;
      const int32_t c = P.start[b] + __builtin_popcountll(m & ((1ULL << lane) - 1));
      const int32_t o = P.perm[c];
      const int l[3] = {lane % bx, (lane / bx) % by, lane / (bx * by)};
      for (int d = 0; d < 6; ++d) {
        int q[3] = {l[0] + kDir[d][0], l[1] + kDir[d][1], l[2] + kDir[d][2]};
        int32_t tb = b;
        for (int a = 0; a < 3; ++a) {
          if (q[a] < 0 || q[a] >= bd[a]) {
            tb = P.nbb[(size_t)b * 6 + d];
            q[a] = (q[a] + bd[a]) % bd[a];
          }
        }
        int32_t expect = -1;
        if (tb >= 0) {
          const int cell = q[0] + bx * (q[1] + by * q[2]);
          const uint64_t tm = P.mask[tb];
          if ((tm >> cell) & 1)
            expect = P.perm[P.start[tb] + __builtin_popcountll(tm & ((1ULL << cell) - 1))];
        }
        const int32_t given = nb[(int64_t)o * nb_stride + d];
        if (given == -1 && expect >= 0) {
          P.cut[(size_t)b * 6 + d] |= 1ULL << lane;
        } else if (given != expect) {
          err = "neighbour table entry is not the adjacent lattice voxel at voxel " +
                std::to_string(o) + " direction " + std::to_string(d);
          return -2;
        }
      }
    }
  }
  return 0;
}

}  // namespace phs

#include "phs_internal.h"

extern "C" int phs_plan_build(int32_t n, const int32_t* ijk, const int32_t* nb, int32_t nb_stride,
                              const int32_t* brick, int32_t* perm, int32_t* n_bricks,
                              uint64_t* masks) {
  if (!brick || !perm || !n_bricks) return phs::fail(PHS_ERR_ARG, "NULL argument");
  if (nb && nb_stride < 6) return phs::fail(PHS_ERR_ARG, "nb_stride must be >= 6");
  phs::BrickPlan P;
  std::string err;
  const int rc = phs::build_plan(n, ijk, nb, nb_stride, brick[0], brick[1], brick[2], P, err);
  if (rc == -2) return phs::fail(PHS_ERR_UNSUPPORTED, err);
  if (rc) return phs::fail(PHS_ERR_ARG, err);
  std::copy(P.perm.begin(), P.perm.end(), perm);
  *n_bricks = P.n_bricks;
  if (masks) std::copy(P.mask.begin(), P.mask.end(), masks);
  return PHS_OK;
}
