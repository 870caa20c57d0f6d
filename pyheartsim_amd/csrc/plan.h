// Brick plan built on the host (see plan.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace phs {

struct BrickPlan {
  int bx = 4, by = 4, bz = 4;
  int32_t n_bricks = 0;
  std::vector<int32_t> perm;    // compact index -> caller voxel id
  std::vector<int32_t> inv;     // caller voxel id -> compact index
  std::vector<uint64_t> mask;   // occupied cells per brick
  std::vector<int32_t> start;   // first compact index per brick (plan-local)
  std::vector<int32_t> nbb;     // [n_bricks][6] face-neighbour bricks, -1 = none
  std::vector<int32_t> bcoord;  // [n_bricks][3] brick lattice coordinates
  // [n_bricks][6] lanes whose lattice face neighbour in direction d exists but is absent from
  // the caller's table (a cut link: that row of the operator leaves it out)
  std::vector<uint64_t> cut;
};

// Returns 0, or -1 (bad input) / -2 (a table entry that is not the adjacent lattice voxel)
// with `err` set.
int build_plan(int32_t n, const int32_t* ijk, const int32_t* nb, int32_t nb_stride, int bx, int by,
               int bz, BrickPlan& plan, std::string& err);

}  // namespace phs
