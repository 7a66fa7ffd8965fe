// Host-side tree structures shared by the tree build, moments and traversal.
#pragma once
#include <stdint.h>

#include <vector>

#include "internal.h"

namespace ofw {

struct HostTree {
  int64_t m = 0, count = 0;
  int depth = 0;
  // creation numbering of the level build (vortex_tree.py:244-268)
  std::vector<int32_t> lo, hi, left, right;
  // the reference's numbering (vortex_tree.py:271-306)
  std::vector<int32_t> ref_lo, ref_hi, ref_left, ref_right;
  // pre-order layout (left child first): the visit order of the reference's stack walk
  std::vector<int32_t> pre_ref, pre_skip, pre_parent, pre_depth;
};

int build_tree(Ctx& c, int64_t m, const double* mx, const double* my, const double* mz, int64_t leaf,
               int32_t* order, HostTree& t);
int finish_tree(HostTree& t);

}  // namespace ofw
