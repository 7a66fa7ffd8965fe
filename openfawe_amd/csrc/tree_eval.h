// Traversal tasks and the tree-evaluation driver.
#pragma once
#include <stdint.h>

#include <vector>

#include "internal.h"
#include "tree.h"
#include "tree_records.h"

namespace ofw {

static constexpr int MAX_ANC = 14;

// One frontier subtree of the traversal: its root (pre-order index) and the root's ancestors, root first.
struct EvalTask {
  int32_t node;
  int32_t nanc;
  int32_t anc[MAX_ANC];
};

void frontier_tasks(const HostTree& t, int depth, std::vector<EvalTask>& tasks);

int tree_prep(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths, double* mx,
              double* my, double* mz);
int tree_moments(Ctx& c, const HostTree& t, const int32_t* order, const double* starts, const double* directions,
                 const double* lengths, const double* gamma, const double* radius, int model, double clearance,
                 TreeDevice& td, bool want_outputs);
int tree_evaluate(Ctx& c, const HostTree& t, const TreeDevice& td, const int32_t* order, int precision, int64_t n,
                  const double* points, int model, int raise_singular, const int32_t* ex_ptr,
                  const int32_t* ex_seg, double theta, int expansion_order, int split_depth, double* out,
                  int8_t* flags, int64_t* far_counts, int64_t* near_counts, std::vector<EvalTask>& tasks);

}  // namespace ofw
