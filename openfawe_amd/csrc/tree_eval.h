// Traversal tasks and the tree-evaluation driver.
#pragma once
#include <stdint.h>

#include <vector>

#include "internal.h"
#include "tree.h"
#include "tree_records.h"

namespace ofw {

static constexpr int MAX_ANC = 14;

// Frontier subtrees near their tree's targets (whose walks open most of their nodes) are split EXTRA_DEPTH levels
// deeper than the others, so that no single walk sets the walk kernel's end.
static constexpr int EXTRA_DEPTH = 3;

// One frontier subtree of the traversal: its root (pre-order index), the root's ancestors (root of the tree first),
// and its interaction lists: one of list_cap entries per group of 64 targets, from list_off.
struct EvalTask {
  int32_t node;
  int32_t nanc;
  int32_t anc[MAX_ANC];
  int32_t list_off;
  int32_t list_cap;
};

// Per-call shape of the targets of a forest: targets of tree k are [target_off[k], target_off[k+1]).
struct EvalShape {
  int K = 1;
  int64_t n = 0;
  std::vector<int64_t> target_off;  // K+1
  std::vector<int64_t> seg_off;     // K+1
  int split_depth = 0;
  // near-field work-item slots: < 0 = sized from the device count of this call (one synchronisation, the
  // count is left in the context's stats[6]); >= 0 = fixed (no synchronisation; more work items overflow)
  int64_t chunk_slots = -1;
};

int tree_prep(Ctx& c, int64_t m, const double* starts, const double* directions, const double* lengths, double* mx,
              double* my, double* mz);
// Node quantities and traversal records of a built forest; `offsets` are the trees' segment offsets (K+1).
int tree_moments(Ctx& c, int buf, const TreeArrays& t, const std::vector<int64_t>& offsets, const double* starts,
                 const double* directions, const double* lengths, const double* gamma, const double* radius, int model,
                 double clearance, TreeDevice& td, bool want_outputs, bool want_rec64);
int tree_evaluate(Ctx& c, int buf, const TreeArrays& t, const TreeDevice& td, const EvalShape& shape, int precision,
                  const double* points, int model, int raise_singular, const int32_t* ex_ptr, const int32_t* ex_seg,
                  int64_t n_ex, double theta, int expansion_order, double* out, int8_t* flags, int64_t* far_counts,
                  int64_t* near_counts);
// Frontier depth that gives the traversal enough wavefronts: `groups` = target groups of 64 over all trees.
int choose_split_depth(const Ctx& c, int64_t groups, int K);

// One tree sum over K independent snapshots (K = 1 for ofw_tree_sum).  With a fixed near-field capacity
// (chunk_slots >= 0) it only enqueues: no host synchronisation, no memset, and no allocation once the
// context's buffers have the sizes of an earlier identical call, so the whole sequence can be captured in a
// graph.  With chunk_slots < 0 the capacity is read from the device once, after the walk.
struct TreeCall {
  int precision = OFW_MIXED;
  int K = 1;
  int64_t n = 0, m = 0;
  std::vector<int64_t> toff, soff;
  const double *points = nullptr, *starts = nullptr, *directions = nullptr, *lengths = nullptr;
  const double *gamma = nullptr, *radius = nullptr;
  int model = 0, raise_singular = 0;
  const int64_t *ex_ptr = nullptr, *ex_seg = nullptr;
  int64_t n_ex = 0;          // exclusion entries, or (ex_capacity) an upper bound the device checks
  bool ex_capacity = false;
  double theta = 0.5;
  int order = 2;
  int leaf = 32;
  double clearance = 0.0;
  int split_depth = 0;
  int64_t chunk_slots = -1;  // near-field work-item slots (< 0: sized from this call's count)
  double* out = nullptr;
  int8_t* flags = nullptr;
  int64_t *far_counts = nullptr, *near_counts = nullptr;
};

// Enqueue a whole tree sum (midpoints, build, node quantities, traversal), or its two halves: the build and node
// quantities need the segments only, the traversal the targets and their exclusions as well.
int tree_pipeline(Ctx& c, const TreeCall& tc, TreeArrays& t);
int tree_pipeline_build(Ctx& c, const TreeCall& tc, TreeArrays& t, TreeDevice& td);
int tree_pipeline_eval(Ctx& c, const TreeCall& tc, TreeArrays& t, const TreeDevice& td);

}  // namespace ofw
