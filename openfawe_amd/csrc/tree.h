// Device-side tree (a forest of one or more trees over disjoint segment ranges) and its build.
#pragma once
#include <stdint.h>

#include <vector>

#include "internal.h"

namespace ofw {

// Device counters of a build (one int32 array; graph-capturable: nothing is read back during the build).
enum {
  C_NODES = 0,      // nodes created (global ids are dense from 0; some heap slots may stay unused)
  C_LEVEL,          // roots of the forest (the first top level)
  C_DUP,            // the list build found a run of equal coordinates too long to re-order: sort build instead
  C_TILES,          // tiles of the current level
  C_BIG,            // nodes of the current level sorted by tiles
  C_BOTTOM,         // subtree roots of the bottom phase
  C_HEAP,           // heap slots reserved by the bottom phase
  C_OVERFLOW,       // capacity exceeded (checked by the host after the call)
  C_FINAL,          // existing nodes (after the finish)
  C_ITEMS,          // moment items
  C_TASKS,          // traversal subtrees (all trees)
  C_WORK,           // traversal work items
  C_DEPTH,          // largest node depth
  C_CHUNKS,         // near-field work items of the traversal
  C_SUBTREES,       // subtrees of the moment pass
  C_STRAND,         // top-phase nodes whose segments stay in the second order buffer
  C_RUNS,           // runs of equal coordinates in the roots' lists (list build)
  C_RUNMEM,         // their members
  C_EXNEED,         // a call's exclusion count when it exceeds the capacity the call assumed (0: it fits)
  C_COUNT = 19
};
static constexpr int LEVEL_MAX = 64;  // top levels of the list build (a tree of 2^28 segments has at most 18)

// A device count, or 0 once a capacity has overflowed: every kernel after an overflow does no work, and the
// host reports the overflow when the call has finished.
__device__ __forceinline__ int live_count(const int32_t* counters, int which) {
  return counters[C_OVERFLOW] ? 0 : counters[which];
}

struct TreeArrays {
  // build inputs
  int K = 1;                     // trees in the forest
  int64_t m = 0;
  int leaf = 32;
  int cap = 0;                   // node capacity
  int tile_cap = 0;
  int bottom_cap = 0;
  int32_t* roots_lo = nullptr;   // (K+1) segment offsets of the trees
  // node table (global ids)
  int32_t *lo, *hi, *left, *right, *parent, *depth, *tree;
  int8_t *axis, *split;
  int32_t* counters;
  int32_t *level_a, *big, *big_alt, *strand, *p_flag;
  int32_t *tile_node, *tile_p0, *tile_p1;
  double* box_partial;
  uint64_t *keyA, *keyB;
  uint32_t *posA, *posB;
  int32_t* tmp;                  // second order buffer of the top phase (odd depths)
  int32_t *bottom, *heap_base;
  // finish: pre-order layout
  int32_t *lo_cnt, *lo_mind, *lo_base, *scan_part;
  int32_t* pre;                  // node id -> pre-order index
  int32_t *p_lo, *p_hi, *p_skip, *p_parent, *p_depth, *p_tree, *p_node, *p_irank;
  int32_t* tree_first;           // (K+1) first pre-order index of each tree
  int32_t* order = nullptr;      // segment permutation (tree order -> segment)
  const double* mid[3] = {};     // the midpoints the tree is built on (segment order), for the traversal plan
  // list build: per-axis segment lists presorted once, split by stable partitions level after level
  int lists = 0;                 // 1: the top phase uses the lists (every node's segments stay in `order`)
  int rs_passes = 0;             // radix passes of the presort (4 key digits + the tree-id digits)
  int rs_tiles = 0;              // radix tiles per axis
  uint32_t* rkey[2] = {};        // 3 x m keys per buffer (axis-major): coordinates rounded to float32
  int32_t* rval[2] = {};         // 3 x m segment ids per buffer; the lists live here after the presort
  unsigned long long* lstat = nullptr;  // tile_cap x 3 look-back status words of the partition (tagged by level)
  int32_t* tickets = nullptr;    // LEVEL_MAX work-unit tickets, one per top level
  // the finish's two single-pass scans (over lo and over the split flags): tickets and look-back status words
  int32_t* scan_ticket = nullptr;
  unsigned long long* scan_status = nullptr;
  int scan_words = 0;
  int32_t* rcount = nullptr;     // 3 x 256 x rs_tiles digit counts of one presort pass, then their scans
  int32_t* rtotal = nullptr;     // 3 x 256 digit totals of the pass
  // runs of equal coordinates (at most 3m members in 3m/2 runs)
  int32_t* run_seg = nullptr;    // members of every run, run after run
  int32_t* run_off = nullptr;    // first member of each run
  int32_t* run_len = nullptr;    // length << 2 | axis
  uint8_t* tie = nullptr;        // per segment: 1 when it is in a run of any axis
  int32_t* inv_o = nullptr;      // per run segment: its position in `order`
  int32_t* invb = nullptr;       // 3 x m, per run segment: its position in each list
  int32_t* ord_inv = nullptr;    // per segment of a bottom subtree: its position in `order`
};

// Copy host int64 values (each below 2^31) to a device int32 array through kernel arguments: no host
// memory is read at execution time, so the copy can be captured in a graph.
int upload_i32(Ctx& c, const std::vector<int64_t>& values, int32_t* dst);

// Exclusive scan of n int32 values (three kernels); the total goes to *total_out when it is not null.
void scan_exclusive(Ctx& c, int n, const int32_t* in, int32_t* out, int32_t* part, int32_t* total_out);
// The same in one launch (single-pass look-back): *ticket and scan_lb_words(n) status words must be zero when it
// starts; a look-back that gives up sets counters[C_OVERFLOW].
int scan_lb_words(int64_t n);
void scan_exclusive_lb(Ctx& c, int n, const int32_t* in, int32_t* out, int32_t* ticket, unsigned long long* status,
                       int32_t* total_out, int32_t* counters);

// Stable LSD radix sort of (key, val) int32 pairs of n items by the low 8 * passes bits of key (8-bit digits);
// returns the buffer index (0 or 1) holding the result.  hist and hist_off hold radix_hist_size(n) entries, part
// the scan partials of that many items.
int radix_sort(Ctx& c, int n, int passes, int32_t* key[2], int32_t* val[2], int32_t* hist, int32_t* hist_off,
               int32_t* part);
// Number of 8-bit passes for keys in [0, top]; the histogram entries of a sort of n items.
int radix_passes(int64_t top);
int radix_hist_size(int n);

// Enqueue-only forms of the network bookkeeping (no host synchronisation, no allocation once the context's
// buffers have the sizes of an earlier identical call), for recorded graphs:
// incident pairs, at most `capacity` written; *status_out -> [0] the 64-bit pair total, [1] a bad edge.
int incident_pairs_enqueue(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, int64_t first,
                           int64_t* pairs, int64_t capacity, const unsigned long long** status_out);
// segment geometry; OFW_STEP_EDGE_INDEX / OFW_STEP_ZERO_LENGTH or-ed into *status.
int network_segments_enqueue(Ctx& c, int64_t n, const double* nodes, int64_t m, const int64_t* edges, double* starts,
                             double* directions, double* lengths, int32_t* status);
// exclusion CSR of target ranges (only pairs below *live when live is given); OFW_STEP_EXCLUSION into *status.
int exclusion_subset_enqueue(Ctx& c, int64_t npairs, const int64_t* pairs, const unsigned long long* live, int64_t lo,
                             int64_t hi, int64_t count, const int64_t* point_map, int64_t m, const int64_t* remap,
                             int segment_ordered, int64_t* ptr_out, int64_t* seg_out, int32_t* status);

// Reserve and lay out every array of a build of m segments in K trees with leaf size `leaf`.
int tree_layout(Ctx& c, int buf, int64_t m, int K, int leaf, TreeArrays& t);
// Enqueue the whole build (no host synchronisation): midpoints mx/my/mz (m), tree segment offsets on the
// host (K+1, for the launch bounds) and on the device (t.roots_lo).
int tree_build(Ctx& c, TreeArrays& t, const double* mx, const double* my, const double* mz,
               const std::vector<int64_t>& offsets);
// Reference numbering of the nodes of tree k (host arrays of its node count), from the device layout.
int tree_reference_arrays(Ctx& c, const TreeArrays& t, std::vector<int64_t>& lo, std::vector<int64_t>& hi,
                          std::vector<int64_t>& left, std::vector<int64_t>& right, std::vector<int32_t>& ref_of_pre,
                          int64_t& count);

}  // namespace ofw
