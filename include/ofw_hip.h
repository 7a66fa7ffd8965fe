/*
 * libofw_hip.so -- MI355X (gfx950) free-wake induced-velocity library: C ABI.
 *
 * The drop-in boundary for OpenFAWE's vortex batch backends.  Each entry point
 * replaces one engine of the reference (cited per declaration) and keeps its
 * argument meaning: targets (n,3), segment starts (m,3), unit directions (m,3),
 * lengths (m), circulations (m), core radii (m) in; velocities (n,3) and
 * per-target flags out.  Flags follow vortex_numba.py:32-34 (0 OK, 1 SINGULAR,
 * 2 NONFINITE; the first flag of a target in its summation order is kept).
 *
 * Conventions
 *   - Every array pointer is a DEVICE pointer (HBM) unless a name says host.
 *     Arrays are C-contiguous; float64 for geometry, int64 for indices as in
 *     the reference's NumPy engines.
 *   - Calls are enqueued on the context's stream and are synchronous on return
 *     (the host may read the outputs after the call returns), like the
 *     reference's engines, which synchronize every launch.
 *   - Errors are returned as status codes, never thrown; ofw_last_error()
 *     gives the message of the calling thread's last failure.
 *   - Results are deterministic: the same inputs and settings give the same
 *     bits on every call (fixed-order reductions, no floating-point atomics).
 *   - A context owns the scratch memory of its calls and is not shared
 *     between host threads.
 */
#ifndef OFW_HIP_H
#define OFW_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFW_ABI_VERSION 3

enum ofw_status {
  OFW_OK = 0,
  OFW_ERR_ARG = 1,         /* invalid argument (the reference raises ValueError) */
  OFW_ERR_HIP = 2,         /* HIP runtime failure */
  OFW_ERR_UNSUPPORTED = 3, /* e.g. a core model the tree code refuses (vortex_tree.py:593-594) */
  OFW_ERR_OVERFLOW = 4     /* traversal limit exceeded (vortex_cuda_tree.py:548-549) */
};

enum ofw_precision {
  OFW_FP64 = 0,  /* FP64 throughout */
  OFW_MIXED = 1, /* FP32 far field and far pairs, FP64 near pairs and accumulation */
  OFW_FP32 = 2   /* as MIXED with an FP32 accumulator */
};

enum ofw_stage {         /* RK stage arithmetic of ofw_stage_combine (wake.py:596-614), NumPy evaluation order */
  OFW_STAGE_STEP = 0,    /* out = x + h*a               Euler step, Heun predictor, RK4 stage 4 */
  OFW_STAGE_HALF = 1,    /* out = x + h*a/2             RK4 stages 2 and 3 */
  OFW_STAGE_HEUN = 2,    /* out = x + h*(a+b)/2         Heun corrector */
  OFW_STAGE_RK4 = 3,     /* out = x + h*(a+2b+2c+d)/6   RK4 combination */
  OFW_STAGE_ADD = 4      /* out = x + a                 ambient plus induced velocity */
};

enum ofw_step_status {          /* bits of a geometry step's status word (ofw_step_run) */
  OFW_STEP_PAIR_OVERFLOW = 1,   /* the incident pairs exceeded the plan's capacity */
  OFW_STEP_EDGE_INDEX = 2,      /* an edge names a node outside [0, n) */
  OFW_STEP_ZERO_LENGTH = 4,     /* a zero-length segment */
  OFW_STEP_EXCLUSION = 8,       /* an exclusion pair outside its ranges */
  OFW_STEP_TREE_OVERFLOW = 16,  /* the tree code exceeded a capacity of its plan */
  OFW_STEP_ROW_FLAG = 32,       /* a wing-row target has a flagged pair (the reference raises) */
  OFW_STEP_RING_FLAG = 64,      /* a ring target has a flagged pair (its 'zero' call must be repeated) */
  OFW_STEP_NONFINITE = 128      /* a non-finite velocity */
};

typedef struct ofw_ctx ofw_ctx;

/* ABI version of the library (OFW_ABI_VERSION). */
int ofw_abi_version(void);

/* Message of the calling thread's last failed call ("" when none). */
const char *ofw_last_error(void);

/* Create a context bound to a HIP stream (hipStream_t; NULL = the null stream) on the current device.
 * Replaces the module-level device state of vortex_cuda.py:53-82 (probe, kernels cache). */
int ofw_ctx_create(ofw_ctx **ctx, void *stream);
int ofw_ctx_destroy(ofw_ctx *ctx);
int ofw_ctx_set_stream(ofw_ctx *ctx, void *stream);

/* Phase timings of the context's last call, in milliseconds (host buffer of `count` floats), measured with
 * HIP events on the context's stream.  Returns the number of phases written.  The direct sum has three
 * phases (segment packing + exclusion conversion, the pair kernel, the chunk reduction).  Replaces the
 * Statistics timings of vortex_cuda.py:372-378 and vortex_cuda_tree.py:429-436. */
int ofw_ctx_timings(ofw_ctx *ctx, float *ms_host, int count);

/* Device probe: writes the device name (host buffer) and its compute-unit count.
 * Replaces vortex_cuda.available / probe_error (vortex_cuda.py:58-82). */
int ofw_device_info(char *name_host, int name_len, int *cu_count_host);

/* Direct sum over all segments for every target, in segment order per target, skipping the pairs of the
 * exclusion CSR (ex_ptr (n+1), ex_seg (e), per-target sorted unique segment ids; vortex_numba.py:228-243).
 * The ids of a target must be sorted and unique, as exclusion_csr makes them: the kernels walk them with a cursor
 * that advances in segment order, so an unsorted row would exclude other pairs than the reference does.
 * Replaces vortex_numba.direct_sum (vortex_numba.py:246-260) and vortex_cuda.cuda_sum (vortex_cuda.py:381-467).
 *   precision  ofw_precision
 *   block      threads per workgroup (32, 64, 128 or 256; CudaSettings.block, vortex.py:103)
 *   model      CoreModel 0..5 (vortex.py:21-27); raise_singular: singularity == "raise"
 *   flag_seg   optional (may be NULL): per target, the segment index of its first flagged pair, -1 if none
 * out (n,3) float64, flags (n) int8. */
int ofw_direct_sum(ofw_ctx *ctx, int precision, int block, int64_t n, const double *points, int64_t m,
                   const double *starts, const double *directions, const double *lengths, const double *gamma,
                   const double *radius, int model, int raise_singular, const int64_t *ex_ptr,
                   const int64_t *ex_seg, double *out, int8_t *flags, int32_t *flag_seg);

/* Bytes of the static scratch of ofw_direct_sum_static: one device array of this size per GPU, held from the
 * first context created on that GPU until the process ends. */
#define OFW_STATIC_SCRATCH_BYTES (128LL << 20)

/* ofw_direct_sum without allocating device memory.  It has no counterpart in the reference (whose engines
 * allocate on every call); it is for callers that cannot allocate: tight serving loops and harnesses that
 * forbid hipMalloc.  Same arguments, results and flags as ofw_direct_sum, and the same bits whenever the call
 * fits the scratch with ofw_direct_sum's own chunk count.  Its scratch (the packed segment records, 160 B per
 * segment, and the chunk partials, 28 B per target and segment chunk) is carved from a static device array of
 * OFW_STATIC_SCRATCH_BYTES on the context's GPU instead of the context's buffers; the chunk count is lowered
 * until the partials fit, and a call that does not fit with one chunk returns OFW_ERR_ARG (at most about
 * 800,000 segments).  It calls no hipMalloc or hipFree.  Every call on one GPU shares that array, which keeps the
 * last call's contents: calls that overlap on one GPU (several host threads, or several contexts and streams)
 * race.  Synchronous on return, like ofw_direct_sum. */
int ofw_direct_sum_static(ofw_ctx *ctx, int precision, int block, int64_t n, const double *points, int64_t m,
                          const double *starts, const double *directions, const double *lengths,
                          const double *gamma, const double *radius, int model, int raise_singular,
                          const int64_t *ex_ptr, const int64_t *ex_seg, double *out, int8_t *flags,
                          int32_t *flag_seg);

/* Counters of the context's last call (host buffer of `count` int64): [0] direct-sum segment chunks,
 * [1] segments per chunk, [2] traversal subtrees, [3] traversal split depth, [4] tree nodes, [5] tree depth,
 * [6] near-field work items of the traversal, [7] runs of the tree pipeline the last ofw_tree_sum took (2 when the
 * work-item slots kept from the call before were too few, or the call had more exclusion entries than the
 * capacity the context kept from its earlier calls).
 * Returns the number written. */
int ofw_ctx_stats(ofw_ctx *ctx, int64_t *stats_host, int count);

/* Median-split binary tree over segment midpoints mid (m,3), element for element the reference's
 * vortex_tree.build_tree (vortex_tree.py:208-241; serial form :51-115): order (m), and lo, hi, left,
 * right (node arrays of capacity 4*(m//leaf_size+1)+1, reference numbering, -1 children for leaves).
 * The node count is written to *count_host. */
int ofw_build_tree(ofw_ctx *ctx, int64_t m, const double *mid, int64_t leaf_size, int64_t *order, int64_t *lo,
                   int64_t *hi, int64_t *left, int64_t *right, int64_t *count_host);

/* Node quantities of the tree built on the segments: centre (c,3), bound (c), core (c), m0 (c,3),
 * m1 (c,3,3), m2 (c,3,3,3), w1 (c,3), t2 (c,3), w2 (c,3,3) in the reference numbering (arrays of the
 * capacity of ofw_build_tree).  Replaces vortex_tree.node_moments + node_contractions
 * (vortex_tree.py:309-376, 425-447) and vortex_cuda_moments.device_node_quantities
 * (vortex_cuda_moments.py:206-230).  Centre, bound and core are bit-exact; the moments are sums in
 * another order (equal to rounding). */
int ofw_tree_nodes(ofw_ctx *ctx, int64_t m, const double *starts, const double *directions, const double *lengths,
                   const double *gamma, const double *radius, int64_t leaf_size, double *centre, double *bound,
                   double *core, double *m0, double *m1, double *m2, double *w1, double *t2, double *w2,
                   int64_t *count_host);

/* Tree-code sum: build, node quantities and traversal on the GPU.  Replaces vortex_tree.tree_sum
 * (vortex_tree.py:588-626) and vortex_cuda_tree.cuda_tree_sum (vortex_cuda_tree.py:439-595).
 * theta, expansion_order, leaf_size, clearance: TreeSettings (vortex.py:30-55).  model must be 0, 1 or 2
 * (OFW_ERR_UNSUPPORTED otherwise, as vortex_tree.py:593-594 when a core is present).  The accepted nodes
 * are the reference's (FP64 acceptance test in every precision).  split_depth: depth of the traversal's
 * subtree frontier (-1 = chosen from the target count).  far_counts / near_counts (optional, may be NULL):
 * accepted nodes and evaluated near pairs per target.  The context reads the exclusion count ex_ptr[n] back
 * once and keeps a capacity (a quarter larger); later calls check it on the device instead of reading it,
 * and a call beyond it runs its pipeline again with the exact count. */
int ofw_tree_sum(ofw_ctx *ctx, int precision, int64_t n, const double *points, int64_t m, const double *starts,
                 const double *directions, const double *lengths, const double *gamma, const double *radius,
                 int model, int raise_singular, const int64_t *ex_ptr, const int64_t *ex_seg, double theta,
                 int expansion_order, int64_t leaf_size, double clearance, int split_depth, double *out,
                 int8_t *flags, int64_t *far_counts, int64_t *near_counts);

/* Batched tree sum over K independent wake snapshots packed back to back (no counterpart in the reference,
 * whose engines take one snapshot per call, vortex_tree.py:588 and vortex_cuda_tree.py:439).  Snapshot k owns
 * targets [target_offsets_host[k], target_offsets_host[k+1]) of points and segments
 * [segment_offsets_host[k], segment_offsets_host[k+1]) of the segment arrays (host int64 arrays of K+1
 * non-decreasing offsets from 0).  ex_ptr (n+1) / ex_seg: optional exclusion CSR over all targets with global
 * segment ids (NULL for none).  Other arguments as ofw_tree_sum; split_depth -1 = chosen from the total target
 * count.  Each snapshot's result equals ofw_tree_sum on that snapshot alone with the same split_depth, bit for
 * bit.  Creation runs the call once (sizing all scratch and checking every capacity) and, when use_graph is
 * set, records its whole launch sequence (build, node quantities, traversal and reduction of every snapshot)
 * in a HIP graph.  The pointers are fixed at creation; ofw_tree_batch_run replays the call on the current
 * contents of the same arrays, on the stream the context had at creation, synchronously.  Writes to the arrays
 * made on another stream must be ordered before the run by the caller (an event the creation stream waits on).
 * A replay whose inputs need more near-field work-item slots than the plan recorded returns OFW_ERR_OVERFLOW
 * (the velocities are then not valid); a new plan sizes them again. */
typedef struct ofw_tree_batch ofw_tree_batch;
int ofw_tree_batch_create(ofw_ctx *ctx, ofw_tree_batch **plan, int precision, int64_t K,
                          const int64_t *target_offsets_host, const int64_t *segment_offsets_host,
                          const double *points, const double *starts, const double *directions,
                          const double *lengths, const double *gamma, const double *radius, int model,
                          int raise_singular, const int64_t *ex_ptr, const int64_t *ex_seg, double theta,
                          int expansion_order, int64_t leaf_size, double clearance, int split_depth, int use_graph,
                          double *out, int8_t *flags, int64_t *far_counts, int64_t *near_counts);
int ofw_tree_batch_run(ofw_tree_batch *plan);
/* Timings (ms) of the plan's last run: one phase (the graph) when recorded, else the phases of ofw_tree_sum. */
int ofw_tree_batch_timings(ofw_tree_batch *plan, float *ms_host, int count);
/* Counters of the plan's last run, as ofw_ctx_stats. */
int ofw_tree_batch_stats(ofw_tree_batch *plan, int64_t *stats_host, int count);
int ofw_tree_batch_destroy(ofw_tree_batch *plan);

/* Unit-circulation velocity of every segment at every target: out (n,m,3), flags (n) (first flag in
 * segment order).  Replaces vortex_numba.pair_matrix (vortex_numba.py:271-287).  n <= 65,535. */
int ofw_pair_matrix(ofw_ctx *ctx, int64_t n, const double *points, int64_t m, const double *starts,
                    const double *directions, const double *lengths, const double *radius, int model,
                    int raise_singular, double *out, int8_t *flags);

/* out[p, j, :] = sum over k in [column_ptr[j], column_ptr[j+1]) of column_sign[k] * pairs[p, column_segment[k], :],
 * added sequentially from 0.0 in k order.  Replaces vortex_numba.assemble_columns (vortex_numba.py:290-308). */
int ofw_assemble_columns(ofw_ctx *ctx, int64_t n, int64_t m, const double *pairs, int64_t columns,
                         const int64_t *column_ptr, const int64_t *column_segment, const double *column_sign,
                         double *out);

/* Segment geometry from end points (the preparation of vortex.py:460-465): directions (m,3) and lengths (m)
 * from starts (m,3) and ends (m,3).  Returns OFW_ERR_ARG ("zero-length segment") when a length is 0. */
int ofw_segment_geometry(ofw_ctx *ctx, int64_t m, const double *starts, const double *ends, double *directions,
                         double *lengths);

/* Adjoint of the direct sum (no counterpart in the reference, which has no derivatives of the sum; its
 * VSM Jacobian vsm.py:333-349 consumes dV/dGamma built from pair_matrix).  Given grad_out = dL/dV (n,3):
 * grad_gamma (m) = dL/dGamma, grad_starts / grad_ends (m,3) = dL/d(segment end points), and optionally
 * grad_points (n,3) = dL/d(targets) and the fused forward velocities out (n,3) (either may be NULL).
 * Segment s runs from starts[s] to ends[s]; models 0-5 (vortex.py:21-27), the Rosenhead-Moore core included.
 * precision: OFW_FP64, or OFW_MIXED (pairs more than 4 core radii from the segment and away from its line in
 * FP32, near pairs and every accumulation in FP64; the criteria of the mixed direct sum).  The exclusion CSR is
 * that of ofw_direct_sum (sorted unique segment ids per target). */
int ofw_direct_vjp(ofw_ctx *ctx, int precision, int64_t n, const double *points, int64_t m, const double *starts,
                   const double *ends, const double *gamma, const double *radius, int model, const int64_t *ex_ptr,
                   const int64_t *ex_seg, const double *grad_out, double *grad_gamma, double *grad_starts,
                   double *grad_ends, double *grad_points, double *out);

/* Directed ring-edge aggregation of a wake lattice (replaces network_numba.pack_ring_edges, network_numba.py:
 * 112-117, and the loop of wake.pack_rings, wake.py:300-307).  Ring (i, j), 0 <= i < depth - 1, of panel j with
 * columns panel_edges[j] = (left, right) runs (i,left) -> (i,right) -> (i+1,right) -> (i+1,left) over node ids
 * row * width + column with circulation strengths[i * panels + j].  Every edge is stored once as (lo, hi), lo < hi,
 * with the sum, from 0.0 and in ring order, of +strength for each ring that runs it lo -> hi and -strength
 * otherwise; edges come in the order of their first appearance and exact-zero sums are omitted.  Sets *count_out
 * (host) to the number of edges and, when it is at most capacity, writes edges (capacity,2) int64 and gamma
 * (capacity) float64 (device); OFW_ERR_OVERFLOW when the count exceeds capacity (nothing written), OFW_ERR_ARG for
 * a panel column outside [0, width) or a panel joining a column to itself.  Synchronises the stream once. */
int ofw_pack_ring_edges(ofw_ctx *ctx, int64_t depth, int64_t width, int64_t panels, const int64_t *panel_edges,
                        const double *strengths, int64_t *edges, double *gamma, int64_t capacity, int64_t *count_out);

/* Material core radius of every wake edge (replaces cores.edge_radii, cores.py:129-149).  Node ids (edges (m,2)
 * plus `offset`) are row * width + column over the rows bound, trailing edge, free rows.  A spanwise edge of free
 * row r >= 2 takes span[(r-2) * panels + j] for the (last) panel j of panel_edges (panels,2) joining its columns; a
 * streamwise edge between rows r and r+1 at column c with r >= 1 takes stream[(r-1) * width + c]; every other
 * bound or trailing-edge edge takes `initial`.  span (span_rows, panels) and stream (stream_rows, width) are the
 * WakeCores arrays.  OFW_ERR_ARG for an edge that is neither spanwise nor streamwise, a spanwise edge without a
 * panel, or a row beyond the core arrays.  Synchronises the stream once. */
int ofw_edge_radii(ofw_ctx *ctx, int64_t m, const int64_t *edges, int64_t offset, int64_t width, int64_t panels,
                   const int64_t *panel_edges, int64_t span_rows, const double *span, int64_t stream_rows,
                   const double *stream, double initial, double *out);

/* Incident-exclusion pairs of a wake network (replaces network_numba.incident_pairs, network_numba.py:242-253,
 * and the loop of wake.py:427-438): for the free points nodes[first:] (n,3), every (point, segment) pair in
 * which the point's coordinates equal an end of segment s (edges (m,2)) bit for bit, in segment order, the
 * start end before the end end, and for one end the coinciding points in increasing order; a node with a
 * NaN coordinate never matches.  Point indices count from `first`.  Sets *count_out (host) to the number of
 * pairs and, when it is at most capacity, writes them to pairs (capacity,2) int64 (device); returns
 * OFW_ERR_OVERFLOW when count > capacity (nothing written: call again with room for *count_out pairs) and
 * OFW_ERR_ARG when an edge names a node outside [0, n), n or m exceeds 2^29 or the count exceeds 2^31 - 1.
 * Synchronises the stream once (to read the count). */
int ofw_incident_pairs(ofw_ctx *ctx, int64_t n, const double *nodes, int64_t m, const int64_t *edges, int64_t first,
                       int64_t *pairs, int64_t capacity, int64_t *count_out);

/* Segment preparation of segments_velocity (vortex.py:458-465) from a node/edge network: starts (m,3) =
 * nodes[edges[:,0]], delta = nodes[edges[:,1]] - starts, lengths (m) = hypot(hypot(dx,dy),dz), directions (m,3) =
 * delta / lengths.  OFW_ERR_ARG ("connectivity index out of range", "zero-length segment") as the reference.
 * Synchronises the stream once (for the errors). */
int ofw_network_segments(ofw_ctx *ctx, int64_t n, const double *nodes, int64_t m, const int64_t *edges,
                         double *starts, double *directions, double *lengths);

/* Exclusion CSR of target ranges (the subsets of wake.py:440-487 followed by exclusion_csr, vortex_numba.py:
 * 180-243): from (point, segment) pairs (npairs,2) in any order, the pairs whose point p lies in [lo, hi), as
 * output point p - lo, or point_map[p - lo] when point_map (hi-lo) is given (negative: dropped; several target
 * ranges packed into one sum), and, when remap (m) is given, with segments mapped through it (negative: dropped).
 * ptr_out (count+1) and seg_out (capacity npairs) receive per output point its sorted unique segments; count must
 * be hi - lo without a point map.  segment_ordered != 0 promises that the kept pairs' (remapped) segments are
 * non-decreasing in pair order (as ofw_incident_pairs writes them), which saves the sort by segment.  OFW_ERR_ARG
 * ("exclusion index out of range") when a kept pair names a segment outside [0, m) or the map an output point at
 * or beyond count. */
int ofw_exclusion_subset(ofw_ctx *ctx, int64_t npairs, const int64_t *pairs, int64_t lo, int64_t hi, int64_t count,
                         const int64_t *point_map, int64_t m, const int64_t *remap, int segment_ordered,
                         int64_t *ptr_out, int64_t *seg_out);

/* Stage arithmetic of convect_geometry over count doubles (enum ofw_stage; b, c, d may be NULL where unused). */
int ofw_stage_combine(ofw_ctx *ctx, int mode, int64_t count, double h, const double *x, const double *a,
                      const double *b, const double *c, const double *d, double *out);

/* A free-wake geometry step recorded as one graph (no counterpart in the reference, whose convect_geometry,
 * wake.py:586-616, calls its engines stage by stage).  The plan advances the free points points_in (k,3) by
 * `substeps` steps of the integrator (0 Euler, 1 Heun, 2 RK4; dt_s / substeps each) into points_out (k,3), with
 * the stage velocity of _network_convection (wake.py:404-491): the uniform inflow (device, 3 doubles) plus, at the
 * targets, the tree sum over the network concatenate(bound (first,3), free points) with edges (m,2), gamma, radius
 * (m), the incident segments excluded; targets are the wing rows [0, rows) and the ring nodes [ring_lo, ring_hi)
 * of the free points, every other free point keeps the inflow.  drift (k,3; NULL: none) is added to every free
 * point's stage velocity after that sum: the rings' sheet drift (wake.py:582-588), zero outside the drifting ring
 * nodes.  The bound row moves by time fraction times displacement (first,3; NULL: fixed).  Every pointer is
 * kept: a run reads the arrays' current contents.
 * engine 0 is the tree code (theta, expansion_order, leaf_size, clearance; models 0-2), engine 1 the direct sum
 * (block as ofw_direct_sum; models 0-5); both in `precision`.  Creating runs the step twice (measuring the pair and
 * list capacities, then with them) and records it when use_graph != 0; stats_host (3 int64, may be NULL) receives
 * the status of the last run and the two capacities. */
typedef struct ofw_step ofw_step;
int ofw_step_create(ofw_ctx *ctx, ofw_step **plan, int precision, int64_t first, const double *bound,
                    const double *displacement, int64_t k, const double *points_in, double *points_out, int64_t m,
                    const int64_t *edges, const double *gamma, const double *radius, int model, int64_t rows,
                    int64_t ring_lo, int64_t ring_hi, double theta, int expansion_order, int64_t leaf_size,
                    double clearance, const double *inflow, const double *drift, double time, double dt_s,
                    int substeps, int integrator, int engine, int block, int use_graph, int64_t *stats_host);
/* Run the step (one graph launch) and read its status block (one synchronisation): status_host[0] the
 * ofw_step_status bits, [1] the largest incident-pair count, [2] the largest list work-item count.  A capacity bit
 * asks for a new plan; a flag, zero-length or non-finite bit for the stage-by-stage step, which raises or repeats a
 * ring call as the reference. */
int ofw_step_run(ofw_step *plan, int64_t *status_host);
int ofw_step_timings(ofw_step *plan, float *ms_host);   /* the last run's time on the stream, ms */
int ofw_step_destroy(ofw_step *plan);

#ifdef __cplusplus
}
#endif
#endif /* OFW_HIP_H */
