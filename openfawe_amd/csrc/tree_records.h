// Device-side record layouts of the direct sum and the tree code.
#pragma once
#include <stdint.h>

#include <vector>

#include "common.h"

namespace ofw {

// Pack one segment's FP64 and FP32 pair records (slots in common.h) from the reference's engine inputs; either
// output may be null.
__device__ __forceinline__ void pack_record(int s, const double* __restrict__ starts,
                                            const double* __restrict__ directions,
                                            const double* __restrict__ lengths, const double* __restrict__ gamma,
                                            const double* __restrict__ radius, int model, double* __restrict__ r,
                                            float* __restrict__ f) {
  const double L = lengths[s], g = gamma[s], rad = radius[s];
  const double invL = 1.0 / L;
  const double g4 = (g / L) / OFW_FOUR_PI;
  const double rc = rad / L;
  const double rc2 = rc * rc;
  const bool cored = rad > 0.0 && model != OFW_MODEL_NONE;
  const bool ratio_model = model == OFW_MODEL_RANKINE || model == OFW_MODEL_LAMB || model == OFW_MODEL_VATISTAS;
  const double exact = (cored && ratio_model && !(rc >= OFW_FAST_RC_MIN)) ? 1.0 : 0.0;
  if (r) {
  r[R_SX] = starts[3 * (size_t)s];
  r[R_SY] = starts[3 * (size_t)s + 1];
  r[R_SZ] = starts[3 * (size_t)s + 2];
  r[R_DX] = directions[3 * (size_t)s];
  r[R_DY] = directions[3 * (size_t)s + 1];
  r[R_DZ] = directions[3 * (size_t)s + 2];
  r[R_L] = L;
  r[R_INVL] = invL;
  r[R_G4] = g4;
  r[R_RC] = rc;
  r[R_RC2] = cored ? rc2 : 0.0;
  r[R_GAMMA] = g;
  r[R_RADIUS] = rad;
  r[R_EXACT] = exact;
  r[R_IRC2] = cored ? 1.0 / rc2 : 0.0;
  r[R_PAD1] = 0.0;
  }
  if (!f) return;
  // a segment whose factors leave the normal FP32 range always takes the FP64 path
  const double ag = fabs(g4);
  const bool f32_ok = exact == 0.0 && invL < 1e30 && invL > 1e-30 && ag < 1e30 && (ag == 0.0 || ag > 1e-30) &&
                      (!cored || (rc2 < 1e30 && rc2 > 1e-30));
  f[F_DX] = (float)directions[3 * (size_t)s];
  f[F_DY] = (float)directions[3 * (size_t)s + 1];
  f[F_DZ] = (float)directions[3 * (size_t)s + 2];
  f[F_INVL] = (float)invL;
  f[F_G4] = (float)g4;
  f[F_RC2] = cored ? (float)rc2 : 0.0f;
  f[F_EXACT] = f32_ok ? 0.0f : 1.0f;
  f[F_PAD0] = 0.0f;
}

// The tree near field's FP32 record (Near32, common.h) of segment s from the engine inputs; the length vector is
// the FP64 product direction * length, rounded.
__device__ __forceinline__ void pack_near32(int s, const double* __restrict__ directions,
                                            const double* __restrict__ lengths, const double* __restrict__ gamma,
                                            const double* __restrict__ radius, int model, float* __restrict__ f) {
#pragma clang fp contract(off)
  const double L = lengths[s], g = gamma[s], rad = radius[s];
  const double gp = g / OFW_FOUR_PI;
  const bool cored = rad > 0.0 && model != OFW_MODEL_NONE;
  const double lr = L * rad;
  const double lr2 = cored ? lr * lr : 0.0;
  const double L2 = L * L;
  const bool ratio_model = model == OFW_MODEL_RANKINE || model == OFW_MODEL_LAMB || model == OFW_MODEL_VATISTAS;
  const bool tiny_core = cored && ratio_model && !(rad / L >= OFW_FAST_RC_MIN);  // the FP64 pair's exact path
  const double ag = fabs(gp);
  const bool f32_ok = !tiny_core && L2 >= OFW_NEAR_L2_MIN && L2 <= OFW_NEAR_L2_MAX &&
                      (ag == 0.0 || (ag > 1e-30 && ag < 1e30)) && (!cored || (lr2 > 1e-30 && lr2 < 1e30));
  f[N_LX] = (float)(directions[3 * (size_t)s] * L);
  f[N_LY] = (float)(directions[3 * (size_t)s + 1] * L);
  f[N_LZ] = (float)(directions[3 * (size_t)s + 2] * L);
  f[N_L2] = (float)L2;
  f[N_GP] = (float)gp;
  f[N_QLR2] = (float)(OFW_LAMB_Q * lr2);
  f[N_EXACT] = f32_ok ? 0.0f : 1.0f;
  f[N_BAND] = cored ? (float)((double)OFW_LAMB_BAND32 * OFW_LAMB_Q * lr2) : -1.0f;
}

// Moment inputs in tree order (SoA): start, length vector, radius, gamma.  The end point and alpha = gamma x
// length vector are formed where they are read, by the same FP64 operations as vortex_tree.py:608-611.
struct MomentInputs {
  double* s[3];
  double* l[3];
  double* rad;
  double* gam;
  __device__ __forceinline__ double end(int c, int k) const {
#pragma clang fp contract(off)
    return s[c][k] + l[c][k];
  }
  __device__ __forceinline__ double alpha(int c, int k) const {
#pragma clang fp contract(off)
    return gam[k] * l[c][k];
  }
};

// Where a tree-order pair record comes from: the segment of tree position k is order[k] in the engine inputs.
struct SegSource {
  const int32_t* order = nullptr;
  const double *starts = nullptr, *directions = nullptr, *lengths = nullptr, *gamma = nullptr, *radius = nullptr;
  int model = 0;
};

// Optional full per-node outputs (pre-order node index), as node_moments / node_contractions return them
struct NodeOutputs {
  double* bound = nullptr;
  double* m0 = nullptr;
  double* m1 = nullptr;
  double* m2 = nullptr;
  double* w1 = nullptr;
  double* t2 = nullptr;
  double* w2 = nullptr;
};

// Acceptance record of a node (8 doubles, 64 bytes): centre, bounding radius, clearance * largest core, and
// (lo, hi), (skip, tree), (right child or -1, 0) as int32 pairs in the last three slots.
enum { A_CX = 0, A_CY, A_CZ, A_BOUND, A_CLEAR, A_INTS, A_INTS2, A_INTS3, ACC_REC = 8 };

// Far-field record (48 values): m0 (3), m1 (9, row-major), w1 (3), symmetric m2 (3 x 6: 00 01 02 11 12 22,
// the off-diagonal entries doubled), t2 (3), w2 (9, row-major), pad
enum { FR_M0 = 0, FR_M1 = 3, FR_W1 = 12, FR_M2 = 15, FR_T2 = 33, FR_W2 = 36, FR_PAD0 = 45, FR_PAD1 = 46, FR_PAD2 = 47,
       FAR_REC = 48 };

struct TreeDevice {
  int count = 0;
  double* rec64 = nullptr;  // FP64 pair records in tree order (FP64 evaluations only)
  float* rec32 = nullptr;   // FP32 near-field records (Near32) in tree order
  const double* start[3] = {};  // segment starts in tree order
  const double* lenv[3] = {};   // segment length vectors in tree order
  SegSource src;            // the engine inputs, for the FP64 record of one pair
  double* centre = nullptr;  // pre-order
  double* core = nullptr;
  double* acc = nullptr;
  double* far64 = nullptr;
  float* far32 = nullptr;
  NodeOutputs out;
  std::vector<int32_t> h_items[3], h_nodes[5];  // host staging kept alive until the call synchronises
};

}  // namespace ofw
