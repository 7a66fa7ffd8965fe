// Shared definitions of the MI355X (gfx950) free-wake velocity library.
//
// Segment records.  The per-segment inputs of the reference engine
// (starts, directions, lengths, gamma, radius; vortex_numba.py:246) are packed
// once per call into 16-double records (128 B, one cache line pair) so that a
// wavefront that walks segments in lock-step reads each record with scalar
// loads (s_load_dwordx16) into SGPRs: the segment operand of every pair is
// wave-uniform, the target operand is per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OFW_FLAG_OK 0
#define OFW_FLAG_SINGULAR 1
#define OFW_FLAG_NONFINITE 2

#define OFW_MODEL_NONE 0
#define OFW_MODEL_RANKINE 1
#define OFW_MODEL_LAMB 2
#define OFW_MODEL_VATISTAS 3
#define OFW_MODEL_OFFSET 4
#define OFW_MODEL_RM 5

// FP64 segment record slots
enum {
  R_SX = 0, R_SY, R_SZ, R_DX, R_DY, R_DZ, R_L, R_INVL, R_G4, R_RC, R_RC2, R_GAMMA, R_RADIUS, R_EXACT,
  R_IRC2, R_PAD1, REC64 = 16
};
// FP32 segment record slots (16 floats, 64 B)
enum {
  F_DX = 0, F_DY, F_DZ, F_INVL, F_G4, F_RC2, F_EXACT, F_PAD0, REC32 = 8
};

#define OFW_AXIS_EPS (8.0 * 2.220446049250313e-16)   // 8 * np.finfo(np.float64).eps, vortex_numba.py:28
#define OFW_LAMB 1.25643

// Threshold bands of the fast FP64 pair.  The axis guard (h <= 8 eps max(1, |a|, |b|), vortex_numba.py:59) and the
// Lamb-Oseen cutoff (h / rc <= sqrt(10 / 1.25643), vortex_numba.py:95-97) are decided on squares only where the
// squares are more than 1e-12 (relative) from the threshold; the squares and the reference's hypot norms differ
// by a few ulps (below 1e-15), so the decision there is the reference's.  Inside the band the decision is taken
// with the reference's operations (hypot norms, h / rc) on the wavefront's ballot.
#define OFW_BAND_HI (1.0 + 1e-12)
#define OFW_BAND_LO (1.0 - 1e-12)
#define OFW_LAMB_Q (10.0 / OFW_LAMB)                  // (h / rc)^2 at the cutoff
#define OFW_FOUR_PI (4.0 * 3.141592653589793)        // 4 * np.pi, vortex_numba.py:30

// Fast FP64 path limits: outside them the pair is evaluated operation by operation as the reference
// does (pair_exact64).  |a| <= 1e100 segment lengths keeps every square finite and normal; a core of
// at least 1e-50 segment lengths keeps (h/rc)^2 finite.
#define OFW_FAST_A2_MAX 1e200
#define OFW_FAST_A2_MIN 1e-200
#define OFW_FAST_RC_MIN 1e-50

// Mixed precision: a pair goes to FP64 when the target is beside the segment within NEAR_BESIDE of its line,
// or within NEAR_AXIS of its line anywhere (distances relative to max(L, |p - s|, |p - e|), as the axis guard
// scales them).  NEAR_AXIS is 1e4 times the FP32 rounding of the cross product, so every pair the reference's
// FP64 axis guard catches (8 eps64 relative) reaches the FP64 path and takes the same branch there.  A target
// on the line's extension beyond either end needs no FP64: there the denominator does not cancel and the pair
// velocity vanishes with the distance to the line.  Core pairs: where the cross product is formed in FP32
// (the direct sum), a target within NEAR_CORES core radii of the segment takes FP64; where it is formed in
// FP64 (the tree's near field), only a target within LAMB_BAND32 of the Lamb-Oseen cutoff does.
#define OFW_NEAR_CORES2 16.0f          // (4 core radii)^2
#define OFW_NEAR_BESIDE2 1e-6f         // (1e-3)^2
#define OFW_NEAR_AXIS2 1e-8f           // (1e-4)^2
#define OFW_LAMB_BAND32 1e-4f          // relative band of q about the Lamb-Oseen cutoff that takes FP64
#define OFW_F32_A2_MAX 1e30f

struct PairOut {
  double x, y, z;
  int flag;
};

__device__ __forceinline__ double ofw_norm3(double x, double y, double z) { return hypot(hypot(x, y), z); }

// Operation-by-operation restatement of vortex_numba.py:42-109 (pair_velocity; OpenFAWE, Apache-2.0, see NOTICE);
// the FP64 fallback for
// extreme ranges (huge |a|, tiny cores).  Returns the pair velocity and its flag by value (nothing passes
// through memory on the call).
__device__ __noinline__ PairOut pair_exact64(double px, double py, double pz, const double* __restrict__ r,
                                             int model, int raise_singular) {
#pragma clang fp contract(off)
  const double sx = r[R_SX], sy = r[R_SY], sz = r[R_SZ];
  const double dx = r[R_DX], dy = r[R_DY], dz = r[R_DZ];
  const double length = r[R_L], gamma = r[R_GAMMA], radius = r[R_RADIUS];
  double ax = (px - sx) / length, ay = (py - sy) / length, az = (pz - sz) / length;
  double t = (ax * dx + ay * dy) + az * dz;
  double cx = dy * az - dz * ay, cy = dz * ax - dx * az, cz = dx * ay - dy * ax;
  double h = ofw_norm3(cx, cy, cz), na = ofw_norm3(ax, ay, az), nb = ofw_norm3(ax - dx, ay - dy, az - dz);
  if (!(isfinite(h) && isfinite(na) && isfinite(nb) && isfinite(t))) return PairOut{0.0, 0.0, 0.0, OFW_FLAG_NONFINITE};
  double big = na > nb ? na : nb;
  double scale = 1.0 > big ? 1.0 : big;
  if (h <= OFW_AXIS_EPS * scale) {
    if (raise_singular && t >= 0.0 && t <= 1.0) return PairOut{0.0, 0.0, 0.0, OFW_FLAG_SINGULAR};
    return PairOut{0.0, 0.0, 0.0, OFW_FLAG_OK};
  }
  if (gamma == 0.0) return PairOut{0.0, 0.0, 0.0, OFW_FLAG_OK};
  double product = na * nb;
  double dot = h * h + t * (t - 1.0);
  double plus = product + dot;
  if (dot < 0.0) plus = (h * h) / (product - dot);
  double denominator = product * plus;
  double rc = radius / length;
  double factor = 1.0;
  if (radius > 0.0 && model != 0) {
    if (model == OFW_MODEL_RM) {
      na = hypot(na, rc);
      nb = hypot(nb, rc);
      product = na * nb;
      double transverse2 = h * h + rc * rc;
      dot = transverse2 + t * (t - 1.0);
      plus = product + dot;
      if (dot < 0.0) plus = transverse2 / (product - dot);
      denominator = product * plus;
    } else if (model == OFW_MODEL_OFFSET) {
      denominator = denominator + rc * rc;
    } else {
      double ratio = h / rc;
      if (!isfinite(ratio)) return PairOut{0.0, 0.0, 0.0, OFW_FLAG_NONFINITE};
      if (model == OFW_MODEL_RANKINE) {
        double limited = ratio < 1.0 ? ratio : 1.0;
        factor = limited * limited;
      } else if (model == OFW_MODEL_LAMB) {
        if (ratio <= sqrt(10.0 / OFW_LAMB)) factor = -expm1(-OFW_LAMB * (ratio * ratio));
      } else if (model == OFW_MODEL_VATISTAS) {
        if (ratio < 1.0) {
          double q = ratio * ratio;
          factor = q / hypot(1.0, q);
        } else {
          double inverse = 1.0 / ratio;
          factor = 1.0 / hypot(1.0, inverse * inverse);
        }
      }
    }
  }
  double coefficient = (gamma / length) / OFW_FOUR_PI * factor * (na + nb) / denominator;
  if (!(isfinite(denominator) && isfinite(coefficient) && isfinite(product) && isfinite(dot)))
    return PairOut{0.0, 0.0, 0.0, OFW_FLAG_NONFINITE};
  return PairOut{coefficient * cx, coefficient * cy, coefficient * cz, OFW_FLAG_OK};
}

// The reference's axis guard with its own operations: h <= 8 eps max(1, max(|a|, |a - d|)) with hypot norms.
__device__ __noinline__ bool axis_reference(double ax, double ay, double az, double bx, double by, double bz,
                                           double cx, double cy, double cz) {
#pragma clang fp contract(off)
  const double h = ofw_norm3(cx, cy, cz), na = ofw_norm3(ax, ay, az), nb = ofw_norm3(bx, by, bz);
  const double big = na > nb ? na : nb;
  return h <= OFW_AXIS_EPS * (1.0 > big ? 1.0 : big);
}

// The reference's Lamb-Oseen cutoff test with its own operations: ratio = h / rc <= sqrt(10 / 1.25643).
__device__ __noinline__ bool lamb_inside_reference(double cx, double cy, double cz, double rc) {
#pragma clang fp contract(off)
  const double ratio = ofw_norm3(cx, cy, cz) / rc;
  return ratio <= sqrt(10.0 / OFW_LAMB);
}

// Reciprocal and square root for normal, positive FP64 operands: the hardware estimates (v_rcp_f64,
// v_rsq_f64) refined by Newton-Raphson steps to within an ulp or two, instead of the IEEE-exact division and
// square-root sequences with their range scaling.  Every operation is written out (no contraction), so the
// result does not depend on the kernel the function is inlined into.
__device__ __forceinline__ double ofw_rcp(double x) {
#pragma clang fp contract(off)
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double ofw_sqrt(double x) {
#pragma clang fp contract(off)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// One Newton-Raphson step from the hardware estimates (relative error 2^-24 or less): 2^-48 or less, a few
// ulps, for the pair's reciprocal and norms (within the direct sum's 1e-12 bound by three orders of magnitude,
// and far inside the 1e-12 bands of its decisions).
__device__ __forceinline__ double ofw_rcp1(double x) {
#pragma clang fp contract(off)
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double ofw_sqrt1(double x) {
#pragma clang fp contract(off)
  const double y = __builtin_amdgcn_rsq(x);
  const double g = x * y;
  const double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, 0.5 * y, g);
}

__device__ __forceinline__ double ofw_rsqrt(double x) {
#pragma clang fp contract(off)
  double y = __builtin_amdgcn_rsq(x);
  double xy = x * y;
  double e = __builtin_fma(-xy, y, 1.0);  // 1 - x y^2
  y = __builtin_fma(0.5 * y, e, y);
  xy = x * y;
  e = __builtin_fma(-xy, y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}

// Fast FP64 pair: the reference formula (vortex_numba.py:46-109) reorganised for the VALU: a per-segment
// reciprocal length and circulation factor, squared norms instead of hypot, the axis test and the core
// ratios on squares, one reciprocal per pair.  The same decisions and flags (the thresholds where a decision jumps
// are decided within their bands by the reference's operations), taken with selects so that a
// wavefront evaluating one segment for 64 targets does not diverge; only the rare paths (the Lamb-Oseen core
// inside its cutoff, extreme ranges) branch, on the wavefront's ballot.  A lane with use == false evaluates
// the pair but adds nothing (an excluded pair) and reports no flag.  a = (p - s)/L and the cross product
// d x a are rounded exactly as the reference rounds them (they decide h near a segment's line, where the
// cross product cancels); the other quantities agree with it to a few ulps (one Newton step from the hardware
// reciprocal and square-root estimates).  The arithmetic is written out operation by operation (no
// contraction), so every kernel that inlines it (direct sum, tree near field, pair matrix) gives the same bits
// for the same pair.  Pairs outside the fast range take pair_exact64.
template <int MODEL>
__device__ __forceinline__ int pair_fast64(double px, double py, double pz, const double* __restrict__ r,
                                           int raise_singular, bool use, double& ux, double& uy, double& uz) {
#pragma clang fp contract(off)
  const double L = r[R_L], invL = r[R_INVL];
  const double dx = r[R_DX], dy = r[R_DY], dz = r[R_DZ];
  // the correctly rounded quotient: the product with the correctly rounded reciprocal, corrected once by
  // its exact remainder (Markstein)
  const double qx = px - r[R_SX], qy = py - r[R_SY], qz = pz - r[R_SZ];
  double ax = qx * invL, ay = qy * invL, az = qz * invL;
  ax = __builtin_fma(__builtin_fma(-ax, L, qx), invL, ax);
  ay = __builtin_fma(__builtin_fma(-ay, L, qy), invL, ay);
  az = __builtin_fma(__builtin_fma(-az, L, qz), invL, az);
  const double cx = dy * az - dz * ay, cy = dz * ax - dx * az, cz = dx * ay - dy * ax;
  const double na2 = __builtin_fma(az, az, __builtin_fma(ay, ay, ax * ax));
  const double bx = ax - dx, by = ay - dy, bz = az - dz;
  const double nb2 = __builtin_fma(bz, bz, __builtin_fma(by, by, bx * bx));
  const double h2 = __builtin_fma(cz, cz, __builtin_fma(cy, cy, cx * cx));
  const double t = __builtin_fma(az, dz, __builtin_fma(ay, dy, ax * dx));
  const bool exact_seg = (MODEL == OFW_MODEL_RANKINE || MODEL == OFW_MODEL_LAMB || MODEL == OFW_MODEL_VATISTAS)
                             ? r[R_EXACT] != 0.0 : false;
  const bool slow = use && (!(fmax(na2, nb2) <= OFW_FAST_A2_MAX) || !(fmin(na2, nb2) >= OFW_FAST_A2_MIN) || exact_seg);
  const double na = ofw_sqrt1(na2), nb = ofw_sqrt1(nb2);
  const double mx = fmax(1.0, fmax(na, nb));
  // axis guard: certain outside the band (one multiply more than a plain test), exact inside it
  bool axis = h2 <= (OFW_AXIS_EPS * OFW_AXIS_EPS * OFW_BAND_HI) * (mx * mx);
  if (__builtin_amdgcn_ballot_w64(use && !slow && axis) != 0 && axis &&
      !(h2 < (OFW_AXIS_EPS * OFW_AXIS_EPS * OFW_BAND_LO) * (mx * mx)))
    axis = axis_reference(ax, ay, az, bx, by, bz, cx, cy, cz);
  const double g4 = r[R_G4];
  double sa = na, sb = nb;
  double product = na * nb;
  double dot = __builtin_fma(t, t - 1.0, h2);
  double factor = 1.0;
  double coef;
  if (MODEL == OFW_MODEL_OFFSET || MODEL == OFW_MODEL_RM) {
    double plus = dot < 0.0 ? h2 * ofw_rcp1(product - dot) : product + dot;
    double den = product * plus;
    const double rc2 = r[R_RC2];
    if (r[R_RADIUS] > 0.0) {
      if (MODEL == OFW_MODEL_OFFSET) {
        den = den + rc2;
      } else {
        sa = ofw_sqrt1(na2 + rc2);
        sb = ofw_sqrt1(nb2 + rc2);
        product = sa * sb;
        const double tr2 = h2 + rc2;
        dot = __builtin_fma(t, t - 1.0, tr2);
        plus = dot < 0.0 ? tr2 * ofw_rcp1(product - dot) : product + dot;
        den = product * plus;
      }
    }
    coef = ((g4 * factor) * (sa + sb)) * ofw_rcp1(den);
  } else {
    if (MODEL != OFW_MODEL_NONE) {
      const bool cored = r[R_RADIUS] > 0.0;
      const double q = h2 * r[R_IRC2];  // (h / rc)^2
      if (MODEL == OFW_MODEL_RANKINE) {
        factor = cored ? fmin(q, 1.0) : 1.0;
      } else if (MODEL == OFW_MODEL_LAMB) {
        // inside the cutoff only (a few pairs per target at most): a branch on the wavefront's ballot, where a
        // q within the band of the cutoff is decided by the reference's test
        bool inside = use && cored && q <= OFW_LAMB_Q * OFW_BAND_HI;
        if (__builtin_amdgcn_ballot_w64(inside) != 0 && inside) {
          if (!(q < OFW_LAMB_Q * OFW_BAND_LO)) inside = lamb_inside_reference(cx, cy, cz, r[R_RC]);
          if (inside) factor = -expm1(-OFW_LAMB * q);
        }
      } else {  // Vatistas: q / sqrt(1 + q^2) below one core radius, 1 / sqrt(1 + q^-2) beyond
        const double iq = q < 1.0 ? q : ofw_rcp1(q);
        const double rs = ofw_rsqrt(__builtin_fma(iq, iq, 1.0));
        factor = cored ? (q < 1.0 ? q * rs : rs) : 1.0;
      }
    }
    // den = product * plus with plus = h^2 / (product - dot) when dot < 0 (the reference's rationalised form):
    // one reciprocal of product * h^2 there, of product * (product + dot) elsewhere
    const bool rat = dot < 0.0;
    const double dd = rat ? product * h2 : product * (product + dot);
    const double xm = rat ? product - dot : 1.0;
    coef = (((g4 * factor) * (sa + sb)) * xm) * ofw_rcp1(dd);
  }
  bool add = use && !slow && !axis && g4 != 0.0;
  const bool finite = isfinite(coef);
  int flag = OFW_FLAG_OK;
  // flags are rare (a target on a segment's line, a non-finite coefficient): worked out on the ballot only
  if (__builtin_amdgcn_ballot_w64((use && !slow && axis) || (add && !finite)) != 0) {
    if (use && !slow && axis && raise_singular && t >= 0.0 && t <= 1.0) flag = OFW_FLAG_SINGULAR;
    if (add && !finite) flag = OFW_FLAG_NONFINITE;
  }
  add = add && finite;
  // the pair velocity is rounded before it is added, as in the reference's ordered sum; a pair not added adds
  // +0.0 (the sums start at +0.0, so this leaves them as they are)
  const double cf = add ? coef : 0.0;
  ux = ux + cf * cx;
  uy = uy + cf * cy;
  uz = uz + cf * cz;
  if (__builtin_amdgcn_ballot_w64(slow) != 0 && slow) {
    const PairOut e = pair_exact64(px, py, pz, r, MODEL, raise_singular);
    if (e.flag == OFW_FLAG_OK) {
      ux += e.x;
      uy += e.y;
      uz += e.z;
    }
    flag = e.flag;
  }
  return flag;
}

template <int MODEL>
__device__ __forceinline__ int pair_fast64(double px, double py, double pz, const double* __restrict__ r,
                                           int raise_singular, double& ux, double& uy, double& uz) {
  return pair_fast64<MODEL>(px, py, pz, r, raise_singular, true, ux, uy, uz);
}

// FP32 pair for the mixed / fp32 precisions.  rel = p - s rounded from FP64.  Returns true when the pair
// is near (must be evaluated in FP64, see OFW_NEAR_*) or outside the FP32 range; otherwise v is the FP32
// pair velocity (zero for a segment without circulation).  The arithmetic runs unconditionally, so that a
// wavefront evaluating one segment against 64 targets takes no divergent branch except the rare core factor
// of a target close to the line beyond an end of the segment.
struct Rec32 {
  float dx, dy, dz, invL, g4, rc2, exact, pad;
};

// The 32-byte FP32 record in registers (two 16-byte loads).
__device__ __forceinline__ Rec32 load_rec32(const float* __restrict__ f) {
  const float4 a = reinterpret_cast<const float4*>(f)[0], b = reinterpret_cast<const float4*>(f)[1];
  return Rec32{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// The FP32 pair (vortex_numba.py:42-109 in FP32, the rationalised denominator where r1.r2 < 0); it returns true
// when the pair must be evaluated in FP64 instead.  r = target - start (unscaled).  With XC the caller gives the
// cross product c = d x (r / L) formed in FP64 and rounded (cin): h is then exact to FP32 rounding however close
// the target is to the line, so core pairs stay in FP32 and only a thin band about the Lamb-Oseen cutoff takes
// FP64 (the decision there is the reference's).  Without it, c is formed in FP32 from rounded r / L and d, with
// an error of eps32 |r| / L in h, and the pairs within 4 core radii of the segment take FP64.
template <int MODEL, bool XC = false>
__device__ __forceinline__ bool pair_f32(float rx, float ry, float rz, const Rec32& f, float& vx, float& vy,
                                         float& vz, float cxi = 0.0f, float cyi = 0.0f, float czi = 0.0f) {
  const float invL = f.invL;
  const float dx = f.dx, dy = f.dy, dz = f.dz;
  const float ax = rx * invL, ay = ry * invL, az = rz * invL;
  const float cx = XC ? cxi : dy * az - dz * ay, cy = XC ? cyi : dz * ax - dx * az, cz = XC ? czi : dx * ay - dy * ax;
  const float na2 = ax * ax + ay * ay + az * az;
  const float bx = ax - dx, by = ay - dy, bz = az - dz;
  const float nb2 = bx * bx + by * by + bz * bz;
  const float h2 = cx * cx + cy * cy + cz * cz;
  const float m2 = fmaxf(1.0f, fmaxf(na2, nb2));
  const float rc2 = f.rc2;
  const float t = (ax * dx + ay * dy) + az * dz;
  // every test evaluated (bitwise combination, no short-circuit): no branch per pair
  const bool beside = (t >= 0.0f) & (t <= 1.0f);
  // FP64 for the pairs the reference's single-precision kernels mark suspect (within 1e-4 of the line,
  // vortex_cuda.py:110-124), beside the segment within 1e-3 of its line, or outside the FP32 range
  bool near = (!(m2 <= OFW_F32_A2_MAX)) | (h2 <= OFW_NEAR_AXIS2 * m2) | (f.exact != 0.0f) |
              (beside & (h2 <= OFW_NEAR_BESIDE2 * m2));
  if (MODEL != OFW_MODEL_NONE) {
    if (XC) {
      if (MODEL == OFW_MODEL_LAMB) {
        // the cutoff decision q <= 10/1.25643 within 1e-4 of it (q exact to a few FP32 roundings)
        const float qc = (float)OFW_LAMB_Q;
        near = near | ((rc2 > 0.0f) & (fabsf(h2 - qc * rc2) <= OFW_LAMB_BAND32 * qc * rc2));
      }
    } else {
      // squared distance to the segment itself (in segment lengths)
      const float dseg2 = beside ? h2 : (t < 0.0f ? na2 : nb2);
      near = near | (dseg2 < OFW_NEAR_CORES2 * rc2);
    }
  }
  const float g4 = f.g4;
  float sa = __builtin_amdgcn_sqrtf(na2), sb = __builtin_amdgcn_sqrtf(nb2);
  float product = sa * sb;
  float dot = h2 + t * (t - 1.0f);
  float factor = 1.0f;
  float coef;
  if (MODEL == OFW_MODEL_OFFSET || MODEL == OFW_MODEL_RM) {
    float plus = dot < 0.0f ? h2 * __builtin_amdgcn_rcpf(product - dot) : product + dot;
    float den = product * plus;
    if (MODEL == OFW_MODEL_OFFSET) {
      den = den + rc2;
    } else if (rc2 > 0.0f) {
      sa = __builtin_amdgcn_sqrtf(na2 + rc2);
      sb = __builtin_amdgcn_sqrtf(nb2 + rc2);
      product = sa * sb;
      const float tr2 = h2 + rc2;
      dot = tr2 + t * (t - 1.0f);
      plus = dot < 0.0f ? tr2 * __builtin_amdgcn_rcpf(product - dot) : product + dot;
      den = product * plus;
    }
    coef = g4 * (sa + sb) * __builtin_amdgcn_rcpf(den);
  } else {
    if (MODEL == OFW_MODEL_RANKINE || MODEL == OFW_MODEL_LAMB) {
      // core factors of q = (h / rc)^2 (vortex_numba.py:79-101) within 4 core radii of the line (beyond them
      // both factors are 1)
      const bool core = !near && rc2 > 0.0f && h2 < OFW_NEAR_CORES2 * rc2;
      if (__builtin_amdgcn_ballot_w64(core) != 0 && core) {
        const float q = h2 * __builtin_amdgcn_rcpf(rc2);
        if (MODEL == OFW_MODEL_RANKINE) factor = fminf(q, 1.0f);
        else if (q <= (float)(10.0 / OFW_LAMB)) factor = -expm1f(-(float)OFW_LAMB * q);
      }
    } else if (MODEL == OFW_MODEL_VATISTAS && rc2 > 0.0f) {
      // q / sqrt(1 + q^2) = 1 / sqrt(1 + q^-2) for every q > 0
      const float iq = rc2 * __builtin_amdgcn_rcpf(h2);
      factor = __builtin_amdgcn_rsqf(1.0f + iq * iq);
    }
    // den = product * plus, plus = h^2 / (product - dot) where dot < 0 (the rationalised form): one reciprocal,
    // of product * h^2 there and of product * (product + dot) elsewhere
    const bool rat = dot < 0.0f;
    const float dd = rat ? product * h2 : product * (product + dot);
    const float xm = rat ? product - dot : 1.0f;
    coef = g4 * factor * (sa + sb) * xm * __builtin_amdgcn_rcpf(dd);
  }
  vx = coef * cx;
  vy = coef * cy;
  vz = coef * cz;
  return near;
}

// FP32 record of the tree's near field (8 floats, 32 B): the length vector l, L^2, Gamma / (4 pi), Q (L rc)^2 with
// rc the core radius in metres and Q = 10/1.25643 the Lamb-Oseen cutoff of (h / rc)^2 (0 without a core), the FP64
// flag, and the half-width of the band about the cutoff that takes FP64 (-1 without a core).  The near field evaluates the pair in
// metres: u = Gamma / (4 pi) K(r1, r2) (l x r1), K = (|r1| + |r2|) / (|r1| |r2| (|r1| |r2| + r1.r2)), which is the
// formula of vortex_numba.py:46-109 in segment lengths (a = r1 / L, c = d x a = (l x r1) / L^2,
// g4 = Gamma / (4 pi L), K(a, b) = L^3 K(r1, r2)) with the powers of L cancelled; its thresholds carry the same
// powers of L (h^2 <= k max(1, |a|^2, |b|^2) is |l x r1|^2 <= k L^2 max(L^2, |r1|^2, |r2|^2)).
enum { N_LX = 0, N_LY, N_LZ, N_L2, N_GP, N_QLR2, N_EXACT, N_BAND, NEAR32 = 8 };
struct Near32 {
  float lx, ly, lz, l2, gp, qlr2, exact, band;
};
// Segments of length in [1e-6, 1e6] (every threshold product stays a normal float) with |Gamma| / (4 pi) and
// (L rc)^2 inside [1e-30, 1e30] take the FP32 evaluation; targets beyond 1e9 of a segment's ends take FP64.
#define OFW_NEAR_L2_MIN 1e-12
#define OFW_NEAR_L2_MAX 1e12
#define OFW_NEAR_R2_MAX 1e18f

__device__ __forceinline__ Near32 load_near32(const float* __restrict__ f) {
  const float4 a = reinterpret_cast<const float4*>(f)[0], b = reinterpret_cast<const float4*>(f)[1];
  return Near32{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// The tree near field's FP32 pair in metres: r1 = target - start, C = l x r1 (formed in FP64 by the caller and
// rounded, so h keeps its FP32 accuracy however close the target is to the line).  coef is the factor of C in the
// velocity; the return value is true when the pair must be evaluated in FP64 instead: within NEAR_AXIS of the
// line anywhere, within NEAR_BESIDE of it beside the segment, within LAMB_BAND32 of the Lamb-Oseen cutoff, or
// outside the FP32 range (the criteria of pair_f32 with the cross product given).  Models NONE, RANKINE and LAMB
// (the tree code's).
template <int MODEL>
__device__ __forceinline__ bool pair_near32(float rx, float ry, float rz, float cx, float cy, float cz,
                                            const Near32& f, float& coef) {
  const float r1 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
  const float bx = rx - f.lx, by = ry - f.ly, bz = rz - f.lz;
  const float r2 = __builtin_fmaf(bz, bz, __builtin_fmaf(by, by, bx * bx));
  const float h2 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
  const float t = __builtin_fmaf(rz, f.lz, __builtin_fmaf(ry, f.ly, rx * f.lx));  // r1 . l
  const float m2 = fmaxf(f.l2, fmaxf(r1, r2));
  const float lm = f.l2 * m2;
  // every test evaluated (bitwise combination, no short-circuit): no branch per pair
  const bool beside = (t >= 0.0f) & (t <= f.l2);
  bool near = (!(m2 <= OFW_NEAR_R2_MAX)) | (h2 <= OFW_NEAR_AXIS2 * lm) | (f.exact != 0.0f) |
              (beside & (h2 <= OFW_NEAR_BESIDE2 * lm));
  // the cutoff decision q = h2 / (L rc)^2 <= 10/1.25643 within 1e-4 of it (no core: band -1, never)
  if (MODEL == OFW_MODEL_LAMB) near = near | (fabsf(h2 - f.qlr2) <= f.band);
  float factor = 1.0f;
  if (MODEL == OFW_MODEL_RANKINE || MODEL == OFW_MODEL_LAMB) {
    // core factors of q = (h / rc)^2 (vortex_numba.py:79-101) within 4 core radii of the line (beyond them both
    // factors are 1; no core: qlr2 = 0, never)
    const float qc = (float)OFW_LAMB_Q;
    const bool core = !near && h2 < (float)(16.0 / OFW_LAMB_Q) * f.qlr2;
    if (__builtin_amdgcn_ballot_w64(core) != 0 && core) {
      const float q = qc * (h2 * __builtin_amdgcn_rcpf(f.qlr2));
      if (MODEL == OFW_MODEL_RANKINE) factor = fminf(q, 1.0f);
      else if (q <= (float)(10.0 / OFW_LAMB)) factor = -expm1f(-(float)OFW_LAMB * q);
    }
  }
  const float sa = __builtin_amdgcn_sqrtf(r1), sb = __builtin_amdgcn_sqrtf(r2);
  const float product = sa * sb;
  const float dot = r1 - t;  // r1 . r2
  // |r1| |r2| + r1.r2 = |l x r1|^2 / (|r1| |r2| - r1.r2) where r1.r2 < 0 (the rationalised form): one reciprocal
  const bool rat = dot < 0.0f;
  const float dd = rat ? product * h2 : product * (product + dot);
  const float xm = rat ? product - dot : 1.0f;
  coef = f.gp * factor * (sa + sb) * xm * __builtin_amdgcn_rcpf(dd);
  return near;
}

// Deterministic first-flag bookkeeping: a lane keeps the first flagged segment in its own order.
struct FlagState {
  int flag;
  int seg;
};

__device__ __forceinline__ void note_flag(FlagState& st, int f, int s) {
  if (f != OFW_FLAG_OK && st.flag == OFW_FLAG_OK) {
    st.flag = f;
    st.seg = s;
  }
}

#define OFW_CHECK_LAUNCH() (hipGetLastError())
