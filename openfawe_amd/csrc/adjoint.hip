// Unit-circulation pair matrix, ordered column assembly (the VSM influence operator) and the adjoint
// (vector-Jacobian product) of the direct sum.
//
// Layout: a wavefront owns 64 consecutive segments (one per lane) and walks the targets in order, so the
// target is wave-uniform (scalar loads) and the per-segment outputs are written coalesced.  The adjoint
// accumulates dV/dGamma, dV/dA and dV/dB of its segment over the targets in target order, and the fused
// forward velocities are reduced over the wavefront per target (fixed butterfly) and over wavefronts in
// segment order by a second kernel: deterministic.
#include <climits>
#include <vector>

#include "common.h"
#include "internal.h"
#include "tree_records.h"

namespace ofw {

// ---------------------------------------------------------------- pair matrix (vortex_numba.py:271-287)
template <int MODEL>
__global__ __launch_bounds__(256) void k_pair_matrix(int n, const double* __restrict__ pts, int m,
                                                     const double* __restrict__ rec, int raise_singular,
                                                     double* __restrict__ out, int32_t* __restrict__ code) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  const bool valid = s < m;
  const double px = pts[3 * (size_t)p], py = pts[3 * (size_t)p + 1], pz = pts[3 * (size_t)p + 2];
  double ux = 0.0, uy = 0.0, uz = 0.0;
  int f = OFW_FLAG_OK;
  if (valid) f = pair_fast64<MODEL>(px, py, pz, rec + (size_t)s * REC64, raise_singular, ux, uy, uz);
  if (valid) {
    double* o = out + 3 * ((size_t)p * m + s);
    o[0] = ux;
    o[1] = uy;
    o[2] = uz;
  }
  // first flagged segment of the target: an exact minimum, independent of the schedule
  if (valid && f != OFW_FLAG_OK) atomicMin(&code[p], s * 4 + f);
}

__global__ void k_code_init(int n, int32_t* __restrict__ code) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) code[p] = INT_MAX;
}

__global__ void k_code_to_flag(int n, const int32_t* __restrict__ code, int8_t* __restrict__ flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) flags[p] = code[p] == INT_MAX ? 0 : (int8_t)(code[p] & 3);
}

// ---------------------------------------------------------------- columns (vortex_numba.py:290-308)
__global__ void k_assemble_columns(int n, int m, const double* __restrict__ pairs, int columns,
                                   const int64_t* __restrict__ ptr, const int64_t* __restrict__ seg,
                                   const double* __restrict__ sign, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (j >= columns) return;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k) {
    const double* v = pairs + 3 * ((size_t)p * m + seg[k]);
    const double sg = sign[k];
    ux += sg * v[0];
    uy += sg * v[1];
    uz += sg * v[2];
  }
  double* o = out + 3 * ((size_t)p * columns + j);
  o[0] = ux;
  o[1] = uy;
  o[2] = uz;
}

// ---------------------------------------------------------------- adjoint of the direct sum
// u = Gamma/(4 pi) * K(r1, r2) * g(rho^2) * w, with r1 = x - A, r2 = x - B, w = r1 x r2, e = B - A = r1 - r2,
// K = (|r1| + |r2|) / (|r1||r2| D), D = |r1||r2| + r1.r2 (the reference's kernel in physical units,
// vortex_numba.py:65-105; D is the rationalised w^2 / (|r1||r2| - r1.r2) when r1.r2 < 0); rho^2 = |w|^2/(|e|^2 R^2)
// the squared core ratio.  The offset core replaces |r1||r2| D by |r1||r2| D + R^2 |e|^2.  The Rosenhead-Moore core
// (vortex_numba.py:74-83) is the same K with regularised norms n = sqrt(|r|^2 + R^2) and D = n1 n2 + r1.r2 + R^2
// (rationalised: (w^2 + R^2 |e|^2) / (n1 n2 - r1.r2 - R^2)), so its gradient is the plain one in those variables.
template <typename T>
struct PairGrad {
  T kx, ky, kz;      // unit-circulation velocity (dV/dGamma)
  T g1x, g1y, g1z;   // d(u . gu)/d r1
  T g2x, g2y, g2z;   // d(u . gu)/d r2
};

// Reciprocal and reciprocal square root: the refined hardware estimates (FP64, normal operands) or the
// hardware estimates (FP32); IEEE-exact operations when FAST is off.
template <typename T, bool FAST>
struct Arith;
template <>
struct Arith<double, true> {
  static __device__ __forceinline__ double rcp(double x) { return ofw_rcp(x); }
  static __device__ __forceinline__ double rsq(double x) { return ofw_rsqrt(x); }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double expm1_(double x) { return expm1(x); }
};
template <>
struct Arith<double, false> {
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double rsq(double x) { return 1.0 / __builtin_sqrt(x); }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double expm1_(double x) { return expm1(x); }
};
template <>
struct Arith<float, true> {
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
  static __device__ __forceinline__ float expm1_(float x) { return expm1f(x); }
};

// Per-segment constants of the adjoint, in the evaluation type.
template <typename T>
struct SegConst {
  T ex, ey, ez;   // B - A
  T L2, R2;       // |B - A|^2, R^2
  T iL2, iL2R2;   // 1 / L2, 1 / (L2 R2)
  T c;            // Gamma / (4 pi)
  bool cored;
};

template <int MODEL, typename T, bool FAST>
__device__ __forceinline__ bool pair_grad(T r1x, T r1y, T r1z, T r2x, T r2y, T r2z, const SegConst<T>& sc,
                                          T gux, T guy, T guz, PairGrad<T>& o) {
  typedef Arith<T, FAST> F;
  const T wx = r1y * r2z - r1z * r2y, wy = r1z * r2x - r1x * r2z, wz = r1x * r2y - r1y * r2x;
  T s1 = r1x * r1x + r1y * r1y + r1z * r1z;
  T s2 = r2x * r2x + r2y * r2y + r2z * r2z;
  const T w2 = wx * wx + wy * wy + wz * wz;
  const T L2 = sc.L2, R2 = sc.R2;
  const T mx = fmax(L2, fmax(s1, s2));
  // the axis guard of the reference (h <= 8 eps max(1, |a|, |b|) in units of the length)
  if (w2 <= (T)(OFW_AXIS_EPS * OFW_AXIS_EPS) * L2 * mx) return false;
  T dotp = r1x * r2x + r1y * r2y + r1z * r2z;
  const bool rm = MODEL == OFW_MODEL_RM && sc.cored;
  if (rm) {
    s1 += R2;
    s2 += R2;
    dotp += R2;
  }
  const T i1 = F::rsq(s1), i2 = F::rsq(s2);
  const T n1 = s1 * i1, n2 = s2 * i2;
  const T P = n1 * n2;
  const T is = F::rcp(n1 + n2);
  const T a1 = i1 * is, a2 = i2 * is;
  T K, l1x, l1y, l1z, l2x, l2y, l2z;
  if (MODEL == OFW_MODEL_OFFSET && sc.cored) {
    // Q = n1 n2 D + R^2 L2
    const T D = dotp < (T)0 ? w2 * F::rcp(P - dotp) : P + dotp;
    const T iQ = F::rcp(P * D + R2 * L2);
    K = (n1 + n2) * iQ;
    const T q1 = n2 * D * i1 + n2 * n2, q2 = n1 * D * i2 + n1 * n1;
    const T r2e = (T)2 * R2;
    l1x = a1 * r1x - (q1 * r1x + P * r2x + r2e * sc.ex) * iQ;
    l1y = a1 * r1y - (q1 * r1y + P * r2y + r2e * sc.ey) * iQ;
    l1z = a1 * r1z - (q1 * r1z + P * r2z + r2e * sc.ez) * iQ;
    l2x = a2 * r2x - (q2 * r2x + P * r1x - r2e * sc.ex) * iQ;
    l2y = a2 * r2y - (q2 * r2y + P * r1y - r2e * sc.ey) * iQ;
    l2z = a2 * r2z - (q2 * r2z + P * r1z - r2e * sc.ez) * iQ;
  } else {
    // 1 / D, from the rationalised form where r1.r2 < 0
    const T iD = dotp < (T)0 ? (P - dotp) * F::rcp(rm ? w2 + R2 * L2 : w2) : F::rcp(P + dotp);
    K = (n1 + n2) * (i1 * i2) * iD;
    const T e1 = a1 - i1 * i1 - (n2 * i1) * iD, e2 = a2 - i2 * i2 - (n1 * i2) * iD;
    l1x = e1 * r1x - r2x * iD;
    l1y = e1 * r1y - r2y * iD;
    l1z = e1 * r1z - r2z * iD;
    l2x = e2 * r2x - r1x * iD;
    l2y = e2 * r2y - r1y * iD;
    l2z = e2 * r2z - r1z * iD;
  }
  // core factor and its derivative in rho^2
  T g = 1, gp = 0, rho2 = 0;
  if (sc.cored && (MODEL == OFW_MODEL_RANKINE || MODEL == OFW_MODEL_LAMB || MODEL == OFW_MODEL_VATISTAS)) {
    rho2 = w2 * sc.iL2R2;
    if (MODEL == OFW_MODEL_RANKINE) {
      if (rho2 < (T)1) {
        g = rho2;
        gp = 1;
      }
    } else if (MODEL == OFW_MODEL_LAMB) {
      if (rho2 <= (T)OFW_LAMB_Q) {
        g = -F::expm1_((T)(-OFW_LAMB) * rho2);
        gp = (T)OFW_LAMB * F::exp_((T)(-OFW_LAMB) * rho2);
      }
    } else {
      const T is2 = F::rsq((T)1 + rho2 * rho2);   // 1 / sqrt(1 + rho^4)
      g = rho2 * is2;
      gp = is2 * is2 * is2;
    }
  }
  const T c4 = (T)(1.0 / OFW_FOUR_PI);
  const T Kg = K * g;
  o.kx = c4 * Kg * wx;
  o.ky = c4 * Kg * wy;
  o.kz = c4 * Kg * wz;
  const T cKg = sc.c * Kg;
  const T S = wx * gux + wy * guy + wz * guz;
  const T A = cKg * S;  // u . gu
  // d S / d r1 = r2 x gu, d S / d r2 = gu x r1
  o.g1x = A * l1x + cKg * (r2y * guz - r2z * guy);
  o.g1y = A * l1y + cKg * (r2z * gux - r2x * guz);
  o.g1z = A * l1z + cKg * (r2x * guy - r2y * gux);
  o.g2x = A * l2x + cKg * (guy * r1z - guz * r1y);
  o.g2y = A * l2y + cKg * (guz * r1x - gux * r1z);
  o.g2z = A * l2z + cKg * (gux * r1y - guy * r1x);
  if (gp != (T)0) {
    // d rho^2 / d r1 = 2 (r2 x w)/(L2 R^2) - 2 rho^2 e / L2; d rho^2 / d r2 = 2 (w x r1)/(L2 R^2) + 2 rho^2 e / L2
    const T B = sc.c * K * S * gp * (T)2;
    const T t1x = r2y * wz - r2z * wy, t1y = r2z * wx - r2x * wz, t1z = r2x * wy - r2y * wx;
    const T t2x = wy * r1z - wz * r1y, t2y = wz * r1x - wx * r1z, t2z = wx * r1y - wy * r1x;
    const T re = rho2 * sc.iL2;
    o.g1x += B * (t1x * sc.iL2R2 - re * sc.ex);
    o.g1y += B * (t1y * sc.iL2R2 - re * sc.ey);
    o.g1z += B * (t1z * sc.iL2R2 - re * sc.ez);
    o.g2x += B * (t2x * sc.iL2R2 + re * sc.ex);
    o.g2y += B * (t2y * sc.iL2R2 + re * sc.ey);
    o.g2z += B * (t2z * sc.iL2R2 + re * sc.ez);
  }
  return true;
}

// Mixed precision: a pair is evaluated in FP64 when the target is within 4 core radii of the segment, within
// 1e-4 (relative) of its line, beside it within 1e-3 of the line, or outside the FP32 range: the criteria of
// the forward mixed pair (OFW_NEAR_*), in physical units.
__device__ __forceinline__ bool vjp_near32(float r1x, float r1y, float r1z, float r2x, float r2y, float r2z,
                                           const SegConst<float>& sc) {
  const float wx = r1y * r2z - r1z * r2y, wy = r1z * r2x - r1x * r2z, wz = r1x * r2y - r1y * r2x;
  const float s1 = r1x * r1x + r1y * r1y + r1z * r1z;
  const float s2 = r2x * r2x + r2y * r2y + r2z * r2z;
  const float w2 = wx * wx + wy * wy + wz * wz;
  const float mx = fmaxf(sc.L2, fmaxf(s1, s2));
  const float te = r1x * sc.ex + r1y * sc.ey + r1z * sc.ez;  // t * L2
  const bool beside = te >= 0.0f && te <= sc.L2;
  const float lm = sc.L2 * mx;
  bool near = !(mx <= 1e30f * sc.L2) || !(lm <= 1e36f) || w2 <= OFW_NEAR_AXIS2 * lm ||
              (beside && w2 <= OFW_NEAR_BESIDE2 * lm);
  if (sc.cored) {
    const float dseg2 = beside ? w2 * sc.iL2 : (te < 0.0f ? s1 : s2);
    near = near || dseg2 < OFW_NEAR_CORES2 * sc.R2;
  }
  return near;
}

// Butterfly sum over the wavefront (every lane ends with the same bits).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Lanes own segments (64 per wavefront), the wavefront walks the targets [p0, p1) of its target chunk in order
// (blockIdx.y): each target and its dL/dV are wave-uniform.  Per segment the gradients over the chunk's targets
// are accumulated in target order (seg_part[chunk][s][7]); per (target, segment wavefront) the forward velocity
// and dL/dx are summed over the wavefront by a fixed butterfly (part_u / part_gx[wave][p - pb][3], pb the first
// target of the launch).
template <int MODEL, bool MIXED, bool FWD, bool GX>
__global__ __launch_bounds__(256) void k_direct_vjp(int pb, int pe, int tchunk, const double* __restrict__ pts, int m,
                                                    const double* __restrict__ starts,
                                                    const double* __restrict__ ends,
                                                    const double* __restrict__ gamma,
                                                    const double* __restrict__ radius,
                                                    const int64_t* __restrict__ ex_ptr,
                                                    const int64_t* __restrict__ ex_seg,
                                                    const double* __restrict__ grad_out,
                                                    double* __restrict__ seg_part, double* __restrict__ part_u,
                                                    double* __restrict__ part_gx) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = s >> 6;
  if (wave * 64 >= m) return;  // a whole wavefront beyond the segments (no partial slot of its own)
  const int p0 = pb + blockIdx.y * tchunk;
  const int p1 = min(pe, p0 + tchunk);
  const int nb = pe - pb;
  const bool valid = s < m;
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, gam = 0;
  SegConst<double> sc{};
  if (valid) {
    ax = starts[3 * (size_t)s];
    ay = starts[3 * (size_t)s + 1];
    az = starts[3 * (size_t)s + 2];
    bx = ends[3 * (size_t)s];
    by = ends[3 * (size_t)s + 1];
    bz = ends[3 * (size_t)s + 2];
    const double rad = radius[s];
    sc.ex = bx - ax;
    sc.ey = by - ay;
    sc.ez = bz - az;
    sc.L2 = sc.ex * sc.ex + sc.ey * sc.ey + sc.ez * sc.ez;
    sc.R2 = rad * rad;
    sc.cored = MODEL != OFW_MODEL_NONE && rad > 0.0;
    sc.iL2 = 1.0 / sc.L2;
    sc.iL2R2 = sc.cored ? 1.0 / (sc.L2 * sc.R2) : 0.0;
    gam = gamma[s];
    sc.c = gam / OFW_FOUR_PI;
  }
  // FP32 constants and the segment's own FP32-range test (a segment outside it takes FP64 for every pair)
  SegConst<float> sf{};
  bool seg64 = true;
  if (MIXED) {
    sf.ex = (float)sc.ex;
    sf.ey = (float)sc.ey;
    sf.ez = (float)sc.ez;
    sf.L2 = (float)sc.L2;
    sf.R2 = (float)sc.R2;
    sf.iL2 = (float)sc.iL2;
    sf.iL2R2 = (float)sc.iL2R2;
    sf.c = (float)sc.c;
    sf.cored = sc.cored;
    const double ac = fabs(sc.c);
    seg64 = !(sc.L2 >= 1e-30 && sc.L2 <= 1e30) || !(ac < 1e30 && (ac == 0.0 || ac > 1e-30)) ||
            (sc.cored && !(sc.iL2R2 >= 1e-30 && sc.iL2R2 <= 1e30));
  }
  double gg = 0.0, gax = 0, gay = 0, gaz = 0, gbx = 0, gby = 0, gbz = 0;
  for (int p = p0; p < p1; ++p) {
    const double px = pts[3 * (size_t)p], py = pts[3 * (size_t)p + 1], pz = pts[3 * (size_t)p + 2];
    const double gux = grad_out[3 * (size_t)p], guy = grad_out[3 * (size_t)p + 1], guz = grad_out[3 * (size_t)p + 2];
    bool use = valid;
    for (int64_t k = ex_ptr[p]; k < ex_ptr[p + 1]; ++k) use = use && ex_seg[k] != s;
    const double r1x = px - ax, r1y = py - ay, r1z = pz - az;
    const double r2x = px - bx, r2y = py - by, r2z = pz - bz;
    PairGrad<double> o{};
    bool ok = false, need64 = use;
    if (MIXED) {
      const float f1x = (float)r1x, f1y = (float)r1y, f1z = (float)r1z;
      const float f2x = (float)r2x, f2y = (float)r2y, f2z = (float)r2z;
      need64 = use && (seg64 || vjp_near32(f1x, f1y, f1z, f2x, f2y, f2z, sf));
      PairGrad<float> of{};
      const bool okf = use && !need64 &&
                       pair_grad<MODEL, float, true>(f1x, f1y, f1z, f2x, f2y, f2z, sf, (float)gux, (float)guy,
                                                     (float)guz, of);
      if (okf) {
        o = PairGrad<double>{of.kx, of.ky, of.kz, of.g1x, of.g1y, of.g1z, of.g2x, of.g2y, of.g2z};
        ok = true;
      }
    }
    if (!MIXED || __builtin_amdgcn_ballot_w64(need64) != 0) {
      // the refined reciprocals need normal operands: extreme ranges take the IEEE operations
      const double s1 = r1x * r1x + r1y * r1y + r1z * r1z, s2 = r2x * r2x + r2y * r2y + r2z * r2z;
      const bool extreme = need64 && !(fmin(fmin(s1, s2), sc.L2) >= 1e-280 && fmax(fmax(s1, s2), sc.L2) <= 1e280);
      if (need64 && !extreme)
        ok = pair_grad<MODEL, double, true>(r1x, r1y, r1z, r2x, r2y, r2z, sc, gux, guy, guz, o);
      if (__builtin_amdgcn_ballot_w64(extreme) != 0 && extreme)
        ok = pair_grad<MODEL, double, false>(r1x, r1y, r1z, r2x, r2y, r2z, sc, gux, guy, guz, o);
    }
    if (ok) {
      gg += o.kx * gux + o.ky * guy + o.kz * guz;
      gax -= o.g1x;
      gay -= o.g1y;
      gaz -= o.g1z;
      gbx -= o.g2x;
      gby -= o.g2y;
      gbz -= o.g2z;
    } else {
      o = PairGrad<double>{};
    }
    const size_t slot = 3 * ((size_t)wave * nb + (p - pb));
    if (FWD) {
      const double vx = wave_sum(gam * o.kx), vy = wave_sum(gam * o.ky), vz = wave_sum(gam * o.kz);
      if ((threadIdx.x & 63) == 0) {
        part_u[slot] = vx;
        part_u[slot + 1] = vy;
        part_u[slot + 2] = vz;
      }
    }
    if (GX) {
      const double vx = wave_sum(o.g1x + o.g2x), vy = wave_sum(o.g1y + o.g2y), vz = wave_sum(o.g1z + o.g2z);
      if ((threadIdx.x & 63) == 0) {
        part_gx[slot] = vx;
        part_gx[slot + 1] = vy;
        part_gx[slot + 2] = vz;
      }
    }
  }
  if (valid) {
    double* q = seg_part + 7 * ((size_t)blockIdx.y * m + s);
    q[0] = gg;
    q[1] = gax;
    q[2] = gay;
    q[3] = gaz;
    q[4] = gbx;
    q[5] = gby;
    q[6] = gbz;
  }
}

// Per segment: the target chunks' gradients added in chunk order, onto the previous target blocks' sum.
__global__ void k_vjp_segments(int m, int chunks, int accumulate, const double* __restrict__ seg_part,
                               double* __restrict__ grad_gamma, double* __restrict__ grad_a,
                               double* __restrict__ grad_b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  double v[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) v[q] = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* p = seg_part + 7 * ((size_t)c * m + s);
#pragma unroll
    for (int q = 0; q < 7; ++q) v[q] += p[q];
  }
  if (accumulate) {
    v[0] += grad_gamma[s];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[1 + q] += grad_a[3 * (size_t)s + q];
      v[4 + q] += grad_b[3 * (size_t)s + q];
    }
  }
  grad_gamma[s] = v[0];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    grad_a[3 * (size_t)s + q] = v[1 + q];
    grad_b[3 * (size_t)s + q] = v[4 + q];
  }
}

// Per target: the segment wavefronts' sums (part[wave][p][3]) added by LANES lanes per target, lane j taking the
// wavefronts j, j + LANES, ... in order and the lanes combined by a fixed butterfly (deterministic).
template <int LANES>
__global__ __launch_bounds__(256) void k_vjp_targets(int nb, int waves, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = g / LANES;
  const int j = g % LANES;
  double x = 0.0, y = 0.0, z = 0.0;
  if (p < nb) {
    for (int w = j; w < waves; w += LANES) {
      const double* v = part + 3 * ((size_t)w * nb + p);
      x += v[0];
      y += v[1];
      z += v[2];
    }
  }
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    y += __shfl_xor(y, o);
    z += __shfl_xor(z, o);
  }
  if (p >= nb || j != 0) return;
  out[3 * (size_t)p] = x;
  out[3 * (size_t)p + 1] = y;
  out[3 * (size_t)p + 2] = z;
}

// vortex.py:460-465: delta = B - A, L = hypot(hypot(dx, dy), dz), direction = delta / L
__global__ void k_segment_geometry(int m, const double* __restrict__ a, const double* __restrict__ b,
                                   double* __restrict__ dir, double* __restrict__ len, int32_t* __restrict__ zero) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const double dx = b[3 * (size_t)s] - a[3 * (size_t)s];
  const double dy = b[3 * (size_t)s + 1] - a[3 * (size_t)s + 1];
  const double dz = b[3 * (size_t)s + 2] - a[3 * (size_t)s + 2];
  const double L = hypot(hypot(dx, dy), dz);
  if (L == 0.0) atomicOr(zero, 1);
  len[s] = L;
  dir[3 * (size_t)s] = dx / L;
  dir[3 * (size_t)s + 1] = dy / L;
  dir[3 * (size_t)s + 2] = dz / L;
}

}  // namespace ofw

using ofw::set_error;

extern "C" {

int ofw_pair_matrix(ofw_ctx* x, int64_t n, const double* points, int64_t m, const double* starts,
                    const double* directions, const double* lengths, const double* radius, int model,
                    int raise_singular, double* out, int8_t* flags) {
  if (!x || n < 0 || m < 0 || model < 0 || model > 5 || n > 65535 || m > ((int64_t)1 << 28)) {
    set_error("ofw_pair_matrix: invalid arguments (at most 65,535 targets per call)");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0) return OFW_OK;
  // unit circulations: the pair records are packed with gamma = 1
  OFW_HIP(c.rec64.reserve((size_t)(m > 0 ? m : 1) * REC64 * sizeof(double)));
  OFW_HIP(c.rec32.reserve((size_t)(m > 0 ? m : 1) * REC32 * sizeof(float)));
  OFW_HIP(c.part.reserve((size_t)(m > 0 ? m : 1) * sizeof(double) + (size_t)n * sizeof(int32_t)));
  double* ones = (double*)c.part.ptr;
  int32_t* code = (int32_t*)((char*)c.part.ptr + (size_t)(m > 0 ? m : 1) * sizeof(double));
  if (m > 0) {
    std::vector<double> host((size_t)m, 1.0);
    OFW_HIP(hipMemcpyAsync(ones, host.data(), (size_t)m * 8, hipMemcpyHostToDevice, c.stream));
    int st = ofw::pack_segments(c, m, starts, directions, lengths, ones, radius, model, (double*)c.rec64.ptr,
                                (float*)c.rec32.ptr);
    if (st) return st;
    OFW_HIP(hipStreamSynchronize(c.stream));
  }
  hipLaunchKernelGGL(ofw::k_code_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, code);
  if (m > 0) {
    dim3 grid((unsigned)((m + 255) / 256), (unsigned)n);
    const double* rec = (const double*)c.rec64.ptr;
    switch (model) {
      case 0: hipLaunchKernelGGL(ofw::k_pair_matrix<0>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
      case 1: hipLaunchKernelGGL(ofw::k_pair_matrix<1>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
      case 2: hipLaunchKernelGGL(ofw::k_pair_matrix<2>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
      case 3: hipLaunchKernelGGL(ofw::k_pair_matrix<3>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
      case 4: hipLaunchKernelGGL(ofw::k_pair_matrix<4>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
      default: hipLaunchKernelGGL(ofw::k_pair_matrix<5>, grid, dim3(256), 0, c.stream, (int)n, points, (int)m, rec, raise_singular, out, code); break;
    }
  }
  hipLaunchKernelGGL(ofw::k_code_to_flag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, code,
                     flags);
  OFW_HIP(hipGetLastError());
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

int ofw_segment_geometry(ofw_ctx* x, int64_t m, const double* starts, const double* ends, double* directions,
                         double* lengths) {
  if (!x || m < 0 || m > ((int64_t)1 << 28)) {
    set_error("ofw_segment_geometry: invalid arguments");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (m == 0) return OFW_OK;
  OFW_HIP(c.pflag.reserve(sizeof(int32_t)));
  int32_t* zero = (int32_t*)c.pflag.ptr;
  OFW_HIP(hipMemsetAsync(zero, 0, sizeof(int32_t), c.stream));
  hipLaunchKernelGGL(ofw::k_segment_geometry, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m,
                     starts, ends, directions, lengths, zero);
  OFW_HIP(hipGetLastError());
  int32_t host = 0;
  OFW_HIP(hipMemcpyAsync(&host, zero, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  OFW_HIP(hipStreamSynchronize(c.stream));
  if (host) {
    set_error("zero-length segment");
    return OFW_ERR_ARG;
  }
  return OFW_OK;
}

int ofw_assemble_columns(ofw_ctx* x, int64_t n, int64_t m, const double* pairs, int64_t columns,
                         const int64_t* column_ptr, const int64_t* column_segment, const double* column_sign,
                         double* out) {
  if (!x || n < 0 || m < 0 || columns < 0 || n > 65535) {
    set_error("ofw_assemble_columns: invalid arguments (at most 65,535 targets per call)");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (n == 0 || columns == 0) return OFW_OK;
  hipLaunchKernelGGL(ofw::k_assemble_columns, dim3((unsigned)((columns + 127) / 128), (unsigned)n), dim3(128), 0,
                     c.stream, (int)n, (int)m, pairs, (int)columns, column_ptr, column_segment, column_sign, out);
  OFW_HIP(hipGetLastError());
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

int ofw_direct_vjp(ofw_ctx* x, int precision, int64_t n, const double* points, int64_t m, const double* starts,
                   const double* ends, const double* gamma, const double* radius, int model, const int64_t* ex_ptr,
                   const int64_t* ex_seg, const double* grad_out, double* grad_gamma, double* grad_starts,
                   double* grad_ends, double* grad_points, double* out) {
  if (!x || n < 0 || m < 0 || model < 0 || model > 5 || n > ((int64_t)1 << 30) || m > ((int64_t)1 << 28) ||
      !(precision == OFW_FP64 || precision == OFW_MIXED)) {
    set_error("ofw_direct_vjp: invalid arguments (precision OFW_FP64 or OFW_MIXED, model 0-5)");
    return OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  OFW_HIP(c.mark());
  if (m == 0) {
    if (out && n) OFW_HIP(hipMemsetAsync(out, 0, (size_t)n * 24, c.stream));
    if (grad_points && n) OFW_HIP(hipMemsetAsync(grad_points, 0, (size_t)n * 24, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    return OFW_OK;
  }
  const bool fwd = out != nullptr, gx = grad_points != nullptr;
  const int64_t segwaves = (m + 63) / 64;
  const int64_t want = (int64_t)c.cu_count * 4 * 8;  // wavefronts: several per SIMD
  // per-target partials (segwaves x targets of a block x 24 B each) stay within 256 MiB: targets go in blocks
  const int64_t per_target = segwaves * 24 * ((fwd ? 1 : 0) + (gx ? 1 : 0));
  int64_t block_n = per_target > 0 ? ((int64_t)256 << 20) / per_target : n;
  if (block_n < 1) block_n = 1;
  if (block_n > n) block_n = n;
  int64_t max_chunks = 1;
  {
    int64_t tc = (want + segwaves - 1) / segwaves;
    max_chunks = tc < 1 ? 1 : tc;
  }
  OFW_HIP(c.part.reserve((size_t)per_target * (size_t)(block_n > 0 ? block_n : 1) + 256));
  OFW_HIP(c.vjp.reserve((size_t)max_chunks * (size_t)m * 7 * sizeof(double)));
  double* part_u = (double*)c.part.ptr;
  double* part_gx = part_u + (fwd ? 3 * segwaves * block_n : 0);
  double* seg_part = (double*)c.vjp.ptr;
  const int lanes = segwaves <= 8 ? 1 : segwaves <= 32 ? 4 : segwaves <= 128 ? 16 : 64;
  OFW_HIP(c.mark());
  for (int64_t pb = 0; pb < n || (n == 0 && pb == 0); pb += block_n > 0 ? block_n : 1) {
    const int64_t pe = n == 0 ? 0 : (pb + block_n < n ? pb + block_n : n);
    const int64_t nbk = pe - pb;
    int64_t chunks = max_chunks;
    const int64_t most = nbk / 4 > 0 ? nbk / 4 : 1;
    if (chunks > most) chunks = most;
    int64_t tchunk = nbk > 0 ? (nbk + chunks - 1) / chunks : 1;
    chunks = nbk > 0 ? (nbk + tchunk - 1) / tchunk : 0;
    if (chunks > 0) {
      dim3 grid((unsigned)((m + 255) / 256), (unsigned)chunks);
#define OFW_VJP(MODEL, MIX)                                                                                       \
  do {                                                                                                            \
    if (fwd && gx)                                                                                                \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, MIX, true, true>), grid, dim3(256), 0, c.stream, (int)pb,      \
                         (int)pe, (int)tchunk, points, (int)m, starts, ends, gamma, radius, ex_ptr, ex_seg, grad_out, \
                         seg_part, part_u, part_gx);                                                              \
    else if (fwd)                                                                                                 \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, MIX, true, false>), grid, dim3(256), 0, c.stream, (int)pb,     \
                         (int)pe, (int)tchunk, points, (int)m, starts, ends, gamma, radius, ex_ptr, ex_seg, grad_out, \
                         seg_part, part_u, part_gx);                                                              \
    else if (gx)                                                                                                  \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, MIX, false, true>), grid, dim3(256), 0, c.stream, (int)pb,     \
                         (int)pe, (int)tchunk, points, (int)m, starts, ends, gamma, radius, ex_ptr, ex_seg, grad_out, \
                         seg_part, part_u, part_gx);                                                              \
    else                                                                                                          \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, MIX, false, false>), grid, dim3(256), 0, c.stream, (int)pb,    \
                         (int)pe, (int)tchunk, points, (int)m, starts, ends, gamma, radius, ex_ptr, ex_seg, grad_out, \
                         seg_part, part_u, part_gx);                                                              \
  } while (0)
#define OFW_VJP_P(MODEL)                        \
  do {                                          \
    if (precision == OFW_MIXED) OFW_VJP(MODEL, true); \
    else OFW_VJP(MODEL, false);                 \
  } while (0)
      switch (model) {
        case 0: OFW_VJP_P(0); break;
        case 1: OFW_VJP_P(1); break;
        case 2: OFW_VJP_P(2); break;
        case 3: OFW_VJP_P(3); break;
        case 4: OFW_VJP_P(4); break;
        default: OFW_VJP_P(5); break;
      }
#undef OFW_VJP_P
#undef OFW_VJP
      OFW_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(ofw::k_vjp_segments, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c.stream, (int)m,
                       (int)chunks, pb > 0 ? 1 : 0, seg_part, grad_gamma, grad_starts, grad_ends);
    OFW_HIP(hipGetLastError());
    for (int q = 0; q < 2 && nbk > 0; ++q) {
      double* dst = q == 0 ? out : grad_points;
      const double* src = q == 0 ? part_u : part_gx;
      if (!dst) continue;
      const dim3 rg((unsigned)((nbk * lanes + 255) / 256));
      switch (lanes) {
        case 1: hipLaunchKernelGGL(ofw::k_vjp_targets<1>, rg, dim3(256), 0, c.stream, (int)nbk, (int)segwaves, src, dst + 3 * pb); break;
        case 4: hipLaunchKernelGGL(ofw::k_vjp_targets<4>, rg, dim3(256), 0, c.stream, (int)nbk, (int)segwaves, src, dst + 3 * pb); break;
        case 16: hipLaunchKernelGGL(ofw::k_vjp_targets<16>, rg, dim3(256), 0, c.stream, (int)nbk, (int)segwaves, src, dst + 3 * pb); break;
        default: hipLaunchKernelGGL(ofw::k_vjp_targets<64>, rg, dim3(256), 0, c.stream, (int)nbk, (int)segwaves, src, dst + 3 * pb); break;
      }
      OFW_HIP(hipGetLastError());
    }
    if (n == 0) break;
  }
  OFW_HIP(c.mark());
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

}  // extern "C"
