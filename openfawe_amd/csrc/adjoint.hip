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
// u = Gamma/(4 pi) * K(r1, r2) * g(rho^2) * w, with r1 = x - A, r2 = x - B, w = r1 x r2,
// K = (|r1| + |r2|) / (|r1||r2| D), D = |r1||r2| + r1.r2 (the reference's kernel in physical units,
// vortex_numba.py:65-105); rho^2 = |w|^2 / (|B-A|^2 R^2) the squared core ratio.  The offset core
// replaces |r1||r2| D by |r1||r2| D + R^2 |B-A|^2.
struct PairGrad {
  double kx, ky, kz;      // unit-circulation velocity (dV/dGamma)
  double g1x, g1y, g1z;   // d(u . gu)/d r1
  double g2x, g2y, g2z;   // d(u . gu)/d r2
};

template <int MODEL>
__device__ __forceinline__ bool pair_grad(double r1x, double r1y, double r1z, double r2x, double r2y, double r2z,
                                          double radius, double gamma, double gux, double guy, double guz,
                                          PairGrad& o) {
  const double wx = r1y * r2z - r1z * r2y, wy = r1z * r2x - r1x * r2z, wz = r1x * r2y - r1y * r2x;
  const double ex = r1x - r2x, ey = r1y - r2y, ez = r1z - r2z;  // B - A
  const double L2 = ex * ex + ey * ey + ez * ez;
  const double n1 = __builtin_sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
  const double n2 = __builtin_sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
  const double w2 = wx * wx + wy * wy + wz * wz;
  const double mx = fmax(L2, fmax(n1 * n1, n2 * n2));
  // the axis guard of the reference (h <= 8 eps max(1, |a|, |b|) in units of the length)
  if (w2 <= (OFW_AXIS_EPS * OFW_AXIS_EPS) * L2 * mx) return false;
  const double dotp = r1x * r2x + r1y * r2y + r1z * r2z;
  const double P = n1 * n2;
  const double D = dotp < 0.0 ? w2 / (P - dotp) : P + dotp;
  double Q = P * D;
  const double R2 = radius * radius;
  const bool cored = MODEL != OFW_MODEL_NONE && radius > 0.0;
  if (MODEL == OFW_MODEL_OFFSET && cored) Q += R2 * L2;
  const double K = (n1 + n2) / Q;
  // core factor and its derivative in rho^2
  double g = 1.0, gp = 0.0, rho2 = 0.0;
  if (cored && (MODEL == OFW_MODEL_RANKINE || MODEL == OFW_MODEL_LAMB || MODEL == OFW_MODEL_VATISTAS)) {
    rho2 = w2 / (L2 * R2);
    if (MODEL == OFW_MODEL_RANKINE) {
      if (rho2 < 1.0) {
        g = rho2;
        gp = 1.0;
      }
    } else if (MODEL == OFW_MODEL_LAMB) {
      if (rho2 <= 10.0 / OFW_LAMB) {
        const double ex2 = exp(-OFW_LAMB * rho2);
        g = -expm1(-OFW_LAMB * rho2);
        gp = OFW_LAMB * ex2;
      }
    } else {
      const double q2 = rho2 * rho2;
      const double s = __builtin_sqrt(1.0 + q2);
      g = rho2 / s;
      gp = 1.0 / (s * (1.0 + q2));
    }
  }
  const double c4 = 1.0 / OFW_FOUR_PI;
  o.kx = c4 * K * g * wx;
  o.ky = c4 * K * g * wy;
  o.kz = c4 * K * g * wz;
  const double c = gamma * c4;
  const double S = wx * gux + wy * guy + wz * guz;
  // gradient of ln K
  double l1x, l1y, l1z, l2x, l2y, l2z;
  {
    const double a1 = 1.0 / (n1 * (n1 + n2)), a2 = 1.0 / (n2 * (n1 + n2));
    if (MODEL == OFW_MODEL_OFFSET && cored) {
      // Q = n1 n2 D + R^2 L^2
      const double q1 = n2 * D / n1 + n2 * n2, q2 = n1 * D / n2 + n1 * n1;
      const double iQ = 1.0 / Q;
      l1x = a1 * r1x - (q1 * r1x + P * r2x + 2.0 * R2 * ex) * iQ;
      l1y = a1 * r1y - (q1 * r1y + P * r2y + 2.0 * R2 * ey) * iQ;
      l1z = a1 * r1z - (q1 * r1z + P * r2z + 2.0 * R2 * ez) * iQ;
      l2x = a2 * r2x - (q2 * r2x + P * r1x - 2.0 * R2 * ex) * iQ;
      l2y = a2 * r2y - (q2 * r2y + P * r1y - 2.0 * R2 * ey) * iQ;
      l2z = a2 * r2z - (q2 * r2z + P * r1z - 2.0 * R2 * ez) * iQ;
    } else {
      const double b1 = 1.0 / (n1 * n1), b2 = 1.0 / (n2 * n2), iD = 1.0 / D;
      const double e1 = a1 - b1 - (n2 / n1) * iD, e2 = a2 - b2 - (n1 / n2) * iD;
      l1x = e1 * r1x - r2x * iD;
      l1y = e1 * r1y - r2y * iD;
      l1z = e1 * r1z - r2z * iD;
      l2x = e2 * r2x - r1x * iD;
      l2y = e2 * r2y - r1y * iD;
      l2z = e2 * r2z - r1z * iD;
    }
  }
  // d S / d r1 = r2 x gu, d S / d r2 = gu x r1
  const double s1x = r2y * guz - r2z * guy, s1y = r2z * gux - r2x * guz, s1z = r2x * guy - r2y * gux;
  const double s2x = guy * r1z - guz * r1y, s2y = guz * r1x - gux * r1z, s2z = gux * r1y - guy * r1x;
  const double A = c * K * g * S;  // u . gu
  o.g1x = A * l1x + c * K * g * s1x;
  o.g1y = A * l1y + c * K * g * s1y;
  o.g1z = A * l1z + c * K * g * s1z;
  o.g2x = A * l2x + c * K * g * s2x;
  o.g2y = A * l2y + c * K * g * s2y;
  o.g2z = A * l2z + c * K * g * s2z;
  if (gp != 0.0) {
    // d rho^2 / d r1 = 2 (r2 x w)/(L^2 R^2) - 2 rho^2 (r1 - r2)/L^2; d/d r2 = 2 (w x r1)/(L^2 R^2) + 2 rho^2 (r1 - r2)/L^2
    const double iL2R2 = 1.0 / (L2 * R2), iL2 = 1.0 / L2;
    const double t1x = r2y * wz - r2z * wy, t1y = r2z * wx - r2x * wz, t1z = r2x * wy - r2y * wx;
    const double t2x = wy * r1z - wz * r1y, t2y = wz * r1x - wx * r1z, t2z = wx * r1y - wy * r1x;
    const double B = c * K * S * gp * 2.0;
    o.g1x += B * (t1x * iL2R2 - rho2 * ex * iL2);
    o.g1y += B * (t1y * iL2R2 - rho2 * ey * iL2);
    o.g1z += B * (t1z * iL2R2 - rho2 * ez * iL2);
    o.g2x += B * (t2x * iL2R2 + rho2 * ex * iL2);
    o.g2y += B * (t2y * iL2R2 + rho2 * ey * iL2);
    o.g2z += B * (t2z * iL2R2 + rho2 * ez * iL2);
  }
  return true;
}

template <int MODEL, bool FWD, bool GX>
__global__ __launch_bounds__(256) void k_direct_vjp(int n, const double* __restrict__ pts, int m,
                                                    const double* __restrict__ starts,
                                                    const double* __restrict__ ends,
                                                    const double* __restrict__ gamma,
                                                    const double* __restrict__ radius,
                                                    const int32_t* __restrict__ ex_ptr,
                                                    const int32_t* __restrict__ ex_seg,
                                                    const double* __restrict__ grad_out,
                                                    double* __restrict__ grad_gamma, double* __restrict__ grad_a,
                                                    double* __restrict__ grad_b, double* __restrict__ part_u,
                                                    double* __restrict__ part_gx) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = s >> 6;
  const int lane = threadIdx.x & 63;
  if (wave * 64 >= m) return;  // a whole wavefront beyond the segments (no partial slot of its own)
  const bool valid = s < m;
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, gam = 0, rad = 0;
  if (valid) {
    ax = starts[3 * (size_t)s];
    ay = starts[3 * (size_t)s + 1];
    az = starts[3 * (size_t)s + 2];
    bx = ends[3 * (size_t)s];
    by = ends[3 * (size_t)s + 1];
    bz = ends[3 * (size_t)s + 2];
    gam = gamma[s];
    rad = radius[s];
  }
  double gg = 0.0, gax = 0, gay = 0, gaz = 0, gbx = 0, gby = 0, gbz = 0;
  for (int p = 0; p < n; ++p) {
    const double px = pts[3 * (size_t)p], py = pts[3 * (size_t)p + 1], pz = pts[3 * (size_t)p + 2];
    const double gux = grad_out[3 * (size_t)p], guy = grad_out[3 * (size_t)p + 1], guz = grad_out[3 * (size_t)p + 2];
    bool use = valid;
    for (int k = ex_ptr[p]; k < ex_ptr[p + 1]; ++k) use = use && ex_seg[k] != s;
    PairGrad o{};
    bool ok = false;
    if (use) ok = pair_grad<MODEL>(px - ax, py - ay, pz - az, px - bx, py - by, pz - bz, rad, gam, gux, guy, guz, o);
    if (ok) {
      gg += o.kx * gux + o.ky * guy + o.kz * guz;
      gax -= o.g1x;
      gay -= o.g1y;
      gaz -= o.g1z;
      gbx -= o.g2x;
      gby -= o.g2y;
      gbz -= o.g2z;
    } else {
      o = PairGrad{};
    }
    if (FWD) {
      double v[3] = {gam * o.kx, gam * o.ky, gam * o.kz};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
      }
      if (lane == 0) {
        double* u = part_u + 3 * ((size_t)wave * n + p);
        u[0] = v[0];
        u[1] = v[1];
        u[2] = v[2];
      }
    }
    if (GX) {
      double v[3] = {o.g1x + o.g2x, o.g1y + o.g2y, o.g1z + o.g2z};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
      }
      if (lane == 0) {
        double* u = part_gx + 3 * ((size_t)wave * n + p);
        u[0] = v[0];
        u[1] = v[1];
        u[2] = v[2];
      }
    }
  }
  if (valid) {
    grad_gamma[s] = gg;
    grad_a[3 * (size_t)s] = gax;
    grad_a[3 * (size_t)s + 1] = gay;
    grad_a[3 * (size_t)s + 2] = gaz;
    grad_b[3 * (size_t)s] = gbx;
    grad_b[3 * (size_t)s + 1] = gby;
    grad_b[3 * (size_t)s + 2] = gbz;
  }
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

__global__ void k_wave_reduce(int n, int waves, const double* __restrict__ part, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double x = 0.0, y = 0.0, z = 0.0;
  for (int w = 0; w < waves; ++w) {
    const double* v = part + 3 * ((size_t)w * n + p);
    x += v[0];
    y += v[1];
    z += v[2];
  }
  out[3 * (size_t)p] = x;
  out[3 * (size_t)p + 1] = y;
  out[3 * (size_t)p + 2] = z;
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

int ofw_direct_vjp(ofw_ctx* x, int64_t n, const double* points, int64_t m, const double* starts, const double* ends,
                   const double* gamma, const double* radius, int model, const int64_t* ex_ptr,
                   const int64_t* ex_seg, const double* grad_out, double* grad_gamma, double* grad_starts,
                   double* grad_ends, double* grad_points, double* out) {
  if (!x || n < 0 || m < 0 || model < 0 || model > 4 || n > ((int64_t)1 << 30) || m > ((int64_t)1 << 28)) {
    set_error(model == 5 ? "ofw_direct_vjp: the Rosenhead-Moore core has no adjoint in this library"
                         : "ofw_direct_vjp: invalid arguments");
    return model == 5 ? OFW_ERR_UNSUPPORTED : OFW_ERR_ARG;
  }
  ofw::Ctx& c = x->c;
  c.nev = 0;
  if (m == 0) {
    if (out && n) OFW_HIP(hipMemsetAsync(out, 0, (size_t)n * 24, c.stream));
    if (grad_points && n) OFW_HIP(hipMemsetAsync(grad_points, 0, (size_t)n * 24, c.stream));
    OFW_HIP(hipStreamSynchronize(c.stream));
    return OFW_OK;
  }
  OFW_HIP(c.mark());
  int32_t *p32 = nullptr, *s32 = nullptr;
  int st = ofw::csr_to_i32(c, n, ex_ptr, ex_seg, &p32, &s32);
  if (st) return st;
  const int waves = (int)((m + 63) / 64);
  const size_t part_bytes = (size_t)waves * (size_t)(n > 0 ? n : 1) * 24;
  OFW_HIP(c.part.reserve(2 * part_bytes));
  double* part_u = (double*)c.part.ptr;
  double* part_gx = (double*)((char*)c.part.ptr + part_bytes);
  const bool fwd = out != nullptr, gx = grad_points != nullptr;
  dim3 grid((unsigned)((m + 255) / 256));
#define OFW_VJP(MODEL)                                                                                           \
  do {                                                                                                           \
    if (fwd && gx)                                                                                               \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, true, true>), grid, dim3(256), 0, c.stream, (int)n, points,  \
                         (int)m, starts, ends, gamma, radius, p32, s32, grad_out, grad_gamma, grad_starts,       \
                         grad_ends, part_u, part_gx);                                                            \
    else if (fwd)                                                                                                \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, true, false>), grid, dim3(256), 0, c.stream, (int)n, points, \
                         (int)m, starts, ends, gamma, radius, p32, s32, grad_out, grad_gamma, grad_starts,       \
                         grad_ends, part_u, part_gx);                                                            \
    else if (gx)                                                                                                 \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, false, true>), grid, dim3(256), 0, c.stream, (int)n, points, \
                         (int)m, starts, ends, gamma, radius, p32, s32, grad_out, grad_gamma, grad_starts,       \
                         grad_ends, part_u, part_gx);                                                            \
    else                                                                                                         \
      hipLaunchKernelGGL((ofw::k_direct_vjp<MODEL, false, false>), grid, dim3(256), 0, c.stream, (int)n,        \
                         points, (int)m, starts, ends, gamma, radius, p32, s32, grad_out, grad_gamma,            \
                         grad_starts, grad_ends, part_u, part_gx);                                               \
  } while (0)
  OFW_HIP(c.mark());
  switch (model) {
    case 0: OFW_VJP(0); break;
    case 1: OFW_VJP(1); break;
    case 2: OFW_VJP(2); break;
    case 3: OFW_VJP(3); break;
    default: OFW_VJP(4); break;
  }
#undef OFW_VJP
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  if (n > 0 && fwd)
    hipLaunchKernelGGL(ofw::k_wave_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, waves,
                       part_u, out);
  if (n > 0 && gx)
    hipLaunchKernelGGL(ofw::k_wave_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (int)n, waves,
                       part_gx, grad_points);
  OFW_HIP(hipGetLastError());
  OFW_HIP(c.mark());
  OFW_HIP(hipStreamSynchronize(c.stream));
  return OFW_OK;
}

}  // extern "C"
