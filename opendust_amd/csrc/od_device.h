// od_device.h -- device-side building blocks of the gfx950 ion-flow engine:
// counter-based RNG, the ion-ion pair law in FP32 and FP64, the external
// (dust, trap, field) forces and the reinjection tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "opendust_amd.h"

namespace od {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter = {ion, step, sim, stream | sub << 8}, key = seed.
// ---------------------------------------------------------------------------
struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ void philox_round(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k0;
  const uint32_t n2 = hi0 ^ c[3] ^ k1;
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

__host__ __device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                        uint64_t seed) {
  uint32_t c[4] = {c0, c1, c2, c3};
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  u32x4 o;
  o.v[0] = c[0];
  o.v[1] = c[1];
  o.v[2] = c[2];
  o.v[3] = c[3];
  return o;
}

// stream ids of the per-ion draws (low byte of counter word 3)
enum : uint32_t {
  S_FACE = 0,
  S_TOPV = 1,
  S_TOPX = 2,
  S_BOTV = 3,
  S_BOTX = 4,
  S_SIDEX = 5,
  S_SIDEV = 6,
  S_COLL = 7,
  S_TRAP = 16,
};

// 53-bit uniform in [0, 1) from two words (res53 construction).
__host__ __device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)k * (1.0 / 9007199254740992.0);
}

// Box-Muller pair from four words: u1 in (0, 1], u2 in [0, 1).
__host__ __device__ __forceinline__ void box_muller(const u32x4& r, double& g1, double& g2) {
  const double u1 = u53(r.v[0], r.v[1]) + (1.0 / 9007199254740992.0);
  const double u2 = u53(r.v[2], r.v[3]);
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586 * u2;
  g1 = rad * cos(ang);
  g2 = rad * sin(ang);
}

// ---------------------------------------------------------------------------
// Reinjection inverse CDF, evaluated exactly as the reference expression
// (sum over all segments of the masked linear pieces; opendust.py:702-723).
// ---------------------------------------------------------------------------
__host__ __device__ inline double table_eval(const od_table& t, double u) {
  double s = 0.0;
  for (int k = 0; k < t.n_seg; ++k) {
    const double sel = (t.half[k] - fabs(u - t.center[k]) >= 0.0) ? 1.0 : 0.0;
    s = s + (t.v0[k] + (u - t.xi0[k]) * t.slope[k]) * sel;
  }
  return s;
}

// ---------------------------------------------------------------------------
// Ion-ion pair law.  U(r) = kqq exp(-r/lambda)/(r + r_0) (opendust.py:472-479).
// With rho = r/lambda, rho0 = r_0/lambda, s = 1/(rho + rho0):
//   F_i = (kqq/lambda^2) sum_j exp(-rho) s (1 + s) / rho * (x_i - x_j)/lambda.
//
// FP32 form, coordinates pre-scaled by L/lambda (L = log2 e) so the exponent is
// a bare v_exp_f32: with d = u_j - u_i, r2 = |d|^2 + eps, r = sqrt(r2),
//   w = 1/(r (r + a)),  a = rho0 L,   t = r w + 1/L,   h = 2^-r w t
//   F_i = -(kqq/lambda^2) L^2 sum_j h d.
// 13 full-rate VALU ops + 3 transcendental (sqrt, exp2, rcp) per pair.  The
// eps term makes the self pair (d = 0) contribute exactly 0 instead of NaN, and
// ions parked at PARK32 contribute exactly 0 because 2^-r underflows to 0.
// ---------------------------------------------------------------------------
constexpr float L2E_F = 1.4426950408889634f;
constexpr float EPS32 = 1e-30f;
constexpr float PARK32 = 1e15f;
constexpr double EPS64 = 1e-300;
constexpr double PARK64 = 1e15;

__device__ __forceinline__ void pair32(float xi, float yi, float zi, float xj, float yj, float zj, float a,
                                       float linv, float& ax, float& ay, float& az) {
  const float dx = xj - xi;
  const float dy = yj - yi;
  const float dz = zj - zi;
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, EPS32)));
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float e = __builtin_amdgcn_exp2f(-r);
  const float w = __builtin_amdgcn_rcpf(__builtin_fmaf(a, r, r2));
  const float t = __builtin_fmaf(r, w, linv);
  const float h = e * w * t;
  ax = __builtin_fmaf(h, dx, ax);
  ay = __builtin_fmaf(h, dy, ay);
  az = __builtin_fmaf(h, dz, az);
}

// FP64 form in lambda units: h = exp(-r) w (1 + r w), w = 1/(r (r + rho0)),
// F_i = -(kqq/lambda^2) sum_j h d.
__device__ __forceinline__ void pair64(double xi, double yi, double zi, double xj, double yj, double zj,
                                       double rho0, double& ax, double& ay, double& az) {
  const double dx = xj - xi;
  const double dy = yj - yi;
  const double dz = zj - zi;
  const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, EPS64)));
  const double r = sqrt(r2);
  const double e = exp(-r);
  const double w = 1.0 / (r * (r + rho0));
  const double t = fma(r, w, 1.0);
  const double h = e * w * t;
  ax = fma(h, dx, ax);
  ay = fma(h, dy, ay);
  az = fma(h, dz, az);
}

// Truncated / switched forms.  Inside the cutoff the energy is U(r) S(x) with
// S = 1 - 10x^3 + 15x^4 - 6x^5, x = (r - r_s)/(r_c - r_s) clamped to [0, 1]; the
// force is -(U' S + U S').  In the FP32 units above, with e = 2^-r and w as in
// pair32, U S'/r contributes -e w dS/dr (dS/dr in the same units), so
// h -> h S - e w S'.  Beyond the cutoff h = 0 (no branch: a select).
template <bool SW>
__device__ __forceinline__ void pair32_cut(float xi, float yi, float zi, float xj, float yj, float zj, float a,
                                           float linv, float rc, float rs, float invw, float& ax, float& ay,
                                           float& az) {
  const float dx = xj - xi;
  const float dy = yj - yi;
  const float dz = zj - zi;
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, EPS32)));
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float e = __builtin_amdgcn_exp2f(-r);
  const float w = __builtin_amdgcn_rcpf(__builtin_fmaf(a, r, r2));
  const float t = __builtin_fmaf(r, w, linv);
  float h = e * w * t;
  if constexpr (SW) {
    const float x = __builtin_fminf(__builtin_fmaxf((r - rs) * invw, 0.0f), 1.0f);
    const float x2 = x * x;
    const float S = __builtin_fmaf(x2 * x, __builtin_fmaf(x, __builtin_fmaf(x, -6.0f, 15.0f), -10.0f), 1.0f);
    const float omx = 1.0f - x;
    const float dS = -30.0f * x2 * omx * omx * invw;
    h = __builtin_fmaf(h, S, -(e * w * dS));
  }
  h = (r < rc) ? h : 0.0f;
  ax = __builtin_fmaf(h, dx, ax);
  ay = __builtin_fmaf(h, dy, ay);
  az = __builtin_fmaf(h, dz, az);
}

__device__ __forceinline__ void pair64_cut(double xi, double yi, double zi, double xj, double yj, double zj,
                                           double rho0, bool sw, double rc, double rs, double invw, double& ax,
                                           double& ay, double& az) {
  const double dx = xj - xi;
  const double dy = yj - yi;
  const double dz = zj - zi;
  const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, EPS64)));
  const double r = sqrt(r2);
  if (!(r < rc)) return;
  const double e = exp(-r);
  const double w = 1.0 / (r * (r + rho0));
  double h = e * w * fma(r, w, 1.0);
  if (sw && r > rs) {
    const double x = (r - rs) * invw;
    const double S = 1.0 - 10.0 * x * x * x + 15.0 * x * x * x * x - 6.0 * x * x * x * x * x;
    const double dS = (-30.0 * x * x + 60.0 * x * x * x - 30.0 * x * x * x * x) * invw;
    h = h * S - e * w * dS;
  }
  ax = fma(h, dx, ax);
  ay = fma(h, dy, ay);
  az = fma(h, dz, az);
}

// Kahan step: add a block partial p into the compensated sum (s, c).
__device__ __forceinline__ void kahan_add(float& s, float& c, float p) {
  const float y = p - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// ---------------------------------------------------------------------------
// External per-ion forces in OpenMM units (kJ/mol/nm), FP64:
//   dust -> ion Coulomb  F = q_p k_dust (x - x_p)/|x - x_p|^3   (opendust.py:349-360)
//   trap                 F = -grad(C sum alpha_m rho^a h^b)     (opendust.py:550-577)
//   field                F_z = E q_s conv                        (opendust.py:579-585)
// The trap gradient is written with G = (1/rho) dP/drho and K = dP/dh, both
// polynomials, so it is regular on the axis.
// ---------------------------------------------------------------------------
struct ExtParams {
  int n_dust;
  int field_driven;
  double inv_lam;
  double trap_scale;  // -trap_C / lambda
  double alpha[15];
  double k_dust;
  double field_force;
  double dx[OD_MAX_DUST], dy[OD_MAX_DUST], dz[OD_MAX_DUST], dr[OD_MAX_DUST];
};

__device__ __forceinline__ void ext_force(const ExtParams& P, const double* __restrict__ q, double x, double y,
                                          double z, double& fx, double& fy, double& fz) {
  fx = 0.0;
  fy = 0.0;
  fz = 0.0;
  for (int p = 0; p < P.n_dust; ++p) {
    const double ddx = x - P.dx[p], ddy = y - P.dy[p], ddz = z - P.dz[p];
    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
    const double r = sqrt(r2);
    const double c = q[p] * P.k_dust / (r2 * r);
    fx += c * ddx;
    fy += c * ddy;
    fz += c * ddz;
  }
  const double* a = P.alpha;
  const double X = x * P.inv_lam, Y = y * P.inv_lam, h = z * P.inv_lam;
  const double p2 = X * X + Y * Y;
  const double p4 = p2 * p2, p6 = p4 * p2;
  const double h2 = h * h, h4 = h2 * h2, h6 = h4 * h2;
  const double G = 2.0 * a[1] + 2.0 * a[3] * h2 + 4.0 * a[4] * p2 + 4.0 * a[6] * p2 * h2 + 2.0 * a[7] * h4 +
                   6.0 * a[8] * p4 + 4.0 * a[10] * p2 * h4 + 6.0 * a[11] * p4 * h2 + 2.0 * a[12] * h6 +
                   8.0 * a[13] * p6;
  const double K = h * (2.0 * a[2] + 2.0 * a[3] * p2 + 4.0 * a[5] * h2 + 2.0 * a[6] * p4 + 4.0 * a[7] * p2 * h2 +
                        6.0 * a[9] * h4 + 4.0 * a[10] * p4 * h2 + 2.0 * a[11] * p6 + 6.0 * a[12] * p2 * h4 +
                        8.0 * a[14] * h6);
  fx += P.trap_scale * G * X;
  fy += P.trap_scale * G * Y;
  fz += P.trap_scale * K;
  if (P.field_driven) fz += P.field_force;
}

}  // namespace od
