// od_engine.hip -- gfx950 kernels and the C ABI of the ion-flow engine.
//
// One step replaces one integrator.step(1) of the reference's OpenMM program
// (opendust.py:839-1037) plus the host-side charge update and dust-force
// assembly (opendust.py:1117-1384):
//
//   k_kick_drift   collisions, first half-kick, drift, absorption/outflow flags
//   k_scan/k_list  ordered compaction of the flagged ions (ion-index order)
//   k_pair32/64    ion-ion force at the drifted positions; flagged ions are
//                  parked far away as sources, so this is the sum over kept ions
//   k_kick2        + flagged-source correction, second half-kick, absorbed
//                  counts and momenta, Philox reinjection, orbit sums
//   k_rows         full force rows of the reinjected ions at their new place
//   k_inflow       force of the reinjected ions on everybody else
//   k_dust         FP64 reductions, charge update, per-step dust records
//
// The reference evaluates the ion-ion force twice per step (OpenMM recomputes
// forces after the reinjection writes x).  Here the second evaluation costs
// O(N * n_reinjected) instead of O(N^2) and involves no cancellation: the
// kept-source sum is shared, only the flagged sources differ.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "od_device.h"

using namespace od;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// message setter for the library's other translation units
int od_set_error_message(int code, const char* msg) { return set_err(code, "%s", msg); }

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return set_err(OD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define CHKRET(expr)      \
  do {                    \
    int r_ = (expr);      \
    if (r_ != OD_OK) return r_; \
  } while (0)

// ---------------------------------------------------------------------------
// device parameters (one copy per context, read-only in kernels)
// ---------------------------------------------------------------------------
struct DevParams {
  ExtParams ext;
  int n, n_pad, n_dust, distribution, use_bottom;
  // position of this simulation in the (possibly batched) arrays
  int ion_off;   // first global ion index
  int stride;    // SoA stride of X, V, F (all simulations)
  int blk0, blk1;  // its 256-ion blocks [blk0, blk1)
  int q_off;     // its dust charges in q[]
  uint32_t sim_id;  // Philox counter word 2
  uint32_t dust_mask;  // flag bits of the dusts
  uint64_t seed;
  double R, H, half_H, dt, half_dt, m_s, v_T, v_fl, lam;
  double pos_scale;   // nm -> kernel units (FP32: L/lambda, FP64: 1/lambda)
  double pair_scale;  // kernel accumulator -> kJ/mol/nm
  float a32, linv32;
  double rho0;
  // truncation (cutoff mode): FP32 values in L/lambda units, FP64 in lambda units
  int use_cut, use_switch;
  float rc32, rs32, invw32;
  double rc64, rs64, invw64;
  // cell grid (nm)
  double cell_h;
  int nx, ny, nz, n_cells;
  uint64_t thr_top, thr_side, thr_coll;
  int dust_method[OD_MAX_DUST];
  double si_k, si_e, si_q_s, si_lam, si_dt, si_kB, si_Te, si_Tn, si_E;
  double gamma_e_pref, oml_A, oml_B, oml_u2, oml_erf_u, oml_tail, collect_scale;
  double si_dx[OD_MAX_DUST], si_dy[OD_MAX_DUST], si_dz[OD_MAX_DUST], si_dr[OD_MAX_DUST];
  od_table top, bottom, side, vz;
};

template <class T>
struct Pos;
template <>
struct Pos<float> {
  using V = float4;
  __device__ static V make(double x, double y, double z, double s) {
    return make_float4((float)(x * s), (float)(y * s), (float)(z * s), 0.0f);
  }
  __device__ static V park() { return make_float4(PARK32, PARK32, PARK32, 0.0f); }
};
template <>
struct Pos<double> {
  using V = double4;
  __device__ static V make(double x, double y, double z, double s) { return make_double4(x * s, y * s, z * s, 0.0); }
  __device__ static V park() { return make_double4(PARK64, PARK64, PARK64, 0.0); }
};

// Accumulator of one force component: FP32 with tile-wise Kahan, or FP64.
template <class T>
struct Acc;
template <>
struct Acc<float> {
  float a = 0.f, s = 0.f, c = 0.f;
  __device__ void flush() {
    kahan_add(s, c, a);
    a = 0.f;
  }
  __device__ double total() const { return (double)s - (double)c; }
};
template <>
struct Acc<double> {
  double a = 0.0;
  __device__ void flush() {}
  __device__ double total() const { return a; }
};

template <class T>
__device__ __forceinline__ void pair_acc(const DevParams& P, const typename Pos<T>::V& i, const typename Pos<T>::V& j,
                                         Acc<T>& ax, Acc<T>& ay, Acc<T>& az);
template <>
__device__ __forceinline__ void pair_acc<float>(const DevParams& P, const float4& i, const float4& j, Acc<float>& ax,
                                                Acc<float>& ay, Acc<float>& az) {
  if (!P.use_cut)
    pair32(i.x, i.y, i.z, j.x, j.y, j.z, P.a32, P.linv32, ax.a, ay.a, az.a);
  else if (P.use_switch)
    pair32_cut<true>(i.x, i.y, i.z, j.x, j.y, j.z, P.a32, P.linv32, P.rc32, P.rs32, P.invw32, ax.a, ay.a, az.a);
  else
    pair32_cut<false>(i.x, i.y, i.z, j.x, j.y, j.z, P.a32, P.linv32, P.rc32, P.rs32, P.invw32, ax.a, ay.a, az.a);
}
template <>
__device__ __forceinline__ void pair_acc<double>(const DevParams& P, const double4& i, const double4& j,
                                                 Acc<double>& ax, Acc<double>& ay, Acc<double>& az) {
  if (!P.use_cut)
    pair64(i.x, i.y, i.z, j.x, j.y, j.z, P.rho0, ax.a, ay.a, az.a);
  else
    pair64_cut(i.x, i.y, i.z, j.x, j.y, j.z, P.rho0, P.use_switch != 0, P.rc64, P.rs64, P.invw64, ax.a, ay.a,
               az.a);
}

// The untruncated law with its constants read once (registers), for loops
// that only run in all-pairs mode: no per-pair branch on the cutoff flags.
template <class T>
struct PlainLaw;
template <>
struct PlainLaw<float> {
  float a, linv;
  __device__ explicit PlainLaw(const DevParams& P) : a(P.a32), linv(P.linv32) {}
  __device__ __forceinline__ void operator()(const float4& i, const float4& j, Acc<float>& ax, Acc<float>& ay,
                                             Acc<float>& az) const {
    pair32(i.x, i.y, i.z, j.x, j.y, j.z, a, linv, ax.a, ay.a, az.a);
  }
};
template <>
struct PlainLaw<double> {
  double rho0;
  __device__ explicit PlainLaw(const DevParams& P) : rho0(P.rho0) {}
  __device__ __forceinline__ void operator()(const double4& i, const double4& j, Acc<double>& ax, Acc<double>& ay,
                                             Acc<double>& az) const {
    pair64(i.x, i.y, i.z, j.x, j.y, j.z, rho0, ax.a, ay.a, az.a);
  }
};

// parameters of the simulation a 256-ion block belongs to (blk_sim == null: one simulation)
__device__ __forceinline__ const DevParams* sim_of_block(const DevParams* Pall, const int* blk_sim, int blk) {
  return blk_sim ? Pall + blk_sim[blk] : Pall;
}

// deterministic block sum of a double over a 256-thread block (value valid on all threads)
__device__ double block_sum256(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  const double r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// positions -> kernel coordinates
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_load_pos(const double* __restrict__ X, int n, int n_pad, double s,
                                                  typename Pos<T>::V* __restrict__ pi,
                                                  typename Pos<T>::V* __restrict__ pj) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pad) return;
  typename Pos<T>::V p = (i < n) ? Pos<T>::make(X[i], X[n_pad + i], X[2 * n_pad + i], s) : Pos<T>::park();
  pi[i] = p;
  pj[i] = p;
}

// ---------------------------------------------------------------------------
// all-pairs ion-ion force, FP32 with tile-compensated sums.
// Block = 256 threads, IB i-ions per thread (stride 256).  j runs over one
// chunk of the padded source array; source positions are wave-uniform, so they
// arrive through scalar loads and feed the VALU as SGPR operands.  Blocks are
// ordered chunk-fastest, so consecutive blocks (which the dispatcher spreads
// over the 8 XCDs) stream different chunks and each XCD's L2 holds one chunk.
// out: [n_chunks][3][n_pad] partial sums already scaled to kJ/mol/nm.
// ---------------------------------------------------------------------------
constexpr int TJ = 64;  // Kahan flush period (sources)

template <int IB>
__global__ __launch_bounds__(256) void k_pair32(const float4* __restrict__ pi, const float4* __restrict__ pj,
                                                int n_pad, int n_chunks, int chunk_len, float a, float linv,
                                                double scale, double* __restrict__ out) {
  const int chunk = blockIdx.x % n_chunks;
  const int iblk = blockIdx.x / n_chunks;
  const int ibase = iblk * (256 * IB) + threadIdx.x;
  float xi[IB], yi[IB], zi[IB];
  float ax[IB], ay[IB], az[IB], sx[IB], sy[IB], sz[IB], cx[IB], cy[IB], cz[IB];
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    const float4 p = pi[ibase + k * 256];
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    sx[k] = sy[k] = sz[k] = 0.f;
    cx[k] = cy[k] = cz[k] = 0.f;
  }
  const int j0 = chunk * chunk_len;
  const int j1 = min(j0 + chunk_len, n_pad);
  for (int jt = j0; jt < j1; jt += TJ) {
#pragma unroll
    for (int k = 0; k < IB; ++k) ax[k] = ay[k] = az[k] = 0.f;
#pragma unroll 8
    for (int jj = 0; jj < TJ; ++jj) {
      const float4 p = pj[jt + jj];
#pragma unroll
      for (int k = 0; k < IB; ++k) pair32(xi[k], yi[k], zi[k], p.x, p.y, p.z, a, linv, ax[k], ay[k], az[k]);
    }
#pragma unroll
    for (int k = 0; k < IB; ++k) {
      kahan_add(sx[k], cx[k], ax[k]);
      kahan_add(sy[k], cy[k], ay[k]);
      kahan_add(sz[k], cz[k], az[k]);
    }
  }
  double* o = out + (size_t)chunk * 3 * n_pad;
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    const int i = ibase + k * 256;
    o[i] = ((double)sx[k] - (double)cx[k]) * scale;
    o[n_pad + i] = ((double)sy[k] - (double)cy[k]) * scale;
    o[2 * n_pad + i] = ((double)sz[k] - (double)cz[k]) * scale;
  }
}

// all-pairs ion-ion force, plain FP64 (lambda units)
template <int IB>
__global__ __launch_bounds__(256) void k_pair64(const double4* __restrict__ pi, const double4* __restrict__ pj,
                                                int n_pad, int n_chunks, int chunk_len, double rho0, double scale,
                                                double* __restrict__ out) {
  const int chunk = blockIdx.x % n_chunks;
  const int iblk = blockIdx.x / n_chunks;
  const int ibase = iblk * (256 * IB) + threadIdx.x;
  double xi[IB], yi[IB], zi[IB], ax[IB], ay[IB], az[IB];
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    const double4 p = pi[ibase + k * 256];
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    ax[k] = ay[k] = az[k] = 0.0;
  }
  const int j0 = chunk * chunk_len;
  const int j1 = min(j0 + chunk_len, n_pad);
#pragma unroll 4
  for (int j = j0; j < j1; ++j) {
    const double4 p = pj[j];
#pragma unroll
    for (int k = 0; k < IB; ++k) pair64(xi[k], yi[k], zi[k], p.x, p.y, p.z, rho0, ax[k], ay[k], az[k]);
  }
  double* o = out + (size_t)chunk * 3 * n_pad;
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    const int i = ibase + k * 256;
    o[i] = ax[k] * scale;
    o[n_pad + i] = ay[k] * scale;
    o[2 * n_pad + i] = az[k] * scale;
  }
}

// F = sum over chunks (fixed order)
__global__ __launch_bounds__(256) void k_chunk_reduce(const double* __restrict__ part, int n_chunks, int n3,
                                                      double* __restrict__ F) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n3) return;
  double s = part[i];
  for (int c = 1; c < n_chunks; ++c) s += part[(size_t)c * n3 + i];
  F[i] = s;
}

// ---------------------------------------------------------------------------
// Symmetric all-pairs ion-ion force, FP32 (every unordered pair once).
//
// Ions are cut into blocks of SYM_TB.  A workgroup owns one target block I
// and `ch` consecutive source blocks J > I (upper triangle); pairs inside a
// block are done by k_pair32_diag.  Inside a workgroup (SYM_TB / (64 MT)
// waves) each wave keeps MT target sub-blocks of 64 ions in registers (MT = 8
// by default; 4 and 16 are autotuner variants) and streams 64-ion source
// sub-blocks.  Within a 16-lane row, step s pairs lane c's targets with the
// source of lane (c - s) mod 16: the source position is read through a DPP
// row-rotate folded into the subtraction, and the source's reaction
// accumulator is rotated by one lane per step, so after 16 steps it is home.
// Rows exchange sources (xor 16 / xor 32) between the four phases, which
// covers all 64 x 64 pairs.  The waves' reactions on a source sub-block
// are summed in a fixed order through LDS and stored as one partial per
// (workgroup, source); target sums are Kahan-compensated across sub-blocks.
// k_sym_reduce adds diagonal, target and source partials in a fixed order.
// 16 + 3 VALU per unordered pair (3 of them transcendental) plus one
// 3-register rotate per MT pairs.
// ---------------------------------------------------------------------------
constexpr int SYM_TB = 1024;  // ions per block of the triangular schedule

template <int S>
__device__ __forceinline__ float dpp_ror(float v) {
  if constexpr (S == 0) {
    return v;
  } else {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v),
                                                                 __builtin_bit_cast(int, v), 0x120 + S, 0xF,
                                                                 0xF, true));
  }
}

template <int MT>
struct SymT {
  float x[MT], y[MT], z[MT];
  float ax[MT], ay[MT], az[MT];
};

// pair-law parameters of the symmetric kernels (FP32, units r_D_e / log2 e)
struct SymLaw {
  float a, linv;
  float rc, rs, invw;  // truncation (cell-list kernel)
};

// Empty asm with a read-write register operand: the value must exist at this
// point, so the optimiser can neither hoist a step's source reads above it nor
// sink an accumulator chain below it (either blows up register pressure).
__device__ __forceinline__ void pin(float& v) { asm volatile("" : "+v"(v)); }

// d = (source rotated by S lanes within each row of 16) - target, for four of
// the wave's target sub-blocks: the rotation rides on the subtraction as a DPP
// modifier.  s_nop 1 keeps the two wait states a VALU write needs before a
// DPP read of the same register.
template <int S>
__device__ __forceinline__ void sym_diffs4(const float* tx, const float* ty, const float* tz, float xj, float yj,
                                           float zj, float* dx, float* dy, float* dz) {
  if constexpr (S == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dx[k] = xj - tx[k];
      dy[k] = yj - ty[k];
      dz[k] = zj - tz[k];
    }
  } else {
    asm("s_nop 1\n\t"
        "v_sub_f32_dpp %0, %12, %15 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %1, %13, %16 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %2, %14, %17 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %3, %12, %18 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %4, %13, %19 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %5, %14, %20 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %6, %12, %21 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %7, %13, %22 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %8, %14, %23 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %9, %12, %25 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %10, %13, %26 row_ror:%c24 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %11, %14, %27 row_ror:%c24 row_mask:0xf bank_mask:0xf"
        : "=&v"(dx[0]), "=&v"(dy[0]), "=&v"(dz[0]), "=&v"(dx[1]), "=&v"(dy[1]), "=&v"(dz[1]), "=&v"(dx[2]),
          "=&v"(dy[2]), "=&v"(dz[2]), "=&v"(dx[3]), "=&v"(dy[3]), "=&v"(dz[3])
        : "v"(xj), "v"(yj), "v"(zj), "v"(tx[0]), "v"(ty[0]), "v"(tz[0]), "v"(tx[1]), "v"(ty[1]),
          "v"(tz[1]), "v"(tx[2]), "v"(ty[2]), "v"(tz[2]), "i"(S), "v"(tx[3]), "v"(ty[3]), "v"(tz[3]));
  }
}

template <int S, int MT>
__device__ __forceinline__ void sym_diffs(const SymT<MT>& T, float xj, float yj, float zj, float* dx, float* dy,
                                          float* dz) {
  static_assert(MT % 4 == 0, "targets per wave in groups of four");
#pragma unroll
  for (int g = 0; g < MT; g += 4)
    sym_diffs4<S>(T.x + g, T.y + g, T.z + g, xj, yj, zj, dx + g, dy + g, dz + g);
}

// CUT: 0 all pairs, 1 hard truncation at rc, 2 OpenMM switch from rs to rc
template <int S, int MT, int CUT>
__device__ __forceinline__ void sym_step(SymT<MT>& T, float xj, float yj, float zj, float& bx, float& by,
                                         float& bz, const SymLaw& L) {
  float DX[MT], DY[MT], DZ[MT];
  sym_diffs<S>(T, xj, yj, zj, DX, DY, DZ);
#pragma unroll
  for (int k = 0; k < MT; ++k) {
    const float dx = DX[k];
    const float dy = DY[k];
    const float dz = DZ[k];
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, EPS32)));
    const float r = __builtin_amdgcn_sqrtf(r2);
    const float e = __builtin_amdgcn_exp2f(-r);
    const float w = __builtin_amdgcn_rcpf(__builtin_fmaf(L.a, r, r2));
    const float t = __builtin_fmaf(r, w, L.linv);
    float h = e * w * t;
    if constexpr (CUT == 2) {
      const float x = __builtin_fminf(__builtin_fmaxf((r - L.rs) * L.invw, 0.0f), 1.0f);
      const float x2 = x * x;
      const float Sx = __builtin_fmaf(x2 * x, __builtin_fmaf(x, __builtin_fmaf(x, -6.0f, 15.0f), -10.0f), 1.0f);
      const float omx = 1.0f - x;
      const float dS = -30.0f * x2 * omx * omx * L.invw;
      h = __builtin_fmaf(h, Sx, -(e * w * dS));
    }
    if constexpr (CUT >= 1) h = (r < L.rc) ? h : 0.0f;
    T.ax[k] = __builtin_fmaf(h, dx, T.ax[k]);
    T.ay[k] = __builtin_fmaf(h, dy, T.ay[k]);
    T.az[k] = __builtin_fmaf(h, dz, T.az[k]);
    bx = __builtin_fmaf(-h, dx, bx);
    by = __builtin_fmaf(-h, dy, by);
    bz = __builtin_fmaf(-h, dz, bz);
  }
  bx = dpp_ror<1>(bx);
  by = dpp_ror<1>(by);
  bz = dpp_ror<1>(bz);
  pin(bx);
  pin(by);
  pin(bz);
#pragma unroll
  for (int k = 0; k < MT; ++k) {
    pin(T.ax[k]);
    pin(T.ay[k]);
    pin(T.az[k]);
  }
}

__device__ __forceinline__ void sym_xchg(int mask, float& xj, float& yj, float& zj, float& bx, float& by,
                                         float& bz) {
  xj = __shfl_xor(xj, mask, 64);
  yj = __shfl_xor(yj, mask, 64);
  zj = __shfl_xor(zj, mask, 64);
  bx = __shfl_xor(bx, mask, 64);
  by = __shfl_xor(by, mask, 64);
  bz = __shfl_xor(bz, mask, 64);
}

template <int MT, int CUT, int... S>
__device__ __forceinline__ void sym_phase(std::integer_sequence<int, S...>, SymT<MT>& T, float xj, float yj,
                                          float zj, float& bx, float& by, float& bz, const SymLaw& L) {
  (sym_step<S, MT, CUT>(T, xj, yj, zj, bx, by, bz, L), ...);
}

// The 64 x 64 pairs of the wave's source sub-block (xj, yj, zj, one per lane)
// with its MT target sub-blocks: four phases of 16 rotation steps, rows
// exchanging sources between phases.  On return (bx, by, bz) hold the
// reaction on the lane's own source and T.a* the targets' accumulators.
template <int MT, int CUT>
__device__ __forceinline__ void sym_tile(SymT<MT>& T, float xj, float yj, float zj, float& bx, float& by, float& bz,
                                         const SymLaw& L) {
  const auto seq = std::make_integer_sequence<int, 16>{};
  sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
  sym_xchg(16, xj, yj, zj, bx, by, bz);
  sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
  sym_xchg(32, xj, yj, zj, bx, by, bz);
  sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
  sym_xchg(16, xj, yj, zj, bx, by, bz);
  sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
  bx = __shfl_xor(bx, 32, 64);
  by = __shfl_xor(by, 32, 64);
  bz = __shfl_xor(bz, 32, 64);
}

// Work-group (I, k) of the launch-order table wg_tab covers target block I and
// source blocks I+1+k*ch ..; its partials live in slot wg_base[I] + k, so the
// launch order (longest first) does not change any sum.
// tp: [n_wg][3][SYM_TB] FP64 target partials; sp: [n_wg][3][ch * SYM_TB] FP32 source partials.
// Batched simulations: blk_end[I] (first block past I's simulation) and
// blk_a[I] (its a) per block; null for one simulation (n_blocks, a).
// Tile shape MT (64-target sub-blocks per wave; the work-group has
// SYM_TB / (64 MT) waves) and WPE (waves per SIMD the register allocation
// must allow) are template parameters; the autotuner picks among the
// instantiated variants (sym_variants).  The next source sub-block's
// positions are loaded while the current one is processed.
template <int MT, int WPE>
__global__ __launch_bounds__(SYM_TB / MT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pair32_sym(
    const float4* __restrict__ pos, int n_blocks, int ch, const int2* __restrict__ wg_tab,
    const int* __restrict__ wg_base, float a, float linv, const int* __restrict__ blk_end,
    const float* __restrict__ blk_a, double* __restrict__ tp, float* __restrict__ sp) {
  constexpr int NW = SYM_TB / (64 * MT);
  __shared__ double red[NW][3][64];
  const int2 ik = wg_tab[blockIdx.x];
  const int I = ik.x;
  const size_t slot = (size_t)wg_base[I] + ik.y;
  const int J0 = I + 1 + ik.y * ch;
  const int J1 = min(J0 + ch, blk_end ? blk_end[I] : n_blocks);
  if (blk_a) a = blk_a[I];
  const int nsub = (J1 - J0) * (SYM_TB / 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t sp_len = (size_t)ch * SYM_TB;
  SymT<MT> T;
  float sx[MT], sy[MT], sz[MT], cx[MT], cy[MT], cz[MT];
#pragma unroll
  for (int k = 0; k < MT; ++k) {
    const float4 p = pos[(size_t)I * SYM_TB + wave * (64 * MT) + k * 64 + lane];
    T.x[k] = p.x;
    T.y[k] = p.y;
    T.z[k] = p.z;
    sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
  }
  const float4* src = pos + (size_t)J0 * SYM_TB + lane;
  float* spw = sp + slot * 3 * sp_len;
  const SymLaw L{a, linv, 0.f, 0.f, 0.f};
  float4 p = src[0];
  for (int b = 0; b < nsub; ++b) {
    const float4 pn = src[(size_t)min(b + 1, nsub - 1) * 64];
    float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
    for (int k = 0; k < MT; ++k) T.ax[k] = T.ay[k] = T.az[k] = 0.f;
    sym_tile<MT, 0>(T, p.x, p.y, p.z, bx, by, bz, L);
#pragma unroll
    for (int k = 0; k < MT; ++k) {
      kahan_add(sx[k], cx[k], T.ax[k]);
      kahan_add(sy[k], cy[k], T.ay[k]);
      kahan_add(sz[k], cz[k], T.az[k]);
    }
    red[wave][0][lane] = bx;
    red[wave][1][lane] = by;
    red[wave][2][lane] = bz;
    __syncthreads();
    for (int e = threadIdx.x; e < 192; e += 64 * NW) {
      const int c = e >> 6, l = e & 63;
      double v = red[0][c][l];
#pragma unroll
      for (int w = 1; w < NW; ++w) v += red[w][c][l];
      spw[c * sp_len + (size_t)b * 64 + l] = (float)v;
    }
    __syncthreads();
    p = pn;
  }
  double* tpw = tp + slot * 3 * SYM_TB;
#pragma unroll
  for (int k = 0; k < MT; ++k) {
    const int t = wave * (64 * MT) + k * 64 + lane;
    tpw[t] = (double)sx[k] - (double)cx[k];
    tpw[SYM_TB + t] = (double)sy[k] - (double)cy[k];
    tpw[2 * SYM_TB + t] = (double)sz[k] - (double)cz[k];
  }
}

using SymKernel = void (*)(const float4*, int, int, const int2*, const int*, float, float, const int*, const float*,
                           double*, float*);
// Variants of the symmetric kernel the autotuner chooses from (od_config.sym_tile
// - 1): tile shape MT and the waves per SIMD the register budget must allow.
// As compiled for gfx950: <8,3> 145 VGPRs (3 waves per SIMD), <8,4> 128 (4, a
// few registers spilled), <4,5> 84 (5), <16,2> 256 (2, a few spilled), <16,1>
// 264 (1).
struct SymVariant {
  SymKernel fn;
  int mt, wpe;
};
static const SymVariant sym_variants[] = {
    {k_pair32_sym<8, 3>, 8, 3},   {k_pair32_sym<8, 4>, 8, 4},   {k_pair32_sym<4, 5>, 4, 5},
    {k_pair32_sym<16, 2>, 16, 2}, {k_pair32_sym<16, 1>, 16, 1},
};
constexpr int SYM_N_VARIANTS = (int)(sizeof(sym_variants) / sizeof(sym_variants[0]));

// pairs inside one block (ordered, sources via scalar loads); dg: [3][n_pad]
__global__ __launch_bounds__(256) void k_pair32_diag(const float4* __restrict__ pos, int n_pad, float a,
                                                     float linv, const float* __restrict__ blk_a,
                                                     double* __restrict__ dg) {
  const int X = blockIdx.x / (SYM_TB / 256), q = blockIdx.x % (SYM_TB / 256);
  if (blk_a) a = blk_a[X];
  const int i = X * SYM_TB + q * 256 + threadIdx.x;
  const float4 pi = pos[i];
  float ax = 0.f, ay = 0.f, az = 0.f, sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  const float4* src = pos + (size_t)X * SYM_TB;
  for (int jt = 0; jt < SYM_TB; jt += TJ) {
    ax = ay = az = 0.f;
#pragma unroll 8
    for (int jj = 0; jj < TJ; ++jj) {
      const float4 p = src[jt + jj];
      pair32(pi.x, pi.y, pi.z, p.x, p.y, p.z, a, linv, ax, ay, az);
    }
    kahan_add(sx, cx, ax);
    kahan_add(sy, cy, ay);
    kahan_add(sz, cz, az);
  }
  dg[i] = (double)sx - (double)cx;
  dg[n_pad + i] = (double)sy - (double)cy;
  dg[2 * n_pad + i] = (double)sz - (double)cz;
}

// F = scale * (diagonal + target partials + source partials), fixed order.
// Batched simulations: blk_begin[X] (first block of X's simulation) and
// blk_scale[X] per block; null for one simulation (0, scale).
__global__ __launch_bounds__(256) void k_sym_reduce(const double* __restrict__ dg, const double* __restrict__ tp,
                                                    const float* __restrict__ sp, const int* __restrict__ wg_base,
                                                    int n_blocks, int ch, int n_pad, double scale,
                                                    const int* __restrict__ blk_begin,
                                                    const double* __restrict__ blk_scale, double* __restrict__ F) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pad) return;
  const int X = i / SYM_TB, off = i % SYM_TB;
  const int Jb = blk_begin ? blk_begin[X] : 0;
  if (blk_scale) scale = blk_scale[X];
  const size_t sp_len = (size_t)ch * SYM_TB;
  const int k0 = wg_base[X], k1 = wg_base[X + 1];
  // the three components walk the partials together (three independent
  // chains of loads in flight); each component's order is J ascending, then w
  double s0 = dg[i], s1 = dg[(size_t)n_pad + i], s2 = dg[2 * (size_t)n_pad + i];
#pragma unroll 4
  for (int J = Jb; J < X; ++J) {
    const int d = X - J - 1;
    const size_t wg = (size_t)wg_base[J] + d / ch;
    const float* q = sp + wg * 3 * sp_len + (size_t)(d % ch) * SYM_TB + off;
    s0 += (double)q[0];
    s1 += (double)q[sp_len];
    s2 += (double)q[2 * sp_len];
  }
#pragma unroll 4
  for (int w = k0; w < k1; ++w) {
    const double* t = tp + (size_t)w * 3 * SYM_TB + off;
    s0 += t[0];
    s1 += t[SYM_TB];
    s2 += t[2 * SYM_TB];
  }
  F[i] = s0 * scale;
  F[(size_t)n_pad + i] = s1 * scale;
  F[2 * (size_t)n_pad + i] = s2 * scale;
}

// ---------------------------------------------------------------------------
// step kernels
// ---------------------------------------------------------------------------
constexpr uint32_t FLAG_OUT_R = 1u << 16;
constexpr uint32_t FLAG_OUT_Z = 1u << 17;

// Collisions (fieldDriven, opendust.py:869-873), first half-kick, drift, flags
// (opendust.py:864-910).  pj gets the drifted position, or the parking spot
// for flagged ions, which are excluded as sources of the main pair kernel.
template <class T>
__global__ __launch_bounds__(256) void k_kick_drift(const DevParams* __restrict__ Pall, const int* __restrict__ blk_sim,
                                                    double* __restrict__ X, double* __restrict__ V,
                                                    const double* __restrict__ F, const double* __restrict__ qall,
                                                    const uint32_t* __restrict__ st,
                                                    typename Pos<T>::V* __restrict__ pi,
                                                    typename Pos<T>::V* __restrict__ pj, uint32_t* __restrict__ flags,
                                                    int* __restrict__ block_cnt, int* __restrict__ err) {
  const DevParams* P = sim_of_block(Pall, blk_sim, blockIdx.x);
  const double* q = qall + P->q_off;
  const int np = P->stride;
  const int i = blockIdx.x * 256 + threadIdx.x;  // global index
  const int li = i - P->ion_off;                 // index within the simulation
  int flagged = 0;
  if (li < P->n) {
    double x = X[i], y = X[np + i], z = X[2 * np + i];
    double vx = V[i], vy = V[np + i], vz = V[2 * np + i];
    const uint32_t step = st[0];
    if (P->distribution == OD_DIST_FIELD_DRIVEN) {
      const u32x4 r = philox4x32_10((uint32_t)li, step, P->sim_id, S_COLL, P->seed);
      const bool c0 = r.v[0] < P->thr_coll, c1 = r.v[1] < P->thr_coll, c2 = r.v[2] < P->thr_coll;
      if (c0 || c1 || c2) {
        double g0, g1, g2, gd;
        box_muller(philox4x32_10((uint32_t)li, step, P->sim_id, S_COLL | (1u << 8), P->seed), g0, g1);
        box_muller(philox4x32_10((uint32_t)li, step, P->sim_id, S_COLL | (2u << 8), P->seed), g2, gd);
        if (c0) vx = P->v_T * g0;
        if (c1) vy = P->v_T * g1;
        if (c2) vz = P->v_T * g2;
      }
    }
    double fx, fy, fz;
    ext_force(P->ext, q, x, y, z, fx, fy, fz);
    fx += F[i];
    fy += F[np + i];
    fz += F[2 * np + i];
    vx = vx + P->half_dt * fx / P->m_s;
    vy = vy + P->half_dt * fy / P->m_s;
    vz = vz + P->half_dt * fz / P->m_s;
    x = x + P->dt * vx;
    y = y + P->dt * vy;
    z = z + P->dt * vz;
    uint32_t fl = 0;
    for (int p = 0; p < P->n_dust; ++p) {
      const double dx = x - P->ext.dx[p], dy = y - P->ext.dy[p], dz = z - P->ext.dz[p];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (-d + P->ext.dr[p] >= 0.0) fl |= 1u << p;
    }
    if (-P->R + sqrt(x * x + y * y) >= 0.0) fl |= FLAG_OUT_R;
    if (-P->half_H + fabs(z) >= 0.0) fl |= FLAG_OUT_Z;
    if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(vx) && isfinite(vy) && isfinite(vz)))
      atomicOr(err, 1);
    X[i] = x;
    X[np + i] = y;
    X[2 * np + i] = z;
    V[i] = vx;
    V[np + i] = vy;
    V[2 * np + i] = vz;
    flags[i] = fl;
    const typename Pos<T>::V p = Pos<T>::make(x, y, z, P->pos_scale);
    pi[i] = p;
    pj[i] = fl ? Pos<T>::park() : p;
    flagged = fl != 0;
  }
  const int cnt = __syncthreads_count(flagged);
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = cnt;
}

// exclusive scan of the per-block flagged counts (one block)
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ block_cnt, int nb, int* __restrict__ block_off,
                                               int* __restrict__ n_flag) {
  __shared__ int s[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = min(t * per, nb), b1 = min(b0 + per, nb);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += block_cnt[b];
  s[t] = local;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = (t >= o) ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = s[t] - local;  // exclusive prefix of this thread's range
  for (int b = b0; b < b1; ++b) {
    block_off[b] = run;
    run += block_cnt[b];
  }
  if (t == 1023) n_flag[0] = s[1023];
}

// Exclusive scan of long int arrays in three launches: tile sums (1024
// entries per block), a one-block scan of the tile sums (k_scan), and the
// scan inside each tile plus its offset.  Integer sums: exact in any order.
__global__ __launch_bounds__(256) void k_tile_sum(const int* __restrict__ in, int n, int* __restrict__ tsum) {
  __shared__ int red[4];
  const int base = blockIdx.x * 1024;
  int v = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = base + r * 256 + (int)threadIdx.x;
    if (i < n) v += in[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) tsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_tile_scan(const int* __restrict__ in, int n, const int* __restrict__ toff,
                                                   int* __restrict__ out) {
  __shared__ int wsum[4];
  const int i0 = blockIdx.x * 1024 + 4 * (int)threadIdx.x;
  int v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : 0;
  const int t = v[0] + v[1] + v[2] + v[3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int run = toff[blockIdx.x] + incl - t;
  for (int w = 0; w < wave; ++w) run += wsum[w];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (i0 + k < n) out[i0 + k] = run;
    run += v[k];
  }
}

// ordered list of flagged ions: slot = offset of the block + rank inside it
template <class T>
__global__ __launch_bounds__(256) void k_list(const uint32_t* __restrict__ flags, const int* __restrict__ block_cnt,
                                              const int* __restrict__ block_off, int n,
                                              const typename Pos<T>::V* __restrict__ pi, int* __restrict__ slot_of,
                                              int* __restrict__ list_ion, typename Pos<T>::V* __restrict__ list_old) {
  if (block_cnt[blockIdx.x] == 0) return;
  __shared__ int wcnt[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool f = (i < n) && flags[i] != 0u;
  const unsigned long long m = __ballot(f);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r_in_wave = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wcnt[wave] = __popcll(m);
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wcnt[w];
  if (f) {
    const int s = block_off[blockIdx.x] + before + r_in_wave;
    slot_of[i] = s;
    list_ion[s] = i;
    list_old[s] = pi[i];
  }
}

// Reinjection of one flagged ion (opendust.py:923-975), FP64.
// face: 1 top, 2 side, 3 bottom.
__device__ int reinject(const DevParams& P, uint32_t ion, uint32_t step, double& x, double& y, double& z, double& vx,
                        double& vy, double& vz) {
  const double two_pi = 2.0 * 3.141592653589793;
  const uint32_t uf = philox4x32_10(ion, step, P.sim_id, S_FACE, P.seed).v[0];
  int face;
  if ((uint64_t)uf < P.thr_top)
    face = 1;
  else if ((uint64_t)uf < P.thr_side)
    face = 2;
  else
    face = 3;
  if (face == 1 || face == 3) {
    const bool top = face == 1;
    double g1, g2;
    box_muller(philox4x32_10(ion, step, P.sim_id, top ? S_TOPV : S_BOTV, P.seed), g1, g2);
    const u32x4 rz = philox4x32_10(ion, step, P.sim_id, (top ? S_TOPV : S_BOTV) | (1u << 8), P.seed);
    const double uz = u53(rz.v[0], rz.v[1]);
    vx = P.v_T * g1;
    vy = P.v_T * g2;
    vz = table_eval(top ? P.top : P.bottom, uz);
    const u32x4 rx = philox4x32_10(ion, step, P.sim_id, top ? S_TOPX : S_BOTX, P.seed);
    const double ux = u53(rx.v[0], rx.v[1]), uy = u53(rx.v[2], rx.v[3]);
    const double rr = P.R * sqrt(ux);
    x = rr * cos(two_pi * uy);
    y = rr * sin(two_pi * uy);
    z = top ? -P.half_H : P.half_H;
  } else {
    const u32x4 rx = philox4x32_10(ion, step, P.sim_id, S_SIDEX, P.seed);
    const double uy = u53(rx.v[0], rx.v[1]), uz = u53(rx.v[2], rx.v[3]);
    x = P.R * cos(two_pi * uy);
    y = P.R * sin(two_pi * uy);
    z = uz * P.H - P.half_H;
    double gy, gz;
    box_muller(philox4x32_10(ion, step, P.sim_id, S_SIDEV, P.seed), gy, gz);
    const u32x4 ru = philox4x32_10(ion, step, P.sim_id, S_SIDEV | (1u << 8), P.seed);
    const double us = u53(ru.v[0], ru.v[1]);
    const double vs = table_eval(P.side, us);
    vx = -vs * (x / P.R) + P.v_T * gy * (y / P.R);
    vy = -vs * (y / P.R) - P.v_T * gy * (x / P.R);
    if (P.distribution == OD_DIST_FIELD_DRIVEN)
      vz = table_eval(P.vz, us);
    else
      vz = gz * P.v_T + P.v_fl;
  }
  return face;
}

// Second half-kick with the flagged-source correction, absorbed counts and
// momenta (opendust.py:911-921), reinjection, orbit sums at the final
// positions (opendust.py:977-1037).  partials: [nb][OD_MAX_DUST][8].
template <class T>
__global__ __launch_bounds__(256) void k_kick2(const DevParams* __restrict__ Pall, const int* __restrict__ blk_sim,
                                               const int* __restrict__ block_off, int nb_all,
                                               double* __restrict__ X, double* __restrict__ V,
                                               const double* __restrict__ F, const double* __restrict__ qall,
                                               const uint32_t* __restrict__ st,
                                               const typename Pos<T>::V* __restrict__ pi,
                                               typename Pos<T>::V* __restrict__ pj,
                                               const uint32_t* __restrict__ flags, const int* __restrict__ n_flag,
                                               const int* __restrict__ slot_of,
                                               const typename Pos<T>::V* __restrict__ list_old,
                                               typename Pos<T>::V* __restrict__ list_new, uint8_t* __restrict__ face_out,
                                               const double* __restrict__ rowD, const double* __restrict__ G,
                                               double* __restrict__ partials) {
  __shared__ double lds[4];
  const DevParams* P = sim_of_block(Pall, blk_sim, blockIdx.x);
  const double* q = qall + P->q_off;
  const int np = P->stride;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int li = i - P->ion_off;
  const bool valid = li < P->n;
  const typename Pos<T>::V own = valid ? pi[i] : Pos<T>::park();
  Acc<T> ax, ay, az;
  // this simulation's flagged ions occupy list slots [l0, l1); with G the
  // flagged sources' share was computed in cell order (cutoff mode)
  const int l0 = block_off[P->blk0];
  const int l1 = G ? l0 : ((P->blk1 < nb_all) ? block_off[P->blk1] : n_flag[0]);
  const PlainLaw<T> law(*P);  // the list loop runs in all-pairs mode only
  for (int s0 = l0; s0 < l1; s0 += TJ) {
    const int s1 = min(s0 + TJ, l1);
    for (int s = s0; s < s1; ++s) law(own, list_old[s], ax, ay, az);
    ax.flush();
    ay.flush();
    az.flush();
  }
  uint32_t fl = 0;
  double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0, px = 0, py = 0, pz = 0;
  if (valid) {
    fl = flags[i];
    x = X[i];
    y = X[np + i];
    z = X[2 * np + i];
    vx = V[i];
    vy = V[np + i];
    vz = V[2 * np + i];
    double fx, fy, fz;
    ext_force(P->ext, q, x, y, z, fx, fy, fz);
    if (fl & P->dust_mask) {
      // absorbed: its full row at the drifted state (the main kernel parks flagged ions)
      const int s = slot_of[i];
      fx += rowD[3 * s];
      fy += rowD[3 * s + 1];
      fz += rowD[3 * s + 2];
    } else {
      // flagged ions that only left the domain are reinjected below, whatever this kick gives
      if (G) {
        fx += F[i] + G[i];
        fy += F[np + i] + G[np + i];
        fz += F[2 * np + i] + G[2 * np + i];
      } else {
        fx += F[i] + ax.total() * P->pair_scale;
        fy += F[np + i] + ay.total() * P->pair_scale;
        fz += F[2 * np + i] + az.total() * P->pair_scale;
      }
    }
    vx = vx + P->half_dt * fx / P->m_s;
    vy = vy + P->half_dt * fy / P->m_s;
    vz = vz + P->half_dt * fz / P->m_s;
    px = vx;
    py = vy;
    pz = vz;
    int face = 0;
    if (fl) {
      face = reinject(*P, (uint32_t)li, st[0], x, y, z, vx, vy, vz);
      X[i] = x;
      X[np + i] = y;
      X[2 * np + i] = z;
      const typename Pos<T>::V p = Pos<T>::make(x, y, z, P->pos_scale);
      pj[i] = p;
      list_new[slot_of[i]] = p;
    }
    V[i] = vx;
    V[np + i] = vy;
    V[2 * np + i] = vz;
    face_out[i] = (uint8_t)face;
  }
  for (int p = 0; p < P->n_dust; ++p) {
    const double in = (valid && ((fl >> p) & 1u)) ? 1.0 : 0.0;
    double ox = 0, oy = 0, oz = 0;
    if (valid) {
      const double xp = P->ext.dx[p], yp = P->ext.dy[p], zp = P->ext.dz[p], lam = P->lam;
      const double dx = x - xp, dy = y - yp, dz = z - zp;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double e = exp(-r / lam);
      const double r3 = r * r * r, r2 = r * r, l2 = lam * lam;
      ox = e * (xp - x) / r3 * l2 + e * (xp - x) / r2 * lam;
      oy = e * (yp - y) / r3 * l2 + e * (yp - y) / r2 * lam;
      oz = e * (zp - z) / r3 * l2 + e * (zp - z) / r2 * lam;
    }
    const double v[7] = {in, in * px, in * py, in * pz, ox, oy, oz};
    double* dst = partials + ((size_t)blockIdx.x * OD_MAX_DUST + p) * 8;
    for (int k = 0; k < 7; ++k) {
      const double s = block_sum256(v[k], lds);
      if (threadIdx.x == 0) dst[k] = s;
    }
  }
}

// Full force rows of listed ions against all sources `pj`: the reinjected ions
// at the post-reinjection state (need_mask = 0), or the absorbed ions at the
// drifted state (need_mask = dust bits: other listed ions are skipped).
// A work-group of RowShape<T>::TB threads takes RowShape<T>::G consecutive
// list slots and streams its simulation's sources once for all of them (each
// source load feeds G pairs); slots of another simulation in the same group
// (batched runs) get a pass of their own.  FP32: 4 slots, 16 waves (about
// 1200 reinjected ions per step at 2^20 give ~300 work-groups, each 16 waves
// deep for latency hiding); FP64: 2 slots, 8 waves (twice the groups for the
// heavier pairs).  Per thread: sources strided by TB, Kahan flush every
// TJ of them; then a fixed-order sum over the work-group.  All-pairs mode only
// (the cell-list mode has k_rows_cell).
template <class T>
struct RowShape;
template <>
struct RowShape<float> {
  static constexpr int G = 4, TB = 1024;
};
template <>
struct RowShape<double> {
  static constexpr int G = 2, TB = 512;
};

template <class T>
__global__ __launch_bounds__(RowShape<T>::TB) void k_rows(const DevParams* __restrict__ Pall, const int* __restrict__ blk_sim,
                                                 const typename Pos<T>::V* __restrict__ pj,
                                                 const int* __restrict__ n_flag,
                                                 const typename Pos<T>::V* __restrict__ list_new,
                                                 const uint32_t* __restrict__ flags, const int* __restrict__ list_ion,
                                                 uint32_t need_mask, double* __restrict__ rowF) {
  using PV = typename Pos<T>::V;
  constexpr int ROW_G = RowShape<T>::G, ROW_TB = RowShape<T>::TB;
  constexpr int NW = ROW_TB / 64, NV = 3 * ROW_G;
  __shared__ double lds[NW][NV];
  const int nf = n_flag[0];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int s0 = blockIdx.x * ROW_G; s0 < nf; s0 += gridDim.x * ROW_G) {
    uint32_t todo = 0;
#pragma unroll
    for (int k = 0; k < ROW_G; ++k) {
      const int s = s0 + k;
      if (s < nf && (need_mask == 0u || (flags[list_ion[s]] & need_mask) != 0u)) todo |= 1u << k;
    }
    while (todo) {
      const int sim_blk = list_ion[s0 + __builtin_ctz(todo)] >> 8;
      const DevParams* P = sim_of_block(Pall, blk_sim, sim_blk);
      uint32_t mine = 0;
      PV own[ROW_G];
#pragma unroll
      for (int k = 0; k < ROW_G; ++k) {
        own[k] = Pos<T>::park();
        if (((todo >> k) & 1u) && sim_of_block(Pall, blk_sim, list_ion[s0 + k] >> 8) == P) {
          mine |= 1u << k;
          own[k] = list_new[s0 + k];
        }
      }
      todo &= ~mine;
      Acc<T> ax[ROW_G], ay[ROW_G], az[ROW_G];
      const PlainLaw<T> law(*P);
      const int jb = P->ion_off, je = P->ion_off + P->n_pad;
      for (int j0 = jb + (int)threadIdx.x; j0 < je; j0 += ROW_TB * TJ) {
        const int j1 = min(j0 + ROW_TB * TJ, je);
        // four sources in flight per thread: their loads are issued together,
        // so one memory latency is spread over 4 x ROW_G pair evaluations
        int j = j0;
        for (; j + 3 * ROW_TB < j1; j += 4 * ROW_TB) {
          const PV p0 = pj[j], p1 = pj[j + ROW_TB], p2 = pj[j + 2 * ROW_TB], p3 = pj[j + 3 * ROW_TB];
#pragma unroll
          for (int k = 0; k < ROW_G; ++k) law(own[k], p0, ax[k], ay[k], az[k]);
#pragma unroll
          for (int k = 0; k < ROW_G; ++k) law(own[k], p1, ax[k], ay[k], az[k]);
#pragma unroll
          for (int k = 0; k < ROW_G; ++k) law(own[k], p2, ax[k], ay[k], az[k]);
#pragma unroll
          for (int k = 0; k < ROW_G; ++k) law(own[k], p3, ax[k], ay[k], az[k]);
        }
        for (; j < j1; j += ROW_TB) {
          const PV p = pj[j];
#pragma unroll
          for (int k = 0; k < ROW_G; ++k) law(own[k], p, ax[k], ay[k], az[k]);
        }
#pragma unroll
        for (int k = 0; k < ROW_G; ++k) {
          ax[k].flush();
          ay[k].flush();
          az[k].flush();
        }
      }
      double v[NV];
#pragma unroll
      for (int k = 0; k < ROW_G; ++k) {
        v[3 * k + 0] = ax[k].total();
        v[3 * k + 1] = ay[k].total();
        v[3 * k + 2] = az[k].total();
      }
#pragma unroll
      for (int m = 0; m < NV; ++m) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[m] += __shfl_xor(v[m], o, 64);
      }
      if (lane == 0) {
#pragma unroll
        for (int m = 0; m < NV; ++m) lds[wave][m] = v[m];
      }
      __syncthreads();
      if (threadIdx.x < NV) {
        const int m = threadIdx.x, k = m / 3;
        double s = lds[0][m];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += lds[w][m];
        if ((mine >> k) & 1u) rowF[3 * (s0 + k) + m % 3] = s * P->pair_scale;
      }
      __syncthreads();
    }
  }
}

// Force of the reinjected ions on the kept ions; reinjected ions take their rows.
template <class T>
__global__ __launch_bounds__(256) void k_inflow(const DevParams* __restrict__ Pall, const int* __restrict__ blk_sim,
                                                const int* __restrict__ block_off, int nb_all, double* __restrict__ F,
                                                const typename Pos<T>::V* __restrict__ pj,
                                                const uint32_t* __restrict__ flags, const int* __restrict__ n_flag,
                                                const typename Pos<T>::V* __restrict__ list_new,
                                                const int* __restrict__ slot_of, const double* __restrict__ rowF,
                                                int rows_only) {
  const DevParams* P = sim_of_block(Pall, blk_sim, blockIdx.x);
  const int np = P->stride;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool valid = i - P->ion_off < P->n;
  const typename Pos<T>::V own = valid ? pj[i] : Pos<T>::park();
  Acc<T> ax, ay, az;
  // rows_only: the kept ions' share was added in cell order (cutoff mode)
  const int l0 = block_off[P->blk0];
  const int l1 = rows_only ? l0 : ((P->blk1 < nb_all) ? block_off[P->blk1] : n_flag[0]);
  const PlainLaw<T> law(*P);  // the list loop runs in all-pairs mode only
  for (int s0 = l0; s0 < l1; s0 += TJ) {
    const int s1 = min(s0 + TJ, l1);
    for (int s = s0; s < s1; ++s) law(own, list_new[s], ax, ay, az);
    ax.flush();
    ay.flush();
    az.flush();
  }
  if (!valid) return;
  if (flags[i]) {
    const int s = slot_of[i];
    F[i] = rowF[3 * s + 0];
    F[np + i] = rowF[3 * s + 1];
    F[2 * np + i] = rowF[3 * s + 2];
  } else if (!rows_only) {
    F[i] += ax.total() * P->pair_scale;
    F[np + i] += ay.total() * P->pair_scale;
    F[2 * np + i] += az.total() * P->pair_scale;
  }
}

// Dust records: FP64 block-partial reduction (fixed order), charge update
// (opendust.py:1117-1202), dust forces (opendust.py:1204-1384).
// One block per simulation (grid = number of simulations); records go to
// rec[(step - base) * n_sims + sim][nd_max].  The step counter is advanced by
// block 0 when `advance` is set (one simulation), else by k_step_advance.
__global__ __launch_bounds__(256) void k_dust(const DevParams* __restrict__ Pall, const double* __restrict__ partials,
                                              double* __restrict__ qall, uint32_t* __restrict__ st,
                                              od_dust_record* __restrict__ rec, int nd_max, int advance) {
  __shared__ double lds[4];
  __shared__ double tot[OD_MAX_DUST][8];
  const DevParams* P = Pall + blockIdx.x;
  double* q = qall + P->q_off;
  const int nd = P->n_dust;
  for (int p = 0; p < nd; ++p) {
    for (int k = 0; k < 7; ++k) {
      double s = 0.0;
      for (int b = P->blk0 + (int)threadIdx.x; b < P->blk1; b += 256)
        s += partials[((size_t)b * OD_MAX_DUST + p) * 8 + k];
      s = block_sum256(s, lds);
      if (threadIdx.x == 0) tot[p][k] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const uint32_t i = st[0];
  const uint32_t ri = i - st[1];
  const double pi_ = 3.141592653589793;
  double qn[OD_MAX_DUST];
  for (int p = 0; p < nd; ++p) {
    const double qp = q[p];
    qn[p] = qp;
    const int m = P->dust_method[p];
    if (i > 0 && (m == OD_CHARGE_CONSISTENT || m == OD_CHARGE_OML)) {
      const double rp = P->si_dr[p];
      const double phi = P->si_k * qp / rp;
      const double gamma_e = P->gamma_e_pref * exp(-P->si_e * phi / P->si_Te / P->si_kB);
      const double S = rp * rp * 4.0 * pi_;
      double dq;
      if (m == OD_CHARGE_CONSISTENT) {
        dq = tot[p][0] * P->si_q_s + gamma_e * P->si_dt * S * P->si_e;
      } else {
        const double chi = P->si_e * phi / P->si_Tn / P->si_kB;
        const double gamma_i = P->oml_A * ((P->oml_B + chi / P->oml_u2) * P->oml_erf_u + P->oml_tail);
        dq = gamma_i * P->si_dt * S * fabs(P->si_e) + gamma_e * P->si_dt * S * P->si_e;
      }
      qn[p] = qp + dq;
    }
  }
  const double lam = P->si_lam, k = P->si_k, qs = P->si_q_s;
  const double* al = P->ext.alpha;
  for (int p = 0; p < nd; ++p) {
    od_dust_record r;
    r.q = qn[p];
    const double xp = P->si_dx[p], yp = P->si_dy[p], zp = P->si_dz[p];
    double fdx = 0, fdy = 0, fdz = 0;
    for (int j = 0; j < nd; ++j) {
      if (j == p) continue;
      const double dx = xp - P->si_dx[j], dy = yp - P->si_dy[j], dz = zp - P->si_dz[j];
      const double rr = sqrt(dx * dx + dy * dy + dz * dz);
      const double r3 = rr * rr * rr;
      fdx += k * qn[p] * qn[j] * dx / r3;
      fdy += k * qn[p] * qn[j] * dy / r3;
      fdz += k * qn[p] * qn[j] * dz / r3;
    }
    r.force_dust[0] = fdx;
    r.force_dust[1] = fdy;
    r.force_dust[2] = fdz;
    // in-domain (orbit) part
    const double cq = qn[p] * k * qs;
    const double l2 = lam * lam;
    double fo[3];
    for (int c = 0; c < 3; ++c) fo[c] = tot[p][4 + c] * cq / l2;
    // out-of-domain part: analytic -grad of the trap polynomial at the dust
    const double rr = sqrt(xp * xp + yp * yp);
    const double z = zp;
    const double r2 = rr * rr, z2 = z * z, l4 = l2 * l2, l6 = l4 * l2, l8 = l4 * l4;
    const double r3 = r2 * rr, r5 = r3 * r2, r7 = r5 * r2, r4 = r2 * r2, r6 = r4 * r2;
    const double z3 = z2 * z, z4 = z2 * z2, z5 = z4 * z, z6 = z4 * z2, z7 = z6 * z;
    double fR = (-2 * al[1] * rr) / l2;
    fR += (-2 * al[3] * rr * z2 - 4 * al[4] * r3) / l4;
    fR += (-4 * al[6] * r3 * z2 - 2 * al[7] * rr * z4 - 6 * al[8] * r5) / l6;
    fR += (-4 * al[10] * r3 * z4 - 6 * al[11] * r5 * z2 - 2 * al[12] * rr * z6 - 8 * al[13] * r7) / l8;
    fR *= k * qn[p] * qs / lam;
    double fH = (-2 * al[2] * z) / l2;
    fH += (-2 * al[3] * r2 * z - 4 * al[5] * z3) / l4;
    fH += (-2 * al[6] * r4 * z - 4 * al[7] * r2 * z3 - 6 * al[9] * z5) / l6;
    fH += (-4 * al[10] * r4 * z3 - 2 * al[11] * r6 * z - 6 * al[12] * r2 * z5 - 8 * al[14] * z7) / l8;
    fH *= k * qn[p] * qs / lam;
    double ofx = 0, ofy = 0;
    if (rr != 0.0) {
      ofx = fR * xp / rr;
      ofy = fR * yp / rr;
    }
    r.force_orbit[0] = fo[0] + ofx;
    r.force_orbit[1] = fo[1] + ofy;
    r.force_orbit[2] = fo[2] + fH;
    for (int c = 0; c < 3; ++c) {
      r.force_collect[c] = tot[p][1 + c] * P->collect_scale;
      r.orbit_sum[c] = tot[p][4 + c];
      r.momentum_sum[c] = tot[p][1 + c];
    }
    r.force_field[0] = 0.0;
    r.force_field[1] = 0.0;
    r.force_field[2] = (P->distribution == OD_DIST_FIELD_DRIVEN) ? qn[p] * P->si_E : 0.0;
    r.absorbed = (int64_t)tot[p][0];
    rec[((size_t)ri * gridDim.x + blockIdx.x) * nd_max + p] = r;
  }
  for (int p = 0; p < nd; ++p) q[p] = qn[p];
  if (advance) st[0] = i + 1;
}

__global__ void k_step_advance(uint32_t* __restrict__ st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st[0] = st[0] + 1;
}

// ---------------------------------------------------------------------------
// trap background potential (Monte Carlo, opendust.py:499-523)
// ---------------------------------------------------------------------------
__host__ __device__ inline void trap_sample(double R, double H, uint64_t seed, int64_t k, double& x, double& y,
                                            double& z) {
  const u32x4 a = philox4x32_10((uint32_t)k, (uint32_t)((uint64_t)k >> 32), 0u, S_TRAP, seed);
  const u32x4 b = philox4x32_10((uint32_t)k, (uint32_t)((uint64_t)k >> 32), 0u, S_TRAP | (1u << 8), seed);
  const double r = sqrt(u53(a.v[0], a.v[1]));
  const double theta = u53(a.v[2], a.v[3]);
  x = R * r * cos(2.0 * 3.141592653589793 * theta);
  y = R * r * sin(2.0 * 3.141592653589793 * theta);
  z = -H / 2.0 + (H / 2.0 - (-H / 2.0)) * u53(b.v[0], b.v[1]);
}

__global__ __launch_bounds__(256) void k_trap_samples(double R, double H, uint64_t seed, int64_t n,
                                                      double* __restrict__ xs, double* __restrict__ ys,
                                                      double* __restrict__ zs) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  double x, y, z;
  trap_sample(R, H, seed, k, x, y, z);
  xs[k] = x;
  ys[k] = y;
  zs[k] = z;
}

__global__ __launch_bounds__(256) void k_trap_grid(const double* __restrict__ rho, const double* __restrict__ hh,
                                                   int n_h, const double* __restrict__ xs,
                                                   const double* __restrict__ ys, const double* __restrict__ zs,
                                                   int64_t n, int64_t chunk, int n_chunks, double* __restrict__ part) {
  __shared__ double lds[4];
  const int gp = blockIdx.x / n_chunks, c = blockIdx.x % n_chunks;
  const double gR = rho[gp / n_h], gH = hh[gp % n_h];
  const int64_t k0 = (int64_t)c * chunk, k1 = min(n, k0 + chunk);
  double s = 0.0;
  for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const double dx = gR - xs[k], dz = gH - zs[k], y = ys[k];
    const double d = sqrt(dx * dx + y * y + dz * dz);
    s += exp(-d) / (d + 0.001);
  }
  s = block_sum256(s, lds);
  if (threadIdx.x == 0) part[(size_t)gp * n_chunks + c] = s;
}

__global__ __launch_bounds__(256) void k_trap_finish(const double* __restrict__ part, int n_gp, int n_chunks,
                                                     double n_mc, double* __restrict__ out) {
  const int gp = blockIdx.x * 256 + threadIdx.x;
  if (gp >= n_gp) return;
  double s = 0.0;
  for (int c = 0; c < n_chunks; ++c) s += part[(size_t)gp * n_chunks + c];
  out[gp] = s / n_mc;
}

// ---------------------------------------------------------------------------
// Cutoff mode: cell list built every step by integer binning and a stable,
// deterministic LSD radix sort (8-bit digits, two passes for 16-bit cell ids;
// within a digit, items keep their input order), then a pair kernel over the
// 5x5x5 cell stencil (cells of cutoff/2).
// ---------------------------------------------------------------------------
constexpr int RS_ITEMS = 1024;  // items per radix-sort block (256 threads x 4 rounds)

__device__ __forceinline__ int cell_of(const DevParams& P, double x, double y, double z) {
  int ix = (int)floor((x + P.R) / P.cell_h);
  int iy = (int)floor((y + P.R) / P.cell_h);
  int iz = (int)floor((z + P.half_H) / P.cell_h);
  ix = min(max(ix, 0), P.nx - 1);
  iy = min(max(iy, 0), P.ny - 1);
  iz = min(max(iz, 0), P.nz - 1);
  return ix + P.nx * (iy + P.ny * iz);
}

// key = cell of the ion (FP64 position), or n_cells for ions left out (flagged,
// or padding); value = ion index
__global__ __launch_bounds__(256) void k_bin(const DevParams* __restrict__ P, const double* __restrict__ X,
                                             const uint32_t* __restrict__ flags, uint32_t* __restrict__ key,
                                             uint32_t* __restrict__ val) {
  const int np = P->n_pad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  uint32_t k = (uint32_t)P->n_cells;
  if (i < P->n && (flags == nullptr || flags[i] == 0u)) k = (uint32_t)cell_of(*P, X[i], X[np + i], X[2 * np + i]);
  key[i] = k;
  val[i] = (uint32_t)i;
}

// per-block digit histogram, digit-major: hist[d * nblk + b]
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t* __restrict__ key, int n, int shift, int nblk,
                                                 int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_ITEMS;
  for (int r = 0; r < RS_ITEMS / 256; ++r) {
    const int i = base + r * 256 + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// stable scatter: rank = (digit base of this block) + (same-digit items earlier in the block)
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ key_in,
                                                    const uint32_t* __restrict__ val_in, int n, int shift, int nblk,
                                                    const int* __restrict__ off, uint32_t* __restrict__ key_out,
                                                    uint32_t* __restrict__ val_out) {
  __shared__ int run[256];
  __shared__ int wc[4][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  run[t] = off[t * nblk + blockIdx.x];
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int base = blockIdx.x * RS_ITEMS;
  for (int r = 0; r < RS_ITEMS / 256; ++r) {
    for (int w = 0; w < 4; ++w) wc[w][t] = 0;
    __syncthreads();
    const int i = base + r * 256 + t;
    const bool valid = i < n;
    const uint32_t k = valid ? key_in[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long m = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    if (valid && rank == 0) wc[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
      key_out[pos] = k;
      val_out[pos] = val_in[i];
    }
    __syncthreads();
    run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
    __syncthreads();
  }
}

// cell_start[c] = first position of the sorted keys with key >= c, c <= n_cells
// (keys of ions left out are n_cells): position i fills the cells between
// the previous key and its own.
__global__ __launch_bounds__(256) void k_cell_starts(const uint32_t* __restrict__ skey, int n, int n_cells,
                                                     int* __restrict__ cell_start) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > n) return;
  const int prev = (i == 0) ? -1 : min((int)skey[i - 1], n_cells);
  const int cur = (i == n) ? n_cells : min((int)skey[i], n_cells);
  for (int cc = prev + 1; cc <= cur; ++cc) cell_start[cc] = i;
}

__global__ __launch_bounds__(256) void k_cell_chunks(const int* __restrict__ cell_start, int n_cells,
                                                     int* __restrict__ chunks) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < n_cells) chunks[c] = (cell_start[c + 1] - cell_start[c] + 63) / 64;
}

// work items (cell, 64-target chunk), packed as cell | chunk << 16
__global__ __launch_bounds__(256) void k_cell_items(const int* __restrict__ chunks, const int* __restrict__ item_off,
                                                    int n_cells, int* __restrict__ items) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n_cells) return;
  const int o = item_off[c];
  for (int k = 0; k < chunks[c]; ++k) items[o + k] = c | (k << 16);
}

// sorted source positions (kept ions) and their ion ids
template <class T>
__global__ __launch_bounds__(256) void k_cell_gather(const uint32_t* __restrict__ skey,
                                                     const uint32_t* __restrict__ sval, int n_pad, int n_cells,
                                                     const typename Pos<T>::V* __restrict__ pos,
                                                     typename Pos<T>::V* __restrict__ spos) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_pad) return;
  spos[p] = ((int)skey[p] < n_cells) ? pos[sval[p]] : Pos<T>::park();
}

// Truncated pair law with its parameters held in registers (SW: OpenMM switch).
template <class T, int SW>
struct CutLaw;
template <int SW>
struct CutLaw<float, SW> {
  float a, linv, rc, rs, invw;
  __device__ explicit CutLaw(const DevParams& P)
      : a(P.a32), linv(P.linv32), rc(P.rc32), rs(P.rs32), invw(P.invw32) {}
  __device__ __forceinline__ void operator()(const float4& i, const float4& j, Acc<float>& ax, Acc<float>& ay,
                                             Acc<float>& az) const {
    pair32_cut<SW != 0>(i.x, i.y, i.z, j.x, j.y, j.z, a, linv, rc, rs, invw, ax.a, ay.a, az.a);
  }
};
template <int SW>
struct CutLaw<double, SW> {
  double rho0, rc, rs, invw;
  __device__ explicit CutLaw(const DevParams& P) : rho0(P.rho0), rc(P.rc64), rs(P.rs64), invw(P.invw64) {}
  __device__ __forceinline__ void operator()(const double4& i, const double4& j, Acc<double>& ax,
                                             Acc<double>& ay, Acc<double>& az) const {
    pair64_cut(i.x, i.y, i.z, j.x, j.y, j.z, rho0, SW != 0, rc, rs, invw, ax.a, ay.a, az.a);
  }
};

// Forces on one wave's 64 targets (one cell, one chunk) from the 5x5x5 stencil.
// The item index is made wave-uniform explicitly, so cell bounds and source
// positions arrive through scalar loads and feed the VALU as SGPR operands;
// the law's parameters live in registers; pairs beyond the cutoff are masked.
template <class T, int SW>
__global__ __launch_bounds__(256) void k_pair_cell(const DevParams* __restrict__ P,
                                                   const typename Pos<T>::V* __restrict__ spos,
                                                   const uint32_t* __restrict__ sval,
                                                   const int* __restrict__ cell_start, const int* __restrict__ items,
                                                   const int* __restrict__ n_items, double* __restrict__ F) {
  const CutLaw<T, SW> law(*P);
  const int lane = threadIdx.x & 63;
  const int np = P->n_pad;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const double scale = P->pair_scale;
  const int nit = n_items[0];
  const int nw = gridDim.x * 4;
  for (int it = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6)); it < nit; it += nw) {
    const int packed = items[it];
    const int c = packed & 0xFFFF, k = packed >> 16;
    const int t0 = cell_start[c] + k * 64;
    const int tp = t0 + lane;
    const bool valid = tp < cell_start[c + 1];
    const typename Pos<T>::V own = valid ? spos[tp] : Pos<T>::park();
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    Acc<T> ax, ay, az;
    for (int z = max(cz - 2, 0); z <= min(cz + 2, nz - 1); ++z) {
      for (int y = max(cy - 2, 0); y <= min(cy + 2, ny - 1); ++y) {
        const int row = nx * (y + ny * z);
        const int j0 = cell_start[row + max(cx - 2, 0)];
        const int j1 = cell_start[row + min(cx + 2, nx - 1) + 1];
        for (int jt = j0; jt < j1; jt += TJ) {
          const int je = min(jt + TJ, j1);
          int j = jt;
          for (; j + 4 <= je; j += 4) {
            const typename Pos<T>::V p0 = spos[j], p1 = spos[j + 1], p2 = spos[j + 2], p3 = spos[j + 3];
            law(own, p0, ax, ay, az);
            law(own, p1, ax, ay, az);
            law(own, p2, ax, ay, az);
            law(own, p3, ax, ay, az);
          }
          for (; j < je; ++j) law(own, spos[j], ax, ay, az);
          ax.flush();
          ay.flush();
          az.flush();
        }
      }
    }
    if (valid) {
      const int i = (int)sval[tp];
      F[i] = ax.total() * scale;
      F[np + i] = ay.total() * scale;
      F[2 * np + i] = az.total() * scale;
    }
  }
}

// ---------------------------------------------------------------------------
// Symmetric cell-list force (cutoff mode, FP32c): every unordered pair within
// the stencil once.  Work-group = one wave = (target cell c, forward set k):
// k < 12 are the stencil rows (dy, dz) with dz > 0 or (dz == 0, dy > 0), each
// 5 cells wide in x (contiguous in cell order); k = 12 is cells c+1, c+2 of
// c's own row.  The cell's targets are taken 256 at a time (4 sub-blocks of
// 64 in registers), the set's sources 64 at a time, each 64 x 64 tile by the
// rotation scheme of the all-pairs kernel with the truncated law.  Partials:
//   target  tp[k][3][n_pad]  (FP64, sorted position; k = 13: own-cell pairs)
//   source  sp[slot][3][n_pad] (FP32), slot = 5 k + (source cell x - cx + 2)
//           for k < 12, 60 + (source cell x - cx - 1) for k = 12,
// so no two work-groups write the same entry; a work-group adds its later
// target groups' reactions to its own earlier ones in order.  The pairs
// inside one cell are done ordered (k_symcell_self), and k_symcell_reduce
// adds every ion's entries in a fixed order.
// ---------------------------------------------------------------------------
constexpr int SC_MT = 4;            // target sub-blocks per wave
constexpr int SC_TG = 64 * SC_MT;   // targets per group
constexpr int SC_NK = 13;           // forward source sets per cell
constexpr int SC_NS = 62;           // source partial slots

// forward stencil row k < 12 -> (dy, dz)
__device__ __forceinline__ void sc_row(int k, int& dy, int& dz) {
  if (k < 2) {
    dy = k + 1;
    dz = 0;
  } else {
    dz = 1 + (k - 2) / 5;
    dy = (k - 2) % 5 - 2;
  }
}

// source range [s0, s1) of set k of cell (cx, cy, cz); empty if outside the grid
__device__ __forceinline__ void sc_sources(const int* __restrict__ cell_start, int nx, int ny, int nz, int cx,
                                           int cy, int cz, int k, int& s0, int& s1) {
  s0 = s1 = 0;
  if (k < 12) {
    int dy, dz;
    sc_row(k, dy, dz);
    const int ry = cy + dy, rz = cz + dz;
    if (ry < 0 || ry >= ny || rz >= nz) return;
    const int row = nx * (ry + ny * rz);
    s0 = cell_start[row + max(cx - 2, 0)];
    s1 = cell_start[row + min(cx + 2, nx - 1) + 1];
  } else {
    if (cx + 1 >= nx) return;
    const int c = cx + nx * (cy + ny * cz);
    s0 = cell_start[c + 1];
    s1 = cell_start[c + min(2, nx - 1 - cx) + 1];
  }
}

template <int CUT>
__global__ __launch_bounds__(64) void k_pair32_symcell(const DevParams* __restrict__ P,
                                                       const float4* __restrict__ spos,
                                                       const uint32_t* __restrict__ skey,
                                                       const int* __restrict__ cell_start,
                                                       double* __restrict__ tp, float* __restrict__ sp) {
  const int nx = P->nx, ny = P->ny, nz = P->nz, np = P->n_pad;
  const int c = blockIdx.x / SC_NK, k = blockIdx.x % SC_NK;
  const int lane = threadIdx.x;
  const int t0 = cell_start[c], t1 = cell_start[c + 1];
  if (t0 == t1) return;
  const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
  int s0, s1;
  sc_sources(cell_start, nx, ny, nz, cx, cy, cz, k, s0, s1);
  const SymLaw L{P->a32, P->linv32, P->rc32, P->rs32, P->invw32};
  const float4 park = make_float4(PARK32, PARK32, PARK32, 0.f);
  for (int g0 = t0; g0 < t1; g0 += SC_TG) {
    SymT<SC_MT> T;
    float sx[SC_MT], sy[SC_MT], sz[SC_MT], ccx[SC_MT], ccy[SC_MT], ccz[SC_MT];
#pragma unroll
    for (int m = 0; m < SC_MT; ++m) {
      const int t = g0 + m * 64 + lane;
      const float4 q = (t < t1) ? spos[t] : park;
      T.x[m] = q.x;
      T.y[m] = q.y;
      T.z[m] = q.z;
      sx[m] = sy[m] = sz[m] = ccx[m] = ccy[m] = ccz[m] = 0.f;
    }
    for (int b0 = s0; b0 < s1; b0 += 64) {
      const int j = b0 + lane;
      const bool jv = j < s1;
      const float4 q = jv ? spos[j] : park;
      float bx = 0.f, by = 0.f, bz = 0.f;
#pragma unroll
      for (int m = 0; m < SC_MT; ++m) T.ax[m] = T.ay[m] = T.az[m] = 0.f;
      sym_tile<SC_MT, CUT>(T, q.x, q.y, q.z, bx, by, bz, L);
#pragma unroll
      for (int m = 0; m < SC_MT; ++m) {
        kahan_add(sx[m], ccx[m], T.ax[m]);
        kahan_add(sy[m], ccy[m], T.ay[m]);
        kahan_add(sz[m], ccz[m], T.az[m]);
      }
      if (jv) {
        const int jx = (int)(skey[j] % (uint32_t)nx);
        const int slot = (k < 12) ? 5 * k + (jx - cx + 2) : 60 + (jx - cx - 1);
        float* o = sp + (size_t)slot * 3 * np + j;
        if (g0 == t0) {
          o[0] = bx;
          o[np] = by;
          o[2 * (size_t)np] = bz;
        } else {
          o[0] += bx;
          o[np] += by;
          o[2 * (size_t)np] += bz;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < SC_MT; ++m) {
      const int t = g0 + m * 64 + lane;
      if (t < t1) {
        double* o = tp + (size_t)k * 3 * np + t;
        o[0] = (double)sx[m] - (double)ccx[m];
        o[np] = (double)sy[m] - (double)ccy[m];
        o[2 * (size_t)np] = (double)sz[m] - (double)ccz[m];
      }
    }
  }
}

// pairs inside each cell, ordered (each target against every source of its
// cell), into target slot 13; sources are work-group-uniform
template <int SW>
__global__ __launch_bounds__(256) void k_symcell_self(const DevParams* __restrict__ P, const float4* __restrict__ spos,
                                                      const int* __restrict__ cell_start, double* __restrict__ tp) {
  const CutLaw<float, SW> law(*P);
  const int np = P->n_pad;
  const int c = blockIdx.x;
  const int t0 = cell_start[c], t1 = cell_start[c + 1];
  for (int t = t0 + (int)threadIdx.x; t - (int)threadIdx.x < t1; t += 256) {
    const bool tv = t < t1;
    const float4 own = tv ? spos[t] : spos[t0];
    Acc<float> ax, ay, az;
    for (int jt = t0; jt < t1; jt += TJ) {
      const int je = min(jt + TJ, t1);
      for (int j = jt; j < je; ++j) law(own, spos[j], ax, ay, az);
      ax.flush();
      ay.flush();
      az.flush();
    }
    if (tv) {
      double* o = tp + (size_t)13 * 3 * np + t;
      o[0] = ax.total();
      o[np] = ay.total();
      o[2 * (size_t)np] = az.total();
    }
  }
}

// F[ion] = pair_scale * (its 14 target partials + the source partials of every
// work-group whose set holds it), fixed order
__global__ __launch_bounds__(256) void k_symcell_reduce(const DevParams* __restrict__ P,
                                                        const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        const int* __restrict__ cell_start,
                                                        const double* __restrict__ tp, const float* __restrict__ sp,
                                                        double* __restrict__ F) {
  const int nx = P->nx, ny = P->ny, nz = P->nz, np = P->n_pad;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= cell_start[P->n_cells]) return;
  const int c = (int)skey[p];
  const int jx = c % nx, jy = (c / nx) % ny, jz = c / (nx * ny);
  double s[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = 0.0;
    // every work-group (c, k) of a non-empty cell wrote its targets' entries,
    // zero when its set is empty; slot 13 holds the own-cell pairs
    for (int k = 0; k < 14; ++k) v += tp[((size_t)k * 3 + q) * np + p];
    s[q] = v;
  }
  // reactions: work-groups (c', k) whose set k holds this ion's cell
  for (int k = 0; k < SC_NK; ++k) {
    int dy = 0, dz = 0, dlo, dhi;
    if (k < 12) {
      sc_row(k, dy, dz);
      dlo = -2;
      dhi = 2;
    } else {
      dlo = 1;
      dhi = 2;
    }
    const int cy = jy - dy, cz = jz - dz;
    if (cy < 0 || cy >= ny || cz < 0 || cz >= nz) continue;
    for (int d = dlo; d <= dhi; ++d) {
      const int cx = jx - d;
      if (cx < 0 || cx >= nx) continue;
      const int cc = cx + nx * (cy + ny * cz);
      if (cell_start[cc] == cell_start[cc + 1]) continue;
      const int slot = (k < 12) ? 5 * k + (d + 2) : 60 + (d - 1);
#pragma unroll
      for (int q = 0; q < 3; ++q) s[q] += (double)sp[((size_t)slot * 3 + q) * np + p];
    }
  }
  const int i = (int)sval[p];
  const double sc = P->pair_scale;
  F[i] = s[0] * sc;
  F[np + i] = s[1] * sc;
  F[2 * (size_t)np + i] = s[2] * sc;
}

// Rows of listed ions in cutoff mode: kept sources from the cell stencil of the
// ion's cell (FP64 position X) plus the other listed ions `others`.
template <class T>
__global__ __launch_bounds__(256) void k_rows_cell(const DevParams* __restrict__ P, const double* __restrict__ X,
                                                   const typename Pos<T>::V* __restrict__ spos,
                                                   const int* __restrict__ cell_start, const int* __restrict__ n_flag,
                                                   const typename Pos<T>::V* __restrict__ own_list,
                                                   const typename Pos<T>::V* __restrict__ others,
                                                   const uint32_t* __restrict__ flags, const int* __restrict__ list_ion,
                                                   uint32_t need_mask, double* __restrict__ rowF) {
  __shared__ double lds[4];
  const int nf = n_flag[0];
  const int np = P->n_pad;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  for (int s = blockIdx.x; s < nf; s += gridDim.x) {
    const int ion = list_ion[s];
    if (need_mask != 0u && (flags[ion] & need_mask) == 0u) continue;
    const typename Pos<T>::V own = own_list[s];
    const int c = cell_of(*P, X[ion], X[np + ion], X[2 * np + ion]);
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    Acc<T> ax, ay, az;
    for (int z = max(cz - 2, 0); z <= min(cz + 2, nz - 1); ++z) {
      for (int y = max(cy - 2, 0); y <= min(cy + 2, ny - 1); ++y) {
        const int row = nx * (y + ny * z);
        const int j0 = cell_start[row + max(cx - 2, 0)];
        const int j1 = cell_start[row + min(cx + 2, nx - 1) + 1];
        for (int j = j0 + (int)threadIdx.x; j < j1; j += 256) pair_acc<T>(*P, own, spos[j], ax, ay, az);
        ax.flush();
        ay.flush();
        az.flush();
      }
    }
    for (int j = threadIdx.x; j < nf; j += 256) pair_acc<T>(*P, own, others[j], ax, ay, az);
    ax.flush();
    ay.flush();
    az.flush();
    const double fx = block_sum256(ax.total(), lds);
    const double fy = block_sum256(ay.total(), lds);
    const double fz = block_sum256(az.total(), lds);
    if (threadIdx.x == 0) {
      rowF[3 * s + 0] = fx * P->pair_scale;
      rowF[3 * s + 1] = fy * P->pair_scale;
      rowF[3 * s + 2] = fz * P->pair_scale;
    }
  }
}

// Share of the listed (flagged) sources in the force on the binned (kept)
// ions, cutoff mode.  Targets run in cell order, so a wave's 64 targets are
// close together: a listed source is evaluated only if it lies within the
// cutoff (plus a relative margin of 1e-5) of the wave's bounding box, decided
// 64 sources at a time by a ballot; skipped sources are beyond the cutoff of
// every target, where the law gives exactly zero.
// out[ion] = (add ? out[ion] : 0) + sum * pair_scale.
template <class T, int SW>
__global__ __launch_bounds__(256) void k_corr_cell(const DevParams* __restrict__ P,
                                                   const typename Pos<T>::V* __restrict__ spos,
                                                   const uint32_t* __restrict__ sval,
                                                   const int* __restrict__ cell_start, const int* __restrict__ n_flag,
                                                   const typename Pos<T>::V* __restrict__ list, int add,
                                                   double* __restrict__ out) {
  using PV = typename Pos<T>::V;
  const CutLaw<T, SW> law(*P);
  const int np = P->n_pad;
  const int nbin = cell_start[P->n_cells];
  const int wave0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 256 + (threadIdx.x & ~63)));
  if (wave0 >= nbin) return;
  const int lane = threadIdx.x & 63;
  const int p = wave0 + lane;
  const bool valid = p < nbin;
  const PV own = valid ? spos[p] : spos[wave0];
  T lox = own.x, loy = own.y, loz = own.z, hix = own.x, hiy = own.y, hiz = own.z;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lox = min(lox, __shfl_xor(lox, o, 64));
    loy = min(loy, __shfl_xor(loy, o, 64));
    loz = min(loz, __shfl_xor(loz, o, 64));
    hix = max(hix, __shfl_xor(hix, o, 64));
    hiy = max(hiy, __shfl_xor(hiy, o, 64));
    hiz = max(hiz, __shfl_xor(hiz, o, 64));
  }
  const T rcm = (sizeof(T) == 4 ? (T)P->rc32 : (T)P->rc64) * (T)1.00001;
  const T rcm2 = rcm * rcm;
  Acc<T> ax, ay, az;
  int cnt = 0;
  const int nf = n_flag[0];
  for (int s0 = 0; s0 < nf; s0 += 64) {
    const int s = s0 + lane;
    const PV q = (s < nf) ? list[s] : Pos<T>::park();
    const T dx = max(max(lox - q.x, q.x - hix), (T)0);
    const T dy = max(max(loy - q.y, q.y - hiy), (T)0);
    const T dz = max(max(loz - q.z, q.z - hiz), (T)0);
    uint64_t m = __ballot(s < nf && dx * dx + dy * dy + dz * dz <= rcm2);
    while (m) {
      const int l = __builtin_ctzll(m);
      m &= m - 1;
      PV qs;
      qs.x = __shfl(q.x, l, 64);
      qs.y = __shfl(q.y, l, 64);
      qs.z = __shfl(q.z, l, 64);
      qs.w = 0;
      law(own, qs, ax, ay, az);
      if (++cnt == TJ) {
        ax.flush();
        ay.flush();
        az.flush();
        cnt = 0;
      }
    }
  }
  ax.flush();
  ay.flush();
  az.flush();
  if (valid) {
    const int i = (int)sval[p];
    const double sc = P->pair_scale;
    out[i] = (add ? out[i] : 0.0) + ax.total() * sc;
    out[np + i] = (add ? out[np + i] : 0.0) + ay.total() * sc;
    out[2 * np + i] = (add ? out[2 * np + i] : 0.0) + az.total() * sc;
  }
}

// in-range and candidate pair counts over the current cell list (the same
// distance test as k_pair_cell); integer atomics, so the totals are exact
template <class T>
__global__ __launch_bounds__(256) void k_cell_pair_count(const DevParams* __restrict__ P,
                                                         const typename Pos<T>::V* __restrict__ spos,
                                                         const int* __restrict__ cell_start,
                                                         const int* __restrict__ items, const int* __restrict__ n_items,
                                                         unsigned long long* __restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const int nit = n_items[0];
  const int nw = gridDim.x * 4;
  unsigned long long in_range = 0, cand = 0;
  for (int it = blockIdx.x * 4 + (threadIdx.x >> 6); it < nit; it += nw) {
    const int packed = items[it];
    const int c = packed & 0xFFFF, k = packed >> 16;
    const int tp = cell_start[c] + k * 64 + lane;
    const bool valid = tp < cell_start[c + 1];
    const typename Pos<T>::V own = valid ? spos[tp] : Pos<T>::park();
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    for (int z = max(cz - 2, 0); z <= min(cz + 2, nz - 1); ++z)
      for (int y = max(cy - 2, 0); y <= min(cy + 2, ny - 1); ++y) {
        const int row = nx * (y + ny * z);
        const int j0 = cell_start[row + max(cx - 2, 0)];
        const int j1 = cell_start[row + min(cx + 2, nx - 1) + 1];
        for (int j = j0; j < j1; ++j) {
          if (!valid || j == tp) continue;
          const typename Pos<T>::V q = spos[j];
          bool in;
          if constexpr (sizeof(T) == 4) {
            const float dx = q.x - own.x, dy = q.y - own.y, dz = q.z - own.z;
            const float r = __builtin_amdgcn_sqrtf(
                __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, EPS32))));
            in = r < P->rc32;
          } else {
            const double dx = q.x - own.x, dy = q.y - own.y, dz = q.z - own.z;
            in = sqrt(fma(dx, dx, fma(dy, dy, fma(dz, dz, EPS64)))) < P->rc64;
          }
          in_range += in ? 1 : 0;
          cand += 1;
        }
      }
  }
  atomicAdd(&counts[0], in_range);
  atomicAdd(&counts[1], cand);
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct od_ctx {
  od_config cfg;
  int dev = 0;
  int prec = OD_PREC_FP32C;
  hipStream_t stream = nullptr;
  int n = 0, n_pad = 0, nb = 0, n_chunks = 1, chunk_len = 0, ib = 2, rows_grid = 1;
  size_t pos_bytes = 16;
  DevParams* dP = nullptr;
  double *X = nullptr, *V = nullptr, *F = nullptr, *Fpart = nullptr, *q = nullptr, *partials = nullptr,
         *rowF = nullptr;
  void *pi = nullptr, *pj = nullptr, *list_old = nullptr, *list_new = nullptr;
  uint32_t *flags = nullptr, *st = nullptr;
  uint8_t* face = nullptr;
  int *block_cnt = nullptr, *block_off = nullptr, *n_flag = nullptr, *slot_of = nullptr, *list_ion = nullptr,
      *err = nullptr;
  od_dust_record* rec = nullptr;
  int rec_cap = 1024;
  int64_t host_step = 0;
  bool have_ions = false;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_steps = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double pair_ms = 0.0;
  int64_t pair_launches = 0;
  // symmetric pair kernel
  int sym = 0, n_blocks = 0, sym_ch = 0, n_wg = 0, sym_var = 0;
  int2* wg_tab = nullptr;
  int* wg_base = nullptr;
  double *sym_tp = nullptr, *sym_dg = nullptr, *rowD = nullptr;
  float* sym_sp = nullptr;
  // cutoff mode: cell list; Gcut: flagged sources' share of the drift-state force;
  // sc_tp, sc_sp: partials of the symmetric cell kernel (FP32c)
  double* Gcut = nullptr;
  double* sc_tp = nullptr;
  float* sc_sp = nullptr;
  int *scan_tsum = nullptr, *scan_toff = nullptr;
  int cut = 0, n_cells = 0, rs_nblk = 0, max_items = 0;
  uint32_t *ckey[2] = {nullptr, nullptr}, *cval[2] = {nullptr, nullptr};
  int *rs_hist = nullptr, *rs_off = nullptr, *rs_total = nullptr, *cell_start = nullptr,
      *cell_chunks = nullptr, *item_off = nullptr, *items = nullptr, *n_items = nullptr;
  void* spos = nullptr;
};

static int fill_params(const od_config* c, int n_pad, DevParams* P) {
  std::memset(P, 0, sizeof(*P));
  const double L = 1.4426950408889634;
  P->n = c->n_ions;
  P->n_pad = n_pad;
  P->n_dust = c->n_dust;
  P->distribution = c->distribution;
  P->use_bottom = c->use_bottom;
  P->dust_mask = (c->n_dust > 0) ? ((1u << c->n_dust) - 1u) : 0u;
  P->ion_off = 0;
  P->stride = n_pad;
  P->blk0 = 0;
  P->blk1 = n_pad / 256;
  P->q_off = 0;
  P->sim_id = (uint32_t)c->sim_id;
  P->seed = c->seed;
  P->R = c->R;
  P->H = c->H;
  P->half_H = c->H / 2.0;
  P->dt = c->dt;
  P->half_dt = 0.5 * c->dt;
  P->m_s = c->m_s;
  P->v_T = c->v_T;
  P->v_fl = c->v_fl;
  P->lam = c->r_D_e;
  P->rho0 = c->r_0 / c->r_D_e;
  const double C = c->kqq / (c->r_D_e * c->r_D_e);  // kJ/mol/nm
  if (c->precision == OD_PREC_FP32C) {
    P->pos_scale = L / c->r_D_e;
    P->pair_scale = -C * L * L;
  } else {
    P->pos_scale = 1.0 / c->r_D_e;
    P->pair_scale = -C;
  }
  if (c->cutoff > 0.0) {
    P->use_cut = 1;
    P->use_switch = (c->switch_dist > 0.0 && c->switch_dist < c->cutoff) ? 1 : 0;
    const double rc = c->cutoff / c->r_D_e, rs = P->use_switch ? c->switch_dist / c->r_D_e : rc;
    P->rc64 = rc;
    P->rs64 = rs;
    P->invw64 = P->use_switch ? 1.0 / (rc - rs) : 0.0;
    P->rc32 = (float)(rc * L);
    P->rs32 = (float)(rs * L);
    P->invw32 = P->use_switch ? (float)(1.0 / ((rc - rs) * L)) : 0.0f;
    P->cell_h = c->cutoff / 2.0;
    P->nx = (int)ceil(2.0 * c->R / P->cell_h);
    P->ny = P->nx;
    P->nz = (int)ceil(c->H / P->cell_h);
    P->n_cells = P->nx * P->ny * P->nz;
  }
  P->a32 = (float)(P->rho0 * L);
  P->linv32 = (float)(1.0 / L);
  P->thr_top = c->thr_top;
  P->thr_side = c->use_bottom ? c->thr_side : 4294967296ull;
  P->thr_coll = c->thr_coll;
  ExtParams& E = P->ext;
  E.n_dust = c->n_dust;
  E.field_driven = c->distribution == OD_DIST_FIELD_DRIVEN;
  E.inv_lam = 1.0 / c->r_D_e;
  E.trap_scale = -c->trap_C / c->r_D_e;
  for (int m = 0; m < 15; ++m) E.alpha[m] = c->trap_alpha[m];
  E.k_dust = c->k_dust;
  E.field_force = c->field_force;
  for (int p = 0; p < c->n_dust; ++p) {
    E.dx[p] = c->dust_x[p];
    E.dy[p] = c->dust_y[p];
    E.dz[p] = c->dust_z[p];
    E.dr[p] = c->dust_r[p];
    P->dust_method[p] = c->dust_method[p];
    P->si_dx[p] = c->si_dust_x[p];
    P->si_dy[p] = c->si_dust_y[p];
    P->si_dz[p] = c->si_dust_z[p];
    P->si_dr[p] = c->si_dust_r[p];
  }
  P->si_k = c->si_k;
  P->si_e = c->si_e;
  P->si_q_s = c->si_q_s;
  P->si_lam = c->si_r_D_e;
  P->si_dt = c->si_dt;
  P->si_kB = c->si_kB;
  P->si_Te = c->si_Te;
  P->si_Tn = c->si_Tn;
  P->si_E = c->si_E;
  P->gamma_e_pref = c->gamma_e_pref;
  P->oml_A = c->oml_A;
  P->oml_B = c->oml_B;
  P->oml_u2 = c->oml_u2;
  P->oml_erf_u = c->oml_erf_u;
  P->oml_tail = c->oml_tail;
  P->collect_scale = c->collect_scale;
  P->top = c->top;
  P->bottom = c->bottom;
  P->side = c->side;
  P->vz = c->vz;
  return OD_OK;
}

static void free_ctx(od_ctx* c) {
  if (!c) return;
  hipSetDevice(c->dev);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  void* bufs[] = {c->dP, c->X, c->V, c->F, c->Fpart, c->q, c->partials, c->rowF, c->pi, c->pj,
                  c->list_old, c->list_new, c->flags, c->st, c->face, c->block_cnt, c->block_off, c->n_flag,
                  c->slot_of, c->list_ion, c->err, c->rec, c->wg_tab, c->wg_base, c->sym_tp, c->sym_dg,
                  c->rowD, c->sym_sp, c->Gcut, c->scan_tsum, c->scan_toff, c->sc_tp, c->sc_sp, c->ckey[0], c->ckey[1], c->cval[0], c->cval[1], c->rs_hist,
                  c->rs_off, c->rs_total, c->cell_start, c->cell_chunks, c->item_off, c->items,
                  c->n_items, c->spos};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

template <class P>
static int dalloc(P** p, size_t bytes) {
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return OD_OK;
}

// launch geometry of the pair kernel
static void pair_geometry(int n_pad, int ib, int j_split, int* n_chunks, int* chunk_len) {
  const int n_iblk = n_pad / (256 * ib);
  int c = j_split;
  if (c <= 0) {
    // enough blocks to fill 256 CUs several times, chunks of >= 1024 sources
    c = 1;
    while (n_iblk * c < 4096 && c < 64 && n_pad / (2 * c) >= 1024) c *= 2;
  }
  c = std::max(1, std::min(c, n_pad / TJ));
  int len = (n_pad + c - 1) / c;
  len = (len + TJ - 1) / TJ * TJ;
  *n_chunks = c;
  *chunk_len = len;
}

// Work-groups of the triangular schedule: target block X, source blocks
// X+1+k*ch .. X+(k+1)*ch clipped at be[X] (the first block past X's
// simulation), k = 0, 1, ...
static int64_t sym_wg_count(const std::vector<int>& be, int ch) {
  int64_t n = 0;
  for (int X = 0; X < (int)be.size(); ++X) n += (be[X] - 1 - X + ch - 1) / ch;
  return n;
}

static double sym_bytes(const std::vector<int>& be, int ch) {
  return (double)sym_wg_count(be, ch) * 3 * SYM_TB * (4.0 * ch + 8.0);
}

// (Re)builds a context's schedule for `ch` source blocks per work-group:
// partial slots in row-major (X, k) order (wg_base), launch order longest
// work-group first (ties in row-major order), so the short work-groups that
// end each row fill the tail of the launch.  C: od_ctx or od_batch.
template <class C>
static int sym_build(C* c, const std::vector<int>& be, int ch) {
  const int NB = (int)be.size();
  const int64_t n_wg = sym_wg_count(be, ch);
  if (n_wg > (1ll << 31) - 1) return set_err(OD_ERR_ARG, "symmetric pair kernel: too many workgroups");
  std::vector<int2> tab;
  std::vector<int> base(NB + 1), len;
  tab.reserve(n_wg);
  len.reserve(n_wg);
  for (int X = 0; X < NB; ++X) {
    base[X] = (int)tab.size();
    const int nk = (be[X] - 1 - X + ch - 1) / ch;
    for (int k = 0; k < nk; ++k) {
      tab.push_back(make_int2(X, k));
      len.push_back(std::min(ch, be[X] - (X + 1 + k * ch)));
    }
  }
  base[NB] = (int)tab.size();
  std::vector<int> order(tab.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len[a] > len[b]; });
  std::vector<int2> launch(tab.size());
  for (size_t i = 0; i < order.size(); ++i) launch[i] = tab[order[i]];
  void* old[] = {c->wg_tab, c->wg_base, c->sym_tp, c->sym_sp};
  for (void* q : old)
    if (q) hipFree(q);
  c->wg_tab = nullptr;
  c->wg_base = nullptr;
  c->sym_tp = nullptr;
  c->sym_sp = nullptr;
  const size_t nw = std::max<size_t>(launch.size(), 1);
  int rc = OD_OK;
  if ((rc = dalloc(&c->wg_tab, nw * sizeof(int2))) || (rc = dalloc(&c->wg_base, base.size() * sizeof(int))) ||
      (rc = dalloc(&c->sym_tp, nw * 3 * SYM_TB * 8)) || (rc = dalloc(&c->sym_sp, nw * 3 * ch * SYM_TB * 4)))
    return rc;
  if (!launch.empty()) HIPCHK(hipMemcpy(c->wg_tab, launch.data(), launch.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->wg_base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
  c->n_blocks = NB;
  c->sym_ch = ch;
  c->n_wg = (int)n_wg;
  return OD_OK;
}

// default kernel variant: the configured one, else 8 target sub-blocks per
// wave (2 waves per work-group) when the schedule has at least 2048
// work-groups (4 waves per SIMD), else 4 (4 waves per work-group: at 2^15 ions
// the 496 work-groups then still give about 2 waves per SIMD).  Measured on
// MI355X at 2^15 ions: 0.33 ms per evaluation with 4, 0.64 ms with 8.
static int sym_default_variant(int64_t n_wg, int configured) {
  if (configured > 0) return configured - 1;
  return n_wg >= 2048 ? 0 : 2;
}

// default source blocks per work-group: the largest of 16, 8, .. that still
// gives >= 4096 work-groups (8 per XCD CU set), or the configured value
static int sym_default_chunk(const std::vector<int>& be, int configured) {
  if (configured > 0) return configured;
  int ch = 16;
  while (ch > 1 && sym_wg_count(be, ch) < 4096) ch /= 2;
  return ch;
}

static int setup_cells(od_ctx* c, const od_config* cfg, const DevParams& hp) {
  c->cut = 0;
  if (!(cfg->cutoff > 0.0)) return OD_OK;
  if ((double)hp.nx * hp.ny * hp.nz > 65535.0)
    return set_err(OD_ERR_ARG, "cutoff %g nm gives %d x %d x %d cells; at most 65535 are supported", cfg->cutoff,
                   hp.nx, hp.ny, hp.nz);
  c->cut = 1;
  c->n_cells = hp.n_cells;
  c->rs_nblk = (c->n_pad + RS_ITEMS - 1) / RS_ITEMS;
  c->max_items = c->n_pad / 64 + c->n_cells + 1;
  const size_t np = c->n_pad;
  int rc = OD_OK;
  if ((rc = dalloc(&c->ckey[0], np * 4)) || (rc = dalloc(&c->ckey[1], np * 4)) || (rc = dalloc(&c->cval[0], np * 4)) ||
      (rc = dalloc(&c->cval[1], np * 4)) || (rc = dalloc(&c->rs_hist, (size_t)256 * c->rs_nblk * 4)) ||
      (rc = dalloc(&c->rs_off, (size_t)256 * c->rs_nblk * 4)) || (rc = dalloc(&c->rs_total, 4)) ||
      (rc = dalloc(&c->cell_start, (size_t)(c->n_cells + 1) * 4)) ||
      (rc = dalloc(&c->cell_chunks, (size_t)(c->n_cells + 1) * 4)) ||
      (rc = dalloc(&c->item_off, (size_t)(c->n_cells + 1) * 4)) || (rc = dalloc(&c->items, (size_t)c->max_items * 4)) ||
      (rc = dalloc(&c->n_items, 4)) || (rc = dalloc(&c->spos, np * c->pos_bytes)) ||
      (rc = dalloc(&c->Gcut, 3 * np * 8)))
    return rc;
  const size_t n_tiles = (std::max<size_t>((size_t)256 * c->rs_nblk, (size_t)c->n_cells + 1) + 1023) / 1024;
  if ((rc = dalloc(&c->scan_tsum, n_tiles * 4)) || (rc = dalloc(&c->scan_toff, n_tiles * 4))) return rc;
  if (c->prec == OD_PREC_FP32C &&
      ((rc = dalloc(&c->sc_tp, (size_t)14 * 3 * np * 8)) || (rc = dalloc(&c->sc_sp, (size_t)SC_NS * 3 * np * 4))))
    return rc;
  return OD_OK;
}

static int setup_sym(od_ctx* c, const od_config* cfg) {
  c->sym = 0;
  if (c->prec != OD_PREC_FP32C || cfg->pair_mode == OD_PAIR_ROWS || cfg->cutoff > 0.0) return OD_OK;
  const int T = c->n_pad / SYM_TB;
  const bool forced = cfg->pair_mode == OD_PAIR_SYM;
  // automatic choice: symmetric from 2^15 ions (measured on MI355X, one step:
  // 2^15 0.47 ms symmetric vs 0.62 ms rows, 2^14 0.30 vs 0.22 ms)
  if (T < 2 || (!forced && c->n_pad < (1 << 15))) return OD_OK;
  const std::vector<int> be(T, T);
  const int ch = sym_default_chunk(be, cfg->sym_chunk);
  if (!forced && sym_bytes(be, ch) > 48e9) return OD_OK;
  int rc = OD_OK;
  if ((rc = sym_build(c, be, ch)) || (rc = dalloc(&c->sym_dg, (size_t)3 * c->n_pad * 8))) return rc;
  c->sym_var = sym_default_variant(c->n_wg, cfg->sym_tile);
  c->sym = 1;
  return OD_OK;
}

static int check_config(const od_config* c) {
  if (!c) return set_err(OD_ERR_ARG, "null config");
  if (c->abi_version != OD_ABI_VERSION)
    return set_err(OD_ERR_ARG, "ABI version %d, library expects %d", c->abi_version, OD_ABI_VERSION);
  if (c->struct_size != (int32_t)sizeof(od_config))
    return set_err(OD_ERR_ARG, "od_config layout mismatch: caller %d bytes, library %d bytes", c->struct_size,
                   (int)sizeof(od_config));
  if (c->n_ions < 1) return set_err(OD_ERR_ARG, "n_ions must be >= 1 (got %d)", c->n_ions);
  if (c->n_ions > (1 << 28)) return set_err(OD_ERR_ARG, "n_ions %d above the supported 2^28", c->n_ions);
  if (c->n_dust < 0 || c->n_dust > OD_MAX_DUST)
    return set_err(OD_ERR_ARG, "n_dust %d outside [0, %d]", c->n_dust, OD_MAX_DUST);
  if (c->precision != OD_PREC_FP32C && c->precision != OD_PREC_FP64)
    return set_err(OD_ERR_ARG, "unknown precision %d", c->precision);
  if (c->distribution != OD_DIST_MAXWELLIAN && c->distribution != OD_DIST_FIELD_DRIVEN)
    return set_err(OD_ERR_ARG, "unknown distribution %d", c->distribution);
  if (c->sym_tile < 0 || c->sym_tile > SYM_N_VARIANTS)
    return set_err(OD_ERR_ARG, "sym_tile %d outside [0, %d]", c->sym_tile, SYM_N_VARIANTS);
  if (c->sym_chunk < 0) return set_err(OD_ERR_ARG, "sym_chunk %d is negative", c->sym_chunk);
  const od_table* ts[4] = {&c->top, &c->bottom, &c->side, &c->vz};
  for (const od_table* t : ts)
    if (t->n_seg < 0 || t->n_seg > OD_MAX_TABLE) return set_err(OD_ERR_ARG, "table length %d invalid", t->n_seg);
  if (!(c->R > 0 && c->H > 0 && c->dt > 0 && c->r_D_e > 0 && c->m_s > 0))
    return set_err(OD_ERR_ARG, "R, H, dt, r_D_e, m_s must be positive");
  return OD_OK;
}

extern "C" int od_create(const od_config* cfg, od_ctx** out) {
  if (!out) return set_err(OD_ERR_ARG, "null output pointer");
  *out = nullptr;
  CHKRET(check_config(cfg));
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return set_err(OD_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (cfg->device < 0 || cfg->device >= ndev)
    return set_err(OD_ERR_ARG, "device %d outside [0, %d)", cfg->device, ndev);
  od_ctx* c = new od_ctx();
  c->cfg = *cfg;
  c->dev = cfg->device;
  c->prec = cfg->precision;
  c->n = cfg->n_ions;
  c->n_pad = (c->n + 1023) / 1024 * 1024;
  c->nb = c->n_pad / 256;
  c->ib = (c->prec == OD_PREC_FP32C) ? 2 : 1;
  pair_geometry(c->n_pad, c->ib, cfg->j_split, &c->n_chunks, &c->chunk_len);
  c->pos_bytes = (c->prec == OD_PREC_FP32C) ? sizeof(float4) : sizeof(double4);
  c->rows_grid = 1024;
  int rc = OD_OK;
  auto fail = [&](int r) {
    free_ctx(c);
    return r;
  };
  if (hipSetDevice(c->dev) != hipSuccess) return fail(set_err(OD_ERR_HIP, "hipSetDevice(%d) failed", c->dev));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "hipEventCreate failed"));
  const size_t np = c->n_pad;
  if ((rc = dalloc(&c->dP, sizeof(DevParams))) || (rc = dalloc(&c->X, 3 * np * 8)) ||
      (rc = dalloc(&c->V, 3 * np * 8)) || (rc = dalloc(&c->F, 3 * np * 8)) ||
      (rc = dalloc(&c->Fpart, (c->n_chunks > 1 ? c->n_chunks : 1) * 3 * np * 8)) ||
      (rc = dalloc(&c->q, OD_MAX_DUST * 8)) || (rc = dalloc(&c->partials, (size_t)c->nb * OD_MAX_DUST * 8 * 8)) ||
      (rc = dalloc(&c->rowF, 3 * np * 8)) || (rc = dalloc(&c->pi, np * c->pos_bytes)) ||
      (rc = dalloc(&c->pj, np * c->pos_bytes)) || (rc = dalloc(&c->list_old, np * c->pos_bytes)) ||
      (rc = dalloc(&c->list_new, np * c->pos_bytes)) || (rc = dalloc(&c->flags, np * 4)) ||
      (rc = dalloc(&c->st, 4 * 4)) || (rc = dalloc(&c->face, np)) || (rc = dalloc(&c->block_cnt, c->nb * 4)) ||
      (rc = dalloc(&c->block_off, c->nb * 4)) || (rc = dalloc(&c->n_flag, 4)) ||
      (rc = dalloc(&c->slot_of, np * 4)) || (rc = dalloc(&c->list_ion, np * 4)) || (rc = dalloc(&c->err, 4)) ||
      (rc = dalloc(&c->rec, (size_t)c->rec_cap * std::max(1, cfg->n_dust) * sizeof(od_dust_record))) ||
      (rc = dalloc(&c->rowD, 3 * np * 8)) || (rc = setup_sym(c, cfg)))
    return fail(rc);
  DevParams hp;
  fill_params(cfg, c->n_pad, &hp);
  if ((rc = setup_cells(c, cfg, hp))) return fail(rc);
  double q0[OD_MAX_DUST] = {0};
  for (int p = 0; p < cfg->n_dust; ++p) q0[p] = cfg->dust_q0[p];
  uint32_t st0[4] = {0, 0, 0, 0};
  if (hipMemcpy(c->dP, &hp, sizeof(hp), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->q, q0, sizeof(q0), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->st, st0, sizeof(st0), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->X, 0, 3 * np * 8) != hipSuccess || hipMemset(c->V, 0, 3 * np * 8) != hipSuccess ||
      hipMemset(c->F, 0, 3 * np * 8) != hipSuccess || hipMemset(c->err, 0, 4) != hipSuccess ||
      hipMemset(c->n_flag, 0, 4) != hipSuccess || hipMemset(c->flags, 0, np * 4) != hipSuccess ||
      hipMemset(c->face, 0, np) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "device initialisation failed"));
  *out = c;
  return OD_OK;
}

extern "C" int od_destroy(od_ctx* c) {
  free_ctx(c);
  return OD_OK;
}

// Pair force of sources pj on targets pi into F, optionally timed.  The
// symmetric kernel treats one array as both targets and sources: it is only
// used when pi == pj.
static int launch_pair(od_ctx* c, const void* pi, const void* pj, double* F, bool timed) {
  const double L = 1.4426950408889634;
  const double C = c->cfg.kqq / (c->cfg.r_D_e * c->cfg.r_D_e);
  if (c->sym && pi == pj) {
    const float a = (float)(c->cfg.r_0 / c->cfg.r_D_e * L), linv = (float)(1.0 / L);
    if (timed) HIPCHK(hipEventRecord(c->ev0, c->stream));
    const SymVariant& V = sym_variants[c->sym_var];
    hipLaunchKernelGGL(V.fn, dim3(c->n_wg), dim3(SYM_TB / V.mt), 0, c->stream, (const float4*)pj, c->n_blocks,
                       c->sym_ch, (const int2*)c->wg_tab, (const int*)c->wg_base, a, linv, (const int*)nullptr,
                       (const float*)nullptr, c->sym_tp, c->sym_sp);
    if (timed) HIPCHK(hipEventRecord(c->ev1, c->stream));
    hipLaunchKernelGGL(k_pair32_diag, dim3(c->n_blocks * (SYM_TB / 256)), dim3(256), 0, c->stream,
                       (const float4*)pj, c->n_pad, a, linv, (const float*)nullptr, c->sym_dg);
    hipLaunchKernelGGL(k_sym_reduce, dim3(c->n_pad / 256), dim3(256), 0, c->stream, c->sym_dg, c->sym_tp,
                       c->sym_sp, c->wg_base, c->n_blocks, c->sym_ch, c->n_pad, -C * L * L, (const int*)nullptr,
                       (const double*)nullptr, F);
    return OD_OK;
  }
  double* out = (c->n_chunks > 1) ? c->Fpart : F;
  const int n_iblk = c->n_pad / (256 * c->ib);
  const int grid = n_iblk * c->n_chunks;
  if (timed) HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (c->prec == OD_PREC_FP32C) {
    const float a = (float)(c->cfg.r_0 / c->cfg.r_D_e * L), linv = (float)(1.0 / L);
    hipLaunchKernelGGL(k_pair32<2>, dim3(grid), dim3(256), 0, c->stream, (const float4*)pi, (const float4*)pj,
                       c->n_pad, c->n_chunks, c->chunk_len, a, linv, -C * L * L, out);
  } else {
    hipLaunchKernelGGL(k_pair64<1>, dim3(grid), dim3(256), 0, c->stream, (const double4*)pi, (const double4*)pj,
                       c->n_pad, c->n_chunks, c->chunk_len, c->cfg.r_0 / c->cfg.r_D_e, -C, out);
  }
  if (timed) HIPCHK(hipEventRecord(c->ev1, c->stream));
  if (c->n_chunks > 1) {
    const int n3 = 3 * c->n_pad;
    hipLaunchKernelGGL(k_chunk_reduce, dim3((n3 + 255) / 256), dim3(256), 0, c->stream, c->Fpart, c->n_chunks, n3,
                       F);
  }
  return OD_OK;
}

// exclusive scan of n ints (in -> out, sum -> total) on the context's stream
static void launch_scan(od_ctx* c, const int* in, int n, int* out, int* total) {
  if (n <= 16384) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, in, n, out, total);
    return;
  }
  const int nt = (n + 1023) / 1024;
  hipLaunchKernelGGL(k_tile_sum, dim3(nt), dim3(256), 0, c->stream, in, n, c->scan_tsum);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, (const int*)c->scan_tsum, nt, c->scan_toff, total);
  hipLaunchKernelGGL(k_tile_scan, dim3(nt), dim3(256), 0, c->stream, in, n, (const int*)c->scan_toff, out);
}

// cell list of the ions at FP64 positions X (flagged ions left out when flags != null)
template <class T>
static int build_cells_t(od_ctx* c, const double* X, const uint32_t* flags, const void* pos) {
  using PV = typename Pos<T>::V;
  const int np = c->n_pad, nb = c->nb, nc = c->n_cells;
  hipLaunchKernelGGL(k_bin, dim3(nb), dim3(256), 0, c->stream, c->dP, X, flags, c->ckey[0], c->cval[0]);
  for (int pass = 0; pass < 2; ++pass) {
    const int in = pass, out = 1 - pass, shift = 8 * pass;
    hipLaunchKernelGGL(k_rs_hist, dim3(c->rs_nblk), dim3(256), 0, c->stream, c->ckey[in], np, shift, c->rs_nblk,
                       c->rs_hist);
    launch_scan(c, c->rs_hist, 256 * c->rs_nblk, c->rs_off, c->rs_total);
    hipLaunchKernelGGL(k_rs_scatter, dim3(c->rs_nblk), dim3(256), 0, c->stream, c->ckey[in], c->cval[in], np, shift,
                       c->rs_nblk, c->rs_off, c->ckey[out], c->cval[out]);
  }
  const int gc = (nc + 256) / 256;
  hipLaunchKernelGGL(k_cell_starts, dim3(np / 256 + 1), dim3(256), 0, c->stream, c->ckey[0], np, nc,
                     c->cell_start);
  hipLaunchKernelGGL(k_cell_chunks, dim3(gc), dim3(256), 0, c->stream, c->cell_start, nc, c->cell_chunks);
  launch_scan(c, c->cell_chunks, nc, c->item_off, c->n_items);
  hipLaunchKernelGGL(k_cell_items, dim3(gc), dim3(256), 0, c->stream, c->cell_chunks, c->item_off, nc, c->items);
  hipLaunchKernelGGL(k_cell_gather<T>, dim3(nb), dim3(256), 0, c->stream, c->ckey[0], c->cval[0], np, nc,
                     (const PV*)pos, (PV*)c->spos);
  return OD_OK;
}

// Ion-ion force on every ion from every other (kept) ion: the all-pairs kernels,
// or, in cutoff mode, the cell list + stencil kernel.  X: FP64 positions of the
// same state as pos; flags: ions to leave out (cutoff mode only; may be null).
template <class T>
static int launch_forces_t(od_ctx* c, const double* X, const void* pos, const uint32_t* flags, double* F,
                           bool timed) {
  using PV = typename Pos<T>::V;
  if (!c->cut) return launch_pair(c, pos, pos, F, timed);
  CHKRET(build_cells_t<T>(c, X, flags, pos));
  const int grid = std::max(1, std::min(4096, c->max_items / 4 + 1));
  const bool sw = c->cfg.switch_dist > 0.0 && c->cfg.switch_dist < c->cfg.cutoff;
  if (timed) HIPCHK(hipEventRecord(c->ev0, c->stream));
  if constexpr (std::is_same<T, float>::value) {
    // symmetric: stencil half-sets, own-cell pairs, per-ion reduction
    const float4* sp4 = (const float4*)c->spos;
    if (sw) {
      hipLaunchKernelGGL((k_pair32_symcell<2>), dim3(c->n_cells * SC_NK), dim3(64), 0, c->stream, c->dP, sp4,
                         c->ckey[0], c->cell_start, c->sc_tp, c->sc_sp);
      hipLaunchKernelGGL((k_symcell_self<1>), dim3(c->n_cells), dim3(256), 0, c->stream, c->dP, sp4, c->cell_start,
                         c->sc_tp);
    } else {
      hipLaunchKernelGGL((k_pair32_symcell<1>), dim3(c->n_cells * SC_NK), dim3(64), 0, c->stream, c->dP, sp4,
                         c->ckey[0], c->cell_start, c->sc_tp, c->sc_sp);
      hipLaunchKernelGGL((k_symcell_self<0>), dim3(c->n_cells), dim3(256), 0, c->stream, c->dP, sp4, c->cell_start,
                         c->sc_tp);
    }
    hipLaunchKernelGGL(k_symcell_reduce, dim3(c->n_pad / 256), dim3(256), 0, c->stream, c->dP, c->ckey[0],
                       c->cval[0], c->cell_start, (const double*)c->sc_tp, (const float*)c->sc_sp, F);
  } else if (sw) {
    hipLaunchKernelGGL((k_pair_cell<T, 1>), dim3(grid), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->cval[0], c->cell_start, c->items, c->n_items, F);
  } else {
    hipLaunchKernelGGL((k_pair_cell<T, 0>), dim3(grid), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->cval[0], c->cell_start, c->items, c->n_items, F);
  }
  if (timed) HIPCHK(hipEventRecord(c->ev1, c->stream));
  return OD_OK;
}

static int launch_forces(od_ctx* c, const double* X, const void* pos, const uint32_t* flags, double* F, bool timed) {
  return (c->prec == OD_PREC_FP32C) ? launch_forces_t<float>(c, X, pos, flags, F, timed)
                                    : launch_forces_t<double>(c, X, pos, flags, F, timed);
}

// flagged sources' share of the force on the binned ions, in cell order
template <class T>
static void launch_corr_cell(od_ctx* c, const typename Pos<T>::V* list, int add, double* out) {
  using PV = typename Pos<T>::V;
  if (c->cfg.switch_dist > 0.0 && c->cfg.switch_dist < c->cfg.cutoff)
    hipLaunchKernelGGL((k_corr_cell<T, 1>), dim3(c->nb), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->cval[0], c->cell_start, c->n_flag, list, add, out);
  else
    hipLaunchKernelGGL((k_corr_cell<T, 0>), dim3(c->nb), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->cval[0], c->cell_start, c->n_flag, list, add, out);
}

template <class T>
static int launch_step_t(od_ctx* c, bool timed) {
  using PV = typename Pos<T>::V;
  const int nb = c->nb;
  hipLaunchKernelGGL(k_kick_drift<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->X, c->V,
                     c->F, c->q, c->st, (PV*)c->pi, (PV*)c->pj, c->flags, c->block_cnt, c->err);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, c->block_cnt, nb, c->block_off, c->n_flag);
  hipLaunchKernelGGL(k_list<T>, dim3(nb), dim3(256), 0, c->stream, c->flags, c->block_cnt, c->block_off, c->n,
                     (const PV*)c->pi, c->slot_of, c->list_ion, (PV*)c->list_old);
  // flagged ions are parked in pj, as sources and as targets (cutoff mode:
  // left out of the cell list)
  CHKRET(launch_forces_t<T>(c, c->X, c->pj, c->flags, c->F, timed));
  const uint32_t dust_bits = (c->cfg.n_dust > 0) ? ((1u << c->cfg.n_dust) - 1u) : 0u;
  if (dust_bits) {
    if (c->cut)
      hipLaunchKernelGGL(k_rows_cell<T>, dim3(c->rows_grid), dim3(256), 0, c->stream, c->dP, c->X,
                         (const PV*)c->spos, c->cell_start, c->n_flag, (const PV*)c->list_old,
                         (const PV*)c->list_old, c->flags, c->list_ion, dust_bits, c->rowD);
    else
      hipLaunchKernelGGL(k_rows<T>, dim3(c->rows_grid), dim3(RowShape<T>::TB), 0, c->stream, c->dP, (const int*)nullptr,
                         (const PV*)c->pi, c->n_flag, (const PV*)c->list_old, c->flags, c->list_ion, dust_bits,
                         c->rowD);
  }
  if (c->cut) launch_corr_cell<T>(c, (const PV*)c->list_old, 0, c->Gcut);
  hipLaunchKernelGGL(k_kick2<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->block_off, nb,
                     c->X, c->V, c->F, c->q, c->st,
                     (const PV*)c->pi, (PV*)c->pj, c->flags, c->n_flag, c->slot_of, (const PV*)c->list_old,
                     (PV*)c->list_new, c->face, c->rowD, c->cut ? (const double*)c->Gcut : (const double*)nullptr,
                     c->partials);
  if (c->cut)
    hipLaunchKernelGGL(k_rows_cell<T>, dim3(c->rows_grid), dim3(256), 0, c->stream, c->dP, c->X, (const PV*)c->spos,
                       c->cell_start, c->n_flag, (const PV*)c->list_new, (const PV*)c->list_new, c->flags,
                       c->list_ion, 0u, c->rowF);
  else
    hipLaunchKernelGGL(k_rows<T>, dim3(c->rows_grid), dim3(RowShape<T>::TB), 0, c->stream, c->dP, (const int*)nullptr,
                       (const PV*)c->pj, c->n_flag, (const PV*)c->list_new, c->flags, c->list_ion, 0u, c->rowF);
  if (c->cut) launch_corr_cell<T>(c, (const PV*)c->list_new, 1, c->F);
  hipLaunchKernelGGL(k_inflow<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->block_off, nb,
                     c->F, (const PV*)c->pj, c->flags, c->n_flag, (const PV*)c->list_new, c->slot_of, c->rowF,
                     c->cut);
  hipLaunchKernelGGL(k_dust, dim3(1), dim3(256), 0, c->stream, c->dP, c->partials, c->q, c->st, c->rec,
                     std::max(1, c->cfg.n_dust), 1);
  return OD_OK;
}

static int launch_step(od_ctx* c, bool timed) {
  return (c->prec == OD_PREC_FP32C) ? launch_step_t<float>(c, timed) : launch_step_t<double>(c, timed);
}

template <class T>
static int load_positions_t(od_ctx* c, const double* X, void* pi, void* pj) {
  using PV = typename Pos<T>::V;
  const double s = (c->prec == OD_PREC_FP32C) ? 1.4426950408889634 / c->cfg.r_D_e : 1.0 / c->cfg.r_D_e;
  hipLaunchKernelGGL(k_load_pos<T>, dim3(c->nb), dim3(256), 0, c->stream, X, c->n, c->n_pad, s, (PV*)pi, (PV*)pj);
  return OD_OK;
}

static int load_positions(od_ctx* c, const double* X, void* pi, void* pj) {
  return (c->prec == OD_PREC_FP32C) ? load_positions_t<float>(c, X, pi, pj) : load_positions_t<double>(c, X, pi, pj);
}

static void aos_to_soa(const double* a, int n, int n_pad, std::vector<double>& s) {
  s.assign(3 * (size_t)n_pad, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) s[(size_t)k * n_pad + i] = a[3 * (size_t)i + k];
}

extern "C" int od_set_ions(od_ctx* c, const double* x, const double* v, int64_t n) {
  if (!c || !x || !v) return set_err(OD_ERR_ARG, "null argument");
  if (n != c->n) return set_err(OD_ERR_ARG, "od_set_ions: %lld ions, context holds %d", (long long)n, c->n);
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(v[i]))
      return set_err(OD_ERR_NONFINITE, "od_set_ions: non-finite value at ion %lld", (long long)(i / 3));
  HIPCHK(hipSetDevice(c->dev));
  std::vector<double> sx, sv;
  aos_to_soa(x, c->n, c->n_pad, sx);
  aos_to_soa(v, c->n, c->n_pad, sv);
  HIPCHK(hipMemcpyAsync(c->X, sx.data(), sx.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->V, sv.data(), sv.size() * 8, hipMemcpyHostToDevice, c->stream));
  CHKRET(load_positions(c, c->X, c->pi, c->pj));
  CHKRET(launch_forces(c, c->X, c->pj, nullptr, c->F, false));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_ions = true;
  return OD_OK;
}

extern "C" int od_get_ions(od_ctx* c, double* x, double* v) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  std::vector<double> s(3 * (size_t)c->n_pad);
  if (x) {
    HIPCHK(hipMemcpyAsync(s.data(), c->X, s.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->n; ++i)
      for (int k = 0; k < 3; ++k) x[3 * (size_t)i + k] = s[(size_t)k * c->n_pad + i];
  }
  if (v) {
    HIPCHK(hipMemcpyAsync(s.data(), c->V, s.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * (size_t)i + k] = s[(size_t)k * c->n_pad + i];
  }
  return OD_OK;
}

// Replaces the context's parameters (dust positions, radii, charge methods,
// plasma constants, trap coefficients, tables, seed) between runs while the
// ion state stays on the device.  Sizes and geometry that buffers or the cell
// list depend on must not change.  Dust charges keep their current values
// (od_set_dust_charges sets them).
extern "C" int od_update_config(od_ctx* c, const od_config* cfg) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  CHKRET(check_config(cfg));
  const od_config& o = c->cfg;
  if (cfg->n_ions != o.n_ions || cfg->n_dust != o.n_dust || cfg->precision != o.precision ||
      cfg->device != o.device || cfg->pair_mode != o.pair_mode || cfg->j_split != o.j_split ||
      cfg->cutoff != o.cutoff || cfg->switch_dist != o.switch_dist || cfg->R != o.R || cfg->H != o.H)
    return set_err(OD_ERR_ARG,
                   "od_update_config: ion count, dust count, precision, device, pair kernel, cutoff and the "
                   "cylinder must stay the same");
  HIPCHK(hipSetDevice(c->dev));
  DevParams hp;
  CHKRET(fill_params(cfg, c->n_pad, &hp));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->dP, &hp, sizeof(hp), hipMemcpyHostToDevice));
  const int keep_chunk = c->cfg.sym_chunk, keep_tile = c->cfg.sym_tile;
  c->cfg = *cfg;
  c->cfg.sym_chunk = keep_chunk;
  c->cfg.sym_tile = keep_tile;
  // captured steps hold kernel arguments derived from the old parameters
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
  return OD_OK;
}

extern "C" int od_set_dust_charges(od_ctx* c, const double* q) {
  if (!c || !q) return set_err(OD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  double h[OD_MAX_DUST] = {0};
  for (int p = 0; p < c->cfg.n_dust; ++p) h[p] = q[p];
  HIPCHK(hipMemcpyAsync(c->q, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OD_OK;
}

extern "C" int od_get_dust_charges(od_ctx* c, double* q) {
  if (!c || !q) return set_err(OD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  double h[OD_MAX_DUST];
  HIPCHK(hipMemcpyAsync(h, c->q, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int p = 0; p < c->cfg.n_dust; ++p) q[p] = h[p];
  return OD_OK;
}

extern "C" int od_set_step(od_ctx* c, int64_t step) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (step < 0 || step > 0xFFFFFFFFll) return set_err(OD_ERR_ARG, "step %lld outside u32 range", (long long)step);
  HIPCHK(hipSetDevice(c->dev));
  uint32_t s = (uint32_t)step;
  HIPCHK(hipMemcpyAsync(c->st, &s, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->host_step = step;
  return OD_OK;
}

static int build_graph(od_ctx* c, int steps) {
  if (c->gexec && c->graph_steps == steps) return OD_OK;
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (c->graph) {
    hipGraphDestroy(c->graph);
    c->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = OD_OK;
  for (int s = 0; s < steps && rc == OD_OK; ++s) rc = launch_step(c, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != OD_OK) return rc;
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  c->graph = g;
  HIPCHK(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
  c->graph_steps = steps;
  return OD_OK;
}

extern "C" int od_run(od_ctx* c, int64_t nsteps, od_dust_record* out, int32_t flags) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (nsteps < 0) return set_err(OD_ERR_ARG, "negative step count");
  if (c->host_step + nsteps > 0xFFFFFFFFll) return set_err(OD_ERR_ARG, "step counter would exceed u32");
  HIPCHK(hipSetDevice(c->dev));
  const bool profile = (flags & OD_RUN_PROFILE_PAIR) != 0;
  const bool graph = (flags & OD_RUN_GRAPH) != 0 && !profile;
  const int nd = c->cfg.n_dust;
  int64_t done = 0;
  while (done < nsteps) {
    const int chunk = (int)std::min<int64_t>(c->rec_cap, nsteps - done);
    uint32_t base = (uint32_t)c->host_step;
    HIPCHK(hipMemcpyAsync(c->st + 1, &base, 4, hipMemcpyHostToDevice, c->stream));
    int left = chunk;
    if (graph && chunk >= 4) {
      const int gs = std::min(chunk, 32);
      CHKRET(build_graph(c, gs));
      while (left >= gs) {
        HIPCHK(hipGraphLaunch(c->gexec, c->stream));
        left -= gs;
      }
    }
    for (; left > 0; --left) {
      CHKRET(launch_step(c, profile));
      HIPCHK(hipGetLastError());
      if (profile) {
        HIPCHK(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->pair_ms += ms;
        c->pair_launches += 1;
      }
    }
    if (out && nd > 0)
      HIPCHK(hipMemcpyAsync(out + done * nd, c->rec, (size_t)chunk * nd * sizeof(od_dust_record),
                            hipMemcpyDeviceToHost, c->stream));
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->host_step += chunk;
    done += chunk;
    if (herr) return set_err(OD_ERR_NONFINITE, "non-finite ion position or velocity before step %lld",
                             (long long)c->host_step);
  }
  return OD_OK;
}

extern "C" int od_get_cells(od_ctx* c, int32_t* cell_of_ion, int32_t* sorted_ions, int64_t* n_binned,
                            int32_t* dims) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (!c->cut) return set_err(OD_ERR_STATE, "od_get_cells: context has no cutoff (no cell list)");
  HIPCHK(hipSetDevice(c->dev));
  std::vector<uint32_t> key(c->n_pad), val(c->n_pad);
  HIPCHK(hipMemcpyAsync(key.data(), c->ckey[0], c->n_pad * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(val.data(), c->cval[0], c->n_pad * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int64_t nb = 0;
  for (int p = 0; p < c->n_pad; ++p) {
    const uint32_t i = val[p];
    if ((int)i < c->n && cell_of_ion) cell_of_ion[i] = (int32_t)key[p];
    if ((int)key[p] < c->n_cells) {
      if (sorted_ions) sorted_ions[nb] = (int32_t)i;
      ++nb;
    }
  }
  if (n_binned) *n_binned = nb;
  if (dims) {
    DevParams hp;
    fill_params(&c->cfg, c->n_pad, &hp);
    dims[0] = hp.nx;
    dims[1] = hp.ny;
    dims[2] = hp.nz;
  }
  return OD_OK;
}

extern "C" int od_cutoff_pair_count(od_ctx* c, int64_t* in_range, int64_t* candidates) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (!c->cut) return set_err(OD_ERR_STATE, "od_cutoff_pair_count: context has no cutoff");
  HIPCHK(hipSetDevice(c->dev));
  unsigned long long* d = nullptr;
  CHKRET(dalloc(&d, 16));
  unsigned long long h[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(d, h, 16, hipMemcpyHostToDevice, c->stream);
  const int grid = std::max(1, std::min(4096, c->max_items / 4 + 1));
  if (e == hipSuccess) {
    if (c->prec == OD_PREC_FP32C)
      hipLaunchKernelGGL(k_cell_pair_count<float>, dim3(grid), dim3(256), 0, c->stream, c->dP,
                         (const float4*)c->spos, c->cell_start, c->items, c->n_items, d);
    else
      hipLaunchKernelGGL(k_cell_pair_count<double>, dim3(grid), dim3(256), 0, c->stream, c->dP,
                         (const double4*)c->spos, c->cell_start, c->items, c->n_items, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "od_cutoff_pair_count: %s", hipGetErrorString(e));
  if (in_range) *in_range = (int64_t)h[0];
  if (candidates) *candidates = (int64_t)h[1];
  return OD_OK;
}

extern "C" int od_sym_choice(od_ctx* c, int32_t* sym_chunk, int32_t* sym_tile) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (sym_chunk) *sym_chunk = c->sym ? c->sym_ch : 0;
  if (sym_tile) *sym_tile = c->sym ? c->sym_var + 1 : 0;
  return OD_OK;
}

extern "C" int od_pair_info(od_ctx* c, int32_t* mode, int64_t* pairs_per_launch) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  const int64_t n = c->n;
  if (mode) *mode = c->cut ? OD_PAIR_CELL : (c->sym ? OD_PAIR_SYM : OD_PAIR_ROWS);
  if (c->cut) {
    if (pairs_per_launch) *pairs_per_launch = -1;  // data dependent: od_cutoff_pair_count
    return OD_OK;
  }
  if (pairs_per_launch) {
    if (c->sym) {
      // ordered pairs between different blocks (the main kernel's share)
      int64_t diag = 0;
      for (int X = 0; X < c->n_blocks; ++X) {
        const int64_t nx = std::max<int64_t>(0, std::min<int64_t>(SYM_TB, n - (int64_t)X * SYM_TB));
        diag += nx * (nx - 1);
      }
      *pairs_per_launch = n * (n - 1) - diag;
    } else {
      *pairs_per_launch = n * (n - 1);
    }
  }
  return OD_OK;
}

extern "C" int od_pair_timing(od_ctx* c, double* total_ms, int64_t* launches, int32_t reset) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (total_ms) *total_ms = c->pair_ms;
  if (launches) *launches = c->pair_launches;
  if (reset) {
    c->pair_ms = 0.0;
    c->pair_launches = 0;
  }
  return OD_OK;
}

extern "C" int od_get_step_info(od_ctx* c, uint32_t* flags, uint8_t* face, int64_t* n_reinjected) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (flags) HIPCHK(hipMemcpyAsync(flags, c->flags, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
  if (face) HIPCHK(hipMemcpyAsync(face, c->face, (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
  int nf = 0;
  HIPCHK(hipMemcpyAsync(&nf, c->n_flag, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (n_reinjected) *n_reinjected = nf;
  return OD_OK;
}

extern "C" int od_pair_forces(od_ctx* c, const double* x, double* f) {
  if (!c || !x || !f) return set_err(OD_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  const size_t np = c->n_pad;
  double *dX = nullptr, *dF = nullptr;
  void *pi = nullptr, *pj = nullptr;
  int rc = OD_OK;
  if ((rc = dalloc(&dX, 3 * np * 8)) || (rc = dalloc(&dF, 3 * np * 8)) || (rc = dalloc(&pi, np * c->pos_bytes)) ||
      (rc = dalloc(&pj, np * c->pos_bytes))) {
    hipFree(dX);
    hipFree(dF);
    hipFree(pi);
    hipFree(pj);
    return rc;
  }
  std::vector<double> sx;
  aos_to_soa(x, c->n, c->n_pad, sx);
  std::vector<double> sf(3 * np);
  hipError_t e = hipMemcpyAsync(dX, sx.data(), sx.size() * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    rc = load_positions(c, dX, pi, pj);
    if (rc == OD_OK) rc = launch_forces(c, dX, pj, nullptr, dF, false);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(sf.data(), dF, sf.size() * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  hipFree(dX);
  hipFree(dF);
  hipFree(pi);
  hipFree(pj);
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "od_pair_forces: %s", hipGetErrorString(e));
  if (rc != OD_OK) return rc;
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < 3; ++k) f[3 * (size_t)i + k] = sf[(size_t)k * np + i];
  return OD_OK;
}

extern "C" int od_trap_potential_grid(int32_t device, double R, double H, int32_t n_r, const double* rho,
                                      int32_t n_h, const double* h, int64_t n_mc, uint64_t seed, double* out) {
  if (!rho || !h || !out) return set_err(OD_ERR_ARG, "null argument");
  if (n_r < 1 || n_h < 1 || n_mc < 1) return set_err(OD_ERR_ARG, "grid and sample counts must be positive");
  if ((int64_t)n_r * n_h > (1 << 22)) return set_err(OD_ERR_ARG, "trap grid too large");
  int ndev = 0;
  hipError_t e0 = hipGetDeviceCount(&ndev);
  if (e0 != hipSuccess || ndev == 0) return set_err(OD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return set_err(OD_ERR_ARG, "device %d outside [0, %d)", device, ndev);
  HIPCHK(hipSetDevice(device));
  const int n_gp = n_r * n_h;
  const int64_t chunk = 1 << 16;
  const int n_chunks = (int)((n_mc + chunk - 1) / chunk);
  if ((int64_t)n_gp * n_chunks > (1ll << 31) - 1) return set_err(OD_ERR_ARG, "trap grid x samples too large");
  double *xs = nullptr, *ys = nullptr, *zs = nullptr, *drho = nullptr, *dh = nullptr, *part = nullptr,
         *dout = nullptr;
  int rc = OD_OK;
  auto cleanup = [&]() {
    hipFree(xs);
    hipFree(ys);
    hipFree(zs);
    hipFree(drho);
    hipFree(dh);
    hipFree(part);
    hipFree(dout);
  };
  if ((rc = dalloc(&xs, n_mc * 8)) || (rc = dalloc(&ys, n_mc * 8)) || (rc = dalloc(&zs, n_mc * 8)) ||
      (rc = dalloc(&drho, n_r * 8)) || (rc = dalloc(&dh, n_h * 8)) ||
      (rc = dalloc(&part, (size_t)n_gp * n_chunks * 8)) || (rc = dalloc(&dout, (size_t)n_gp * 8))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpy(drho, rho, n_r * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dh, h, n_h * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_trap_samples, dim3((unsigned)((n_mc + 255) / 256)), dim3(256), 0, 0, R, H, seed, n_mc, xs,
                       ys, zs);
    hipLaunchKernelGGL(k_trap_grid, dim3(n_gp * n_chunks), dim3(256), 0, 0, drho, dh, n_h, xs, ys, zs, n_mc, chunk,
                       n_chunks, part);
    hipLaunchKernelGGL(k_trap_finish, dim3((n_gp + 255) / 256), dim3(256), 0, 0, part, n_gp, n_chunks, (double)n_mc,
                       dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n_gp * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  cleanup();
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "od_trap_potential_grid: %s", hipGetErrorString(e));
  return OD_OK;
}

extern "C" int od_trap_samples_host(double R, double H, uint64_t seed, int64_t k0, int64_t n, double* xyz) {
  if (!xyz || n < 0 || k0 < 0) return set_err(OD_ERR_ARG, "bad argument");
  for (int64_t k = 0; k < n; ++k) trap_sample(R, H, seed, k0 + k, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
  return OD_OK;
}

extern "C" int od_abi_sizes(int64_t* config_size, int64_t* record_size) {
  if (config_size) *config_size = (int64_t)sizeof(od_config);
  if (record_size) *record_size = (int64_t)sizeof(od_dust_record);
  return OD_OK;
}

extern "C" const char* od_last_error(void) { return g_err.c_str(); }

extern "C" const char* od_build_info(void) { return "opendust_amd engine, gfx950, ABI " "1"; }

// ---------------------------------------------------------------------------
// Batched parameter sweeps: many independent simulations with uneven ion and
// dust counts advanced together.  Their ions are concatenated (each padded to a
// multiple of 1024), a block -> simulation table selects each 256-ion block's
// parameter record, and the all-pairs work of all simulations is one launch
// over a list of items (simulation, 512 targets, source chunk) so small and
// large simulations fill the GPU together.  The step is the same sequence of
// kernels as a single simulation's, each launched once for all simulations,
// and can be replayed as a captured hipGraph.
// ---------------------------------------------------------------------------
constexpr int BATCH_SRC_CHUNK = 8192;  // sources per pair item

struct PairItem {
  int sim;       // simulation
  int tgt;       // first of the 512 targets (global ion index)
  int src0, src1;  // source range (global ion indices)
};

template <class T>
__global__ __launch_bounds__(256) void k_pair_items(const DevParams* __restrict__ Pall,
                                                    const PairItem* __restrict__ items,
                                                    const typename Pos<T>::V* __restrict__ pj,
                                                    double* __restrict__ part) {
  const PairItem it = items[blockIdx.x];
  const DevParams* P = Pall + it.sim;
  const typename Pos<T>::V a = pj[it.tgt + threadIdx.x];
  const typename Pos<T>::V b = pj[it.tgt + 256 + threadIdx.x];
  Acc<T> ax0, ay0, az0, ax1, ay1, az1;
  const PlainLaw<T> law(*P);  // batched sweeps run all-pairs forces
  for (int jt = it.src0; jt < it.src1; jt += TJ) {
#pragma unroll 8
    for (int j = jt; j < jt + TJ; ++j) {
      const typename Pos<T>::V p = pj[j];
      law(a, p, ax0, ay0, az0);
      law(b, p, ax1, ay1, az1);
    }
    ax0.flush();
    ay0.flush();
    az0.flush();
    ax1.flush();
    ay1.flush();
    az1.flush();
  }
  double* o = part + (size_t)blockIdx.x * 3 * 512;
  o[threadIdx.x] = ax0.total();
  o[512 + threadIdx.x] = ay0.total();
  o[1024 + threadIdx.x] = az0.total();
  o[256 + threadIdx.x] = ax1.total();
  o[512 + 256 + threadIdx.x] = ay1.total();
  o[1024 + 256 + threadIdx.x] = az1.total();
}

// F = pair_scale(sim) * sum of the ion's items (fixed order); group_items[g] =
// {first item, count} of the 512-ion group g
__global__ __launch_bounds__(256) void k_items_reduce(const DevParams* __restrict__ Pall,
                                                      const int* __restrict__ blk_sim,
                                                      const int2* __restrict__ group_items,
                                                      const double* __restrict__ part, int n_tot,
                                                      double* __restrict__ F) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_tot) return;
  const DevParams* P = Pall + blk_sim[blockIdx.x];
  const int2 gi = group_items[i >> 9];
  const int off = i & 511;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int k = 0; k < gi.y; ++k) s += part[((size_t)(gi.x + k) * 3 + c) * 512 + off];
    F[(size_t)c * n_tot + i] = s * P->pair_scale;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_batch_load_pos(const DevParams* __restrict__ Pall,
                                                        const int* __restrict__ blk_sim, const double* __restrict__ X,
                                                        typename Pos<T>::V* __restrict__ pi,
                                                        typename Pos<T>::V* __restrict__ pj) {
  const DevParams* P = Pall + blk_sim[blockIdx.x];
  const int np = P->stride;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const typename Pos<T>::V p = (i - P->ion_off < P->n) ? Pos<T>::make(X[i], X[np + i], X[2 * np + i], P->pos_scale)
                                                      : Pos<T>::park();
  pi[i] = p;
  pj[i] = p;
}

struct od_batch {
  int n_sims = 0, dev = 0, prec = OD_PREC_FP32C;
  hipStream_t stream = nullptr;
  std::vector<od_config> cfgs;
  std::vector<int> off, npad;
  int n_tot = 0, nb = 0, n_items = 0, rows_grid = 1024, rec_cap = 128;
  size_t pos_bytes = 16;
  DevParams* dP = nullptr;
  int* blk_sim = nullptr;
  PairItem* items = nullptr;
  int2* group_items = nullptr;
  double *X = nullptr, *V = nullptr, *F = nullptr, *q = nullptr, *partials = nullptr, *rowF = nullptr,
         *rowD = nullptr, *part = nullptr;
  void *pi = nullptr, *pj = nullptr, *list_old = nullptr, *list_new = nullptr;
  uint32_t *flags = nullptr, *st = nullptr;
  uint8_t* face = nullptr;
  int *block_cnt = nullptr, *block_off = nullptr, *n_flag = nullptr, *slot_of = nullptr, *list_ion = nullptr,
      *err = nullptr;
  od_dust_record* rec = nullptr;
  int64_t host_step = 0;
  bool stale = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_steps = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double pair_ms = 0.0;
  int64_t pair_launches = 0, pairs_per_launch = 0;
  // symmetric FP32 pair kernel over all simulations (1024-ion blocks)
  int sym = 0, sym_ch = 1, n_wg = 0, n_blocks = 0, sym_var = 0;
  std::vector<int> blk_end_host;
  int2* wg_tab = nullptr;
  int *wg_base = nullptr, *blk_begin = nullptr, *blk_end = nullptr;
  float* blk_a = nullptr;
  double *blk_scale = nullptr, *sym_tp = nullptr, *sym_dg = nullptr;
  float* sym_sp = nullptr;
};

static void free_batch(od_batch* b) {
  if (!b) return;
  hipSetDevice(b->dev);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  void* bufs[] = {b->dP, b->blk_sim, b->items, b->group_items, b->X, b->V, b->F, b->q, b->partials, b->rowF,
                  b->rowD, b->part, b->pi, b->pj, b->list_old, b->list_new, b->flags, b->st, b->face, b->block_cnt,
                  b->block_off, b->n_flag, b->slot_of, b->list_ion, b->err, b->rec, b->wg_tab, b->wg_base,
                  b->blk_begin, b->blk_end, b->blk_a, b->blk_scale, b->sym_tp, b->sym_dg, b->sym_sp};
  for (void* p : bufs)
    if (p) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int od_batch_create(const od_config* cfgs, int32_t n_sims, od_batch** out) {
  if (!out || !cfgs) return set_err(OD_ERR_ARG, "null argument");
  *out = nullptr;
  if (n_sims < 1 || n_sims > 4096) return set_err(OD_ERR_ARG, "n_sims %d outside [1, 4096]", n_sims);
  for (int s = 0; s < n_sims; ++s) {
    CHKRET(check_config(&cfgs[s]));
    if (cfgs[s].precision != cfgs[0].precision || cfgs[s].device != cfgs[0].device)
      return set_err(OD_ERR_ARG, "all simulations of a batch share precision and device");
    if (cfgs[s].cutoff > 0.0) return set_err(OD_ERR_ARG, "batched sweeps run all-pairs forces (no cutoff)");
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return set_err(OD_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (cfgs[0].device < 0 || cfgs[0].device >= ndev) return set_err(OD_ERR_ARG, "device %d out of range", cfgs[0].device);
  od_batch* b = new od_batch();
  auto fail = [&](int r) {
    free_batch(b);
    return r;
  };
  b->n_sims = n_sims;
  b->dev = cfgs[0].device;
  b->prec = cfgs[0].precision;
  b->pos_bytes = (b->prec == OD_PREC_FP32C) ? sizeof(float4) : sizeof(double4);
  b->cfgs.assign(cfgs, cfgs + n_sims);
  int64_t tot = 0;
  for (int s = 0; s < n_sims; ++s) {
    const int np = (cfgs[s].n_ions + 1023) / 1024 * 1024;
    b->off.push_back((int)tot);
    b->npad.push_back(np);
    tot += np;
  }
  if (tot > (1 << 28)) return fail(set_err(OD_ERR_ARG, "batch too large (%lld ions)", (long long)tot));
  b->n_tot = (int)tot;
  b->nb = b->n_tot / 256;
  // parameter records and tables
  std::vector<DevParams> hp(n_sims);
  std::vector<int> blk_sim(b->nb);
  std::vector<PairItem> items;
  std::vector<int2> groups(b->n_tot / 512);
  for (int s = 0; s < n_sims; ++s) {
    fill_params(&cfgs[s], b->npad[s], &hp[s]);
    hp[s].ion_off = b->off[s];
    hp[s].stride = b->n_tot;
    hp[s].blk0 = b->off[s] / 256;
    hp[s].blk1 = (b->off[s] + b->npad[s]) / 256;
    hp[s].q_off = s * OD_MAX_DUST;
    for (int k = hp[s].blk0; k < hp[s].blk1; ++k) blk_sim[k] = s;
    for (int t = 0; t < b->npad[s]; t += 512) {
      groups[(b->off[s] + t) / 512] = make_int2((int)items.size(), 0);
      for (int j = 0; j < b->npad[s]; j += BATCH_SRC_CHUNK) {
        PairItem it;
        it.sim = s;
        it.tgt = b->off[s] + t;
        it.src0 = b->off[s] + j;
        it.src1 = b->off[s] + std::min(b->npad[s], j + BATCH_SRC_CHUNK);
        items.push_back(it);
        groups[(b->off[s] + t) / 512].y += 1;
      }
    }
  }
  b->n_items = (int)items.size();
  b->sym = (b->prec == OD_PREC_FP32C && cfgs[0].pair_mode != OD_PAIR_ROWS) ? 1 : 0;
  const size_t np = b->n_tot;
  int rc = OD_OK;
  if (hipSetDevice(b->dev) != hipSuccess) return fail(set_err(OD_ERR_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "hipStreamCreate failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "hipEventCreate failed"));
  for (int s = 0; s < n_sims; ++s) b->pairs_per_launch += (int64_t)cfgs[s].n_ions * (cfgs[s].n_ions - 1);
  if ((rc = dalloc(&b->dP, n_sims * sizeof(DevParams))) || (rc = dalloc(&b->blk_sim, b->nb * 4)) ||
      (rc = dalloc(&b->items, items.size() * sizeof(PairItem))) ||
      (rc = dalloc(&b->group_items, groups.size() * sizeof(int2))) || (rc = dalloc(&b->X, 3 * np * 8)) ||
      (rc = dalloc(&b->V, 3 * np * 8)) || (rc = dalloc(&b->F, 3 * np * 8)) ||
      (rc = dalloc(&b->q, (size_t)n_sims * OD_MAX_DUST * 8)) ||
      (rc = dalloc(&b->partials, (size_t)b->nb * OD_MAX_DUST * 8 * 8)) || (rc = dalloc(&b->rowF, 3 * np * 8)) ||
      (rc = dalloc(&b->rowD, 3 * np * 8)) || (rc = dalloc(&b->part, b->sym ? 8 : (size_t)b->n_items * 3 * 512 * 8)) ||
      (rc = dalloc(&b->pi, np * b->pos_bytes)) || (rc = dalloc(&b->pj, np * b->pos_bytes)) ||
      (rc = dalloc(&b->list_old, np * b->pos_bytes)) || (rc = dalloc(&b->list_new, np * b->pos_bytes)) ||
      (rc = dalloc(&b->flags, np * 4)) || (rc = dalloc(&b->st, 16)) || (rc = dalloc(&b->face, np)) ||
      (rc = dalloc(&b->block_cnt, b->nb * 4)) || (rc = dalloc(&b->block_off, b->nb * 4)) ||
      (rc = dalloc(&b->n_flag, 4)) || (rc = dalloc(&b->slot_of, np * 4)) || (rc = dalloc(&b->list_ion, np * 4)) ||
      (rc = dalloc(&b->err, 4)) ||
      (rc = dalloc(&b->rec, (size_t)b->rec_cap * n_sims * OD_MAX_DUST * sizeof(od_dust_record))))
    return fail(rc);
  std::vector<double> q0((size_t)n_sims * OD_MAX_DUST, 0.0);
  for (int s = 0; s < n_sims; ++s)
    for (int p = 0; p < cfgs[s].n_dust; ++p) q0[(size_t)s * OD_MAX_DUST + p] = cfgs[s].dust_q0[p];
  uint32_t st0[4] = {0, 0, 0, 0};
  if (hipMemcpy(b->dP, hp.data(), n_sims * sizeof(DevParams), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->blk_sim, blk_sim.data(), b->nb * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->items, items.data(), items.size() * sizeof(PairItem), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->group_items, groups.data(), groups.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->q, q0.data(), q0.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->st, st0, sizeof(st0), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->X, 0, 3 * np * 8) != hipSuccess || hipMemset(b->V, 0, 3 * np * 8) != hipSuccess ||
      hipMemset(b->F, 0, 3 * np * 8) != hipSuccess || hipMemset(b->err, 0, 4) != hipSuccess ||
      hipMemset(b->n_flag, 0, 4) != hipSuccess || hipMemset(b->flags, 0, np * 4) != hipSuccess ||
      hipMemset(b->face, 0, np) != hipSuccess ||
      hipMemset(b->rec, 0, (size_t)b->rec_cap * n_sims * OD_MAX_DUST * sizeof(od_dust_record)) != hipSuccess)
    return fail(set_err(OD_ERR_HIP, "batch initialisation failed"));
  if (b->sym) {
    // triangular schedule over every simulation's 1024-ion blocks
    const int NB = b->n_tot / SYM_TB;
    std::vector<int> bb(NB), be(NB);
    std::vector<float> ba(NB);
    std::vector<double> bs(NB);
    for (int s = 0; s < n_sims; ++s)
      for (int X = b->off[s] / SYM_TB; X < (b->off[s] + b->npad[s]) / SYM_TB; ++X) {
        bb[X] = b->off[s] / SYM_TB;
        be[X] = (b->off[s] + b->npad[s]) / SYM_TB;
        ba[X] = hp[s].a32;
        bs[X] = hp[s].pair_scale;
      }
    // ordered pairs of real ions the symmetric launch covers (different blocks)
    b->pairs_per_launch = 0;
    for (int s = 0; s < n_sims; ++s) {
      const int64_t n = cfgs[s].n_ions;
      int64_t diag = 0;
      for (int64_t x0 = 0; x0 < n; x0 += SYM_TB) {
        const int64_t m = std::min<int64_t>(SYM_TB, n - x0);
        diag += m * (m - 1);
      }
      b->pairs_per_launch += n * (n - 1) - diag;
    }
    b->blk_end_host = be;
    if ((rc = sym_build(b, be, sym_default_chunk(be, cfgs[0].sym_chunk))) || (rc = dalloc(&b->blk_begin, NB * 4)) ||
        (rc = dalloc(&b->blk_end, NB * 4)) || (rc = dalloc(&b->blk_a, NB * 4)) ||
        (rc = dalloc(&b->blk_scale, NB * 8)) || (rc = dalloc(&b->sym_dg, 3 * np * 8)))
      return fail(rc);
    if (hipMemcpy(b->blk_begin, bb.data(), NB * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->blk_end, be.data(), NB * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->blk_a, ba.data(), NB * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->blk_scale, bs.data(), NB * 8, hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(OD_ERR_HIP, "batch initialisation failed"));
    b->sym_var = sym_default_variant(b->n_wg, cfgs[0].sym_tile);
  }
  *out = b;
  return OD_OK;
}

extern "C" int od_batch_destroy(od_batch* b) {
  free_batch(b);
  return OD_OK;
}

extern "C" int od_batch_set_ions(od_batch* b, int32_t sim, const double* x, const double* v, int64_t n) {
  if (!b || !x || !v) return set_err(OD_ERR_ARG, "null argument");
  if (sim < 0 || sim >= b->n_sims) return set_err(OD_ERR_ARG, "simulation %d out of range", sim);
  if (n != b->cfgs[sim].n_ions)
    return set_err(OD_ERR_ARG, "simulation %d holds %d ions, got %lld", sim, b->cfgs[sim].n_ions, (long long)n);
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(v[i])) return set_err(OD_ERR_NONFINITE, "non-finite ion state");
  HIPCHK(hipSetDevice(b->dev));
  const size_t np = b->n_tot;
  for (int k = 0; k < 3; ++k) {
    std::vector<double> cx(n), cv(n);
    for (int64_t i = 0; i < n; ++i) {
      cx[i] = x[3 * i + k];
      cv[i] = v[3 * i + k];
    }
    HIPCHK(hipMemcpy(b->X + k * np + b->off[sim], cx.data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->V + k * np + b->off[sim], cv.data(), n * 8, hipMemcpyHostToDevice));
  }
  b->stale = true;
  return OD_OK;
}

extern "C" int od_batch_get_ions(od_batch* b, int32_t sim, double* x, double* v) {
  if (!b) return set_err(OD_ERR_ARG, "null batch");
  if (sim < 0 || sim >= b->n_sims) return set_err(OD_ERR_ARG, "simulation %d out of range", sim);
  HIPCHK(hipSetDevice(b->dev));
  const int n = b->cfgs[sim].n_ions;
  const size_t np = b->n_tot;
  std::vector<double> c(n);
  for (int k = 0; k < 3; ++k) {
    if (x) {
      HIPCHK(hipMemcpy(c.data(), b->X + k * np + b->off[sim], n * 8, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) x[3 * (size_t)i + k] = c[i];
    }
    if (v) {
      HIPCHK(hipMemcpy(c.data(), b->V + k * np + b->off[sim], n * 8, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) v[3 * (size_t)i + k] = c[i];
    }
  }
  return OD_OK;
}

template <class T>
static void batch_pair(od_batch* b, bool timed, double* Fout) {
  using PV = typename Pos<T>::V;
  if (b->sym) {
    const float linv = (float)(1.0 / 1.4426950408889634);
    if (timed) hipEventRecord(b->ev0, b->stream);
    if (b->n_wg > 0)
      hipLaunchKernelGGL(sym_variants[b->sym_var].fn, dim3(b->n_wg), dim3(SYM_TB / sym_variants[b->sym_var].mt), 0,
                         b->stream, (const float4*)b->pj, b->n_blocks,
                         b->sym_ch, (const int2*)b->wg_tab, (const int*)b->wg_base, 0.0f, linv,
                         (const int*)b->blk_end, (const float*)b->blk_a, b->sym_tp, b->sym_sp);
    if (timed) hipEventRecord(b->ev1, b->stream);
    hipLaunchKernelGGL(k_pair32_diag, dim3(b->n_tot / 256), dim3(256), 0, b->stream, (const float4*)b->pj, b->n_tot,
                       0.0f, linv, (const float*)b->blk_a, b->sym_dg);
    hipLaunchKernelGGL(k_sym_reduce, dim3(b->n_tot / 256), dim3(256), 0, b->stream, b->sym_dg, b->sym_tp,
                       b->sym_sp, b->wg_base, b->n_blocks, b->sym_ch, b->n_tot, 0.0, (const int*)b->blk_begin,
                       (const double*)b->blk_scale, Fout);
    return;
  }
  if (timed) hipEventRecord(b->ev0, b->stream);
  hipLaunchKernelGGL(k_pair_items<T>, dim3(b->n_items), dim3(256), 0, b->stream, b->dP, b->items, (const PV*)b->pj,
                     b->part);
  if (timed) hipEventRecord(b->ev1, b->stream);
  hipLaunchKernelGGL(k_items_reduce, dim3(b->nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->group_items,
                     b->part, b->n_tot, Fout);
}

template <class T>
static int batch_step_t(od_batch* b, bool timed) {
  using PV = typename Pos<T>::V;
  const int nb = b->nb;
  hipLaunchKernelGGL(k_kick_drift<T>, dim3(nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->X, b->V, b->F, b->q,
                     b->st, (PV*)b->pi, (PV*)b->pj, b->flags, b->block_cnt, b->err);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, b->stream, b->block_cnt, nb, b->block_off, b->n_flag);
  hipLaunchKernelGGL(k_list<T>, dim3(nb), dim3(256), 0, b->stream, b->flags, b->block_cnt, b->block_off, b->n_tot,
                     (const PV*)b->pi, b->slot_of, b->list_ion, (PV*)b->list_old);
  batch_pair<T>(b, timed, b->F);
  hipLaunchKernelGGL(k_rows<T>, dim3(b->rows_grid), dim3(RowShape<T>::TB), 0, b->stream, b->dP, b->blk_sim, (const PV*)b->pi,
                     b->n_flag, (const PV*)b->list_old, b->flags, b->list_ion, (uint32_t)((1u << OD_MAX_DUST) - 1u),
                     b->rowD);
  hipLaunchKernelGGL(k_kick2<T>, dim3(nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->block_off, nb, b->X, b->V,
                     b->F, b->q, b->st, (const PV*)b->pi, (PV*)b->pj, b->flags, b->n_flag, b->slot_of,
                     (const PV*)b->list_old, (PV*)b->list_new, b->face, b->rowD, (const double*)nullptr,
                     b->partials);
  hipLaunchKernelGGL(k_rows<T>, dim3(b->rows_grid), dim3(RowShape<T>::TB), 0, b->stream, b->dP, b->blk_sim, (const PV*)b->pj,
                     b->n_flag, (const PV*)b->list_new, b->flags, b->list_ion, 0u, b->rowF);
  hipLaunchKernelGGL(k_inflow<T>, dim3(nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->block_off, nb, b->F,
                     (const PV*)b->pj, b->flags, b->n_flag, (const PV*)b->list_new, b->slot_of, b->rowF, 0);
  hipLaunchKernelGGL(k_dust, dim3(b->n_sims), dim3(256), 0, b->stream, b->dP, b->partials, b->q, b->st, b->rec,
                     (int)OD_MAX_DUST, 0);
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, b->stream, b->st);
  return OD_OK;
}

static int batch_step(od_batch* b, bool timed) {
  return (b->prec == OD_PREC_FP32C) ? batch_step_t<float>(b, timed) : batch_step_t<double>(b, timed);
}

template <class T>
static int batch_forces_t(od_batch* b) {
  using PV = typename Pos<T>::V;
  hipLaunchKernelGGL(k_batch_load_pos<T>, dim3(b->nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->X,
                     (PV*)b->pi, (PV*)b->pj);
  batch_pair<T>(b, false, b->F);
  return OD_OK;
}

extern "C" int od_batch_run(od_batch* b, int64_t nsteps, od_dust_record* out, int32_t flags) {
  if (!b) return set_err(OD_ERR_ARG, "null batch");
  if (nsteps < 0) return set_err(OD_ERR_ARG, "negative step count");
  if (b->host_step + nsteps > 0xFFFFFFFFll) return set_err(OD_ERR_ARG, "step counter would exceed u32");
  HIPCHK(hipSetDevice(b->dev));
  if (b->stale) {
    CHKRET((b->prec == OD_PREC_FP32C) ? batch_forces_t<float>(b) : batch_forces_t<double>(b));
    HIPCHK(hipGetLastError());
    b->stale = false;
  }
  const bool profile = (flags & OD_RUN_PROFILE_PAIR) != 0;
  const bool graph = (flags & OD_RUN_GRAPH) != 0 && !profile;
  int64_t done = 0;
  const size_t rec_step = (size_t)b->n_sims * OD_MAX_DUST;
  while (done < nsteps) {
    const int chunk = (int)std::min<int64_t>(b->rec_cap, nsteps - done);
    uint32_t base = (uint32_t)b->host_step;
    HIPCHK(hipMemcpyAsync(b->st + 1, &base, 4, hipMemcpyHostToDevice, b->stream));
    int left = chunk;
    if (graph && chunk >= 4) {
      const int gs = std::min(chunk, 16);
      if (!(b->gexec && b->graph_steps == gs)) {
        if (b->gexec) hipGraphExecDestroy(b->gexec);
        if (b->graph) hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = OD_OK;
        for (int s = 0; s < gs && rc == OD_OK; ++s) rc = batch_step(b, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != OD_OK) return rc;
        if (e != hipSuccess) return set_err(OD_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        b->graph_steps = gs;
      }
      while (left >= gs) {
        HIPCHK(hipGraphLaunch(b->gexec, b->stream));
        left -= gs;
      }
    }
    for (; left > 0; --left) {
      CHKRET(batch_step(b, profile));
      HIPCHK(hipGetLastError());
      if (profile) {
        HIPCHK(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->pair_ms += ms;
        b->pair_launches += 1;
      }
    }
    if (out)
      HIPCHK(hipMemcpyAsync(out + done * rec_step, b->rec, (size_t)chunk * rec_step * sizeof(od_dust_record),
                            hipMemcpyDeviceToHost, b->stream));
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, b->err, 4, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->host_step += chunk;
    done += chunk;
    if (herr) return set_err(OD_ERR_NONFINITE, "non-finite ion state in the batch before step %lld",
                             (long long)b->host_step);
  }
  return OD_OK;
}

extern "C" int od_batch_get_step_info(od_batch* b, int32_t sim, uint32_t* flags, uint8_t* face) {
  if (!b) return set_err(OD_ERR_ARG, "null batch");
  if (sim < 0 || sim >= b->n_sims) return set_err(OD_ERR_ARG, "simulation %d out of range", sim);
  HIPCHK(hipSetDevice(b->dev));
  const int n = b->cfgs[sim].n_ions;
  if (flags) HIPCHK(hipMemcpy(flags, b->flags + b->off[sim], (size_t)n * 4, hipMemcpyDeviceToHost));
  if (face) HIPCHK(hipMemcpy(face, b->face + b->off[sim], (size_t)n, hipMemcpyDeviceToHost));
  return OD_OK;
}

extern "C" int od_batch_pair_timing(od_batch* b, double* total_ms, int64_t* launches, int64_t* pairs_per_launch,
                                    int32_t reset) {
  if (!b) return set_err(OD_ERR_ARG, "null batch");
  if (total_ms) *total_ms = b->pair_ms;
  if (launches) *launches = b->pair_launches;
  if (pairs_per_launch) *pairs_per_launch = b->pairs_per_launch;
  if (reset) {
    b->pair_ms = 0.0;
    b->pair_launches = 0;
  }
  return OD_OK;
}

// Autotuner of the symmetric pair evaluation (kernel, diagonal, reduction) on
// the current positions.  Stage 1 times each candidate number of source blocks
// per work-group with the configured kernel variant; stage 2 every variant
// (tile shape, waves per SIMD) at the best block count; stage 3 the
// neighbouring block counts with the best variant.  Both choices group the
// partial sums, so runs are reproducible for a given pair, which is returned
// and can be fixed through od_config.sym_chunk and od_config.sym_tile.
static const int kTuneChunks[] = {1, 2, 4, 8, 16, 32};

template <class C, class F>
static int tune_sym(C* c, const std::vector<int>& be, int reps, F&& evaluate, int* chosen_ch, int* chosen_var) {
  double best_ms = 1e300;
  int best_ch = c->sym_ch, best_var = c->sym_var;
  // a captured step holds the schedule's buffers, which sym_build replaces:
  // drop it before the first rebuild, so no error path can leave it behind
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
  auto time_one = [&](int ch, int var) -> int {
    if (sym_wg_count(be, ch) == 0 || sym_bytes(be, ch) > 48e9) return OD_OK;
    if (c->sym_ch != ch) CHKRET(sym_build(c, be, ch));
    c->sym_var = var;
    evaluate();  // warm-up
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) evaluate();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms < best_ms) {
      best_ms = ms;
      best_ch = ch;
      best_var = var;
    }
    return OD_OK;
  };
  const int var0 = c->sym_var;
  for (int ch : kTuneChunks) CHKRET(time_one(ch, var0));
  const int ch1 = best_ch;
  for (int v = 0; v < SYM_N_VARIANTS; ++v)
    if (v != var0) CHKRET(time_one(ch1, v));
  if (best_var != var0) {
    for (int ch : kTuneChunks)
      if (ch != ch1 && (ch == 2 * ch1 || 2 * ch == ch1)) CHKRET(time_one(ch, best_var));
  }
  if (c->sym_ch != best_ch) CHKRET(sym_build(c, be, best_ch));
  c->sym_var = best_var;
  *chosen_ch = best_ch;
  *chosen_var = best_var;
  return OD_OK;
}

extern "C" int od_sym_variants(void) { return SYM_N_VARIANTS; }

extern "C" int od_sym_variant_info(int32_t v, int32_t* targets_per_lane, int32_t* waves_per_simd) {
  if (v < 1 || v > SYM_N_VARIANTS) return set_err(OD_ERR_ARG, "variant %d outside [1, %d]", v, SYM_N_VARIANTS);
  if (targets_per_lane) *targets_per_lane = sym_variants[v - 1].mt;
  if (waves_per_simd) *waves_per_simd = sym_variants[v - 1].wpe;
  return OD_OK;
}

extern "C" int od_tune(od_ctx* c, int32_t reps, int32_t* sym_chunk, int32_t* sym_tile) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (reps < 1 || reps > 100) return set_err(OD_ERR_ARG, "reps %d outside [1, 100]", reps);
  int chosen = 0, var = -1;
  if (c->sym) {
    if (!c->have_ions) return set_err(OD_ERR_STATE, "od_tune needs the ions (call od_set_ions first)");
    HIPCHK(hipSetDevice(c->dev));
    double* Fs = nullptr;
    CHKRET(dalloc(&Fs, (size_t)3 * c->n_pad * 8));
    const std::vector<int> be(c->n_blocks, c->n_blocks);
    const int rc = tune_sym(c, be, reps, [&]() { launch_pair(c, c->pj, c->pj, Fs, false); }, &chosen, &var);
    hipFree(Fs);
    if (rc != OD_OK) return rc;
    // the ion-ion force of the current positions, with the chosen schedule
    CHKRET(load_positions(c, c->X, c->pi, c->pj));
    CHKRET(launch_forces(c, c->X, c->pj, nullptr, c->F, false));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (sym_chunk) *sym_chunk = chosen;
  if (sym_tile) *sym_tile = var + 1;
  return OD_OK;
}

extern "C" int od_batch_tune(od_batch* b, int32_t reps, int32_t* sym_chunk, int32_t* sym_tile) {
  if (!b) return set_err(OD_ERR_ARG, "null batch");
  if (reps < 1 || reps > 100) return set_err(OD_ERR_ARG, "reps %d outside [1, 100]", reps);
  int chosen = 0, var = -1;
  if (b->sym) {
    HIPCHK(hipSetDevice(b->dev));
    if (b->stale)
      hipLaunchKernelGGL(k_batch_load_pos<float>, dim3(b->nb), dim3(256), 0, b->stream, b->dP, b->blk_sim, b->X,
                         (float4*)b->pi, (float4*)b->pj);
    double* Fs = nullptr;
    CHKRET(dalloc(&Fs, (size_t)3 * b->n_tot * 8));
    const int rc = tune_sym(b, b->blk_end_host, reps, [&]() { batch_pair<float>(b, false, Fs); }, &chosen, &var);
    hipFree(Fs);
    if (rc != OD_OK) return rc;
    b->stale = true;  // the next run recomputes the forces with the chosen schedule
  }
  if (sym_chunk) *sym_chunk = chosen;
  if (sym_tile) *sym_tile = var + 1;
  return OD_OK;
}

// ---------------------------------------------------------------------------
// Axial ion-density profile (tools/concentration.cpp:32-101).  One wave per
// node walks the ions in index order; every lane computes its ion's deposit on
// that node with the tool's expressions (FP64, no contraction), and the wave
// adds the non-zero deposits in lane order, so each node's sum is the tool's
// sequential sum bit for bit (an ion deposits at most once on a node, and
// adding an exact zero changes nothing).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_density(const double* __restrict__ z_nm, int64_t N, int n, double L,
                                                const double* __restrict__ zg, double* __restrict__ sums) {
  const int node = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int64_t j0 = 0; j0 < N; j0 += 64) {
    const int64_t j = j0 + lane;
    double v = 0.0;
    if (j < N) {
      const double z = z_nm[j] / 1e9;
      const int index = (int)((0.5 * L + z) * (n - 1) / L);
      if (index == 0 && z >= -0.5 * L) {
        if (node == 1) v = 1.0 - (zg[1] - z) / (zg[1] - zg[0]);
      } else if (index == n - 2) {
        if (node == n - 2) v = 1.0 - (z - zg[n - 2]) / (zg[n - 1] - zg[n - 2]);
      } else if (index < n - 2 && index > 0) {
        if (node == index) v = 1.0 - (z - zg[index]) / (zg[index + 1] - zg[index]);
        if (node == index + 1) v = 1.0 - (zg[index + 1] - z) / (zg[index + 1] - zg[index]);
      }
    }
    uint64_t m = __ballot(v != 0.0);
    while (m) {
      const int l = __builtin_ctzll(m);
      s = s + __shfl(v, l, 64);
      m &= m - 1;
    }
  }
  if (lane == 0) sums[node] = s;
}

static void density_finish(int n, double R, double L, double N_s, const std::vector<double>& sums, double* z_grid,
                           double* conc) {
  std::vector<double> zg(n);
  for (int i = 0; i < n; i++) zg[i] = -L / 2.0 + i * L / (double)(n - 1);
  std::vector<double> c(sums);
  const double pi = 3.14159265358979323846;
  for (int i = 1; i < n - 1; i++) c[i] /= 0.5 * ((zg[i + 1] - zg[i]) + (zg[i] - zg[i - 1])) * (pi * R * R) / N_s;
  c[0] = 2 * c[1] - c[2];
  c[n - 1] = 2 * c[n - 2] - c[n - 3];
  for (int i = 0; i < n; i++) {
    if (z_grid) z_grid[i] = zg[i];
    if (conc) conc[i] = c[i];
  }
}

static int density_run(hipStream_t st, const double* z_dev, int64_t N, int n, double R, double L, double N_s,
                       double* z_grid, double* conc) {
  std::vector<double> zg(n);
  for (int i = 0; i < n; i++) zg[i] = -L / 2.0 + i * L / (double)(n - 1);
  double *d_zg = nullptr, *d_s = nullptr;
  int rc = OD_OK;
  if ((rc = dalloc(&d_zg, (size_t)n * 8)) || (rc = dalloc(&d_s, (size_t)n * 8))) {
    if (d_zg) hipFree(d_zg);
    return rc;
  }
  std::vector<double> sums(n);
  hipError_t e = hipMemcpyAsync(d_zg, zg.data(), (size_t)n * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_density, dim3(n), dim3(64), 0, st, z_dev, N, n, L, (const double*)d_zg, d_s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(sums.data(), d_s, (size_t)n * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d_zg);
  hipFree(d_s);
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "density profile: %s", hipGetErrorString(e));
  density_finish(n, R, L, N_s, sums, z_grid, conc);
  return OD_OK;
}

static int density_check(int n, double R, double L, double N_s) {
  if (n < 3 || n > 65536) return set_err(OD_ERR_ARG, "density profile: %d nodes outside [3, 65536]", n);
  if (!(R > 0.0) || !(L > 0.0) || !(N_s > 0.0) || !std::isfinite(R) || !std::isfinite(L) || !std::isfinite(N_s))
    return set_err(OD_ERR_ARG, "density profile: R, L and N_s must be positive and finite");
  return OD_OK;
}

extern "C" int od_density_profile(od_ctx* c, int32_t n_nodes, double R, double L, double N_s, double* z_grid,
                                  double* conc) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  CHKRET(density_check(n_nodes, R, L, N_s));
  if (!c->have_ions) return set_err(OD_ERR_STATE, "density profile: no ions yet (call od_set_ions first)");
  HIPCHK(hipSetDevice(c->dev));
  return density_run(c->stream, c->X + 2 * (size_t)c->n_pad, c->n, n_nodes, R, L, N_s, z_grid, conc);
}

extern "C" int od_density_profile_z(int32_t device, const double* z_nm, int64_t N, int32_t n_nodes, double R,
                                    double L, double N_s, double* z_grid, double* conc) {
  if (!z_nm && N > 0) return set_err(OD_ERR_ARG, "null positions");
  if (N < 0) return set_err(OD_ERR_ARG, "negative ion count");
  CHKRET(density_check(n_nodes, R, L, N_s));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return set_err(OD_ERR_ARG, "device %d not available", device);
  HIPCHK(hipSetDevice(device));
  double* d_z = nullptr;
  CHKRET(dalloc(&d_z, (size_t)std::max<int64_t>(N, 1) * 8));
  int rc = OD_OK;
  if (N > 0 && hipMemcpy(d_z, z_nm, (size_t)N * 8, hipMemcpyHostToDevice) != hipSuccess)
    rc = set_err(OD_ERR_HIP, "density profile: upload failed");
  if (rc == OD_OK) rc = density_run(nullptr, d_z, N, n_nodes, R, L, N_s, z_grid, conc);
  hipFree(d_z);
  return rc;
}
