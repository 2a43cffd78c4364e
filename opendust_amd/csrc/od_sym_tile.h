// od_sym_tile.h -- the register tile of the symmetric FP32 pair kernels
// (all-pairs and cell list): 64 sources per wave against MT 64-target
// sub-blocks, each unordered pair evaluated once, its force added to the
// target and its reaction to the source.  Included by the engine's pair and
// cell kernels and by the VALU issue probe, which times this tile alone.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "od_device.h"

namespace od {

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

// the same for one and for two target sub-blocks
template <int S>
__device__ __forceinline__ void sym_diffs1(const float* tx, const float* ty, const float* tz, float xj, float yj,
                                           float zj, float* dx, float* dy, float* dz) {
  if constexpr (S == 0) {
    dx[0] = xj - tx[0];
    dy[0] = yj - ty[0];
    dz[0] = zj - tz[0];
  } else {
    asm("s_nop 1\n\t"
        "v_sub_f32_dpp %0, %3, %6 row_ror:%c9 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %1, %4, %7 row_ror:%c9 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %2, %5, %8 row_ror:%c9 row_mask:0xf bank_mask:0xf"
        : "=&v"(dx[0]), "=&v"(dy[0]), "=&v"(dz[0])
        : "v"(xj), "v"(yj), "v"(zj), "v"(tx[0]), "v"(ty[0]), "v"(tz[0]), "i"(S));
  }
}

template <int S>
__device__ __forceinline__ void sym_diffs2(const float* tx, const float* ty, const float* tz, float xj, float yj,
                                           float zj, float* dx, float* dy, float* dz) {
  if constexpr (S == 0) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      dx[k] = xj - tx[k];
      dy[k] = yj - ty[k];
      dz[k] = zj - tz[k];
    }
  } else {
    asm("s_nop 1\n\t"
        "v_sub_f32_dpp %0, %6, %9 row_ror:%c15 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %1, %7, %10 row_ror:%c15 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %2, %8, %11 row_ror:%c15 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %3, %6, %12 row_ror:%c15 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %4, %7, %13 row_ror:%c15 row_mask:0xf bank_mask:0xf\n\t"
        "v_sub_f32_dpp %5, %8, %14 row_ror:%c15 row_mask:0xf bank_mask:0xf"
        : "=&v"(dx[0]), "=&v"(dy[0]), "=&v"(dz[0]), "=&v"(dx[1]), "=&v"(dy[1]), "=&v"(dz[1])
        : "v"(xj), "v"(yj), "v"(zj), "v"(tx[0]), "v"(ty[0]), "v"(tz[0]), "v"(tx[1]), "v"(ty[1]), "v"(tz[1]),
          "i"(S));
  }
}

template <int S, int MT>
__device__ __forceinline__ void sym_diffs(const SymT<MT>& T, float xj, float yj, float zj, float* dx, float* dy,
                                          float* dz) {
  static_assert(MT == 1 || MT == 2 || MT % 4 == 0, "targets per wave: 1, 2 or a multiple of four sub-blocks");
  if constexpr (MT == 1) {
    sym_diffs1<S>(T.x, T.y, T.z, xj, yj, zj, dx, dy, dz);
  } else if constexpr (MT == 2) {
    sym_diffs2<S>(T.x, T.y, T.z, xj, yj, zj, dx, dy, dz);
  } else {
#pragma unroll
    for (int g = 0; g < MT; g += 4)
      sym_diffs4<S>(T.x + g, T.y + g, T.z + g, xj, yj, zj, dx + g, dy + g, dz + g);
  }
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

// The same pair step with every full-rate operand in a VGPR: the source
// coordinates are rotated by one lane per step with the reactions (three more
// v_mov_b32_dpp, shared by the MT target sub-blocks) instead of riding on the
// subtractions, and the law's constants live in VGPRs (SymLaw filled through
// vgpr()).  Free of the DPP asm blocks, the compiler fits 8 target sub-blocks
// in 127 VGPRs (4 waves per SIMD) where the DPP form takes 145 (3 waves per
// SIMD); at equal occupancy the two forms run at the same rate (the register
// tile alone, `bench.py --issue-probe`), so occupancy is what makes it faster.
template <int MT, int CUT>
__device__ __forceinline__ void sym_step_v(SymT<MT>& T, float& xs, float& ys, float& zs, float& bx, float& by,
                                           float& bz, const SymLaw& L, float eps) {
#pragma unroll
  for (int k = 0; k < MT; ++k) {
    const float dx = xs - T.x[k];
    const float dy = ys - T.y[k];
    const float dz = zs - T.z[k];
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps)));
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
  xs = dpp_ror<1>(xs);
  ys = dpp_ror<1>(ys);
  zs = dpp_ror<1>(zs);
  bx = dpp_ror<1>(bx);
  by = dpp_ror<1>(by);
  bz = dpp_ror<1>(bz);
  pin(xs);
  pin(ys);
  pin(zs);
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

// a copy of v the compiler must keep in a VGPR (its uses then take VGPR operands)
__device__ __forceinline__ float vgpr(float v) {
  asm volatile("" : "+v"(v));
  return v;
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

template <int MT, int CUT>
__device__ __forceinline__ void sym_phase_v(SymT<MT>& T, float& xs, float& ys, float& zs, float& bx, float& by,
                                            float& bz, const SymLaw& L, float eps) {
#pragma unroll
  for (int st = 0; st < 16; ++st) sym_step_v<MT, CUT>(T, xs, ys, zs, bx, by, bz, L, eps);
}

// v rotated by one lane across the whole wave (DPP wave_ror:1): lane i
// receives lane i - 1's value, lane 0 lane 63's.
__device__ __forceinline__ float wave_ror1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, v), __builtin_bit_cast(int, v),
                                                               0x13C, 0xF, 0xF, false));
}

// The 64 x 64 pairs of a wave's source sub-block with the sources read from
// LDS instead of rotated through registers.  The wave's private LDS region
// holds its 64 sources twice (ls[j] = ls[j + 64] = source j); at step s lane
// l reads source (l - s) mod 64 at ls[l + 64 - s] (the step in the immediate
// offset), and the reaction sums travel with the sources: one wave_ror:1 per
// component and step, so after 64 steps lane l holds the reaction on its own
// source.  Per step: MT pair terms, 3 DPP moves, one LDS read.
// PF: read step s + 1's source while step s is evaluated (3 more VGPRs and
// one address instruction per step).  Measured on MI355X: the all-pairs
// kernel with 8 target sub-blocks went from 276.2 to 269.9 ms at 2^20; the
// cell kernel with 2 target sub-blocks ran slower with it, 52.2 -> 57.0 ms
// at 2^22, and does without.
// ls: this lane's entry ls + lane + 64.  (bx, by, bz): added to.
template <int MT, int CUT, bool PF = false>
__device__ __forceinline__ void sym_tile_lds(SymT<MT>& T, const float4* ls, float& bx, float& by, float& bz,
                                             const SymLaw& L) {
  const float eps = vgpr(EPS32);
  // PF: the read after the last step, ls[-64], is the lane's own entry
  // ls[lane] (in bounds, unused)
  float4 qn = ls[0];
  for (int s0 = 0; s0 < 64; s0 += 16) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      float4 q;
      if constexpr (PF) {
        q = qn;
        qn = ls[-(s0 + s + 1)];
      } else {
        q = ls[-(s0 + s)];
      }
#pragma unroll
      for (int k = 0; k < MT; ++k) {
        const float dx = q.x - T.x[k];
        const float dy = q.y - T.y[k];
        const float dz = q.z - T.z[k];
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps)));
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
      bx = wave_ror1(bx);
      by = wave_ror1(by);
      bz = wave_ror1(bz);
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
  }
}

// The 64 x 64 pairs of the wave's source sub-block (xj, yj, zj, one per lane)
// with its MT target sub-blocks: four phases of 16 rotation steps, rows
// exchanging sources between phases.  On return (bx, by, bz) hold the
// reaction on the lane's own source (added to what they held) and T.a* the
// targets' accumulators.
// VG: the VGPR-operand step (sym_step_v; L's constants must be VGPRs).
template <int MT, int CUT, int VG = 0>
__device__ __forceinline__ void sym_tile(SymT<MT>& T, float xj, float yj, float zj, float& bx, float& by, float& bz,
                                         const SymLaw& L) {
  if constexpr (VG) {
    const float eps = vgpr(EPS32);
    sym_phase_v<MT, CUT>(T, xj, yj, zj, bx, by, bz, L, eps);
    sym_xchg(16, xj, yj, zj, bx, by, bz);
    sym_phase_v<MT, CUT>(T, xj, yj, zj, bx, by, bz, L, eps);
    sym_xchg(32, xj, yj, zj, bx, by, bz);
    sym_phase_v<MT, CUT>(T, xj, yj, zj, bx, by, bz, L, eps);
    sym_xchg(16, xj, yj, zj, bx, by, bz);
    sym_phase_v<MT, CUT>(T, xj, yj, zj, bx, by, bz, L, eps);
  } else {
    const auto seq = std::make_integer_sequence<int, 16>{};
    sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
    sym_xchg(16, xj, yj, zj, bx, by, bz);
    sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
    sym_xchg(32, xj, yj, zj, bx, by, bz);
    sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
    sym_xchg(16, xj, yj, zj, bx, by, bz);
    sym_phase<MT, CUT>(seq, T, xj, yj, zj, bx, by, bz, L);
  }
  bx = __shfl_xor(bx, 32, 64);
  by = __shfl_xor(by, 32, 64);
  bz = __shfl_xor(bz, 32, 64);
}

}  // namespace od
