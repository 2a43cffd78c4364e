// od_issue_probe.hip -- VALU issue-rate probe for gfx950 (measurement only,
// not part of the engine library; `python bench.py --issue-probe` builds and
// runs it).  Every kernel runs 8 waves per SIMD over all CUs and loops `iters`
// times over a fixed instruction mix per lane:
//   fma<NF>      NF independent v_fma_f32 (8 chains): the full-rate ceiling
//   trans<OP>    3 transcendental ops (exp2 / rcp / sqrt) on 4 chains, alone
//                and after 16 FMAs
//   law<T>       the pair term of the ion-ion kernels for 4 targets per lane
//                and a source that moves every iteration: 3 sub, 3 fma (r^2),
//                sqrt, exp2, fma + rcp, fma, 2 mul, 3 fma (accumulation) per
//                pair; law<0> replaces the three transcendentals by full-rate
//                ops, so the two times give their cost in that mix.
//   form<F>      16 register-only instructions of one encoding per iteration
//                (8 chains), written in asm so the compiler cannot change the
//                form: the issue cost of VGPR, SGPR, DPP and transcendental
//                operands (cycles per wave64 instruction per SIMD).
//   sym<VG>      the symmetric kernels' register tile (od_sym_tile.h) alone:
//                8 target sub-blocks per lane against a 64-source sub-block
//                per iteration, no memory traffic; VG = 1 the VGPR-operand
//                step, 0 the DPP/SGPR-operand step.  Its pair rate is the
//                ceiling of k_pair32_sym's inner loop.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "od_sym_tile.h"

template <int NF, int NT, int OP>
__global__ __launch_bounds__(256) void probe_mix(float* out, int iters, float a) {
  float f[8], t[4];
#pragma unroll
  for (int k = 0; k < 8; ++k) f[k] = 0.001f * (threadIdx.x + k);
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = 0.25f + 0.001f * (threadIdx.x + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < NF; ++k) f[k % 8] = __builtin_fmaf(f[k % 8], a, 0.5f);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if constexpr (OP == 0) t[k % 4] = __builtin_amdgcn_exp2f(-t[k % 4]);
      if constexpr (OP == 1) t[k % 4] = __builtin_amdgcn_rcpf(t[k % 4] + 1.0f);
      if constexpr (OP == 2) t[k % 4] = __builtin_amdgcn_sqrtf(t[k % 4]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += f[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) s += t[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int TRANS>
__global__ __launch_bounds__(256) void probe_law(float* out, int iters, float a, float c) {
  float tx[4], ty[4], tz[4], ax[4], ay[4], az[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    tx[k] = 0.01f * (threadIdx.x + 7 * k);
    ty[k] = 0.02f * (threadIdx.x + 3 * k);
    tz[k] = -0.015f * (threadIdx.x + 5 * k);
    ax[k] = ay[k] = az[k] = 0.f;
  }
  float xs = 0.3f + 0.001f * blockIdx.x, ys = -0.2f, zs = 0.1f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float dx = xs - tx[k], dy = ys - ty[k], dz = zs - tz[k];
      const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, 1e-30f)));
      float r, e, w;
      if constexpr (TRANS) {
        r = __builtin_amdgcn_sqrtf(r2);
        e = __builtin_amdgcn_exp2f(-r);
        w = __builtin_amdgcn_rcpf(__builtin_fmaf(a, r, r2));
      } else {
        r = __builtin_fmaf(r2, 0.5f, 0.5f);
        e = __builtin_fmaf(r, -0.25f, 1.0f);
        w = __builtin_fmaf(__builtin_fmaf(a, r, r2), -0.125f, 1.0f);
      }
      const float t = __builtin_fmaf(r, w, c);
      const float h = e * w * t;
      ax[k] = __builtin_fmaf(h, dx, ax[k]);
      ay[k] = __builtin_fmaf(h, dy, ay[k]);
      az[k] = __builtin_fmaf(h, dz, az[k]);
    }
    xs += 1e-4f;
    ys -= 2e-4f;
    zs += 3e-4f;
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) s += ax[k] + ay[k] + az[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

#define OD_FORM16(I)                                                                          \
  asm volatile(I(0) I(1) I(2) I(3) I(4) I(5) I(6) I(7) I(0) I(1) I(2) I(3) I(4) I(5) I(6) I(7) \
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)  \
               : "v"(b), "v"(c), "s"(sb))
#define OD_FMAC_V(k) "v_fmac_f32_e32 %" #k ", %8, %9\n"
#define OD_FMA_V(k) "v_fma_f32 %" #k ", %" #k ", %8, %9\n"
#define OD_MUL_V(k) "v_mul_f32_e32 %" #k ", %8, %" #k "\n"
#define OD_SUB_DPP(k) "v_sub_f32_dpp %" #k ", %8, %" #k " row_ror:1 row_mask:0xf bank_mask:0xf\n"
#define OD_EXP(k) "v_exp_f32_e64 %" #k ", -%" #k "\n"
#define OD_FMA_S(k) "v_fma_f32 %" #k ", %" #k ", %10, 0.5\n"
#define OD_FMAC_S(k) "v_fmac_f32_e32 %" #k ", %10, %9\n"
#define OD_MIX(k) "v_fmac_f32_e32 %" #k ", %8, %9\n"
// 12 fmac + 4 exp: chains 3 and 7 take the exponentials
#define OD_FORM_MIX                                                                                          \
  asm volatile(OD_FMAC_V(0) OD_FMAC_V(1) OD_FMAC_V(2) OD_EXP(3) OD_FMAC_V(4) OD_FMAC_V(5) OD_FMAC_V(6) OD_EXP(7) \
                   OD_FMAC_V(0) OD_FMAC_V(1) OD_FMAC_V(2) OD_EXP(3) OD_FMAC_V(4) OD_FMAC_V(5) OD_FMAC_V(6)   \
                       OD_EXP(7)                                                                             \
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)            \
               : "v"(b), "v"(c), "s"(sb))

template <int F>
__global__ __launch_bounds__(256) void probe_form(float* out, int iters, float sb) {
  float a0 = 0.1f + threadIdx.x * 1e-3f, a1 = a0 + 0.01f, a2 = a0 + 0.02f, a3 = a0 + 0.03f, a4 = a0 + 0.04f,
        a5 = a0 + 0.05f, a6 = a0 + 0.06f, a7 = a0 + 0.07f;
  const float b = 0.9999f, c = 0.25f;
  for (int i = 0; i < iters; ++i) {
    if constexpr (F == 0) OD_FORM16(OD_FMAC_V);   // v_fmac_f32_e32, VGPR operands
    if constexpr (F == 1) OD_FORM16(OD_FMA_V);    // v_fma_f32 (VOP3), VGPR operands
    if constexpr (F == 2) OD_FORM16(OD_MUL_V);    // v_mul_f32_e32, VGPR operands
    if constexpr (F == 3) OD_FORM16(OD_SUB_DPP);  // v_sub_f32 with a DPP row rotation
    if constexpr (F == 4) OD_FORM16(OD_EXP);      // v_exp_f32
    if constexpr (F == 5) OD_FORM16(OD_FMA_S);    // v_fma_f32 with an SGPR and an inline constant
    if constexpr (F == 6) OD_FORM_MIX;            // 12 v_fmac_f32 (VGPR) + 4 v_exp_f32
    if constexpr (F == 7) OD_FORM16(OD_FMAC_S);   // v_fmac_f32_e32 with an SGPR operand
  }
  out[blockIdx.x * 256 + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

template <int VG>
__global__ __launch_bounds__(128) void probe_sym(float* out, int iters, float a, float linv) {
  using namespace od;
  SymT<8> T;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    T.x[m] = 0.01f * (threadIdx.x + 64 * m);
    T.y[m] = -0.007f * (threadIdx.x + 32 * m);
    T.z[m] = 0.005f * (threadIdx.x + 16 * m);
    T.ax[m] = T.ay[m] = T.az[m] = 0.f;
  }
  const SymLaw L = VG ? SymLaw{vgpr(a), vgpr(linv), 0.f, 0.f, 0.f} : SymLaw{a, linv, 0.f, 0.f, 0.f};
  const int lane = threadIdx.x & 63;
  // VG == 2: each wave's sources twice in LDS (sym_tile_lds, read one step ahead)
  __shared__ float4 lsrc[VG == 2 ? 2 : 1][VG == 2 ? 128 : 1];
  float xj = 0.3f + 1e-3f * lane, yj = -0.2f + 2e-3f * lane, zj = 0.1f - 1e-3f * lane;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int i = 0; i < iters; ++i) {
    float bx = 0.f, by = 0.f, bz = 0.f;
    if constexpr (VG == 2) {
      float4* ls = lsrc[threadIdx.x >> 6];
      ls[lane] = ls[lane + 64] = make_float4(xj, yj, zj, 0.f);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      sym_tile_lds<8, 0, true>(T, ls + lane + 64, bx, by, bz, L);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    } else {
      sym_tile<8, 0, VG>(T, xj, yj, zj, bx, by, bz, L);
    }
    sx += bx;
    sy += by;
    sz += bz;
    xj += 1e-4f;
    yj -= 2e-4f;
    zj += 3e-4f;
  }
  float s = sx + sy + sz;
#pragma unroll
  for (int m = 0; m < 8; ++m) s += T.ax[m] + T.ay[m] + T.az[m];
  out[blockIdx.x * 128 + threadIdx.x] = s;
}

template <class K>
static double time_ms(K launch, int reps) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) launch();
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

int main() {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 1;
  const int blocks = p.multiProcessorCount * 8;  // 256-thread blocks: 32 waves per CU, 8 per SIMD
  const int iters = 20000, reps = 5;
  const double waves = (double)blocks * 4;
  float* d = nullptr;
  // output: one float per thread of the largest launch (form and mix: blocks x 256; sym: at most 12 x CUs x 128)
  if (hipMalloc(&d, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
  printf("{\"device\": \"%s\", \"cus\": %d, \"clock_khz\": %d, \"lds_per_cu\": %zu, \"waves_per_simd\": 8, \"iters\": %d,\n",
         p.gcnArchName, p.multiProcessorCount, p.clockRate, (size_t)p.maxSharedMemoryPerMultiProcessor, iters);
  auto mix = [&](const char* name, auto kern, int n_instr, bool last = false) {
    const double ms = time_ms([&]() { hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, d, iters, 0.999f); },
                              reps);
    printf(" \"%s\": {\"ms\": %.4f, \"wave_instr_per_s\": %.5e}%s\n", name, ms, waves * iters * n_instr / (ms * 1e-3),
           last ? "" : ",");
  };
  mix("fma16", probe_mix<16, 0, 0>, 16);
  mix("fma19", probe_mix<19, 0, 0>, 19);
  mix("exp4", probe_mix<0, 4, 0>, 4);
  mix("rcp4", probe_mix<0, 4, 1>, 4);
  mix("sqrt4", probe_mix<0, 4, 2>, 4);
  mix("fma16_exp3", probe_mix<16, 3, 0>, 19);
  mix("fma16_rcp3", probe_mix<16, 3, 1>, 19);
  mix("fma16_sqrt3", probe_mix<16, 3, 2>, 19);
  for (int tr = 0; tr < 2; ++tr) {
    const double ms = time_ms(
        [&]() {
          if (tr)
            hipLaunchKernelGGL(probe_law<1>, dim3(blocks), dim3(256), 0, 0, d, iters, 1.44e-3f, 0.6931f);
          else
            hipLaunchKernelGGL(probe_law<0>, dim3(blocks), dim3(256), 0, 0, d, iters, 1.44e-3f, 0.6931f);
        },
        reps);
    printf(" \"law_%s\": {\"ms\": %.4f, \"pairs_per_s\": %.5e},\n", tr ? "trans" : "fullrate", ms,
           waves * 64 * iters * 4 / (ms * 1e-3));
  }
  // instruction forms, 8 waves per SIMD
  const double simd_hz = 4.0 * p.multiProcessorCount * p.clockRate * 1e3;
  auto form = [&](const char* name, auto kern) {
    const double ms =
        time_ms([&]() { hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, d, iters, 0.9999f); }, reps);
    const double rate = waves * iters * 16 / (ms * 1e-3);
    printf(" \"%s\": {\"ms\": %.4f, \"wave_instr_per_s\": %.5e, \"cycles_per_wave_instr_per_simd\": %.3f},\n",
           name, ms, rate, simd_hz / rate);
  };
  form("form_fmac_vgpr", probe_form<0>);
  form("form_fma_vop3_vgpr", probe_form<1>);
  form("form_mul_vgpr", probe_form<2>);
  form("form_sub_dpp", probe_form<3>);
  form("form_exp", probe_form<4>);
  form("form_fma_sgpr_inline", probe_form<5>);
  form("form_12fmac_4exp", probe_form<6>);
  form("form_fmac_sgpr", probe_form<7>);
  // the symmetric register tile: 2-wave blocks, exactly 2 wps blocks per CU
  // (wps waves per SIMD): each block reserves LDS it does not use, so that no
  // more than 2 wps blocks fit in a CU's LDS, and the grid holds 2 wps blocks
  // per CU (the forms use at most 80 VGPRs, so up to 6 waves per SIMD fit);
  // the LDS-source form's own 4 KiB of sources come out of that reserve
  const int sym_iters = 100;
  const char* sym_names[3] = {"dpp_sgpr", "vgpr", "lds_sources"};
  for (int vg = 0; vg < 3; ++vg) {
    printf(" \"sym_tile_%s\": {", sym_names[vg]);
    const int wps_list[3] = {3, 4, 6};
    for (int w = 0; w < 3; ++w) {
      const int per_cu = 2 * wps_list[w];
      const int sym_blocks = p.multiProcessorCount * per_cu;
      const size_t lds = (((size_t)p.maxSharedMemoryPerMultiProcessor / per_cu) & ~(size_t)1023) -
                         (vg == 2 ? 2 * 128 * sizeof(float4) : 0);
      const double ms = time_ms(
          [&]() {
            if (vg == 2)
              hipLaunchKernelGGL(probe_sym<2>, dim3(sym_blocks), dim3(128), lds, 0, d, sym_iters, 0.05f, 0.6931f);
            else if (vg == 1)
              hipLaunchKernelGGL(probe_sym<1>, dim3(sym_blocks), dim3(128), lds, 0, d, sym_iters, 0.05f, 0.6931f);
            else
              hipLaunchKernelGGL(probe_sym<0>, dim3(sym_blocks), dim3(128), lds, 0, d, sym_iters, 0.05f, 0.6931f);
          },
          reps);
      const double unordered = (double)sym_blocks * 2 * sym_iters * 64.0 * 8 * 64;
      printf("\"%d_waves_per_simd\": {\"ms\": %.4f, \"unordered_pairs_per_s\": %.5e, \"lds_per_block\": %zu}%s",
             wps_list[w], ms, unordered / (ms * 1e-3), lds, w < 2 ? ", " : "");
    }
    printf("}%s\n", vg == 2 ? "" : ",");
  }
  printf("}\n");
  (void)hipFree(d);
  return 0;
}
