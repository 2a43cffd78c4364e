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
#include <hip/hip_runtime.h>

#include <cstdio>

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
  if (hipMalloc(&d, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
  printf("{\"device\": \"%s\", \"cus\": %d, \"clock_khz\": %d, \"waves_per_simd\": 8, \"iters\": %d,\n", p.gcnArchName,
         p.multiProcessorCount, p.clockRate, iters);
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
    printf(" \"law_%s\": {\"ms\": %.4f, \"pairs_per_s\": %.5e}%s\n", tr ? "trans" : "fullrate", ms,
           waves * 64 * iters * 4 / (ms * 1e-3), tr ? "" : ",");
  }
  printf("}\n");
  (void)hipFree(d);
  return 0;
}
