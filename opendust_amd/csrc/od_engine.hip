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
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "od_device.h"
#include "od_sym_tile.h"

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

// per-device address of the allocation-free entry points' scratch (defined
// with od_density_profile_noalloc below)
static int scratch_address(int dev, double** out);

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
  // cell grid (nm): columns of col_w in x and y, slabs of slab_h in z
  double col_w, slab_h;
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

#include "od_kernels_pair.inc"
#include "od_kernels_step.inc"
#include "od_kernels_trap.inc"
#include "od_kernels_cell.inc"

// Partial memory of one band of the symmetric schedule of a single context
// (default of od_config.sym_band_mib): rows are grouped into consecutive bands
// of at most this many bytes of partials, launched and reduced one after
// another (one band, the usual case, is one launch and one reduction).
// Batches run one band.
constexpr double SYM_BAND_BYTES = 48e9;

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
  // per-stage timing of one step (OD_RUN_PROFILE_STAGES): events recorded on
  // the stream at the start of each stage, kst[i] = stage of mark i
  static constexpr int KMARKS = 24;
  bool kprof = false;
  int kn = 0;
  hipEvent_t kev[KMARKS] = {};
  int kst[KMARKS] = {};
  double stage_ms[OD_N_STAGES] = {};
  int64_t stage_steps = 0;
  // symmetric pair kernel; bands of rows: band b = rows [band_row[b],
  // band_row[b+1]), its work-groups at launch-table and slot offset band_slot[b]
  int sym = 0, n_blocks = 0, sym_ch = 0, n_wg = 0, sym_var = 0, sym_R = 1;  // sym_R: target blocks per row group
  double band_cap = SYM_BAND_BYTES;
  int2* wg_tab = nullptr;
  int* wg_base = nullptr;
  double *sym_tp = nullptr, *sym_dg = nullptr, *rowD = nullptr, *sym_acc = nullptr;
  float* sym_sp = nullptr;
  std::vector<int> band_row, band_slot;
  // cutoff mode: cell list; Gcut: flagged sources' share of the drift-state force;
  // sc_tp, sc_sp: partials of the symmetric cell kernel (FP32c)
  double* Gcut = nullptr;
  double* sc_tp = nullptr;
  float* sc_sp = nullptr;
  int *scan_tsum = nullptr, *scan_toff = nullptr;
  int cut = 0, n_cells = 0, n_cols = 0, rs_nblk = 0, max_items = 0, rs_passes = 2;
  int nx = 0, ny = 0, nz = 0;  // cell grid
  int sc_var = 0;  // symmetric cell kernel variant (index into cell_variant_mt)
  uint32_t *ckey[2] = {nullptr, nullptr}, *cval[2] = {nullptr, nullptr};
  uint32_t *skey = nullptr, *sval = nullptr;  // sorted keys/values: the buffers the last radix pass wrote
  int *rs_hist = nullptr, *rs_off = nullptr, *rs_total = nullptr, *cell_start = nullptr,
      *cell_chunks = nullptr, *item_off = nullptr, *n_items = nullptr;
  int2* items = nullptr;
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
    P->col_w = c->cutoff / 4.0;
    P->slab_h = c->cutoff / 16.0;
    const double nxd = ceil(2.0 * c->R / P->col_w), nzd = ceil(c->H / P->slab_h);
    // sizes beyond the cell maximum are rejected by the caller (setup_cells)
    const bool fits = nxd * nxd * nzd <= (double)OD_MAX_CELLS;
    P->nx = fits ? (int)nxd : 1;
    P->ny = P->nx;
    P->nz = fits ? (int)nzd : 1;
    P->n_cells = fits ? P->nx * P->ny * P->nz : -1;
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
  (void)hipSetDevice(c->dev);  // teardown: nothing to report a failure to
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  void* bufs[] = {c->dP, c->X, c->V, c->F, c->Fpart, c->q, c->partials, c->rowF, c->pi, c->pj,
                  c->list_old, c->list_new, c->flags, c->st, c->face, c->block_cnt, c->block_off, c->n_flag,
                  c->slot_of, c->list_ion, c->err, c->rec, c->wg_tab, c->wg_base, c->sym_tp, c->sym_dg,
                  c->rowD, c->sym_sp, c->sym_acc, c->Gcut, c->scan_tsum, c->scan_toff, c->sc_tp, c->sc_sp, c->ckey[0], c->ckey[1], c->cval[0], c->cval[1], c->rs_hist,
                  c->rs_off, c->rs_total, c->cell_start, c->cell_chunks, c->item_off, c->items,
                  c->n_items, c->spos};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->kev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
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
// work-groups of the schedule with `ch` source blocks per work-group and row
// groups of R blocks (group g: rows g R ..; sources from its first row + 1)
static int64_t sym_wg_count(const std::vector<int>& be, int ch, int R = 1) {
  int64_t n = 0;
  for (int X = 0; X < (int)be.size(); X += R) n += std::max(0, (be[X] - 1 - X + ch - 1) / ch);
  return n;
}

static double sym_bytes(const std::vector<int>& be, int ch, int R = 1) {
  return (double)sym_wg_count(be, ch, R) * 3 * SYM_TB * (4.0 * ch + 8.0 * R);
}


// Launch order of one band's work-groups (indices into the group-major table
// `tab`), for a device whose dispatcher deals work-group i to XCD i mod n_xcc.
// * Body: the work-groups with the full R ch (target block, source block)
//   pairs.  Cut into tiles of SYM_TILE_BLOCKS target blocks (row groups
//   g R ..) by SYM_TILE_BLOCKS source blocks (by each work-group's first
//   source block J0), tiles in row-major order, (g, k) order inside a tile;
//   that sequence is split into n_xcc contiguous runs of equal count and XCD x
//   is dealt run x.  The work-groups an XCD holds at a time then share their
//   target and source blocks (about 2 MB of positions per tile against a 4 MB
//   L2 per XCD), so a block is fetched past the L2 about once per tile instead
//   of once per work-group; equal counts of equal work keep the XCDs level.
// * Tail: the partial work-groups (the first of a row group, where rows skip
//   the source blocks J <= I, and the last, clipped at the end of the
//   simulation), longest first, so they fill the end of the launch.
// A list-scheduling model (in-order dispatch to 8 XCDs x 64 resident
// work-groups, time proportional to pairs) puts this
// order 0.7 % above the ideal makespan at 2^20 with R = 4 and 8 source blocks
// per work-group (longest-first in group-major order: 0.9 %; the tiles
// without the separate tail: 3.2 %).  The order does not change what any
// work-group computes or where it writes (slots follow the table).
static std::vector<int2> sym_launch_order(const std::vector<int2>& tab, const std::vector<int>& be, int w0, int w1,
                                          int ch, int R, int n_xcc) {
  constexpr int SYM_TILE_BLOCKS = 64;
  const int tile_groups = std::max(1, SYM_TILE_BLOCKS / R);
  const int tile_cols = std::max(ch, SYM_TILE_BLOCKS / ch * ch);
  const int NB = (int)be.size();
  // (target block, source block) pairs the work-group evaluates
  auto work = [&](int i) {
    const int X = tab[i].x * R, J0 = X + 1 + tab[i].y * ch, J1 = std::min(J0 + ch, be[X]);
    int w = 0;
    for (int r = 0; r < R && X + r < NB; ++r) w += std::max(0, J1 - std::max(J0, X + r + 1));
    return w;
  };
  std::vector<int> body, tail;
  std::vector<int> wt(w1 - w0);
  for (int i = w0; i < w1; ++i) {
    wt[i - w0] = work(i);
    (wt[i - w0] == R * ch ? body : tail).push_back(i);
  }
  auto key = [&](int i) {
    const int g = tab[i].x, j0 = g * R + 1 + tab[i].y * ch;
    return std::make_pair(g / tile_groups, j0 / tile_cols);
  };
  std::stable_sort(body.begin(), body.end(), [&](int a, int b) { return key(a) < key(b); });
  std::stable_sort(tail.begin(), tail.end(), [&](int a, int b) { return wt[a - w0] > wt[b - w0]; });
  const int n = (int)body.size(), nx = std::max(1, n_xcc);
  std::vector<int> off(nx + 1, 0);
  for (int x = 0; x < nx; ++x) off[x + 1] = off[x] + n / nx + (x < n % nx);
  std::vector<int2> out;
  out.reserve(w1 - w0);
  for (int j = 0; j <= n / nx; ++j)
    for (int x = 0; x < nx; ++x)
      if (off[x] + j < off[x + 1]) out.push_back(tab[body[off[x] + j]]);
  for (int i : tail) out.push_back(tab[i]);
  return out;
}

// XCDs of device `dev` (1 where the attribute is unavailable)
static int device_xcc_count(int dev) {
  int n = 1;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || n < 1) {
    (void)hipGetLastError();  // clear the failed query; one XCD is a valid order
    n = 1;
  }
  return n;
}

// (Re)builds a context's schedule for `ch` source blocks per work-group and
// row groups of R target blocks (R > 1: one simulation only): partial slots in
// group-major (g, k) order (wg_base, indexed by group); groups gathered into
// bands of at most band_cap bytes of partials; each band launched in the
// XCD-aware tile order of sym_launch_order.
// C: od_ctx or od_batch.
template <class C>
static int sym_build(C* c, const std::vector<int>& be, int ch, double band_cap, int R) {
  const int NB = (int)be.size();
  const int NG = (NB + R - 1) / R;
  const int64_t n_wg = sym_wg_count(be, ch, R);
  if (n_wg > (1ll << 31) - 1) return set_err(OD_ERR_ARG, "symmetric pair kernel: too many workgroups");
  std::vector<int2> tab;
  std::vector<int> base(NG + 1);
  tab.reserve(n_wg);
  for (int g = 0; g < NG; ++g) {
    const int X = g * R;
    base[g] = (int)tab.size();
    const int nk = std::max(0, (be[X] - 1 - X + ch - 1) / ch);
    for (int k = 0; k < nk; ++k) tab.push_back(make_int2(g, k));
  }
  base[NG] = (int)tab.size();
  // bands of row groups: each holds at least one group and at most band_cap
  // bytes of partials unless a single group exceeds it
  const double wg_bytes = 3.0 * SYM_TB * (4.0 * ch + 8.0 * R);
  std::vector<int> band_row{0};
  for (int g = 0; g < NG; ++g) {
    const double rows_bytes = (double)(base[g + 1] - base[band_row.back()]) * wg_bytes;
    if (g > band_row.back() && rows_bytes > band_cap) band_row.push_back(g);
  }
  band_row.push_back(NG);
  std::vector<int2> launch;
  launch.reserve(tab.size());
  int max_band = 1;
  const int n_xcc = device_xcc_count(c->dev);
  for (size_t b = 0; b + 1 < band_row.size(); ++b) {
    const int w0 = base[band_row[b]], w1 = base[band_row[b + 1]];
    const std::vector<int2> order = sym_launch_order(tab, be, w0, w1, ch, R, n_xcc);
    launch.insert(launch.end(), order.begin(), order.end());
    max_band = std::max(max_band, w1 - w0);
  }
  void* old[] = {c->wg_tab, c->wg_base, c->sym_tp, c->sym_sp};
  for (void* q : old)
    if (q) (void)hipFree(q);
  c->wg_tab = nullptr;
  c->wg_base = nullptr;
  c->sym_tp = nullptr;
  c->sym_sp = nullptr;
  const size_t nw = std::max<size_t>(launch.size(), 1), nb_w = (size_t)max_band;
  int rc = OD_OK;
  if ((rc = dalloc(&c->wg_tab, nw * sizeof(int2))) || (rc = dalloc(&c->wg_base, base.size() * sizeof(int))) ||
      (rc = dalloc(&c->sym_tp, nb_w * R * 3 * SYM_TB * 8)) ||
      (rc = dalloc(&c->sym_sp, nb_w * 3 * ch * SYM_TB * 4)))
    return rc;
  c->band_row = band_row;
  c->band_slot.assign(band_row.size(), 0);
  for (size_t b = 0; b < band_row.size(); ++b) c->band_slot[b] = base[band_row[b]];
  if (!launch.empty()) HIPCHK(hipMemcpy(c->wg_tab, launch.data(), launch.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->wg_base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
  c->n_blocks = NB;
  c->sym_ch = ch;
  c->sym_R = R;
  c->n_wg = (int)n_wg;
  return OD_OK;
}

// default kernel variant with one target block per work-group (batches, and
// single simulations too small for row groups, see setup_sym): the
// configured one, else 8 target sub-blocks per wave (2 waves per work-group)
// with the sources read from LDS (variant 11) when the schedule has at least
// 2048 work-groups (4 waves per SIMD), else 4 (4 waves per work-group: at
// 2^15 ions the 496 work-groups then still give about 2 waves per SIMD).
// Measured on MI355X at 2^15 ions: 0.33 ms per evaluation with 4, 0.64 ms with
// 8; at 2^20, 4 source blocks per work-group: 308.5 ms with the DPP/SGPR
// operand step, 287.7 ms with the VGPR operand step, 278.8 ms with the
// sources from LDS.
static int sym_default_variant(int64_t n_wg, int configured) {
  if (configured > 0) return configured - 1;
  return n_wg >= 2048 ? 10 : 2;
}

// default source blocks per work-group: the largest of 4, 2, 1 that still
// gives >= 4096 work-groups (8 per XCD CU set), or the configured value.
// Measured at 2^20 with the default variant: 291.1 ms with 16, 287.7 with 4,
// 289.7 with 2.
static int sym_default_chunk(const std::vector<int>& be, int configured) {
  if (configured > 0) return configured;
  int ch = 4;
  while (ch > 1 && sym_wg_count(be, ch) < 4096) ch /= 2;
  return ch;
}

static int setup_cells(od_ctx* c, const od_config* cfg, const DevParams& hp) {
  c->cut = 0;
  if (!(cfg->cutoff > 0.0)) return OD_OK;
  if (hp.n_cells < 0)
    return set_err(OD_ERR_ARG, "cutoff %g nm in a cylinder R = %g, H = %g nm gives more than %d cells", cfg->cutoff,
                   cfg->R, cfg->H, OD_MAX_CELLS);
  c->cut = 1;
  c->n_cells = hp.n_cells;
  c->n_cols = hp.nx * hp.ny;
  c->nx = hp.nx;
  c->ny = hp.ny;
  c->nz = hp.nz;
  // 8-bit digits over the keys 0 .. n_cells (n_cells marks ions left out)
  c->rs_passes = 1;
  while (c->rs_passes < 4 && ((uint64_t)c->n_cells >> (8 * c->rs_passes)) != 0) ++c->rs_passes;
  c->rs_nblk = (c->n_pad + RS_ITEMS - 1) / RS_ITEMS;
  c->max_items = c->n_pad / 64 + c->n_cols + 1;
  const size_t np = c->n_pad;
  int rc = OD_OK;
  if ((rc = dalloc(&c->ckey[0], np * 4)) || (rc = dalloc(&c->ckey[1], np * 4)) || (rc = dalloc(&c->cval[0], np * 4)) ||
      (rc = dalloc(&c->cval[1], np * 4)) || (rc = dalloc(&c->rs_hist, (size_t)256 * c->rs_nblk * 4)) ||
      (rc = dalloc(&c->rs_off, (size_t)256 * c->rs_nblk * 4)) || (rc = dalloc(&c->rs_total, 4)) ||
      (rc = dalloc(&c->cell_start, (size_t)(c->n_cells + 1) * 4)) ||
      (rc = dalloc(&c->cell_chunks, (size_t)(c->n_cols + 1) * 4)) ||
      (rc = dalloc(&c->item_off, (size_t)(c->n_cols + 1) * 4)) ||
      (rc = dalloc(&c->items, (size_t)c->max_items * sizeof(int2))) || (rc = dalloc(&c->n_items, 4)) ||
      (rc = dalloc(&c->spos, np * c->pos_bytes)) || (rc = dalloc(&c->Gcut, 3 * np * 8)))
    return rc;
  c->skey = c->ckey[c->rs_passes & 1];
  c->sval = c->cval[c->rs_passes & 1];
  const size_t n_tiles = (std::max<size_t>((size_t)256 * c->rs_nblk, (size_t)c->n_cols + 1) + 1023) / 1024;
  if ((rc = dalloc(&c->scan_tsum, n_tiles * 4)) || (rc = dalloc(&c->scan_toff, n_tiles * 4))) return rc;
  if (c->prec == OD_PREC_FP32C) {
    if ((rc = dalloc(&c->sc_tp, (size_t)CC_NSET * 3 * np * 8)) ||
        (rc = dalloc(&c->sc_sp, (size_t)CC_NSET * 3 * np * 4)))
      return rc;
    c->sc_var = cell_variant_of(cfg->sym_tile);
    if (c->sc_var < 0) return set_err(OD_ERR_ARG, "cutoff mode: sym_tile %d outside [0, %d]", cfg->sym_tile, CC_N_VARIANTS);
  }
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
  int ch = sym_default_chunk(be, cfg->sym_chunk);
  c->band_cap = cfg->sym_band_mib > 0 ? cfg->sym_band_mib * 1048576.0 : SYM_BAND_BYTES;
  int rc = OD_OK;
  int var = sym_default_variant(sym_wg_count(be, ch), cfg->sym_tile);
  // default for large N: row groups of 4 target blocks with 8 source blocks
  // per work-group and the sources read from LDS (variant 12), when that
  // still gives >= 4096 work-groups of 8 waves.  Measured at 2^20: 279.3 ms
  // per launch (288.5 with the VGPR-operand step, variant 10), 4.3 GB of HBM
  // traffic per launch against 16.9 GB with one target block and 4 source
  // blocks per work-group.
  if (cfg->sym_tile == 0) {
    const int ch4 = cfg->sym_chunk > 0 ? cfg->sym_chunk : 8;
    if (sym_wg_count(be, ch4, 4) >= 4096) {
      var = SYM_GROUPED_DEFAULT;
      ch = ch4;
    }
  }
  if ((rc = sym_build(c, be, ch, c->band_cap, sym_variants[var].rows)) ||
      (rc = dalloc(&c->sym_dg, (size_t)3 * c->n_pad * 8)) || (rc = dalloc(&c->sym_acc, (size_t)3 * c->n_pad * 8)))
    return rc;
  c->sym_var = var;
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
  {
    // per-device address of the allocation-free entry points' scratch array
    double* sp = nullptr;
    if ((rc = scratch_address(c->dev, &sp))) return fail(rc);
  }
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

// start of a stage of the step (stage profiling only; stage -1 ends the step)
static void mark(od_ctx* c, int stage) {
  if (!c->kprof || c->kn >= od_ctx::KMARKS) return;
  if (hipEventRecord(c->kev[c->kn], c->stream) != hipSuccess) return;
  c->kst[c->kn++] = stage;
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
    const int nbands = (int)c->band_row.size() - 1;
    if (nbands == 1) {
      hipLaunchKernelGGL(V.fn, dim3(c->n_wg), dim3(V.rows * SYM_TB / V.mt), 0, c->stream, (const float4*)pj,
                         c->n_blocks, c->sym_ch, (const int2*)c->wg_tab, (const int*)c->wg_base, 0, a, linv,
                         (const int*)nullptr, (const float*)nullptr, c->sym_tp, c->sym_sp);
      if (timed) HIPCHK(hipEventRecord(c->ev1, c->stream));
      hipLaunchKernelGGL(k_pair32_diag, dim3(c->n_blocks * (SYM_TB / 256)), dim3(256), 0, c->stream,
                         (const float4*)pj, c->n_pad, a, linv, (const float*)nullptr, c->sym_dg);
      hipLaunchKernelGGL(k_sym_reduce, dim3(c->n_pad / 256), dim3(256), 0, c->stream, c->sym_dg, c->sym_tp,
                         c->sym_sp, c->wg_base, c->n_blocks, c->sym_ch, c->sym_R, c->n_pad, -C * L * L,
                         (const int*)nullptr, (const double*)nullptr, F);
      return OD_OK;
    }
    // bands of row groups, each launched and reduced into acc in order (the
    // timed region then holds the band reductions between the launches)
    hipLaunchKernelGGL(k_pair32_diag, dim3(c->n_blocks * (SYM_TB / 256)), dim3(256), 0, c->stream,
                       (const float4*)pj, c->n_pad, a, linv, (const float*)nullptr, c->sym_dg);
    for (int b = 0; b < nbands; ++b) {
      const int w0 = c->band_slot[b], nw = c->band_slot[b + 1] - w0;
      if (nw > 0)
        hipLaunchKernelGGL(V.fn, dim3(nw), dim3(V.rows * SYM_TB / V.mt), 0, c->stream, (const float4*)pj,
                           c->n_blocks, c->sym_ch, (const int2*)(c->wg_tab + w0), (const int*)c->wg_base, w0, a,
                           linv, (const int*)nullptr, (const float*)nullptr, c->sym_tp, c->sym_sp);
      if (timed && b == nbands - 1) HIPCHK(hipEventRecord(c->ev1, c->stream));
      hipLaunchKernelGGL(k_sym_band_reduce, dim3(c->n_pad / 256), dim3(256), 0, c->stream, (const double*)c->sym_dg,
                         (const double*)c->sym_tp, (const float*)c->sym_sp, (const int*)c->wg_base, c->band_row[b],
                         c->band_row[b + 1], w0, c->sym_ch, c->sym_R, c->n_pad, -C * L * L, b == 0 ? 1 : 0,
                         b == nbands - 1 ? 1 : 0, c->sym_acc, F);
    }
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
  for (int pass = 0; pass < c->rs_passes; ++pass) {
    const int in = pass & 1, out = 1 - in, shift = 8 * pass;
    hipLaunchKernelGGL(k_rs_hist, dim3(c->rs_nblk), dim3(256), 0, c->stream, c->ckey[in], np, shift, c->rs_nblk,
                       c->rs_hist);
    launch_scan(c, c->rs_hist, 256 * c->rs_nblk, c->rs_off, c->rs_total);
    hipLaunchKernelGGL(k_rs_scatter, dim3(c->rs_nblk), dim3(256), 0, c->stream, c->ckey[in], c->cval[in], np, shift,
                       c->rs_nblk, c->rs_off, c->ckey[out], c->cval[out]);
  }
  hipLaunchKernelGGL(k_cell_starts, dim3(nc / 256 + 1), dim3(256), 0, c->stream, c->skey, np, nc, c->cell_start);
  const int gc = (c->n_cols + 255) / 256;
  hipLaunchKernelGGL(k_cell_chunks, dim3(gc), dim3(256), 0, c->stream, c->cell_start, c->n_cols, c->nz,
                     c->cell_chunks);
  launch_scan(c, c->cell_chunks, c->n_cols, c->item_off, c->n_items);
  hipLaunchKernelGGL(k_cell_items, dim3(gc), dim3(256), 0, c->stream, c->cell_chunks, c->item_off, c->n_cols,
                     c->items);
  hipLaunchKernelGGL(k_cell_gather<T>, dim3(nb), dim3(256), 0, c->stream, c->skey, c->sval, np, nc,
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
  if (!c->cut) {
    mark(c, OD_STAGE_PAIR);
    return launch_pair(c, pos, pos, F, timed);
  }
  mark(c, OD_STAGE_CELLS);
  CHKRET(build_cells_t<T>(c, X, flags, pos));
  mark(c, OD_STAGE_PAIR);
  const int grid = std::max(1, std::min(4096, c->max_items / 4 + 1));
  const bool sw = c->cfg.switch_dist > 0.0 && c->cfg.switch_dist < c->cfg.cutoff;
  if (timed) HIPCHK(hipEventRecord(c->ev0, c->stream));
  if constexpr (std::is_same<T, float>::value) {
    // symmetric: forward neighbour columns and the own column, per-ion reduction
    const float4* sp4 = (const float4*)c->spos;
    const dim3 g(c->n_cols * CC_NSET);
    const int ki = c->sc_var * 2 + (sw ? 1 : 0);
    hipLaunchKernelGGL(cell_sym_kernels[ki], g, dim3(64), 0, c->stream, c->dP, sp4, c->skey, c->cell_start, c->sc_tp,
                       c->sc_sp);
    hipLaunchKernelGGL(k_symcol_reduce, dim3(c->n_pad / 256), dim3(256), 0, c->stream, c->dP, c->skey, c->sval,
                       c->cell_start, 64 * cell_variant_mt[c->sc_var], (const double*)c->sc_tp, (const float*)c->sc_sp, F);
  } else if (sw) {
    hipLaunchKernelGGL((k_pair_cell<T, 1>), dim3(grid), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->skey, c->sval, c->cell_start, c->items, c->n_items, F);
  } else {
    hipLaunchKernelGGL((k_pair_cell<T, 0>), dim3(grid), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->skey, c->sval, c->cell_start, c->items, c->n_items, F);
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
                       c->sval, c->cell_start, c->n_flag, list, add, out);
  else
    hipLaunchKernelGGL((k_corr_cell<T, 0>), dim3(c->nb), dim3(256), 0, c->stream, c->dP, (const PV*)c->spos,
                       c->sval, c->cell_start, c->n_flag, list, add, out);
}

template <class T>
static int launch_step_t(od_ctx* c, bool timed) {
  using PV = typename Pos<T>::V;
  const int nb = c->nb;
  mark(c, OD_STAGE_KICK_DRIFT);
  hipLaunchKernelGGL(k_kick_drift<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->X, c->V,
                     c->F, c->q, c->st, (PV*)c->pi, (PV*)c->pj, c->flags, c->block_cnt, c->err);
  mark(c, OD_STAGE_LIST);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c->stream, c->block_cnt, nb, c->block_off, c->n_flag);
  hipLaunchKernelGGL(k_list<T>, dim3(nb), dim3(256), 0, c->stream, c->flags, c->block_cnt, c->block_off, c->n,
                     (const PV*)c->pi, c->slot_of, c->list_ion, (PV*)c->list_old);
  // flagged ions are parked in pj, as sources and as targets (cutoff mode:
  // left out of the cell list)
  CHKRET(launch_forces_t<T>(c, c->X, c->pj, c->flags, c->F, timed));
  const uint32_t dust_bits = (c->cfg.n_dust > 0) ? ((1u << c->cfg.n_dust) - 1u) : 0u;
  mark(c, OD_STAGE_ROWS_ABSORBED);
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
  if (c->cut) {
    mark(c, OD_STAGE_CORR);
    launch_corr_cell<T>(c, (const PV*)c->list_old, 0, c->Gcut);
  }
  mark(c, OD_STAGE_KICK2);
  hipLaunchKernelGGL(k_kick2<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->block_off, nb,
                     c->X, c->V, c->F, c->q, c->st,
                     (const PV*)c->pi, (PV*)c->pj, c->flags, c->n_flag, c->slot_of, (const PV*)c->list_old,
                     (PV*)c->list_new, c->face, c->rowD, c->cut ? (const double*)c->Gcut : (const double*)nullptr,
                     c->partials);
  mark(c, OD_STAGE_ROWS_REINJECTED);
  if (c->cut)
    hipLaunchKernelGGL(k_rows_cell<T>, dim3(c->rows_grid), dim3(256), 0, c->stream, c->dP, c->X, (const PV*)c->spos,
                       c->cell_start, c->n_flag, (const PV*)c->list_new, (const PV*)c->list_new, c->flags,
                       c->list_ion, 0u, c->rowF);
  else
    hipLaunchKernelGGL(k_rows<T>, dim3(c->rows_grid), dim3(RowShape<T>::TB), 0, c->stream, c->dP, (const int*)nullptr,
                       (const PV*)c->pj, c->n_flag, (const PV*)c->list_new, c->flags, c->list_ion, 0u, c->rowF);
  if (c->cut) {
    mark(c, OD_STAGE_CORR);
    launch_corr_cell<T>(c, (const PV*)c->list_new, 1, c->F);
  }
  mark(c, OD_STAGE_INFLOW);
  hipLaunchKernelGGL(k_inflow<T>, dim3(nb), dim3(256), 0, c->stream, c->dP, (const int*)nullptr, c->block_off, nb,
                     c->F, (const PV*)c->pj, c->flags, c->n_flag, (const PV*)c->list_new, c->slot_of, c->rowF,
                     c->cut);
  mark(c, OD_STAGE_DUST);
  hipLaunchKernelGGL(k_dust, dim3(1), dim3(256), 0, c->stream, c->dP, c->partials, c->q, c->st, c->rec,
                     std::max(1, c->cfg.n_dust), 1);
  mark(c, -1);
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
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
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
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  if (c->graph) {
    (void)hipGraphDestroy(c->graph);
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
  const bool stages = (flags & OD_RUN_PROFILE_STAGES) != 0;
  const bool graph = (flags & OD_RUN_GRAPH) != 0 && !profile && !stages;
  if (stages)
    for (hipEvent_t& e : c->kev)
      if (!e) HIPCHK(hipEventCreate(&e));
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
      c->kprof = stages;
      c->kn = 0;
      const int rc = launch_step(c, profile);
      c->kprof = false;
      CHKRET(rc);
      HIPCHK(hipGetLastError());
      if (profile) {
        HIPCHK(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->pair_ms += ms;
        c->pair_launches += 1;
      }
      if (stages && c->kn > 0) {
        HIPCHK(hipEventSynchronize(c->kev[c->kn - 1]));
        for (int i = 0; i + 1 < c->kn; ++i) {
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, c->kev[i], c->kev[i + 1]));
          if (c->kst[i] >= 0) c->stage_ms[c->kst[i]] += ms;
        }
        c->stage_steps += 1;
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
  HIPCHK(hipMemcpyAsync(key.data(), c->skey, c->n_pad * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(val.data(), c->sval, c->n_pad * 4, hipMemcpyDeviceToHost, c->stream));
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

// Ordered pair terms the context's cell kernel evaluates on a cell list
// (cs: cell starts, key: sorted cell ids), padding lanes and slots included:
// the FP64 kernel evaluates 64 lanes x every source of the 77 slab ranges of
// each 64-target chunk; the symmetric FP32 kernel's slots each give two
// ordered terms (force and reaction), its own-group tiles one.
static int64_t cell_candidates(const od_ctx* c, const std::vector<int>& cs, const std::vector<uint32_t>& key) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  auto range = [&](int bx, int by, int blo, int bhi, int nbz, int64_t& s0, int64_t& s1) {
    s0 = s1 = 0;
    if (bx < 0 || bx >= nx || by < 0 || by >= ny) return;
    const int cb = nz * (bx + nx * by);
    s0 = cs[cb + std::max(blo - nbz, 0)];
    s1 = cs[cb + std::min(bhi + nbz, nz - 1) + 1];
  };
  auto up64 = [](int64_t v) { return (v + 63) / 64 * 64; };
  int64_t total = 0;
  for (int A = 0; A < c->n_cols; ++A) {
    const int cA = A * nz, a0 = cs[cA], a1 = cs[cA + nz];
    if (a0 == a1) continue;
    const int ax = A % nx, ay = A / nx;
    if (c->prec == OD_PREC_FP32C) {
      const int MTv = cell_variant_mt[c->sc_var];
      const bool split = cell_variant_split[c->sc_var];
      const int G = 64 * MTv;
      for (int k = 0; k < CC_NSET; ++k) {
        const bool own = k == CC_NFWD;
        int dx = 0, dy = 0;
        if (!own) cc_fwd(k, dx, dy);
        const int bx = ax + dx, by = ay + dy;
        if (bx < 0 || bx >= nx || by >= ny) continue;
        const int cb = nz * (bx + nx * by);
        if (!own && cs[cb] == cs[cb + nz]) continue;
        const int nbz = cc_nbz(dx, dy);
        for (int t0 = a0; t0 < a1; t0 += G) {
          const int t1 = std::min(t0 + G, a1);
          const int blo = (int)key[t0] - cA, bhi = (int)key[t1 - 1] - cA;
          int64_t s0, s1;
          range(bx, by, blo, bhi, nbz, s0, s1);
          if (own) {
            s0 = t1;
            total += (int64_t)G * up64(t1 - t0);
          }
          if (s1 <= s0) continue;
          if (!split) {
            total += 2 * (int64_t)G * up64(s1 - s0);
            continue;
          }
          // each 64-source chunk meets only the sub-blocks that reach it
          int b[4], e[4], lo = 0, hi = 0;
          if (MTv == 2)
            cc_sub_ranges<2>(cs.data(), key.data(), cA, cb, nbz, nz, own, t0, t1, b, e);
          else
            cc_sub_ranges<4>(cs.data(), key.data(), cA, cb, nbz, nz, own, t0, t1, b, e);
          for (int64_t jb = s0; jb < s1; jb += 64) {
            if (MTv == 2)
              cc_chunk_run<2>(b, e, (int)jb, lo, hi);
            else
              cc_chunk_run<4>(b, e, (int)jb, lo, hi);
            total += 2 * 64 * 64 * (int64_t)(hi - lo + 1);
          }
        }
      }
    } else {
      for (int t0 = a0; t0 < a1; t0 += 64) {
        const int tl = std::min(t0 + 63, a1 - 1);
        const int blo = (int)key[t0] - cA, bhi = (int)key[tl] - cA;
        for (int k = 0; k < CC_NOFF; ++k) {
          int dx, dy;
          cc_off(k, dx, dy);
          int64_t s0, s1;
          range(ax + dx, ay + dy, blo, bhi, cc_nbz(dx, dy), s0, s1);
          total += 64 * (s1 - s0);
        }
      }
    }
  }
  return total;
}

extern "C" int od_cutoff_pair_count(od_ctx* c, int64_t* in_range, int64_t* candidates) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (!c->cut) return set_err(OD_ERR_STATE, "od_cutoff_pair_count: context has no cutoff");
  HIPCHK(hipSetDevice(c->dev));
  unsigned long long* d = nullptr;
  CHKRET(dalloc(&d, 8));
  unsigned long long h = 0;
  hipError_t e = hipMemcpyAsync(d, &h, 8, hipMemcpyHostToDevice, c->stream);
  const int grid = std::max(1, std::min(4096, c->max_items / 4 + 1));
  if (e == hipSuccess) {
    if (c->prec == OD_PREC_FP32C)
      hipLaunchKernelGGL(k_cell_pair_count<float>, dim3(grid), dim3(256), 0, c->stream, c->dP,
                         (const float4*)c->spos, c->skey, c->cell_start, c->items, c->n_items, d);
    else
      hipLaunchKernelGGL(k_cell_pair_count<double>, dim3(grid), dim3(256), 0, c->stream, c->dP,
                         (const double4*)c->spos, c->skey, c->cell_start, c->items, c->n_items, d);
    e = hipGetLastError();
  }
  std::vector<int> cs(c->n_cells + 1);
  std::vector<uint32_t> key(c->n_pad);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(cs.data(), c->cell_start, cs.size() * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(key.data(), c->skey, key.size() * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_err(OD_ERR_HIP, "od_cutoff_pair_count: %s", hipGetErrorString(e));
  if (in_range) *in_range = (int64_t)h;
  if (candidates) *candidates = cell_candidates(c, cs, key);
  return OD_OK;
}

extern "C" int od_sym_choice(od_ctx* c, int32_t* sym_chunk, int32_t* sym_tile, int32_t* bands) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (c->cut && c->prec == OD_PREC_FP32C) {
    // symmetric cell kernel: no source blocks, the variant 1 .. CC_N_VARIANTS
    if (sym_chunk) *sym_chunk = 0;
    if (sym_tile) *sym_tile = c->sc_var + 1;
    if (bands) *bands = 0;
    return OD_OK;
  }
  if (sym_chunk) *sym_chunk = c->sym ? c->sym_ch : 0;
  if (sym_tile) *sym_tile = c->sym ? c->sym_var + 1 : 0;
  if (bands) *bands = c->sym ? (int32_t)c->band_row.size() - 1 : 0;
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

extern "C" int od_stage_timing(od_ctx* c, double* ms, int64_t* steps, int32_t reset) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  for (int i = 0; i < OD_N_STAGES; ++i)
    if (ms) ms[i] = c->stage_ms[i];
  if (steps) *steps = c->stage_steps;
  if (reset) {
    for (double& v : c->stage_ms) v = 0.0;
    c->stage_steps = 0;
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
    (void)hipFree(dX);
    (void)hipFree(dF);
    (void)hipFree(pi);
    (void)hipFree(pj);
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
  (void)hipFree(dX);
  (void)hipFree(dF);
  (void)hipFree(pi);
  (void)hipFree(pj);
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
    (void)hipFree(xs);
    (void)hipFree(ys);
    (void)hipFree(zs);
    (void)hipFree(drho);
    (void)hipFree(dh);
    (void)hipFree(part);
    (void)hipFree(dout);
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

#define OD_STR_(x) #x
#define OD_STR(x) OD_STR_(x)
extern "C" const char* od_build_info(void) { return "opendust_amd engine, gfx950, ABI " OD_STR(OD_ABI_VERSION); }

#include "od_batch.inc"

// Autotuner of the symmetric pair evaluation (kernel, diagonal, reduction) on
// the current positions.  Stage 1 times each candidate number of source blocks
// per work-group with the configured kernel variant; stage 2 every variant
// (tile shape, waves per SIMD) at the best block count; stage 3 the
// neighbouring block counts with the best variant.  Both choices group the
// partial sums, so runs are reproducible for a given pair, which is returned
// and can be fixed through od_config.sym_chunk and od_config.sym_tile.
static const int kTuneChunks[] = {1, 2, 4, 8, 16, 32};

// band_cap: the band size schedules are built with (batches: unlimited, one
// band).  Candidates whose partials exceed skip_cap in all (48 GB) are
// skipped, which keeps tuning short; larger schedules keep their default.
template <class C, class F>
static int tune_sym(C* c, const std::vector<int>& be, int reps, F&& evaluate, int* chosen_ch, int* chosen_var,
                    double band_cap, double skip_cap) {
  double best_ms = 1e300;
  int best_ch = c->sym_ch, best_var = c->sym_var;
  struct Cand {
    double ms, bytes;
    int ch, var;
  };
  std::vector<Cand> cands;
  // a captured step holds the schedule's buffers, which sym_build replaces:
  // drop it before the first rebuild, so no error path can leave it behind
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
  auto time_one = [&](int ch, int var) -> int {
    const int R = sym_variants[var].rows;
    // row groups need one simulation (a group must not straddle two)
    if (R > 1 && std::is_same<C, od_batch>::value) return OD_OK;
    if (sym_wg_count(be, ch, R) == 0 || sym_bytes(be, ch, R) > skip_cap) return OD_OK;
    if (c->sym_ch != ch || c->sym_R != R) CHKRET(sym_build(c, be, ch, band_cap, R));
    c->sym_var = var;
    evaluate();  // warm-up
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) evaluate();
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    cands.push_back({(double)ms, sym_bytes(be, ch, R), ch, var});
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
  // candidates within 1 % of the fastest count as ties (this timing's noise
  // is a few tenths of a percent): among them, keep the one whose schedule
  // writes the fewest partial bytes (HBM traffic, the bandwidth left to other
  // work; faster first among equal bytes)
  double sel_bytes = 1e300, sel_ms = 1e300;
  for (const Cand& k : cands)
    if (k.ms <= best_ms * 1.01 && (k.bytes < sel_bytes || (k.bytes == sel_bytes && k.ms < sel_ms))) {
      sel_bytes = k.bytes;
      sel_ms = k.ms;
      best_ch = k.ch;
      best_var = k.var;
    }
  if (c->sym_ch != best_ch || c->sym_R != sym_variants[best_var].rows)
    CHKRET(sym_build(c, be, best_ch, band_cap, sym_variants[best_var].rows));
  c->sym_var = best_var;
  *chosen_ch = best_ch;
  *chosen_var = best_var;
  return OD_OK;
}

extern "C" int od_sym_variants(void) { return SYM_N_VARIANTS; }

extern "C" int od_sym_variant_info(int32_t v, int32_t* targets_per_lane, int32_t* waves_per_simd,
                                   int32_t* vgpr_operands, int32_t* target_blocks) {
  if (v < 1 || v > SYM_N_VARIANTS) return set_err(OD_ERR_ARG, "variant %d outside [1, %d]", v, SYM_N_VARIANTS);
  if (targets_per_lane) *targets_per_lane = sym_variants[v - 1].mt;
  if (waves_per_simd) *waves_per_simd = sym_variants[v - 1].wpe;
  if (vgpr_operands) *vgpr_operands = sym_variants[v - 1].vgpr_step;
  if (target_blocks) *target_blocks = sym_variants[v - 1].rows;
  return OD_OK;
}

// Autotuner of the symmetric cell kernel (cutoff mode, FP32c): times `reps`
// whole force evaluations (binning included) on the current positions for
// each variant (cell_sym_kernels) and keeps the fastest.
static int tune_cells(od_ctx* c, int reps, int* chosen_var) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
  double* Fs = nullptr;
  CHKRET(dalloc(&Fs, (size_t)3 * c->n_pad * 8));
  int rc = load_positions(c, c->X, c->pi, c->pj);
  double best_ms = 1e300;
  int best = c->sc_var;
  for (int v = 0; v < CC_N_VARIANTS && rc == OD_OK; ++v) {
    c->sc_var = v;
    rc = launch_forces(c, c->X, c->pj, nullptr, Fs, false);  // warm-up
    if (rc == OD_OK && hipEventRecord(c->ev0, c->stream) != hipSuccess) rc = set_err(OD_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && rc == OD_OK; ++r) rc = launch_forces(c, c->X, c->pj, nullptr, Fs, false);
    float ms = 0.f;
    if (rc == OD_OK && (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
                        hipGetLastError() != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess))
      rc = set_err(OD_ERR_HIP, "timing the cell kernel variants failed");
    if (rc == OD_OK && ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  (void)hipFree(Fs);
  c->sc_var = best;
  *chosen_var = best;
  return rc;
}

extern "C" int od_tune(od_ctx* c, int32_t reps, int32_t* sym_chunk, int32_t* sym_tile) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (reps < 1 || reps > 100) return set_err(OD_ERR_ARG, "reps %d outside [1, 100]", reps);
  int chosen = 0, var = -1;
  if (c->cut && c->prec == OD_PREC_FP32C) {
    if (!c->have_ions) return set_err(OD_ERR_STATE, "od_tune needs the ions (call od_set_ions first)");
    HIPCHK(hipSetDevice(c->dev));
    CHKRET(tune_cells(c, reps, &var));
    CHKRET(load_positions(c, c->X, c->pi, c->pj));
    CHKRET(launch_forces(c, c->X, c->pj, nullptr, c->F, false));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->sym) {
    if (!c->have_ions) return set_err(OD_ERR_STATE, "od_tune needs the ions (call od_set_ions first)");
    HIPCHK(hipSetDevice(c->dev));
    double* Fs = nullptr;
    CHKRET(dalloc(&Fs, (size_t)3 * c->n_pad * 8));
    const std::vector<int> be(c->n_blocks, c->n_blocks);
    const int rc = tune_sym(c, be, reps, [&]() { launch_pair(c, c->pj, c->pj, Fs, false); }, &chosen, &var,
                            c->band_cap, SYM_BAND_BYTES);
    (void)hipFree(Fs);
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
    const int rc = tune_sym(b, b->blk_end_host, reps, [&]() { batch_pair<float>(b, false, Fs); }, &chosen, &var,
                            1e300, SYM_BAND_BYTES);
    (void)hipFree(Fs);
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
    if (d_zg) (void)hipFree(d_zg);
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
  (void)hipFree(d_zg);
  (void)hipFree(d_s);
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

// ---------------------------------------------------------------------------
// Allocation-free density profile.  Its scratch (the node positions and the
// per-node sums k_density writes) lives in one file-scope device array; each
// GPU holds its own copy, whose address is looked up once per device (at
// od_create, or at the first call on a device) and kept in g_scratch_ptr.
// Host code uses only that address, never the array's name.  Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU race.  Nothing
// needs clearing: each call uploads all node positions before the kernel and
// the kernel writes every node's sum.
// ---------------------------------------------------------------------------
constexpr int OD_NOALLOC_MAX_NODES = 65536;
static __device__ __align__(256) double scratch[2 * OD_NOALLOC_MAX_NODES];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device; null until looked up

// address of this device's copy of `scratch`; the device must be current
static int scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(OD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return OD_OK;
}

extern "C" int od_density_profile_noalloc(od_ctx* c, int32_t n_nodes, double R, double L, double N_s,
                                          double* z_grid, double* conc) {
  if (!c) return set_err(OD_ERR_ARG, "null context");
  if (n_nodes > OD_NOALLOC_MAX_NODES)
    return set_err(OD_ERR_ARG, "od_density_profile_noalloc: %d nodes exceed the scratch maximum of %d", n_nodes,
                   OD_NOALLOC_MAX_NODES);
  CHKRET(density_check(n_nodes, R, L, N_s));
  if (!c->have_ions) return set_err(OD_ERR_STATE, "density profile: no ions yet (call od_set_ions first)");
  HIPCHK(hipSetDevice(c->dev));
  double* sp = nullptr;
  CHKRET(scratch_address(c->dev, &sp));
  const int n = n_nodes;
  std::vector<double> zg(n), sums(n);
  for (int i = 0; i < n; i++) zg[i] = -L / 2.0 + i * L / (double)(n - 1);
  double* d_zg = sp;
  double* d_s = sp + OD_NOALLOC_MAX_NODES;
  HIPCHK(hipMemcpyAsync(d_zg, zg.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_density, dim3(n), dim3(64), 0, c->stream, (const double*)(c->X + 2 * (size_t)c->n_pad),
                     (int64_t)c->n, n, L, (const double*)d_zg, d_s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(sums.data(), d_s, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  density_finish(n, R, L, N_s, sums, z_grid, conc);
  return OD_OK;
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
  (void)hipFree(d_z);
  return rc;
}
