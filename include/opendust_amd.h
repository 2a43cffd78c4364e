/*
 * opendust_amd.h -- C ABI of the MI355X (gfx950) ion-flow engine.
 *
 * The engine replaces the per-timestep hot path of OpenDust (opendust.py, the
 * OpenMM System/CustomIntegrator built in OpenDust.simulate and driven by
 * integrator.step(1) at opendust.py:1059-1061) with hand-written HIP kernels.
 * The Python layer (opendust_amd/opendust.py) keeps OpenDust's public classes
 * and calls these functions through ctypes.
 *
 * Units inside the engine are the ones the reference hands to OpenMM
 * (opendust.py:274-312): lengths in nm, time in ps, masses in amu, energies in
 * kJ/mol, forces in kJ/mol/nm, velocities in nm/ps.  Dust charges and the
 * per-step dust records are SI (C, N), as in OpenDust.dustParticles.
 *
 * Conventions: every function returns 0 on success and a non-zero status on
 * failure; od_last_error() then returns a thread-local message.  Host buffers
 * are owned by the caller and never retained.  One context = one HIP device +
 * one HIP stream; a context is not thread-safe, distinct contexts are.
 */
#ifndef OPENDUST_AMD_H
#define OPENDUST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OD_ABI_VERSION 3
#define OD_MAX_DUST 8
#define OD_MAX_TABLE 512
#define OD_MAX_CELLS (1 << 27) /* cutoff mode: cells of the grid below */

/* status codes */
#define OD_OK 0
#define OD_ERR_ARG 1
#define OD_ERR_HIP 2
#define OD_ERR_NONFINITE 3
#define OD_ERR_STATE 4

/* distribution types (OpenDust distributionType, opendust.py:300-305) */
#define OD_DIST_MAXWELLIAN 0
#define OD_DIST_FIELD_DRIVEN 1

/* pair-force precision */
#define OD_PREC_FP32C 0 /* FP32 pair arithmetic, compensated (Kahan) per-ion sums */
#define OD_PREC_FP64 1  /* plain FP64 pair arithmetic */

/* all-pairs kernel (FP32 path; the FP64 path always uses the row kernel) */
#define OD_PAIR_AUTO 0 /* symmetric when the problem is large enough, else rows */
#define OD_PAIR_ROWS 1 /* every ordered pair, one target row per thread */
#define OD_PAIR_SYM 2  /* every unordered pair once (Newton's third law) */
#define OD_PAIR_CELL 3 /* reported by od_pair_info when a cutoff is set: cell list + stencil */

/* charge methods (opendust.py:1120, :1153; any other string keeps q) */
#define OD_CHARGE_GIVEN 0
#define OD_CHARGE_CONSISTENT 1
#define OD_CHARGE_OML 2

/* Piecewise-linear inverse CDF in the form the reference emits as an OpenMM
 * expression (opendust.py:702-723, :764-785):
 *   V(u) = sum_k (v0[k] + (u - xi0[k]) * slope[k]) * step(half[k] - |u - center[k]|)
 * with step(t) = 1 for t >= 0.  The constants are the %.8f-rounded values the
 * reference prints into the expression.  n_seg == 0 means "table absent". */
typedef struct od_table {
  int32_t n_seg;
  int32_t pad_;
  double v0[OD_MAX_TABLE];
  double xi0[OD_MAX_TABLE];
  double slope[OD_MAX_TABLE];
  double center[OD_MAX_TABLE];
  double half[OD_MAX_TABLE];
} od_table;

/* Everything the step needs; filled by the Python host layer. */
typedef struct od_config {
  int32_t abi_version;   /* = OD_ABI_VERSION */
  int32_t struct_size;   /* = sizeof(od_config) as the caller laid it out */
  int32_t n_ions;        /* N superions */
  int32_t n_dust;        /* 0 .. OD_MAX_DUST */
  int32_t distribution;  /* OD_DIST_* */
  int32_t precision;     /* OD_PREC_* */
  int32_t device;        /* HIP device ordinal */
  int32_t use_bottom;    /* reference applies the bottom-face selects only if p_b != 0 (opendust.py:939) */
  int32_t j_split;       /* source chunks of the row pair kernel; 0 = automatic */
  int32_t pair_mode;     /* OD_PAIR_*: which all-pairs kernel */
  int32_t sym_chunk;     /* symmetric kernel: 1024-ion source blocks per workgroup; 0 = automatic */
  int32_t sim_id;        /* Philox counter word 2 (index of the simulation in a sweep; 0 alone) */
  int32_t sym_tile;      /* symmetric kernel variant (tile shape, waves per SIMD, operand form),
                            1..od_sym_variants(); 0 = automatic (one simulation: variant 12, row
                            groups of 4 target blocks, sources read from LDS, with 8 source blocks
                            per work-group when that gives at least 4096 work-groups; else variant
                            11 with at least 2048 work-groups, else 3).  Cutoff mode, FP32c: the
                            symmetric cell kernel's 64-target sub-blocks per wave, 1, 2, 3 = 1, 2, 4
                            with sources rotated through registers, 4, 5, 6 = 1, 2, 4 with sources
                            read from LDS; 0 = 5 */
  int32_t sym_band_mib;  /* symmetric kernel: partials per band of rows of its triangular schedule, MiB;
                            0 = 48 GB (larger schedules, about 2^21 ions and up, run in several bands) */
  uint64_t seed;         /* Philox4x32-10 key */

  /* geometry, integration (OpenMM units) */
  double R, H, dt;       /* nm, nm, ps */
  double r_D_e, r_0;     /* nm */
  double m_s;            /* superion mass, amu */
  double v_T, v_fl;      /* nm/ps */

  /* force prefactors (OpenMM units) */
  double kqq;            /* kq_sq_s of the ion-ion law, kJ/mol*nm (opendust.py:472-479) */
  double k_dust;         /* k*conv*q_s of the dust->ion law: U = q_p*k_dust/r (opendust.py:351-359) */
  double trap_C;         /* prefactor of the trap polynomial, kJ/mol (opendust.py:552-557) */
  double trap_alpha[15]; /* fitted coefficients alphaTrap[0..14] (opendust.py:547) */
  double field_force;    /* E*q_s*conv, kJ/mol/nm (fieldDriven, opendust.py:580-585) */

  /* Optional truncation of the ion-ion law (opendust.py:481-485 uses OpenMM's
   * CutoffNonPeriodic with cutoff 3.2 r_D_e and a switch from 3 r_D_e).
   * cutoff <= 0: all pairs.  switch_dist in (0, cutoff): the energy is
   * multiplied by S(x) = 1 - 10x^3 + 15x^4 - 6x^5, x = (r - switch)/(cutoff - switch),
   * and the force is -d(U S)/dr; otherwise hard truncation at r < cutoff.
   * Sources are found through a cell list of columns w = cutoff/4 wide in x
   * and y and slabs h = cutoff/16 high in z (nm):
   *   nx = ny = ceil(2R/w), nz = ceil(H/h), ix = clamp(floor((x + R)/w), 0, nx-1),
   *   iy likewise, iz = clamp(floor((z + H/2)/h), 0, nz-1), cell = iz + nz (ix + nx iy),
   * at most OD_MAX_CELLS cells. */
  double cutoff;         /* nm */
  double switch_dist;    /* nm */

  /* Philox decision thresholds on a u32 draw, precomputed from p_t, p_s, p_b, dt*w_c */
  uint64_t thr_top;      /* face = top    if u32 <  thr_top */
  uint64_t thr_side;     /* face = side   if thr_top <= u32 < thr_side, else bottom */
  uint64_t thr_coll;     /* collision on a DOF if u32 < thr_coll (fieldDriven) */

  /* dust particles: nm for the ion side, SI for the dust-force assembly */
  double dust_x[OD_MAX_DUST], dust_y[OD_MAX_DUST], dust_z[OD_MAX_DUST], dust_r[OD_MAX_DUST];
  double dust_q0[OD_MAX_DUST];      /* initial charge, C */
  int32_t dust_method[OD_MAX_DUST]; /* OD_CHARGE_* */
  double si_dust_x[OD_MAX_DUST], si_dust_y[OD_MAX_DUST], si_dust_z[OD_MAX_DUST], si_dust_r[OD_MAX_DUST];

  /* SI constants of the host-side per-step physics (opendust.py:1117-1384) */
  double si_k;           /* 1/(4 pi eps_0) */
  double si_e;           /* electron charge (negative) */
  double si_q_s;         /* superion charge |e|*N_s */
  double si_r_D_e;       /* m */
  double si_dt;          /* s */
  double si_kB, si_Te, si_Tn;
  double si_E;           /* external field, V/m (fieldDriven) */
  double gamma_e_pref;   /* n_inf*(T_e*k_B/pi/m_e/2)**0.5 (opendust.py:1136-1140) */
  double oml_A;          /* n_inf*v_ti*u/4 */
  double oml_B;          /* (1 + 1/(2u*2)) */
  double oml_u2;         /* u**2 */
  double oml_erf_u;      /* erf(u) */
  double oml_tail;       /* 1/u/sqrt(pi)*exp(-u**2) */
  double collect_scale;  /* (nm/ps -> m/s) * m_i * N_s / d_t (opendust.py:1344-1361) */

  od_table top, bottom, side, vz;
} od_config;

/* Per-step, per-dust record (SI).  Field names follow OpenDust.dustParticles[p]. */
typedef struct od_dust_record {
  double q;                  /* dustParticles[p].q[i] */
  double force_orbit[3];     /* forceIonsOrbit[i] */
  double force_collect[3];   /* forceIonsCollect[i] */
  double force_dust[3];      /* forceDust[i] */
  double force_field[3];     /* forceExternalField[i] (fieldDriven) */
  double orbit_sum[3];       /* raw sum over ions (reference's orbit_f_*_p / 3) */
  double momentum_sum[3];    /* raw sum of v over absorbed ions, nm/ps (moment_f_*_p / 3) */
  int64_t absorbed;          /* ions inside this dust this step (delta_q_p / 3) */
} od_dust_record;

typedef struct od_ctx od_ctx;

/* Replaces mm.Context(system, integrator, platform, properties)
 * (opendust.py:1039-1042): allocates device state for cfg->n_ions ions. */
int od_create(const od_config* cfg, od_ctx** out);

/* Replaces context.setPositions / setVelocities (opendust.py:1043-1044).
 * x, v: [n][3] doubles, nm and nm/ps.  Recomputes the ion-ion force. */
int od_set_ions(od_ctx* ctx, const double* x, const double* v, int64_t n);

/* Replaces context.getState(getPositions, getVelocities) (opendust.py:1074-1078, :1103-1105). */
int od_get_ions(od_ctx* ctx, double* x, double* v);

/* Replaces the context's parameters between runs, keeping the ion state on the
 * device: dust positions, radii and charge methods, plasma constants, trap
 * coefficients, reinjection tables and seed may change; the ion and dust
 * counts, precision, device, pair kernel, cutoff and cylinder must not.  Dust
 * charges keep their current values.  Serves dust-dynamics loops, which the
 * reference runs as one full OpenDust(...).simulate() per dust step
 * (examples/DustDynamics/TwoParticles/launch.py:113-160: new context, ions
 * re-read from RESTART, trap refitted). */
int od_update_config(od_ctx* ctx, const od_config* cfg);

/* Replaces context.setParameter("q<p>", ...) (opendust.py:1151, :1202): q[n_dust], C. */
int od_set_dust_charges(od_ctx* ctx, const double* q);
int od_get_dust_charges(od_ctx* ctx, double* q);

/* Sets the loop index i of the next step (drives the "i > 0" charge rule and the
 * Philox step counter).  od_create starts at 0. */
int od_set_step(od_ctx* ctx, int64_t step);

/* Replaces nsteps iterations of the loop body at opendust.py:1059-1384 without
 * file output: integrator.step(1) plus the per-dust charge update and force
 * assembly.  out: caller-owned [nsteps][n_dust] records (may be NULL).
 * flags: OD_RUN_* below. */
#define OD_RUN_GRAPH 1          /* replay the step as a captured hipGraph */
#define OD_RUN_PROFILE_PAIR 2   /* time every ion-ion pair launch with HIP events */
#define OD_RUN_PROFILE_STAGES 4 /* time every stage of the step with HIP events (od_stage_timing) */
int od_run(od_ctx* ctx, int64_t nsteps, od_dust_record* out, int32_t flags);

/* Accumulated HIP-event time of the main pair launches timed with
 * OD_RUN_PROFILE_PAIR since the last reset, and their count. */
int od_pair_timing(od_ctx* ctx, double* total_ms, int64_t* launches, int32_t reset);

/* Stages of one step, as od_stage_timing reports them (measurement support,
 * no counterpart in the reference). */
#define OD_STAGE_KICK_DRIFT 0       /* collisions, first half-kick, drift, flags (k_kick_drift) */
#define OD_STAGE_LIST 1             /* ordered list of the flagged ions (k_scan, k_list) */
#define OD_STAGE_CELLS 2            /* cutoff mode: binning, radix sort, cell starts, work items, gather */
#define OD_STAGE_PAIR 3             /* main ion-ion force evaluation and its reduction */
#define OD_STAGE_ROWS_ABSORBED 4    /* force rows of the ions inside a dust (drifted state) */
#define OD_STAGE_CORR 5             /* cutoff mode: the flagged ions' share as sources (both states) */
#define OD_STAGE_KICK2 6            /* correction, second half-kick, counts, reinjection, orbit sums (k_kick2) */
#define OD_STAGE_ROWS_REINJECTED 7  /* force rows of the reinjected ions */
#define OD_STAGE_INFLOW 8           /* the reinjected ions' share of everybody's force (k_inflow) */
#define OD_STAGE_DUST 9             /* FP64 reductions, charge update, dust records (k_dust) */
#define OD_N_STAGES 10

/* Accumulated HIP-event time per stage (ms[OD_N_STAGES]) of the steps run
 * with OD_RUN_PROFILE_STAGES since the last reset, and their count. */
int od_stage_timing(od_ctx* ctx, double* ms, int64_t* steps, int32_t reset);

/* Which all-pairs kernel the context runs (OD_PAIR_ROWS or OD_PAIR_SYM) and
 * how many ordered ion pairs its main launch (the one OD_RUN_PROFILE_PAIR
 * times) covers: N(N-1) for rows, N(N-1) minus the intra-block pairs for the
 * symmetric kernel.  No counterpart in the reference (measurement support). */
int od_pair_info(od_ctx* ctx, int32_t* mode, int64_t* pairs_per_launch);

/* The symmetric kernel's current schedule: source blocks per work-group,
 * kernel variant (1-based), as configured, defaulted or tuned, and the number
 * of bands of rows its triangular schedule runs in (od_config.sym_band_mib);
 * 0, 0 and 0 when the context does not run it.  Cutoff mode, FP32c: 0, the
 * symmetric cell kernel's variant (1, 2, 3: 1, 2, 4 target sub-blocks per
 * wave) and 0.  Measurement support. */
int od_sym_choice(od_ctx* ctx, int32_t* sym_chunk, int32_t* sym_tile, int32_t* bands);

/* Autotuner of the symmetric FP32 pair kernel (contexts that run it; others
 * return *sym_chunk = *sym_tile = 0 and change nothing; cutoff mode, FP32c:
 * it times the symmetric cell kernel's six variants on the current ions, keeps
 * the fastest and returns *sym_chunk = 0, *sym_tile = the variant).  It times the whole
 * pair evaluation `reps` times per candidate on the current ions: first 1, 2,
 * 4, 8, 16 and 32 source blocks per work-group with the configured kernel
 * variant, then every kernel variant (tile shape: 4, 8 or 16 64-target
 * sub-blocks per wave; register budget: waves per SIMD; operand form; target
 * blocks per work-group, one simulation only) at the best of those, then the
 * neighbouring block counts with the best variant.  Among the candidates
 * within 1 % of the fastest (counted as ties) it keeps the one
 * whose schedule writes the fewest partial bytes, and recomputes the ion-ion
 * force of the current positions with it.
 * Both choices group the partial sums, so results are reproducible for a
 * given pair: *sym_chunk and *sym_tile return it, and od_config.sym_chunk and
 * od_config.sym_tile fix it.  Call after od_set_ions.  No counterpart in the
 * reference (OpenMM picks its kernels internally). */
int od_tune(od_ctx* ctx, int32_t reps, int32_t* sym_chunk, int32_t* sym_tile);

/* Number of symmetric kernel variants, and the tile shape (64-target
 * sub-blocks per wave), waves per SIMD, operand form (0: sources rotated
 * through DPP operands of the subtractions, SGPR constants; 1: sources and
 * reactions rotated by DPP moves, every full-rate instruction with VGPR
 * operands; 2: sources read from the wave's LDS copy, reactions rotated by
 * DPP moves, VGPR operands) and
 * target blocks per work-group (> 1: row groups that share each source
 * partial, one simulation only, not in batches) of variant v (1-based) as
 * compiled.  Measurement support. */
int od_sym_variants(void);
int od_sym_variant_info(int32_t v, int32_t* targets_per_lane, int32_t* waves_per_simd, int32_t* vgpr_operands,
                        int32_t* target_blocks);

/* Cutoff mode only: the cell list of the last force evaluation.  cell_of_ion[n]
 * (binned ions: their cell; ions left out: the number of cells), sorted_ions
 * (binned ions in cell order, ties by ion index), their count, and the grid
 * dims[3].  Any pointer may be NULL.  No counterpart in the reference (OpenMM
 * keeps its neighbour list internal). */
int od_get_cells(od_ctx* ctx, int32_t* cell_of_ion, int32_t* sorted_ions, int64_t* n_binned, int32_t* dims);

/* Cutoff mode only, for the current cell list (measurement support):
 * ordered ion pairs within the cutoff, and the ordered pair terms the cell
 * kernel evaluates, padding included (FP64: 64 lanes x the sources of the
 * slab ranges of each 64-target chunk; FP32c symmetric kernel: two terms per
 * evaluated pair slot, force and reaction, one per slot of a group's own
 * pairs). */
int od_cutoff_pair_count(od_ctx* ctx, int64_t* in_range, int64_t* candidates);

/* Diagnostics of the last step (for parity tests): per-ion flag word
 * (bit p < 16: inside dust p; bit 16: outside in r; bit 17: outside in z),
 * per-ion reinjection face (0 none, 1 top, 2 side, 3 bottom), and the count. */
int od_get_step_info(od_ctx* ctx, uint32_t* flags, uint8_t* face, int64_t* n_reinjected);

/* Ion-ion force of the reference law on a given state, with the engine's
 * kernels and precision (opendust.py:472-479): x [n][3] nm -> f [n][3] kJ/mol/nm. */
int od_pair_forces(od_ctx* ctx, const double* x, double* f);

/* Background potential of a uniform ion cylinder on the (rho, h) grid, Monte
 * Carlo with n_mc Philox samples (opendust.py:499-523).  R, H, rho[n_r], h[n_h]
 * in units of r_D_e (grid points (rho, 0, h)).
 * out: [n_r][n_h] mean of exp(-d)/(d + 0.001). */
int od_trap_potential_grid(int32_t device, double R, double H, int32_t n_r, const double* rho,
                           int32_t n_h, const double* h, int64_t n_mc, uint64_t seed, double* out);

/* Sample coordinates of the Monte Carlo above (for checking the sampler):
 * k in [k0, k0 + n) -> xyz[n][3], units of r_D_e.  Pure host computation. */
int od_trap_samples_host(double R, double H, uint64_t seed, int64_t k0, int64_t n, double* xyz);

/* Axial ion-density profile of the context's current ions, computed on the
 * device: replaces writing an XYZ frame and running tools/concentration.cpp
 * (:32-101) on it.  n_nodes nodes z_grid[i] = -L/2 + i L/(n_nodes-1) (m);
 * conc[i] in 1/m^3 from the tool's linear deposits (an ion in the first or
 * last interval feeds only node 1 or n_nodes-2), divided by
 * 0.5 (dz_left + dz_right) pi R^2 / N_s, end nodes extrapolated linearly.
 * R, L (the domain radius and height) in m, N_s ions per superion.  Sums are
 * the tool's sequential sums in ion order, bit for bit. */
int od_density_profile(od_ctx* ctx, int32_t n_nodes, double R, double L, double N_s, double* z_grid, double* conc);

/* od_density_profile without any device allocation, for callers that cannot
 * allocate (tight dust-dynamics or serving loops, code that must not call
 * hipMalloc).  No counterpart in the reference.  Same arguments and results as
 * od_density_profile (bit-identical); n_nodes is at most 65536, the size of
 * the library's per-device scratch array (1 MiB, held until the process ends).
 * Calls that overlap on one GPU share that array and race. */
int od_density_profile_noalloc(od_ctx* ctx, int32_t n_nodes, double R, double L, double N_s, double* z_grid,
                               double* conc);

/* The same for host positions z_nm[N] (nm), e.g. one frame of an XYZ file. */
int od_density_profile_z(int32_t device, const double* z_nm, int64_t N, int32_t n_nodes, double R, double L,
                         double N_s, double* z_grid, double* conc);

/* ---- output files (host code, no device use) ----------------------------
 * The reference's text formats, every number written as Python's "{}".format
 * writes a float (shortest round-trip digits; fixed notation for decimal
 * exponents -4..15, else scientific with a two-digit exponent).
 * Replace the per-ion Python loops at opendust.py:1080-1088 (XYZ frame,
 * appended: "N", "frame", then "Ar\tx\ty\tz" per ion), :1089-1099 (CSV frame
 * via csv.writer: header x,y,z,density,mass, CRLF line ends) and :1100-1115
 * (RESTART: "x\ty\tz\tvx\tvy\tvz" per ion).  pos, vel: [n][3] doubles. */
int od_write_xyz(const char* path, const double* pos, int64_t n, int64_t frame);
int od_write_csv(const char* path, const double* pos, int64_t n, double density, double mass);
int od_write_restart(const char* path, const double* pos, const double* vel, int64_t n);

/* ---- batched parameter sweeps ------------------------------------------
 * Many independent simulations advanced together on one device: one launch of
 * each step kernel covers all of them, and the all-pairs work of all
 * simulations is a single launch over (simulation, target block, source chunk)
 * items.  Replaces a sequential Python loop of OpenDust(...).simulate() calls
 * (examples/Piel2017/Figure5/launchPiel2017Figure5.py:30); the reference has no
 * batch interface.  Every simulation has its own od_config (any ion and dust
 * count, plasma and geometry); all share precision and device, and run
 * all-pairs forces.  cfg.sim_id is the simulation's Philox counter word, so a
 * batched simulation draws the same random numbers as an od_ctx created with
 * the same od_config. */
typedef struct od_batch od_batch;

int od_batch_create(const od_config* cfgs, int32_t n_sims, od_batch** out);

/* Initial ions of simulation sim: x, v [n][3] (n = its cfg.n_ions). */
int od_batch_set_ions(od_batch* b, int32_t sim, const double* x, const double* v, int64_t n);
int od_batch_get_ions(od_batch* b, int32_t sim, double* x, double* v);

/* nsteps steps of every simulation.  out: caller-owned
 * [nsteps][n_sims][OD_MAX_DUST] records (slots p >= a simulation's n_dust are
 * zero; may be NULL).  flags: OD_RUN_GRAPH, OD_RUN_PROFILE_PAIR. */
int od_batch_run(od_batch* b, int64_t nsteps, od_dust_record* out, int32_t flags);

/* Flags and faces of simulation sim's ions in the last step (as od_get_step_info). */
int od_batch_get_step_info(od_batch* b, int32_t sim, uint32_t* flags, uint8_t* face);

/* HIP-event time of the batched pair launches timed with OD_RUN_PROFILE_PAIR
 * since the last reset, their count, and the ordered pairs of real ions one
 * launch covers: sum over simulations of n(n-1), less the pairs inside one
 * 1024-ion block when the symmetric FP32 kernel runs (fp32c and pair_mode !=
 * OD_PAIR_ROWS in the first configuration).  Measurement support. */
int od_batch_pair_timing(od_batch* b, double* total_ms, int64_t* launches, int64_t* pairs_per_launch,
                         int32_t reset);

/* od_tune for a batch (all simulations share one schedule; the choice is
 * fixed through the first configuration's sym_chunk and sym_tile).  Call
 * after the od_batch_set_ions calls; the next od_batch_run recomputes the
 * forces. */
int od_batch_tune(od_batch* b, int32_t reps, int32_t* sym_chunk, int32_t* sym_tile);

int od_batch_destroy(od_batch* b);

/* sizeof(od_config), sizeof(od_dust_record) as compiled (ABI check, no GPU use). */
int od_abi_sizes(int64_t* config_size, int64_t* record_size);

int od_destroy(od_ctx* ctx);

/* Thread-local message of the last failure on this thread. */
const char* od_last_error(void);

/* Version string of the build (arch, ABI). */
const char* od_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* OPENDUST_AMD_H */
