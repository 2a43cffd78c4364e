/*
 * od_oracle.c -- FP64 CPU restatement of OpenDust's per-step ion-flow algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the engine is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it.  The engine never calls it.
 *
 * It follows the reference directly (src/opendust/opendust.py), in OpenMM units
 * (nm, ps, amu, kJ/mol) and in the reference's own formula shapes:
 *   ion-ion force      opendust.py:472-479   F = -dU/dr r_hat, U = kqq e^{-r/lambda}/(r+r_0)
 *   dust -> ion        opendust.py:349-360   U = q_p k' q_s / r
 *   trap               opendust.py:550-577   U = C sum alpha_m (rho/lambda)^a (z/lambda)^b
 *   field              opendust.py:579-585   U = -E' q_s z
 *   collisions         opendust.py:869-873
 *   KDK                opendust.py:864-868 / :874-877
 *   flags              opendust.py:894-910
 *   counts, momenta    opendust.py:911-921
 *   reinjection        opendust.py:923-975 (tables: opendust.py:606-837, built on the host)
 *   orbit sums         opendust.py:977-1037
 *   charge update      opendust.py:1117-1202
 *   dust forces        opendust.py:1204-1384
 * The ion-ion force is evaluated twice per step by brute force: at the drifted
 * positions for the second half-kick, and again at the post-reinjection
 * positions for the next step's first half-kick, as OpenMM does after the
 * reinjection writes x.
 *
 * Randomness follows the engine's counter-based specification (Philox4x32-10,
 * Salmon et al. SC'11), restated here: counter {ion, step, sim_id,
 * stream | sub<<8}, key = seed; u53(a,b) = ((a>>5)*2^26 + (b>>6)) 2^-53;
 * Box-Muller on (u53(w0,w1) + 2^-53, u53(w2,w3)).
 *
 * Also: the trap Monte Carlo (opendust.py:499-523), the cell binning
 * specification of the cutoff mode (include/opendust_amd.h), and the axial
 * density profile of tools/concentration.cpp:32-101.
 *
 * How each part is pinned (tests/test_oracle_pinning.py, tests/test_host.py):
 * Philox by published known-answer vectors and rocRAND; pair, switched-cutoff,
 * dust, trap and field forces as minus the gradient of the reference's
 * energies (finite differences); KDK, flags, counts, momenta and the step
 * ordering by closed forms on crafted states; reinjection by its sampling
 * distributions; charge updates by textbook limits; dust forces by Coulomb
 * gradients; the trap grid by quadrature; the density profile by linear
 * interpolation identities.  OpenMM's own kernels and random streams are not
 * reproducible without OpenMM: that level is parity unpinned (the end-to-end
 * statistics of the reference's example runs are the check there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "opendust_amd.h"

/* ---------------------------------------------------------------- Philox */
void orc_philox(const uint32_t ctr[4], uint64_t key, uint32_t out[4]) {
  uint32_t x[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {(uint32_t)(key & 0xffffffffu), (uint32_t)(key >> 32)};
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    const uint64_t m0 = (uint64_t)x[0] * 0xD2511F53u;
    const uint64_t m1 = (uint64_t)x[2] * 0xCD9E8D57u;
    const uint32_t y0 = (uint32_t)(m1 >> 32) ^ x[1] ^ k[0];
    const uint32_t y1 = (uint32_t)m1;
    const uint32_t y2 = (uint32_t)(m0 >> 32) ^ x[3] ^ k[1];
    const uint32_t y3 = (uint32_t)m0;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
    x[3] = y3;
  }
  memcpy(out, x, sizeof(x));
}

enum { ST_FACE = 0, ST_TOPV = 1, ST_TOPX = 2, ST_BOTV = 3, ST_BOTX = 4, ST_SIDEX = 5, ST_SIDEV = 6, ST_COLL = 7,
       ST_TRAP = 16 };

/* counter {ion, step, simulation, stream | sub << 8}, key = seed */
static void draw(const od_config* c, uint32_t ion, uint32_t step, uint32_t stream, uint32_t sub, uint32_t out[4]) {
  const uint32_t ctr[4] = {ion, step, (uint32_t)c->sim_id, stream | (sub << 8)};
  orc_philox(ctr, c->seed, out);
}

static double unif53(uint32_t a, uint32_t b) {
  return ldexp((double)(((uint64_t)(a >> 5) << 26) + (uint64_t)(b >> 6)), -53);
}

static void gauss_pair(const uint32_t w[4], double* g1, double* g2) {
  const double u1 = unif53(w[0], w[1]) + ldexp(1.0, -53);
  const double u2 = unif53(w[2], w[3]);
  const double rad = sqrt(-2.0 * log(u1));
  *g1 = rad * cos(6.283185307179586 * u2);
  *g2 = rad * sin(6.283185307179586 * u2);
}

/* piecewise-linear inverse CDF in the reference's expression form */
double orc_table_eval(const od_table* t, double u) {
  double s = 0.0;
  for (int k = 0; k < t->n_seg; ++k) {
    const double line = t->v0[k] + (u - t->xi0[k]) * t->slope[k];
    const double mask = (t->half[k] - fabs(u - t->center[k]) >= 0.0) ? 1.0 : 0.0;
    s = s + line * mask;
  }
  return s;
}

/* ------------------------------------------------------------- forces */
static const int MONO[15][2] = {{0, 0}, {2, 0}, {0, 2}, {2, 2}, {4, 0}, {0, 4}, {4, 2}, {2, 4},
                                {6, 0}, {0, 6}, {4, 4}, {6, 2}, {2, 6}, {8, 0}, {0, 8}};

/* ion-ion force on rows [i0, i1) (kJ/mol/nm).  With cfg->cutoff > 0 only
 * pairs with r < cutoff interact; with 0 < switch_dist < cutoff the energy is
 * U(r) S(x), S = 1 - 10x^3 + 15x^4 - 6x^5, x = (r - switch)/(cutoff - switch)
 * (OpenMM's switching function, opendust.py:483-485), and the force is
 * -d(U S)/dr. */
void orc_pair_forces(const od_config* c, int n, const double* x, double* f, int i0, int i1, int nthreads) {
  const double lam = c->r_D_e, r0 = c->r_0, kqq = c->kqq;
  const int cut = c->cutoff > 0.0;
  const int sw = cut && c->switch_dist > 0.0 && c->switch_dist < c->cutoff;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int i = i0; i < i1; ++i) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = xi - x[3 * j], dy = yi - x[3 * j + 1], dz = zi - x[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (cut && !(r < c->cutoff)) continue;
      const double rr = r + r0;
      /* -dU/dr for U = kqq e^{-r/lambda} / (r + r_0) */
      double mag = kqq * exp(-r / lam) * (1.0 / (lam * rr) + 1.0 / (rr * rr));
      if (sw && r > c->switch_dist) {
        const double width = c->cutoff - c->switch_dist;
        const double t = (r - c->switch_dist) / width;
        const double S = 1.0 - 10.0 * t * t * t + 15.0 * t * t * t * t - 6.0 * t * t * t * t * t;
        const double dSdr = (-30.0 * t * t + 60.0 * t * t * t - 30.0 * t * t * t * t) / width;
        const double U = kqq * exp(-r / lam) / rr;
        mag = mag * S - U * dSdr;
      }
      fx += mag * dx / r;
      fy += mag * dy / r;
      fz += mag * dz / r;
    }
    f[3 * (i - i0)] = fx;
    f[3 * (i - i0) + 1] = fy;
    f[3 * (i - i0) + 2] = fz;
  }
}

/* Test support for the truncated law: for rows i in [i0, i1), the number of
 * other ions j whose distance lies within rel * cutoff of the cutoff, i.e. the
 * pairs whose inclusion a relative rounding of r by up to rel can flip.  The
 * distance is computed as in orc_pair_forces. */
void orc_cut_edge_count(const od_config* c, int n, const double* x, int i0, int i1, double rel, int* out,
                        int nthreads) {
  const double lo = c->cutoff * (1.0 - rel), hi = c->cutoff * (1.0 + rel);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int i = i0; i < i1; ++i) {
    int m = 0;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = xi - x[3 * j], dy = yi - x[3 * j + 1], dz = zi - x[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      m += (r >= lo && r <= hi) ? 1 : 0;
    }
    out[i - i0] = m;
  }
}

/* dust Coulomb + trap + field force on one ion, q[] = current dust charges */
void orc_ext_force(const od_config* c, const double* q, const double* xi, double* fo) {
  double f[3] = {0.0, 0.0, 0.0};
  for (int p = 0; p < c->n_dust; ++p) {
    const double d[3] = {xi[0] - c->dust_x[p], xi[1] - c->dust_y[p], xi[2] - c->dust_z[p]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) f[k] += q[p] * c->k_dust * d[k] / (r * r * r);
  }
  /* trap: F = -C grad sum alpha_m (rho/lam)^a (z/lam)^b */
  const double lam = c->r_D_e;
  const double rho = sqrt(xi[0] * xi[0] + xi[1] * xi[1]);
  const double z = xi[2];
  double dPdrho = 0.0, dPdz = 0.0;
  for (int m = 1; m < 15; ++m) {
    const int a = MONO[m][0], b = MONO[m][1];
    if (a > 0) dPdrho += c->trap_alpha[m] * a * pow(rho, a - 1) * pow(z, b) / pow(lam, a + b);
    if (b > 0) dPdz += c->trap_alpha[m] * b * pow(rho, a) * pow(z, b - 1) / pow(lam, a + b);
  }
  const double FR = -c->trap_C * dPdrho, FZ = -c->trap_C * dPdz;
  if (rho > 0.0) {
    f[0] += FR * xi[0] / rho;
    f[1] += FR * xi[1] / rho;
  }
  f[2] += FZ;
  if (c->distribution == OD_DIST_FIELD_DRIVEN) f[2] += c->field_force;
  fo[0] = f[0];
  fo[1] = f[1];
  fo[2] = f[2];
}

/* total force on every ion (pair + external) */
static void total_forces(const od_config* c, int n, const double* x, const double* q, double* f, int nthreads) {
  orc_pair_forces(c, n, x, f, 0, n, nthreads);
  for (int i = 0; i < n; ++i) {
    double e[3];
    orc_ext_force(c, q, x + 3 * i, e);
    f[3 * i] += e[0];
    f[3 * i + 1] += e[1];
    f[3 * i + 2] += e[2];
  }
}

/* ------------------------------------------------------- reinjection */
/* returns face: 1 top, 2 side, 3 bottom */
int orc_reinject(const od_config* c, uint32_t ion, uint32_t step, double* x, double* v) {
  uint32_t w[4];
  const double pi2 = 2.0 * 3.141592653589793;
  draw(c, ion, step, ST_FACE, 0, w);
  const uint64_t u = w[0];
  const uint64_t thr_side = c->use_bottom ? c->thr_side : 4294967296ull;
  int face = (u < c->thr_top) ? 1 : (u < thr_side ? 2 : 3);
  if (face == 1 || face == 3) {
    const int top = face == 1;
    const od_table* tab = top ? &c->top : &c->bottom;
    double g1, g2;
    draw(c, ion, step, top ? ST_TOPV : ST_BOTV, 0, w);
    gauss_pair(w, &g1, &g2);
    draw(c, ion, step, top ? ST_TOPV : ST_BOTV, 1, w);
    const double uz = unif53(w[0], w[1]);
    v[0] = c->v_T * g1;
    v[1] = c->v_T * g2;
    v[2] = orc_table_eval(tab, uz);
    draw(c, ion, step, top ? ST_TOPX : ST_BOTX, 0, w);
    const double ux = unif53(w[0], w[1]), uy = unif53(w[2], w[3]);
    x[0] = c->R * sqrt(ux) * cos(pi2 * uy);
    x[1] = c->R * sqrt(ux) * sin(pi2 * uy);
    x[2] = top ? -c->H / 2 : c->H / 2;
  } else {
    draw(c, ion, step, ST_SIDEX, 0, w);
    const double uy = unif53(w[0], w[1]), uz = unif53(w[2], w[3]);
    x[0] = c->R * cos(pi2 * uy);
    x[1] = c->R * sin(pi2 * uy);
    x[2] = uz * c->H - c->H / 2;
    double gy, gz;
    draw(c, ion, step, ST_SIDEV, 0, w);
    gauss_pair(w, &gy, &gz);
    draw(c, ion, step, ST_SIDEV, 1, w);
    const double us = unif53(w[0], w[1]);
    const double vs = orc_table_eval(&c->side, us);
    v[0] = -vs * (x[0] / c->R) + c->v_T * gy * (x[1] / c->R);
    v[1] = -vs * (x[1] / c->R) - c->v_T * gy * (x[0] / c->R);
    v[2] = (c->distribution == OD_DIST_FIELD_DRIVEN) ? orc_table_eval(&c->vz, us) : gz * c->v_T + c->v_fl;
  }
  return face;
}

/* collisions of one ion's velocity (fieldDriven) */
void orc_collide(const od_config* c, uint32_t ion, uint32_t step, double* v) {
  uint32_t w[4];
  draw(c, ion, step, ST_COLL, 0, w);
  int hit[3];
  for (int k = 0; k < 3; ++k) hit[k] = (uint64_t)w[k] < c->thr_coll;
  if (!(hit[0] || hit[1] || hit[2])) return;
  double g[4];
  draw(c, ion, step, ST_COLL, 1, w);
  gauss_pair(w, &g[0], &g[1]);
  draw(c, ion, step, ST_COLL, 2, w);
  gauss_pair(w, &g[2], &g[3]);
  for (int k = 0; k < 3; ++k)
    if (hit[k]) v[k] = c->v_T * g[k];
}

/* ------------------------------------------------- dust observables */
/* flags of one ion at x (bit p: inside dust p, bit 16: rho >= R, bit 17: |z| >= H/2) */
uint32_t orc_flags(const od_config* c, const double* x) {
  uint32_t fl = 0;
  for (int p = 0; p < c->n_dust; ++p) {
    const double dx = x[0] - c->dust_x[p], dy = x[1] - c->dust_y[p], dz = x[2] - c->dust_z[p];
    if (-sqrt(dx * dx + dy * dy + dz * dz) + c->dust_r[p] >= 0.0) fl |= 1u << p;
  }
  if (-c->R + sqrt(x[0] * x[0] + x[1] * x[1]) >= 0.0) fl |= 1u << 16;
  if (-c->H / 2 + fabs(x[2]) >= 0.0) fl |= 1u << 17;
  return fl;
}

/* orbit term of one ion for dust p (opendust.py:989-1037), summed by the caller */
void orc_orbit_term(const od_config* c, int p, const double* x, double* o) {
  const double lam = c->r_D_e;
  const double xp[3] = {c->dust_x[p], c->dust_y[p], c->dust_z[p]};
  const double r = sqrt((x[0] - xp[0]) * (x[0] - xp[0]) + (x[1] - xp[1]) * (x[1] - xp[1]) +
                        (x[2] - xp[2]) * (x[2] - xp[2]));
  const double e = exp(-r / lam);
  for (int k = 0; k < 3; ++k)
    o[k] = e * (xp[k] - x[k]) / (r * r * r) * (lam * lam) + e * (xp[k] - x[k]) / (r * r) * lam;
}

/* charge update of one dust (opendust.py:1117-1202); returns q[i] */
double orc_charge(const od_config* c, int p, int64_t step, double q_prev, int64_t absorbed) {
  const int m = c->dust_method[p];
  if (step <= 0 || (m != OD_CHARGE_CONSISTENT && m != OD_CHARGE_OML)) return q_prev;
  const double rp = c->si_dust_r[p];
  const double phi = c->si_k * q_prev / rp;
  const double gamma_e = c->gamma_e_pref * exp(-c->si_e * phi / c->si_Te / c->si_kB);
  const double S = rp * rp * 4 * 3.141592653589793;
  double dq;
  if (m == OD_CHARGE_CONSISTENT) {
    dq = (double)absorbed * c->si_q_s + gamma_e * c->si_dt * S * c->si_e;
  } else {
    const double chi = c->si_e * phi / c->si_Tn / c->si_kB;
    const double gamma_i = c->oml_A * ((c->oml_B + chi / c->oml_u2) * c->oml_erf_u + c->oml_tail);
    dq = gamma_i * c->si_dt * S * fabs(c->si_e) + gamma_e * c->si_dt * S * c->si_e;
  }
  return q_prev + dq;
}

/* out-of-domain trap force on dust p with charge qp (opendust.py:1258-1321), SI */
void orc_out_domain_force(const od_config* c, int p, double qp, double* f) {
  const double lam = c->si_r_D_e;
  const double x = c->si_dust_x[p], y = c->si_dust_y[p], z = c->si_dust_z[p];
  const double r = sqrt(x * x + y * y);
  double dPdr = 0.0, dPdz = 0.0; /* derivatives of sum alpha_m (r/lam)^a (z/lam)^b */
  for (int m = 1; m < 15; ++m) {
    const int a = MONO[m][0], b = MONO[m][1];
    if (a > 0) dPdr += c->trap_alpha[m] * a * pow(r, a - 1) * pow(z, b) / pow(lam, a + b);
    if (b > 0) dPdz += c->trap_alpha[m] * b * pow(r, a) * pow(z, b - 1) / pow(lam, a + b);
  }
  const double pref = c->si_k * qp * c->si_q_s / lam; /* energy scale of the trap seen by charge qp */
  const double fR = -pref * dPdr;
  const double fH = -pref * dPdz;
  f[0] = (r == 0.0) ? 0.0 : fR * x / r;
  f[1] = (r == 0.0) ? 0.0 : fR * y / r;
  f[2] = fH;
}

/* assemble the per-step records from the raw sums; q_new[] = charges after the update */
static void assemble(const od_config* c, const double* q_new, const double* cnt, const double* mom,
                     const double* orb, od_dust_record* rec) {
  const int nd = c->n_dust;
  for (int p = 0; p < nd; ++p) {
    od_dust_record* r = &rec[p];
    memset(r, 0, sizeof(*r));
    r->q = q_new[p];
    for (int j = 0; j < nd; ++j) {
      if (j == p) continue;
      const double d[3] = {c->si_dust_x[p] - c->si_dust_x[j], c->si_dust_y[p] - c->si_dust_y[j],
                           c->si_dust_z[p] - c->si_dust_z[j]};
      const double rr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int k = 0; k < 3; ++k) r->force_dust[k] += c->si_k * q_new[p] * q_new[j] * d[k] / (rr * rr * rr);
    }
    double fo[3];
    orc_out_domain_force(c, p, q_new[p], fo);
    for (int k = 0; k < 3; ++k) {
      r->orbit_sum[k] = orb[3 * p + k];
      r->momentum_sum[k] = mom[3 * p + k];
      r->force_orbit[k] = orb[3 * p + k] * q_new[p] * c->si_k * c->si_q_s / (c->si_r_D_e * c->si_r_D_e) + fo[k];
      r->force_collect[k] = mom[3 * p + k] * c->collect_scale;
    }
    r->force_field[2] = (c->distribution == OD_DIST_FIELD_DRIVEN) ? q_new[p] * c->si_E : 0.0;
    r->absorbed = (int64_t)cnt[p];
  }
}

/* ----------------------------------------------------------------- step */
/* One step of the reference loop body.  x, v: [n][3] (in/out), q: [n_dust]
 * (in: q[i-1], out: q[i]).  rec: [n_dust].  flags_out/face_out may be NULL. */
int orc_step(const od_config* c, int n, double* x, double* v, double* q, int64_t step, od_dust_record* rec,
             uint32_t* flags_out, uint8_t* face_out, int nthreads) {
  const int nd = c->n_dust;
  const uint32_t st = (uint32_t)step;
  double* f = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  if (!f) return 1;
  if (c->distribution == OD_DIST_FIELD_DRIVEN)
    for (int i = 0; i < n; ++i) orc_collide(c, (uint32_t)i, st, v + 3 * i);
  total_forces(c, n, x, q, f, nthreads);
  for (int i = 0; i < 3 * n; ++i) v[i] = v[i] + 0.5 * c->dt * f[i] / c->m_s;
  for (int i = 0; i < 3 * n; ++i) x[i] = x[i] + c->dt * v[i];
  total_forces(c, n, x, q, f, nthreads);
  for (int i = 0; i < 3 * n; ++i) v[i] = v[i] + 0.5 * c->dt * f[i] / c->m_s;
  double cnt[OD_MAX_DUST] = {0}, mom[3 * OD_MAX_DUST] = {0}, orb[3 * OD_MAX_DUST] = {0};
  for (int i = 0; i < n; ++i) {
    const uint32_t fl = orc_flags(c, x + 3 * i);
    if (flags_out) flags_out[i] = fl;
    for (int p = 0; p < nd; ++p)
      if ((fl >> p) & 1u) {
        cnt[p] += 1.0;
        for (int k = 0; k < 3; ++k) mom[3 * p + k] += v[3 * i + k];
      }
    int face = 0;
    if (fl) face = orc_reinject(c, (uint32_t)i, st, x + 3 * i, v + 3 * i);
    if (face_out) face_out[i] = (uint8_t)face;
  }
  for (int i = 0; i < n; ++i)
    for (int p = 0; p < nd; ++p) {
      double o[3];
      orc_orbit_term(c, p, x + 3 * i, o);
      for (int k = 0; k < 3; ++k) orb[3 * p + k] += o[k];
    }
  double qn[OD_MAX_DUST];
  for (int p = 0; p < nd; ++p) qn[p] = orc_charge(c, p, step, q[p], (int64_t)cnt[p]);
  assemble(c, qn, cnt, mom, orb, rec);
  for (int p = 0; p < nd; ++p) q[p] = qn[p];
  free(f);
  return 0;
}

/* nsteps steps starting at loop index step0; recs: [nsteps][n_dust] */
int orc_run(const od_config* c, int n, double* x, double* v, double* q, int64_t step0, int64_t nsteps,
            od_dust_record* recs, int nthreads) {
  for (int64_t s = 0; s < nsteps; ++s) {
    const int rc = orc_step(c, n, x, v, q, step0 + s, recs + s * c->n_dust, NULL, NULL, nthreads);
    if (rc) return rc;
  }
  return 0;
}

/* --------------------------------------------------- trap Monte Carlo */
void orc_trap_sample(double R, double H, uint64_t seed, int64_t k, double* xyz) {
  uint32_t a[4], b[4];
  const uint32_t ca[4] = {(uint32_t)k, (uint32_t)((uint64_t)k >> 32), 0u, ST_TRAP};
  const uint32_t cb[4] = {(uint32_t)k, (uint32_t)((uint64_t)k >> 32), 0u, ST_TRAP | (1u << 8)};
  orc_philox(ca, seed, a);
  orc_philox(cb, seed, b);
  const double r = sqrt(unif53(a[0], a[1]));
  const double theta = unif53(a[2], a[3]);
  xyz[0] = R * r * cos(2.0 * 3.141592653589793 * theta);
  xyz[1] = R * r * sin(2.0 * 3.141592653589793 * theta);
  xyz[2] = -H / 2.0 + H * unif53(b[0], b[1]);
}

/* out[n_r][n_h] = mean over n_mc samples of exp(-d)/(d + 0.001) (opendust.py:515-523) */
int orc_trap_grid(double R, double H, int n_r, const double* rho, int n_h, const double* h, int64_t n_mc,
                  uint64_t seed, double* out, int nthreads) {
  double* s = (double*)malloc(sizeof(double) * 3 * (size_t)n_mc);
  if (!s) return 1;
  for (int64_t k = 0; k < n_mc; ++k) orc_trap_sample(R, H, seed, k, s + 3 * k);
#ifdef _OPENMP
#pragma omp parallel for collapse(2) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int i = 0; i < n_r; ++i)
    for (int j = 0; j < n_h; ++j) {
      double acc = 0.0;
      for (int64_t k = 0; k < n_mc; ++k) {
        const double d = sqrt((rho[i] - s[3 * k]) * (rho[i] - s[3 * k]) + s[3 * k + 1] * s[3 * k + 1] +
                              (h[j] - s[3 * k + 2]) * (h[j] - s[3 * k + 2]));
        acc += exp(-d) / (d + 0.001);
      }
      out[i * n_h + j] = acc / (double)n_mc;
    }
  free(s);
  return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* layout check against the ctypes mirror */
int64_t orc_config_size(void) { return (int64_t)sizeof(od_config); }

/* cell of a position under the engine's binning specification (opendust_amd.h):
 * columns of cutoff/4 in x and y, slabs of cutoff/16 in z, z fastest */
int orc_cell_of(const od_config* c, const double* x, int* dims) {
  const double w = c->cutoff / 4.0, h = c->cutoff / 16.0;
  const int nx = (int)ceil(2.0 * c->R / w), ny = nx, nz = (int)ceil(c->H / h);
  int ix = (int)floor((x[0] + c->R) / w), iy = (int)floor((x[1] + c->R) / w), iz = (int)floor((x[2] + c->H / 2.0) / h);
  ix = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
  iz = iz < 0 ? 0 : (iz >= nz ? nz - 1 : iz);
  if (dims) {
    dims[0] = nx;
    dims[1] = ny;
    dims[2] = nz;
  }
  return iz + nz * (ix + nx * iy);
}

/* Axial ion-density profile of one frame (tools/concentration.cpp:32-101):
 * n nodes z_grid[i] = -L/2 + i L/(n-1); every ion (z in nm, converted as the
 * tool does, z/1e9) deposits linear (cloud-in-cell) weights on the two nodes
 * around it, in ion order, except that an ion in the first interval only feeds
 * node 1 and one in the last interval only node n-2; ions below -L/2 or at or
 * beyond the last node are ignored.  Interior nodes are divided by
 * 0.5 (dz_left + dz_right) pi R^2 / N_s, the end nodes extrapolated linearly.
 * R, L in m; out: z_grid[n] (m), conc[n] (1/m^3). */
void orc_density_profile(const double* z_nm, int64_t N, int n, double R, double L, double N_s, double* z_grid,
                         double* conc) {
  for (int i = 0; i < n; i++) {
    z_grid[i] = -L / 2.0 + i * L / (double)(n - 1);
    conc[i] = 0.0;
  }
  for (int64_t j = 0; j < N; j++) {
    const double z = z_nm[j] / 1e9;
    const int index = (int)((0.5 * L + z) * (n - 1) / L);
    if (index == 0 && z >= -0.5 * L) {
      conc[1] += 1.0 - (z_grid[1] - z) / (z_grid[1] - z_grid[0]);
    } else if (index == n - 2) {
      conc[n - 2] += 1.0 - (z - z_grid[n - 2]) / (z_grid[n - 1] - z_grid[n - 2]);
    } else if (index < n - 2 && index > 0) {
      conc[index] += 1.0 - (z - z_grid[index]) / (z_grid[index + 1] - z_grid[index]);
      conc[index + 1] += 1.0 - (z_grid[index + 1] - z) / (z_grid[index + 1] - z_grid[index]);
    }
  }
  const double pi = 3.14159265358979323846;
  for (int i = 1; i < n - 1; i++)
    conc[i] /= 0.5 * ((z_grid[i + 1] - z_grid[i]) + (z_grid[i] - z_grid[i - 1])) * (pi * R * R) / N_s;
  conc[0] = 2 * conc[1] - conc[2];
  conc[n - 1] = 2 * conc[n - 2] - conc[n - 3];
}
