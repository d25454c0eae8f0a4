/*
 * FP64 C port of braincell's staggered step for the CPU baseline.
 * Test infrastructure only (timed CPU baseline and second oracle route).
 *
 * Per step and cell, in the reference order (braincell/quad/_staggered.py:126-154):
 *   1. Ca pool current at (V_n, gates_n, Ci_n)   (cell.py:2261-2280)
 *   2. linearisation: linear = -sum(g_eff)/Cm, const = f - V linear
 *      (_staggered.py:763-795, cell.py:1907-1963)
 *   3. DHS numeric state (_staggered.py:449-513), leaf-to-root elimination over
 *      the level-chunked edge list (comp_triang_raw, :641-657) and either the
 *      recursive-doubling back substitution (comp_backsub_raw, :742-760) or
 *      the equivalent sequential sweep x_i = s_i + e_i x_parent (faster on a CPU)
 *   4. Ca pool by exp-Euler with the cached current (ion/calcium.py:499-503)
 *   5. gates by exp-Euler at V_{n+1} (_exp_euler.py:437-475)
 * Cells are independent: OpenMP parallel loop over cells.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int C, P, n_edges, n_levels, n_jump;
  const int* dyn;       /* [C] CV -> row */
  const int* parent;    /* [P] parent row or -1 */
  const double* diag;   /* [P] ms^-1 */
  const double* lowers; /* [P] */
  const double* uppers; /* [P] */
  const int* edges;     /* [2*n_edges] (child, parent) */
  const int* level_off; /* [n_levels + 1] */
  const int* jumps;     /* [n_jump * (P + 1)] */
  const double* cm;     /* [C] uF/cm^2 */
  const double* g;      /* [C * 8] mS/cm^2: Na, KDR, IL, KL, CaL, CaT, AHP, HCN */
  const double* camask; /* [C] */
  const double* clamp;  /* [C] uA/cm^2 while the clamp is on */
  int clamp_on, clamp_off;
  int soma_cv;
  double* V;            /* [C] in/out */
  double* S;            /* [NS * C] in/out, slots as the HIP kernel */
} port_cell;

typedef struct {
  double E_Na, E_K, E_IL, na_vsh, na_phi, kdr_vsh, kdr_phi, cal_vsh, cal_phi_p, cal_phi_q;
  double ca_drive, ca_tau_inv, ca_rest, ca_Co, ca_rtzf;
} port_const;

static double exprel_(double x) {
  if (fabs(x) < 1e-8) return 1.0 + x * (0.5 + x / 6.0);
  return expm1(x) / x;
}
static double ee(double y, double f, double lam, double dt) { return y + dt * exprel_(dt * lam) * f; }
static double gab(double x, double a, double b, double phi, double dt) {
  return ee(x, phi * (a * (1.0 - x) - b * x), -phi * (a + b), dt);
}
static double git(double x, double inf, double tau, double phi, double dt) {
  return ee(x, phi * (inf - x) / tau, -phi / tau, dt);
}

/* Slots for the channel set Na | KDR | IL | CaL | pool: Na p,q; KDR p; CaL p,q; Ci. */
static void step_cell(const port_cell* c, const port_const* k, double dt, int n, int sequential, double* w) {
  const int C = c->C, P = c->P;
  double *diags = w, *solves = w + (P + 1), *lo = w + 2 * (P + 1), *up = w + 3 * (P + 1);
  double *ica = w + 4 * (P + 1), *lin = ica + C, *cst = lin + C, *le = cst + C, *se = le + (P + 1);
  double* V = c->V;
  double* S = c->S;
  double *na_p = S, *na_q = S + C, *kd_p = S + 2 * C, *cl_p = S + 3 * C, *cl_q = S + 4 * C, *ci = S + 5 * C;
  const int clamp_on = (n >= c->clamp_on && n < c->clamp_off);
  for (int i = 0; i < C; ++i) {
    const double* g = c->g + 8 * i;
    double v = V[i];
    double eca = k->ca_rtzf * log(k->ca_Co / ci[i]);
    double gna = g[0] * na_p[i] * na_p[i] * na_p[i] * na_q[i];
    double p2 = kd_p[i] * kd_p[i];
    double gk = g[1] * p2 * p2;
    double gl = g[2];
    double gca = g[4] * cl_p[i] * cl_p[i] * cl_q[i];
    double I = gna * (k->E_Na - v) + gk * (k->E_K - v) + gl * (k->E_IL - v) + gca * (eca - v);
    ica[i] = c->camask[i] * gca * (eca - v);
    if (clamp_on) I += c->clamp[i];
    double f = I / c->cm[i];
    lin[i] = -(gna + gk + gl + gca) / c->cm[i];
    cst[i] = f - v * lin[i];
  }
  for (int r = 0; r < P; ++r) {
    diags[r] = c->diag[r] * dt;
    solves[r] = 0.0;
    lo[r] = c->lowers[r] * dt;
    up[r] = c->uppers[r] * dt;
  }
  diags[P] = 1.0;
  solves[P] = 0.0;
  lo[P] = 0.0;
  up[P] = 0.0;
  for (int i = 0; i < C; ++i) {
    int r = c->dyn[i];
    diags[r] += 1.0 - dt * lin[i];
    solves[r] = V[i] + dt * cst[i];
  }
  for (int l = 0; l < c->n_levels; ++l)
    for (int e = c->level_off[l]; e < c->level_off[l + 1]; ++e) {
      int ch = c->edges[2 * e], pa = c->edges[2 * e + 1];
      double m = up[ch] / diags[ch];
      diags[pa] -= lo[ch] * m;
      solves[pa] -= solves[ch] * m;
    }
  if (sequential) {
    /* rows are parent-before-child (peel-descending order) */
    for (int r = 0; r < P; ++r) {
      int p = c->parent[r];
      double x = solves[r] / diags[r];
      if (p >= 0) x += (-lo[r] / diags[r]) * se[p];
      se[r] = x;
    }
  } else {
    lo[0] = 0.0;
    for (int r = 0; r <= P; ++r) {
      le[r] = -lo[r] / diags[r];
      se[r] = solves[r] / diags[r];
    }
    double* tmp_s = ica + 0; /* reuse: not needed after the solve */
    (void)tmp_s;
    for (int j = 0; j < c->n_jump; ++j) {
      const int* kp = c->jumps + (size_t)j * (P + 1);
      /* pointer jumping reads the previous iterate: two passes via diags/solves as scratch */
      for (int r = 0; r <= P; ++r) {
        solves[r] = se[r] + le[r] * se[kp[r]];
        diags[r] = le[r] * le[kp[r]];
      }
      memcpy(se, solves, sizeof(double) * (P + 1));
      memcpy(le, diags, sizeof(double) * (P + 1));
    }
  }
  for (int i = 0; i < C; ++i) {
    double v = se[c->dyn[i]];
    double cold = ci[i];
    double drive = ica[i] * k->ca_drive;
    if (drive < 0.0) drive = 0.0;
    double cn = ee(cold, drive + (k->ca_rest - cold) * k->ca_tau_inv, -k->ca_tau_inv, dt);
    ci[i] = cn;
    double vp = v - k->na_vsh;
    double ap = 1.28 / exprel_(-(vp - 13.0) / 4.0), bp = 1.4 / exprel_((vp - 40.0) / 5.0);
    double aq = 0.128 * exp(-(vp - 17.0) / 18.0), bq = 4.0 / (1.0 + exp(-(vp - 40.0) / 5.0));
    na_p[i] = gab(na_p[i], ap, bp, k->na_phi, dt);
    na_q[i] = gab(na_q[i], aq, bq, k->na_phi, dt);
    double vk = v - k->kdr_vsh;
    double ak = 0.16 / exprel_(-(vk - 15.0) / 5.0), bk = 0.5 * exp(-(vk - 10.0) / 40.0);
    kd_p[i] = gab(kd_p[i], ak, bk, k->kdr_phi, dt);
    double vc = v - k->cal_vsh;
    double pi = 1.0 / (1.0 + exp(-(vc + 10.0) / 4.0));
    double pt = 0.4 + 0.7 / (exp(-(vc + 5.0) / 15.0) + exp((vc + 5.0) / 15.0));
    double qi = 1.0 / (1.0 + exp((vc + 25.0) / 2.0));
    double qt = 300.0 + 100.0 / (exp((vc + 40.0) / 9.5) + exp(-(vc + 40.0) / 9.5));
    cl_p[i] = git(cl_p[i], pi, pt, k->cal_phi_p, dt);
    cl_q[i] = git(cl_q[i], qi, qt, k->cal_phi_q, dt);
    V[i] = v;
  }
}

/* Runs n_steps steps for every cell; soma voltages go to trace[step * n_cells + cell].
 * Returns 0, or -1 when a workspace allocation fails. */
int port_run(int n_cells, const port_cell* cells, const port_const* k, double dt, int step0, int n_steps,
             double* trace, int sequential, int threads) {
  int err = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int ci = 0; ci < n_cells; ++ci) {
    const port_cell* c = &cells[ci];
    double* w = (double*)malloc(sizeof(double) * (6 * (size_t)(c->P + 1) + 3 * (size_t)c->C));
    if (!w) {
      err = -1;
      continue;
    }
    for (int s = 0; s < n_steps; ++s) {
      step_cell(c, k, dt, step0 + s, sequential, w);
      if (trace) trace[(size_t)s * n_cells + ci] = c->V[c->soma_cv];
    }
    free(w);
  }
  return err;
}

int port_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
