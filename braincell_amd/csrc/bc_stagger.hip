// Persistent staggered-step kernel for gfx950.
//
// One workgroup integrates one *pack*: a forest of whole cells whose DHS
// rows fit in LDS.  Per step:
//
//   A. (one thread per CV, gate/pool states in registers) linearise the
//      membrane current at V_n: write diag = 1 + dt/Cm * sum(g_eff) and
//      rhs = dt/Cm * (sum(g_eff (E - V_n)) + I_clamp) into LDS; remember the
//      Ca pool current at (V_n, gates_n, Ci_n).  Boundary rows get diag 0 and
//      rhs 0 (+ any clamp on them).
//   B. Hines solve on the row tree in LDS for dV = V_{n+1} - V_n.  The
//      reference solves (I + dt A - dt L) V_{n+1} = V_n + dt (f - L V_n); the
//      same system written for dV has the right-hand side dt (f - A V_n), so
//      the axial defect dt A V_n is folded in edge by edge (from voltage
//      differences) while a thread walks its chain.  In FP32 this keeps the
//      solve's rounding relative to dV rather than V: the fat soma couples
//      its CVs with dt g / C ~ 1e3, which costs three digits otherwise.
//      Rows are grouped into chains (runs whose parent has one child); a
//      thread eliminates a whole chain bottom-to-top in registers, passing
//      its top row's contribution to the parent junction through LDS.
//      Chains are processed level by level, leaves first; back-substitution
//      runs the levels in reverse.  Boundary-row voltages persist between
//      steps (the constraint rows are re-imposed by every solve).
//   C. (same CV threads) V += dV; Ca pool by exp-Euler with the cached
//      current, then every gate by exp-Euler at V_{n+1} with E_Ca and Ci at
//      n+1; write probe voltages to the trace.
//
// Reference semantics: braincell/quad/_staggered.py:126-154 (step order),
// :449-513 (numeric state), :641-657 (elimination), :742-760 (back
// substitution, recomputed here by a sequential sweep that yields the same
// x_i = s_i + e_i x_parent), _multi_compartment/cell.py:1907-1963
// (linearizer), :2157-2217 (family order), :2261-2280 (Ca current cache).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc_channels.h"
#include "bc_common.h"

namespace {

template <typename real>
struct StaggerArgs {
  const bc_pack* packs;
  const uint32_t* cv_info;
  const uint32_t* bnd_info;
  const int32_t* chain_off;
  const int32_t* chain_top;
  const int32_t* kid_off;
  const int32_t* kids;
  const int32_t* level_off;
  const real* lo;
  const real* up;
  const real* cls_tab;
  const int32_t* clamp_row;
  const int32_t* clamp_on;
  const int32_t* clamp_off;
  const real* clamp_amp;
  const int32_t* probe_row;
  const int32_t* probe_col;
  real* V;
  real* Vbnd;
  real* state;
  int64_t n_cv_total;
  real* trace;
  int64_t trace_ld;
  real* rec;  // optional: [(n_steps + 1) * (NS + 1) * n_cv_total] step-start V and states
  uint32_t* status;
  int32_t step0;
  int32_t n_steps;
  real dt;
  int32_t max_rows, max_chains, max_cls;
  bc_chan_const<real> cc;
};

constexpr uint32_t kRowMask = 0xFFFFu;
constexpr uint32_t kProbeBit = 1u << 30;
constexpr uint32_t kClampBit = 1u << 31;

template <typename real>
__device__ __forceinline__ real clamp_current(const StaggerArgs<real>& a, const bc_pack& P, int row, int gstep) {
  real I = real(0);
  for (int i = 0; i < P.n_clamp; ++i) {
    int e = P.clamp0 + i;
    if (a.clamp_row[e] == row && gstep >= a.clamp_on[e] && gstep < a.clamp_off[e]) I += a.clamp_amp[e];
  }
  return I;
}

// Linearisation of one CV (phase A): G = sum(g_eff), GE = sum(g_eff (E - V)).
// Returns the Ca pool current (masked to the pool's region).
template <typename real, uint32_t CH>
__device__ __forceinline__ real linearise(const bc_chan_const<real>& cc, const real* cp, const real* s, real V,
                                          real& G, real& GE) {
  typedef bc_math<real> M;
  G = real(0);
  GE = real(0);
  real ica = real(0);
  real eca = cc.E_Ca_fixed;
  if constexpr ((CH & BC_CH_POOL) != 0) eca = cc.ca_rtzf * M::lg(bc_div<real>(cc.ca_Co, s[bc_slot_ca(CH)]));
  if constexpr ((CH & BC_CH_NA) != 0) {
    real p = s[bc_slot_na(CH)], q = s[bc_slot_na(CH) + 1];
    real g = cp[BC_CLS_G_NA] * (p * p * p) * q;
    G += g;
    GE += g * (cc.E_Na - V);
  }
  if constexpr ((CH & BC_CH_KDR) != 0) {
    real p = s[bc_slot_kdr(CH)];
    real p2 = p * p;
    real g = cp[BC_CLS_G_KDR] * (p2 * p2);
    G += g;
    GE += g * (cc.E_K - V);
  }
  if constexpr ((CH & BC_CH_IL) != 0) {
    real g = cp[BC_CLS_G_IL];
    G += g;
    GE += g * (cc.E_IL - V);
  }
  if constexpr ((CH & BC_CH_KL) != 0) {
    real g = cp[BC_CLS_G_KL];
    G += g;
    GE += g * (cc.E_K - V);
  }
  if constexpr ((CH & BC_CH_CAL) != 0) {
    real p = s[bc_slot_cal(CH)], q = s[bc_slot_cal(CH) + 1];
    real g = cp[BC_CLS_G_CAL] * (p * p) * q;
    G += g;
    GE += g * (eca - V);
    ica += g * (eca - V);
  }
  if constexpr ((CH & BC_CH_CAT) != 0) {
    real p = s[bc_slot_cat(CH)], q = s[bc_slot_cat(CH) + 1];
    real g = cp[BC_CLS_G_CAT] * (p * p) * q;
    G += g;
    GE += g * (eca - V);
    ica += g * (eca - V);
  }
  if constexpr ((CH & BC_CH_AHP) != 0) {
    real p = s[bc_slot_ahp(CH)];
    real g = cp[BC_CLS_G_AHP] * (p * p);
    G += g;
    GE += g * (cc.E_K - V);
  }
  if constexpr ((CH & BC_CH_HCN) != 0) {
    real p = s[bc_slot_hcn(CH)];
    real g = cp[BC_CLS_G_HCN] * p;
    G += g;
    GE += g * (cc.hcn_E - V);
  }
  return ica * cp[BC_CLS_CAMASK];
}

// Phase C: Ca pool then gates at V_{n+1}.
// The exp-Euler update of every gate is exact for its linear ODE at frozen
// V, so it is written in closed form x_inf + (x - x_inf) e^{dt lam}.  The Ca
// pool's lam = -1/tau is constant on this path: ca_b = dt exprel(-dt/tau).
template <typename real, uint32_t CH>
__device__ __forceinline__ void advance_states(const bc_chan_const<real>& cc, real* s, real V, real ica, real dt,
                                               real ca_b) {
  typedef bc_math<real> M;
  real ci = real(0);
  if constexpr ((CH & BC_CH_POOL) != 0) {
    real c = s[bc_slot_ca(CH)];
    real drive = ica * cc.ca_drive;
    drive = drive > real(0) ? drive : real(0);
    real f = drive + (cc.ca_rest - c) * cc.ca_tau_inv;
    ci = c + ca_b * f;
    s[bc_slot_ca(CH)] = ci;
  }
  if constexpr ((CH & BC_CH_NA) != 0) {
    real ap, bp, aq, bq;
    bc_na_rates<real>(V, cc.na_vsh, ap, bp, aq, bq);
    s[bc_slot_na(CH)] = bc_gate_ab_cf<real>(s[bc_slot_na(CH)], ap, bp, cc.na_phi, dt);
    s[bc_slot_na(CH) + 1] = bc_gate_ab_cf<real>(s[bc_slot_na(CH) + 1], aq, bq, cc.na_phi, dt);
  }
  if constexpr ((CH & BC_CH_KDR) != 0) {
    real ap, bp;
    bc_kdr_rates<real>(V, cc.kdr_vsh, ap, bp);
    s[bc_slot_kdr(CH)] = bc_gate_ab_cf<real>(s[bc_slot_kdr(CH)], ap, bp, cc.kdr_phi, dt);
  }
  if constexpr ((CH & BC_CH_CAL) != 0) {
    real pi, pt, qi, qt;
    bc_cal_rates<real>(V, cc.cal_vsh, pi, pt, qi, qt);
    s[bc_slot_cal(CH)] = bc_gate_it_cf<real>(s[bc_slot_cal(CH)], pi, pt, cc.cal_phi_p, dt);
    s[bc_slot_cal(CH) + 1] = bc_gate_it_cf<real>(s[bc_slot_cal(CH) + 1], qi, qt, cc.cal_phi_q, dt);
  }
  if constexpr ((CH & BC_CH_CAT) != 0) {
    real pi, pt, qi, qt;
    bc_cat_rates<real>(V, cc.cat_vsh, pi, pt, qi, qt);
    s[bc_slot_cat(CH)] = bc_gate_it_cf<real>(s[bc_slot_cat(CH)], pi, pt, cc.cat_phi_p, dt);
    s[bc_slot_cat(CH) + 1] = bc_gate_it_cf<real>(s[bc_slot_cat(CH) + 1], qi, qt, cc.cat_phi_q, dt);
  }
  if constexpr ((CH & BC_CH_AHP) != 0) {
    real a = cc.ahp_alpha * M::pw(ci, cc.ahp_n);
    s[bc_slot_ahp(CH)] = bc_gate_ab_cf<real>(s[bc_slot_ahp(CH)], a, cc.ahp_beta, cc.ahp_phi, dt);
  }
  if constexpr ((CH & BC_CH_HCN) != 0) {
    real pi, pt;
    bc_hcn_rates<real>(V, pi, pt);
    s[bc_slot_hcn(CH)] = bc_gate_it_cf<real>(s[bc_slot_hcn(CH)], pi, pt, cc.hcn_phi, dt);
  }
}

// Step-start snapshot for checkpoint-and-recompute: slot 0 = V, slots 1.. =
// gate/pool states, laid out [step][slot][cv].
template <typename real, int MAXC, int NS>
__device__ __forceinline__ void record_state(const StaggerArgs<real>& a, const bc_pack& P, int tid, int NT, int step,
                                             const real* vv, const real (*st)[NS], const uint32_t* info) {
  const int64_t base = (int64_t)step * (NS + 1) * a.n_cv_total;
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    if (info[i] != 0xFFFFFFFFu) {
      const int64_t g = (int64_t)P.cv0 + tid + i * NT;
      a.rec[base + g] = vv[i];
#pragma unroll
      for (int s = 0; s < NS; ++s) a.rec[base + (int64_t)(s + 1) * a.n_cv_total + g] = st[i][s];
    }
  }
}

template <typename real, uint32_t CH, int MAXC>
__global__ void __launch_bounds__(256) stagger_kernel(StaggerArgs<real> a) {
  constexpr int NS = bc_nstates(CH) > 0 ? bc_nstates(CH) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  real* bd = reinterpret_cast<real*>(smem_raw);
  real* br = bd + a.max_rows;
  real* xv = br + a.max_rows;  // V_n at every slot
  real* LO = xv + a.max_rows;  // dt * A[slot, parent]
  real* UP = LO + a.max_rows;  // dt * A[parent, slot]
  real* kd = UP + a.max_rows;
  real* kr = kd + a.max_chains;
  real* cls = kr + a.max_chains;
  // chain metadata, 16-bit: first slot of each chain (+ end), top parent slot,
  // kid list offsets (+ end), kid chains, level offsets (+ end)
  int16_t* m_j0 = reinterpret_cast<int16_t*>(cls + a.max_cls * BC_NCLS);
  int16_t* m_tp = m_j0 + (a.max_chains + 1);
  int16_t* m_k0 = m_tp + a.max_chains;
  int16_t* m_kid = m_k0 + (a.max_chains + 1);
  int16_t* m_lv = m_kid + a.max_chains;

  const bc_pack P = a.packs[blockIdx.x];
  const int tid = threadIdx.x;
  const int NT = blockDim.x;
  for (int i = tid; i < P.n_cls * BC_NCLS; i += NT) cls[i] = a.cls_tab[(int64_t)P.cls0 * BC_NCLS + i];
  for (int i = tid; i < P.n_rows; i += NT) {
    LO[i] = a.lo[(int64_t)P.row0 + i];
    UP[i] = a.up[(int64_t)P.row0 + i];
  }
  const int kbase = a.kid_off[P.ch0];
  for (int i = tid; i <= P.n_ch; i += NT) {
    m_j0[i] = (int16_t)(a.chain_off[P.ch0 + i] - P.row0);
    m_k0[i] = (int16_t)(a.kid_off[P.ch0 + i] - kbase);
    if (i < P.n_ch) m_tp[i] = (int16_t)a.chain_top[P.ch0 + i];
  }
  for (int i = tid; i < a.kid_off[P.ch0 + P.n_ch] - kbase; i += NT) m_kid[i] = (int16_t)a.kids[kbase + i];
  for (int i = tid; i <= P.n_lv; i += NT) m_lv[i] = (int16_t)a.level_off[P.lv0 + i];

  real st[MAXC][NS];
  real vv[MAXC];
  real ica[MAXC];
  uint32_t info[MAXC];
  int pcol[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    int c = tid + i * NT;
    info[i] = 0xFFFFFFFFu;
    pcol[i] = -1;
    vv[i] = real(0);
    ica[i] = real(0);
#pragma unroll
    for (int s = 0; s < NS; ++s) st[i][s] = real(0);
    if (c < P.n_cv) {
      int64_t g = (int64_t)P.cv0 + c;
      info[i] = a.cv_info[g];
      vv[i] = a.V[g];
      xv[a.cv_info[g] & kRowMask] = vv[i];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < bc_nstates(CH)) st[i][s] = a.state[(int64_t)s * a.n_cv_total + g];
      if (info[i] & kProbeBit) {
        int row = (int)(info[i] & kRowMask);
        for (int e = 0; e < P.n_probe; ++e)
          if (a.probe_row[P.probe0 + e] == row) pcol[i] = a.probe_col[P.probe0 + e];
      }
    }
  }
  for (int i = tid; i < P.n_bnd; i += NT) xv[a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu] = a.Vbnd[P.bnd0 + i];
  __syncthreads();

  const real dt = a.dt;
  const real ca_b = dt * bc_exprel<real>(-dt * a.cc.ca_tau_inv);
  uint32_t bad = 0;
  for (int step = 0; step < a.n_steps; ++step) {
    const int gstep = a.step0 + step;
    if (a.rec) record_state<real, MAXC, NS>(a, P, tid, NT, step, vv, st, info);
    // ---- phase A: linearise at V_n (midpoint rows)
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] != 0xFFFFFFFFu) {
        const int row = (int)(info[i] & kRowMask);
        const real* cp = cls + ((info[i] >> 16) & 0x3FFFu) * BC_NCLS;
        real G, GE;
        ica[i] = linearise<real, CH>(a.cc, cp, st[i], vv[i], G, GE);
        real Icl = (info[i] & kClampBit) ? clamp_current<real>(a, P, row, gstep) : real(0);
        const real dtcm = cp[BC_CLS_DTCM];
        bd[row] = real(1) + dtcm * G;
        br[row] = dtcm * (GE + Icl);
      }
    }
    // boundary (algebraic) rows
    for (int i = tid; i < P.n_bnd; i += NT) {
      uint32_t b = a.bnd_info[P.bnd0 + i];
      int row = (int)(b & 0x7FFFFFFFu);
      bd[row] = real(0);
      br[row] = (b & kClampBit) ? clamp_current<real>(a, P, row, gstep) : real(0);
    }
    __syncthreads();

    // ---- phase B: Hines elimination, leaves to root, one chain per thread.
    // Diagonal rebuilt from the zero row sums of the axial operator:
    // d_j = e_j - lo_j with e_j = diag_j - sum over children c of
    // up_c e_c / d_c >= 0, so no step subtracts nearly equal numbers.
    for (int l = 0; l < P.n_lv; ++l) {
      const int k0 = m_lv[l], k1 = m_lv[l + 1];
      for (int k = k0 + tid; k < k1; k += NT) {
        const int j0 = m_j0[k], j1 = m_j0[k + 1];
        const int tp = m_tp[k];
        real e = bd[j0];
        real r = br[j0];
        const int q0 = m_k0[k], q1 = m_k0[k + 1];
        for (int q = q0; q < q1; ++q) {
          const int kk = m_kid[q];
          e -= kd[kk];
          r -= kr[kk];
        }
        real vr = xv[j0];
        real kdc, krc, up, dl;
        for (int j = j0;;) {
          // axial defect of the edge to the parent: dt A V_n terms
          const int pr = (j + 1 < j1) ? j + 1 : tp;
          const real vp = pr >= 0 ? xv[pr] : vr;
          const real lo = LO[j];
          up = UP[j];
          dl = vr - vp;
          r += lo * dl;
          const real inv = bc_rcp<real>(e - lo);
          bd[j] = inv;
          br[j] = r;
          kdc = up * e * inv;
          krc = up * r * inv;
          if (++j >= j1) break;
          vr = vp;
          e = bd[j] - kdc;
          r = (br[j] - krc) - up * dl;
        }
        kd[k] = kdc;
        kr[k] = krc + up * dl;
      }
      __syncthreads();
    }
    // ---- back substitution, root to leaves
    for (int l = P.n_lv - 1; l >= 0; --l) {
      const int k0 = m_lv[l], k1 = m_lv[l + 1];
      for (int k = k0 + tid; k < k1; k += NT) {
        const int j0 = m_j0[k], j1 = m_j0[k + 1];
        const int tp = m_tp[k];
        real xp = tp >= 0 ? br[tp] : real(0);
        for (int j = j1 - 1; j >= j0; --j) {
          const real x = (br[j] - LO[j] * xp) * bd[j];
          br[j] = x;
          xp = x;
        }
      }
      __syncthreads();
    }

    // ---- phase C: V += dV; Ca pool and gates at V_{n+1}
    for (int i = tid; i < P.n_bnd; i += NT) {
      const int row = (int)(a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu);
      xv[row] += br[row];
    }
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] != 0xFFFFFFFFu) {
        const int row = (int)(info[i] & kRowMask);
        const real Vn = vv[i] + br[row];
        xv[row] = Vn;
        advance_states<real, CH>(a.cc, st[i], Vn, ica[i], dt, ca_b);
        vv[i] = Vn;
        bad |= (Vn == Vn && fabs(Vn) < real(1e30)) ? 0u : 1u;
        if (pcol[i] >= 0) a.trace[(int64_t)step * a.trace_ld + pcol[i]] = Vn;
      }
    }
  }

  if (a.rec) record_state<real, MAXC, NS>(a, P, tid, NT, a.n_steps, vv, st, info);
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    if (info[i] != 0xFFFFFFFFu) {
      int64_t g = (int64_t)P.cv0 + tid + i * NT;
      a.V[g] = vv[i];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < bc_nstates(CH)) a.state[(int64_t)s * a.n_cv_total + g] = st[i][s];
    }
  }
  for (int i = tid; i < P.n_bnd; i += NT) a.Vbnd[P.bnd0 + i] = xv[a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu];
  if (bad) atomicOr(a.status, 1u);
}

// ------------------------------------------------------------------ dispatch
constexpr uint32_t kSetPassive = BC_CH_IL;
constexpr uint32_t kSetHH = BC_CH_NA | BC_CH_KDR | BC_CH_IL;
constexpr uint32_t kSetHHCaL = BC_CH_NA | BC_CH_KDR | BC_CH_IL | BC_CH_CAL | BC_CH_POOL;
constexpr uint32_t kSetAll = BC_CH_NA | BC_CH_KDR | BC_CH_IL | BC_CH_KL | BC_CH_CAL | BC_CH_CAT | BC_CH_AHP |
                             BC_CH_HCN | BC_CH_POOL;

typedef void (*kernel_f32)(StaggerArgs<float>);
typedef void (*kernel_f64)(StaggerArgs<double>);

template <typename real, uint32_t CH>
void* pick_maxc(int maxc) {
  switch (maxc) {
    case 1: return (void*)&stagger_kernel<real, CH, 1>;
    case 2: return (void*)&stagger_kernel<real, CH, 2>;
    case 4: return (void*)&stagger_kernel<real, CH, 4>;
    case 8: return (void*)&stagger_kernel<real, CH, 8>;
    case 16: return (void*)&stagger_kernel<real, CH, 16>;
    default: return nullptr;
  }
}

template <typename real>
void* pick_kernel(uint32_t ch, int maxc) {
  switch (ch) {
    case kSetPassive: return pick_maxc<real, kSetPassive>(maxc);
    case kSetHH: return pick_maxc<real, kSetHH>(maxc);
    case kSetHHCaL: return pick_maxc<real, kSetHHCaL>(maxc);
    case kSetAll: return pick_maxc<real, kSetAll>(maxc);
    default: return nullptr;
  }
}

template <typename real>
int fill_args(const bc_stagger_desc* d, StaggerArgs<real>* a) {
  a->packs = d->packs;
  a->cv_info = d->cv_info;
  a->bnd_info = d->bnd_info;
  a->chain_off = d->chain_off;
  a->chain_top = d->chain_top;
  a->kid_off = d->kid_off;
  a->kids = d->kids;
  a->level_off = d->level_off;
  a->lo = static_cast<const real*>(d->lo);
  a->up = static_cast<const real*>(d->up);
  a->cls_tab = static_cast<const real*>(d->cls_tab);
  a->clamp_row = d->clamp_row;
  a->clamp_on = d->clamp_on;
  a->clamp_off = d->clamp_off;
  a->clamp_amp = static_cast<const real*>(d->clamp_amp);
  a->probe_row = d->probe_row;
  a->probe_col = d->probe_col;
  a->V = static_cast<real*>(d->V);
  a->Vbnd = static_cast<real*>(d->Vbnd);
  a->state = static_cast<real*>(d->state);
  a->n_cv_total = d->n_cv_total;
  a->trace = static_cast<real*>(d->trace);
  a->trace_ld = d->trace_ld;
  a->rec = static_cast<real*>(d->rec);
  a->status = d->status;
  a->step0 = d->step0;
  a->n_steps = d->n_steps;
  a->dt = static_cast<real>(d->dt);
  a->max_rows = d->max_rows;
  a->max_chains = d->max_chains;
  a->max_cls = d->max_cls;
  if (d->chan_const == nullptr) return bc::fail(BC_INVALID_ARG, "chan_const is NULL");
  a->cc = *static_cast<const bc_chan_const<real>*>(d->chan_const);
  return BC_OK;
}

int validate(const bc_stagger_desc* d) {
  if (d == nullptr) return bc::fail(BC_INVALID_ARG, "desc is NULL");
  if (d->precision != 32 && d->precision != 64) return bc::fail(BC_INVALID_ARG, "precision must be 32 or 64");
  if (d->n_packs < 0) return bc::fail(BC_INVALID_ARG, "n_packs < 0");
  if (d->threads < 64 || d->threads > 256 || d->threads % 64 != 0)
    return bc::fail(BC_INVALID_ARG, "threads must be 64, 128, 192 or 256 (got %d)", d->threads);
  if (d->n_steps < 0) return bc::fail(BC_INVALID_ARG, "n_steps < 0");
  if (d->max_rows <= 0 || d->max_rows > 65536) return bc::fail(BC_INVALID_LAYOUT, "max_rows out of range");
  if (d->max_chains <= 0 || d->max_chains > 32767 || d->max_cls <= 0 || d->max_cls > 16384)
    return bc::fail(BC_INVALID_LAYOUT, "max_chains/max_cls out of range");
  if (d->max_rows > 32767 || d->max_levels <= 0 || d->max_levels > 32767)
    return bc::fail(BC_INVALID_LAYOUT, "max_rows/max_levels out of range for 16-bit chain metadata");
  if (!(d->dt > 0.0)) return bc::fail(BC_INVALID_ARG, "dt must be > 0");
  if (d->n_packs > 0 && (d->packs == nullptr || d->cv_info == nullptr || d->V == nullptr || d->Vbnd == nullptr ||
                         d->status == nullptr))
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  return BC_OK;
}

template <typename real>
int launch(const bc_stagger_desc* d, hipStream_t s, int32_t step0, int32_t n_steps, void* trace) {
  void* fn = pick_kernel<real>(d->channels, d->cv_per_thread);
  if (fn == nullptr)
    return bc::fail(BC_UNSUPPORTED, "no kernel for channels=0x%x cv_per_thread=%d", d->channels, d->cv_per_thread);
  StaggerArgs<real> a;
  int rc = fill_args<real>(d, &a);
  if (rc != BC_OK) return rc;
  a.step0 = step0;
  a.n_steps = n_steps;
  a.trace = static_cast<real*>(trace);
  size_t lds = (size_t)(5 * d->max_rows + 2 * d->max_chains + d->max_cls * BC_NCLS) * sizeof(real) +
               (size_t)(4 * d->max_chains + 3 + d->max_levels + 1) * sizeof(int16_t);
  if (lds > 160 * 1024) return bc::fail(BC_INVALID_LAYOUT, "pack needs %zu bytes of LDS (> 160 KiB)", lds);
  if (lds > 64 * 1024) BC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (d->n_packs == 0 || n_steps == 0) return BC_OK;
  void* args[] = {&a};
  BC_HIP(hipLaunchKernel(fn, dim3(d->n_packs), dim3(d->threads), args, lds, s));
  return BC_OK;
}

int launch_any(const bc_stagger_desc* d, hipStream_t s, int32_t step0, int32_t n_steps, void* trace) {
  int rc = validate(d);
  if (rc != BC_OK) return rc;
  if (d->precision == 32) return launch<float>(d, s, step0, n_steps, trace);
  return launch<double>(d, s, step0, n_steps, trace);
}

struct GraphHandle {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

}  // namespace

extern "C" {

int bc_version(void) { return 100; }

const char* bc_last_error(void) { return bc::err_buf(); }

int bc_stagger_supported(int precision, uint32_t channels, int cv_per_thread) {
  if (precision == 32) return pick_kernel<float>(channels, cv_per_thread) != nullptr;
  if (precision == 64) return pick_kernel<double>(channels, cv_per_thread) != nullptr;
  return 0;
}

int bc_stagger_run(const bc_stagger_desc* desc, void* stream) {
  int rc = launch_any(desc, static_cast<hipStream_t>(stream), desc ? desc->step0 : 0, desc ? desc->n_steps : 0,
                      desc ? desc->trace : nullptr);
  if (rc != BC_OK) return rc;
  BC_HIP(hipGetLastError());
  return bc::ok();
}

int bc_stagger_graph_create(const bc_stagger_desc* desc, int32_t n_chunks, int32_t steps_per_chunk, void* stream,
                            void** graph_out) {
  if (graph_out == nullptr) return bc::fail(BC_INVALID_ARG, "graph_out is NULL");
  *graph_out = nullptr;
  int rc = validate(desc);
  if (rc != BC_OK) return rc;
  if (n_chunks <= 0 || steps_per_chunk <= 0) return bc::fail(BC_INVALID_ARG, "n_chunks/steps_per_chunk must be > 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t esz = desc->precision == 32 ? sizeof(float) : sizeof(double);
  BC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int c = 0; c < n_chunks; ++c) {
    char* tr = static_cast<char*>(desc->trace);
    void* trc = tr ? (void*)(tr + (size_t)c * steps_per_chunk * desc->trace_ld * esz) : nullptr;
    rc = launch_any(desc, s, desc->step0 + c * steps_per_chunk, steps_per_chunk, trc);
    if (rc != BC_OK) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
  }
  GraphHandle* h = new GraphHandle();
  hipError_t e = hipStreamEndCapture(s, &h->graph);
  if (e != hipSuccess) {
    delete h;
    return bc::fail(BC_HIP_ERROR, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(h->graph);
    delete h;
    return bc::fail(BC_HIP_ERROR, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *graph_out = h;
  return bc::ok();
}

int bc_graph_launch(void* graph, void* stream) {
  if (graph == nullptr) return bc::fail(BC_INVALID_ARG, "graph is NULL");
  GraphHandle* h = static_cast<GraphHandle*>(graph);
  BC_HIP(hipGraphLaunch(h->exec, static_cast<hipStream_t>(stream)));
  return bc::ok();
}

int bc_graph_destroy(void* graph) {
  if (graph == nullptr) return bc::ok();
  GraphHandle* h = static_cast<GraphHandle*>(graph);
  (void)hipGraphExecDestroy(h->exec);
  (void)hipGraphDestroy(h->graph);
  delete h;
  return bc::ok();
}

}  // extern "C"
