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
  int32_t max_rows, max_chains, max_cls, max_levels;
  uint32_t flags;  // BC_STEP_VOLTAGE_ONLY arrives as BC_STEP_SKIP_GATES with ca_b = 0
  real ca_b;       // Ca pool increment factor dt exprel(-dt / tau); 0 freezes the pool
  bc_chan_const<real> cc;
  // synapses and spike delivery (network runs; unused by the plain kernels)
  const int32_t* syn_pack_off;
  const int32_t* syn_row;
  const int32_t* syn_row_off;
  const real* syn_par;
  real* syn_state;
  const int32_t* con_off;
  const int32_t* con_pre;
  const int32_t* con_delay;
  const real* con_w;
  const int32_t* cv_cell;
  uint8_t* spk;
  int32_t spk_len, n_cells;
  // external drive (NetStim trains, bound inputs): NULL ext_slot = none
  const int32_t* ext_slot;
  const real* ext_drive;
  int32_t n_ext, ext_step0, ext_steps;
  real* syn_trace;  // A, B of every synapse after each step, or NULL
  int32_t syn_n;    // synapses (row length of syn_trace / 2, of syn_drive)
  const real* syn_drive;  // [n_steps * syn_n] drive of the launch's steps, or NULL (none)
  int32_t max_syn, max_syn_rows;  // LDS bound of a pack's synapses / synapse rows
  real v_th;
  // chain tables in LDS form (bc_stagger_desc.chain_rec ..)
  const int4* chain_rec;
  const int16_t* chain_par;
  const int16_t* kid_pad;
  const int16_t* slot_link;
  const int16_t* slot_chain;
  const int16_t* chain_work;  // optional work order of the per-chain passes
  const uint32_t* cv_aux;     // per CV: slot link | parent chain of its chain << 16
  const uint32_t* bnd_aux;    // per boundary row: the same
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
  typedef bc_fastmath<real> M;
  G = real(0);
  GE = real(0);
  real ica = real(0);
  real eca = cc.E_Ca_fixed;
  if constexpr ((CH & BC_CH_POOL) != 0) eca = cc.ca_rtzf * M::lg(bc_div<real>(cc.ca_Co, s[bc_slot_ca(CH)]));
  if constexpr ((CH & BC_CH_NA_ANY) != 0) {
    real p = s[bc_slot_na(CH)], q = s[bc_slot_na(CH) + 1];
    real g = cp[BC_CLS_G_NA] * (p * p * p) * q;
    G += g;
    GE += g * (cc.E_Na - V);
  }
  if constexpr ((CH & BC_CH_K4_ANY) != 0) {
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
                                               real ca_b, uint32_t flags) {
  typedef bc_fastmath<real> M;
  real ci = real(0), ci_gates = real(0);
  if constexpr ((CH & BC_CH_POOL) != 0) {
    real c = s[bc_slot_ca(CH)];
    real drive = ica * cc.ca_drive;
    drive = drive > real(0) ? drive : real(0);
    real f = drive + (cc.ca_rest - c) * cc.ca_tau_inv;
    ci = c + ca_b * f;
    s[bc_slot_ca(CH)] = ci;
    ci_gates = (flags & BC_STEP_INTEGRATION_ORDER) ? c : ci;
  }
  if (flags & BC_STEP_SKIP_GATES) return;
  if constexpr ((CH & BC_CH_NA) != 0) {
    real ap, bp, aq, bq;
    bc_na_rates<real, M>(V, cc.na_vsh, ap, bp, aq, bq);
    s[bc_slot_na(CH)] = bc_gate_ab_cf<real, M>(s[bc_slot_na(CH)], ap, bp, cc.na_phi, dt);
    s[bc_slot_na(CH) + 1] = bc_gate_ab_cf<real, M>(s[bc_slot_na(CH) + 1], aq, bq, cc.na_phi, dt);
  }
  if constexpr ((CH & BC_CH_KDR) != 0) {
    real ap, bp;
    bc_kdr_rates<real, M>(V, cc.kdr_vsh, ap, bp);
    s[bc_slot_kdr(CH)] = bc_gate_ab_cf<real, M>(s[bc_slot_kdr(CH)], ap, bp, cc.kdr_phi, dt);
  }
  if constexpr ((CH & BC_CH_NAHH) != 0) {
    real ap, bp, aq, bq;
    bc_nahh_rates<real, M>(V, cc.na_vsh, ap, bp, aq, bq);
    s[bc_slot_na(CH)] = bc_gate_ab_cf<real, M>(s[bc_slot_na(CH)], ap, bp, cc.na_phi, dt);
    s[bc_slot_na(CH) + 1] = bc_gate_ab_cf<real, M>(s[bc_slot_na(CH) + 1], aq, bq, cc.na_phi, dt);
  }
  if constexpr ((CH & BC_CH_KHH) != 0) {
    real ap, bp;
    bc_khh_rates<real, M>(V, cc.kdr_vsh, ap, bp);
    s[bc_slot_kdr(CH)] = bc_gate_ab_cf<real, M>(s[bc_slot_kdr(CH)], ap, bp, cc.kdr_phi, dt);
  }
  if constexpr ((CH & BC_CH_CAL) != 0) {
    real pi, pt, qi, qt;
    bc_cal_rates<real, M>(V, cc.cal_vsh, pi, pt, qi, qt);
    s[bc_slot_cal(CH)] = bc_gate_it_cf<real, M>(s[bc_slot_cal(CH)], pi, pt, cc.cal_phi_p, dt);
    s[bc_slot_cal(CH) + 1] = bc_gate_it_cf<real, M>(s[bc_slot_cal(CH) + 1], qi, qt, cc.cal_phi_q, dt);
  }
  if constexpr ((CH & BC_CH_CAT) != 0) {
    real pi, pt, qi, qt;
    bc_cat_rates<real, M>(V, cc.cat_vsh, pi, pt, qi, qt);
    s[bc_slot_cat(CH)] = bc_gate_it_cf<real, M>(s[bc_slot_cat(CH)], pi, pt, cc.cat_phi_p, dt);
    s[bc_slot_cat(CH) + 1] = bc_gate_it_cf<real, M>(s[bc_slot_cat(CH) + 1], qi, qt, cc.cat_phi_q, dt);
  }
  if constexpr ((CH & BC_CH_AHP) != 0) {
    real a = cc.ahp_alpha * M::pw(ci_gates, cc.ahp_n);
    s[bc_slot_ahp(CH)] = bc_gate_ab_cf<real, M>(s[bc_slot_ahp(CH)], a, cc.ahp_beta, cc.ahp_phi, dt);
  }
  if constexpr ((CH & BC_CH_HCN) != 0) {
    real pi, pt;
    bc_hcn_rates<real, M>(V, pi, pt);
    s[bc_slot_hcn(CH)] = bc_gate_it_cf<real, M>(s[bc_slot_hcn(CH)], pi, pt, cc.hcn_phi, dt);
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

constexpr int kWave = 64;  // gfx950 wavefront

// One DHS row in LDS: after phase A d = diagonal, r = right-hand side; the
// elimination overwrites d with 1 / pivot and r with the reduced rhs, the
// back substitution r with the solution.  lo / up = dt A[row, parent] and
// dt A[parent, row].  One 16-byte (FP32) load fetches a row.
template <typename real>
struct alignas(4 * sizeof(real)) HRow {
  real d, r, lo, up;
};
// Contribution of an eliminated chain to the row it hangs from.
template <typename real>
struct alignas(2 * sizeof(real)) KPair {
  real d, r;
};

// Chain record (16 bytes): x = j0 | j1 << 16 (slot range), y = top | n_kids
// << 16 (top = slot the chain hangs from, -1 for a root), z = kid0 | kid1
// << 16, w = offset of the chain's kid list (kids past the second).
struct ChainRec {
  int j0, j1, tp, nk, k0, k1, q0;
};
__device__ __forceinline__ ChainRec unpack_chain(int4 m) {
  ChainRec c;
  c.j0 = m.x & 0xFFFF;
  c.j1 = (int)((unsigned)m.x >> 16);
  c.tp = (m.y << 16) >> 16;
  c.nk = (int)((unsigned)m.y >> 16);
  c.k0 = m.z & 0xFFFF;
  c.k1 = (int)((unsigned)m.z >> 16);
  c.q0 = m.w;
  return c;
}

// Orders LDS accesses of the lanes of one wavefront across this point (LDS
// operations of a wavefront complete in issue order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Axial part of the delta-V right-hand side of a row, dt (A V_n) with the
// sign of the defect: lo_row (V_row - V_parent) minus up_c (V_c - V_row) for
// the chain child c = row - 1 (a chain's bottom row gets its kid chains'
// terms from their elimination instead).  Branch-free: a bottom row reads
// its own slot in place of row - 1 and discards it.
template <typename real>
__device__ __forceinline__ real axial_defect(int row, real V, const HRow<real>* R, const real* xv, uint32_t link) {
  const unsigned u = link & 0xFFFFu;
  const int par = (int)(u & 0x7FFFu);
  const bool bottom = (u & 0x8000u) != 0;
  const int below = bottom ? row : row - 1;
  const real own = R[row].lo * (V - xv[par]);
  const real kid = R[below].up * (xv[below] - V);
  return bottom ? own : own - kid;
}

// Rows of a chain are walked in groups of kGroup whose rows are loaded
// together, the next group's loads issued before the current group is
// computed, so LDS latency is paid once per group.  A chain of len rows is
// split into a head of len mod kGroup rows, done first, and len / kGroup
// full groups.  Only the head checks which of its rows belong to the chain
// (the others are computed on rows of other chains or padding, keep the
// running values and store to the dummy rows); the full groups run without
// any check.  Loads may reach up to 2 kGroup - 1 rows past either end of a
// chain, which the padding rows before row 0 and after the last row cover.
constexpr int kGroup = 4;

// Boundary rows per thread whose metadata is kept in registers (more are
// handled by a loop that reads it from global memory).
constexpr int kBndRegs = 2;

// Forward elimination in the delta-V form.  (sd, sr) is what the rows below
// subtract from the current row: the kids' contributions at the bottom,
// then the previous row's.
template <typename real>
__device__ __forceinline__ void eliminate_row(const HRow<real>& w, HRow<real>* dst, bool valid, real& sd, real& sr) {
  const real e = w.d - sd;
  const real r = w.r - sr;
  const real inv = bc_rcp<real>(e - w.lo);
  const real ui = w.up * inv;
  dst->d = inv;
  dst->r = r;
  sd = valid ? ui * e : sd;
  sr = valid ? ui * r : sr;
}

// Rows j .. j + kGroup - 1, all inside the chain.
template <typename real>
__device__ __forceinline__ void eliminate_full(const HRow<real>* g, HRow<real>* Rj, real& sd, real& sr) {
#pragma unroll
  for (int i = 0; i < kGroup; ++i) eliminate_row<real>(g[i], Rj + i, true, sd, sr);
}

// The chain's first rem < kGroup rows (j0 .. j0 + rem - 1).  The wavefront
// computes as many head rows as the largest rem among its active lanes (a
// uniform count, so a lane alone on the critical path computes only its own
// head rows).
template <typename real>
__device__ __forceinline__ void eliminate_head(const HRow<real>* g, int rem, HRow<real>* Rj, HRow<real>* Rd, real& sd,
                                               real& sr) {
  static_assert(kGroup == 4, "the head holds at most three rows");
  const bool any1 = __ballot(rem >= 1) != 0;
  const bool any2 = __ballot(rem >= 2) != 0;
  const bool any3 = __ballot(rem >= 3) != 0;
  if (any1) eliminate_row<real>(g[0], (rem > 0 ? Rj : Rd) + 0, rem > 0, sd, sr);
  if (any2) eliminate_row<real>(g[1], (rem > 1 ? Rj : Rd) + 1, rem > 1, sd, sr);
  if (any3) eliminate_row<real>(g[2], (rem > 2 ? Rj : Rd) + 2, rem > 2, sd, sr);
}

// Elimination through chain composites.  Eliminating row j maps what the
// rows below subtract, (sd, sr), to what row j passes up:
//   sd' = up (d - sd) / (d - lo - sd),   sr' = up (r - sr) / (d - lo - sd),
// a linear-fractional map.  In homogeneous coordinates (x, y, z), sd = x/z,
// sr = y/z, it is linear:
//   x' = -up x + up d z,   y' = -up y + up r z,   z' = -x + (d - lo) z,
// so a chain's rows compose into one map (x, y, z) = M (sd, sr, 1) with
//   x = a sd + b,   y = c sd + f sr + g,   z = h sd + k.
// With up, lo < 0 and d >= 0 the entries keep the signs a, k > 0, b, h <= 0
// row after row, so every product term of a, b, h, k and of x, z (sd <= 0)
// adds values of one sign: no cancellation.  M is rescaled to k = 1 after
// every group of rows (homogeneous coordinates).  The top edge's axial
// term, added to what the chain passes up, is folded into c and g.
//   pass 1 (every chain at once): the chain's composite;
//   pass 2 (dataflow over the chain tree, last-arriving kid climbs): one
//          composite application per chain, so the critical path costs
//          O(1) per chain instead of O(rows);
//   pass 3 (every chain at once): the rows' pivots and reduced right-hand
//          sides from the chain's input, then the chain's affine back
//          substitution.
template <typename real>
struct alignas(4 * sizeof(real)) Comp {
  real a, b, h, c, f, g, pad0, pad1;  // 32 bytes (FP32): two 16-byte loads
};

template <typename real>
__device__ __forceinline__ void comp_row(const HRow<real>& w, bool valid, real& a, real& b, real& h, real& k, real& c,
                                         real& f, real& g) {
  const real dl = w.d - w.lo;
  const real a2 = w.up * (w.d * h - a);
  const real b2 = w.up * (w.d * k - b);
  const real c2 = w.up * (w.r * h - c);
  const real g2 = w.up * (w.r * k - g);
  const real h2 = dl * h - a;
  const real k2 = dl * k - b;
  const real f2 = -w.up * f;
  a = valid ? a2 : a;
  b = valid ? b2 : b;
  c = valid ? c2 : c;
  g = valid ? g2 : g;
  h = valid ? h2 : h;
  k = valid ? k2 : k;
  f = valid ? f2 : f;
}

template <typename real>
__device__ __forceinline__ void chain_composite(int k, const int4* CR, const HRow<real>* R, const real* xv,
                                                Comp<real>* CM, KPair<real>* KK) {
  const ChainRec c = unpack_chain(CR[k]);
  real a = real(1), b = real(0), h = real(0), kk = real(1), cc = real(0), f = real(1), g = real(0);
  HRow<real> w[kGroup], nx[kGroup];
#pragma unroll
  for (int i = 0; i < kGroup; ++i) w[i] = R[c.j0 + i];
  const int tq = c.tp >= 0 ? c.tp : c.j1 - 1;
  const real td = R[c.j1 - 1].up * (xv[c.j1 - 1] - xv[tq]);
  for (int j = c.j0; j < c.j1; j += kGroup) {
    // the next group is loaded before this one is used (rows past the chain
    // land in other chains or padding and are discarded)
#pragma unroll
    for (int i = 0; i < kGroup; ++i) nx[i] = R[j + kGroup + i];
#pragma unroll
    for (int i = 0; i < kGroup; ++i) comp_row<real>(w[i], j + i < c.j1, a, b, h, kk, cc, f, g);
    const real ik = bc_rcp<real>(kk);
    a *= ik;
    b *= ik;
    h *= ik;
    cc *= ik;
    f *= ik;
    g *= ik;
    kk = real(1);
#pragma unroll
    for (int i = 0; i < kGroup; ++i) w[i] = nx[i];
  }
  Comp<real> m;
  m.a = a;
  m.b = b;
  m.h = h;
  m.c = cc + td * h;
  m.f = f;
  m.g = g + td;
  CM[k] = m;
  if (c.nk == 0) {  // a leaf chain's input is zero: its contribution is (b, g)
    KPair<real> o;
    o.d = m.b;
    o.r = m.g;
    KK[k] = o;
  }
}

// Sum of the kid chains' contributions (fixed order: first kid, second kid,
// then the rest of the kid list), the input of chain c's bottom row.
template <typename real>
__device__ __forceinline__ KPair<real> kid_sum(const ChainRec& c, const int16_t* m_kid, const KPair<real>* KK) {
  // both first kids are loaded unconditionally (an absent kid's index is 0
  // and its value is dropped), so the two loads share one round trip
  const KPair<real> t0 = KK[c.k0];
  const KPair<real> t1 = KK[c.k1];
  KPair<real> s;
  s.d = c.nk > 1 ? t0.d + t1.d : (c.nk > 0 ? t0.d : real(0));
  s.r = c.nk > 1 ? t0.r + t1.r : (c.nk > 0 ? t0.r : real(0));
  for (int q = 2; q < c.nk; ++q) {
    const KPair<real> t = KK[m_kid[c.q0 + q]];
    s.d += t.d;
    s.r += t.r;
  }
  return s;
}

template <typename real>
__device__ __forceinline__ KPair<real> apply_composite(const Comp<real>& m, const KPair<real>& s) {
  const real iz = bc_rcp<real>(m.h * s.d + real(1));
  KPair<real> o;
  o.d = (m.a * s.d + m.b) * iz;
  o.r = (m.c * s.d + m.f * s.r + m.g) * iz;
  return o;
}

// Pass 3 of the elimination: chain k's rows from its kids' contributions
// (R[j].d becomes 1 / pivot, R[j].r the reduced right-hand side).
template <typename real>
__device__ __forceinline__ void eliminate_rows(int k, const int4* CR, int dummy, const int16_t* m_kid, HRow<real>* R,
                                               const KPair<real>* KK) {
  const ChainRec c = unpack_chain(CR[k]);
  const int rem = (c.j1 - c.j0) & (kGroup - 1);
  int j = c.j0 + rem;
  HRow<real> gh[kGroup - 1], ga[kGroup], gb[kGroup];
#pragma unroll
  for (int i = 0; i < kGroup - 1; ++i) gh[i] = R[c.j0 + i];
#pragma unroll
  for (int i = 0; i < kGroup; ++i) ga[i] = R[j + i];
  const KPair<real> s = kid_sum<real>(c, m_kid, KK);
  real sd = s.d, sr = s.r;
  eliminate_head<real>(gh, rem, R + c.j0, R + dummy, sd, sr);
  while (j < c.j1) {
#pragma unroll
    for (int i = 0; i < kGroup; ++i) gb[i] = R[j + kGroup + i];
    eliminate_full<real>(ga, R + j, sd, sr);
    j += kGroup;
    if (j >= c.j1) break;
#pragma unroll
    for (int i = 0; i < kGroup; ++i) ga[i] = R[j + kGroup + i];
    eliminate_full<real>(gb, R + j, sd, sr);
    j += kGroup;
  }
}

// Back substitution in affine form.  Inside chain k the solution of row j
// is x_j = (r_j - lo_j x_{j+1}) / d_j, so x_j = alpha_j + beta_j X where X is
// the solution of the row the chain hangs from (zero for a root):
//   alpha_j = (r_j - lo_j alpha_{j+1}) / d_j,  beta_j = -lo_j beta_{j+1} / d_j,
// starting from alpha = 0, beta = 1 above the top (beta = 0 for a root).
// Every chain computes its (alpha, beta) at once, independently of the
// others; the chains' X then follow from pointer jumping over the chain
// tree.  R[j].d / R[j].r are overwritten with beta_j / alpha_j, and the
// bottom row's pair is returned through KK[k] = (alpha, beta).  The walk is
// top to bottom: a checked head of len mod kGroup rows at the top, then
// full groups.
template <typename real>
__device__ __forceinline__ void affine_row(const HRow<real>& w, HRow<real>* dst, bool valid, real& al, real& be) {
  const real a2 = (w.r - w.lo * al) * w.d;
  const real b2 = -w.lo * be * w.d;
  dst->d = b2;
  dst->r = a2;
  al = valid ? a2 : al;
  be = valid ? b2 : be;
}

template <typename real>
__device__ __forceinline__ void chain_affine(int k, int dummy, const int4* CR, HRow<real>* R) {
  const ChainRec c = unpack_chain(CR[k]);
  real al = real(0), be = c.tp >= 0 ? real(1) : real(0);
  const int rem = (c.j1 - c.j0) & (kGroup - 1);
  int j = c.j1 - 1 - rem;  // top row of the first full group
  HRow<real> gh[kGroup - 1], ga[kGroup], gb[kGroup];
#pragma unroll
  for (int i = 0; i < kGroup - 1; ++i) gh[i] = R[c.j1 - 1 - i];
#pragma unroll
  for (int i = 0; i < kGroup; ++i) ga[i] = R[j - i];
  HRow<real>* const Rt = R + (c.j1 - 1);
  HRow<real>* const Rd = R + dummy + (kGroup - 1);
#pragma unroll
  for (int i = 0; i < kGroup - 1; ++i) affine_row<real>(gh[i], (i < rem ? Rt : Rd) - i, i < rem, al, be);
  while (j >= c.j0) {
#pragma unroll
    for (int i = 0; i < kGroup; ++i) gb[i] = R[j - kGroup - i];
#pragma unroll
    for (int i = 0; i < kGroup; ++i) affine_row<real>(ga[i], R + (j - i), true, al, be);
    j -= kGroup;
    if (j < c.j0) break;
#pragma unroll
    for (int i = 0; i < kGroup; ++i) ga[i] = R[j - kGroup - i];
#pragma unroll
    for (int i = 0; i < kGroup; ++i) affine_row<real>(gb[i], R + (j - i), true, al, be);
    j -= kGroup;
  }
}

// Solution of one row after the affine pass and the pointer jumping.
// meta's high half is the chain the row's chain hangs from (0xFFFF: none).
template <typename real>
__device__ __forceinline__ real row_solution(int row, const HRow<real>* R, uint32_t meta, const KPair<real>* Xb) {
  const uint32_t p = meta >> 16;
  const real X = p != 0xFFFFu ? Xb[p].d : real(0);
  const HRow<real> w = R[row];
  return w.r + w.d * X;
}

// Rows after the pack's rows: the first kGroup are the dummy rows that head
// rows outside their chain write their (unused) results to (never read, so
// all threads share them); with the rest they pad the row array so a chain
// walk may load up to 2 kGroup - 1 rows past its last row without clamping
// (loads past a chain's end are discarded).
constexpr int kDummySlots = 8;
// Padding rows before row 0: a downward chain walk may load up to
// 2 kGroup - 1 rows below its first row.
constexpr int kPadBefore = 8;

// LDS carve-up (bytes), shared with the launch code: rows and chains are
// rounded up to multiples of 4 so every array stays naturally aligned.  The
// chain composites of the elimination and the pointer-jumping buffers of
// the back substitution (kk2, ptr, ptr2) are live in different passes and
// share one region.
struct LdsPlan {
  size_t rows, recs, kk, cmp, kk2, ptr, ptr2, xv, cls, kid, work, lv, dfw, total;
};
template <typename real>
__host__ __device__ inline LdsPlan lds_plan(int max_rows, int max_chains, int max_cls, int max_levels) {
  const size_t mr = (size_t)((max_rows + 3) & ~3), mc = (size_t)((max_chains + 3) & ~3);
  LdsPlan p;
  p.rows = kPadBefore * sizeof(HRow<real>);  // row 0; kPadBefore padding rows precede it
  p.recs = p.rows + (mr + kDummySlots) * sizeof(HRow<real>);
  p.kk = p.recs + mc * sizeof(int4);
  p.cmp = p.kk + mc * sizeof(KPair<real>);
  p.kk2 = p.cmp;
  p.ptr = p.kk2 + mc * sizeof(KPair<real>);
  p.ptr2 = p.ptr + mc * sizeof(int16_t);
  p.xv = p.cmp + mc * sizeof(Comp<real>);
  p.cls = p.xv + mr * sizeof(real);
  p.kid = p.cls + (size_t)max_cls * BC_NCLS * sizeof(real);
  p.work = p.kid + mc * sizeof(int16_t);
  p.lv = p.work + mc * sizeof(int16_t);
  p.dfw = p.lv + (size_t)((max_levels + 4) & ~3) * sizeof(int16_t);
  p.total = p.dfw + mc * sizeof(uint32_t);
  return p;
}

constexpr int kSynPar = BC_SYN_NPAR;  // parameters per synapse (bc_synapse_desc.par)

// Drive of every synapse at every step of a launch, before the launch:
// drive[k * syn_n + s] = D of synapse s at global step step0 + k, the sum
// over its incoming contacts of w_c spike(pre_c, step - delay_c), then its
// external drive.  Launches are no longer than the shortest delay, so every
// spike read here was written by an earlier launch.  One wavefront per
// synapse and block of kDriveSteps steps (blockIdx.y), whose history loads
// are all issued before they are summed: lane l takes contacts l, l + 64,
// ... in order and a fixed butterfly sums the lanes, so the result does not
// depend on scheduling.
constexpr int kDriveSteps = 8;

template <typename real>
__global__ void __launch_bounds__(256) synapse_drive(StaggerArgs<real> a, real* drive, int32_t step0,
                                                     int32_t n_steps, int32_t k_base) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (s >= a.syn_n) return;  // whole wavefronts
  const int c0 = a.spk != nullptr ? a.con_off[s] : 0;
  const int c1 = a.spk != nullptr ? a.con_off[s + 1] : 0;
  const int e = a.ext_slot != nullptr ? a.ext_slot[s] : -1;
  {
    const int k0 = k_base + (int)blockIdx.y * kDriveSteps;
    real dr[kDriveSteps];
#pragma unroll
    for (int q = 0; q < kDriveSteps; ++q) dr[q] = real(0);
    for (int c = c0 + lane; c < c1; c += 64) {
      const int pre = a.con_pre[c];
      const int dl = a.con_delay[c];
      const real w = a.con_w[c];
      uint8_t hit[kDriveSteps];
#pragma unroll
      for (int q = 0; q < kDriveSteps; ++q) {
        const int src = step0 + k0 + q - dl;
        hit[q] = (k0 + q < n_steps && src >= 0) ? a.spk[(int64_t)(src % a.spk_len) * a.n_cells + pre] : 0;
      }
#pragma unroll
      for (int q = 0; q < kDriveSteps; ++q)
        if (hit[q]) dr[q] += w;
    }
#pragma unroll
    for (int q = 0; q < kDriveSteps; ++q) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) dr[q] += __shfl_xor(dr[q], off);
      const int k = k0 + q;
      if (k < n_steps) {
        const int ke = step0 + k - a.ext_step0;  // arrivals first, then the external drive
        if (e >= 0 && ke >= 0 && ke < a.ext_steps) dr[q] += a.ext_drive[(int64_t)ke * a.n_ext + e];
        if (lane == 0) drive[(int64_t)k * a.syn_n + s] = dr[q];
      }
    }
  }
}

// A pack's synapses in LDS for the whole launch (after the pack's tree):
// per synapse its parameters (kSynPar reals), states A, B and the
// conductance k g its row sees this step; per synapse row its DHS row and
// the offset of its first synapse (plus one end offset).
struct SynPlan {
  size_t par, st, kg, row, off, total;
};
template <typename real>
__host__ __device__ inline SynPlan syn_plan(size_t base, int max_syn, int max_syn_rows) {
  const size_t ms = (size_t)(max_syn > 0 ? max_syn : 0), mr = (size_t)(max_syn_rows > 0 ? max_syn_rows : 0);
  SynPlan p;
  p.par = (base + 15) & ~(size_t)15;
  p.st = p.par + ms * kSynPar * sizeof(real);
  p.kg = p.st + ms * 2 * sizeof(real);
  p.row = p.kg + ms * sizeof(real);
  p.off = p.row + mr * sizeof(int32_t);
  p.total = p.off + (mr + 1) * sizeof(int32_t);
  return p;
}

// One synapse at one step: the drive D of the step (events), the
// conductance k g the solve sees, and the states' decay / integration.
template <typename real>
__device__ __forceinline__ void synapse_update(const real* q, real* st, real* kg_out, real drive, real dt) {
  const int kind = (int)q[2];
  real A = st[0];
  real B = st[1];
  if (kind == BC_SYN_EXP) {
    // events jump A and B; the solve sees g = B - A; both then decay
    A += drive * q[3];
    B += drive * q[4];
    *kg_out = q[0] * (B - A);
    A -= q[5] * A;
    B -= q[6] * B;
  } else {
    // the solve sees g = B at its step-start value; the events of the
    // step drive the ODEs, integrated from the step-start states
    *kg_out = q[0] * B;
    if (kind == BC_SYN_MARKOV1) {  // g' = aT D (1 - g) - beta g
      const real u = q[3] * drive;
      B = bc_expeuler<real>(B, u * (real(1) - B) - q[4] * B, -(u + q[4]), dt);
    } else {  // NMDA: g' = a1 x (1 - g) - b1 g, x' = a2T D (1 - x) - b2 x (A = x, B = g)
      const real u = q[5] * drive;
      const real gn = bc_expeuler<real>(B, q[3] * A * (real(1) - B) - q[4] * B, -(q[3] * A + q[4]), dt);
      A = bc_expeuler<real>(A, u * (real(1) - A) - q[6] * A, -(u + q[6]), dt);
      B = gn;
    }
  }
  st[0] = A;
  st[1] = B;
}

template <typename real, uint32_t CH, int MAXC, bool SYN, bool SWEEP = false>
__global__ void __launch_bounds__(256) stagger_kernel(StaggerArgs<real> a) {
  constexpr int NS = bc_nstates(CH) > 0 ? bc_nstates(CH) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const LdsPlan lp = lds_plan<real>(a.max_rows, a.max_chains, a.max_cls, a.max_levels);
  HRow<real>* R = reinterpret_cast<HRow<real>*>(smem_raw + lp.rows);
  int4* CR = reinterpret_cast<int4*>(smem_raw + lp.recs);
  KPair<real>* KK = reinterpret_cast<KPair<real>*>(smem_raw + lp.kk);
  real* xv = reinterpret_cast<real*>(smem_raw + lp.xv);  // V_n at every slot
  real* cls = reinterpret_cast<real*>(smem_raw + lp.cls);
  int16_t* m_kid = reinterpret_cast<int16_t*>(smem_raw + lp.kid);
  int16_t* m_work = reinterpret_cast<int16_t*>(smem_raw + lp.work);  // chain order of the per-chain passes
  int16_t* m_lv = reinterpret_cast<int16_t*>(smem_raw + lp.lv);
  KPair<real>* KK2 = reinterpret_cast<KPair<real>*>(smem_raw + lp.kk2);
  Comp<real>* CM = reinterpret_cast<Comp<real>*>(smem_raw + lp.cmp);  // chain composites (aliases KK2, ptr)
  // per chain: its parent chain (low half; 0xFFFF for a root)
  uint32_t* m_dfw = reinterpret_cast<uint32_t*>(smem_raw + lp.dfw);
  int16_t* m_ptr = reinterpret_cast<int16_t*>(smem_raw + lp.ptr);  // pointer jumping (two buffers)
  int16_t* m_ptr2 = reinterpret_cast<int16_t*>(smem_raw + lp.ptr2);

  const bc_pack P = a.packs[blockIdx.x];
  const int tid = threadIdx.x;
  const int NT = blockDim.x;
  for (int i = tid; i < P.n_cls * BC_NCLS; i += NT) cls[i] = a.cls_tab[(int64_t)P.cls0 * BC_NCLS + i];
  // the pack's rows and chain tables, copied with independent coalesced
  // loads (the host precomputed their LDS form).  The rows go in batches of
  // kCopy per thread, every load of a batch issued before its LDS stores, so
  // the batch pays one global-memory latency instead of kCopy (a lane past
  // the pack's rows reloads its first row and stores nothing).
  {
    constexpr int kCopy = 10;
    for (int i0 = tid; i0 < P.n_rows; i0 += kCopy * NT) {
      real l[kCopy], u[kCopy];
#pragma unroll
      for (int q = 0; q < kCopy; ++q) {
        const int i = i0 + q * NT;
        const int64_t g = (int64_t)P.row0 + (i < P.n_rows ? i : 0);
        l[q] = a.lo[g];
        u[q] = a.up[g];
      }
#pragma unroll
      for (int q = 0; q < kCopy; ++q) {
        const int i = i0 + q * NT;
        if (i < P.n_rows) {
          R[i].lo = l[q];
          R[i].up = u[q];
        }
      }
    }
  }
  {
    // the chain tables in batches of kCh entries per thread, loads first
    constexpr int kCh = 2;
    for (int i0 = tid; i0 < P.n_ch; i0 += kCh * NT) {
      int4 cr[kCh];
      int16_t cp[kCh], ck[kCh], cw[kCh];
#pragma unroll
      for (int q = 0; q < kCh; ++q) {
        const int i = i0 + q * NT;
        const int g = P.ch0 + (i < P.n_ch ? i : 0);
        cr[q] = a.chain_rec[g];
        cp[q] = a.chain_par[g];
        ck[q] = a.kid_pad[g];
        cw[q] = a.chain_work ? a.chain_work[g] : (int16_t)i;
      }
#pragma unroll
      for (int q = 0; q < kCh; ++q) {
        const int i = i0 + q * NT;
        if (i < P.n_ch) {
          CR[i] = cr[q];
          m_dfw[i] = (uint32_t)(uint16_t)cp[q];
          m_kid[i] = ck[q];
          m_work[i] = cw[q];
        }
      }
    }
  }
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
  // boundary-row voltages in batches of kBnd per thread, loads first
  {
    constexpr int kBnd = 2;
    for (int i0 = tid; i0 < P.n_bnd; i0 += kBnd * NT) {
      uint32_t bi[kBnd];
      real bv[kBnd];
#pragma unroll
      for (int q = 0; q < kBnd; ++q) {
        const int i = i0 + q * NT;
        const int g = P.bnd0 + (i < P.n_bnd ? i : 0);
        bi[q] = a.bnd_info[g];
        bv[q] = a.Vbnd[g];
      }
#pragma unroll
      for (int q = 0; q < kBnd; ++q)
        if (i0 + q * NT < P.n_bnd) xv[bi[q] & 0x7FFFFFFFu] = bv[q];
    }
  }
  // the pack's synapses (contiguous: synapse rows r0 .., synapses s0 ..)
  int syn_s0 = 0, syn_n = 0, syn_nr = 0;
  uint32_t syn_bad = 0;
  real* sy_par = nullptr;
  real* sy_st = nullptr;
  real* sy_kg = nullptr;
  int32_t* sy_row = nullptr;
  int32_t* sy_off = nullptr;
  if constexpr (SYN) {
    const SynPlan sp = syn_plan<real>(lp.total, a.max_syn, a.max_syn_rows);
    sy_par = reinterpret_cast<real*>(smem_raw + sp.par);
    sy_st = reinterpret_cast<real*>(smem_raw + sp.st);
    sy_kg = reinterpret_cast<real*>(smem_raw + sp.kg);
    sy_row = reinterpret_cast<int32_t*>(smem_raw + sp.row);
    sy_off = reinterpret_cast<int32_t*>(smem_raw + sp.off);
    const int r0 = a.syn_pack_off[blockIdx.x];
    syn_nr = a.syn_pack_off[blockIdx.x + 1] - r0;
    syn_s0 = a.syn_row_off[r0];
    syn_n = a.syn_row_off[r0 + syn_nr] - syn_s0;
    if (syn_n > a.max_syn || syn_nr > a.max_syn_rows) {  // the tables do not fit the LDS plan
      syn_bad = 4u;
      syn_n = syn_nr = 0;
    }
    for (int i = tid; i < syn_n * kSynPar; i += NT) sy_par[i] = a.syn_par[(int64_t)syn_s0 * kSynPar + i];
    for (int i = tid; i < 2 * syn_n; i += NT) sy_st[i] = a.syn_state[2 * (int64_t)syn_s0 + i];
    for (int i = tid; i <= syn_nr; i += NT) {
      if (i < syn_nr) sy_row[i] = a.syn_row[r0 + i];
      sy_off[i] = a.syn_row_off[r0 + i] - syn_s0;
    }
  }
  __syncthreads();

  // Static per-row metadata the steps need, kept in registers: for every CV
  // its slot link (low half: the parent slot, a root's top pointing at
  // itself, | 0x8000 at a chain's bottom slot, whose children are the kid
  // chains rather than slot - 1) and the chain its row's solution hangs from
  // (high half, 0xFFFF for a root chain); for the first kBndRegs boundary
  // rows of the thread the link (bl) and the row (bits 0-14), the clamp bit
  // (bit 15) and the chain its solution hangs from (high half) (bx).  The
  // steps then read each operand in one LDS round trip for all of the
  // thread's rows.  Links and slot chains come from global memory (further
  // boundary rows read them there every step).
  uint32_t aux[MAXC];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    aux[i] = 0;
    if (info[i] != 0xFFFFFFFFu) {
      aux[i] = a.cv_aux[(int64_t)P.cv0 + tid + i * NT];
    }
  }
  uint32_t bx[kBndRegs], bl[kBndRegs];
#pragma unroll
  for (int b = 0; b < kBndRegs; ++b) {
    const int i = tid + b * NT;
    bx[b] = 0;
    bl[b] = 0;
    if (i < P.n_bnd) {
      const uint32_t u = a.bnd_info[P.bnd0 + i];
      const int row = (int)(u & 0x7FFFFFFFu);
      const uint32_t x = a.bnd_aux[P.bnd0 + i];
      bx[b] = (uint32_t)row | ((u & kClampBit) ? 0x8000u : 0u) | (x & 0xFFFF0000u);
      bl[b] = x & 0xFFFFu;
    }
  }

  const int dummy = (a.max_rows + 3) & ~3;
  const real dt = a.dt;
  const real ca_b = a.ca_b;
  uint32_t bad = syn_bad;
  // drive of the thread's first synapse, loaded one step ahead so that its
  // latency hides behind the rest of the step
  real drv_next = real(0);
  if constexpr (SYN) {
    if (tid < syn_n && a.syn_drive != nullptr && a.n_steps > 0) drv_next = a.syn_drive[syn_s0 + tid];
  }
  for (int step = 0; step < a.n_steps; ++step) {
    const int gstep = a.step0 + step;
    if (a.rec) record_state<real, MAXC, NS>(a, P, tid, NT, step, vv, st, info);
    // ---- phase A: linearise at V_n.  Every operand is read before the
    // first row is written, so the reads of all rows share their latency.
    real dA[MAXC], dB[kBndRegs];
#pragma unroll
    for (int i = 0; i < MAXC; ++i)
      dA[i] = info[i] != 0xFFFFFFFFu ? axial_defect<real>((int)(info[i] & kRowMask), vv[i], R, xv, aux[i]) : real(0);
#pragma unroll
    for (int b = 0; b < kBndRegs; ++b) {
      const int row = (int)(bx[b] & 0x7FFFu);
      dB[b] = tid + b * NT < P.n_bnd ? axial_defect<real>(row, xv[row], R, xv, bl[b]) : real(0);
    }
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] != 0xFFFFFFFFu) {
        const int row = (int)(info[i] & kRowMask);
        const real* cp = cls + ((info[i] >> 16) & 0x3FFFu) * BC_NCLS;
        real G, GE;
        ica[i] = linearise<real, CH>(a.cc, cp, st[i], vv[i], G, GE);
        const real Icl = (info[i] & kClampBit) ? clamp_current<real>(a, P, row, gstep) : real(0);
        const real dtcm = cp[BC_CLS_DTCM];
        R[row].d = real(1) + dtcm * G;
        R[row].r = dtcm * (GE + Icl) + dA[i];
      }
    }
    // boundary (algebraic) rows
#pragma unroll
    for (int b = 0; b < kBndRegs; ++b) {
      if (tid + b * NT < P.n_bnd) {
        const int row = (int)(bx[b] & 0x7FFFu);
        R[row].d = real(0);
        R[row].r = ((bx[b] & 0x8000u) ? clamp_current<real>(a, P, row, gstep) : real(0)) + dB[b];
      }
    }
    for (int i = tid + kBndRegs * NT; i < P.n_bnd; i += NT) {
      uint32_t b = a.bnd_info[P.bnd0 + i];
      int row = (int)(b & 0x7FFFFFFFu);
      const real D = axial_defect<real>(row, xv[row], R, xv, a.bnd_aux[P.bnd0 + i] & 0xFFFFu);
      R[row].d = real(0);
      R[row].r = ((b & kClampBit) ? clamp_current<real>(a, P, row, gstep) : real(0)) + D;
    }
    if constexpr (SYN) {
      // the step's synapse events and state updates (independent of V: the
      // current reaches the rows below, after the barrier)
      if (tid < syn_n) {
        const real drv = drv_next;
        if (a.syn_drive != nullptr && step + 1 < a.n_steps)
          drv_next = a.syn_drive[(int64_t)(step + 1) * a.syn_n + syn_s0 + tid];
        synapse_update<real>(sy_par + tid * kSynPar, sy_st + 2 * tid, sy_kg + tid, drv, dt);
        if (a.syn_trace != nullptr) {
          real* o = a.syn_trace + (int64_t)(gstep - a.step0) * (2 * (int64_t)a.syn_n) + 2 * (int64_t)(syn_s0 + tid);
          o[0] = sy_st[2 * tid];
          o[1] = sy_st[2 * tid + 1];
        }
      }
      for (int i = tid + NT; i < syn_n; i += NT) {
        const real drv = a.syn_drive != nullptr ? a.syn_drive[(int64_t)step * a.syn_n + syn_s0 + i] : real(0);
        synapse_update<real>(sy_par + i * kSynPar, sy_st + 2 * i, sy_kg + i, drv, dt);
        if (a.syn_trace != nullptr) {
          real* o = a.syn_trace + (int64_t)(gstep - a.step0) * (2 * (int64_t)a.syn_n) + 2 * (int64_t)(syn_s0 + i);
          o[0] = sy_st[2 * i];
          o[1] = sy_st[2 * i + 1];
        }
      }
    }
    __syncthreads();
    if constexpr (SYN) {
      // each synapse row adds its synapses' conductance k g and current
      // k g (e - V_n) to its DHS row, synapses in table order
      for (int j = tid; j < syn_nr; j += NT) {
        const int row = sy_row[j];
        real gs = real(0), gse = real(0);
        for (int i = sy_off[j]; i < sy_off[j + 1]; ++i) {
          const real kg = sy_kg[i];
          gs += kg;
          gse += kg * sy_par[i * kSynPar + 1];
        }
        R[row].d += gs;
        R[row].r += gse - gs * xv[row];
      }
      __syncthreads();
    }

    // ---- phase B: Hines elimination through chain composites.  Diagonal
    // rebuilt from the zero row sums of the axial operator: d_j = e_j - lo_j
    // with e_j = diag_j - sum over children c of up_c e_c / d_c >= 0, so no
    // step subtracts nearly equal numbers.
    // Pass 1: every chain's composite map (and a leaf chain's contribution).
    // Rounds of NT positions of the work order (longest chains first within
    // the leaf and the non-leaf part), every other round in reverse: the
    // threads whose first chain is long take the short ones of a second
    // round.
    for (int r = 0, base = 0; base < P.n_ch; ++r, base += NT) {
      const int i = base + ((r & 1) ? NT - 1 - tid : tid);
      if (i < P.n_ch) chain_composite<real>(m_work[i], CR, R, xv, CM, KK);
    }
    __syncthreads();
    // Pass 2 in wavefront 0: the chain levels in order (chains are sorted by
    // level, level 0 being the leaves), each chain applying its composite to
    // the sum of its kids' contributions; a wavefront's LDS operations
    // complete in order, so a level only needs the wavefront's own sync
    // before the next one reads it.  Chain values do not depend on the
    // schedule: results are deterministic.  Meanwhile the other wavefronts
    // run pass 3 and the affine back substitution of the leaf chains (whose
    // input is zero); they write rows only, never what pass 2 reads.
    const int n_leaf = m_lv[1] - m_lv[0];
    const int lw = NT > kWave ? kWave : 0;  // first thread of the leaf crew
    if (tid < kWave) {
      // the sweep is the pack's critical path: its wavefront gets issue
      // priority over the other wavefronts of its SIMD while it runs
      __builtin_amdgcn_s_setprio(2);
      // the top level holds root chains only (a parent sits above its
      // kids), whose contributions no chain reads: the sweep ends below it
      for (int lev = 1; lev < P.n_lv - 1; ++lev) {
        const int c1 = m_lv[lev + 1];
        for (int k = m_lv[lev] + tid; k < c1; k += kWave) {
          const int4 rec = CR[k];
          const Comp<real> m = CM[k];
          KK[k] = apply_composite<real>(m, kid_sum<real>(unpack_chain(rec), m_kid, KK));
        }
        wave_sync();
      }
      __builtin_amdgcn_s_setprio(0);
    }
    if (lw > 0 && tid >= lw) {
      for (int i = tid - lw; i < n_leaf; i += NT - lw) {
        const int k = m_work[i];
        eliminate_rows<real>(k, CR, dummy, m_kid, R, KK);
        chain_affine<real>(k, dummy, CR, R);
      }
    }
    __syncthreads();
    // Pass 3 and back substitution of the other chains: every chain's rows
    // from its input, then its affine map x_j = alpha_j + beta_j X into the
    // rows (X = the solution of the row the chain hangs from).
    for (int i = (lw > 0 ? n_leaf : 0) + tid; i < P.n_ch; i += NT) {
      const int k = m_work[i];
      eliminate_rows<real>(k, CR, dummy, m_kid, R, KK);
      chain_affine<real>(k, dummy, CR, R);
    }
    __syncthreads();
    // the chains' bottom solutions X_k = alpha_k + beta_k X_parent(k), from
    // each chain's bottom-row (alpha, beta) and its parent chain
    const KPair<real>* Xb;
    if constexpr (SWEEP) {
      // sequential variant: wavefront 0 walks the chain levels top-down
      // (a parent chain's level is above its kids'), one chain per lane;
      // level 0 (the leaf chains) is skipped: only parents are read
      if (tid < kWave) {
        for (int lev = P.n_lv - 1; lev >= 1; --lev) {
          const int c1 = m_lv[lev + 1];
          for (int k = m_lv[lev] + tid; k < c1; k += kWave) {
            const HRow<real> w = R[CR[k].x & 0xFFFF];
            const uint32_t q = m_dfw[k] & 0xFFFFu;
            KPair<real> v;
            v.d = q == 0xFFFFu ? w.r : w.r + w.d * KK[q].d;
            v.r = real(0);
            KK[k] = v;
          }
          wave_sync();
        }
      }
      __syncthreads();
      Xb = KK;
    } else {
    // recursive doubling: pointer jumping over the chain tree.  Only parent
    // chains' solutions are read (row_solution), and a parent is never a
    // leaf: the leaf chains (level 0, the first n_leaf) take no part
    for (int k = n_leaf + tid; k < P.n_ch; k += NT) {
      const HRow<real> w = R[CR[k].x & 0xFFFF];
      KPair<real> v;
      v.d = w.r;
      v.r = w.d;
      KK2[k] = v;
      m_ptr[k] = (int16_t)(m_dfw[k] & 0xFFFFu);
    }
    __syncthreads();
    // solution of row j of chain k: x_j = alpha_j + beta_j X, X = the solution
    // at the bottom of k's parent chain (Xb[parent].d after the jumping)
    {
      // (A, B, ptr): x_bottom(k) = A + B x_bottom(ptr); ptr < 0 once resolved
      KPair<real>* A0 = KK2;
      KPair<real>* A1 = KK;
      int16_t* p0 = m_ptr;
      int16_t* p1 = m_ptr2;
      // a non-leaf chain's depth below its root chain is at most n_lv - 2
      // (levels are heights, a non-leaf chain's is at least 1), so
      // ceil(log2(n_lv - 1)) rounds resolve every pointer; one still open
      // afterwards means an inconsistent layout and raises the error flag
      // instead of hanging
      int rounds = 0;
      while ((1 << rounds) < P.n_lv - 1) ++rounds;
      for (int round = 0; round < rounds; ++round) {
        int open = 0;
        for (int k = n_leaf + tid; k < P.n_ch; k += NT) {
          const int q = p0[k];
          KPair<real> v = A0[k];
          int qn = q;
          if (q >= 0) {
            const KPair<real> u = A0[q];
            v.d = v.d + v.r * u.d;
            v.r = v.r * u.r;
            qn = p0[q];
            open |= qn >= 0;
          }
          A1[k] = v;
          p1[k] = (int16_t)qn;
        }
        __syncthreads();
        KPair<real>* ta = A0;
        A0 = A1;
        A1 = ta;
        int16_t* tq = p0;
        p0 = p1;
        p1 = tq;
        if (round == rounds - 1 && open) bad |= 2u;
      }
      Xb = A0;
    }
    }


    // ---- phase C: V += dV; Ca pool and gates at V_{n+1}.  All row
    // solutions are read before the new voltages are written.
#pragma unroll
    for (int i = 0; i < MAXC; ++i)
      dA[i] = info[i] != 0xFFFFFFFFu ? row_solution<real>((int)(info[i] & kRowMask), R, aux[i], Xb) : real(0);
#pragma unroll
    for (int b = 0; b < kBndRegs; ++b) {
      const int row = (int)(bx[b] & 0x7FFFu);
      dB[b] = tid + b * NT < P.n_bnd ? xv[row] + row_solution<real>(row, R, bx[b], Xb) : real(0);
    }
    for (int i = tid + kBndRegs * NT; i < P.n_bnd; i += NT) {
      const int row = (int)(a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu);
      xv[row] += row_solution<real>(row, R, a.bnd_aux[P.bnd0 + i] & 0xFFFF0000u, Xb);
    }
#pragma unroll
    for (int b = 0; b < kBndRegs; ++b)
      if (tid + b * NT < P.n_bnd) xv[bx[b] & 0x7FFFu] = dB[b];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] != 0xFFFFFFFFu) {
        const real Vn = vv[i] + dA[i];
        xv[info[i] & kRowMask] = Vn;
        advance_states<real, CH>(a.cc, st[i], Vn, ica[i], dt, ca_b, a.flags);
        if constexpr (SYN) {
          if (a.spk != nullptr && vv[i] < a.v_th && Vn >= a.v_th)
            a.spk[(int64_t)(gstep % a.spk_len) * a.n_cells + a.cv_cell[(int64_t)P.cv0 + tid + i * NT]] = 1;
        }
        vv[i] = Vn;
        bad |= (Vn == Vn && fabs(Vn) < real(1e30)) ? 0u : 1u;
        if (pcol[i] >= 0) a.trace[(int64_t)step * a.trace_ld + pcol[i]] = Vn;
      }
    }
    __syncthreads();  // phase A reads the voltages of other threads' rows
  }

  if (a.rec) record_state<real, MAXC, NS>(a, P, tid, NT, a.n_steps, vv, st, info);
  // with the gates skipped only V and the Ca pool changed: the gates' HBM
  // copies already hold their values
  const bool gates_out = !(a.flags & BC_STEP_SKIP_GATES);
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    if (info[i] != 0xFFFFFFFFu) {
      int64_t g = (int64_t)P.cv0 + tid + i * NT;
      a.V[g] = vv[i];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const bool pool = (CH & BC_CH_POOL) != 0 && s == bc_slot_ca(CH);
        if (s < bc_nstates(CH) && (gates_out || pool)) a.state[(int64_t)s * a.n_cv_total + g] = st[i][s];
      }
    }
  }
  for (int i = tid; i < P.n_bnd; i += NT) a.Vbnd[P.bnd0 + i] = xv[a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu];
  if constexpr (SYN) {
    for (int i = tid; i < 2 * syn_n; i += NT) a.syn_state[2 * (int64_t)syn_s0 + i] = sy_st[i];
  }
  if (bad) atomicOr(a.status, bad);
}

// ------------------------------------------------------------------ dispatch
constexpr uint32_t kSetPassive = BC_CH_IL;
constexpr uint32_t kSetHH = BC_CH_NA | BC_CH_KDR | BC_CH_IL;
constexpr uint32_t kSetHHCaL = BC_CH_NA | BC_CH_KDR | BC_CH_IL | BC_CH_CAL | BC_CH_POOL;
constexpr uint32_t kSetAll = BC_CH_NA | BC_CH_KDR | BC_CH_IL | BC_CH_KL | BC_CH_CAL | BC_CH_CAT | BC_CH_AHP |
                             BC_CH_HCN | BC_CH_POOL;
constexpr uint32_t kSetHH1952 = BC_CH_NAHH | BC_CH_KHH | BC_CH_IL;

typedef void (*kernel_f32)(StaggerArgs<float>);
typedef void (*kernel_f64)(StaggerArgs<double>);

template <typename real, uint32_t CH>
void* pick_maxc(int maxc, bool syn, bool sweep) {
  if (sweep) {  // the sequential back substitution: FP32 kernels without synapses
    if constexpr (sizeof(real) != 4) return nullptr;
    if (syn) return nullptr;
    switch (maxc) {
      case 1: return (void*)&stagger_kernel<real, CH, 1, false, true>;
      case 2: return (void*)&stagger_kernel<real, CH, 2, false, true>;
      case 4: return (void*)&stagger_kernel<real, CH, 4, false, true>;
      case 8: return (void*)&stagger_kernel<real, CH, 8, false, true>;
      case 16: return (void*)&stagger_kernel<real, CH, 16, false, true>;
      default: return nullptr;
    }
  }
  if (syn) {  // network runs: up to 8 CVs per thread
    switch (maxc) {
      case 1: return (void*)&stagger_kernel<real, CH, 1, true>;
      case 2: return (void*)&stagger_kernel<real, CH, 2, true>;
      case 4: return (void*)&stagger_kernel<real, CH, 4, true>;
      case 8: return (void*)&stagger_kernel<real, CH, 8, true>;
      default: return nullptr;
    }
  }
  switch (maxc) {
    case 1: return (void*)&stagger_kernel<real, CH, 1, false>;
    case 2: return (void*)&stagger_kernel<real, CH, 2, false>;
    case 4: return (void*)&stagger_kernel<real, CH, 4, false>;
    case 8: return (void*)&stagger_kernel<real, CH, 8, false>;
    case 16: return (void*)&stagger_kernel<real, CH, 16, false>;
    default: return nullptr;
  }
}

template <typename real>
void* pick_kernel(uint32_t ch, int maxc, bool syn = false, bool sweep = false) {
  switch (ch) {
    case kSetPassive: return pick_maxc<real, kSetPassive>(maxc, syn, sweep);
    case kSetHH: return pick_maxc<real, kSetHH>(maxc, syn, sweep);
    case kSetHHCaL: return pick_maxc<real, kSetHHCaL>(maxc, syn, sweep);
    case kSetAll: return pick_maxc<real, kSetAll>(maxc, syn, sweep);
    case kSetHH1952: return pick_maxc<real, kSetHH1952>(maxc, syn, sweep);
    default: return nullptr;
  }
}

// Clears the spike-history slots of steps step0 .. step0 + n_steps - 1.
__global__ void spike_clear(uint8_t* spk, int spk_len, int n_cells, int step0, int n_steps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_steps * n_cells) return;
  const int slot = (int)((step0 + i / n_cells) % spk_len);
  spk[(int64_t)slot * n_cells + i % n_cells] = 0;
}

// Threshold crossings on the launch's trace rows: one thread per trace
// column, rows read in order (coalesced across the columns of a row), in
// batches of kScanRows independent loads so that a batch pays one memory
// latency.
constexpr int kScanRows = 16;
template <typename real>
__global__ void __launch_bounds__(256) spike_scan(const real* __restrict__ trace, int64_t ld, int32_t n_steps,
                                                  int32_t step0, real v_th, real* __restrict__ v_last,
                                                  int32_t* __restrict__ count, int32_t* __restrict__ steps,
                                                  int32_t cap) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ld) return;
  real v = v_last[col];
  int32_t c = count[col];
  int32_t* out = steps + col * cap;
  for (int32_t n0 = 0; n0 < n_steps; n0 += kScanRows) {
    real x[kScanRows];
#pragma unroll
    for (int j = 0; j < kScanRows; ++j) x[j] = n0 + j < n_steps ? trace[(int64_t)(n0 + j) * ld + col] : real(0);
#pragma unroll
    for (int j = 0; j < kScanRows; ++j) {
      if (n0 + j < n_steps) {
        if (v < v_th && x[j] >= v_th) {
          if (c < cap) out[c] = step0 + n0 + j;
          ++c;
        }
        v = x[j];
      }
    }
  }
  v_last[col] = v;
  count[col] = c;
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
  // voltage-only = gates skipped and the pool frozen (ca_b = 0 below)
  a->flags = d->flags | ((d->flags & BC_STEP_VOLTAGE_ONLY) ? BC_STEP_SKIP_GATES : 0u);
  a->max_rows = d->max_rows;
  a->max_chains = d->max_chains;
  a->max_cls = d->max_cls;
  a->max_levels = d->max_levels;
  if (d->chan_const == nullptr) return bc::fail(BC_INVALID_ARG, "chan_const is NULL");
  a->cc = *static_cast<const bc_chan_const<real>*>(d->chan_const);
  a->ca_b = (d->flags & BC_STEP_VOLTAGE_ONLY) ? real(0) : a->dt * bc_exprel<real>(-a->dt * a->cc.ca_tau_inv);
  const bc_synapse_desc* y = d->syn;
  a->syn_pack_off = y ? y->pack_off : nullptr;
  a->syn_row = y ? y->row : nullptr;
  a->syn_row_off = y ? y->row_off : nullptr;
  a->syn_par = y ? static_cast<const real*>(y->par) : nullptr;
  a->syn_state = y ? static_cast<real*>(y->state) : nullptr;
  a->con_off = y ? y->con_off : nullptr;
  a->con_pre = y ? y->con_pre : nullptr;
  a->con_delay = y ? y->con_delay : nullptr;
  a->con_w = y ? static_cast<const real*>(y->con_w) : nullptr;
  a->cv_cell = y ? y->cv_cell : nullptr;
  a->spk = y ? y->spk : nullptr;
  a->spk_len = y ? y->spk_len : 1;
  a->n_cells = y ? y->n_cells : 0;
  a->v_th = y ? static_cast<real>(y->v_th) : real(0);
  a->ext_slot = y ? y->ext_slot : nullptr;
  a->ext_drive = y ? static_cast<const real*>(y->ext_drive) : nullptr;
  a->n_ext = y ? y->n_ext : 0;
  a->ext_step0 = y ? y->ext_step0 : 0;
  a->ext_steps = y ? y->ext_steps : 0;
  a->syn_trace = y ? static_cast<real*>(y->syn_trace) : nullptr;
  a->syn_n = y ? y->n_syn : 0;
  // no ring and no external drive: no synapse ever receives drive
  a->syn_drive = (y && (y->spk || y->ext_slot)) ? static_cast<const real*>(y->drive) : nullptr;
  a->max_syn = y ? y->max_syn : 0;
  a->max_syn_rows = y ? y->max_syn_rows : 0;
  a->chain_rec = reinterpret_cast<const int4*>(d->chain_rec);
  a->chain_par = d->chain_par;
  a->kid_pad = d->kid_pad;
  a->slot_link = d->slot_link;
  a->slot_chain = d->slot_chain;
  a->chain_work = d->chain_work;
  a->cv_aux = d->cv_aux;
  a->bnd_aux = d->bnd_aux;
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
  if ((d->flags & ~(BC_STEP_SKIP_GATES | BC_STEP_INTEGRATION_ORDER | BC_STEP_VOLTAGE_ONLY | BC_STEP_BACKSUB_SWEEP)) != 0)
    return bc::fail(BC_INVALID_ARG, "unknown flags 0x%x", d->flags);
  if (d->n_packs > 0 && (d->packs == nullptr || d->cv_info == nullptr || d->V == nullptr || d->Vbnd == nullptr ||
                         d->status == nullptr || d->chain_rec == nullptr || d->chain_par == nullptr ||
                         d->kid_pad == nullptr || d->slot_link == nullptr || d->slot_chain == nullptr ||
                         d->cv_aux == nullptr || d->bnd_aux == nullptr))
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  if (const bc_spike_desc* k = d->spk) {
    if (k->cap < 0) return bc::fail(BC_INVALID_ARG, "spike desc: cap < 0");
    if (d->trace_ld <= 0) return bc::fail(BC_INVALID_ARG, "spike desc: detectors are trace columns (trace_ld > 0)");
    if (k->v_last == nullptr || k->count == nullptr || (k->cap > 0 && k->steps == nullptr))
      return bc::fail(BC_INVALID_ARG, "spike desc: required pointer is NULL");
  }
  if (const bc_synapse_desc* y = d->syn) {
    if (y->pack_off == nullptr || y->row == nullptr || y->row_off == nullptr || y->par == nullptr ||
        y->state == nullptr)
      return bc::fail(BC_INVALID_ARG, "synapse desc: required pointer is NULL");
    if (y->spk != nullptr) {
      if (y->n_cells <= 0 || y->min_delay < 1 || y->max_delay < y->min_delay)
        return bc::fail(BC_INVALID_ARG, "synapse desc: n_cells > 0 and 1 <= min_delay <= max_delay required");
      if (y->con_off == nullptr || y->con_pre == nullptr || y->con_delay == nullptr || y->con_w == nullptr ||
          y->cv_cell == nullptr)
        return bc::fail(BC_INVALID_ARG, "synapse desc: contact tables are NULL with a spike ring");
    }
    if (y->n_syn < 0) return bc::fail(BC_INVALID_ARG, "synapse desc: n_syn < 0");
    if (y->syn_trace != nullptr && y->n_syn <= 0)
      return bc::fail(BC_INVALID_ARG, "synapse desc: a synapse trace needs n_syn > 0");
    if (y->n_syn > 0 && (y->max_syn < 1 || y->max_syn > y->n_syn || y->max_syn_rows < 1 ||
                         y->max_syn_rows > y->max_syn))
      return bc::fail(BC_INVALID_ARG, "synapse desc: need 1 <= max_syn_rows <= max_syn <= n_syn (got %d, %d, %d)",
                      y->max_syn_rows, y->max_syn, y->n_syn);
    if (y->n_syn > 0 && (y->spk != nullptr || y->ext_slot != nullptr) && y->drive == nullptr)
      return bc::fail(BC_INVALID_ARG, "synapse desc: the drive workspace is NULL");
    if (y->ext_slot != nullptr && (y->n_ext <= 0 || y->ext_steps < 0 || y->ext_drive == nullptr))
      return bc::fail(BC_INVALID_ARG, "synapse desc: external drive needs n_ext > 0, ext_steps >= 0 and ext_drive");
  }
  return BC_OK;
}

template <typename real>
int launch(const bc_stagger_desc* d, hipStream_t s, int32_t step0, int32_t n_steps, void* trace) {
  const bool sweep = (d->flags & BC_STEP_BACKSUB_SWEEP) != 0;
  void* fn = pick_kernel<real>(d->channels, d->cv_per_thread, d->syn != nullptr, sweep);
  if (fn == nullptr)
    return bc::fail(BC_UNSUPPORTED, "no kernel for channels=0x%x cv_per_thread=%d%s%s", d->channels,
                    d->cv_per_thread, d->syn ? " with synapses" : "",
                    sweep ? " with the sequential back substitution (FP32 kernels without synapses)" : "");
  StaggerArgs<real> a;
  int rc = fill_args<real>(d, &a);
  if (rc != BC_OK) return rc;
  a.step0 = step0;
  a.n_steps = n_steps;
  a.trace = static_cast<real*>(trace);
  size_t lds = lds_plan<real>(d->max_rows, d->max_chains, d->max_cls, d->max_levels).total;
  if (d->syn != nullptr) lds = syn_plan<real>(lds, d->syn->max_syn, d->syn->max_syn_rows).total;
  if (lds > 160 * 1024) return bc::fail(BC_INVALID_LAYOUT, "pack needs %zu bytes of LDS (> 160 KiB)", lds);
  if (lds > 64 * 1024) BC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (d->n_packs == 0 || n_steps == 0) return BC_OK;
  if (d->spk != nullptr && trace == nullptr) return bc::fail(BC_INVALID_ARG, "spike output needs a trace buffer");
  if (const bc_synapse_desc* y = d->syn; y != nullptr && y->spk != nullptr) {
    if (n_steps > y->min_delay || y->spk_len < y->max_delay + n_steps || step0 < 0)
      return bc::fail(BC_INVALID_ARG,
                      "network launch of %d steps from step %d: need n_steps <= min_delay (%d), "
                      "spk_len (%d) >= max_delay (%d) + n_steps and step0 >= 0",
                      n_steps, step0, y->min_delay, y->spk_len, y->max_delay);
    const int64_t n = (int64_t)n_steps * y->n_cells;
    hipLaunchKernelGGL(spike_clear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y->spk, y->spk_len,
                       y->n_cells, step0, n_steps);
  }
  if (const bc_synapse_desc* y = d->syn; y != nullptr && y->n_syn > 0 && a.syn_drive != nullptr) {
    // the grid's y extent (step blocks) is at most 65535 per launch
    const int64_t n_blk = ((int64_t)n_steps + kDriveSteps - 1) / kDriveSteps;
    for (int64_t b0 = 0; b0 < n_blk; b0 += 65535) {
      const unsigned nb = (unsigned)(n_blk - b0 < 65535 ? n_blk - b0 : 65535);
      hipLaunchKernelGGL(synapse_drive<real>, dim3((unsigned)((y->n_syn + 3) / 4), nb), dim3(256), 0, s, a,
                         static_cast<real*>(y->drive), step0, n_steps, (int32_t)(b0 * kDriveSteps));
    }
  }
  void* args[] = {&a};
  BC_HIP(hipLaunchKernel(fn, dim3(d->n_packs), dim3(d->threads), args, lds, s));
  if (const bc_spike_desc* k = d->spk) {
    hipLaunchKernelGGL(spike_scan<real>, dim3((unsigned)((d->trace_ld + 255) / 256)), dim3(256), 0, s,
                       static_cast<const real*>(trace), d->trace_ld, n_steps, step0, static_cast<real>(k->v_th),
                       static_cast<real*>(k->v_last), k->count, k->steps, k->cap);
  }
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
  // move the executable graph to the device now, not at its first launch
  e = hipGraphUpload(h->exec, s);
  if (e != hipSuccess) {
    (void)hipGraphExecDestroy(h->exec);
    (void)hipGraphDestroy(h->graph);
    delete h;
    return bc::fail(BC_HIP_ERROR, "hipGraphUpload: %s", hipGetErrorString(e));
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
