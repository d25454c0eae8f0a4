// Hand-written adjoint of the staggered step (reverse mode, one workgroup
// per pack, steps of one checkpoint segment in reverse order).
//
// Forward step n (braincell/quad/_staggered.py:126-154, family order):
//   X      = M(G_n)^-1 R_n        DHS solve; V_{n+1} = X at CV rows
//            M_ii = 1 + a_ii + dt/Cm G_i,  R_i = V_i + dt/Cm (sum g_eff E + I_clamp)
//   Ci'    = Ci + dt exprel(-dt/tau) (max(k I_Ca(V_n, gates_n, Ci_n), 0) + (C_rest - Ci)/tau)
//   x'     = x + dt exprel(u(V_{n+1}, Ci')) f(x, V_{n+1}, Ci')      every gate
// Reverse (given adjoints of V_{n+1}, gates_{n+1}, Ci_{n+1}):
//   gates: xbar_n = xbar' e^u;  Xbar += xbar' dx'/dV;  Cibar' += xbar'_AHP dx'/dCi'
//   pool:  Cibar_n = Cibar' e^{-dt/tau};  I_Ca bar = Cibar' dt exprel(-dt/tau) k [k I_Ca > 0]
//   solve: Rbar = M^-T Xbar (same tree, lowers and uppers swapped);
//          Vbar_n += Rbar;  g_eff bar = dt/Cm Rbar (E - X);  E_Ca bar -> Cibar_n
//   g_eff = g_max prod(gates^p): -> gate adjoints and d loss / d g_max (per
//          mechanism class, i.e. per painted region of each cell).
// Step-start states come from the forward kernel's recording (rec).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc_channels.h"
#include "bc_common.h"

namespace {

template <typename real>
struct AdjArgs {
  const bc_pack* packs;
  const uint32_t* cv_info;
  const uint32_t* bnd_info;
  const int32_t* chain_off;
  const int32_t* chain_top;
  const int32_t* kid_off;
  const int32_t* kids;
  const int32_t* level_off;
  const int4* chain_rec;
  const int16_t* chain_par;
  const int16_t* kid_pad;
  const real* lo;
  const real* up;
  const real* cls_tab;
  const int32_t* probe_row;
  const int32_t* probe_col;
  int64_t n_cv_total;
  const real* rec;
  const real* trace_bar;
  int64_t trace_ld;
  real* Vbar;
  real* Sbar;
  real* cls_grad;
  int32_t n_steps;
  real dt;
  int32_t max_rows, max_chains, max_cls, max_levels;
  bc_chan_const<real> cc;
};

// Composite of a chain's transposed elimination (see tchain_composite):
// (sd, sr) -> ((a sd + b) / (h sd + 1), (c sd + f sr + g) / (h sd + 1)).
template <typename real>
struct alignas(4 * sizeof(real)) TComp {
  real a, b, h, c, f, g, pad0, pad1;
};

// LDS of one workgroup: the pack's chain records in the forward kernel's
// form {j0 | j1 << 16, top & 0xFFFF | n_kids << 16, kid0 | kid1 << 16,
// first kid entry} (bc_stagger_desc.chain_rec) and the chains' elimination
// composites first (16-byte aligned), then the rows' diagonal, rhs, lowers
// and uppers, the chains' kid contributions, the mechanism classes and the
// per-wavefront partial sums of the gradient reduction, then the pack's kid
// list, parent chains and level offsets (int16), and two more per-chain
// (real, real) buffers and two int16 pointer arrays for the back
// substitution's pointer jumping.  The transposed solve walks LDS only.
constexpr int kRedSlots = 16;  // 4 wavefronts x 4 gradient columns
__host__ __device__ inline size_t adj_lds_bytes(int max_rows, int max_chains, int max_cls, int max_levels,
                                                size_t real_size) {
  const size_t mc = (size_t)((max_chains + 3) & ~3);
  return mc * 16 + mc * 8 * real_size +
         (size_t)(4 * max_rows + 4 * max_chains + max_cls * BC_NCLS + kRedSlots) * real_size +
         (4 * mc + (size_t)((max_levels + 4) & ~3)) * sizeof(int16_t);
}

// Orders LDS accesses of the lanes of one wavefront across this point.
__device__ __forceinline__ void adj_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transposed elimination of row j: with (sd, sr) what the rows below
// subtract, e = d - sd, r = rhs - sr, the pivot is e - lo and the row passes
// up (up e, lo r) / (e - lo).  In homogeneous coordinates sd = x/z,
// sr = y/z the step is linear,
//   x' = -up x + up d z,   y' = -lo y + lo r z,   z' = -x + (d - lo) z,
// so a chain's rows compose into one map (the forward solve's composite with
// lo in place of up in the right-hand-side part), rescaled to z's sd-free
// term = 1 after every group of four rows.
template <typename real>
__device__ __forceinline__ void tcomp_row(real d, real r, real lo, real up, bool valid, real& a, real& b, real& h,
                                          real& k, real& c, real& f, real& g) {
  const real dl = d - lo;
  const real a2 = up * (d * h - a);
  const real b2 = up * (d * k - b);
  const real c2 = lo * (r * h - c);
  const real g2 = lo * (r * k - g);
  const real h2 = dl * h - a;
  const real k2 = dl * k - b;
  const real f2 = -lo * f;
  a = valid ? a2 : a;
  b = valid ? b2 : b;
  c = valid ? c2 : c;
  g = valid ? g2 : g;
  h = valid ? h2 : h;
  k = valid ? k2 : k;
  f = valid ? f2 : f;
}

// Pass 1: chain k's composite; a leaf chain (input zero) also stores its
// contribution (b, g).
template <typename real>
__device__ __forceinline__ void tchain_composite(int k, const int4* CR, const real* bd, const real* br, const real* LO,
                                                 const real* UP, TComp<real>* CMP, real* kd, real* kr) {
  const int4 cr = CR[k];
  const int j0 = cr.x & 0xFFFF, j1 = (int)((unsigned)cr.x >> 16);
  const int nk = (int)((unsigned)cr.y >> 16);
  real a = real(1), b = real(0), h = real(0), kk = real(1), c = real(0), f = real(1), g = real(0);
  for (int j = j0; j < j1; j += 4) {
    real gd[4], gr[4], gl[4], gu[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int jj = j + i < j1 ? j + i : j1 - 1;
      gd[i] = bd[jj];
      gr[i] = br[jj];
      gl[i] = LO[jj];
      gu[i] = UP[jj];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) tcomp_row<real>(gd[i], gr[i], gl[i], gu[i], j + i < j1, a, b, h, kk, c, f, g);
    const real ik = bc_rcp<real>(kk);
    a *= ik;
    b *= ik;
    h *= ik;
    c *= ik;
    f *= ik;
    g *= ik;
    kk = real(1);
  }
  TComp<real> m;
  m.a = a;
  m.b = b;
  m.h = h;
  m.c = c;
  m.f = f;
  m.g = g;
  m.pad0 = m.pad1 = real(0);
  CMP[k] = m;
  if (nk == 0) {
    kd[k] = b;
    kr[k] = g;
  }
}

// Sum of chain k's kid contributions in a fixed order (first kid, second
// kid, then the rest of the kid list).
template <typename real>
__device__ __forceinline__ void tkid_sum(const int4& cr, const int16_t* m_kid, const real* kd, const real* kr,
                                         real& sd, real& sr) {
  const int nk = (int)((unsigned)cr.y >> 16);
  const int q0 = cr.z & 0xFFFF, q1 = (int)((unsigned)cr.z >> 16);
  const real d0 = kd[q0], r0 = kr[q0], d1 = kd[q1], r1 = kr[q1];
  sd = nk > 1 ? d0 + d1 : (nk > 0 ? d0 : real(0));
  sr = nk > 1 ? r0 + r1 : (nk > 0 ? r0 : real(0));
  for (int q = 2; q < nk; ++q) {
    const int kq = m_kid[cr.w + q];
    sd += kd[kq];
    sr += kr[kq];
  }
}

// Pass 3: chain k's rows from its input (sd, sr): bd[j] becomes the inverse
// pivot, br[j] the reduced right-hand side.  Rows are read four at a time
// before any of them is written.
template <typename real>
__device__ __forceinline__ void trows(int k, const int4* CR, const int16_t* m_kid, real* bd, real* br, const real* LO,
                                      const real* UP, const real* kd, const real* kr) {
  const int4 cr = CR[k];
  const int j0 = cr.x & 0xFFFF, j1 = (int)((unsigned)cr.x >> 16);
  real sd, sr;
  tkid_sum<real>(cr, m_kid, kd, kr, sd, sr);
  for (int j = j0; j < j1; j += 4) {
    real glo[4], gup[4], gd[4], gr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int jj = j + i < j1 ? j + i : j1 - 1;
      glo[i] = LO[jj];
      gup[i] = UP[jj];
      gd[i] = bd[jj];
      gr[i] = br[jj];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (j + i < j1) {
        const real e = gd[i] - sd;
        const real r = gr[i] - sr;
        const real inv = bc_rcp<real>(e - glo[i]);
        bd[j + i] = inv;
        br[j + i] = r;
        sd = gup[i] * e * inv;
        sr = glo[i] * r * inv;
      }
    }
  }
}

constexpr uint32_t kRowMask = 0xFFFFu;
constexpr uint32_t kProbeBit = 1u << 30;
constexpr uint32_t kNone = 0xFFFFFFFFu;

template <typename real, uint32_t CH, int MAXC>
__global__ void __launch_bounds__(256) stagger_adjoint_kernel(AdjArgs<real> a) {
  constexpr int NS = bc_nstates(CH) > 0 ? bc_nstates(CH) : 1;
  constexpr int NG = 4;  // gradient columns: Na, KDR, IL, CaL
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int mc = (a.max_chains + 3) & ~3;
  int4* CR = reinterpret_cast<int4*>(smem_raw);
  TComp<real>* CMP = reinterpret_cast<TComp<real>*>(smem_raw + (size_t)mc * 16);  // elimination composites
  real* bd = reinterpret_cast<real*>(smem_raw + (size_t)mc * 16 + (size_t)mc * sizeof(TComp<real>));
  real* br = bd + a.max_rows;
  real* LO = br + a.max_rows;
  real* UP = LO + a.max_rows;
  real* kd = UP + a.max_rows;
  real* kr = kd + a.max_chains;
  real* kd2 = kr + a.max_chains;  // pointer jumping: second (A, B) buffer
  real* kr2 = kd2 + a.max_chains;
  real* cls = kr2 + a.max_chains;
  real* red = cls + a.max_cls * BC_NCLS;  // [wavefront][gradient column]
  int16_t* m_kid = reinterpret_cast<int16_t*>(red + kRedSlots);
  int16_t* m_par = m_kid + mc;  // parent chain, -1 for a root
  int16_t* m_ptr = m_par + mc;  // pointer jumping (two buffers)
  int16_t* m_ptr2 = m_ptr + mc;
  int16_t* m_lv = m_ptr2 + mc;

  const bc_pack P = a.packs[blockIdx.x];
  const int tid = threadIdx.x;
  const int NT = blockDim.x;
  for (int i = tid; i < P.n_cls * BC_NCLS; i += NT) cls[i] = a.cls_tab[(int64_t)P.cls0 * BC_NCLS + i];
  for (int i = tid; i < P.n_rows; i += NT) {
    LO[i] = a.lo[(int64_t)P.row0 + i];
    UP[i] = a.up[(int64_t)P.row0 + i];
  }
  for (int k = tid; k < P.n_ch; k += NT) {
    CR[k] = a.chain_rec[P.ch0 + k];
    m_kid[k] = a.kid_pad[P.ch0 + k];
    m_par[k] = a.chain_par[P.ch0 + k];
  }
  for (int i = tid; i <= P.n_lv; i += NT) m_lv[i] = (int16_t)a.level_off[P.lv0 + i];
  const bc_chan_const<real>& cc = a.cc;
  const real dt = a.dt;
  const int64_t ncv = a.n_cv_total;
  const int64_t slab = (int64_t)(bc_nstates(CH) + 1) * ncv;

  uint32_t info[MAXC];
  int pcol[MAXC];
  real vb[MAXC];
  real sb[MAXC][NS];
  real gg[MAXC][NG];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int c = tid + i * NT;
    info[i] = kNone;
    pcol[i] = -1;
    vb[i] = real(0);
#pragma unroll
    for (int s = 0; s < NS; ++s) sb[i][s] = real(0);
#pragma unroll
    for (int k = 0; k < NG; ++k) gg[i][k] = real(0);
    if (c < P.n_cv) {
      const int64_t g = (int64_t)P.cv0 + c;
      info[i] = a.cv_info[g];
      vb[i] = a.Vbar[g];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < bc_nstates(CH)) sb[i][s] = a.Sbar[(int64_t)s * ncv + g];
      if (info[i] & kProbeBit) {
        const int row = (int)(info[i] & kRowMask);
        for (int e = 0; e < P.n_probe; ++e)
          if (a.probe_row[P.probe0 + e] == row) pcol[i] = a.probe_col[P.probe0 + e];
      }
    }
  }
  __syncthreads();

  // Recorded states.  Up to 4 FP32 CVs per thread (kCarry) they live in
  // registers: cur = the start of step n (V, then the states), xn = V after
  // step n; the next step's record is loaded while this step's transposed
  // solve runs, and V after step n - 1 is this step's starting V, so every
  // record is read from HBM once.  Larger thread shares read the record
  // where it is used (a second record set would not fit the registers).
  constexpr int NR = bc_nstates(CH) + 1;
  constexpr bool kCarry = MAXC * NR * sizeof(real) <= 4 * 8 * sizeof(float);
  constexpr int CM = kCarry ? MAXC : 1, CR_ = kCarry ? NR : 1;
  real cur[CM][CR_], xn[CM];
  if constexpr (kCarry) {
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      xn[i] = real(0);
#pragma unroll
      for (int k = 0; k < NR; ++k) cur[i][k] = real(0);
      if (info[i] != kNone && a.n_steps > 0) {
        const int64_t g = (int64_t)P.cv0 + tid + i * NT;
        const real* r0 = a.rec + (int64_t)(a.n_steps - 1) * slab;
#pragma unroll
        for (int k = 0; k < NR; ++k) cur[i][k] = r0[(int64_t)k * ncv + g];
        xn[i] = a.rec[(int64_t)a.n_steps * slab + g];
      }
    }
  }
  // record value k (0: V, 1 + s: state s) at the start of step n, and V
  // after step n, of the thread's CV i
  auto rec0 = [&](int n, int i, int k) -> real {
    if constexpr (kCarry) return cur[kCarry ? i : 0][kCarry ? k : 0];
    else return a.rec[(int64_t)n * slab + (int64_t)k * ncv + (int64_t)P.cv0 + tid + i * NT];
  };
  auto rec1 = [&](int n, int i) -> real {
    if constexpr (kCarry) return xn[kCarry ? i : 0];
    else return a.rec[(int64_t)(n + 1) * slab + (int64_t)P.cv0 + tid + i * NT];
  };

  for (int n = a.n_steps - 1; n >= 0; --n) {
    // ---- phase 1: gates and Ca pool backward; transposed-solve inputs
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] == kNone) continue;
      const int row = (int)(info[i] & kRowMask);
      const real* cp = cls + ((info[i] >> 16) & 0x3FFFu) * BC_NCLS;
      const real X = rec1(n, i);
      real s0[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) s0[s] = (s < bc_nstates(CH)) ? rec0(n, i, s + 1) : real(0);
      if (pcol[i] >= 0) vb[i] += a.trace_bar[(int64_t)n * a.trace_ld + pcol[i]];
      real Xbar = vb[i];
      if constexpr ((CH & BC_CH_NA) != 0) {
        real ap, bp, aq, bq, dap, dbp, daq, dbq, dxx, dxv;
        bc_na_rates_d<real, bc_fastmath<real>>(X, cc.na_vsh, ap, bp, aq, bq, dap, dbp, daq, dbq);
        bc_gate_ab_grad<real>(s0[bc_slot_na(CH)], ap, bp, dap, dbp, cc.na_phi, dt, dxx, dxv);
        Xbar += sb[i][bc_slot_na(CH)] * dxv;
        sb[i][bc_slot_na(CH)] *= dxx;
        bc_gate_ab_grad<real>(s0[bc_slot_na(CH) + 1], aq, bq, daq, dbq, cc.na_phi, dt, dxx, dxv);
        Xbar += sb[i][bc_slot_na(CH) + 1] * dxv;
        sb[i][bc_slot_na(CH) + 1] *= dxx;
      }
      if constexpr ((CH & BC_CH_KDR) != 0) {
        real ap, bp, dap, dbp, dxx, dxv;
        bc_kdr_rates_d<real, bc_fastmath<real>>(X, cc.kdr_vsh, ap, bp, dap, dbp);
        bc_gate_ab_grad<real>(s0[bc_slot_kdr(CH)], ap, bp, dap, dbp, cc.kdr_phi, dt, dxx, dxv);
        Xbar += sb[i][bc_slot_kdr(CH)] * dxv;
        sb[i][bc_slot_kdr(CH)] *= dxx;
      }
      if constexpr ((CH & BC_CH_CAL) != 0) {
        real pi, pt, qi, qt, dpi, dpt, dqi, dqt, dxx, dxv;
        bc_cal_rates_d<real, bc_fastmath<real>>(X, cc.cal_vsh, pi, pt, qi, qt, dpi, dpt, dqi, dqt);
        bc_gate_it_grad<real>(s0[bc_slot_cal(CH)], pi, pt, dpi, dpt, cc.cal_phi_p, dt, dxx, dxv);
        Xbar += sb[i][bc_slot_cal(CH)] * dxv;
        sb[i][bc_slot_cal(CH)] *= dxx;
        bc_gate_it_grad<real>(s0[bc_slot_cal(CH) + 1], qi, qt, dqi, dqt, cc.cal_phi_q, dt, dxx, dxv);
        Xbar += sb[i][bc_slot_cal(CH) + 1] * dxv;
        sb[i][bc_slot_cal(CH) + 1] *= dxx;
      }
      const real V0 = rec0(n, i, 0);
      if constexpr ((CH & BC_CH_POOL) != 0) {
        const int sc = bc_slot_ca(CH);
        const real ci = s0[sc];
        const real cib = sb[i][sc];  // adjoint of Ci_{n+1}
        const real u = -dt * cc.ca_tau_inv;
        const real ec = real(1) + u * bc_qexprel<real>(u);
        const real bcf = dt * bc_qexprel<real>(u);
        const real eca = cc.ca_rtzf * bc_fastmath<real>::lg(bc_qdiv<real>(cc.ca_Co, ci));
        real ica = real(0), gcal = real(0), pp = real(0), qq = real(0), gmax = real(0);
        if constexpr ((CH & BC_CH_CAL) != 0) {
          pp = s0[bc_slot_cal(CH)];
          qq = s0[bc_slot_cal(CH) + 1];
          gmax = cp[BC_CLS_G_CAL];
          gcal = gmax * (pp * pp) * qq;
          ica = gcal * (eca - V0);
        }
        const real mask = cp[BC_CLS_CAMASK];
        const real raw = mask * ica * cc.ca_drive;
        const real Ib = raw > real(0) ? cib * bcf * cc.ca_drive * mask : real(0);
        real cin = cib * ec;
        if constexpr ((CH & BC_CH_CAL) != 0) {
          vb[i] = -Ib * gcal;  // start of Vbar_n (the solve adds Rbar below)
          cin += Ib * gcal * bc_qdiv<real>(-cc.ca_rtzf, ci);
          const real drv = Ib * (eca - V0);  // d/d g_eff
          sb[i][bc_slot_cal(CH)] += drv * gmax * real(2) * pp * qq;
          sb[i][bc_slot_cal(CH) + 1] += drv * gmax * pp * pp;
          gg[i][3] += drv * (pp * pp) * qq;
        } else {
          vb[i] = real(0);
        }
        sb[i][sc] = cin;
      } else {
        vb[i] = real(0);
      }
      // transposed solve inputs (M has the same diagonal)
      real G = real(0);
      if constexpr ((CH & BC_CH_NA) != 0) {
        const real p = s0[bc_slot_na(CH)];
        G += cp[BC_CLS_G_NA] * (p * p * p) * s0[bc_slot_na(CH) + 1];
      }
      if constexpr ((CH & BC_CH_KDR) != 0) {
        const real p2 = s0[bc_slot_kdr(CH)] * s0[bc_slot_kdr(CH)];
        G += cp[BC_CLS_G_KDR] * (p2 * p2);
      }
      if constexpr ((CH & BC_CH_IL) != 0) G += cp[BC_CLS_G_IL];
      if constexpr ((CH & BC_CH_CAL) != 0) {
        const real p = s0[bc_slot_cal(CH)];
        G += cp[BC_CLS_G_CAL] * (p * p) * s0[bc_slot_cal(CH) + 1];
      }
      bd[row] = real(1) + cp[BC_CLS_DTCM] * G;
      br[row] = Xbar;
    }
    for (int i = tid; i < P.n_bnd; i += NT) {
      const int row = (int)(a.bnd_info[P.bnd0 + i] & 0x7FFFFFFFu);
      bd[row] = real(0);
      br[row] = real(0);
    }
    // the record of step n - 1, in flight during the solve
    real nxt[CM][CR_];
#pragma unroll
    for (int i = 0; i < CM; ++i) {
#pragma unroll
      for (int k = 0; k < CR_; ++k) nxt[i][k] = real(0);
      if (kCarry && info[i] != kNone && n > 0) {
        const int64_t g = (int64_t)P.cv0 + tid + i * NT;
        const real* rp = a.rec + (int64_t)(n - 1) * slab;
#pragma unroll
        for (int k = 0; k < CR_; ++k) nxt[i][k] = rp[(int64_t)k * ncv + g];
      }
    }
    __syncthreads();

    // ---- phase 2: Rbar = M^-T Xbar.  M^T has the same diagonal (rebuilt as
    // e - lo, as in the forward solve); the off-diagonals swap, so the
    // child's rhs contribution uses lowers and the back substitution uppers.
    // Elimination through chain composites: pass 1 composes every chain's
    // rows into one map (all chains at once); pass 2 applies one map per
    // chain level by level in wavefront 0 (chains are sorted by level, level
    // 0 being the leaves, whose contributions pass 1 already stored) while
    // the other wavefronts run pass 3 on the leaf chains; pass 3 then
    // eliminates every other chain's rows from its kids' contributions.
    // Every chain's arithmetic is the same whichever thread takes it.
    {
      for (int k = tid; k < P.n_ch; k += NT) tchain_composite<real>(k, CR, bd, br, LO, UP, CMP, kd, kr);
      __syncthreads();
      const int lv0 = m_lv[0], lv1 = m_lv[1];
      const int lw = NT > 64 ? 64 : 0;  // first thread of the leaf crew
      if (tid < 64) {
        // the top level holds root chains only, whose contributions no chain
        // reads: the sweep ends below it
        for (int lev = 1; lev < P.n_lv - 1; ++lev) {
          const int c1 = m_lv[lev + 1];
          for (int k = m_lv[lev] + tid; k < c1; k += 64) {
            const int4 cr = CR[k];
            const TComp<real> m = CMP[k];
            real sd, sr;
            tkid_sum<real>(cr, m_kid, kd, kr, sd, sr);
            const real iz = bc_rcp<real>(m.h * sd + real(1));
            kd[k] = (m.a * sd + m.b) * iz;
            kr[k] = (m.c * sd + m.f * sr + m.g) * iz;
          }
          adj_wave_sync();
        }
      }
      if (lw > 0 && tid >= lw)
        for (int k = lv0 + tid - lw; k < lv1; k += NT - lw) trows<real>(k, CR, m_kid, bd, br, LO, UP, kd, kr);
      __syncthreads();
      for (int k = (lw > 0 ? lv1 : lv0) + tid; k < P.n_ch; k += NT) trows<real>(k, CR, m_kid, bd, br, LO, UP, kd, kr);
    }
    __syncthreads();
    // Back substitution in affine form (the transposed system's x_j =
    // (r_j - up_j x_{j+1}) d_j inside a chain, d_j the stored inverse
    // pivot): every chain at once computes x_j = alpha_j + beta_j X, X the
    // solution of the row it hangs from (its parent chain's bottom row),
    // alpha and beta overwriting br and bd; the chains' bottom solutions then
    // follow by pointer jumping over the chain tree, and every row takes its
    // solution.  Rows are read four at a time.
    for (int k = tid; k < P.n_ch; k += NT) {
      const int4 c = CR[k];
      const int j0 = c.x & 0xFFFF, j1 = (int)((unsigned)c.x >> 16);
      const int tp = (c.y << 16) >> 16;
      real al = real(0), be = tp >= 0 ? real(1) : real(0);
      for (int j = j1 - 1; j >= j0; j -= 4) {
        real gb[4], gu[4], gd[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int jj = j - i >= j0 ? j - i : j0;
          gb[i] = br[jj];
          gu[i] = UP[jj];
          gd[i] = bd[jj];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (j - i >= j0) {
            const real a2 = (gb[i] - gu[i] * al) * gd[i];
            const real b2 = -gu[i] * be * gd[i];
            br[j - i] = a2;
            bd[j - i] = b2;
            al = a2;
            be = b2;
          }
        }
      }
      kd[k] = al;
      kr[k] = be;
      m_ptr[k] = m_par[k];
    }
    __syncthreads();
    const real* Xb;
    {
      // (A, B, ptr): x_bottom(k) = A + B x_bottom(ptr); ptr < 0 once resolved.
      // A forest of at most 32767 chains resolves in 16 rounds.
      real *A0 = kd, *B0 = kr, *A1 = kd2, *B1 = kr2;
      int16_t *p0 = m_ptr, *p1 = m_ptr2;
      for (int round = 0; round <= 16; ++round) {
        int open = 0;
        for (int k = tid; k < P.n_ch; k += NT) {
          const int q = p0[k];
          real va = A0[k], vb = B0[k];
          int qn = q;
          if (q >= 0) {
            va = va + vb * A0[q];
            vb = vb * B0[q];
            qn = p0[q];
            open |= qn >= 0;
          }
          A1[k] = va;
          B1[k] = vb;
          p1[k] = (int16_t)qn;
        }
        const int more = __syncthreads_or(open);
        real* t = A0; A0 = A1; A1 = t;
        t = B0; B0 = B1; B1 = t;
        int16_t* tq = p0; p0 = p1; p1 = tq;
        if (!more) break;
      }
      Xb = A0;
    }
    for (int k = tid; k < P.n_ch; k += NT) {
      const int4 c = CR[k];
      const int j0 = c.x & 0xFFFF, j1 = (int)((unsigned)c.x >> 16);
      const int par = m_par[k];
      const real X = par >= 0 ? Xb[par] : real(0);
      for (int j = j0; j < j1; ++j) br[j] = br[j] + bd[j] * X;
    }
    __syncthreads();

    // ---- phase 3: linearisation backward
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] == kNone) continue;
      const int row = (int)(info[i] & kRowMask);
      const real* cp = cls + ((info[i] >> 16) & 0x3FFFu) * BC_NCLS;
      const real X = rec1(n, i);
      const real Rb = br[row];
      const real w = cp[BC_CLS_DTCM] * Rb;  // d loss / d (dt/Cm * g_eff * E) per unit
      vb[i] += Rb;
      if constexpr ((CH & BC_CH_NA) != 0) {
        const real p = rec0(n, i, bc_slot_na(CH) + 1), q = rec0(n, i, bc_slot_na(CH) + 2);
        const real gbar = w * (cc.E_Na - X);
        const real gm = cp[BC_CLS_G_NA];
        sb[i][bc_slot_na(CH)] += gbar * gm * real(3) * p * p * q;
        sb[i][bc_slot_na(CH) + 1] += gbar * gm * p * p * p;
        gg[i][0] += gbar * p * p * p * q;
      }
      if constexpr ((CH & BC_CH_KDR) != 0) {
        const real p = rec0(n, i, bc_slot_kdr(CH) + 1);
        const real gbar = w * (cc.E_K - X);
        sb[i][bc_slot_kdr(CH)] += gbar * cp[BC_CLS_G_KDR] * real(4) * p * p * p;
        gg[i][1] += gbar * (p * p) * (p * p);
      }
      if constexpr ((CH & BC_CH_IL) != 0) gg[i][2] += w * (cc.E_IL - X);
      if constexpr ((CH & BC_CH_CAL) != 0) {
        const real p = rec0(n, i, bc_slot_cal(CH) + 1), q = rec0(n, i, bc_slot_cal(CH) + 2);
        real eca = cc.E_Ca_fixed;
        if constexpr ((CH & BC_CH_POOL) != 0) {
          const real ci = rec0(n, i, bc_slot_ca(CH) + 1);
          eca = cc.ca_rtzf * bc_fastmath<real>::lg(bc_qdiv<real>(cc.ca_Co, ci));
          const real geff = cp[BC_CLS_G_CAL] * (p * p) * q;
          sb[i][bc_slot_ca(CH)] += w * geff * bc_qdiv<real>(-cc.ca_rtzf, ci);
        }
        const real gbar = w * (eca - X);
        const real gm = cp[BC_CLS_G_CAL];
        sb[i][bc_slot_cal(CH)] += gbar * gm * real(2) * p * q;
        sb[i][bc_slot_cal(CH) + 1] += gbar * gm * p * p;
        gg[i][3] += gbar * (p * p) * q;
      }
    }
    if constexpr (kCarry) {
#pragma unroll
      for (int i = 0; i < CM; ++i) {
        xn[i] = cur[i][0];
#pragma unroll
        for (int k = 0; k < CR_; ++k) cur[i][k] = nxt[i][k];
      }
    }
  }

  // carried adjoints out
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    if (info[i] == kNone) continue;
    const int64_t g = (int64_t)P.cv0 + tid + i * NT;
    a.Vbar[g] = vb[i];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (s < bc_nstates(CH)) a.Sbar[(int64_t)s * ncv + g] = sb[i][s];
  }
  // parameter gradients per mechanism class, reduced in a fixed order (the
  // result does not depend on scheduling): every thread sums its own CVs in
  // register order, a butterfly sums the lanes of a wavefront, thread 0 adds
  // the wavefronts in order.  A class belongs to one pack and segment
  // launches run in order, so thread 0's update of its class is the only one.
  const int lane = tid & 63, wave = tid >> 6, n_waves = NT >> 6;
  constexpr int gcol[NG] = {BC_CLS_G_NA, BC_CLS_G_KDR, BC_CLS_G_IL, BC_CLS_G_CAL};
  for (int c = 0; c < P.n_cls; ++c) {
    real sm[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) sm[k] = real(0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      if (info[i] != kNone && (int)((info[i] >> 16) & 0x3FFFu) == c) {
#pragma unroll
        for (int k = 0; k < NG; ++k) sm[k] += gg[i][k];
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int k = 0; k < NG; ++k) sm[k] += __shfl_xor(sm[k], off);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NG; ++k) red[wave * NG + k] = sm[k];
    }
    __syncthreads();
    if (tid == 0) {
      for (int k = 0; k < NG; ++k) {
        real t = red[k];
        for (int w = 1; w < n_waves; ++w) t += red[w * NG + k];
        a.cls_grad[(int64_t)(P.cls0 + c) * BC_NCLS + gcol[k]] += t;
      }
    }
    __syncthreads();
  }
}

constexpr uint32_t kSetPassive = BC_CH_IL;
constexpr uint32_t kSetHH = BC_CH_NA | BC_CH_KDR | BC_CH_IL;
constexpr uint32_t kSetHHCaL = BC_CH_NA | BC_CH_KDR | BC_CH_IL | BC_CH_CAL | BC_CH_POOL;

template <typename real, uint32_t CH>
void* adj_pick_maxc(int maxc) {
  switch (maxc) {
    case 1: return (void*)&stagger_adjoint_kernel<real, CH, 1>;
    case 2: return (void*)&stagger_adjoint_kernel<real, CH, 2>;
    case 4: return (void*)&stagger_adjoint_kernel<real, CH, 4>;
    case 8: return (void*)&stagger_adjoint_kernel<real, CH, 8>;
    default: return nullptr;
  }
}

template <typename real>
void* adj_pick(uint32_t ch, int maxc) {
  switch (ch) {
    case kSetPassive: return adj_pick_maxc<real, kSetPassive>(maxc);
    case kSetHH: return adj_pick_maxc<real, kSetHH>(maxc);
    case kSetHHCaL: return adj_pick_maxc<real, kSetHHCaL>(maxc);
    default: return nullptr;
  }
}

template <typename real>
int adj_launch(const bc_adjoint_desc* d, hipStream_t s) {
  const bc_stagger_desc* f = d->layout;
  void* fn = adj_pick<real>(f->channels, f->cv_per_thread);
  if (fn == nullptr)
    return bc::fail(BC_UNSUPPORTED, "no adjoint kernel for channels=0x%x cv_per_thread=%d", f->channels,
                    f->cv_per_thread);
  AdjArgs<real> a;
  a.packs = f->packs;
  a.cv_info = f->cv_info;
  a.bnd_info = f->bnd_info;
  a.chain_off = f->chain_off;
  a.chain_top = f->chain_top;
  a.kid_off = f->kid_off;
  a.kids = f->kids;
  a.level_off = f->level_off;
  a.chain_rec = reinterpret_cast<const int4*>(f->chain_rec);
  a.chain_par = f->chain_par;
  a.kid_pad = f->kid_pad;
  a.lo = static_cast<const real*>(f->lo);
  a.up = static_cast<const real*>(f->up);
  a.cls_tab = static_cast<const real*>(f->cls_tab);
  a.probe_row = f->probe_row;
  a.probe_col = f->probe_col;
  a.n_cv_total = f->n_cv_total;
  a.rec = static_cast<const real*>(d->rec);
  a.trace_bar = static_cast<const real*>(d->trace_bar);
  a.trace_ld = f->trace_ld;
  a.Vbar = static_cast<real*>(d->Vbar);
  a.Sbar = static_cast<real*>(d->Sbar);
  a.cls_grad = static_cast<real*>(d->cls_grad);
  a.n_steps = d->n_steps;
  a.dt = static_cast<real>(f->dt);
  a.max_rows = f->max_rows;
  a.max_chains = f->max_chains;
  a.max_cls = f->max_cls;
  a.max_levels = f->max_levels;
  a.cc = *static_cast<const bc_chan_const<real>*>(f->chan_const);
  if (f->max_rows > 0xFFFF || f->max_chains > 0x7FFF)
    return bc::fail(BC_INVALID_LAYOUT, "adjoint pack too large (max_rows %d, max_chains %d)", f->max_rows,
                    f->max_chains);
  const size_t lds = adj_lds_bytes(f->max_rows, f->max_chains, f->max_cls, f->max_levels, sizeof(real));
  if (lds > 160 * 1024) return bc::fail(BC_INVALID_LAYOUT, "adjoint pack needs %zu bytes of LDS", lds);
  if (lds > 64 * 1024) BC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (f->n_packs == 0 || d->n_steps == 0) return BC_OK;
  void* args[] = {&a};
  BC_HIP(hipLaunchKernel(fn, dim3(f->n_packs), dim3(f->threads), args, lds, s));
  BC_HIP(hipGetLastError());
  return BC_OK;
}

}  // namespace

extern "C" {

int bc_stagger_adjoint(const bc_adjoint_desc* d, void* stream) {
  if (d == nullptr || d->layout == nullptr) return bc::fail(BC_INVALID_ARG, "desc/layout is NULL");
  const bc_stagger_desc* f = d->layout;
  if (f->precision != 32 && f->precision != 64) return bc::fail(BC_INVALID_ARG, "precision must be 32 or 64");
  if (f->threads < 64 || f->threads > 256 || f->threads % 64 != 0)
    return bc::fail(BC_INVALID_ARG, "threads must be 64, 128, 192 or 256");
  if (d->n_steps < 0) return bc::fail(BC_INVALID_ARG, "n_steps < 0");
  if (f->chan_const == nullptr) return bc::fail(BC_INVALID_ARG, "chan_const is NULL");
  if (f->flags & (BC_STEP_SKIP_GATES | BC_STEP_VOLTAGE_ONLY))
    return bc::fail(BC_UNSUPPORTED, "the adjoint differentiates the full step: BC_STEP_SKIP_GATES and "
                                    "BC_STEP_VOLTAGE_ONLY are not supported");
  if (f->syn != nullptr) return bc::fail(BC_UNSUPPORTED, "the adjoint does not differentiate synapses");
  if (f->n_packs > 0 && (f->chain_rec == nullptr || f->chain_par == nullptr || f->kid_pad == nullptr ||
                         f->level_off == nullptr))
    return bc::fail(BC_INVALID_ARG, "layout chain tables (chain_rec, chain_par, kid_pad, level_off) are NULL");
  if (f->n_packs > 0 && (d->rec == nullptr || d->Vbar == nullptr || d->Sbar == nullptr || d->cls_grad == nullptr ||
                         (f->trace_ld > 0 && d->trace_bar == nullptr)))
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  int rc = f->precision == 32 ? adj_launch<float>(d, static_cast<hipStream_t>(stream))
                              : adj_launch<double>(d, static_cast<hipStream_t>(stream));
  if (rc != BC_OK) return rc;
  return bc::ok();
}

int bc_adjoint_supported(int precision, uint32_t channels, int cv_per_thread) {
  if (precision == 32) return adj_pick<float>(channels, cv_per_thread) != nullptr;
  if (precision == 64) return adj_pick<double>(channels, cv_per_thread) != nullptr;
  return 0;
}

}  // extern "C"
