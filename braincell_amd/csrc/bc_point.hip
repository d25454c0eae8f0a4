// Independent exponential-Euler for a population of thalamocortical point
// neurons (one thread per neuron, every state in registers across steps).
//
// Restates ind_exp_euler_step (braincell/quad/_exp_euler.py:217-320,
// :330-478) on SingleCompartment.update (braincell/_single_compartment/base.py:252-278):
// every state y_k (V, nine gates, Ci) takes y_k + dt exprel(dt lam_k) f_k
// with f_k and lam_k = df_k/dy_k evaluated at the old state (Jacobi).  The
// Ca pool's current is evaluated at the old state (no step-start cache on
// this path, braincell/ion/_base.py:490-506), so lam_Ci includes the Nernst
// dependence of E_Ca on Ci while the drive is positive.
//
// State slots: Na p,q | KDR p | CaL p,q | CaT p,q | AHP p | HCN p | Ci.
// Conductance columns: Na, K_Leak, KDR, CaL, CaT, AHP, HCN, IL (mS/cm^2).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc_channels.h"
#include "bc_common.h"
#include "bc_dual.h"

namespace {

template <typename real>
struct PointArgs {
  int64_t n;
  real* V;
  real* S;          // [10 * n]
  const real* g;    // [8 * n]
  const real* Iext; // [n] uA/cm^2
  real* trace;      // [n_steps * n_trace] or null
  int64_t n_trace;
  uint32_t* spikes; // [n] or null
  uint32_t* status;
  int32_t n_steps;
  real dt, Cm, V_th;
  bc_chan_const<real> cc;
};

// Adds d to the pair (hi, lo) without losing the low bits (fast two-sum,
// |hi| >= |d| for state increments): slow gates change by less than half an
// FP32 ulp per step, and plain accumulation would drop those increments.
template <typename real>
__device__ __forceinline__ void acc2(real& hi, real& lo, real d) {
  const real s = hi + d;
  lo += d - (s - hi);
  const real h = s + lo;
  lo -= h - s;
  hi = h;
}

// FP32 math policy (hardware exp / log, v_rcp_f32 quotients, series exprel:
// bc_fastmath, bc_qdiv, bc_qexprel); FP64 keeps the IEEE functions.
template <typename real>
__device__ __forceinline__ real incr(real y, real f, real lam, real dt) {
  return dt * bc_qexprel<real>(dt * lam) * f;
}

template <typename real>
__device__ __forceinline__ real incr_ab(real x, real a, real b, real phi, real dt) {
  return incr<real>(x, phi * (a * (real(1) - x) - b * x), -phi * (a + b), dt);
}

template <typename real>
__device__ __forceinline__ real incr_it(real x, real inf, real tau, real phi, real dt) {
  const real it = bc_qdiv<real>(real(1), tau);
  return incr<real>(x, phi * (inf - x) * it, -phi * it, dt);
}

template <typename real>
__global__ void __launch_bounds__(256) point_kernel(PointArgs<real> a) {
  typedef bc_fastmath<real> M;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int64_t n = a.n;
  real V = a.V[i];
  real s[10], lo[10], vlo = real(0);
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    s[k] = a.S[k * n + i];
    lo[k] = real(0);
  }
  real g[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) g[k] = a.g[k * n + i];
  const real Iext = a.Iext[i];
  const real dt = a.dt;
  const bc_chan_const<real>& cc = a.cc;
  const real invC = real(1) / a.Cm;
  uint32_t nspk = 0, bad = 0;
  for (int step = 0; step < a.n_steps; ++step) {
    const real ci = s[9];
    const real eca = cc.ca_rtzf * M::lg(bc_qdiv<real>(cc.ca_Co, ci));
    // conductances at the old state
    const real g_na = g[0] * (s[0] * s[0] * s[0]) * s[1];
    const real g_kl = g[1];
    const real p2 = s[2] * s[2];
    const real g_kd = g[2] * (p2 * p2);
    const real g_cl = g[3] * (s[3] * s[3]) * s[4];
    const real g_ct = g[4] * (s[5] * s[5]) * s[6];
    const real g_ah = g[5] * (s[7] * s[7]);
    const real g_hc = g[6] * s[8];
    const real g_il = g[7];
    const real G = g_na + g_kl + g_kd + g_cl + g_ct + g_ah + g_hc + g_il;
    const real I = Iext + g_na * (cc.E_Na - V) + (g_kl + g_kd + g_ah) * (cc.E_K - V) + (g_cl + g_ct) * (eca - V) +
                   g_hc * (cc.hcn_E - V) + g_il * (cc.E_IL - V);
    const real dV = incr<real>(V, I * invC, -G * invC, dt);
    real d[10];
    // Ca pool with its own (old-state) current
    {
      const real gca = g_cl + g_ct;
      const real raw = gca * (eca - V) * cc.ca_drive;
      const real drive = raw > real(0) ? raw : real(0);
      const real f = drive + (cc.ca_rest - ci) * cc.ca_tau_inv;
      real lam = -cc.ca_tau_inv;
      if (raw > real(0)) lam -= bc_qdiv<real>(cc.ca_drive * gca * cc.ca_rtzf, ci);
      d[9] = incr<real>(ci, f, lam, dt);
    }
    {
      real ap, bp, aq, bq;
      bc_na_rates<real, M>(V, cc.na_vsh, ap, bp, aq, bq);
      d[0] = incr_ab<real>(s[0], ap, bp, cc.na_phi, dt);
      d[1] = incr_ab<real>(s[1], aq, bq, cc.na_phi, dt);
    }
    {
      real ap, bp;
      bc_kdr_rates<real, M>(V, cc.kdr_vsh, ap, bp);
      d[2] = incr_ab<real>(s[2], ap, bp, cc.kdr_phi, dt);
    }
    {
      real pi, pt, qi, qt;
      bc_cal_rates<real, M>(V, cc.cal_vsh, pi, pt, qi, qt);
      d[3] = incr_it<real>(s[3], pi, pt, cc.cal_phi_p, dt);
      d[4] = incr_it<real>(s[4], qi, qt, cc.cal_phi_q, dt);
    }
    {
      real pi, pt, qi, qt;
      bc_cat_rates<real, M>(V, cc.cat_vsh, pi, pt, qi, qt);
      d[5] = incr_it<real>(s[5], pi, pt, cc.cat_phi_p, dt);
      d[6] = incr_it<real>(s[6], qi, qt, cc.cat_phi_q, dt);
    }
    {
      const real al = cc.ahp_alpha * M::pw(ci, cc.ahp_n);
      d[7] = incr_ab<real>(s[7], al, cc.ahp_beta, cc.ahp_phi, dt);
    }
    {
      real pi, pt;
      bc_hcn_rates<real, M>(V, pi, pt);
      d[8] = incr_it<real>(s[8], pi, pt, cc.hcn_phi, dt);
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) acc2<real>(s[k], lo[k], d[k]);
    real Vn = V;
    acc2<real>(Vn, vlo, dV);
    nspk += (V < a.V_th && Vn >= a.V_th) ? 1u : 0u;
    bad |= (Vn == Vn && fabs(Vn) < real(1e30)) ? 0u : 1u;
    V = Vn;
    if (i < a.n_trace) a.trace[(int64_t)step * a.n_trace + i] = V;
  }
  a.V[i] = V;
#pragma unroll
  for (int k = 0; k < 10; ++k) a.S[k * n + i] = s[k];
  if (a.spikes) a.spikes[i] += nspk;
  if (bad) atomicOr(a.status, 1u);
}

// Gate terms of the linearly implicit steps: f, df/dV (from the dual rates)
// and lambda = df/dx.
template <typename real>
__device__ __forceinline__ void ab_terms(real x, bc_dual<real> a, bc_dual<real> b, real phi, real& f, real& fv,
                                         real& lam) {
  typedef bc_dual<real> D;
  const D F = D(phi) * (a * D(real(1) - x) - b * D(x));
  f = F.v;
  fv = F.d;
  lam = -phi * (a.v + b.v);
}

template <typename real>
__device__ __forceinline__ void it_terms(real x, bc_dual<real> inf, bc_dual<real> tau, real phi, real& f, real& fv,
                                         real& lam) {
  typedef bc_dual<real> D;
  const D it = bc_rcp<D>(tau);
  const D F = D(phi) * (inf - D(x)) * it;
  f = F.v;
  fv = F.d;
  lam = -phi * it.v;
}

// Linearly implicit steps over the 11 states y = (V, nine gates, Ci):
// (I - theta dt J) dy = dt f(y_n), y_{n+1} = y_n + dy, with J the full
// Jacobian at y_n.  theta = 1 is braincell's backward_euler
// (braincell/quad/_backward_euler.py:31-64: one solve with the Jacobian at
// the step start); theta = 1/2 is the Newton iterate of its implicit_euler
// (braincell/quad/_implicit.py:43-120, order 2, started from y_n:
// y_n - (I - dt/2 J)^-1 (-dt f(y_n))).  J is sparse: V couples to every state
// (conductance partials times driving force, and Ci through the Nernst
// E_Ca), each gate to V and itself (AHP: to Ci instead of V), Ci to V, the
// four Ca gates and itself while the pool's drive is positive.  The gates
// are eliminated onto (V, Ci) and the remaining 2x2 system is solved by
// Cramer's rule; gate pivots 1 + theta dt phi (a + b) are >= 1.
template <typename real>
__global__ void __launch_bounds__(256) point_implicit_kernel(PointArgs<real> a, real theta) {
  typedef bc_fastmath<real> M;
  typedef bc_dual<real> D;
  typedef bc_dmath<real, bc_fastmath<real>> DM;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int64_t n = a.n;
  real V = a.V[i];
  real s[10], lo[10], vlo = real(0);
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    s[k] = a.S[k * n + i];
    lo[k] = real(0);
  }
  real g[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) g[k] = a.g[k * n + i];
  const real Iext = a.Iext[i];
  const real dt = a.dt;
  const real h = theta * dt;
  const bc_chan_const<real>& cc = a.cc;
  const real invC = real(1) / a.Cm;
  uint32_t nspk = 0, bad = 0;
  for (int step = 0; step < a.n_steps; ++step) {
    const real ci = s[9];
    const real eca = cc.ca_rtzf * M::lg(bc_qdiv<real>(cc.ca_Co, ci));
    const real deca = bc_qdiv<real>(-cc.ca_rtzf, ci);  // dE_Ca / dCi
    // conductances and their partial derivatives in the gates
    const real p3 = s[0] * s[0] * s[0];
    const real g_na = g[0] * p3 * s[1];
    const real dna_p = g[0] * real(3) * s[0] * s[0] * s[1], dna_q = g[0] * p3;
    const real n3 = s[2] * s[2] * s[2];
    const real g_kd = g[2] * n3 * s[2], dkd = g[2] * real(4) * n3;
    const real g_cl = g[3] * s[3] * s[3] * s[4];
    const real dcl_p = g[3] * real(2) * s[3] * s[4], dcl_q = g[3] * s[3] * s[3];
    const real g_ct = g[4] * s[5] * s[5] * s[6];
    const real dct_p = g[4] * real(2) * s[5] * s[6], dct_q = g[4] * s[5] * s[5];
    const real g_ah = g[5] * s[7] * s[7], dah = g[5] * real(2) * s[7];
    const real g_hc = g[6] * s[8];
    const real g_kl = g[1], g_il = g[7];
    const real G = g_na + g_kl + g_kd + g_cl + g_ct + g_ah + g_hc + g_il;
    const real I = Iext + g_na * (cc.E_Na - V) + (g_kl + g_kd + g_ah) * (cc.E_K - V) + (g_cl + g_ct) * (eca - V) +
                   g_hc * (cc.hcn_E - V) + g_il * (cc.E_IL - V);
    const real fV = I * invC, JVV = -G * invC;
    real JV[10];  // d fV / d (gate k), k < 9, and d fV / dCi
    JV[0] = dna_p * (cc.E_Na - V) * invC;
    JV[1] = dna_q * (cc.E_Na - V) * invC;
    JV[2] = dkd * (cc.E_K - V) * invC;
    JV[3] = dcl_p * (eca - V) * invC;
    JV[4] = dcl_q * (eca - V) * invC;
    JV[5] = dct_p * (eca - V) * invC;
    JV[6] = dct_q * (eca - V) * invC;
    JV[7] = dah * (cc.E_K - V) * invC;
    JV[8] = g[6] * (cc.hcn_E - V) * invC;
    JV[9] = (g_cl + g_ct) * deca * invC;
    real f[9], fv[9], lam[9];
    const D Vd = D::of(V, real(1));
    {
      D ap, bp, aq, bq;
      bc_na_rates<D, DM>(Vd, D(cc.na_vsh), ap, bp, aq, bq);
      ab_terms<real>(s[0], ap, bp, cc.na_phi, f[0], fv[0], lam[0]);
      ab_terms<real>(s[1], aq, bq, cc.na_phi, f[1], fv[1], lam[1]);
    }
    {
      D ap, bp;
      bc_kdr_rates<D, DM>(Vd, D(cc.kdr_vsh), ap, bp);
      ab_terms<real>(s[2], ap, bp, cc.kdr_phi, f[2], fv[2], lam[2]);
    }
    {
      D pi, pt, qi, qt;
      bc_cal_rates<D, DM>(Vd, D(cc.cal_vsh), pi, pt, qi, qt);
      it_terms<real>(s[3], pi, pt, cc.cal_phi_p, f[3], fv[3], lam[3]);
      it_terms<real>(s[4], qi, qt, cc.cal_phi_q, f[4], fv[4], lam[4]);
    }
    {
      D pi, pt, qi, qt;
      bc_cat_rates<D, DM>(Vd, D(cc.cat_vsh), pi, pt, qi, qt);
      it_terms<real>(s[5], pi, pt, cc.cat_phi_p, f[5], fv[5], lam[5]);
      it_terms<real>(s[6], qi, qt, cc.cat_phi_q, f[6], fv[6], lam[6]);
    }
    // AHP: alpha = alpha0 Ci^n depends on Ci, not on V
    const real al = cc.ahp_alpha * M::pw(ci, cc.ahp_n);
    f[7] = cc.ahp_phi * (al * (real(1) - s[7]) - cc.ahp_beta * s[7]);
    lam[7] = -cc.ahp_phi * (al + cc.ahp_beta);
    fv[7] = cc.ahp_phi * bc_qdiv<real>(al * cc.ahp_n, ci) * (real(1) - s[7]);  // d f / dCi
    {
      D pi, pt;
      bc_hcn_rates<D, DM>(Vd, pi, pt);
      it_terms<real>(s[8], pi, pt, cc.hcn_phi, f[8], fv[8], lam[8]);
    }
    // Ca pool
    const real gca = g_cl + g_ct;
    const real raw = gca * (eca - V) * cc.ca_drive;
    const bool pos = raw > real(0);
    const real fC = (pos ? raw : real(0)) + (cc.ca_rest - ci) * cc.ca_tau_inv;
    const real JCC = -cc.ca_tau_inv + (pos ? cc.ca_drive * gca * deca : real(0));
    const real JCV = pos ? -cc.ca_drive * gca : real(0);
    const real kc = pos ? cc.ca_drive * (eca - V) : real(0);
    const real JC[4] = {kc * dcl_p, kc * dcl_q, kc * dct_p, kc * dct_q};
    // eliminate the gates: dy_k = A_k + B_k dV (AHP: A_7 + B_7 dCi)
    real m11 = real(1) - h * JVV, m12 = -h * JV[9], r1 = dt * fV;
    real m21 = -h * JCV, m22 = real(1) - h * JCC, r2 = dt * fC;
    real A[9], B[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const real piv = bc_qdiv<real>(real(1), real(1) - h * lam[k]);
      A[k] = dt * f[k] * piv;
      B[k] = h * fv[k] * piv;
      if (k == 7) {
        m12 -= h * JV[k] * B[k];
      } else {
        m11 -= h * JV[k] * B[k];
      }
      r1 += h * JV[k] * A[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      m21 -= h * JC[k] * B[3 + k];
      r2 += h * JC[k] * A[3 + k];
    }
    const real idet = bc_qdiv<real>(real(1), m11 * m22 - m12 * m21);
    const real dV = (r1 * m22 - m12 * r2) * idet;
    const real dC = (m11 * r2 - m21 * r1) * idet;
#pragma unroll
    for (int k = 0; k < 9; ++k) acc2<real>(s[k], lo[k], A[k] + B[k] * (k == 7 ? dC : dV));
    acc2<real>(s[9], lo[9], dC);
    real Vn = V;
    acc2<real>(Vn, vlo, dV);
    nspk += (V < a.V_th && Vn >= a.V_th) ? 1u : 0u;
    bad |= (Vn == Vn && fabs(Vn) < real(1e30)) ? 0u : 1u;
    V = Vn;
    if (i < a.n_trace) a.trace[(int64_t)step * a.n_trace + i] = V;
  }
  a.V[i] = V;
#pragma unroll
  for (int k = 0; k < 10; ++k) a.S[k * n + i] = s[k];
  if (a.spikes) a.spikes[i] += nspk;
  if (bad) atomicOr(a.status, 1u);
}

template <typename real>
int launch_point(const bc_point_desc* d, hipStream_t st) {
  if (d->n == 0 || d->n_steps == 0) return BC_OK;
  PointArgs<real> a;
  a.n = d->n;
  a.V = static_cast<real*>(d->V);
  a.S = static_cast<real*>(d->state);
  a.g = static_cast<const real*>(d->g);
  a.Iext = static_cast<const real*>(d->I_ext);
  a.trace = static_cast<real*>(d->trace);
  a.n_trace = d->trace ? d->n_trace : 0;
  a.spikes = d->spikes;
  a.status = d->status;
  a.n_steps = d->n_steps;
  a.dt = static_cast<real>(d->dt);
  a.Cm = static_cast<real>(d->Cm);
  a.V_th = static_cast<real>(d->V_th);
  a.cc = *static_cast<const bc_chan_const<real>*>(d->chan_const);
  const int bs = 256;
  dim3 grid((unsigned)((d->n + bs - 1) / bs));
  if (d->method == BC_POINT_IND_EXP_EULER) {
    hipLaunchKernelGGL(point_kernel<real>, grid, dim3(bs), 0, st, a);
  } else {
    const real theta = d->method == BC_POINT_BACKWARD_EULER ? real(1) : real(0.5);
    hipLaunchKernelGGL(point_implicit_kernel<real>, grid, dim3(bs), 0, st, a, theta);
  }
  BC_HIP(hipGetLastError());
  return BC_OK;
}

}  // namespace

extern "C" {

int bc_point_supported(int precision) { return precision == 32 || precision == 64; }

int bc_point_run(const bc_point_desc* d, void* stream) {
  if (d == nullptr) return bc::fail(BC_INVALID_ARG, "desc is NULL");
  if (d->precision != 32 && d->precision != 64) return bc::fail(BC_INVALID_ARG, "precision must be 32 or 64");
  if (d->n < 0 || d->n_steps < 0) return bc::fail(BC_INVALID_ARG, "n and n_steps must be >= 0");
  if (d->method != BC_POINT_IND_EXP_EULER && d->method != BC_POINT_BACKWARD_EULER &&
      d->method != BC_POINT_IMPLICIT_EULER)
    return bc::fail(BC_INVALID_ARG, "unknown point-neuron method %d", d->method);
  if (!(d->dt > 0.0) || !(d->Cm > 0.0)) return bc::fail(BC_INVALID_ARG, "dt and Cm must be > 0");
  if (d->n > 0 && (d->V == nullptr || d->state == nullptr || d->g == nullptr || d->I_ext == nullptr ||
                   d->status == nullptr || d->chan_const == nullptr))
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  if (d->trace != nullptr && (d->n_trace < 0 || d->n_trace > d->n))
    return bc::fail(BC_INVALID_ARG, "n_trace must lie in [0, n]");
  int rc = d->precision == 32 ? launch_point<float>(d, static_cast<hipStream_t>(stream))
                              : launch_point<double>(d, static_cast<hipStream_t>(stream));
  if (rc != BC_OK) return rc;
  return bc::ok();
}

}  // extern "C"
