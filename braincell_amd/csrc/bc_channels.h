// Channel and ion-pool math shared by the staggered and single-compartment
// kernels.  Each function restates one reference rate method:
//
//   Na_Ba2002   braincell/channel/sodium.py:158-170
//   KDR_Ba2002  braincell/channel/potassium.py:206-214
//   CaL_IS2008  braincell/channel/calcium.py:902-920
//   CaT_HM1992  braincell/channel/calcium.py:345-373
//   AHP_De1994  braincell/channel/potassium_calcium.py:179-185
//   HCN_HM1992  braincell/channel/hyperpolarization_activated.py:124-132
//   CalciumDetailed drive  braincell/ion/calcium.py:499-503
//   Nernst E    braincell/ion/_base.py:462-469
//   exp-Euler   braincell/quad/_exp_euler.py:462-465
//
// Units: mV, ms, mS/cm^2, uA/cm^2, mM.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#define BC_CH_NA   (1u << 0)
#define BC_CH_KDR  (1u << 1)
#define BC_CH_IL   (1u << 2)
#define BC_CH_KL   (1u << 3)
#define BC_CH_CAL  (1u << 4)
#define BC_CH_CAT  (1u << 5)
#define BC_CH_AHP  (1u << 6)
#define BC_CH_HCN  (1u << 7)
#define BC_CH_POOL (1u << 8)
#define BC_CH_NAHH (1u << 9)  // Na_HH1952: the Na gate slots, density column and constants
#define BC_CH_KHH  (1u << 10) // K_HH1952: the KDR gate slot, density column and constants
#define BC_CH_NA_ANY (BC_CH_NA | BC_CH_NAHH)
#define BC_CH_K4_ANY (BC_CH_KDR | BC_CH_KHH)

// Mechanism-class table columns (per painted region of a cell).
#define BC_CLS_DTCM   0   // dt / Cm  (ms cm^2 / uF)
#define BC_CLS_G_NA   1   // mS/cm^2
#define BC_CLS_G_KDR  2
#define BC_CLS_G_IL   3
#define BC_CLS_G_KL   4
#define BC_CLS_G_CAL  5
#define BC_CLS_G_CAT  6
#define BC_CLS_G_AHP  7
#define BC_CLS_G_HCN  8
#define BC_CLS_CAMASK 9   // 1 where the Ca pool is painted
#define BC_NCLS       12

// Number of gate/pool states a channel set carries, and their slots.
__host__ __device__ constexpr int bc_nstates(uint32_t ch) {
  return ((ch & BC_CH_NA_ANY) ? 2 : 0) + ((ch & BC_CH_K4_ANY) ? 1 : 0) + ((ch & BC_CH_CAL) ? 2 : 0) +
         ((ch & BC_CH_CAT) ? 2 : 0) + ((ch & BC_CH_AHP) ? 1 : 0) + ((ch & BC_CH_HCN) ? 1 : 0) +
         ((ch & BC_CH_POOL) ? 1 : 0);
}
__host__ __device__ constexpr int bc_slot_na(uint32_t) { return 0; }
__host__ __device__ constexpr int bc_slot_kdr(uint32_t ch) { return ((ch & BC_CH_NA_ANY) ? 2 : 0); }
__host__ __device__ constexpr int bc_slot_cal(uint32_t ch) { return bc_slot_kdr(ch) + ((ch & BC_CH_K4_ANY) ? 1 : 0); }
__host__ __device__ constexpr int bc_slot_cat(uint32_t ch) { return bc_slot_cal(ch) + ((ch & BC_CH_CAL) ? 2 : 0); }
__host__ __device__ constexpr int bc_slot_ahp(uint32_t ch) { return bc_slot_cat(ch) + ((ch & BC_CH_CAT) ? 2 : 0); }
__host__ __device__ constexpr int bc_slot_hcn(uint32_t ch) { return bc_slot_ahp(ch) + ((ch & BC_CH_AHP) ? 1 : 0); }
__host__ __device__ constexpr int bc_slot_ca(uint32_t ch) { return bc_slot_hcn(ch) + ((ch & BC_CH_HCN) ? 1 : 0); }

// Scalar constants of the channel set (same for every cell of a launch).
// Temperature factors are folded into phi_* on the host.
template <typename real>
struct bc_chan_const {
  real E_Na, E_K, E_IL;
  real na_vsh, na_phi;
  real kdr_vsh, kdr_phi;
  real cal_vsh, cal_phi_p, cal_phi_q;
  real cat_vsh, cat_phi_p, cat_phi_q;
  real ahp_alpha, ahp_beta, ahp_n, ahp_phi;
  real hcn_E, hcn_phi;
  real ca_drive;   // 10 / (2 F d)  (mM/ms per uA/cm^2)
  real ca_tau_inv; // 1 / tau (1/ms)
  real ca_rest;    // C_rest (mM)
  real ca_Co;      // Co (mM)
  real ca_rtzf;    // 1e3 R T / (z F)  (mV)
  real E_Ca_fixed; // used when no pool is present
};

template <typename real> struct bc_math;
template <> struct bc_math<float> {
  static __host__ __device__ __forceinline__ float ex(float x) { return expf(x); }
  static __host__ __device__ __forceinline__ float em1(float x) { return expm1f(x); }
  static __host__ __device__ __forceinline__ float lg(float x) { return logf(x); }
  static __host__ __device__ __forceinline__ float pw(float x, float y) { return powf(x, y); }
  static constexpr float exprel_eps = 1e-4f;
};
template <> struct bc_math<double> {
  static __host__ __device__ __forceinline__ double ex(double x) { return exp(x); }
  static __host__ __device__ __forceinline__ double em1(double x) { return expm1(x); }
  static __host__ __device__ __forceinline__ double lg(double x) { return log(x); }
  static __host__ __device__ __forceinline__ double pw(double x, double y) { return pow(x, y); }
  static constexpr double exprel_eps = 1e-8;
};

// Math policy of the staggered kernel: in FP32 the hardware exponential and
// logarithm (v_exp_f32 / v_log_f32 on exp2 / log2; relative error about
// |x| 2^-24 + 1 ulp), in FP64 the IEEE library functions.
template <typename real> struct bc_fastmath;
template <> struct bc_fastmath<float> {
  static __device__ __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
  static __device__ __forceinline__ float em1(float x) { return expm1f(x); }
  static __device__ __forceinline__ float lg(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
  static __device__ __forceinline__ float pw(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  static constexpr float exprel_eps = 1e-4f;
};
template <> struct bc_fastmath<double> : bc_math<double> {};

// Reciprocal: the hardware v_rcp_f32 (1 ulp) in FP32, an IEEE division in
// FP64.  bc_div(a, b) = a * rcp(b).
template <typename real> __device__ __forceinline__ real bc_rcp(real x);
template <> __device__ __forceinline__ float bc_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double bc_rcp<double>(double x) { return 1.0 / x; }
template <typename real>
__device__ __forceinline__ real bc_div(real a, real b) { return a * bc_rcp<real>(b); }

// (e^x - 1) / x with its removable singularity filled by the series.
template <typename real>
__host__ __device__ __forceinline__ real bc_exprel(real x) {
  if (fabs(x) < bc_math<real>::exprel_eps) return real(1) + x * (real(0.5) + x * real(1.0 / 6.0));
  return bc_math<real>::em1(x) / x;
}

// x / (e^x - 1) = 1 / exprel(x).  FP32: Bernoulli series for |x| <= 1/2
// (truncation below 3e-11, far under an FP32 ulp), x * rcp(e^x - 1) (no
// cancellation there) otherwise.  FP64: the IEEE quotient 1 / exprel(x).
// Other number types (the dual numbers of bc_dual.h) take the FP32 form.
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ real bc_inv_exprel(real x) {
  if constexpr (std::is_same<real, double>::value) return real(1) / bc_exprel<real>(x);
  if (fabs(x) <= real(0.5)) {
    const real x2 = x * x;
    return real(1) - real(0.5) * x +
           x2 * (real(1.0 / 12.0) + x2 * (real(-1.0 / 720.0) + x2 * (real(1.0 / 30240.0) + x2 * real(-1.0 / 1209600.0))));
  }
  return bc_div<real>(x, M::ex(x) - real(1));
}

// y + dt * exprel(dt * lam) * f
template <typename real>
__host__ __device__ __forceinline__ real bc_expeuler(real y, real f, real lam, real dt) {
  return y + dt * bc_exprel<real>(dt * lam) * f;
}

// Gate in alpha/beta form: f = phi (a (1 - x) - b x), lam = -phi (a + b);
// the exp-Euler step is exact for this linear ODE:
// x' = x + dt exprel(dt lam) f = x_inf + (x - x_inf) e^{dt lam}.
template <typename real>
__host__ __device__ __forceinline__ real bc_gate_ab(real x, real a, real b, real phi, real dt) {
  real f = phi * (a * (real(1) - x) - b * x);
  real lam = -phi * (a + b);
  return bc_expeuler<real>(x, f, lam, dt);
}
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ real bc_gate_ab_cf(real x, real a, real b, real phi, real dt) {
  const real s = a + b;
  const real xi = bc_div<real>(a, s);
  return xi + (x - xi) * M::ex(-dt * phi * s);
}

// Gate in inf/tau form: f = phi (inf - x) / tau, lam = -phi / tau.
template <typename real>
__host__ __device__ __forceinline__ real bc_gate_it(real x, real inf, real tau, real phi, real dt) {
  real itau = real(1) / tau;
  real f = phi * (inf - x) * itau;
  real lam = -phi * itau;
  return bc_expeuler<real>(x, f, lam, dt);
}
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ real bc_gate_it_cf(real x, real inf, real tau, real phi, real dt) {
  return inf + (x - inf) * M::ex(-dt * phi * bc_rcp<real>(tau));
}

// Rate functions.  Divisions by literal constants are multiplications by
// their reciprocals; other divisions use bc_div.
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_na_rates(real V, real vsh, real& ap, real& bp, real& aq, real& bq) {
  real Vp = V - vsh;
  ap = real(0.32 * 4.0) * bc_inv_exprel<real, M>(-(Vp - real(13)) * real(0.25));
  bp = real(0.28 * 5.0) * bc_inv_exprel<real, M>((Vp - real(40)) * real(0.2));
  aq = real(0.128) * M::ex(-(Vp - real(17)) * real(1.0 / 18.0));
  bq = bc_div<real>(real(4), real(1) + M::ex(-(Vp - real(40)) * real(0.2)));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_kdr_rates(real V, real vsh, real& ap, real& bp) {
  real Vp = V - vsh;
  ap = real(0.032 * 5.0) * bc_inv_exprel<real, M>(-(Vp - real(15)) * real(0.2));
  bp = real(0.5) * M::ex(-(Vp - real(10)) * real(1.0 / 40.0));
}

// Hodgkin-Huxley 1952 pair (channel/sodium.py:384-398,
// channel/potassium.py:450-456); V' = V - V_sh with V_sh = -45 mV by default.
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_nahh_rates(real V, real vsh, real& ap, real& bp, real& aq, real& bq) {
  real Vp = V - vsh;
  ap = bc_inv_exprel<real, M>(-(Vp - real(5)) * real(0.1));
  bp = real(4) * M::ex(-(Vp + real(20)) * real(1.0 / 18.0));
  aq = real(0.07) * M::ex(-(Vp + real(20)) * real(0.05));
  bq = bc_rcp<real>(real(1) + M::ex(-(Vp - real(10)) * real(0.1)));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_khh_rates(real V, real vsh, real& ap, real& bp) {
  real Vp = V - vsh;
  ap = real(0.1) * bc_inv_exprel<real, M>(-(Vp + real(10)) * real(0.1));
  bp = real(0.125) * M::ex(-(Vp + real(20)) * real(1.0 / 80.0));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_cal_rates(real V, real vsh, real& pi, real& pt, real& qi, real& qt) {
  real Vp = V - vsh;
  pi = bc_rcp<real>(real(1) + M::ex(-(Vp + real(10)) * real(0.25)));
  real e1 = M::ex((Vp + real(5)) * real(1.0 / 15.0));
  pt = real(0.4) + bc_div<real>(real(0.7), bc_rcp<real>(e1) + e1);
  qi = bc_rcp<real>(real(1) + M::ex((Vp + real(25)) * real(0.5)));
  real e2 = M::ex((Vp + real(40)) * real(1.0 / 9.5));
  qt = real(300) + bc_div<real>(real(100), e2 + bc_rcp<real>(e2));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_cat_rates(real V, real vsh, real& pi, real& pt, real& qi, real& qt) {
  real Vp = V - vsh;
  pi = bc_rcp<real>(real(1) + M::ex(-(Vp + real(59)) * real(1.0 / 6.2)));
  pt = bc_rcp<real>(M::ex(-(Vp + real(132)) * real(1.0 / 16.7)) + M::ex((Vp + real(16.8)) * real(1.0 / 18.2))) +
       real(0.612);
  qi = bc_rcp<real>(real(1) + M::ex((Vp + real(83)) * real(0.25)));
  qt = (Vp >= real(-80)) ? (M::ex(-(Vp + real(22)) * real(1.0 / 10.5)) + real(28))
                         : M::ex((Vp + real(467)) * real(1.0 / 66.6));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_hcn_rates(real V, real& pi, real& pt) {
  pi = bc_rcp<real>(real(1) + M::ex((V + real(75)) * real(1.0 / 5.5)));
  pt = bc_rcp<real>(M::ex(real(-0.086) * V - real(14.59)) + M::ex(real(0.0701) * V - real(1.87)));
}

// ----------------------------------------------------------------------------
// Derivatives used by the hand-written adjoint (reverse of one staggered step).

// Quotients and exprel at the precision policy of the hand-written point
// and adjoint kernels.  FP64: the IEEE quotient and bc_exprel.  FP32:
// v_rcp_f32 products, and exprel from its Taylor series for |x| <= 1/2
// (terms to x^7 / 8!, truncation below 1.2e-8 relative) or
// (2^(x log2 e) - 1) / x beyond, where e^x - 1 loses at most a factor 2.6
// in relative accuracy.  The library forms take about twice the
// instructions, and these kernels are bound by instruction issue.
template <typename real>
__device__ __forceinline__ real bc_qdiv(real a, real b) {
  if constexpr (std::is_same<real, double>::value) return a / b;
  else return a * __builtin_amdgcn_rcpf(b);
}

template <typename real>
__device__ __forceinline__ real bc_qexprel(real x) {
  if constexpr (std::is_same<real, double>::value) {
    return bc_exprel<double>(x);
  } else {
    if (fabsf(x) <= 0.5f)
      return 1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f + x * (1.0f / 120.0f +
             x * (1.0f / 720.0f + x * (1.0f / 5040.0f + x * (1.0f / 40320.0f)))))));
    return (__builtin_amdgcn_exp2f(x * 1.4426950408889634f) - 1.0f) * __builtin_amdgcn_rcpf(x);
  }
}

// d/dx exprel(x) = sum_k k x^(k-1) / (k+1)!  =  (x e^x - (e^x - 1)) / x^2.
// FP64: the series for |x| < 1e-3, the closed form beyond (expm1).  FP32:
// the series to k = 8 for |x| <= 1/2 (truncation below 2e-8 relative; the
// closed form cancels there), the closed form with the hardware exponential
// beyond (numerator >= 0.17, error ~1e-6 relative).
template <typename real>
__device__ __forceinline__ real bc_qexprel_d(real x) {
  if constexpr (std::is_same<real, double>::value) {
    if (fabs(x) < 1e-3) return 0.5 + x * (1.0 / 3.0 + x * (0.125 + x * (1.0 / 30.0 + x * (1.0 / 144.0))));
    const double em = expm1(x);
    return (x * (em + 1.0) - em) / (x * x);
  } else {
    if (fabsf(x) <= 0.5f)
      return 0.5f + x * (1.0f / 3.0f + x * (0.125f + x * (1.0f / 30.0f + x * (1.0f / 144.0f +
             x * (1.0f / 840.0f + x * (1.0f / 5760.0f + x * (1.0f / 45360.0f)))))));
    const float ex = __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
    const float ix = __builtin_amdgcn_rcpf(x);
    return (x * ex - (ex - 1.0f)) * (ix * ix);
  }
}

// Alpha/beta gate update x' = x + dt exprel(u) f, u = -dt phi (a + b):
// returns dx'/dx = e^u and dx'/dV given da/dV, db/dV.
template <typename real>
__device__ __forceinline__ void bc_gate_ab_grad(real x, real a, real b, real da, real db, real phi, real dt,
                                                real& dxx, real& dxv) {
  const real u = -dt * phi * (a + b);
  const real er = bc_qexprel<real>(u), erd = bc_qexprel_d<real>(u);
  const real f = phi * (a * (real(1) - x) - b * x);
  const real common = dt * erd * (-dt * phi) * f;
  const real dA = common + dt * er * phi * (real(1) - x);
  const real dB = common - dt * er * phi * x;
  dxx = real(1) + u * er;
  dxv = dA * da + dB * db;
}

// Inf/tau gate update: x' = x + dt exprel(u) phi (inf - x) w, w = 1/tau, u = -dt phi w.
template <typename real>
__device__ __forceinline__ void bc_gate_it_grad(real x, real inf, real tau, real dinf, real dtau, real phi, real dt,
                                                real& dxx, real& dxv) {
  const real w = bc_qdiv<real>(real(1), tau);
  const real u = -dt * phi * w;
  const real er = bc_qexprel<real>(u), erd = bc_qexprel_d<real>(u);
  const real f = phi * (inf - x) * w;
  const real dI = dt * er * phi * w;
  const real dW = dt * erd * (-dt * phi) * f + dt * er * phi * (inf - x);
  dxx = real(1) + u * er;
  dxv = dI * dinf + dW * (-dtau * w * w);
}

// Rates and their V derivatives (the adjoint's rate evaluations); M is the
// exponential policy (bc_math, or bc_fastmath for the hardware exp in FP32).
template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_na_rates_d(real V, real vsh, real& ap, real& bp, real& aq, real& bq, real& dap,
                                              real& dbp, real& daq, real& dbq) {
  const real Vp = V - vsh;
  const real xa = -(Vp - real(13)) * real(0.25), xb = (Vp - real(40)) * real(0.2);
  const real ea = bc_qexprel<real>(xa), eb = bc_qexprel<real>(xb);
  ap = bc_qdiv<real>(real(1.28), ea);
  bp = bc_qdiv<real>(real(1.4), eb);
  dap = bc_qdiv<real>(real(0.32) * bc_qexprel_d<real>(xa), ea * ea);
  dbp = bc_qdiv<real>(real(-0.28) * bc_qexprel_d<real>(xb), eb * eb);
  aq = real(0.128) * M::ex(bc_qdiv<real>(-(Vp - real(17)), real(18)));
  daq = bc_qdiv<real>(-aq, real(18));
  const real e = M::ex(-(Vp - real(40)) * real(0.2));
  bq = bc_qdiv<real>(real(4), real(1) + e);
  dbq = bq * bc_qdiv<real>(e, real(1) + e) * real(0.2);
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_kdr_rates_d(real V, real vsh, real& ap, real& bp, real& dap, real& dbp) {
  const real Vp = V - vsh;
  const real x = -(Vp - real(15)) * real(0.2);
  const real e = bc_qexprel<real>(x);
  ap = bc_qdiv<real>(real(0.16), e);
  dap = bc_qdiv<real>(real(0.032) * bc_qexprel_d<real>(x), e * e);
  bp = real(0.5) * M::ex(bc_qdiv<real>(-(Vp - real(10)), real(40)));
  dbp = bc_qdiv<real>(-bp, real(40));
}

template <typename real, typename M = bc_math<real>>
__device__ __forceinline__ void bc_cal_rates_d(real V, real vsh, real& pi, real& pt, real& qi, real& qt, real& dpi,
                                               real& dpt, real& dqi, real& dqt) {
  const real Vp = V - vsh;
  pi = bc_qdiv<real>(real(1), real(1) + M::ex(-(Vp + real(10)) * real(0.25)));
  dpi = pi * (real(1) - pi) * real(0.25);
  const real e1 = M::ex(bc_qdiv<real>(Vp + real(5), real(15)));
  const real r1 = bc_qdiv<real>(real(1), e1);
  const real s1 = r1 + e1;
  pt = real(0.4) + bc_qdiv<real>(real(0.7), s1);
  dpt = bc_qdiv<real>(real(-0.7) * bc_qdiv<real>(e1 - r1, real(15)), s1 * s1);
  qi = bc_qdiv<real>(real(1), real(1) + M::ex((Vp + real(25)) * real(0.5)));
  dqi = -qi * (real(1) - qi) * real(0.5);
  const real e2 = M::ex(bc_qdiv<real>(Vp + real(40), real(9.5)));
  const real r2 = bc_qdiv<real>(real(1), e2);
  const real s2 = e2 + r2;
  qt = real(300) + bc_qdiv<real>(real(100), s2);
  dqt = bc_qdiv<real>(real(-100) * bc_qdiv<real>(e2 - r2, real(9.5)), s2 * s2);
}
