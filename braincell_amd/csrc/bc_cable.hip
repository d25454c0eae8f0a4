// Linearised backward Euler for passive multi-compartment cells: the
// reference's backward_euler (quad/_backward_euler.py:31-64) applied to a
// Cell through apply_standard_solver_step (quad/_util.py:196-262).  V (CV
// space) is the cell's only differential state, f(V) = I_mem(V) / Cm - K V
// with K the CV axial operator, and jacrev_last_dim (quad/_util.py:265-310)
// gives per CV the column sum J_i = sum_q df_q / dV_i = -a_i - sum_q K_qi.
// One step: V_i += dt f_i / (1 - dt J_i), every f_i from the step-start V.
//
// One workgroup per cell runs all the launch's steps with the cell's state
// in LDS: the voltages (two buffers, one barrier per step) and, copied once
// per launch, its rows of the CSR operator K, a, b and the denominators
// 1 - dt J, so that a step reads global memory only for its clamps.
#include "bc_common.h"

namespace {

template <typename real>
struct CableArgs {
  const int32_t* cell_cv0;
  const int32_t* k_off;
  const int32_t* k_col;
  const real* k_val;
  const real* a;
  const real* b;
  const real* jac;
  const int32_t* clamp_off;
  const int32_t* clamp_cv;
  const int32_t* clamp_on;
  const int32_t* clamp_end;
  const real* clamp_amp;
  const int32_t* probe_off;
  const int32_t* probe_cv;
  const int32_t* probe_col;
  real* trace;
  real* V;
  uint32_t* status;
  int32_t trace_ld, max_cv, max_nnz, step0, n_steps;
  real dt;
};

// Dynamic LDS of one workgroup: V (two buffers), a, b, 1 - dt J, the CSR
// values (real); the row offsets (max_cv + 1) and columns (int32).
template <typename real>
size_t cable_lds_bytes(int max_cv, int max_nnz) {
  return (size_t)(5 * max_cv + max_nnz) * sizeof(real) + (size_t)(max_cv + 1 + max_nnz) * sizeof(int32_t);
}

template <typename real>
__global__ void __launch_bounds__(1024) cable_be_kernel(CableArgs<real> a) {
  extern __shared__ __align__(16) unsigned char bc_cable_lds[];
  real* Va = reinterpret_cast<real*>(bc_cable_lds);
  real* Vb = Va + a.max_cv;
  real* ca = Vb + a.max_cv;
  real* cb = ca + a.max_cv;
  real* den = cb + a.max_cv;
  real* kv = den + a.max_cv;
  int32_t* ko = reinterpret_cast<int32_t*>(kv + a.max_nnz);
  int32_t* kc = ko + a.max_cv + 1;
  const int c = blockIdx.x, tid = threadIdx.x, NT = blockDim.x;
  const int cv0 = a.cell_cv0[c], n = a.cell_cv0[c + 1] - cv0;
  const int e0 = a.k_off[cv0], ne = a.k_off[cv0 + n] - e0;
  const int cl0 = a.clamp_off[c], cl1 = a.clamp_off[c + 1];
  const int pr0 = a.probe_off[c], pr1 = a.probe_off[c + 1];
  if (n > a.max_cv || ne > a.max_nnz) {  // tables larger than the LDS plan: report, touch nothing
    if (tid == 0) atomicOr(a.status, 2u);
    return;
  }
  for (int i = tid; i < n; i += NT) {
    const int g = cv0 + i;
    Va[i] = a.V[g];
    ca[i] = a.a[g];
    cb[i] = a.b[g];
    den[i] = real(1) - a.dt * a.jac[g];
    ko[i] = a.k_off[g] - e0;
  }
  if (tid == 0) ko[n] = ne;
  for (int e = tid; e < ne; e += NT) {
    kv[e] = a.k_val[e0 + e];
    kc[e] = a.k_col[e0 + e];
  }
  __syncthreads();
  uint32_t bad = 0;
  for (int step = 0; step < a.n_steps; ++step) {
    const int gstep = a.step0 + step;
    for (int i = tid; i < n; i += NT) {
      real s = real(0);
      for (int e = ko[i]; e < ko[i + 1]; ++e) s += kv[e] * Va[kc[e]];
      real I = real(0);
      for (int k = cl0; k < cl1; ++k)
        if (a.clamp_cv[k] == i && gstep >= a.clamp_on[k] && gstep < a.clamp_end[k]) I += a.clamp_amp[k];
      const real v = Va[i];
      const real f = (cb[i] - ca[i] * v + I) - s;
      const real vn = v + a.dt * f / den[i];
      Vb[i] = vn;
      bad |= (vn == vn && fabs(vn) < real(1e30)) ? 0u : 1u;
    }
    __syncthreads();
    for (int p = pr0 + tid; p < pr1; p += NT) a.trace[(int64_t)step * a.trace_ld + a.probe_col[p]] = Vb[a.probe_cv[p]];
    real* t = Va;
    Va = Vb;
    Vb = t;
  }
  for (int i = tid; i < n; i += NT) a.V[cv0 + i] = Va[i];
  if (bad) atomicOr(a.status, bad);
}

template <typename real>
int launch_cable(const bc_cable_be_desc* d, hipStream_t st) {
  CableArgs<real> a;
  a.cell_cv0 = d->cell_cv0;
  a.k_off = d->k_off;
  a.k_col = d->k_col;
  a.k_val = static_cast<const real*>(d->k_val);
  a.a = static_cast<const real*>(d->a);
  a.b = static_cast<const real*>(d->b);
  a.jac = static_cast<const real*>(d->jac);
  a.clamp_off = d->clamp_off;
  a.clamp_cv = d->clamp_cv;
  a.clamp_on = d->clamp_on;
  a.clamp_end = d->clamp_end;
  a.clamp_amp = static_cast<const real*>(d->clamp_amp);
  a.probe_off = d->probe_off;
  a.probe_cv = d->probe_cv;
  a.probe_col = d->probe_col;
  a.trace = static_cast<real*>(d->trace);
  a.V = static_cast<real*>(d->V);
  a.status = d->status;
  a.trace_ld = d->trace_ld;
  a.max_cv = d->max_cv;
  a.max_nnz = d->max_nnz;
  a.step0 = d->step0;
  a.n_steps = d->n_steps;
  a.dt = static_cast<real>(d->dt);
  const size_t lds = cable_lds_bytes<real>(d->max_cv, d->max_nnz);
  if (lds > 160 * 1024) return bc::fail(BC_INVALID_LAYOUT, "a cell needs %zu bytes of LDS (> 160 KiB)", lds);
  if (d->n_cells == 0 || d->n_steps == 0) return BC_OK;
  void* fn = reinterpret_cast<void*>(&cable_be_kernel<real>);
  if (lds > 64 * 1024) BC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  void* args[] = {&a};
  BC_HIP(hipLaunchKernel(fn, dim3(d->n_cells), dim3(d->threads), args, lds, st));
  return BC_OK;
}

}  // namespace

extern "C" {

int bc_cable_be_run(const bc_cable_be_desc* d, void* stream) {
  if (d == nullptr) return bc::fail(BC_INVALID_ARG, "desc is NULL");
  if (d->precision != 32 && d->precision != 64) return bc::fail(BC_INVALID_ARG, "precision must be 32 or 64");
  if (d->n_cells < 0 || d->n_steps < 0 || d->step0 < 0)
    return bc::fail(BC_INVALID_ARG, "n_cells, n_steps and step0 must be >= 0");
  if (d->threads < 64 || d->threads > 1024 || d->threads % 64 != 0)
    return bc::fail(BC_INVALID_ARG, "threads must be a multiple of 64 in [64, 1024] (got %d)", d->threads);
  if (!(d->dt > 0.0)) return bc::fail(BC_INVALID_ARG, "dt must be > 0");
  if (d->n_cells > 0 && (d->max_cv <= 0 || d->max_nnz < 0))
    return bc::fail(BC_INVALID_LAYOUT, "max_cv must be > 0 and max_nnz >= 0");
  if (d->n_cells > 0 &&
      (d->cell_cv0 == nullptr || d->k_off == nullptr || d->k_col == nullptr || d->k_val == nullptr ||
       d->a == nullptr || d->b == nullptr || d->jac == nullptr || d->clamp_off == nullptr ||
       d->probe_off == nullptr || d->V == nullptr || d->status == nullptr))
    return bc::fail(BC_INVALID_ARG, "required pointer is NULL");
  if (d->trace == nullptr && d->n_steps > 0 && d->trace_ld > 0)
    return bc::fail(BC_INVALID_ARG, "a trace buffer is required when probes exist");
  int rc = d->precision == 32 ? launch_cable<float>(d, static_cast<hipStream_t>(stream))
                              : launch_cable<double>(d, static_cast<hipStream_t>(stream));
  if (rc != BC_OK) return rc;
  return bc::ok();
}

}  // extern "C"
