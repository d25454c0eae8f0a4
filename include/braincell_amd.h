/*
 * braincell_amd C ABI: MI355X (gfx950) kernels for braincell's staggered
 * multi-compartment step and independent exp-Euler point neurons.
 *
 * Conventions: every entry point returns an int status (BC_OK = 0) and
 * leaves a thread-local message for bc_last_error().  Pointers are device
 * pointers (allocated by the caller, e.g. torch) unless stated otherwise.
 * No entry point allocates device memory except bc_graph_* (which owns a
 * captured hipGraphExec).  A stream argument of NULL means the default
 * stream.  Calls on one stream run in order; calls that share state buffers
 * must not run concurrently.
 */
#ifndef BRAINCELL_AMD_H
#define BRAINCELL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BC_OK = 0,
  BC_INVALID_ARG = 1,
  BC_INVALID_LAYOUT = 2,
  BC_HIP_ERROR = 3,
  BC_NONFINITE = 4,
  BC_UNSUPPORTED = 5
};

/* Channel-set bits (which mechanisms a launch computes). */
#define BC_CH_NA   (1u << 0) /* Na_Ba2002 */
#define BC_CH_KDR  (1u << 1) /* KDR_Ba2002 */
#define BC_CH_IL   (1u << 2) /* IL */
#define BC_CH_KL   (1u << 3) /* K_Leak */
#define BC_CH_CAL  (1u << 4) /* CaL_IS2008 */
#define BC_CH_CAT  (1u << 5) /* CaT_HM1992 */
#define BC_CH_AHP  (1u << 6) /* AHP_De1994 */
#define BC_CH_HCN  (1u << 7) /* HCN_HM1992 */
#define BC_CH_POOL (1u << 8) /* CalciumDetailed */
#define BC_CH_NAHH (1u << 9) /* Na_HH1952 (channel/sodium.py:293-398); never with BC_CH_NA */
#define BC_CH_KHH  (1u << 10) /* K_HH1952 (channel/potassium.py:341-456); never with BC_CH_KDR */

/* Per-pack header: one workgroup integrates one pack (a forest of whole
 * cells).  All *0 fields are global offsets into the corresponding arrays;
 * rows, chains and mechanism classes inside a pack are numbered locally. */
typedef struct bc_pack {
  int32_t cv0, n_cv;       /* CV range: cv_info / state arrays          */
  int32_t row0, n_rows;    /* DHS rows of the pack                      */
  int32_t ch0, n_ch;       /* chain range: chain_off / chain_top / kids */
  int32_t lv0, n_lv;       /* level_off[lv0 .. lv0 + n_lv] (pack chains)*/
  int32_t bnd0, n_bnd;     /* boundary (algebraic) rows: bnd_info       */
  int32_t cls0, n_cls;     /* mechanism classes: cls_tab rows           */
  int32_t clamp0, n_clamp; /* clamp table entries                       */
  int32_t probe0, n_probe; /* probe table entries                       */
} bc_pack;

/* Plain staggered run: n_steps steps of the staggered scheme for every pack.
 *
 * Arrays (global index ranges come from bc_pack):
 *  cv_info[cv]    slot | class << 16 (14 bits) | probe << 30 | clamp << 31
 *  bnd_info[i]    slot | clamp << 31
 *  Rows of a pack are numbered in chain-major "slots": chain k occupies
 *  slots chain_off[k] - row0 .. chain_off[k+1] - row0 - 1 (bottom to top),
 *  chains sorted by level.
 *  chain_off[k]   first slot of chain k as a global row index
 *                 (size n_chains_total + 1)
 *  chain_top[k]   local slot the chain's top hangs from, -1 for a root
 *  kid_off[k]     kids of chain k: kids[kid_off[k] .. kid_off[k+1]-1]
 *                 (local chain indices; size n_chains_total + 1)
 *  level_off[.]   per pack n_lv + 1 local chain offsets of the chain levels
 *                 (level = chain height; level 0 = the leaf chains, where
 *                 the kernel's dataflow elimination starts)
 *  lo/up[row]     dt * lowers, dt * uppers (A[row, parent], A[parent, row],
 *                 ms^-1) per global slot; the axial diagonal is rebuilt
 *                 from them (zero row sums of the axial operator)
 *  cls_tab[c*12+k] mechanism class table (dt/Cm, g per channel, Ca mask)
 *  clamp_row/on/off/amp[i]  local row, active step range [on, off), and
 *                 amplitude (uA/cm^2 on a CV row; mV rhs increment on a
 *                 boundary row)
 *  probe_row/col[i] local CV row and trace column
 *  V[cv], state[s * n_cv_total + cv]  voltage and gate/pool states (in/out)
 *  Vbnd[i]        voltage of boundary row i (bnd_info order; in/out)
 *  trace[step * trace_ld + col]  probe voltages after each step (out)
 *  rec            optional (NULL to skip): step-start snapshots for the
 *                 adjoint, rec[((step * (NS + 1)) + slot) * n_cv_total + cv],
 *                 slot 0 = V, slots 1..NS = states, steps 0 .. n_steps
 *                 (the last one is the final state)
 *  status         OR-ed with 1 when a non-finite voltage appears, with 2
 *                 when the chain tree of a pack does not resolve (an
 *                 inconsistent layout), with 4 when a pack's synapses
 *                 exceed bc_synapse_desc.max_syn / max_syn_rows (out)
 * real = float (precision 32) or double (precision 64). */
/* Synapses and delayed spike delivery for a network run (the synaptic step
 * either side of the cell update: braincell/network/engine.py:474-548,
 * delivery.py:317-410; braincell/synapse/markov.py:65-360).  At the start of
 * step n a synapse receives drive D = sum over its incoming contacts c of
 * w_c * spike(pre_c, n - delay_c).  Every synapse has two states A, B and a
 * kind (par[2]):
 *  BC_SYN_EXP (ExpSyn / Exp2Syn): A, B in uS, g = B - A; A += fa D,
 *    B += fb D; the voltage solve sees g; then A -= ka A, B -= kb B
 *    (ka = dt exprel(-dt / tau_a) / tau_a).  ExpSyn: fa = 0.
 *  BC_SYN_MARKOV1 (AMPA / GABAa): B = g (dimensionless), the solve sees
 *    g_n; then g' = aT D (1 - g) - beta g by exponential Euler.
 *  BC_SYN_MARKOV2 (NMDA): A = x, B = g; the solve sees g_n; then
 *    g' = a1 x (1 - g) - b1 g and x' = a2T D (1 - x) - b2 x, both from the
 *    step-start states (ind_exp_euler, quad/_exp_euler.py:330-470).
 * The current g (e - V) enters the CV's row with conductance k g in the
 * kernel's units (k = dt / (Cm area) uS -> mS/cm^2 for BC_SYN_EXP,
 * dt g_max / Cm for the Markov kinds).  par per synapse: k, e, kind, then
 * BC_SYN_EXP: fa, fb, ka, kb, 0; BC_SYN_MARKOV1: aT, beta, 0, 0, 0;
 * BC_SYN_MARKOV2: a1, b1, a2T, b2, 0.  A cell spikes at step n when any of
 * its CVs crosses v_th upwards (V_n < v_th <= V_{n+1}); the kernel writes
 * that into the history ring spk.  Every launch must satisfy n_steps <=
 * min_delay (each step reads spikes of earlier launches only) and spk_len
 * >= max_delay + n_steps.
 *
 * spk = NULL: no delivery at all (a cell run on its own, whose synapses are
 * driven only by the external drive below): contacts are not read (con_*,
 * cv_cell may be NULL), no spikes are written and the launch length is not
 * bounded by min_delay.
 *
 * External drive (NetStim trains, mech/_point.py:252-311 evaluated by
 * _compute/layouts.py:906-941, and inputs bound with
 * Cell.bind_synapse_input, _multi_compartment/cell.py:403-422, 2426-2487;
 * both summed into the drive by Cell._prepare_runtime_synapse_inputs):
 * when ext_slot is not NULL, synapse s with ext_slot[s] >= 0 adds
 * ext_drive[(n - ext_step0) * n_ext + ext_slot[s]] to D at global step n,
 * for 0 <= n - ext_step0 < ext_steps (real, the payload units of the
 * synapse: uS for BC_SYN_EXP, dimensionless for the Markov kinds).
 *
 * Each launch first gathers D of every synapse for every step of the launch
 * into the caller's workspace drive[k * n_syn + s] (k = step - step0; one
 * wavefront per synapse, its lanes over the contacts in a fixed order and a
 * fixed butterfly sum, then the external drive), which needs only spikes of
 * earlier launches.  The staggered kernel then keeps each pack's synapse
 * states and parameters in LDS for the whole launch (12 reals per synapse
 * and 2 int32 per synapse row beside the pack's tree: max_syn and
 * max_syn_rows bound them); the pack's states go back to state at the end.
 * status is OR-ed with 4 when a pack holds more synapses than max_syn or
 * more synapse rows than max_syn_rows (its synapses are then skipped). */
#define BC_SYN_EXP 0
#define BC_SYN_MARKOV1 1
#define BC_SYN_MARKOV2 2
#define BC_SYN_NPAR 8
typedef struct bc_synapse_desc {
  int32_t n_cells;       /* cells of the layout (columns of spk) */
  int32_t spk_len;       /* steps kept in the spike history ring */
  int32_t min_delay;     /* smallest contact delay in steps (>= 1) */
  int32_t max_delay;     /* largest contact delay in steps */
  double v_th;           /* spike threshold (mV), every cell */
  const int32_t* pack_off;  /* [n_packs + 1] synapse-row range of each pack */
  const int32_t* row;       /* [n_synrow] pack-local DHS row of a synapse row */
  const int32_t* row_off;   /* [n_synrow + 1] synapses of each synapse row */
  const void* par;          /* [n_syn * BC_SYN_NPAR] see above */
  void* state;              /* [2 * n_syn] A, B */
  const int32_t* con_off;   /* [n_syn + 1] incoming contacts of each synapse */
  const int32_t* con_pre;   /* [n_con] presynaptic cell */
  const int32_t* con_delay; /* [n_con] delay in steps */
  const void* con_w;        /* [n_con] weight (uS; dimensionless for the Markov kinds) */
  const int32_t* cv_cell;   /* [n_cv_total] cell of each CV */
  uint8_t* spk;             /* [spk_len * n_cells] spike history ring, or NULL */
  int32_t n_ext;            /* columns of ext_drive */
  int32_t ext_step0;        /* global step of ext_drive's first row */
  int32_t ext_steps;        /* rows of ext_drive */
  const int32_t* ext_slot;  /* [n_syn] column of each synapse (-1: none), or NULL */
  const void* ext_drive;    /* [ext_steps * n_ext] real */
  int32_t n_syn;            /* synapses (the par / state / ext_slot length) */
  void* syn_trace;          /* [n_steps * 2 n_syn] real, or NULL: A, B of every synapse
                               after each step of the launch (step-relative rows) */
  void* drive;              /* [n_steps * n_syn] real workspace of a launch; may be NULL
                               only when spk and ext_slot are both NULL (no drive) */
  int32_t max_syn;          /* most synapses in one pack (>= 1 when n_syn > 0) */
  int32_t max_syn_rows;     /* most synapse rows in one pack (>= 1 when n_syn > 0) */
} bc_synapse_desc;

/* Spike output of a plain run (get_spike, braincell/_base.py:421-431, set
 * every step by Cell._update_dynamics, _multi_compartment/cell.py:2384-2386):
 * every trace column is a detector; step n spikes at a column when its
 * voltage crosses v_th upwards, V_n < v_th <= V_{n+1}.  After each launch a
 * scan over the launch's trace rows appends the global step index n of every
 * crossing: steps[col * cap + count[col]] while count[col] < cap; count[col]
 * always counts.  v_last[col] holds the column's voltage before the launch's
 * first step and is updated to its last row.  Needs a trace (trace_ld > 0). */
typedef struct bc_spike_desc {
  int32_t cap;     /* step slots per column in steps (>= 0) */
  double v_th;     /* threshold (mV) */
  void* v_last;    /* [trace_ld] real, in/out */
  int32_t* count;  /* [trace_ld] in/out */
  int32_t* steps;  /* [trace_ld * cap] out (NULL when cap = 0) */
} bc_spike_desc;

typedef struct bc_stagger_desc {
  int32_t precision;    /* 32 or 64 */
  uint32_t channels;    /* BC_CH_* set compiled into the kernel */
  int32_t n_packs;
  int32_t threads;      /* workgroup size (64, 128, 192 or 256) */
  int32_t cv_per_thread;/* register slots per thread: 1, 2, 4, 8 or 16 */
  int32_t max_rows, max_chains, max_cls; /* LDS sizing (max over packs) */
  int32_t max_levels;
  int64_t n_cv_total;
  const bc_pack* packs;
  const uint32_t* cv_info;
  const uint32_t* bnd_info;
  const int32_t* chain_off;
  const int32_t* chain_top;
  const int32_t* kid_off;
  const int32_t* kids;
  const int32_t* level_off;
  const void* lo;
  const void* up;
  const void* cls_tab;
  const int32_t* clamp_row;
  const int32_t* clamp_on;
  const int32_t* clamp_off;
  const void* clamp_amp;
  const int32_t* probe_row;
  const int32_t* probe_col;
  const void* chan_const; /* host pointer to bc_chan_const<real> (passed by value) */
  void* V;
  void* Vbnd;
  void* state;
  void* trace;
  int64_t trace_ld;
  void* rec;
  uint32_t* status;
  int32_t step0;   /* global index of the first step (clamp windows) */
  int32_t n_steps;
  double dt;       /* ms */
  uint32_t flags;  /* BC_STEP_* */
  const bc_synapse_desc* syn; /* NULL: no synapses */
  /* The chain tables above in the kernel's LDS form, precomputed per layout
   * so that a workgroup copies them with independent loads (all pack-local):
   *  chain_rec[4k..4k+3] per global chain k: j0 | j1 << 16 (slot range),
   *                 top & 0xFFFF | n_kids << 16, kid0 | kid1 << 16 (first
   *                 two kid chains), first kid entry of the pack's kid list
   *  chain_par[k]   parent chain, -1 for a root
   *  kid_pad[ch0 + i] the pack's kid list (entry i; kid_off order), padded
   *                 with zeros to the pack's n_ch entries
   *  slot_link[row] parent slot (a root's top slot: itself) | 0x8000 at
   *                 a chain's bottom slot
   *  slot_chain[row] chain of the slot */
  const int32_t* chain_rec;
  const int16_t* chain_par;
  const int16_t* kid_pad;
  const int16_t* slot_link;
  const int16_t* slot_chain;
  const bc_spike_desc* spk; /* NULL: no spike output */
  /* Optional (NULL: chain order): per global chain position, the pack-local
   * chain that the elimination's per-chain passes take at that position;
   * thread t of a pack takes positions t, t + threads, ...  A permutation of
   * each pack's chains that balances the rows the threads walk. */
  const int16_t* chain_work;
  /* Per CV (cv_info order) and per boundary row (bnd_info order): the row's
   * slot_link | the pack-local chain its chain hangs from << 16 (0xFFFF for
   * a root chain); precomputed so a launch reads them with independent
   * loads. */
  const uint32_t* cv_aux;
  const uint32_t* bnd_aux;
} bc_stagger_desc;

/* bc_stagger_desc.flags */
#define BC_STEP_SKIP_GATES (1u << 0) /* update V and the Ca pool only; gates keep
                                        their step-start values (a closure then
                                        applies the gate update) */
#define BC_STEP_VOLTAGE_ONLY (1u << 2) /* dhs_voltage (quad/_staggered.py:187-320):
                                         the voltage solve alone; gates and the
                                         Ca pool keep their values */
#define BC_STEP_BACKSUB_SWEEP (1u << 3) /* back substitution of the chain tree
                                        by a sequential top-down sweep over
                                        the chain levels in one wavefront,
                                        instead of recursive doubling (pointer
                                        jumping, the form of comp_backsub_raw,
                                        quad/_staggered.py:742-760); which is
                                        faster depends on the trees' shape
                                        (examples/profiling/bench_dhs_levels.py:
                                        122-185), the autotuner times both.
                                        Compiled for the FP32 kernels without
                                        synapses (BC_UNSUPPORTED otherwise) */
#define BC_STEP_INTEGRATION_ORDER (1u << 1) /* ion_channel_update_order
                                        "integration" (_multi_compartment/cell.py:
                                        2137-2155): every node updates from the
                                        old state, so AHP reads Ci_n instead of
                                        the updated Ci_{n+1} ("family") */

/* Library identity and diagnostics. */
int bc_version(void);
const char* bc_last_error(void);
/* Channel set of a precision supported by bc_stagger_run; returns 1/0. */
int bc_stagger_supported(int precision, uint32_t channels, int cv_per_thread);

/* Runs n_steps staggered steps for every pack (one launch).
 * Replaces, per step, braincell's staggered_step
 * (braincell/quad/_staggered.py:47-154): cache_ion_total_currents
 * (_multi_compartment/cell.py:2261-2280), dhs_voltage_step
 * (_staggered.py:187-320: _build_dhs_numeric_state :449-513, comp_triang_raw
 * :641-657, comp_backsub_raw :742-760) and _update_ion_channel_families
 * (cell.py:2157-2217) with ind_exp_euler_step (quad/_exp_euler.py:330-478). */
int bc_stagger_run(const bc_stagger_desc* desc, void* stream);

/* hipGraph of a multi-chunk run: captures n_chunks launches of
 * steps_per_chunk steps each (chunk c starts at desc->step0 + c *
 * steps_per_chunk and writes trace rows from c * steps_per_chunk).
 * Replaces the jitted scan of braincell's run loop
 * (_multi_compartment/run.py:108-134). */
int bc_stagger_graph_create(const bc_stagger_desc* desc, int32_t n_chunks, int32_t steps_per_chunk,
                            void* stream, void** graph_out);
int bc_graph_launch(void* graph, void* stream);
int bc_graph_destroy(void* graph);

/* Reverse pass over one checkpoint segment of n_steps steps (steps are
 * processed last to first).  layout: the forward descriptor (same arrays,
 * dt and channel set; its state pointers are not read).
 *  rec        forward recording of the segment (see bc_stagger_desc.rec)
 *  trace_bar  d loss / d trace for the segment's steps [n_steps * trace_ld]
 *  Vbar, Sbar adjoints of V and of the states at the segment end (in) and
 *             at its start (out), laid out like V / state
 *  cls_grad   += d loss / d cls_tab (per mechanism class; columns of the
 *             channel densities, mS/cm^2)
 * Channel sets: IL; Na+KDR+IL; Na+KDR+IL+CaL+pool.  No counterpart in the
 * reference, which differentiates its JAX scan automatically; this is the
 * hand-written reverse of _staggered.py:126-154. */
typedef struct bc_adjoint_desc {
  const bc_stagger_desc* layout;
  int32_t n_steps;
  const void* rec;
  const void* trace_bar;
  void* Vbar;
  void* Sbar;
  void* cls_grad;
} bc_adjoint_desc;

int bc_adjoint_supported(int precision, uint32_t channels, int cv_per_thread);
int bc_stagger_adjoint(const bc_adjoint_desc* desc, void* stream);

/* Point-neuron population (thalamocortical template: Na_Ba2002, K_Leak,
 * KDR_Ba2002, CaL_IS2008, CaT_HM1992, AHP_De1994, HCN_HM1992, IL and a
 * CalciumDetailed pool) integrated by independent exponential Euler.
 *  state[k * n + i]  k = Na p, Na q, KDR p, CaL p, CaL q, CaT p, CaT q,
 *                    AHP p, HCN p, Ci
 *  g[k * n + i]      k = Na, K_Leak, KDR, CaL, CaT, AHP, HCN, IL (mS/cm^2)
 *  I_ext[i]          injected current density (uA/cm^2)
 *  trace[step * n_trace + i]  V of neurons 0 .. n_trace-1 after each step
 *  spikes[i]         += upward crossings of V_th
 *  method            BC_POINT_IND_EXP_EULER (0): ind_exp_euler, each state
 *                    y_k + dt exprel(dt lam_k) f_k (braincell/quad/
 *                    _exp_euler.py:217-320);
 *                    BC_POINT_BACKWARD_EULER (1): (I - dt J) dy = dt f with
 *                    the full 11-state Jacobian J at the step start
 *                    (braincell/quad/_backward_euler.py:31-64);
 *                    BC_POINT_IMPLICIT_EULER (2): (I - dt/2 J) dy = dt f, the
 *                    Newton iterate of implicit_euler's order-2 residual
 *                    started from y_n (braincell/quad/_implicit.py:43-120) */
#define BC_POINT_IND_EXP_EULER 0
#define BC_POINT_BACKWARD_EULER 1
#define BC_POINT_IMPLICIT_EULER 2
typedef struct bc_point_desc {
  int32_t precision;
  int64_t n;
  int32_t n_steps;
  double dt, Cm, V_th;
  const void* chan_const; /* host pointer to bc_chan_const<real> */
  void* V;
  void* state;
  const void* g;
  const void* I_ext;
  void* trace;
  int64_t n_trace;
  uint32_t* spikes;
  uint32_t* status;
  int32_t method;
} bc_point_desc;

int bc_point_supported(int precision);
/* Replaces SingleCompartment.update (braincell/_single_compartment/
 * base.py:252-278) with solver ind_exp_euler (braincell/quad/
 * _exp_euler.py:217-320), backward_euler (quad/_backward_euler.py:71-160)
 * or implicit_euler (quad/_implicit.py:313-368), per desc->method, for
 * n_steps steps. */
int bc_point_run(const bc_point_desc* desc, void* stream);

/* ------------------------------------------------------------------------
 * Linearised backward Euler on passive multi-compartment cells.  Replaces
 * backward_euler_step (braincell/quad/_backward_euler.py:71-160) with a
 * Cell target: apply_standard_solver_step (quad/_util.py:196-262) stacks
 * the cell's differential states, which only works when V (CV space) is
 * the only one (a gated channel's states live in point space and cannot be
 * stacked with V); f(V) = (I_mem(V) + I_clamp) / Cm - K V
 * (Cell.compute_voltage_derivative, _multi_compartment/cell.py:1892-1999,
 * K = build_cv_axial_operator, quad/_staggered.py:393-413) and
 * jacrev_last_dim (quad/_util.py:265-310) gives per CV the column sum
 * J_i = -a_i - sum_q K_qi.  Per step n, every CV i (global index):
 *   f_i = b_i - a_i V_i + sum of clamp_amp over the cell's clamps on i with
 *         clamp_on <= n < clamp_end - sum_e k_val[e] V[cell_cv0 + k_col[e]]
 *   V_i += dt f_i / (1 - dt jac_i)
 * with a_i = sum g / Cm (1/ms), b_i = sum g E / Cm (mV/ms), clamp_amp in
 * mV/ms (clamps on CV boundary points are not in the CV-space derivative,
 * _multi_compartment/currents.py:166-193).  k_col, clamp_cv and probe_cv
 * are cell-local CV indices; the trace gets V after each step at the
 * probes: trace[step * trace_ld + probe_col].  One workgroup per cell holds
 * its voltages in LDS (2 * max_cv reals, at most 160 KiB).
 * real = float (precision 32) or double (precision 64). */
typedef struct bc_cable_be_desc {
  int32_t precision;
  int32_t n_cells;
  int32_t threads;          /* workgroup size, multiple of 64 in [64, 1024] */
  int32_t max_cv;           /* CVs of the largest cell */
  const int32_t* cell_cv0;  /* [n_cells + 1] CV range of each cell */
  const int32_t* k_off;     /* [n_cv + 1] CSR rows of K */
  const int32_t* k_col;     /* [nnz] */
  const void* k_val;        /* [nnz] real, 1/ms */
  const void* a;            /* [n_cv] real */
  const void* b;            /* [n_cv] real */
  const void* jac;          /* [n_cv] real */
  const int32_t* clamp_off; /* [n_cells + 1] clamps of each cell */
  const int32_t* clamp_cv;  /* [n_clamp] */
  const int32_t* clamp_on;  /* [n_clamp] first active global step */
  const int32_t* clamp_end; /* [n_clamp] first inactive global step */
  const void* clamp_amp;    /* [n_clamp] real */
  const int32_t* probe_off; /* [n_cells + 1] probes of each cell */
  const int32_t* probe_cv;  /* [n_probe] */
  const int32_t* probe_col; /* [n_probe] */
  int32_t trace_ld;         /* trace columns (0: no trace) */
  void* trace;              /* [n_steps * trace_ld] real, out */
  void* V;                  /* [n_cv] real, in/out */
  uint32_t* status;         /* OR-ed with 1 on a non-finite voltage, 2 on a
                               cell larger than max_cv / max_nnz */
  int32_t step0;            /* global index of the first step */
  int32_t n_steps;
  double dt;                /* ms */
  int32_t max_nnz;          /* entries of K in the largest cell's rows (each
                               cell's rows, a, b and J are held in LDS for
                               the launch; a cell beyond max_cv / max_nnz
                               sets status bit 2 and is not stepped) */
} bc_cable_be_desc;
int bc_cable_be_run(const bc_cable_be_desc* desc, void* stream);

/* ------------------------------------------------------------------------
 * Learned gating closure (no counterpart in the reference): an MLP
 * 8 -> 64 -> 64 -> 5 (tanh) that corrects the exponential-Euler update of
 * the 5 gates of the Na/KDR/IL/CaL/pool model at large dt: with delta the
 * plain increment and c the network output (times out_scale), a gate moves
 * by delta (1 + c).
 * Features: (V_{n+1}, Na p, Na q, KDR p, CaL p, CaL q at step start,
 * ln Ci_{n+1}, dt); parameters: W1[8][64] b1[64] W2[64][64] b2[64] W3[64][5]
 * b3[5] (FP32, BC_CLOSURE_NPARAM values); norm = mu[8], sd[8],
 * out_scale[5] (features are (f - mu) / sd, c = out * out_scale).
 * All arrays FP32 device memory. */
#define BC_CLOSURE_NPARAM 5061

/* Training pairs from a fine trajectory: for n compartments, V1 = V and
 * S1 = states at t + dt, S0 = states at t (state arrays [6][ld]: Na p, Na q,
 * KDR p, CaL p, CaL q, Ci).  X[n][8] = raw features, D[n][5] = the plain
 * increment of S0's gates at V1, Y[n][5] = S1 gates minus that plain update. */
typedef struct bc_closure_data_desc {
  int64_t n, ld;
  double dt;
  const void* chan_const; /* host pointer to bc_chan_const<float> */
  const void* V1;
  const void* S0;
  const void* S1;
  void* X;
  void* D;
  void* Y;
} bc_closure_data_desc;
int bc_closure_data(const bc_closure_data_desc* desc, void* stream);

/* Gradient of the minibatch loss
 * mean_b sum_o (Dn[i][o] net(X[i])_o - Yn[i][o])^2, i = idx[b] (X normalised,
 * Dn and Yn = D and Y divided by per-output scales) with respect to P: FP32 forward
 * and backward, per-128-sample partial sums reduced in fixed order (the
 * result is deterministic).  partial: [max_blocks][BC_CLOSURE_NPARAM],
 * loss_part: [max_blocks], grad: [BC_CLOSURE_NPARAM], loss: [1]. */
typedef struct bc_closure_train_desc {
  int64_t batch;
  int32_t max_blocks;
  const void* X;
  const void* Dn;
  const void* Yn;
  const int64_t* idx;
  const void* P;
  void* partial;
  void* loss_part;
  void* grad;
  void* loss;
} bc_closure_train_desc;
int bc_closure_grad(const bc_closure_train_desc* desc, void* stream);

/* bc_closure_grad without caller workspace: the same kernels and result
 * (bitwise), with the per-block partial sums in a fixed device array of the
 * library instead of desc->partial / desc->loss_part (ignored, like
 * max_blocks).  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under hipGraph capture, serving loops,
 * harnesses that forbid hipMalloc); it allocates nothing.  Batches are
 * limited to BC_CLOSURE_NOALLOC_MAX_BATCH samples (the array holds
 * BC_CLOSURE_NOALLOC_MAX_BATCH / 128 blocks x (BC_CLOSURE_NPARAM + 1) floats,
 * 5.2 MB, on every GPU the library runs on, until the process ends); larger
 * batches return BC_INVALID_ARG.  All calls on one GPU share the array:
 * calls that overlap on one GPU (different streams or threads) race. */
#define BC_CLOSURE_NOALLOC_MAX_BATCH 32768
int bc_closure_grad_noalloc(const bc_closure_train_desc* desc, void* stream);

/* One Adam step (bias-corrected, step counts from 1). */
typedef struct bc_closure_adam_desc {
  int64_t step;
  double lr, beta1, beta2, eps;
  void* P;
  void* m;
  void* v;
  const void* grad;
} bc_closure_adam_desc;
int bc_closure_adam(const bc_closure_adam_desc* desc, void* stream);

/* Network outputs out[n][5] = c = net((X - mu) / sd) * out_scale from raw
 * features X[n][8]: mfma = 0 FP32, mfma = 1 BF16 MFMA with FP32 accumulation. */
typedef struct bc_closure_fwd_desc {
  int64_t n;
  int32_t mfma;
  const void* X;
  const void* P;
  const void* norm;
  void* out;
} bc_closure_fwd_desc;
int bc_closure_forward(const bc_closure_fwd_desc* desc, void* stream);

/* Gate update with the closure, in place, after a staggered step of length
 * dt run with BC_STEP_SKIP_GATES: V1 = the new voltage, S = the state
 * ([6][ld]; gates still at their step-start values, Ci updated).  Each gate
 * x becomes clip(x + delta (1 + c), 0, 1): delta = its plain exp-Euler
 * increment at V1, c = the BF16-MFMA network output * out_scale.
 * chan_const: host pointer to bc_chan_const<float>.  flags (optional, uint8
 * [n]): set to 1 for a compartment whose corrected gate, before the clip,
 * lies outside [-flag_tol, 1 + flag_tol] or is not finite (the network
 * over-corrects there); never cleared by the kernel. */
typedef struct bc_closure_apply_desc {
  int64_t n, ld;
  double dt;
  const void* chan_const;
  const void* V1;
  void* S;
  const void* P;
  const void* norm;
  void* flags;
  double flag_tol;
} bc_closure_apply_desc;
int bc_closure_apply(const bc_closure_apply_desc* desc, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BRAINCELL_AMD_H */
