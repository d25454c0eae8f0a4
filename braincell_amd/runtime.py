"""Packed multi-cell runtime for the staggered kernel.

``StaggerLayout`` lowers a list of :class:`~braincell_amd.model.CellModel`
into packs (whole cells grouped per workgroup) and flat arrays;
``StaggerRuntime`` uploads them as torch tensors on the GPU and drives the
C ABI (``bc_stagger_run`` / ``bc_stagger_graph_create``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import rates
from ._native import BcPack, SpikeDesc, StaggerDesc, chan_const_type, check, lib
from .channels import CA_POOL_BIT, CHANNELS, FARADAY, GAS_CONSTANT
from .layout import build_chains_batch, discretize_cell

__all__ = ["StaggerLayout", "StaggerRuntime", "KERNEL_CHANNEL_SETS", "choose_channel_set", "first_step_at"]

CH_ORDER = ("Na_Ba2002", "KDR_Ba2002", "IL", "K_Leak", "CaL_IS2008", "CaT_HM1992", "AHP_De1994", "HCN_HM1992")
CLS_COL = {n: i + 1 for i, n in enumerate(CH_ORDER)}  # column in the class table
# the Hodgkin-Huxley 1952 pair occupies the Na / KDR columns, state slots and
# constants: the compiled sets never hold both pairs
CLS_COL.update({"Na_HH1952": CLS_COL["Na_Ba2002"], "K_HH1952": CLS_COL["KDR_Ba2002"]})
NCLS = 12
B = {n: CHANNELS[n].bit for n in CH_ORDER}
B_HH = {n: CHANNELS[n].bit for n in ("Na_HH1952", "K_HH1952")}
KERNEL_CHANNEL_SETS = (
    B["IL"],
    B["Na_Ba2002"] | B["KDR_Ba2002"] | B["IL"],
    B["Na_Ba2002"] | B["KDR_Ba2002"] | B["IL"] | B["CaL_IS2008"] | CA_POOL_BIT,
    sum(B.values()) | CA_POOL_BIT,
    B_HH["Na_HH1952"] | B_HH["K_HH1952"] | B["IL"],
)
INT32_MAX = 2 ** 31 - 1
REGION_CODES = ("soma", "basal_dendrite", "apical_dendrite", "dendrite", "axon", "custom")


def choose_channel_set(mask: int) -> int:
    for s in KERNEL_CHANNEL_SETS:
        if mask & ~s == 0:
            return s
    raise ValueError(f"no kernel covers channel mask 0x{mask:x}")


def state_slots(chset: int) -> list:
    """(channel, gate index) per state slot, matching ``bc_slot_*``."""
    out = []
    if chset & B["Na_Ba2002"]:
        out += [("Na_Ba2002", 0), ("Na_Ba2002", 1)]
    if chset & B_HH["Na_HH1952"]:
        out += [("Na_HH1952", 0), ("Na_HH1952", 1)]
    if chset & B["KDR_Ba2002"]:
        out += [("KDR_Ba2002", 0)]
    if chset & B_HH["K_HH1952"]:
        out += [("K_HH1952", 0)]
    if chset & B["CaL_IS2008"]:
        out += [("CaL_IS2008", 0), ("CaL_IS2008", 1)]
    if chset & B["CaT_HM1992"]:
        out += [("CaT_HM1992", 0), ("CaT_HM1992", 1)]
    if chset & B["AHP_De1994"]:
        out += [("AHP_De1994", 0)]
    if chset & B["HCN_HM1992"]:
        out += [("HCN_HM1992", 0)]
    if chset & CA_POOL_BIT:
        out += [("Ca", 0)]
    return out


def balanced_work_order(lens: np.ndarray, threads: int, n_leaf: int = None) -> np.ndarray:
    """Order in which the threads of a pack take chains in the per-chain
    passes, balanced so that the threads walk similar numbers of rows.
    With ``n_leaf`` (the pack's leaf chains come first, chains 0 ..
    n_leaf - 1) and more than one wavefront, positions 0 .. n_leaf - 1 hold
    the leaf chains for the threads after the first wavefront (which runs
    the level sweep meanwhile) and the other chains follow for all threads;
    otherwise every chain is ordered for all threads."""
    lens = np.asarray(lens, dtype=np.int64)
    if n_leaf is not None and threads > 64:
        leaves = _snake_order(lens[:n_leaf], threads - 64)
        rest = _snake_order(lens[n_leaf:], threads) + n_leaf
        return np.concatenate([leaves, rest]).astype(np.int16)
    return _snake_order(lens, threads).astype(np.int16)


def _snake_order(lens: np.ndarray, threads: int) -> np.ndarray:
    """Thread t takes positions t, t + threads, ...: the longest chains go
    one per thread, and every further round runs in the opposite direction
    (snake), so the threads with the longest first chains get the shortest
    later ones."""
    order = np.argsort(-lens, kind="stable")
    out = np.empty(lens.size, dtype=np.int64)
    n = lens.size
    for r0 in range(0, n, threads):
        m = min(threads, n - r0)
        chunk = order[r0:r0 + m]
        if (r0 // threads) % 2 == 1:
            # odd rounds: the last (shortest) chain to thread 0
            chunk = chunk[::-1]
        out[r0:r0 + m] = chunk
    return out


def first_step_at(t: float, dt: float) -> int:
    """Smallest n >= 0 with n * dt >= t in float64 (t_n = n dt, ``run.py:64-131``)."""
    if t <= 0.0:
        return 0
    if t / dt > INT32_MAX:
        return INT32_MAX
    n = max(0, int(np.ceil(t / dt)) - 1)
    while n * dt < t:
        n += 1
    while n > 0 and (n - 1) * dt >= t:
        n -= 1
    return n


def _phi(q10, temp, temp_ref):
    return q10 ** ((temp - temp_ref) / 10.0)


@dataclass
class _CellHost:
    geom: object
    chains: object
    cls_rows: np.ndarray  # [n_cls, NCLS] with column 0 = Cm (dt applied later)
    cv_cls: np.ndarray  # [C]
    clamps: list  # (kind, local index (cv or row), t0, t1, amp)
    probes: np.ndarray  # CV indices
    synapses: list = None  # (instance name, mech.Synapse, [(cv, area cm^2, Cm uF/cm^2) per site])
    netstims: list = None  # (mech.NetStim, CV of its point, or -1 on a CV boundary node)


BACKSUB_SWEEP = 8  # BC_STEP_BACKSUB_SWEEP
BACKSUB_KINDS = ("jump", "sweep")


class StaggerLayout:
    """Host-side packed layout of a batch of cells."""

    # back substitution of the chain tree: "jump" = recursive doubling
    # (pointer jumping), "sweep" = a sequential top-down sweep over the chain
    # levels (BC_STEP_BACKSUB_SWEEP); chosen per batch by the autotuner
    backsub = "jump"

    def __init__(self, models, *, threads: int = 256, cv_per_thread: int = None, rows_per_pack: int = None,
                 probes="soma", region_classes: bool = False, probes_per_model=None, max_chain_rows: int = 16):
        if len(models) == 0:
            raise ValueError("need at least one cell")
        self.models = list(models)
        # chains longer than this are cut into pieces (bounds the per-chain
        # passes of the elimination and back substitution)
        self.max_chain_rows = int(max_chain_rows) if max_chain_rows else None
        self.region_classes = bool(region_classes)
        mask = 0
        for m in self.models:
            m.validate()
            for n in m.channels:
                mask |= CHANNELS[n].bit
            if m.ca is not None:
                mask |= CA_POOL_BIT
        self.channel_mask = mask
        self.chset = choose_channel_set(mask)
        self.slots = state_slots(self.chset)
        self.cc = self._channel_constants()
        self._geom_cache = {}
        if probes_per_model is not None:  # one probe spec per model
            if len(probes_per_model) != len(self.models):
                raise ValueError("probes_per_model needs one entry per model")
            self.cells = [self._lower_cell(m, pr) for m, pr in zip(self.models, probes_per_model)]
        else:
            self.cells = [self._lower_cell(m, probes) for m in self.models]
        del self._geom_cache
        self._chain_cells()
        self._configure(threads, cv_per_thread, rows_per_pack)

    def _chain_cells(self):
        """Chains of every cell (cells sharing a discretisation share them)."""
        uniq = {}
        for c in self.cells:
            uniq.setdefault(id(c.geom), c.geom)
        geoms = list(uniq.values())
        chains = build_chains_batch([g.parent_rows for g in geoms], self.max_chain_rows)
        by_geom = {id(g): ch for g, ch in zip(geoms, chains)}
        for c in self.cells:
            c.chains = by_geom[id(c.geom)]

    def rechained(self, max_chain_rows: int = None, **plan):
        """The same lowered cells with chains cut at ``max_chain_rows`` rows
        (and optionally another pack plan: threads, cv_per_thread,
        rows_per_pack)."""
        import copy

        L = copy.copy(self)
        L.max_chain_rows = int(max_chain_rows) if max_chain_rows else None
        L.cells = [copy.copy(c) for c in self.cells]
        L._chain_cells()
        L._configure(plan.get("threads", self.threads), plan.get("cv_per_thread", self.cv_per_thread),
                     plan.get("rows_per_pack", self.rows_per_pack))
        return L

    def _configure(self, threads, cv_per_thread, rows_per_pack):
        """Workgroup size, register slots and pack plan over the lowered cells."""
        self.threads = int(threads)
        if self.threads not in (64, 128, 192, 256):
            raise ValueError("threads must be 64, 128, 192 or 256")
        ncv_max = max(c.geom.n_cv for c in self.cells)
        if cv_per_thread is None:
            cv_per_thread = 1
            while cv_per_thread * self.threads < ncv_max:
                cv_per_thread *= 2
        self.cv_per_thread = int(cv_per_thread)
        if self.cv_per_thread not in (1, 2, 4, 8, 16):
            raise ValueError("cv_per_thread must be 1, 2, 4, 8 or 16")
        cap_cv = self.threads * self.cv_per_thread
        if ncv_max > cap_cv:
            raise ValueError(f"a cell has {ncv_max} CVs > threads*cv_per_thread = {cap_cv}")
        nrow_max = max(c.geom.n_point for c in self.cells)
        self.rows_per_pack = int(max(nrow_max, rows_per_pack or nrow_max))
        self.pack_cells = self._plan_packs(cap_cv)
        self._assemble()

    def repacked(self, *, threads: int = None, cv_per_thread: int = None, rows_per_pack: int = None):
        """The same lowered cells under another workgroup size / pack plan
        (the cells are not discretised again)."""
        import copy

        L = copy.copy(self)
        L._configure(self.threads if threads is None else threads, cv_per_thread, rows_per_pack)
        return L

    def with_backsub(self, kind: str):
        """The same layout with the back-substitution variant ``kind``
        ("jump" or "sweep")."""
        import copy

        if kind not in BACKSUB_KINDS:
            raise ValueError(f"backsub must be one of {BACKSUB_KINDS}, got {kind!r}")
        L = copy.copy(self)
        L.backsub = kind
        return L

    @property
    def step_flags(self) -> int:
        """BC_STEP_* bits the layout adds to every launch."""
        return BACKSUB_SWEEP if self.backsub == "sweep" else 0

    def lds_bytes(self, precision: int = 32) -> int:
        """Dynamic LDS of one workgroup (the kernel's lds_plan)."""
        rs = 8 if precision == 64 else 4
        mr = (self.max_rows + 3) & ~3
        mc = (self.max_chains + 3) & ~3
        rows = (8 + mr + 8) * 4 * rs
        # rows, chain records, chain contributions, chain composites (shared
        # with the pointer-jumping buffers), V per slot, class table, kid
        # list, work order, chain levels, level offsets, dataflow words,
        # queue words
        return int(rows + mc * 16 + mc * 2 * rs + mc * 8 * rs + mr * rs + self.max_cls * 12 * rs
                   + mc * 2 + mc * 2 + mc * 2 + ((self.max_levels + 4) & ~3) * 2 + mc * 4 + 16)

    # ------------------------------------------------------------ constants
    def _channel_constants(self):
        ref = self.models[0]
        params = {}
        for m in self.models:
            for n in m.channels:
                p = m.channel_params(n)
                p = {k: v for k, v in p.items() if k != "g_max"}
                if n in params and params[n] != p:
                    raise ValueError(f"{n}: channel parameters differ between cells of one batch")
                params[n] = p
            if (m.E_Na, m.E_K, m.E_Ca) != (ref.E_Na, ref.E_K, ref.E_Ca):
                raise ValueError("fixed reversal potentials differ between cells of one batch")
        pools = [m.ca for m in self.models if m.ca is not None]
        for ca in pools:
            if any(float(ca[k]) != float(pools[0][k]) for k in ("d", "tau", "C_rest", "Co", "temp", "Ci0")):
                raise ValueError("Ca pool parameters differ between cells of one batch")
        if pools:
            for m in self.models:
                if m.ca is None and any(CHANNELS[n].ion in ("Ca", "KCa") for n in m.channels):
                    raise ValueError("a cell with Ca channels but no Ca pool cannot share a batch with pool cells")
        cc = {k: 0.0 for k in ("E_Na", "E_K", "E_IL", "na_vsh", "na_phi", "kdr_vsh", "kdr_phi", "cal_vsh",
                               "cal_phi_p", "cal_phi_q", "cat_vsh", "cat_phi_p", "cat_phi_q", "ahp_alpha",
                               "ahp_beta", "ahp_n", "ahp_phi", "hcn_E", "hcn_phi", "ca_drive", "ca_tau_inv",
                               "ca_rest", "ca_Co", "ca_rtzf", "E_Ca_fixed")}
        cc["E_Na"], cc["E_K"], cc["E_Ca_fixed"] = ref.E_Na, ref.E_K, ref.E_Ca
        names = CH_ORDER + tuple(B_HH)
        dflt = {n: dict(CHANNELS[n].params) for n in names}
        for n in names:
            dflt[n].update(params.get(n, {}))
        hh = "Na_HH1952" in params or "K_HH1952" in params
        if hh and ("Na_Ba2002" in params or "KDR_Ba2002" in params):
            raise ValueError("Na_HH1952/K_HH1952 and Na_Ba2002/KDR_Ba2002 cannot share a batch")
        na, kd, cl, ct, ah, hc = (dflt[n] for n in (("Na_HH1952", "K_HH1952") if hh else ("Na_Ba2002", "KDR_Ba2002"))
                                  + ("CaL_IS2008", "CaT_HM1992", "AHP_De1994", "HCN_HM1992"))
        cc["E_IL"] = dflt["IL"]["E"]
        cc["na_vsh"], cc["na_phi"] = na["V_sh"], _phi(na["q10"], na["temp"], na["temp_ref"])
        cc["kdr_vsh"], cc["kdr_phi"] = kd["V_sh"], _phi(kd["q10"], kd["temp"], kd["temp_ref"])
        cc["cal_vsh"] = cl["V_sh"]
        cc["cal_phi_p"] = _phi(cl["q10_p"], cl["temp"], cl["temp_ref_p"])
        cc["cal_phi_q"] = _phi(cl["q10_q"], cl["temp"], cl["temp_ref_q"])
        cc["cat_vsh"] = ct["V_sh"]
        cc["cat_phi_p"] = _phi(ct["q10_p"], ct["temp"], ct["temp_ref_p"])
        cc["cat_phi_q"] = _phi(ct["q10_q"], ct["temp"], ct["temp_ref_q"])
        cc["ahp_alpha"], cc["ahp_beta"], cc["ahp_n"], cc["ahp_phi"] = ah["alpha"], ah["beta"], ah["n"], ah["phi"]
        cc["hcn_E"], cc["hcn_phi"] = hc["E"], _phi(hc["q10"], hc["temp"], hc["temp_ref"])
        if pools:
            ca = pools[0]
            cc["ca_drive"] = 10.0 / (2.0 * FARADAY * float(ca["d"]))
            cc["ca_tau_inv"] = 1.0 / float(ca["tau"])
            cc["ca_rest"] = float(ca["C_rest"])
            cc["ca_Co"] = float(ca["Co"])
            cc["ca_rtzf"] = 1e3 * GAS_CONSTANT * float(ca["temp"]) / (2.0 * FARADAY)
        return cc

    # ---------------------------------------------------------------- cells
    def _lower_cell(self, m, probes) -> _CellHost:
        ncvb = m.n_cv_branch()
        cmb, rab = m.per_branch(m.cm), m.per_branch(m.ra)
        # cells that share one morphology object and cable constants share
        # their discretisation (read-only from here on)
        key = (id(m.morph), ncvb.tobytes(), np.asarray(cmb, np.float64).tobytes(), np.asarray(rab, np.float64).tobytes())
        hit = self._geom_cache.get(key)
        if hit is not None and hit[0] is m.morph:
            geom = hit[1]
        else:
            geom = discretize_cell(m.morph, ncvb, cmb, rab)
            self._geom_cache[key] = (m.morph, geom)
        chains = None  # built for all cells at once (_chain_cells)
        nb = m.n_branches
        table = np.zeros((nb, NCLS))
        table[:, 0] = m.per_branch(m.cm)
        for n, spec in m.channels.items():
            table[:, CLS_COL[n]] = np.asarray(spec["g_branch"], dtype=np.float64) * 1e3  # mS/cm^2
        if m.ca is not None:
            table[:, 9] = np.asarray(m.ca["mask_branch"], dtype=np.float64)
        if self.region_classes:  # keep painted regions apart (one class per branch type)
            table[:, 10] = [REGION_CODES.index(b.type) for b in m.morph.branches]
        uniq, inv = np.unique(table, axis=0, return_inverse=True)
        cv_cls = inv.reshape(-1)[geom.cv_branch]
        clamps = []
        for (b, x, t0, dur, amp) in m.clamps:
            kind, idx = geom.locate(int(b), float(x), ncvb)
            clamps.append((kind, idx, float(t0), float(t0) + float(dur), float(amp)))
        if isinstance(probes, str):
            if probes == "soma":
                kind, idx = geom.locate(0, 0.5, ncvb)
                pr = np.asarray([idx if kind == "mid" else 0])
            elif probes == "all":
                pr = np.arange(geom.n_cv)
            elif probes == "none":
                pr = np.zeros(0, dtype=np.int64)
            else:
                raise ValueError(f"unknown probe spec {probes!r}")
        else:
            pr = np.asarray(probes, dtype=np.int64)
        syns = []
        cmb = m.per_branch(m.cm)
        for (name, syn, sites) in getattr(m, "synapses", []) or []:
            loc = []
            for (b, x) in sites:
                kind, idx = geom.locate(int(b), float(x), ncvb)
                if kind != "mid":
                    raise ValueError(f"synapse {name!r} at branch {b}, x = {x} sits on a CV boundary node; "
                                     "place synapses inside a CV")
                loc.append((int(idx), float(geom.cv_area_um2[idx]) * 1e-8, float(cmb[geom.cv_branch[idx]])))
            syns.append((name, syn, loc))
        stims = []
        for (ns, sites) in getattr(m, "netstims", []) or []:
            for (b, x) in sites:
                kind, idx = geom.locate(int(b), float(x), ncvb)
                stims.append((ns, int(idx) if kind == "mid" else -1))
        return _CellHost(geom, chains, uniq, cv_cls, clamps, pr, syns, stims)

    def _plan_packs(self, cap_cv):
        """First-fit decreasing by row count, under the row and CV budgets."""
        order = np.argsort([-c.geom.n_point for c in self.cells], kind="stable")
        packs, rows, cvs = [], [], []
        for i in order.tolist():
            r, c = self.cells[i].geom.n_point, self.cells[i].geom.n_cv
            for p in range(len(packs)):
                if rows[p] + r <= self.rows_per_pack and cvs[p] + c <= cap_cv:
                    packs[p].append(i)
                    rows[p] += r
                    cvs[p] += c
                    break
            else:
                packs.append([i])
                rows.append(r)
                cvs.append(c)
        return packs

    # ------------------------------------------------------------- assemble
    def _assemble(self):
        """Flatten the packs.  Inside a pack, rows are renumbered into
        chain-major *slots*: chain k occupies slots chain_off[k] ..
        chain_off[k+1]-1 bottom to top, chains sorted by level, so the kernel
        walks contiguous LDS.  ``slot_rows`` keeps, per pack, the reference
        row (cell row offset + ``row_to_point_id`` order) of every slot."""
        n_cells = len(self.cells)
        packs = []
        cv_info, bnd_info, bnd_cell = [], [], []
        chain_off, chain_top, kid_cnt, kids, level_off = [0], [], [], [], []
        slot_lo, slot_up, slot_rows_all = [], [], []
        chain_rec_all, chain_par_all, kid_pad_all, slot_link_all, slot_chain_all = [], [], [], [], []
        chain_work_all = []
        cv_aux, bnd_aux = [], []
        cls_tab, clamp_rows, clamp_t, clamp_amp_raw, clamp_kind, clamp_area = [], [], [], [], [], []
        probe_row, probe_col = [], []
        self.cell_region_cls = np.full((n_cells, len(REGION_CODES)), -1, dtype=np.int64)
        self.cell_cv0 = np.zeros(n_cells, dtype=np.int64)
        self.cell_ncv = np.zeros(n_cells, dtype=np.int64)
        self.cell_probe0 = np.zeros(n_cells + 1, dtype=np.int64)
        self.cell_pack = np.zeros(n_cells, dtype=np.int64)
        self.cell_row_off = np.zeros(n_cells, dtype=np.int64)
        self.cell_probe0[1:] = np.cumsum([len(c.probes) for c in self.cells])
        gcv = grow = gch = glv = gbnd = gcls = gclamp = gprobe = 0
        for pi, cells in enumerate(self.pack_cells):
            P = BcPack()
            P.cv0, P.row0, P.ch0, P.lv0, P.bnd0, P.cls0, P.clamp0, P.probe0 = gcv, grow, gch, glv, gbnd, gcls, gclamp, gprobe
            cs = [self.cells[i] for i in cells]
            nrow = np.asarray([c.geom.n_point for c in cs])
            nch = np.asarray([c.chains.n_chains for c in cs])
            row_off = np.concatenate(([0], np.cumsum(nrow)))
            ch_base = np.concatenate(([0], np.cumsum(nch)))
            n_rows, n_ch = int(row_off[-1]), int(ch_base[-1])
            lev = np.concatenate([c.chains.level for c in cs])
            cpos = np.repeat(np.arange(len(cs)), nch)
            kidx = np.concatenate([np.arange(n) for n in nch])
            perm = np.lexsort((kidx, cpos, lev))
            inv = np.empty(n_ch, dtype=np.int64)
            inv[perm] = np.arange(n_ch)
            lens_old = np.concatenate([np.diff(c.chains.off) for c in cs])
            seg_old = np.concatenate([c.chains.off[:-1] + np.sum(nrow[:j]) for j, c in enumerate(cs)])
            allrows = np.concatenate([c.chains.rows + row_off[j] for j, c in enumerate(cs)])
            lens = lens_old[perm]
            new_off = np.concatenate(([0], np.cumsum(lens)))
            idx = np.repeat(seg_old[perm] - new_off[:-1], lens) + np.arange(n_rows)
            slot_rows = allrows[idx]  # pack-local reference row of each slot
            slot_of = np.empty(n_rows, dtype=np.int64)
            slot_of[slot_rows] = np.arange(n_rows)
            lowers = np.concatenate([c.geom.lowers for c in cs])
            uppers = np.concatenate([c.geom.uppers for c in cs])
            slot_lo.append(lowers[slot_rows])
            slot_up.append(uppers[slot_rows])
            slot_rows_all.append(slot_rows)
            tp_old = np.concatenate([np.where(c.chains.top_parent >= 0, c.chains.top_parent + row_off[j], -1)
                                     for j, c in enumerate(cs)])[perm]
            tp_slot = np.where(tp_old >= 0, slot_of[np.maximum(tp_old, 0)], -1)
            chain_top.extend(tp_slot.tolist())
            chain_off.extend((grow + new_off[1:]).tolist())
            # parent chain of each chain = the chain whose bottom slot is its top parent
            chain_at_bottom = np.full(n_rows, -1, dtype=np.int64)
            chain_at_bottom[new_off[:-1]] = np.arange(n_ch)
            parent = np.where(tp_slot >= 0, chain_at_bottom[np.maximum(tp_slot, 0)], -1)
            cnt = np.bincount(parent[parent >= 0], minlength=n_ch)
            kid_cnt.extend(cnt.tolist())
            pk = np.argsort(np.where(parent >= 0, parent, n_ch), kind="stable")[: int(cnt.sum())]
            kids.extend(pk.tolist())
            # the pack's chain metadata in the kernel's LDS form (see bc_stagger_desc)
            q0 = np.concatenate(([0], np.cumsum(cnt)))[:-1]
            pkx = np.concatenate((pk, [0, 0]))
            k0 = np.where(cnt > 0, pkx[q0], 0)
            k1 = np.where(cnt > 1, pkx[np.minimum(q0 + 1, len(pkx) - 1)], 0)
            j0, j1 = new_off[:-1], new_off[1:]
            rec = np.stack([j0 | (j1 << 16), (tp_slot & 0xFFFF) | (cnt << 16), k0 | (k1 << 16), q0], axis=1)
            chain_rec_all.append(rec.astype(np.int32))
            chain_par_all.append(parent.astype(np.int16))
            kp = np.zeros(n_ch, dtype=np.int16)
            kp[: len(pk)] = pk
            kid_pad_all.append(kp)
            chain_work_all.append(balanced_work_order(j1 - j0, self.threads, int(np.sum(lev[perm] == 0))))
            ch_of = np.repeat(np.arange(n_ch), lens)
            link = np.arange(n_rows, dtype=np.int64) + 1
            link[j1 - 1] = np.where(tp_slot >= 0, tp_slot, j1 - 1)
            link[j0] |= 0x8000
            slot_link_all.append(link.astype(np.uint16).view(np.int16))
            slot_chain_all.append(ch_of.astype(np.int16))
            # per slot: its link | the chain its chain hangs from << 16 (0xFFFF: none)
            slot_aux = (link & 0xFFFF) | ((parent[ch_of] & 0xFFFF) << 16)
            lv_sorted = lev[perm]
            nlev = int(lv_sorted.max()) + 1
            level_off.extend(np.searchsorted(lv_sorted, np.arange(nlev + 1), side="left").tolist())
            cv_off = cls_off = 0
            for j, i in enumerate(cells):
                c = self.cells[i]
                g = c.geom
                ro = row_off[j]
                self.cell_cv0[i] = gcv + cv_off
                self.cell_ncv[i] = g.n_cv
                self.cell_pack[i] = pi
                self.cell_row_off[i] = ro
                if np.any(c.cv_cls + cls_off >= (1 << 14)):
                    raise ValueError("too many mechanism classes in one pack")
                cv_slots = slot_of[g.dynamic_rows + ro]
                info = cv_slots.astype(np.uint64) | ((c.cv_cls + cls_off).astype(np.uint64) << np.uint64(16))
                clamp_cv = np.zeros(g.n_cv, dtype=bool)
                clamp_bnd = set()
                for (kind, idx_, t0, t1, amp) in c.clamps:
                    if kind == "mid":
                        clamp_cv[idx_] = True
                        clamp_rows.append(int(cv_slots[idx_]))
                        clamp_area.append(g.cv_area_um2[idx_] * 1e-8)
                    else:
                        sl = int(slot_of[idx_ + ro])
                        clamp_bnd.add(sl)
                        clamp_rows.append(sl)
                        clamp_area.append(0.0)
                    clamp_kind.append(kind)
                    clamp_t.append((t0, t1))
                    clamp_amp_raw.append(amp)
                pflag = np.zeros(g.n_cv, dtype=bool)
                pflag[c.probes] = True
                info |= (pflag.astype(np.uint64) << np.uint64(30)) | (clamp_cv.astype(np.uint64) << np.uint64(31))
                cv_info.append(info.astype(np.uint32))
                cv_aux.append(slot_aux[cv_slots].astype(np.uint32))
                for k, cvi in enumerate(c.probes.tolist()):
                    probe_row.append(int(cv_slots[cvi]))
                    probe_col.append(int(self.cell_probe0[i] + k))
                bslots = slot_of[np.nonzero(~g.row_is_mid)[0] + ro]
                bflag = np.asarray([int(b) in clamp_bnd for b in bslots.tolist()], dtype=np.uint64)
                bnd_info.append((bslots.astype(np.uint64) | (bflag << np.uint64(31))).astype(np.uint32))
                bnd_aux.append(slot_aux[bslots].astype(np.uint32))
                bnd_cell.append(np.full(len(bslots), i, dtype=np.int64))
                cls_tab.append(c.cls_rows)
                if self.region_classes:
                    for code in range(len(REGION_CODES)):
                        hit = np.nonzero(c.cls_rows[:, 10] == code)[0]
                        self.cell_region_cls[i, code] = gcls + cls_off + hit[0] if hit.size else -1
                cv_off += g.n_cv
                cls_off += c.cls_rows.shape[0]
            P.n_cv, P.n_rows, P.n_ch, P.n_lv = cv_off, n_rows, n_ch, nlev
            P.n_bnd = int(sum(len(b) for b in bnd_info)) - gbnd
            P.n_cls = cls_off
            P.n_clamp = len(clamp_rows) - gclamp
            P.n_probe = len(probe_row) - gprobe
            packs.append(P)
            gcv += cv_off
            grow += n_rows
            gch += n_ch
            glv += nlev + 1
            gbnd += P.n_bnd
            gcls += cls_off
            gclamp += P.n_clamp
            gprobe += P.n_probe
        self.packs = (BcPack * len(packs))(*packs)
        self.pack_array = np.frombuffer(bytes(self.packs), dtype=np.int32).reshape(len(packs), 16).copy()
        self.n_packs = len(packs)
        self.n_cv_total = gcv
        self.n_rows_total = grow
        self.cv_info = np.concatenate(cv_info)
        self.bnd_info = np.concatenate(bnd_info) if bnd_info else np.zeros(0, np.uint32)
        self.cv_aux = np.concatenate(cv_aux)
        self.bnd_aux = np.concatenate(bnd_aux) if bnd_aux else np.zeros(0, np.uint32)
        self.bnd_cell = np.concatenate(bnd_cell) if bnd_cell else np.zeros(0, np.int64)
        self.chain_off = np.asarray(chain_off, dtype=np.int32)
        self.chain_top = np.asarray(chain_top, dtype=np.int32)
        self.kid_off = np.concatenate(([0], np.cumsum(kid_cnt))).astype(np.int32)
        self.kids = np.asarray(kids, dtype=np.int32)
        self.level_off = np.asarray(level_off, dtype=np.int32)
        self.chain_rec = np.concatenate(chain_rec_all, axis=0)
        self.chain_par = np.concatenate(chain_par_all)
        self.kid_pad = np.concatenate(kid_pad_all)
        self.chain_work = np.concatenate(chain_work_all)
        self.slot_link = np.concatenate(slot_link_all)
        self.slot_chain = np.concatenate(slot_chain_all)
        self.slot_lo = np.concatenate(slot_lo)
        self.slot_up = np.concatenate(slot_up)
        self.slot_rows = slot_rows_all
        self.cls_tab = np.concatenate(cls_tab, axis=0)
        self.clamp_row = np.asarray(clamp_rows, dtype=np.int32)
        self.clamp_t = np.asarray(clamp_t, dtype=np.float64).reshape(-1, 2)
        self.clamp_amp_raw = np.asarray(clamp_amp_raw, dtype=np.float64)
        self.clamp_kind = clamp_kind
        self.clamp_area = np.asarray(clamp_area, dtype=np.float64)
        self.probe_row = np.asarray(probe_row, dtype=np.int32)
        self.probe_col = np.asarray(probe_col, dtype=np.int32)
        self.n_probes = int(self.cell_probe0[-1])
        p = self.pack_array
        self.max_rows = int(p[:, 3].max())
        self.max_chains = int(p[:, 5].max())
        self.max_cls = int(p[:, 11].max())
        self.max_levels = int(p[:, 7].max())

    # ------------------------------------------------------------ dt-tables
    def dt_tables(self, dt: float):
        """Arrays that depend on dt: dt-scaled axial coefficients, dt/Cm in the
        class table, clamp step windows and clamp right-hand-side terms."""
        cls = self.cls_tab.copy()
        cls[:, 0] = dt / cls[:, 0]
        on = np.asarray([first_step_at(t0, dt) for t0, _ in self.clamp_t], dtype=np.int64)
        off = np.asarray([first_step_at(t1, dt) for _, t1 in self.clamp_t], dtype=np.int64)
        amp = np.empty(len(self.clamp_kind))
        for e, kind in enumerate(self.clamp_kind):
            if kind == "mid":
                amp[e] = self.clamp_amp_raw[e] * 1e-3 / self.clamp_area[e]  # uA/cm^2
            else:
                amp[e] = dt * self.clamp_amp_raw[e] * 1e-3  # mV: dt * I[nA] / 1 uF
        return {"lo": dt * self.slot_lo, "up": dt * self.slot_up, "cls_tab": cls,
                "clamp_on": np.minimum(on, INT32_MAX).astype(np.int32),
                "clamp_off": np.minimum(off, INT32_MAX).astype(np.int32), "clamp_amp": amp}

    # ---------------------------------------------------------- init state
    def initial_state(self, V0=None):
        """reset_state: V at V0 (per cell model unless given), gates at their
        steady state, Ci at its initializer.  Returns (V[n_cv], S[ns, n_cv],
        Vbnd[n_bnd]); boundary rows start at their cell's V0 (any start value
        is consistent: every solve re-imposes the boundary constraints)."""
        V = np.empty(self.n_cv_total)
        v0 = np.empty(len(self.models))
        for i, m in enumerate(self.models):
            v = m.V0 if V0 is None else (V0[i] if np.ndim(V0) else V0)
            v0[i] = v
            V[self.cell_cv0[i]:self.cell_cv0[i] + self.cell_ncv[i]] = v
        Vbnd = v0[self.bnd_cell]
        S = np.zeros((len(self.slots), self.n_cv_total))
        cc = self.cc
        pools = [m.ca for m in self.models if m.ca is not None]
        ci0 = np.full(self.n_cv_total, float(pools[0]["Ci0"]) if pools else 0.0)
        for s, (name, k) in enumerate(self.slots):
            if name == "Na_Ba2002":
                S[s] = rates.na_steady(V, cc["na_vsh"])[k]
            elif name == "Na_HH1952":
                S[s] = rates.nahh_steady(V, cc["na_vsh"])[k]
            elif name == "KDR_Ba2002":
                S[s] = rates.kdr_steady(V, cc["kdr_vsh"])
            elif name == "K_HH1952":
                S[s] = rates.khh_steady(V, cc["kdr_vsh"])
            elif name == "CaL_IS2008":
                S[s] = rates.cal_steady(V, cc["cal_vsh"])[k]
            elif name == "CaT_HM1992":
                S[s] = rates.cat_steady(V, cc["cat_vsh"])[k]
            elif name == "AHP_De1994":
                S[s] = rates.ahp_steady(ci0, cc["ahp_alpha"], cc["ahp_beta"], cc["ahp_n"])
            elif name == "HCN_HM1992":
                S[s] = rates.hcn_steady(V)
            elif name == "Ca":
                S[s] = ci0
        return V, S, Vbnd

    def validate(self):
        """Checks the packed tables a launch reads before anything touches the
        device, so that a corrupted layout (a parent index out of its pack, a
        cycle in the chain tree, a slot used twice, a link that points
        outside its pack ...) raises ``NativeError`` with status
        ``INVALID_LAYOUT`` instead of reaching the kernel.  The kernel keeps
        its own guard (an unresolved chain tree sets status bit 2)."""
        import hashlib

        from ._native import NativeError

        def bad(msg):
            raise NativeError(2, msg)

        # tables already checked and unchanged since: nothing to do
        h = hashlib.blake2b(digest_size=16)
        for arr in (self.pack_array, self.chain_rec, self.chain_off, self.chain_par, self.chain_top, self.slot_link,
                    self.slot_chain, self.cv_info, self.bnd_info, self.level_off, self.kid_pad, self.chain_work,
                    self.probe_row, self.probe_col, self.clamp_row, self.cv_aux, self.bnd_aux):
            h.update(np.ascontiguousarray(arr).view(np.uint8))
        h.update(np.asarray([self.n_cv_total, self.n_rows_total, self.max_rows, self.max_chains, self.n_probes],
                            dtype=np.int64).view(np.uint8))
        digest = h.digest()
        if getattr(self, "_valid_digest", None) == digest:
            return self

        P = self.pack_array
        if P.ndim != 2 or P.shape[1] != 16 or P.shape[0] != self.n_packs:
            bad("pack table has the wrong shape")
        cv0, ncv, row0, nrow, ch0, nch, lv0, nlv = (P[:, i].astype(np.int64) for i in range(8))
        bnd0, nbnd, cls0, ncls, cl0, ncl, pr0, npr = (P[:, i].astype(np.int64) for i in range(8, 16))
        for name, off, cnt, total in (("CV", cv0, ncv, self.n_cv_total), ("row", row0, nrow, self.n_rows_total),
                                      ("chain", ch0, nch, len(self.chain_rec)), ("boundary", bnd0, nbnd, len(self.bnd_info)),
                                      ("class", cls0, ncls, len(self.cls_tab)), ("clamp", cl0, ncl, len(self.clamp_row)),
                                      ("probe", pr0, npr, len(self.probe_row))):
            if np.any(cnt < 0) or np.any(off != np.concatenate(([0], np.cumsum(cnt)[:-1]))) or int(cnt.sum()) != total:
                bad(f"pack {name} ranges are not contiguous over the {name} arrays")
        if np.any(lv0 != np.concatenate(([0], np.cumsum(nlv + 1)[:-1]))) or int((nlv + 1).sum()) != len(self.level_off):
            bad("pack level ranges are not contiguous")
        if np.any(nch < 1) or np.any(nlv < 1) or nrow.max() > self.max_rows or nch.max() > self.max_chains:
            bad("pack sizes exceed the layout maxima")
        pack_of_ch = np.repeat(np.arange(self.n_packs), nch)
        lrow0 = row0[pack_of_ch]
        co = self.chain_off.astype(np.int64)
        j0, j1 = co[:-1] - lrow0, co[1:] - lrow0
        if len(co) != len(self.chain_rec) + 1 or np.any(j1 <= j0) or np.any(j0 < 0) or np.any(j1 > nrow[pack_of_ch]):
            bad("a chain's slot range lies outside its pack")
        if np.any(j0[ch0] != 0) or np.any(j1[ch0 + nch - 1] != nrow):
            bad("chains do not cover their pack's rows")
        rec = self.chain_rec.astype(np.int64)
        if np.any((rec[:, 0] & 0xFFFF) != j0) or np.any((rec[:, 0] >> 16) & 0xFFFF != j1):
            bad("chain records disagree with the chain offsets")
        local = np.arange(len(rec)) - ch0[pack_of_ch]
        par = self.chain_par.astype(np.int64)
        if np.any(par < -1) or np.any(par >= nch[pack_of_ch]) or np.any(par == local):
            bad("a parent chain index lies outside its pack")
        # levels (heights): a parent sits strictly higher than each kid, so
        # the chain tree is acyclic
        lvo = self.level_off.astype(np.int64)
        last = lv0 + nlv
        is_last = np.zeros(len(lvo), dtype=bool)
        is_last[last] = True
        pack_of_lv = np.repeat(np.arange(self.n_packs), nlv + 1)
        inner = np.nonzero(~is_last)[0]
        if np.any(lvo[lv0] != 0) or np.any(lvo[last] != nch) or np.any(lvo[inner + 1] <= lvo[inner]):
            bad("level offsets do not partition their pack's chains")
        flag = np.bincount(lvo[inner] + ch0[pack_of_lv[inner]], minlength=len(rec))
        level = np.cumsum(flag) - (lv0 - np.arange(self.n_packs))[pack_of_ch] - 1
        has_p = par >= 0
        gpar = np.where(has_p, par + ch0[pack_of_ch], 0)
        if np.any(level[gpar][has_p] <= level[has_p]):
            bad("a parent chain is not above its kid (cycle or corrupted parent index)")
        tp = ((rec[:, 1] & 0xFFFF) ^ 0x8000) - 0x8000
        if np.any(np.where(has_p, tp != j0[gpar], tp != -1)):
            bad("a chain's top does not hang from its parent's bottom row")
        nk = (rec[:, 1] >> 16) & 0xFFFF
        cnt = np.bincount(gpar[has_p], minlength=len(rec))
        if np.any(cnt != nk):
            bad("kid counts disagree with the parent indices")
        if np.any(level[nk == 0] != 0) or np.any(level[nk > 0] == 0):
            bad("leaf chains are not exactly level 0")
        # slots: each slot in exactly one chain; links inside the pack
        slot_pack = np.repeat(np.arange(self.n_packs), nrow)
        n_slot = len(self.slot_link)
        if n_slot != self.n_rows_total or len(self.slot_chain) != n_slot:
            bad("slot tables have the wrong length")
        ch_of_slot = np.repeat(np.arange(len(rec)), j1 - j0)
        if np.any(self.slot_chain.astype(np.int64) != local[ch_of_slot]):
            bad("slot chains disagree with the chain ranges")
        link = self.slot_link.astype(np.int64) & 0xFFFF
        lpar = link & 0x7FFF
        pos = np.arange(n_slot) - row0[slot_pack]
        top = pos == (j1[ch_of_slot] - 1)
        exp_par = np.where(top, np.where(has_p[ch_of_slot], tp[ch_of_slot], pos), pos + 1)
        if np.any(lpar != exp_par) or np.any(((link & 0x8000) != 0) != (pos == j0[ch_of_slot])):
            bad("a slot link points outside its chain's parent structure")
        # CV and boundary rows: every slot of a pack is one or the other, once
        cv_pack = np.repeat(np.arange(self.n_packs), ncv)
        cs = self.cv_info.astype(np.int64) & 0xFFFF
        if np.any(cs >= nrow[cv_pack]):
            bad("a CV's slot lies outside its pack")
        if np.any(((self.cv_info.astype(np.int64) >> 16) & 0x3FFF) >= ncls[cv_pack]):
            bad("a CV's mechanism class lies outside its pack")
        bnd_pack = np.repeat(np.arange(self.n_packs), nbnd)
        bs = self.bnd_info.astype(np.int64) & 0x7FFFFFFF
        if np.any(bs >= nrow[bnd_pack]):
            bad("a boundary row's slot lies outside its pack")
        used = np.bincount(np.concatenate([cs + row0[cv_pack], bs + row0[bnd_pack]]), minlength=n_slot)
        if used.size != n_slot or np.any(used != 1):
            bad("a slot is not exactly one CV or boundary row")
        pr_pack = np.repeat(np.arange(self.n_packs), npr)
        if np.any(self.probe_row >= nrow[pr_pack]) or np.any(self.probe_col < 0) or np.any(self.probe_col >= max(self.n_probes, 1)):
            bad("a probe lies outside its pack or the trace")
        cl_pack = np.repeat(np.arange(self.n_packs), ncl)
        if np.any(self.clamp_row < 0) or np.any(self.clamp_row >= nrow[cl_pack]):
            bad("a clamp row lies outside its pack")
        kid_ok = self.kid_pad.astype(np.int64)
        if np.any(kid_ok < 0) or np.any(kid_ok >= nch[pack_of_ch]):
            bad("a kid list entry lies outside its pack")
        slot_aux = (link & 0xFFFF) | ((par[ch_of_slot] & 0xFFFF) << 16)
        if len(self.cv_aux) != self.n_cv_total or np.any(self.cv_aux.astype(np.int64) != slot_aux[cs + row0[cv_pack]]):
            bad("per-CV slot metadata disagrees with the slot links and parent chains")
        if len(self.bnd_aux) != len(self.bnd_info) or np.any(self.bnd_aux.astype(np.int64) != slot_aux[bs + row0[bnd_pack]]):
            bad("per-boundary-row slot metadata disagrees with the slot links and parent chains")
        w = self.chain_work.astype(np.int64)
        if len(w) != len(rec) or np.any(np.sort(w + ch0[pack_of_ch]) != np.arange(len(rec))):
            bad("the chain work order is not a permutation of each pack's chains")
        self._valid_digest = digest
        return self

    def cell_slice(self, i: int) -> slice:
        return slice(int(self.cell_cv0[i]), int(self.cell_cv0[i] + self.cell_ncv[i]))

    def probe_cvs_global(self) -> np.ndarray:
        """Global CV index of every trace column."""
        out = np.empty(self.n_probes, dtype=np.int64)
        for i, c in enumerate(self.cells):
            for k, cv in enumerate(c.probes):
                out[int(self.cell_probe0[i]) + k] = int(self.cell_cv0[i]) + int(cv)
        return out

    def summary(self) -> dict:
        p = self.pack_array
        return {"cells": len(self.cells), "packs": self.n_packs, "cv_total": int(self.n_cv_total),
                "rows_total": int(self.n_rows_total), "max_rows": self.max_rows, "max_levels": int(p[:, 7].max()),
                "threads": self.threads, "cv_per_thread": self.cv_per_thread, "channel_set": hex(self.chset),
                "backsub": self.backsub}


class StaggerRuntime:
    """Device buffers + C-ABI calls for a :class:`StaggerLayout`."""

    def __init__(self, layout: StaggerLayout, dt: float, *, precision: int = 32, device="cuda"):
        import torch

        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        self.torch = torch
        self.L = layout
        self.precision = precision
        self.dtype = torch.float32 if precision == 32 else torch.float64
        self.device = torch.device(device)
        self.dt = float(dt)
        layout.validate()  # INVALID_LAYOUT before any device work
        lib()  # fail loudly when the native library is missing
        if not lib().bc_stagger_supported(precision, layout.chset, layout.cv_per_thread):
            raise RuntimeError(f"no kernel for channel set {hex(layout.chset)} / cv_per_thread "
                               f"{layout.cv_per_thread}")
        t = self._t
        self.packs = t(layout.pack_array.reshape(-1), torch.int32)
        self.cv_info = t(layout.cv_info.view(np.int32), torch.int32)
        self.bnd_info = t(layout.bnd_info.view(np.int32), torch.int32)
        self.chain_off = t(layout.chain_off, torch.int32)
        self.chain_top = t(layout.chain_top, torch.int32)
        self.kid_off = t(layout.kid_off, torch.int32)
        self.kids = t(layout.kids if layout.kids.size else np.zeros(1, np.int32), torch.int32)
        self.level_off = t(layout.level_off, torch.int32)
        self.chain_rec = t(layout.chain_rec.reshape(-1), torch.int32)
        self.chain_par = t(layout.chain_par, torch.int16)
        self.kid_pad = t(layout.kid_pad, torch.int16)
        self.slot_link = t(layout.slot_link, torch.int16)
        self.slot_chain = t(layout.slot_chain, torch.int16)
        self.chain_work = t(layout.chain_work, torch.int16)
        self.cv_aux = t(layout.cv_aux.view(np.int32), torch.int32)
        self.bnd_aux = t(layout.bnd_aux.view(np.int32) if layout.bnd_aux.size else np.zeros(1, np.int32), torch.int32)
        self.probe_row = t(layout.probe_row if layout.probe_row.size else np.zeros(1, np.int32), torch.int32)
        self.probe_col = t(layout.probe_col if layout.probe_col.size else np.zeros(1, np.int32), torch.int32)
        self.clamp_row = t(layout.clamp_row if layout.clamp_row.size else np.zeros(1, np.int32), torch.int32)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._set_dt_tables()
        CC = chan_const_type(precision)
        self._cc = CC(**{k: float(v) for k, v in layout.cc.items()})
        V, S, Vb = layout.initial_state()
        self.V = t(V, self.dtype)
        self.S = t(S.reshape(-1) if S.size else np.zeros(1), self.dtype)
        self.Vbnd = t(Vb if Vb.size else np.zeros(1), self.dtype)
        self.flags = 0  # BC_STEP_* bits for every launch
        self.step_index = 0
        self._graphs = {}
        self._spk = None

    def _t(self, arr, dtype):
        return self.torch.as_tensor(np.ascontiguousarray(arr)).to(device=self.device, dtype=dtype)

    def _set_dt_tables(self):
        tb = self.L.dt_tables(self.dt)
        t = self._t
        self.lo = t(tb["lo"], self.dtype)
        self.up = t(tb["up"], self.dtype)
        self.cls_tab = t(tb["cls_tab"].reshape(-1), self.dtype)
        n = max(1, tb["clamp_on"].size)
        self.clamp_on = t(tb["clamp_on"] if tb["clamp_on"].size else np.zeros(n, np.int32), self.torch.int32)
        self.clamp_off = t(tb["clamp_off"] if tb["clamp_off"].size else np.zeros(n, np.int32), self.torch.int32)
        self.clamp_amp = t(tb["clamp_amp"] if tb["clamp_amp"].size else np.zeros(n), self.dtype)

    def reset_state(self, V0=None):
        V, S, Vb = self.L.initial_state(V0)
        self.V.copy_(self._t(V, self.dtype))
        if S.size:
            self.S.copy_(self._t(S.reshape(-1), self.dtype))
        if Vb.size:
            self.Vbnd.copy_(self._t(Vb, self.dtype))
        self.step_index = 0
        self.status.zero_()
        if self._spk is not None:
            self.clear_spikes()

    # ---------------------------------------------------------------- spikes
    def enable_spikes(self, v_th: float = 0.0, cap: int = 256):
        """Record threshold crossings (V_n < v_th <= V_{n+1}) of every trace
        column in every later launch: up to ``cap`` step indices per column
        and the count of all of them (``get_spike``, ``braincell/_base.py:
        421-431``).  Counting starts now, from the current voltages."""
        if self.L.n_probes == 0:
            raise ValueError("spike detection reads the probe trace: the layout has no probes")
        if cap < 0:
            raise ValueError("cap must be >= 0")
        torch = self.torch
        n = self.L.n_probes
        self._spk_cv = torch.as_tensor(self.L.probe_cvs_global(), device=self.device)
        self._spk_buf = {"v_last": torch.empty(n, dtype=self.dtype, device=self.device),
                         "count": torch.zeros(n, dtype=torch.int32, device=self.device),
                         "steps": torch.full((n, max(cap, 1)), -1, dtype=torch.int32, device=self.device)}
        b = self._spk_buf
        self._spk = SpikeDesc(int(cap), float(v_th), b["v_last"].data_ptr(), b["count"].data_ptr(),
                              b["steps"].data_ptr())
        self.clear_spikes()
        return self

    def clear_spikes(self):
        """Restart the spike record at the current voltages."""
        b = self._spk_buf
        b["v_last"].copy_(self.V[self._spk_cv])
        b["count"].zero_()
        b["steps"].fill_(-1)

    def spike_steps(self):
        """Per trace column, the recorded spike step indices (global step n
        of a crossing V_n < v_th <= V_{n+1}); raises if a column spiked more
        often than the record holds."""
        if self._spk is None:
            raise RuntimeError("spike output is off: call enable_spikes() first")
        cnt = self._spk_buf["count"].cpu().numpy()
        if (cnt > self._spk.cap).any():
            raise RuntimeError(f"{int(cnt.max())} spikes on one column exceed the record of {self._spk.cap} "
                               "per column: enable_spikes(cap=...) with a larger cap")
        st = self._spk_buf["steps"].cpu().numpy()
        return [st[c, :cnt[c]].astype(np.int64) for c in range(cnt.size)]

    def save_state(self):
        """Device-side snapshot of V, the states and the boundary voltages."""
        self._saved = (self.V.clone(), self.S.clone(), self.Vbnd.clone(), self.step_index)

    def restore_state(self):
        """Device-to-device restore of the last :meth:`save_state`."""
        V, S, Vb, n = self._saved
        self.V.copy_(V)
        self.S.copy_(S)
        self.Vbnd.copy_(Vb)
        self.step_index = n
        self.status.zero_()
        if self._spk is not None:
            self.clear_spikes()

    def desc(self, n_steps: int, trace=None, step0: int = None, rec=None) -> StaggerDesc:
        L = self.L
        d = StaggerDesc()
        d.precision = self.precision
        d.channels = L.chset
        d.n_packs = L.n_packs
        d.threads = L.threads
        d.cv_per_thread = L.cv_per_thread
        d.max_rows = L.max_rows
        d.max_chains = L.max_chains
        d.max_cls = L.max_cls
        d.max_levels = L.max_levels
        d.n_cv_total = L.n_cv_total
        p = lambda x: ctypes.c_void_p(x.data_ptr())
        d.packs, d.cv_info, d.bnd_info = p(self.packs), p(self.cv_info), p(self.bnd_info)
        d.chain_off, d.chain_top, d.kid_off, d.kids = p(self.chain_off), p(self.chain_top), p(self.kid_off), p(self.kids)
        d.level_off = p(self.level_off)
        d.chain_rec, d.chain_par, d.kid_pad = p(self.chain_rec), p(self.chain_par), p(self.kid_pad)
        d.slot_link, d.slot_chain = p(self.slot_link), p(self.slot_chain)
        d.chain_work = p(self.chain_work)
        d.cv_aux, d.bnd_aux = p(self.cv_aux), p(self.bnd_aux)
        d.lo, d.up, d.cls_tab = p(self.lo), p(self.up), p(self.cls_tab)
        d.clamp_row, d.clamp_on, d.clamp_off, d.clamp_amp = p(self.clamp_row), p(self.clamp_on), p(self.clamp_off), p(self.clamp_amp)
        d.probe_row, d.probe_col = p(self.probe_row), p(self.probe_col)
        d.chan_const = ctypes.cast(ctypes.pointer(self._cc), ctypes.c_void_p)
        d.V, d.Vbnd, d.state = p(self.V), p(self.Vbnd), p(self.S)
        if trace is not None:
            if trace.dtype != self.dtype or trace.shape[-1] != L.n_probes or not trace.is_contiguous():
                raise ValueError("trace must be a contiguous [n_steps, n_probes] tensor of the runtime dtype")
            if trace.shape[0] < n_steps:
                raise ValueError("trace has fewer rows than steps")
            d.trace = p(trace)
        elif L.n_probes > 0 and n_steps > 0:
            raise ValueError("a trace buffer is required when probes exist")
        d.trace_ld = L.n_probes
        d.rec = p(rec) if rec is not None else None
        d.status = p(self.status)
        d.step0 = self.step_index if step0 is None else int(step0)
        d.n_steps = int(n_steps)
        d.dt = self.dt
        d.flags = self.flags | L.step_flags
        syn = getattr(self, "synapse_desc", None)
        if syn is not None:
            self._syn_drive_workspace(syn, int(n_steps))
        d.syn = ctypes.cast(ctypes.pointer(syn), ctypes.c_void_p) if syn is not None else None
        d.spk = ctypes.cast(ctypes.pointer(self._spk), ctypes.c_void_p) if self._spk is not None else None
        return d

    def _syn_drive_workspace(self, syn, n_steps: int) -> None:
        """The synapse descriptor's drive workspace, at least n_steps x n_syn
        reals.  A grown workspace replaces the pointer; the smaller ones stay
        allocated with the runtime (graphs captured earlier still use them)."""
        need = max(1, n_steps) * max(1, int(syn.n_syn))
        ws = getattr(self, "_syn_drive", [])
        if not ws or ws[-1].numel() < need:
            ws.append(self.torch.empty(need, dtype=self.dtype, device=self.device))
            self._syn_drive = ws
        syn.drive = ws[-1].data_ptr()

    def stream_handle(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def new_trace(self, n_steps: int):
        return self.torch.empty((n_steps, self.L.n_probes), dtype=self.dtype, device=self.device)

    def run(self, n_steps: int, trace=None, chunk: int = None, rec=None):
        """Advance ``n_steps`` steps; returns the probe trace [n_steps, n_probes].
        ``rec`` (one launch only) receives the step-start snapshots."""
        if trace is None:
            trace = self.new_trace(n_steps)
        chunk = n_steps if chunk is None else int(chunk)
        if rec is not None and chunk < n_steps:
            raise ValueError("recording needs a single launch (chunk >= n_steps)")
        done = 0
        syn_trace = getattr(self, "syn_trace", None)  # [n_steps, 2 n_syn] synapse states, or None
        while done < n_steps:
            k = min(chunk, n_steps - done)
            if syn_trace is not None:
                self.synapse_desc.syn_trace = syn_trace[done:].data_ptr()
            d = self.desc(k, trace[done:] if self.L.n_probes else None, rec=rec)
            self._keep = d
            check(lib().bc_stagger_run(ctypes.byref(d), self.stream_handle()))
            done += k
            self.step_index += k
        return trace

    def make_graph(self, n_chunks: int, steps_per_chunk: int, trace):
        """Capture ``n_chunks`` launches into a hipGraph that starts at the
        current step index; replays always restart from that index."""
        d = self.desc(steps_per_chunk, trace)
        h = ctypes.c_void_p()
        # capture needs a non-default stream (default priority)
        cap = self.torch.cuda.Stream(device=self.device)
        self.torch.cuda.synchronize(self.device)
        check(lib().bc_stagger_graph_create(ctypes.byref(d), n_chunks, steps_per_chunk,
                                            ctypes.c_void_p(cap.cuda_stream), ctypes.byref(h)))
        cap.synchronize()
        return _Graph(h, d, cap)

    def soma_trace_cols(self):
        return self.L.cell_probe0[:-1]

    def state_of_cell(self, i: int):
        sl = self.L.cell_slice(i)
        ncv = self.L.n_cv_total
        V = self.V[sl].double().cpu().numpy()
        S = self.S.view(len(self.L.slots), ncv)[:, sl].double().cpu().numpy() if self.L.slots else None
        return V, S

    def check_status(self):
        st = int(self.status.item())
        if st != 0:
            from ._native import NativeError
            if st & 2:
                raise NativeError(4, "chain tree of a pack did not resolve (inconsistent layout)")
            if st & 4:
                raise NativeError(4, "a pack's synapses exceed the synapse LDS plan (inconsistent tables)")
            raise NativeError(4, "non-finite membrane voltage detected")


class _Graph:
    def __init__(self, handle, desc, stream):
        self.h = handle
        self.desc = desc  # keep chan_const pointer alive
        self.stream = stream

    def launch(self, stream):
        check(lib().bc_graph_launch(self.h, stream))

    def __del__(self):
        try:
            if self.h:
                lib().bc_graph_destroy(self.h)
        except Exception:
            pass
