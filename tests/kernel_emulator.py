"""CPU transcription of the HIP staggered kernel's Hines phase (test only).

Runs the chain walker of ``bc_stagger.hip`` (delta-V form) over a packed
``StaggerLayout`` in NumPy so that layout, packing and the algorithm can be
checked against the oracle without a GPU.
"""

import numpy as np


def hines_delta(L, dt, bd_in, br_in, xv, dtype=np.float64):
    """bd_in/br_in/xv: per global row.  Returns dV per global row."""
    tb = L.dt_tables(dt)
    dg, lo_a, up_a = (tb[k].astype(dtype) for k in ("cr_dg", "cr_lo", "cr_up"))
    bd = bd_in.astype(dtype).copy()
    br = br_in.astype(dtype).copy()
    xv = xv.astype(dtype)
    P = L.pack_array
    for p in range(L.n_packs):
        row0, ch0, lv0, nlv, nch = P[p, 2], P[p, 4], P[p, 6], P[p, 7], P[p, 5]
        kd = np.zeros(nch, dtype)
        kr = np.zeros(nch, dtype)
        for lev in range(nlv):
            for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
                gk = ch0 + k
                j0, j1 = L.chain_off[gk], L.chain_off[gk + 1]
                tp = L.chain_top[gk]
                row = L.cr_row[j0]
                vr = xv[row0 + row]
                d = dg[j0] + bd[row0 + row]
                r = br[row0 + row]
                for q in range(L.kid_off[gk], L.kid_off[gk + 1]):
                    d -= kd[L.kids[q]]
                    r -= kr[L.kids[q]]
                j = j0
                while True:
                    pr = L.cr_row[j + 1] if j + 1 < j1 else tp
                    vp = xv[row0 + pr] if pr >= 0 else vr
                    lo, up = lo_a[j], up_a[j]
                    dl = vr - vp
                    r = r + lo * dl
                    inv = dtype(1) / d
                    bd[row0 + row] = inv
                    br[row0 + row] = r
                    m = up * inv
                    j += 1
                    if j >= j1:
                        break
                    row, vr = pr, vp
                    d = (dg[j] + bd[row0 + row]) - lo * m
                    r = (br[row0 + row] - r * m) - up * dl
                kd[k] = lo * m
                kr[k] = r * m + up * dl
        for lev in range(nlv - 1, -1, -1):
            for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
                gk = ch0 + k
                j0, j1 = L.chain_off[gk], L.chain_off[gk + 1]
                tp = L.chain_top[gk]
                xp = br[row0 + tp] if tp >= 0 else dtype(0)
                for j in range(j1 - 1, j0 - 1, -1):
                    row = row0 + L.cr_row[j]
                    x = (br[row] - lo_a[j] * xp) * bd[row]
                    br[row] = x
                    xp = x
    return br


def cv_rows(L):
    """Global row of every CV (pack order)."""
    info = L.cv_info.astype(np.int64)
    rows = np.empty(L.n_cv_total, dtype=np.int64)
    P = L.pack_array
    for p in range(L.n_packs):
        cv0, ncv, row0 = P[p, 0], P[p, 1], P[p, 2]
        rows[cv0:cv0 + ncv] = row0 + (info[cv0:cv0 + ncv] & 0xFFFF)
    return rows


def bnd_rows(L):
    P = L.pack_array
    out = np.empty(len(L.bnd_info), dtype=np.int64)
    for p in range(L.n_packs):
        b0, nb, row0 = P[p, 8], P[p, 9], P[p, 2]
        out[b0:b0 + nb] = row0 + (L.bnd_info[b0:b0 + nb].astype(np.int64) & 0x7FFFFFFF)
    return out
