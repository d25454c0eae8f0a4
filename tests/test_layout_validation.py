"""Host-side layout validation: a corrupted table raises NativeError with
status INVALID_LAYOUT when a runtime is created, before any device work
(CPU only; no kernel is ever launched on a corrupted layout)."""

import copy

import numpy as np
import pytest

from braincell_amd import workloads as W
from braincell_amd._native import NativeError
from braincell_amd.runtime import StaggerLayout, StaggerRuntime


@pytest.fixture(scope="module")
def layout():
    models = [W.uneven_tree_cell(i, cmin=200, cmax=500) for i in range(6)] + [W.single_tree_cell(0)]
    return StaggerLayout(models, rows_per_pack=1400, cv_per_thread=4)


def corrupted(L, name, fn):
    M = copy.copy(L)
    arr = getattr(L, name).copy()
    fn(arr, M)
    setattr(M, name, arr)
    return M


def test_intact_layout_passes(layout):
    assert layout.validate() is layout
    assert layout.validate() is layout  # cached by digest


def _second_pack_chain(L, level_min=1):
    P = L.pack_array
    ch0, nch = int(P[1, 4]), int(P[1, 5])
    lv0 = int(P[1, 6])
    first_nonleaf = ch0 + int(L.level_off[lv0 + level_min])
    return ch0, nch, first_nonleaf


CASES = {
    # a parent index out of its pack
    "parent out of pack": ("chain_par", lambda a, L: a.__setitem__(_second_pack_chain(L)[0], 32000)),
    # a kid pointing at a leaf as its parent (a cycle / below its kid)
    "parent below kid": ("chain_par", lambda a, L: a.__setitem__(_second_pack_chain(L)[2], 0)),
    # a slot link to a row outside the chain structure
    "slot link": ("slot_link", lambda a, L: a.__setitem__(int(L.pack_array[1, 2]) + 1, np.int16(0x7FF0))),
    # a CV slot beyond its pack
    "cv slot": ("cv_info", lambda a, L: a.__setitem__(0, np.uint32(0xFFFF))),
    # two CVs on one slot
    "slot twice": ("cv_info", lambda a, L: a.__setitem__(1, a[0])),
    # a chain record whose range disagrees with the offsets
    "chain record": ("chain_rec", lambda a, L: a.__setitem__((3, 0), a[3, 0] + 1)),
    # a kid count that disagrees with the parents
    "kid count": ("chain_rec", lambda a, L: a.__setitem__((_second_pack_chain(L)[2], 1), a[_second_pack_chain(L)[2], 1] + (1 << 16))),
    # level offsets that do not partition the chains
    "levels": ("level_off", lambda a, L: a.__setitem__(1, a[1] + 1)),
    # a work order that is not a permutation
    "work order": ("chain_work", lambda a, L: a.__setitem__(1, a[0])),
    # a clamp row beyond its pack
    "clamp row": ("clamp_row", lambda a, L: a.__setitem__(0, 30000)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_corrupted_table_is_invalid_layout(layout, case):
    name, fn = CASES[case]
    M = corrupted(layout, name, fn)
    with pytest.raises(NativeError) as e:
        M.validate()
    assert e.value.code == 2
    # a runtime refuses the layout before touching the device
    with pytest.raises(NativeError) as e:
        StaggerRuntime(M, 0.025, precision=32, device="cpu")
    assert e.value.code == 2
    assert "INVALID_LAYOUT" in str(e.value)
