"""The FP64 C port (CPU baseline and the benchmark's accuracy judge) against
the FP64 NumPy oracle on cells of the 4096-tree workload (CPU only): soma
voltage within 1e-8 mV over 2400 steps (60 ms, 50 ms of it under the clamp)
and identical spike steps, for spiking and resting Ca-bearing cells."""

import os

import numpy as np
import pytest

from braincell_amd import workloads as W
from oracle.staggered import StaggeredOracle
from tests.helpers import soma_cv, spike_times

# cell index -> (CVs, soma spikes in the window): 27 (204, 0), 49 (276, 5),
# 30 (388, 16 over 75 ms), 14 (468, 9 over 75 ms)
CELLS = (27, 49, 30, 14)
STEPS = 2400


@pytest.fixture(scope="module")
def port():
    from oracle import build_cport
    from oracle.cport_run import LIB

    if not os.path.exists(LIB):
        build_cport.build(verbose=False)
    from oracle.cport_run import PortBatch

    return PortBatch


@pytest.fixture(scope="module")
def cells():
    models = [W.uneven_tree_cell(i) for i in CELLS]
    scv = [soma_cv(m) for m in models]
    refs = []
    for k, m in enumerate(models):
        o = StaggeredOracle(m)
        o.reset_state(m.V0)
        refs.append(o.run(0.025, STEPS, probe_cvs=[scv[k]])[:, 0])
    return models, scv, refs


@pytest.mark.parametrize("sequential", [True, False], ids=["sequential-backsub", "recursive-doubling"])
def test_port_matches_oracle(port, cells, sequential):
    models, scv, refs = cells
    assert all(m.ca is not None for m in models)
    got = port(models, 0.025, scv).run(STEPS, sequential=sequential, threads=2)
    n_spiking = 0
    for k, m in enumerate(models):
        ref = refs[k]
        assert np.abs(got[:, k] - ref).max() <= 1e-8, (CELLS[k], np.abs(got[:, k] - ref).max())
        st_ref, st_got = spike_times(ref, 0.025), spike_times(got[:, k], 0.025)
        np.testing.assert_array_equal(st_got, st_ref)
        n_spiking += len(st_ref) > 0
    assert n_spiking >= 3
