"""FP32 accuracy of the plain batched step over the full 1000 ms (40 000
steps of 0.025 ms) of the 4096-tree workload, on six of its cells (204 to
1960 CVs, five of them spiking under their clamps), against the FP64 C port
(pinned to the NumPy oracle by test_cport_vs_oracle.py): soma RMSE and
spike times."""

import os

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import soma_cv

pytestmark = pytest.mark.gpu

CELLS = (39, 52, 13, 1, 30, 27)  # 1960, 1936, 1812, 1052, 388, 204 CVs
STEPS = 40_000
DT = 0.025


def crossings(v0, v, th=0.0):
    prev = np.concatenate([[v0], v[:-1]])
    return np.nonzero((prev < th) & (v >= th))[0]


def test_workload_cells_over_1000_ms(cuda):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime
    from oracle.cport_run import LIB, PortBatch

    if not os.path.exists(LIB):
        pytest.skip("FP64 C port not built")
    models = [W.uneven_tree_cell(i) for i in CELLS]
    L = StaggerLayout(models, probes="soma")
    rt = StaggerRuntime(L, DT, precision=32)
    rt.enable_spikes(0.0, cap=2048)
    print(f"horizon: {len(CELLS)} cells x {STEPS} steps on the GPU", flush=True)
    tr = rt.run(STEPS, chunk=200)
    torch.cuda.synchronize()
    rt.check_status()
    got = tr.double().cpu().numpy()
    spk = rt.spike_steps()
    print("horizon: FP64 C port", flush=True)
    ref = PortBatch(models, DT, [soma_cv(m) for m in models]).run(STEPS, sequential=True,
                                                                  threads=PortBatch.max_threads())
    n_spiking = 0
    for k, i in enumerate(CELLS):
        rmse = float(np.sqrt(np.mean((got[:, k] - ref[:, k]) ** 2)))
        st_ref = crossings(models[k].V0, ref[:, k])
        st_got = spk[k]
        print(f"cell {i}: {L.cells[k].geom.n_cv} CVs, soma RMSE {rmse:.3e} mV, spikes {st_got.size} / {st_ref.size}",
              flush=True)
        assert rmse <= 0.05
        assert st_got.size == st_ref.size
        if st_ref.size:
            assert np.abs(st_got - st_ref).max() <= 1  # within one step (dt)
            n_spiking += 1
    assert n_spiking >= 5
