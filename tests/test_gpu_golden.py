"""The HIP staggered step against the committed golden fixtures
(tests/golden/): FP64 to rounding, FP32 to the 0.05 mV soma gate."""

import os

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.golden import make_golden as MG
from tests.helpers import rmse, run_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRACES = [("c0_trace.npz", lambda: W.single_tree_cell(0)), ("ca_cell_trace.npz", MG.ca_cell)]


@pytest.mark.parametrize("name,maker", TRACES)
def test_trace_fixture_fp64(cuda, name, maker):
    g = np.load(os.path.join(HERE, name))
    n = g["soma"].shape[0]
    got, _ = run_gpu([maker()], n, precision=64, probes="all")
    w0 = int(g["window_start"])
    e_all = np.abs(got[w0:] - g["all_cvs"]).max()
    e_soma = np.abs(got[:, int(g["soma_cv"])] - g["soma"]).max()
    print(f"{name}: fp64 max |dV| all CVs {e_all:.2e}, soma {e_soma:.2e} mV")
    assert e_all < 1e-9 and e_soma < 1e-9


@pytest.mark.parametrize("name,maker", TRACES)
def test_trace_fixture_fp32(cuda, name, maker):
    g = np.load(os.path.join(HERE, name))
    n = g["soma"].shape[0]
    got, _ = run_gpu([maker()], n, precision=32, probes="all")
    e = rmse(got[:, int(g["soma_cv"])], g["soma"])
    w0 = int(g["window_start"])
    e_max = np.abs(got[w0:] - g["all_cvs"]).max()
    print(f"{name}: fp32 soma RMSE {e:.2e} mV, max |dV| in the window {e_max:.2e} mV")
    assert e <= 0.05 and e_max <= 0.05


@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 2e-4)])
def test_hines7_one_step(cuda, precision, tol):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    g = np.load(os.path.join(HERE, "hines7.npz"))
    L = StaggerLayout([MG.toy7_model()], probes="all")
    rt = StaggerRuntime(L, float(g["dt"]), precision=precision)
    rt.V.copy_(torch.as_tensor(g["V0"], dtype=rt.V.dtype, device=rt.V.device))
    tr = rt.run(1)
    torch.cuda.synchronize()
    rt.check_status()
    got = tr.double().cpu().numpy()[0]
    err = np.abs(got - g["V1"]).max()
    print(f"hines7 fp{precision}: max |dV| {err:.2e} mV")
    assert err < tol
