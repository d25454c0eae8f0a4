"""Run under an AddressSanitizer runtime (LD_PRELOAD) by test_sanitizers.py.

``abi``: the ASan build of the C-ABI library; every host-side validation
path of the entry points is exercised with bad and complete descriptors (no
GPU is needed: every call returns before a kernel launch, or fails at the
HIP runtime's own device check).
``port <out.npz>``: the ASan/UBSan build of the FP64 C port on two cells of
the 4096-tree workload, both back-substitution forms, results saved for the
parent to compare with the plain build.
"""

import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def abi():
    import braincell_amd._native as N
    from braincell_amd.build import asan_library_path

    N.library_path = asan_library_path
    L = N.lib()
    rc = []
    d = N.StaggerDesc()
    rc.append(L.bc_stagger_run(None, None))  # NULL desc
    d.precision = 16
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))
    d.precision, d.threads = 32, 100
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))
    d.threads, d.max_rows = 256, 0
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))  # INVALID_LAYOUT
    d.max_rows, d.max_chains, d.max_cls, d.max_levels, d.dt = 64, 8, 1, 1, 0.025
    d.flags = 0x80
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))  # unknown flags
    d.flags = 0
    d.channels, d.cv_per_thread = 0x117, 3
    cc = N.chan_const_type(32)()
    d.chan_const = ctypes.cast(ctypes.pointer(cc), ctypes.c_void_p)
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))  # UNSUPPORTED cv_per_thread
    d.cv_per_thread = 8
    k = N.SpikeDesc()
    k.cap = -1
    d.spk = ctypes.cast(ctypes.pointer(k), ctypes.c_void_p)
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))
    d.spk = None
    d.n_packs = 4  # pointers missing
    rc.append(L.bc_stagger_run(ctypes.byref(d), None))
    d.n_packs = 0
    out = ctypes.c_void_p()
    rc.append(L.bc_stagger_graph_create(ctypes.byref(d), 0, 200, None, ctypes.byref(out)))
    rc.append(L.bc_stagger_graph_create(ctypes.byref(d), 1, 200, None, None))
    rc.append(L.bc_graph_launch(None, None))
    rc.append(L.bc_graph_destroy(None))
    rc.append(L.bc_stagger_adjoint(None, None))
    ad = N.AdjointDesc()
    rc.append(L.bc_stagger_adjoint(ctypes.byref(ad), None))
    from braincell_amd.single import PointDesc

    pd = PointDesc()
    pd.precision, pd.n, pd.n_steps, pd.dt, pd.Cm, pd.method = 64, 0, 1, 0.025, 1.0, 7
    rc.append(L.bc_point_run(ctypes.byref(pd), None))
    pd.method = 0
    rc.append(L.bc_point_run(ctypes.byref(pd), None))  # n = 0: nothing to launch
    from braincell_amd.closure import _AdamDesc, _ApplyDesc, _DataDesc, _FwdDesc, _TrainDesc

    for fn, desc in ((L.bc_closure_data, _DataDesc()), (L.bc_closure_grad, _TrainDesc()),
                     (L.bc_closure_grad_noalloc, _TrainDesc()), (L.bc_closure_adam, _AdamDesc()),
                     (L.bc_closure_forward, _FwdDesc()), (L.bc_closure_apply, _ApplyDesc())):
        rc.append(fn(None, None))
        rc.append(fn(ctypes.byref(desc), None))
    t = _TrainDesc(40000, 0, *([ctypes.c_void_p(16)] * 5), None, None, ctypes.c_void_p(16), ctypes.c_void_p(16))
    rc.append(L.bc_closure_grad_noalloc(ctypes.byref(t), None))  # beyond the scratch array
    L.bc_last_error.restype = ctypes.c_char_p
    assert b"exceeds" in L.bc_last_error()
    print("abi return codes", rc, flush=True)
    assert rc[:4] == [1, 1, 1, 2] and rc[4] == 1 and rc[5] == 5 and rc[6] == 1 and rc[7] == 1, rc
    assert rc[8] == 1 and rc[9] == 1 and rc[10] == 1 and rc[11] == 0 and rc[12] == 1 and rc[13] == 1, rc
    assert rc[14] == 1 and rc[15] == 0, rc
    assert all(r == 1 for r in rc[16:]), rc


def port(out):
    import numpy as np

    import oracle.cport_run as CR
    from braincell_amd import workloads as W
    from oracle.build_cport import ASAN_OUT
    from tests.helpers import soma_cv

    CR.LIB = ASAN_OUT
    models = [W.uneven_tree_cell(i) for i in (27, 49)]
    b = CR.PortBatch(models, 0.025, [soma_cv(m) for m in models])
    seq = b.run(400, sequential=True, threads=2)
    b2 = CR.PortBatch(models, 0.025, [soma_cv(m) for m in models])
    rd = b2.run(400, sequential=False, threads=2)
    np.savez(out, seq=seq, rd=rd)
    print("port ok", flush=True)


if __name__ == "__main__":
    if sys.argv[1] == "abi":
        abi()
    else:
        port(sys.argv[2])
