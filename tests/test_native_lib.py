"""The C-ABI library loads, exports what include/braincell_amd.h declares and
rejects bad descriptors before touching the GPU (CPU only)."""

import ctypes
import os
import re

import pytest

from braincell_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bc_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = declared_functions()
    assert "bc_stagger_run" in names and len(names) >= 7
    for n in names:
        assert hasattr(L, n), n


def test_version_and_support_queries():
    L = _native.lib()
    assert L.bc_version() >= 100
    assert L.bc_stagger_supported(32, 0x117, 8) == 1
    assert L.bc_stagger_supported(64, 0x7, 1) == 1
    assert L.bc_stagger_supported(16, 0x7, 1) == 0
    assert L.bc_stagger_supported(32, 0x7, 3) == 0


def test_bad_descriptors_are_rejected_without_gpu():
    L = _native.lib()
    assert L.bc_stagger_run(None, None) == 1
    assert b"NULL" in L.bc_last_error()
    d = _native.StaggerDesc()
    d.precision = 16
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"precision" in L.bc_last_error()
    d.precision = 32
    d.threads = 100
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    d.threads = 256
    d.max_rows = 0
    assert L.bc_stagger_run(ctypes.byref(d), None) == 2
    with pytest.raises(_native.NativeError):
        _native.check(L.bc_stagger_run(ctypes.byref(d), None))
    assert L.bc_graph_launch(None, None) == 1
    assert L.bc_graph_destroy(None) == 0


def test_point_descriptor_method_is_checked_without_gpu():
    from braincell_amd.single import PointDesc

    L = _native.lib()
    d = PointDesc()
    d.precision, d.n, d.n_steps, d.dt, d.Cm = 64, 0, 1, 0.025, 1.0
    d.method = 7
    assert L.bc_point_run(ctypes.byref(d), None) == 1
    assert b"method" in L.bc_last_error()
    d.method = 2
    assert L.bc_point_run(ctypes.byref(d), None) == 0  # n = 0: nothing to launch


def test_spike_descriptor_is_checked_without_gpu():
    L = _native.lib()
    d = _native.StaggerDesc()
    d.precision, d.threads, d.max_rows, d.max_chains, d.max_cls, d.max_levels, d.dt = 32, 256, 64, 8, 1, 1, 0.025
    k = _native.SpikeDesc()
    k.cap = -1
    d.spk = ctypes.cast(ctypes.pointer(k), ctypes.c_void_p)
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"cap" in L.bc_last_error()
    k.cap = 4
    d.trace_ld = 0
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"trace" in L.bc_last_error()
    d.trace_ld = 3
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"NULL" in L.bc_last_error()
    buf = (ctypes.c_int32 * 64)()
    k.v_last = k.count = k.steps = ctypes.cast(buf, ctypes.c_void_p)
    d.channels, d.cv_per_thread = 0x117, 8
    cc = _native.chan_const_type(32)()
    d.chan_const = ctypes.cast(ctypes.pointer(cc), ctypes.c_void_p)
    # a complete descriptor passes validation (without a GPU the call then
    # fails at the runtime's error check, not with INVALID_ARG)
    assert L.bc_stagger_run(ctypes.byref(d), None) in (0, 3)


def test_spike_cap_bounds():
    from braincell_amd.cell import RunResult, spike_cap

    assert spike_cap(1, 0.025) == 1
    assert spike_cap(40000, 0.025) == 508
    assert spike_cap(10, 1.0) == 5
    assert RunResult(time=None, traces={}).spikes == {}


def test_synapse_descriptor_is_checked_without_gpu():
    """The synapse descriptor's LDS plan (max_syn, max_syn_rows) and the
    drive workspace are checked before any launch."""
    L = _native.lib()
    d = _native.StaggerDesc()
    d.precision, d.threads, d.max_rows, d.max_chains, d.max_cls, d.max_levels, d.dt = 32, 64, 64, 8, 1, 1, 0.025
    y = _native.SynapseDesc()
    buf = (ctypes.c_int32 * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    y.pack_off = y.row = y.row_off = y.par = y.state = p
    d.syn = ctypes.cast(ctypes.pointer(y), ctypes.c_void_p)
    y.n_syn = -1
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"n_syn" in L.bc_last_error()
    y.n_syn, y.max_syn, y.max_syn_rows = 10, 0, 0
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    assert b"max_syn" in L.bc_last_error()
    y.max_syn, y.max_syn_rows = 4, 5  # more rows than synapses
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    y.max_syn, y.max_syn_rows = 11, 2  # more than n_syn
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    y.max_syn = 4
    y.ext_slot, y.n_ext, y.ext_steps, y.ext_drive = p, 1, 1, p
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1  # external drive without a drive workspace
    assert b"drive" in L.bc_last_error()
    y.drive = p
    d.channels, d.cv_per_thread = 0x604, 1
    cc = _native.chan_const_type(32)()
    d.chan_const = ctypes.cast(ctypes.pointer(cc), ctypes.c_void_p)
    assert L.bc_stagger_run(ctypes.byref(d), None) in (0, 3)  # valid (no packs: nothing to launch)


def test_sequential_back_substitution_is_fp32_without_synapses():
    # BC_STEP_BACKSUB_SWEEP picks the kernels compiled with the sequential
    # back substitution: FP32 without synapses; other descriptors get
    # BC_UNSUPPORTED before any device work (no packs to launch here)
    L = _native.lib()
    d = _native.StaggerDesc()
    d.precision, d.threads, d.max_rows, d.max_chains, d.max_cls, d.max_levels, d.dt = 32, 256, 64, 8, 1, 1, 0.025
    d.channels, d.cv_per_thread, d.flags = 0x117, 8, 8
    cc = _native.chan_const_type(32)()
    d.chan_const = ctypes.cast(ctypes.pointer(cc), ctypes.c_void_p)
    # supported: past validation and kernel choice (without a GPU the call then
    # fails at the runtime's error check)
    assert L.bc_stagger_run(ctypes.byref(d), None) in (0, 3)
    d.flags = 16  # an unknown bit
    assert L.bc_stagger_run(ctypes.byref(d), None) == 1
    d.flags = 8
    d.precision = 64
    cc64 = _native.chan_const_type(64)()
    d.chan_const = ctypes.cast(ctypes.pointer(cc64), ctypes.c_void_p)
    assert L.bc_stagger_run(ctypes.byref(d), None) == 5
    assert b"sequential back substitution" in L.bc_last_error()
