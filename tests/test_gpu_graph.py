"""Captured SNAP energy/force step (HIP graph replay over capacity-mode
neighbor lists) against the uncaptured call on the same inputs: the
results must be bit-identical on the capturing call, on replays, after
positions change in place, and after a pair-count overflow forces an exact
recomputation and a new capture."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PATHS = {"plain": 0, "adjoint": 1, "adjoint_fp32": 2}


def setup(ipg, ncell=5, seed=11, J=8):
    from ipgpu import workloads
    from ipgpu.device import DeviceArray, dev_system

    s = workloads.snap_bcc(ncell=ncell, a=3.16, jitter=0.05, nspecies=2, seed=seed)
    names = ["W", "Mo"]
    sn = ipg.snap.SNAP(len(s), J, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    beta = np.random.default_rng(seed).normal(scale=0.1, size=2 * sn.ncoeff)
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    d = {
        "pos": DeviceArray.from_host(s.positions),
        "sp": DeviceArray.from_host(sp),
        "beta": DeviceArray.from_host(beta),
        "e": DeviceArray((1,), np.float64),
        "f": DeviceArray((len(s), 3), np.float64),
    }
    dsys = dev_system(d["pos"], d["sp"], len(s), s.box, s.periodic)
    return s, sn, d, dsys


def step(ipg, sn, d, dsys, path):
    from ipgpu._lib import check

    check(ipg.load().ipg_snap_energy_forces_dev(sn._h, C.byref(dsys), 1, C.c_void_p(d["beta"].ptr), 0.25,
                                                PATHS[path], C.c_void_p(d["e"].ptr), C.c_void_p(d["f"].ptr)))
    return d["e"].download()[0], d["f"].download()


def uncaptured(ipg, sn, d, dsys, path):
    from ipgpu import device

    device.set_graphs(False)
    try:
        return step(ipg, sn, d, dsys, path)
    finally:
        device.set_graphs(True)


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("path", ["adjoint_fp32", "adjoint", "plain"])
def test_graph_replay_bit_identical(ipg, path):
    s, sn, d, dsys = setup(ipg)
    ref = uncaptured(ipg, sn, d, dsys, path)
    assert np.isfinite(ref[0]) and np.abs(ref[1]).max() > 0
    got = step(ipg, sn, d, dsys, path)  # exact run, then capture
    info = sn.graph_info()
    assert info["state"] == "ready", info
    assert same(got, ref)
    for _ in range(3):  # replays
        assert same(step(ipg, sn, d, dsys, path), ref)
    assert sn.graph_info()["replays"] == 3
    # positions changed in place: the graph reads the new coordinates
    rng = np.random.default_rng(3)
    X2 = s.positions + rng.normal(scale=0.03, size=s.positions.shape)
    d["pos"].upload(X2)
    got2 = step(ipg, sn, d, dsys, path)
    ref2 = uncaptured(ipg, sn, d, dsys, path)
    assert not same(ref2, ref)
    assert same(got2, ref2)
    info = sn.graph_info()
    assert info["replays"] == 4 and info["overflows"] == 0 and info["captures"] == 1, info


def test_graph_overflow_recomputes_exactly(ipg):
    s, sn, d, dsys = setup(ipg, seed=12)
    path = "adjoint_fp32"
    step(ipg, sn, d, dsys, path)  # capture at the lattice's pair counts
    assert sn.graph_info()["state"] == "ready"
    # squeeze the atoms toward the box centre: far more pairs than the capacities
    c = 0.5 * np.asarray(s.box)
    X3 = c + 0.7 * (s.positions - c)
    d["pos"].upload(X3)
    got = step(ipg, sn, d, dsys, path)  # replay overflows -> exact run -> new capture
    info = sn.graph_info()
    assert info["overflows"] == 1 and info["captures"] == 2, info
    ref3 = uncaptured(ipg, sn, d, dsys, path)
    assert same(got, ref3)
    assert same(step(ipg, sn, d, dsys, path), ref3)  # replay of the new capture
    assert sn.graph_info()["replays"] == 2


def test_graph_with_profiling_events(ipg):
    from ipgpu import device

    s, sn, d, dsys = setup(ipg, seed=13)
    path = "adjoint_fp32"
    ref = uncaptured(ipg, sn, d, dsys, path)
    device.prof_enable(True)
    try:
        device.prof_reset()
        for _ in range(3):
            assert same(step(ipg, sn, d, dsys, path), ref)
        info = sn.graph_info()
        assert info["state"] == "ready", info
        n, ms = device.prof_query("snap_zy_fp32")
        assert n == 3 and ms > 0.0, (n, ms)
    finally:
        device.prof_enable(False)


def test_prof_select_times_only_the_named_kernels(ipg):
    """Events only around the selected kernel, captured and uncaptured; the
    results are those of the unprofiled step."""
    from ipgpu import device

    s, sn, d, dsys = setup(ipg, seed=14)
    path = "adjoint_fp32"
    ref = uncaptured(ipg, sn, d, dsys, path)
    try:
        device.prof_select(["snap_zy_fp32"])
        device.prof_enable(True)
        for graphs in (False, True):
            device.set_graphs(graphs)
            device.prof_reset()
            for _ in range(3):
                assert same(step(ipg, sn, d, dsys, path), ref)
            n, ms = device.prof_query("snap_zy_fp32")
            assert n == 3 and ms > 0.0, (graphs, n, ms)
            assert device.prof_query("snap_ui_fp32")[0] == 0
            assert device.prof_query("snap_dedr_fp32")[0] == 0
        device.prof_select(None)
        device.prof_reset()
        assert same(step(ipg, sn, d, dsys, path), ref)
        assert device.prof_query("snap_ui_fp32")[0] == 1
    finally:
        device.prof_enable(False)
        device.prof_select(None)
        device.set_graphs(True)


@pytest.mark.parametrize("path", ["adjoint_fp32", "plain"])
def test_graph_replay_cell_overflow(ipg, path):
    """1024 atoms, five cells a side: the SNAP list takes the cell-pair row
    kernel.  Moving 70 atoms in place into one cell overfills that kernel's
    per-cell buffers on the next replay; the replay's flag word sends the
    call to the exact (binned) rebuild, whose result equals the uncaptured
    call's bit for bit, and the replays after it stay correct."""
    s, sn, d, dsys = setup(ipg, ncell=8, seed=5)
    ref = uncaptured(ipg, sn, d, dsys, path)
    assert same(step(ipg, sn, d, dsys, path), ref)  # exact run, then capture
    assert same(step(ipg, sn, d, dsys, path), ref)  # replay
    rng = np.random.default_rng(9)
    X2 = s.positions.copy()
    X2[:70] = np.array([11.0, 11.0, 11.0]) + rng.random((70, 3)) * 3.0
    d["pos"].upload(X2)
    got = step(ipg, sn, d, dsys, path)
    ref2 = uncaptured(ipg, sn, d, dsys, path)
    assert same(got, ref2)
    assert sn.graph_info()["overflows"] >= 1
    assert same(step(ipg, sn, d, dsys, path), ref2)
