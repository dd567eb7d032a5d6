"""Batched SNAP descriptors (many configurations in one call) against the
single-configuration GPU path and the oracle.

Tolerances: FP64 batch vs single-configuration FP64 within 1e-12 relative
(same arithmetic, only the launch shapes may differ), vs the oracle within
1e-10 relative; FP32 batch vs FP64 within 2e-5 relative to the largest
entry of B and of dB.  Relative errors are max |a - b| / max |b|."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["W", "Mo"]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make_snap(ipg, J=6, switch=True, bzero=False):
    return ipg.snap.SNAP(100, J, NAMES, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=switch,
                         bzero_flag=bzero, list_mode="cutoff")


def configs(ipg, n=6, nmin=5, nmax=120, seed=11):
    from ipgpu import workloads

    return workloads.snap_batch(nconfig=n, nmin=nmin, nmax=nmax, rcut=5.0, seed=seed)


def single(ipg, g, systems, with_w=False):
    Bs, dBs, Ws = [], [], []
    for s in systems:
        B, dB, W = ipg.snap.compute_all_descriptors(s, g, list_mode="cutoff")
        Bs.append(B)
        dBs.append(dB)
        Ws.append(W)
    out = (np.concatenate(Bs), np.concatenate(dBs))
    return out + (np.stack(Ws),) if with_w else out


def erow_of(g, systems, B, start):
    nc = g.ncoeff
    out = np.zeros((len(systems), nc * len(NAMES)))
    for c, s in enumerate(systems):
        sp = np.array([NAMES.index(x) for x in s.symbols])
        Bc = B[start[c]:start[c + 1]]
        for k in range(len(NAMES)):
            out[c, k * nc:(k + 1) * nc] = Bc[sp == k].sum(0)
    return out


@pytest.mark.parametrize("J,switch,bzero", [(6, True, False), (8, True, True), (3, False, False)])
def test_batch_fp64_matches_single(ipg, J, switch, bzero):
    g = make_snap(ipg, J, switch, bzero)
    systems = configs(ipg, seed=10 + J)
    r = ipg.snap.compute_batch_descriptors(systems, g, precision="fp64")
    B, dB, W = single(ipg, g, systems, with_w=True)
    assert r.B.shape == B.shape and r.dB.shape == dB.shape and r.W.shape == W.shape
    assert rel(r.B, B) < 1e-12
    assert rel(r.dB, dB) < 1e-12
    assert rel(r.W, W) < 1e-12
    assert rel(r.erow, erow_of(g, systems, B, r.start)) < 1e-12
    info = ipg.snap.batch_info(g, "fp64")
    assert info["scan_lanes"] in (16, 32, 64)
    assert info["ui_atoms_per_wave"] in (1, 2)
    assert info["dbdr_threads"] in (64, 128)


def test_batch_vs_oracle(ipg, orc):
    J = 5
    systems = configs(ipg, n=3, nmin=20, nmax=60, seed=21)
    g = make_snap(ipg, J)
    o = orc.Snap(J, 2, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=True)
    r = ipg.snap.compute_batch_descriptors(systems, g)
    for c, s in enumerate(systems):
        sp = np.array([NAMES.index(x) for x in s.symbols], dtype=np.int32)
        off, nbr, R, rv = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
        Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, rv)
        a, b = r.start[c], r.start[c + 1]
        assert rel(r.B[a:b], Bo) < 1e-10
        assert rel(r.dB[a:b], dBo.transpose(0, 2, 1)) < 1e-10
        assert rel(r.W[c], Wo.T) < 1e-10


@pytest.mark.parametrize("J", [6, 8])
def test_batch_fp32_close_to_fp64(ipg, J):
    g = make_snap(ipg, J)
    systems = configs(ipg, n=8, nmin=30, nmax=300, seed=30 + J)
    r64 = ipg.snap.compute_batch_descriptors(systems, g, precision="fp64")
    r32 = ipg.snap.compute_batch_descriptors(systems, g, precision="fp32")
    eB, edB, er, eW = rel(r32.B, r64.B), rel(r32.dB, r64.dB), rel(r32.erow, r64.erow), rel(r32.W, r64.W)
    print("fp32 vs fp64: B %.3g dB %.3g erow %.3g W %.3g" % (eB, edB, er, eW))
    assert eB < 2e-5 and edB < 2e-5 and er < 2e-5 and eW < 2e-5


def test_batch_replay_bit_identical(ipg):
    g = make_snap(ipg, 6)
    systems = configs(ipg, n=5, seed=41)
    for prec in ("fp64", "fp32"):
        r1 = ipg.snap.compute_batch_descriptors(systems, g, precision=prec)
        r2 = ipg.snap.compute_batch_descriptors(systems, g, precision=prec)
        r3 = ipg.snap.compute_batch_descriptors(systems, g, precision=prec)
        for a, b in ((r1, r2), (r1, r3)):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        info = ipg.snap.batch_info(g, prec)
        assert info["graph"] == "ready", info
    assert ipg.snap.batch_info(g, "fp32")["replays"] >= 2


def test_batch_edge_configs(ipg):
    """Empty, single-atom and non-periodic configurations between ordinary ones."""
    from ipgpu.system import AtomsSystem

    g = make_snap(ipg, 4)
    base = configs(ipg, n=3, nmin=20, nmax=80, seed=51)
    rng = np.random.default_rng(5)
    one = AtomsSystem(np.array([[1.0, 2.0, 3.0]]), [10.0, 10.0, 10.0], (True, True, True), ["Mo"], unit="angstrom")
    X = rng.uniform(0, 7.0, size=(25, 3))
    free = AtomsSystem(X, [0.0, 0.0, 0.0], (False, False, False), ["W"] * 25, unit="angstrom")
    empty = AtomsSystem(np.zeros((0, 3)), [10.0, 10.0, 10.0], (True, True, True), [], unit="angstrom")
    systems = [base[0], empty, one, base[1], free, base[2]]
    r = ipg.snap.compute_batch_descriptors(systems, g)
    B, dB, W = single(ipg, g, systems, with_w=True)
    assert list(np.diff(r.start)) == [len(s) for s in systems]
    assert rel(r.B, B) < 1e-12 and rel(r.dB, dB) < 1e-12 and rel(r.W, W) < 1e-12
    assert np.all(r.erow[1] == 0.0) and np.all(r.W[1] == 0.0)
    assert rel(r.erow, erow_of(g, systems, B, r.start)) < 1e-12


def test_batch_layout_change_and_overflow(ipg):
    """A different batch, then the first again, then the first with one
    configuration compressed (more neighbors than the captured capacities)."""
    from ipgpu.system import AtomsSystem

    g = make_snap(ipg, 6)
    A = configs(ipg, n=4, nmin=40, nmax=150, seed=61)
    Bb = configs(ipg, n=7, nmin=10, nmax=90, seed=62)
    for systems in (A, Bb, A):
        r = ipg.snap.compute_batch_descriptors(systems, g)
        B, dB = single(ipg, g, systems)
        assert rel(r.B, B) < 1e-12 and rel(r.dB, dB) < 1e-12
    s0 = A[0]
    dense = AtomsSystem(s0.positions * 0.8, [x * 0.8 for x in s0.box], s0.periodic, s0.symbols, unit="angstrom")
    A2 = [dense] + A[1:]
    r = ipg.snap.compute_batch_descriptors(A2, g)
    B, dB = single(ipg, g, A2)
    assert rel(r.B, B) < 1e-12 and rel(r.dB, dB) < 1e-12


def test_batch_device_entry_point(ipg):
    from ipgpu import _lib
    from ipgpu.device import DeviceArray

    g = make_snap(ipg, 6)
    systems = configs(ipg, n=4, nmin=30, nmax=100, seed=71)
    ref = ipg.snap.compute_batch_descriptors(systems, g)
    n = int(ref.start[-1])
    M = g.ncoeff * 2
    pos = DeviceArray.from_host(np.concatenate([s.positions for s in systems]))
    sp = DeviceArray.from_host(np.concatenate([[NAMES.index(x) for x in s.symbols] for s in systems]).astype(np.int32))
    natoms = np.array([len(s) for s in systems], dtype=np.int64)
    box = np.array([s.box for s in systems], dtype=np.float64)
    per = np.array([[1 if p else 0 for p in s.periodic] for s in systems], dtype=np.int32)
    B = DeviceArray((n, g.ncoeff), np.float64)
    dB = DeviceArray((n, M, 3), np.float64)
    er = DeviceArray((len(systems), M), np.float64)
    W = DeviceArray((len(systems), M, 6), np.float64)
    L = _lib.load()
    for _ in range(2):
        _lib.check(L.ipg_snap_batch_descriptors_dev(g._h, len(systems), natoms.ctypes.data, box.ctypes.data,
                                                    per.ctypes.data, C.c_void_p(pos.ptr), C.c_void_p(sp.ptr),
                                                    _lib.IPG_FP64, C.c_void_p(B.ptr), C.c_void_p(dB.ptr),
                                                    C.c_void_p(er.ptr), C.c_void_p(W.ptr)))
        _lib.check(L.ipg_ctx_synchronize(_lib.default_context()))
        assert np.array_equal(B.download(), ref.B)
        assert np.array_equal(dB.download().transpose(0, 2, 1), ref.dB)
        assert np.array_equal(er.download(), ref.erow)
        assert np.array_equal(W.download().transpose(0, 2, 1), ref.W)


def test_batch_device_species_null_is_species_zero(ipg):
    """species NULL on the device entry point means every atom has species 0."""
    from ipgpu import _lib
    from ipgpu.device import DeviceArray

    from ipgpu.system import AtomsSystem

    g = ipg.snap.SNAP(100, 6, ["W"], 2.5, 0.0, 0.99363, [1.0], [1.0], switch_flag=True, list_mode="cutoff")
    systems = [AtomsSystem(s.positions, s.box, s.periodic, ["W"] * len(s), unit="angstrom")
               for s in configs(ipg, n=3, nmin=30, nmax=90, seed=81)]
    n = sum(len(s) for s in systems)
    pos = DeviceArray.from_host(np.concatenate([s.positions for s in systems]))
    zeros = DeviceArray.from_host(np.zeros(n, np.int32))
    natoms = np.array([len(s) for s in systems], dtype=np.int64)
    box = np.array([s.box for s in systems], dtype=np.float64)
    per = np.ones((len(systems), 3), dtype=np.int32)
    L = _lib.load()
    out = []
    for sp in (None, C.c_void_p(zeros.ptr)):
        B = DeviceArray((n, g.ncoeff), np.float64)
        dB = DeviceArray((n, g.ncoeff, 3), np.float64)
        _lib.check(L.ipg_snap_batch_descriptors_dev(g._h, len(systems), natoms.ctypes.data, box.ctypes.data,
                                                    per.ctypes.data, C.c_void_p(pos.ptr), sp, _lib.IPG_FP64,
                                                    C.c_void_p(B.ptr), C.c_void_p(dB.ptr), None, None))
        _lib.check(L.ipg_ctx_synchronize(_lib.default_context()))
        out.append((B.download(), dB.download()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    Bs = np.concatenate([ipg.snap.compute_local_descriptors(s, g, list_mode="cutoff") for s in systems])
    assert rel(out[0][0], Bs) < 1e-12
