"""Verlet-skin lists on the GPU: an MD-style loop of small random moves,
each step's energy, forces and virial from the skin list against a fresh
evaluation that rebuilds the list (the reference's behaviour), and the
list's coverage of every pair within the cutoff.

Tolerance: 1e-12 relative (FP64; only the summation order differs from the
fresh list's, because the skin list holds extra pairs beyond the cutoff)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def list_pairs(ipg, h, n):
    """(i, j) pairs of a device list, as a set."""
    from ipgpu import _lib

    L = _lib.load()
    npairs = C.c_int64()
    off_p, nbr_p = C.c_void_p(), C.c_void_p()
    _lib.check(L.ipg_nlist_info(h, None, C.byref(npairs), C.byref(off_p), C.byref(nbr_p)))
    off = np.zeros(n + 1, dtype=np.int64)
    nbr = np.zeros(max(npairs.value, 1), dtype=np.int32)
    ctx = _lib.default_context()
    _lib.check(L.ipg_memcpy_d2h(ctx, off.ctypes.data, off_p, off.nbytes))
    _lib.check(L.ipg_memcpy_d2h(ctx, nbr.ctypes.data, nbr_p, 4 * npairs.value))
    i = np.repeat(np.arange(n), np.diff(off))
    return set(zip(i.tolist(), nbr[:npairs.value].tolist()))


def systems(ipg):
    from ipgpu import empirical as E
    from ipgpu import workloads
    from ipgpu.system import angstrom, eV

    s, lj = workloads.lj_argon_fcc(seed=7)
    yield "lj", s, lj
    s2, _ = workloads.mixed_pairs(n_side=(14, 14, 10), seed=8)
    mix = (E.LennardJones(0.0104 * eV, 2.0 * angstrom, 6.0 * angstrom, ("A",))
           + E.Morse(0.1 * eV, 1.5 / angstrom, 2.3 * angstrom, 5.0 * angstrom, ("A", "B"))
           + E.ZBL(18, 54, 1.0, 4.0 * angstrom, ("B",)))
    yield "mixed", s2, mix


@pytest.mark.parametrize("case", ["lj", "mixed"])
def test_verlet_loop_matches_fresh_lists(ipg, orc, case):
    from ipgpu.system import AtomsSystem

    name, s, pot = [x for x in systems(ipg) if x[0] == case][0]
    skin = 1.0  # bohr
    ev = ipg.empirical.VerletEvaluator(s, pot, skin)
    X = s.austrip_positions().copy()
    box = s.austrip_box()
    rng = np.random.default_rng(3)
    rebuilt_steps = []
    for step in range(14):
        X = X + rng.normal(scale=0.06, size=X.shape)
        e, f, v, rebuilt = ev.step(X)
        if rebuilt:
            rebuilt_steps.append(step)
        cur = AtomsSystem(X, box, s.periodic, s.symbols, unit="bohr")
        e0, f0 = ipg.empirical.energy_and_force(cur, pot)
        v0 = ipg.empirical.virial_stress(cur, pot)
        assert abs(e - e0) <= 1e-12 * abs(e0), (step, e, e0)
        assert rel(f, f0) < 1e-12, step
        assert rel(v, v0) < 1e-12, step
        # every pair within the largest cutoff is in the skin list
        rc = pot.get_rcutoff()
        off, j, _ = orc.neighborlist(X, [b if p else np.inf for b, p in zip(box, s.periodic)], rc,
                                     kind=orc.LIST_FULL, cells=True)
        need = set(zip(np.repeat(np.arange(len(X)), np.diff(off)).tolist(), j.tolist()))
        assert need <= list_pairs(ipg, ev._h, len(X)), step
    # the walk moves atoms past skin/2 a few times, not on every step
    assert 1 <= len(rebuilt_steps) < 14, rebuilt_steps
    assert ev.builds == 1 + len(rebuilt_steps)


def test_verlet_box_change_rebuilds(ipg):
    from ipgpu import _lib
    from ipgpu import workloads
    from ipgpu.device import dev_system

    s, pot = workloads.lj_argon_fcc(seed=9)
    ev = ipg.empirical.VerletEvaluator(s, pot, 0.8)
    X = s.austrip_positions()
    e1, f1, _, r1 = ev.step(X)
    assert not r1
    box2 = s.austrip_box() * 1.01
    d2 = dev_system(ev.pos, ev._sp, ev.n, box2, s.periodic)
    r = C.c_int32()
    _lib.check(_lib.load().ipg_nlist_update_dev(ev._h, C.byref(d2), C.byref(r)))
    assert r.value == 1


def test_verlet_rejects_plain_list(ipg):
    from ipgpu import _lib
    from ipgpu import workloads
    from ipgpu.device import DeviceArray, dev_system

    s, pot = workloads.lj_argon_fcc(seed=9)
    pos = DeviceArray.from_host(s.austrip_positions())
    d = dev_system(pos, None, len(s), s.austrip_box(), s.periodic)
    L = _lib.load()
    h = C.c_void_p()
    _lib.check(L.ipg_nlist_build_dev(_lib.default_context(), C.byref(d), pot.get_rcutoff(), 1, C.byref(h)))
    try:
        assert L.ipg_nlist_update_dev(h, C.byref(d), None) == _lib.IPG_ERR_INVALID
    finally:
        L.ipg_nlist_destroy(h)
