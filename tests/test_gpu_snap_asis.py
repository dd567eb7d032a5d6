"""GPU SNAP on the reference's as-is list semantics, against the oracle and
the hand-derived list of tests/golden/snap_list_hand.json.

- REFERENCE-mode lists (utl.jl:38-101 as-is) are not symmetric when the
  species radii differ: B, dB, W and the plain, adjoint and FP32 adjoint
  forces over such lists go through the general reverse rows.
- CUTOFF-mode lists with unequal radii, and a CUTOFF list whose pair
  j -> i is missing while i -> j is kept (r.r within an ulp of the cutoff,
  the two directions rounding differently through mod): the forces use the
  reverse-row fallback, on the exact path and after a graph replay.

Tolerances: B, dB, W within 1e-10 relative (to the largest entry), FP64
energy and forces within 1e-10, FP32 adjoint forces within 1e-4 and energy
within 1e-5 relative; lists bit for bit."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

INF = np.inf


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def first_appearance(symbols):
    seen = []
    for s in symbols:
        if s not in seen:
            seen.append(s)
    return np.array([seen.index(s) for s in symbols], dtype=np.int32)


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
def test_hand_derived_snap_list(ipg, mode):
    with open(os.path.join(GOLDEN, "snap_list_hand.json")) as f:
        v = json.load(f)
    s = ipg.AtomsSystem(v["positions"], v["box"], v["periodic"], v["symbols"], unit="angstrom")
    g = ipg.snap.SNAP(len(s), 2, v["snap_species"], v["rcutfac"], 0.0, 0.99363, v["radii"], [1.0, 1.0],
                      list_mode=mode)
    nl = ipg.snap.snap_neighborlist(s, g)
    want = v[mode]
    off = nl.offsets
    assert [list(nl.j[off[i]:off[i + 1]] + 1) for i in range(len(s))] == want["j_1based"]
    np.testing.assert_array_equal(nl.r, np.array([x for row in want["r"] for x in row]).reshape(-1, 3))
    np.testing.assert_array_equal(nl.R, np.sqrt([x for row in want["rsq"] for x in row]))


def dense_case(ipg, orc, mode, J, seed=11, n=120, switch=True):
    """Random atoms in a 6 A box, three species with radii 0.3, 1.5, 0.9 and
    rcutfac 1: the per-species cutoffs split the inrange result, so the
    REFERENCE list's element lookups (utl.jl:70,86) keep pairs the CUTOFF
    list drops and the other way round, and REFERENCE rows are asymmetric."""
    rng = np.random.default_rng(seed)
    L = np.array([6.0, 6.0, 6.0])
    X = rng.random((n, 3)) * L
    names = ["W", "Mo", "Ta"]
    radii, weight = [0.3, 1.5, 0.9], [1.0, 0.7, 1.3]
    syms = [names[k] for k in rng.integers(0, 3, n)]
    s = ipg.AtomsSystem(X, L, (True,) * 3, syms, unit="angstrom")
    g = ipg.snap.SNAP(n, J, names, 1.0, 0.0, 0.99363, radii, weight, switch_flag=switch, list_mode=mode)
    o = orc.Snap(J, 3, 1.0, 0.0, 0.99363, radii, weight, switch_flag=switch)
    sp = np.array([names.index(x) for x in syms], dtype=np.int32)
    ap = first_appearance(syms)
    om = orc.SNAP_LIST_REFERENCE if mode == "reference" else orc.SNAP_LIST_CUTOFF
    off, nbr, R, r = o.neighborlist(X, L, sp, appear=ap, mode=om)
    return s, g, o, sp, (off, nbr, r)


def asymmetric_rows(off, nbr):
    rows = [set(nbr[off[k]:off[k + 1]].tolist()) for k in range(len(off) - 1)]
    return sum(1 for k in range(len(rows)) for m in rows[k] if k not in rows[m])


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
@pytest.mark.parametrize("J", [4, 8])
def test_unequal_radii_descriptors(ipg, orc, mode, J):
    s, g, o, sp, (off, nbr, r) = dense_case(ipg, orc, mode, J)
    if mode == "reference":
        assert asymmetric_rows(off, nbr) > 0
    assert len(nbr) > 2 * len(s)
    B, dB, W = ipg.snap.compute_all_descriptors(s, g)
    Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    assert rel(W, Wo.T) < 1e-10
    assert rel(ipg.snap.compute_local_descriptors(s, g), Bo) < 1e-10


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
@pytest.mark.parametrize("J", [4, 8])
def test_unequal_radii_energy_forces(ipg, orc, mode, J):
    s, g, o, sp, (off, nbr, r) = dense_case(ipg, orc, mode, J, seed=12)
    beta = np.random.default_rng(J).normal(scale=0.1, size=3 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.1], g)
    ref = o.adjoint_energy_forces(sp, off, nbr, r, beta, 0.1)
    _, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(np.einsum("nmk,m->nk", dBo, beta), ref["forces"]) < 1e-11
    for path in ("plain", "adjoint"):
        e, f = ipg.snap.energy_and_force(s, lbp, path=path)
        assert abs(e - ref["energy"]) <= 1e-10 * abs(ref["energy"]), path
        assert rel(f, ref["forces"]) < 1e-10, path
    e32, f32 = ipg.snap.energy_and_force(s, lbp, path="adjoint_fp32")
    assert abs(e32 - ref["energy"]) <= 1e-5 * abs(ref["energy"])
    assert rel(f32, ref["forces"]) < 1e-4


def asymmetric_cutoff_case(orc, ipg):
    """A W-W pair whose two directions round differently: -get_displacement
    of (i, j) and of (j, i) give squared lengths one ulp apart whenever a
    component's difference goes through mod(x, L) + L.  The W radius is set
    so that cut^2 = (2 r_W rcutfac)^2 sits between them (Mo's larger radius
    keeps both inside the inrange radius cutmax)."""
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=3, a=3.16, jitter=0.05, nspecies=2, seed=41)
    X, L = s.positions, s.box
    sp = np.array([["W", "Mo"].index(x) for x in s.symbols], np.int32)
    best = None
    for i in range(len(X)):
        for j in range(i + 1, len(X)):
            if sp[i] or sp[j]:
                continue
            a = -orc.get_displacement(L, X[i], X[j])
            b = -orc.get_displacement(L, X[j], X[i])
            ra, rb = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], b[0] * b[0] + b[1] * b[1] + b[2] * b[2]
            if ra != rb and 2.5 < np.sqrt(ra) < 2.9 and (best is None or abs(ra - rb) > best[0]):
                best = (abs(ra - rb), min(ra, rb), max(ra, rb))
    assert best is not None
    _, lo, hi = best
    c = np.sqrt(lo)
    for _ in range(64):
        if lo <= c * c < hi:
            break
        c = np.nextafter(c, np.inf) if c * c < lo else np.nextafter(c, -np.inf)
    assert lo <= c * c < hi
    return s, sp, [c / 2.0, 1.5]


@pytest.mark.parametrize("J", [4, 8])
def test_cutoff_missing_reverse_pair(ipg, orc, J):
    s, sp, radii = asymmetric_cutoff_case(orc, ipg)
    names = ["W", "Mo"]
    g = ipg.snap.SNAP(len(s), J, names, 1.0, 0.0, 0.99363, radii, [1.0, 0.8], switch_flag=True,
                      list_mode="cutoff")
    o = orc.Snap(J, 2, 1.0, 0.0, 0.99363, radii, [1.0, 0.8], switch_flag=True)
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
    assert asymmetric_rows(off, nbr) == 1
    nl = ipg.snap.snap_neighborlist(s, g)
    np.testing.assert_array_equal(nl.offsets, off)
    np.testing.assert_array_equal(nl.j, nbr)
    beta = np.random.default_rng(3).normal(scale=0.1, size=2 * g.ncoeff)
    ref = o.adjoint_energy_forces(sp, off, nbr, r, beta, 0.25)
    _, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(np.einsum("nmk,m->nk", dBo, beta), ref["forces"]) < 1e-11
    lbp = ipg.snap.LBasisPotential(beta, [0.25], g)
    for path in ("plain", "adjoint"):
        e, f = ipg.snap.energy_and_force(s, lbp, path=path)
        assert abs(e - ref["energy"]) <= 1e-10 * abs(ref["energy"]), path
        assert rel(f, ref["forces"]) < 1e-10, path
    e32, f32 = ipg.snap.energy_and_force(s, lbp, path="adjoint_fp32")
    assert rel(f32, ref["forces"]) < 1e-4
    # device entry point: the first call captures the step, the second
    # replays it, sees the missing reverse pair and recomputes
    from ipgpu._lib import check
    from ipgpu.device import DeviceArray, dev_system

    d = {"pos": DeviceArray.from_host(s.positions), "sp": DeviceArray.from_host(sp),
         "beta": DeviceArray.from_host(beta), "e": DeviceArray((1,), np.float64),
         "f": DeviceArray((len(s), 3), np.float64)}
    dsys = dev_system(d["pos"], d["sp"], len(s), s.box, s.periodic)
    for path in (1, 2, 0):
        for call in range(3):
            check(ipg.load().ipg_snap_energy_forces_dev(g._h, C.byref(dsys), 1, C.c_void_p(d["beta"].ptr), 0.25,
                                                        path, C.c_void_p(d["e"].ptr), C.c_void_p(d["f"].ptr)))
            f = d["f"].download()
            assert rel(f, ref["forces"]) < (1e-4 if path == 2 else 1e-10), (path, call)
    assert g.graph_info()["replays"] >= 2


def test_device_entry_points_reject_bad_species(ipg):
    """Species indices outside the potential's table on the device entry
    points: the call returns IPG_ERR_INVALID (no kernel reads outside its
    tables; a later call with good species works)."""
    from ipgpu import _lib, workloads
    from ipgpu.device import DeviceArray, dev_system

    L = _lib.load()
    s = workloads.snap_bcc(ncell=3, a=3.16, jitter=0.05, nspecies=2, seed=2)
    names = ["W", "Mo"]
    g = ipg.snap.SNAP(len(s), 4, names, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], list_mode="cutoff")
    good = np.array([names.index(x) for x in s.symbols], np.int32)
    beta = DeviceArray.from_host(np.random.default_rng(0).normal(size=2 * g.ncoeff))
    pos = DeviceArray.from_host(s.positions)
    e, f = DeviceArray((1,), np.float64), DeviceArray((len(s), 3), np.float64)
    for bad_value in (2, -1, 1 << 20):
        bad = good.copy()
        bad[7] = bad_value
        for sp_host, want in ((bad, _lib.IPG_ERR_INVALID), (good, _lib.IPG_OK)):
            sp = DeviceArray.from_host(sp_host)
            dsys = dev_system(pos, sp, len(s), s.box, s.periodic)
            for path in (0, 1, 2):
                rc = L.ipg_snap_energy_forces_dev(g._h, C.byref(dsys), 1, C.c_void_p(beta.ptr), 0.0, path,
                                                  C.c_void_p(e.ptr), C.c_void_p(f.ptr))
                assert rc == want, (bad_value, path, rc)
    # pair potentials: indices outside [0, 64)
    from ipgpu.empirical import _terms_of

    p = ipg.LennardJones(0.0104 * ipg.eV, 2.0 * ipg.angstrom, 4.0 * ipg.angstrom, ["W"])
    terms = (_lib.PairTerm * 1)()
    terms[0] = _terms_of(p)[0][0]._term(1.0, names)
    posb = DeviceArray.from_host(s.austrip_positions())
    v = DeviceArray((6,), np.float64)
    for bad_value in (64, -3):
        bad = good.copy()
        bad[3] = bad_value
        for sp_host, want in ((bad, _lib.IPG_ERR_INVALID), (good, _lib.IPG_OK)):
            sp = DeviceArray.from_host(sp_host)
            dsys = dev_system(posb, sp, len(s), s.austrip_box(), s.periodic)
            rc = L.ipg_pair_energy_forces_dev(_lib.default_context(), C.byref(dsys), terms, 1, 0,
                                              C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr))
            assert rc == want, (bad_value, rc)
            lst = C.c_void_p()
            _lib.check(L.ipg_nlist_build_dev(_lib.default_context(), C.byref(dsys), 4.0 * ipg.angstrom, 1,
                                             C.byref(lst)))
            rc = L.ipg_pair_eval_dev(_lib.default_context(), lst, C.byref(dsys), terms, 1, 0, C.c_void_p(e.ptr),
                                     C.c_void_p(f.ptr), C.c_void_p(v.ptr))
            L.ipg_nlist_destroy(lst)
            assert rc == want, (bad_value, rc)
