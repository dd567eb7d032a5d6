"""GPU neighbor lists vs the oracle: bit-exact neighbor sets and displacements
(src/nnlist.jl:23-46, utl.jl:38-101)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

INF = np.inf


def vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def gpu_list(ipg, pos, box, per, rcut, full=False):
    from ipgpu.nnlist import neighborlist_arrays
    from ipgpu._lib import IPG_LIST_FULL, IPG_LIST_HALF

    return neighborlist_arrays(pos, box, per, rcut, IPG_LIST_FULL if full else IPG_LIST_HALF)


def oracle_list(orc, pos, box, per, rcut, full=False):
    L = [box[d] if per[d] else INF for d in range(3)]
    return orc.neighborlist(pos, L, rcut, kind=orc.LIST_FULL if full else orc.LIST_HALF)


def assert_same(a, b):
    off_a, j_a, d_a = a
    off_b, j_b, d_b = b
    np.testing.assert_array_equal(off_a, off_b)
    np.testing.assert_array_equal(j_a, j_b)
    if d_a is not None and d_b is not None:
        np.testing.assert_array_equal(d_a, d_b)  # bit-exact displacements


def test_reference_unit_vector(ipg):
    v = vectors()["nnlist_unit"]
    s = ipg.AtomsSystem(v["positions_bohr"], v["box_bohr"], v["periodic"], ["Ar"] * 3, unit="bohr")
    nl = ipg.neighborlist(s, v["rcutoff"])
    assert len(nl) == 3
    assert [list(j + 1) for j in nl.j] == v["j_1based"]
    for ri, tr in zip(nl.r, v["r"]):
        if tr:
            np.testing.assert_allclose(ri, np.array(tr), rtol=0, atol=1e-14)


@pytest.mark.parametrize("seed,n,per,rcut", [
    (0, 300, (True, True, True), 2.5),
    (1, 500, (True, False, True), 3.1),
    (2, 400, (False, False, False), 2.0),
    (3, 64, (True, True, True), 7.9),      # rcut > L/2: minimum image only
    (4, 200, (True, True, False), 1.0),
])
def test_random_half_and_full(ipg, orc, seed, n, per, rcut):
    rng = np.random.default_rng(seed)
    box = np.array([9.0, 11.0, 10.0]) if n > 64 else np.array([5.0, 6.0, 7.0])
    pos = rng.random((n, 3)) * box
    pos[::7] += rng.integers(-3, 4, size=pos[::7].shape) * box  # unwrapped images
    for full in (False, True):
        a = gpu_list(ipg, pos, box, per, rcut, full)
        b = oracle_list(orc, pos, box, per, rcut, full)
        assert_same(a, b)


def test_edge_cases(ipg, orc):
    box = np.array([4.0, 4.0, 4.0])
    # empty and single-atom systems
    off, j, d = gpu_list(ipg, np.zeros((0, 3)), box, (True,) * 3, 2.0)
    assert len(off) == 1 and off[0] == 0 and len(j) == 0
    off, j, d = gpu_list(ipg, np.array([[1.0, 2.0, 3.0]]), box, (True,) * 3, 2.0)
    assert list(off) == [0, 0]
    # all atoms on top of each other
    pos = np.tile([[1.0, 1.0, 1.0]], (20, 1))
    assert_same(gpu_list(ipg, pos, box, (True,) * 3, 0.5), oracle_list(orc, pos, box, (True,) * 3, 0.5))
    # pairs exactly at the half box and exactly at rcut
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [0, 1.5, 0], [0, 0, 3.0]])
    for rc in (1.5, 2.0, 1.0):
        for full in (False, True):
            assert_same(gpu_list(ipg, pos, box, (True,) * 3, rc, full), oracle_list(orc, pos, box, (True,) * 3, rc, full))


def test_invalid_inputs(ipg):
    with pytest.raises(ipg.IPGError):
        gpu_list(ipg, np.zeros((3, 3)), np.array([1.0, 1.0, 1.0]), (True,) * 3, -1.0)
    with pytest.raises(ipg.IPGError):
        gpu_list(ipg, np.array([[np.nan, 0, 0]]), np.array([1.0, 1.0, 1.0]), (True,) * 3, 1.0)
    with pytest.raises(ipg.IPGError):
        gpu_list(ipg, np.zeros((2, 3)), np.array([0.0, 1.0, 1.0]), (True,) * 3, 1.0)


def test_large_periodic_lattice(ipg, orc):
    from ipgpu import workloads

    s, p = workloads.lj_argon_fcc()
    pos, box = s.austrip_positions(), s.austrip_box()
    a = gpu_list(ipg, pos, box, (True,) * 3, p.rcutoff)
    b = orc.neighborlist(pos, box, p.rcutoff, kind=orc.LIST_HALF, cells=True)
    assert_same(a, b)
    # FCC a = 5.26 A: shells 12+6+24+12+24 = 78 neighbors within 8.5 A -> ~39 half pairs/atom
    assert 37.0 < len(a[1]) / 1000 < 41.0


@pytest.mark.parametrize("n,per", [(3000, (True, True, True)), (9000, (True, True, True)),
                                   (9000, (True, False, True)), (20000, (True, True, True)),
                                   (40000, (True, True, True)), (40000, (False, True, True))])
def test_binning_paths_by_size(ipg, orc, n, per):
    """The cell binning runs in two launches up to 4096 atoms, three up to
    16384 and five above; every path gives the oracle's lists exactly (half
    and full, liquid-like density, periodic or mixed boundaries)."""
    rng = np.random.default_rng(n)
    L = (n / 0.05) ** (1.0 / 3.0)
    pos = rng.random((n, 3)) * L
    box = [L, L, L]
    for full in (False, True):
        a = gpu_list(ipg, pos, box, per, 5.0, full=full)
        Lo = [box[d] if per[d] else INF for d in range(3)]
        b = orc.neighborlist(pos, Lo, 5.0, kind=orc.LIST_FULL if full else orc.LIST_HALF, cells=True)
        assert_same(a, b)
        assert len(a[1]) > (10 if full else 5) * n


def test_thin_periodic_slab(ipg, orc):
    """40,000 atoms in a periodic slab 8 A thick: one cell across it, fewer
    than the 3 of the stencil, so the scan visits that dimension's cells
    once each (minimum-image prefilter); the lists equal the oracle's."""
    rng = np.random.default_rng(21)
    box = [200.0, 200.0, 8.0]
    pos = rng.random((40000, 3)) * np.array(box)
    for full in (False, True):
        a = gpu_list(ipg, pos, box, (True, True, True), 5.0, full=full)
        b = orc.neighborlist(pos, box, 5.0, kind=orc.LIST_FULL if full else orc.LIST_HALF, cells=True)
        assert_same(a, b)


def _snap_case(ipg, orc, mode, radii, nspecies, seed, n=120, a=3.16, rcutfac=2.0, dense=False):
    rng = np.random.default_rng(seed)
    from ipgpu import workloads

    if dense:
        L = np.array([6.0, 6.0, 6.0])
        X = rng.random((n, 3)) * L
    else:
        X, L = workloads.bcc_positions(4, 4, 4, a)
        X = X[:n] + rng.normal(scale=0.05, size=(n, 3))
    names = ["W", "Mo", "Ta"][:nspecies]
    syms = [names[k] for k in rng.integers(0, nspecies, n)]
    s = ipg.AtomsSystem(X, L, (True,) * 3, syms, unit="angstrom")
    sn = ipg.snap.SNAP(n, 4, names, rcutfac, 0.0, 0.99363, radii, [1.0] * nspecies, list_mode=mode)
    g = ipg.snap.snap_neighborlist(s, sn)
    o = orc.Snap(4, nspecies, rcutfac, 0.0, 0.99363, radii, [1.0] * nspecies)
    sp = np.array([names.index(x) for x in syms], dtype=np.int32)
    uniq = s.unique_symbols()
    ap = np.array([uniq.index(x) for x in syms], dtype=np.int32)
    off, j, R, r = o.neighborlist(X, L, sp, appear=ap,
                                  mode=orc.SNAP_LIST_REFERENCE if mode == "reference" else orc.SNAP_LIST_CUTOFF)
    np.testing.assert_array_equal(g.offsets, off)
    np.testing.assert_array_equal(g.j, j)
    np.testing.assert_array_equal(g.r, r)
    np.testing.assert_allclose(g.R, R, rtol=2e-16, atol=0)
    return len(j)


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
def test_snap_list_equal_radii(ipg, orc, mode):
    npairs = _snap_case(ipg, orc, mode, [1.0, 1.0], 2, 7)
    if mode == "reference":
        assert npairs == 0  # inrange radius sqrt(4.0) = 2.0 < bcc nearest distance 2.74
    else:
        assert npairs > 0


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
def test_snap_list_unequal_radii(ipg, orc, mode):
    # dense random atoms so that the per-species cutoffs split the inrange
    # result: exercises the as-is element lookups of utl.jl:70,86
    npairs = _snap_case(ipg, orc, mode, [0.3, 1.5, 0.9], 3, 11, rcutfac=1.0, dense=True)
    assert npairs > 0


def _snap_list_check(ipg, orc, X, L, syms, names, radii, rcutfac, twojmax=4):
    s = ipg.AtomsSystem(X, L, (True,) * 3, syms, unit="angstrom")
    ns = len(names)
    sn = ipg.snap.SNAP(len(X), twojmax, names, rcutfac, 0.0, 0.99363, radii, [1.0] * ns, list_mode="cutoff")
    g = ipg.snap.snap_neighborlist(s, sn)
    o = orc.Snap(twojmax, ns, rcutfac, 0.0, 0.99363, radii, [1.0] * ns)
    sp = np.array([names.index(x) for x in syms], dtype=np.int32)
    uniq = s.unique_symbols()
    ap = np.array([uniq.index(x) for x in syms], dtype=np.int32)
    off, j, R, r = o.neighborlist(X, L, sp, appear=ap, mode=orc.SNAP_LIST_CUTOFF)
    np.testing.assert_array_equal(g.offsets, off)
    np.testing.assert_array_equal(g.j, j)
    np.testing.assert_array_equal(g.r, r)
    np.testing.assert_allclose(g.R, R, rtol=2e-16, atol=0)
    return np.diff(off)


@pytest.mark.parametrize("seed", [3, 4])
def test_snap_list_cell_rows(ipg, orc, seed):
    """CUTOFF-mode SNAP lists of periodic systems with at least 3 cells a
    side and a few atoms per cell take the cell-pair row kernel (rows
    gathered per home cell, already in j order).  Two species with unequal
    radii, so the species-pair cutoffs drop part of the query-radius hits;
    positions shifted by whole box lengths on some atoms (the exact test
    works on the unwrapped coordinates)."""
    from ipgpu import workloads

    rng = np.random.default_rng(seed)
    X, L = workloads.bcc_positions(8, 8, 8, 3.16)
    X = X + rng.normal(scale=0.15, size=X.shape)
    X[::7] += np.array(L) * rng.integers(-2, 3, size=(len(X[::7]), 3))
    names = ["W", "Mo"]
    syms = [names[k] for k in rng.integers(0, 2, len(X))]
    cnt = _snap_list_check(ipg, orc, X, L, syms, names, [1.0, 0.8], 2.4)
    assert cnt.min() > 5 and cnt.max() < 60


def test_snap_list_cell_rows_overflow_falls_back(ipg, orc):
    """A dense cluster of 70 atoms inside one cell overfills the cell-pair
    kernel's per-cell and per-row buffers; the list is rebuilt by the binned
    scan and still equals the oracle's."""
    from ipgpu import workloads

    rng = np.random.default_rng(5)
    X, L = workloads.bcc_positions(8, 8, 8, 3.16)
    X = X + rng.normal(scale=0.1, size=X.shape)
    cl = np.array([12.0, 12.0, 12.0]) + rng.random((70, 3)) * 3.0
    X = np.vstack([X, cl])
    names = ["W", "Mo"]
    syms = [names[k] for k in rng.integers(0, 2, len(X))]
    cnt = _snap_list_check(ipg, orc, X, L, syms, names, [1.0, 0.8], 2.4)
    assert cnt.max() > 96


def test_snap_list_half_cells_large(ipg, orc):
    """A CUTOFF-mode SNAP list of 40,768 atoms takes the half-cutoff-cell
    scan with the species-pair test.  Expected rows: the oracle's cell list
    at the query radius (PeriodicEuclidean, pinned against brute force)
    filtered by the second test of utl.jl:84-88, r.r of the reference's
    get_displacement against the pair's cutoff squared (the oracle's own
    SNAP list is all-pairs and too slow at this size)."""
    from ipgpu import workloads

    rng = np.random.default_rng(8)
    X, L = workloads.bcc_positions(28, 28, 26, 3.16)
    X = X + rng.normal(scale=0.1, size=X.shape)
    names = ["W", "Mo"]
    radii, rcutfac = [1.0, 0.8], 2.4
    spi = rng.integers(0, 2, len(X))
    syms = [names[k] for k in spi]
    s = ipg.AtomsSystem(X, L, (True,) * 3, syms, unit="angstrom")
    sn = ipg.snap.SNAP(len(X), 4, names, rcutfac, 0.0, 0.99363, radii, [1.0, 1.0], list_mode="cutoff")
    g = ipg.snap.snap_neighborlist(s, sn)
    # cutoffs as utl.jl:42-56 forms them
    cut = np.empty((2, 2))
    for a in range(2):
        for b in range(2):
            c = 2.0 * radii[a] * rcutfac if a == b else (radii[a] + radii[b]) * rcutfac
            cut[a, b] = c * c
    query = max(2.0 * r * rcutfac for r in radii)  # the CUTOFF-mode query radius, cutmax
    off, j, disp = orc.neighborlist(X, L, query, kind=orc.LIST_FULL, cells=True)
    ci = np.repeat(np.arange(len(X)), np.diff(off))
    rsq = (disp[:, 0] * disp[:, 0] + disp[:, 1] * disp[:, 1]) + disp[:, 2] * disp[:, 2]
    keep = rsq <= cut[spi[ci], spi[j]]
    cnt = np.bincount(ci[keep], minlength=len(X))
    np.testing.assert_array_equal(np.diff(g.offsets), cnt)
    np.testing.assert_array_equal(g.j, j[keep])
    np.testing.assert_array_equal(g.r, -disp[keep])
