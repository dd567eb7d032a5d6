"""GPU SNAP vs the oracle.  Tolerances: B and dB within 1e-10 relative (to
the largest entry), FP64 energy/forces 1e-10, FP32 adjoint forces reported
against the FP64 path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = np.inf


def system(ipg, seed=3, ncell=3, a=3.16, jitter=0.05, nspecies=2):
    from ipgpu import workloads

    return workloads.snap_bcc(ncell=ncell, a=a, jitter=jitter, nspecies=nspecies, seed=seed)


def make(ipg, orc, s, J, nspecies=2, switch=False, rcutfac=2.5, radii=None, weight=None, mode="cutoff",
         bzero=False):
    names = ["W", "Mo"][:nspecies]
    radii = radii or [1.0] * nspecies
    weight = weight or [1.0, 0.8][:nspecies]
    g = ipg.snap.SNAP(len(s), J, names, rcutfac, 0.0, 0.99363, radii, weight, switch_flag=switch, bzero_flag=bzero,
                      list_mode=mode)
    o = orc.Snap(J, nspecies, rcutfac, 0.0, 0.99363, radii, weight, switch_flag=switch, bzero_flag=bzero)
    sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
    uniq = s.unique_symbols()
    ap = np.array([uniq.index(x) for x in s.symbols], dtype=np.int32)
    om = orc.SNAP_LIST_CUTOFF if mode == "cutoff" else orc.SNAP_LIST_REFERENCE
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, appear=ap, mode=om)
    return g, o, sp, (off, nbr, r)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 8])
def test_local_descriptors(ipg, orc, J):
    s = system(ipg)
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, J)
    B = ipg.snap.compute_local_descriptors(s, g)
    Bo = o.local_descriptors(sp, off, nbr, r)
    assert B.shape == Bo.shape
    assert rel(B, Bo) < 1e-10


@pytest.mark.parametrize("J,switch,bzero", [(4, False, False), (6, True, False), (8, True, True), (5, False, True)])
def test_all_descriptors(ipg, orc, J, switch, bzero):
    s = system(ipg, seed=5)
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, J, switch=switch, bzero=bzero)
    B, dB, W = ipg.snap.compute_all_descriptors(s, g)
    Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    assert rel(W, Wo.T) < 1e-10


def test_reference_mode_zero_neighbors(ipg, orc):
    # bcc a = 3.32, rcut 4.67: the as-is list radius sqrt(4.67) finds nobody
    s = ipg.snap.__dict__  # noqa: F841 (module import check)
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=3, a=3.32, jitter=0.03, nspecies=1, seed=1)
    g = ipg.snap.SNAP(len(s), 6, ["W"], 4.67, 0.0, 0.99363, [0.5], [1.0], list_mode="reference")
    nl = ipg.snap.snap_neighborlist(s, g)
    assert len(nl.j) == 0
    B = ipg.snap.compute_local_descriptors(s, g)
    o = orc.Snap(6, 1, 4.67, 0.0, 0.99363, [0.5], [1.0])
    off = np.zeros(len(s) + 1, dtype=np.int64)
    Bo = o.local_descriptors(np.zeros(len(s), np.int32), off, np.zeros(0, np.int32), np.zeros((0, 3)))
    assert rel(B, Bo) < 1e-12


@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 8])
def test_energy_forces_paths(ipg, orc, J):
    s = system(ipg, seed=7)
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, J, switch=True)
    rng = np.random.default_rng(J)
    beta = rng.normal(scale=0.1, size=2 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.3], g)
    ref = o.adjoint_energy_forces(sp, off, nbr, r, beta, 0.3)
    # oracle adjoint equals the oracle plain path beta . dB
    Bo, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(np.einsum("nmk,m->nk", dBo, beta), ref["forces"]) < 1e-11
    for path in ("plain", "adjoint"):
        e, f = ipg.snap.energy_and_force(s, lbp, path=path)
        assert abs(e - ref["energy"]) <= 1e-10 * abs(ref["energy"])
        assert rel(f, ref["forces"]) < 1e-10, path
    e32, f32 = ipg.snap.energy_and_force(s, lbp, path="adjoint_fp32")
    assert abs(e32 - ref["energy"]) <= 1e-5 * abs(ref["energy"])
    assert rel(f32, ref["forces"]) < 1e-4


def test_shared_beta_matches_reference_energy(ipg, orc):
    s = system(ipg, seed=8)
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, 4)
    beta = np.random.default_rng(1).normal(scale=0.1, size=g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.0], g)
    e, _ = ipg.snap.energy_and_force(s, lbp)
    Bo = o.local_descriptors(sp, off, nbr, r)
    assert abs(e - Bo.sum(0) @ beta) <= 1e-10 * abs(e)  # linear_potential.jl:20-26


@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 2e-4)])
def test_force_vjp(ipg, orc, precision, tol, J):
    """The pullback's dual force kernel is compiled per twojmax: gbeta
    against the oracle's dB and gpos against central differences of
    gF . F at every level."""
    s = system(ipg, seed=9)
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, J, switch=True)
    rng = np.random.default_rng(2)
    beta = rng.normal(scale=0.1, size=2 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.0], g)
    gF = rng.normal(size=(len(s), 3))
    gb, gp = ipg.snap.force_vjp(s, lbp, gF, precision=precision)
    # gbeta against the oracle's dB: sum_k gF_k . dB[k][:, c]
    _, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    gb_ref = np.einsum("nk,nmk->m", gF, dBo)
    assert rel(gb, gb_ref) < tol
    # gpos against central differences of gF . F(x) (oracle adjoint forces)
    h = 1e-5

    def gdotF(X):
        off2, nbr2, _, r2 = o.neighborlist(X, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
        assert np.array_equal(nbr2, nbr)
        return float(np.sum(gF * o.adjoint_energy_forces(sp, off2, nbr2, r2, beta)["forces"]))

    for (a, k) in [(0, 0), (3, 1), (11, 2), (20, 0)]:
        Xp = s.positions.copy()
        Xp[a, k] += h
        Xm = s.positions.copy()
        Xm[a, k] -= h
        fd = (gdotF(Xp) - gdotF(Xm)) / (2 * h)
        assert abs(gp[a, k] - fd) <= max(tol, 1e-6) * max(np.abs(gp).max(), 1.0), (a, k, gp[a, k], fd)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_force_vjp_without_pairs(ipg, precision):
    """Atoms farther apart than every cutoff: no pairs, so F = 0 for every
    beta and position and the pullback is zero (the force kernels launch
    nothing and the gathers read empty rows)."""
    s = ipg.AtomsSystem(np.array([[1.0, 1.0, 1.0], [9.0, 9.0, 9.0], [1.0, 9.0, 5.0]]), [20.0, 20.0, 20.0],
                        (True, True, True), ["W", "Mo", "W"], unit="angstrom")
    g = ipg.snap.SNAP(len(s), 6, ["W", "Mo"], 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=True,
                      list_mode="cutoff")
    beta = np.random.default_rng(4).normal(scale=0.1, size=2 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.0], g)
    gb, gp = ipg.snap.force_vjp(s, lbp, np.ones((len(s), 3)), precision=precision)
    assert np.all(gb == 0.0) and np.all(gp == 0.0)


@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("bzero", [False, True])
def test_adjoint_fp32_every_level(ipg, orc, J, bzero):
    """The FP32 adjoint path (its Z/Y kernel is compiled per twojmax) against
    the oracle: energy within 1e-5 and forces within 1e-4 of the largest
    force, relative; 40-atom system so one batch of 32 atoms is partial."""
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=3, a=3.16, jitter=0.05, nspecies=2, seed=30 + J)
    keep = np.arange(40)
    s = type(s)(s.positions[keep], s.box, s.periodic, [s.symbols[k] for k in keep], unit="angstrom")
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, J, switch=True, bzero=bzero)
    rng = np.random.default_rng(100 + J)
    beta = rng.normal(scale=0.1, size=2 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.2], g)
    ref = o.adjoint_energy_forces(sp, off, nbr, r, beta, 0.2)
    e32, f32 = ipg.snap.energy_and_force(s, lbp, path="adjoint_fp32")
    assert abs(e32 - ref["energy"]) <= 1e-5 * max(abs(ref["energy"]), 1.0)
    assert rel(f32, ref["forces"]) < 1e-4


def test_cutoff_list_row_overflow_falls_back(ipg, orc):
    """A dense clump in a dilute box overflows the CUTOFF list's row buffer
    (sized from the mean density); the counted list that replaces it gives
    the oracle's descriptors and forces."""
    from ipgpu.system import AtomsSystem

    rng = np.random.default_rng(22)
    L = 40.0
    X = np.concatenate([rng.uniform(0, L, size=(300, 3)), rng.uniform(18.0, 24.0, size=(150, 3))])
    names = ["W", "Mo"]
    syms = [names[k] for k in rng.integers(0, 2, size=len(X))]
    s = AtomsSystem(X, [L, L, L], (True, True, True), syms, unit="angstrom")
    g, o, sp, (off, nbr, r) = make(ipg, orc, s, 4)
    assert np.diff(off).max() > 100  # rows far wider than the mean density suggests
    B, dB, _ = ipg.snap.compute_all_descriptors(s, g, want=("B", "dB"))
    Bo, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    beta = np.random.default_rng(5).normal(size=2 * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, 0.0, g)
    e, f = ipg.snap.energy_and_force(s, lbp, path="adjoint")
    e0, f0 = ipg.snap.energy_and_force(s, lbp, path="plain")
    assert abs(e - e0) <= 1e-10 * abs(e0) and rel(f, f0) < 1e-10
