"""GPU SNAP on the runtime-table kernels: chem_flag = true (with its forced
bnorm_flag, the reference's multi-element arithmetic as written) and
2J = 9, 10, against the oracle, which is pinned for these cases by
tests/test_oracle.py (single-element identity, brute-force mixed-element
B and dB, exact CG and Wigner matrices and dB finite differences at the
top levels).

Tolerances: B, dB and W within 1e-10 relative (to the largest entry),
energy and forces of the plain FP64 route within 1e-10, the virial within
1e-10; the batched descriptors within 1e-12 of the single-configuration
path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IPG_ERR_UNSUPPORTED = -4
NAMES = ["W", "Mo", "Ta"]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def case(ipg, orc, J, nspecies, chem, mode="cutoff", seed=3, bzero=False, wselfall=False, radii=None,
         ncell=3):
    from ipgpu import workloads

    b = workloads.snap_bcc(ncell=ncell, a=3.16, jitter=0.05, nspecies=1, seed=seed)
    names = NAMES[:nspecies]
    syms = [names[k] for k in np.random.default_rng(seed + 100).integers(0, nspecies, len(b))]
    s = ipg.AtomsSystem(b.positions, b.box, b.periodic, syms, unit="angstrom")
    radii = radii or [1.3, 1.1, 0.9][:nspecies]
    weight = [1.0, 0.8, 1.2][:nspecies]
    kw = dict(chem_flag=chem, bnorm_flag=chem, bzero_flag=bzero, switch_flag=True, wselfall_flag=wselfall)
    # REFERENCE mode queries radius sqrt(cutmax) (utl.jl:81): rcutfac 3.2
    # reaches the first bcc shell (2.74 A)
    rcf = 2.4 if mode == "cutoff" else 3.2
    g = ipg.snap.SNAP(len(s), J, names, rcf, 0.0, 0.99363, radii, weight, list_mode=mode, **kw)
    o = orc.Snap(J, nspecies, rcf, 0.0, 0.99363, radii, weight, **kw)
    sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
    uniq = s.unique_symbols()
    ap = np.array([uniq.index(x) for x in s.symbols], dtype=np.int32)
    om = orc.SNAP_LIST_CUTOFF if mode == "cutoff" else orc.SNAP_LIST_REFERENCE
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, appear=ap, mode=om)
    return s, g, o, sp, (off, nbr, r)


@pytest.mark.parametrize("J,ns,mode,bzero,wselfall", [
    (2, 2, "cutoff", False, False), (5, 2, "cutoff", True, False), (8, 2, "cutoff", True, True),
    (4, 3, "cutoff", False, False), (3, 3, "reference", True, False), (6, 2, "reference", False, False)])
def test_chem_all_descriptors(ipg, orc, J, ns, mode, bzero, wselfall):
    s, g, o, sp, (off, nbr, r) = case(ipg, orc, J, ns, True, mode=mode, bzero=bzero, wselfall=wselfall,
                                      radii=[1.0, 1.3, 0.8][:ns])
    assert g.ncoeff == o.ncoeff == o.idxb_max * ns ** 3
    assert off[-1] > 0
    B, dB, W = ipg.snap.compute_all_descriptors(s, g)
    Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    assert rel(W, Wo.T) < 1e-10
    Bl = ipg.snap.compute_local_descriptors(s, g)
    assert np.array_equal(Bl, B)


@pytest.mark.parametrize("J,chem", [(9, False), (10, False), (9, True)])
def test_high_levels_all_descriptors(ipg, orc, J, chem):
    s, g, o, sp, (off, nbr, r) = case(ipg, orc, J, 2, chem, seed=4, bzero=True)
    B, dB, W = ipg.snap.compute_all_descriptors(s, g)
    Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    assert rel(W, Wo.T) < 1e-10


@pytest.mark.parametrize("J,ns,chem,mode", [(4, 2, True, "cutoff"), (8, 3, True, "reference"),
                                            (10, 2, False, "cutoff"), (9, 1, False, "reference")])
def test_plain_energy_forces_virial(ipg, orc, J, ns, chem, mode):
    s, g, o, sp, (off, nbr, r) = case(ipg, orc, J, ns, chem, mode=mode, seed=6, bzero=True)
    beta = np.random.default_rng(J).normal(scale=0.1, size=ns * g.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, [0.25], g)
    eo, fo, _ = o.plain_energy_forces(sp, off, nbr, r, beta)
    e, f = ipg.snap.energy_and_force(s, lbp)  # the default path is the plain one here
    assert abs(e - (eo + 0.25)) <= 1e-10 * abs(eo + 0.25)
    assert off[-1] > 0
    assert rel(f, fo) < 1e-10
    Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    v = ipg.snap.virial_stress(s, lbp)
    assert rel(v, Wo.T @ beta) < 1e-10
    e2, f2, v2 = ipg.snap.energy_force_virial(s, lbp)
    assert abs(e2 - e) <= 1e-12 * abs(e) and rel(f2, f) < 1e-12 and rel(v2, v) < 1e-12


def test_plain_dev_graph_replays(ipg, orc):
    """The captured device step on the runtime-table kernels: replays equal
    the first (exact) call bit for bit and match the oracle."""
    import ctypes as C

    from ipgpu._lib import check
    from ipgpu.device import DeviceArray, dev_system

    s, g, o, sp, (off, nbr, r) = case(ipg, orc, 5, 2, True, seed=8)
    beta = np.random.default_rng(2).normal(scale=0.1, size=2 * g.ncoeff)
    eo, fo, _ = o.plain_energy_forces(sp, off, nbr, r, beta)
    d = {"pos": DeviceArray.from_host(s.positions), "sp": DeviceArray.from_host(sp),
         "beta": DeviceArray.from_host(beta), "e": DeviceArray((1,), np.float64),
         "f": DeviceArray((len(s), 3), np.float64)}
    dsys = dev_system(d["pos"], d["sp"], len(s), s.box, s.periodic)
    outs = []
    for _ in range(4):
        check(ipg.load().ipg_snap_energy_forces_dev(g._h, C.byref(dsys), 1, C.c_void_p(d["beta"].ptr), 0.0, 0,
                                                    C.c_void_p(d["e"].ptr), C.c_void_p(d["f"].ptr)))
        outs.append((d["e"].download()[0], d["f"].download()))
    for e, f in outs[1:]:
        assert e == outs[0][0] and np.array_equal(f, outs[0][1])
    assert abs(outs[0][0] - eo) <= 1e-10 * abs(eo)
    assert rel(outs[0][1], fo) < 1e-10
    info = g.graph_info()
    assert info["state"] == "ready" and info["replays"] >= 2, info


def test_batch_fp64_matches_single(ipg, orc):
    from ipgpu import workloads

    names = ["W", "Mo"]
    systems = workloads.snap_batch(nconfig=4, nmin=10, nmax=60, rcut=5.0, seed=17)
    g = ipg.snap.SNAP(100, 4, names, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], chem_flag=True, bnorm_flag=True,
                      switch_flag=True, list_mode="cutoff")
    r = ipg.snap.compute_batch_descriptors(systems, g, precision="fp64")
    o = orc.Snap(4, 2, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], chem_flag=True, bnorm_flag=True, switch_flag=True)
    for c, s in enumerate(systems):
        B, dB, W = ipg.snap.compute_all_descriptors(s, g, list_mode="cutoff")
        a, b = r.start[c], r.start[c + 1]
        assert rel(r.B[a:b], B) < 1e-12 and rel(r.dB[a:b], dB) < 1e-12 and rel(r.W[c], W) < 1e-12
        sp = np.array([names.index(x) for x in s.symbols], dtype=np.int32)
        off, nbr, R, rv = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
        Bo, dBo, Wo = o.all_descriptors(s.positions, sp, off, nbr, rv)
        assert rel(r.B[a:b], Bo) < 1e-10 and rel(r.dB[a:b], dBo.transpose(0, 2, 1)) < 1e-10
    with pytest.raises(ipg.IPGError) as ei:
        ipg.snap.compute_batch_descriptors(systems, g, precision="fp32")
    assert ei.value.code == IPG_ERR_UNSUPPORTED


def test_generic_refusals(ipg, orc):
    s, g, o, sp, _ = case(ipg, orc, 4, 2, True, seed=9)
    lbp = ipg.snap.LBasisPotential(np.ones(2 * g.ncoeff) * 0.01, [0.0], g)
    for path in ("adjoint", "adjoint_fp32"):
        with pytest.raises(ipg.IPGError) as ei:
            ipg.snap.energy_and_force(s, lbp, path=path)
        assert ei.value.code == IPG_ERR_UNSUPPORTED
    with pytest.raises(ipg.IPGError) as ei:
        ipg.snap.force_vjp(s, lbp, np.ones((len(s), 3)))
    assert ei.value.code == IPG_ERR_UNSUPPORTED
    with pytest.raises(ipg.IPGError) as ei:
        ipg.snap.SNAP(10, 11, ["W"], 2.4, 0.0, 0.99363, [1.0], [1.0])
    assert ei.value.code == IPG_ERR_UNSUPPORTED
