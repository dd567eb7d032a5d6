"""The benchmark workloads at their full sizes, checked through properties
that do not need the oracle at that size, and against the oracle where it
finishes in seconds.

- SNAP 2J=8, 8192-atom two-species bcc supercell (the energy/force bench):
  FP32 adjoint forces within 1e-5 RMS of the FP64 plain forces (in the
  units of beta . B), FP64 adjoint within 1e-10 relative of FP64 plain,
  total force zero (translation invariance) to 1e-9 relative.
- The same workload as a force-matching fit: the pullback of F against
  central differences of the loss in beta and in positions.
- SNAP 2J=6, 2000-atom bcc crystal: B and dB against the oracle (1e-10
  relative) in CUTOFF mode; REFERENCE mode finds no neighbors there, so B is
  the self term alone and dB is zero.
- 512 configurations of 50..2000 atoms, 2J=8 (the batched-descriptor
  bench): FP32 within 2e-5 relative of FP64, energy rows equal to the sums
  of B per configuration and species (1e-12 relative), total dB of each
  configuration zero (1e-9 relative)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_snap_2J8_8192_atoms_paths(ipg):
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=16, a=3.16, jitter=0.05, nspecies=2, seed=5)
    names = ["W", "Mo"]
    sn = ipg.snap.SNAP(len(s), 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    beta = np.random.default_rng(55).normal(scale=0.1, size=2 * sn.ncoeff)
    lbp = ipg.snap.LBasisPotential(beta, 0.0, sn)
    e64, f64 = ipg.snap.energy_and_force(s, lbp, path="plain")
    ea, fa = ipg.snap.energy_and_force(s, lbp, path="adjoint")
    e32, f32 = ipg.snap.energy_and_force(s, lbp, path="adjoint_fp32")
    assert abs(ea - e64) <= 1e-10 * abs(e64) and rel(fa, f64) < 1e-10
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("FP32 adjoint vs FP64 plain: force RMS %.3g, max %.3g, energy %.3g" %
          (rms, np.abs(f32 - f64).max(), abs(e32 - e64)))
    assert rms <= 1e-5
    assert abs(e32 - e64) <= 1e-6 * len(s) * max(1.0, abs(e64) / len(s))
    for f in (f64, fa):
        assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).sum()


def test_snap_2J8_8192_atoms_force_matching_gradient(ipg):
    """Force-matching loss L = mean |F - F_target|^2 over the 8192-atom
    workload, F_target from a second seeded beta.  The pullback of F (FP64)
    at gF = dL/dF must equal central differences of L: in beta (L is
    quadratic in beta, so the difference is exact up to rounding) and in a
    seeded subset of coordinates (h = 1e-4 A).  The FP32 pullback stays
    within 1e-5 of the FP64 one, relative to the largest entry."""
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=16, a=3.16, jitter=0.05, nspecies=2, seed=5)
    names = ["W", "Mo"]
    sn = ipg.snap.SNAP(len(s), 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    rng = np.random.default_rng(56)
    beta = rng.normal(scale=0.1, size=2 * sn.ncoeff)
    beta_t = rng.normal(scale=0.1, size=2 * sn.ncoeff)
    F_t = ipg.snap.energy_and_force(s, ipg.snap.LBasisPotential(beta_t, 0.0, sn), path="adjoint")[1]

    def loss(X=None, b=beta):
        sx = s if X is None else type(s)(X, s.box, s.periodic, s.symbols, unit="angstrom")
        F = ipg.snap.energy_and_force(sx, ipg.snap.LBasisPotential(b, 0.0, sn), path="adjoint")[1]
        return float(np.mean((F - F_t) ** 2)), F

    L0, F = loss()
    gF = 2.0 * (F - F_t) / F.size
    lbp = ipg.snap.LBasisPotential(beta, 0.0, sn)
    gb, gp = ipg.snap.force_vjp(s, lbp, gF, precision="fp64")
    for c in rng.choice(beta.size, 6, replace=False):
        hb = 1e-3
        bp, bm = beta.copy(), beta.copy()
        bp[c] += hb
        bm[c] -= hb
        fd = (loss(b=bp)[0] - loss(b=bm)[0]) / (2 * hb)
        assert abs(gb[c] - fd) <= 1e-8 * np.abs(gb).max(), (c, gb[c], fd)
    h = 1e-4
    for a, k in zip(rng.choice(len(s), 6, replace=False), rng.integers(0, 3, 6)):
        Xp, Xm = s.positions.copy(), s.positions.copy()
        Xp[a, k] += h
        Xm[a, k] -= h
        fd = (loss(Xp)[0] - loss(Xm)[0]) / (2 * h)
        assert abs(gp[a, k] - fd) <= 1e-5 * np.abs(gp).max(), (a, k, gp[a, k], fd)
    gb32, gp32 = ipg.snap.force_vjp(s, lbp, gF, precision="fp32")
    print("FP32 pullback vs FP64: beta %.3g, positions %.3g" % (rel(gb32, gb), rel(gp32, gp)))
    assert rel(gb32, gb) < 1e-5 and rel(gp32, gp) < 1e-5


def test_snap_2J6_2000_atoms_vs_oracle(ipg, orc):
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=10, a=3.32, jitter=0.03, nspecies=1, seed=3)
    sp = np.zeros(len(s), np.int32)
    g = ipg.snap.SNAP(len(s), 6, ["W"], 4.67, 0.0, 0.99363, [0.5], [1.0], list_mode="cutoff")
    o = orc.Snap(6, 1, 4.67, 0.0, 0.99363, [0.5], [1.0])
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
    assert len(nbr) > 8 * len(s)
    B, dB, _ = ipg.snap.compute_all_descriptors(s, g, list_mode="cutoff", want=("B", "dB"))
    Bo, dBo, _ = o.all_descriptors(s.positions, sp, off, nbr, r)
    assert rel(B, Bo) < 1e-10
    assert rel(dB, dBo.transpose(0, 2, 1)) < 1e-10
    # REFERENCE mode: inrange radius sqrt(4.67) < nearest-neighbor distance
    Br, dBr, _ = ipg.snap.compute_all_descriptors(s, g, list_mode="reference", want=("B", "dB"))
    off0, nbr0, _, _ = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_REFERENCE)
    assert len(nbr0) == 0
    Bself, _, _ = o.all_descriptors(s.positions, sp, off0, nbr0, np.zeros((0, 3)))
    assert rel(Br, Bself) < 1e-12
    assert np.all(dBr == 0.0)
    assert np.all(Br == Br[0])  # the self term is the same for every atom


def test_batch_512_configs_properties(ipg):
    from ipgpu import _lib, workloads
    from ipgpu.device import DeviceArray

    names = ["W", "Mo"]
    systems = workloads.snap_batch(nconfig=512, nmin=50, nmax=2000, rcut=5.0, seed=4)
    sn = ipg.snap.SNAP(100, 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    natoms = np.array([len(s) for s in systems], dtype=np.int64)
    n, nconf, M = int(natoms.sum()), len(systems), 2 * sn.ncoeff
    start = np.concatenate([[0], np.cumsum(natoms)])
    spc = np.concatenate([[names.index(x) for x in s.symbols] for s in systems]).astype(np.int32)
    box = np.array([s.box for s in systems], dtype=np.float64)
    per = np.ones((nconf, 3), dtype=np.int32)
    pos = DeviceArray.from_host(np.concatenate([s.positions for s in systems]))
    sp = DeviceArray.from_host(spc)
    L = _lib.load()
    res = {}
    for prec in (_lib.IPG_FP64, _lib.IPG_FP32):
        B = DeviceArray((n, sn.ncoeff), np.float64)
        dB = DeviceArray((n, M, 3), np.float64)
        er = DeviceArray((nconf, M), np.float64)
        _lib.check(L.ipg_snap_batch_descriptors_dev(sn._h, nconf, natoms.ctypes.data, box.ctypes.data,
                                                    per.ctypes.data, C.c_void_p(pos.ptr), C.c_void_p(sp.ptr),
                                                    prec, C.c_void_p(B.ptr), C.c_void_p(dB.ptr),
                                                    C.c_void_p(er.ptr), None))
        _lib.check(L.ipg_ctx_synchronize(_lib.default_context()))
        res[prec] = (B.download(), dB.download(), er.download())
        del B, dB, er
    B64, dB64, er64 = res[_lib.IPG_FP64]
    B32, dB32, er32 = res[_lib.IPG_FP32]
    assert rel(B32, B64) < 2e-5 and rel(dB32, dB64) < 2e-5 and rel(er32, er64) < 2e-5
    # energy rows: sums of B per configuration and species
    rows = np.zeros((nconf, M))
    for c in range(nconf):
        a, b = start[c], start[c + 1]
        for k in range(2):
            rows[c, k * sn.ncoeff:(k + 1) * sn.ncoeff] = B64[a:b][spc[a:b] == k].sum(0)
    assert rel(er64, rows) < 1e-12
    # translation invariance: the dB of each configuration sums to zero
    scale = np.abs(dB64).max()
    for c in range(0, nconf, 37):
        assert np.abs(dB64[start[c]:start[c + 1]].sum(0)).max() <= 1e-9 * scale * (start[c + 1] - start[c])


def test_snap_2J8_8192_atoms_vs_oracle(ipg, orc):
    """The energy/force bench workload against the oracle at full size: the
    CUTOFF list bit for bit (the three-launch binning path), FP64 plain and
    adjoint energy and forces within 1e-10 relative of the oracle's plain
    route (OpenMP, 16 threads at most)."""
    import os

    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=16, a=3.16, jitter=0.05, nspecies=2, seed=5)
    names = ["W", "Mo"]
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    sn = ipg.snap.SNAP(len(s), 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    o = orc.Snap(8, 2, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True)
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, mode=orc.SNAP_LIST_CUTOFF)
    nl = ipg.snap.snap_neighborlist(s, sn)
    np.testing.assert_array_equal(nl.offsets, off)
    np.testing.assert_array_equal(nl.j, nbr)
    np.testing.assert_array_equal(nl.r, r)
    beta = np.random.default_rng(55).normal(scale=0.1, size=2 * sn.ncoeff)
    eo, fo, _ = o.plain_energy_forces(sp, off, nbr, r, beta, nthreads=min(16, os.cpu_count() or 1))
    lbp = ipg.snap.LBasisPotential(beta, 0.0, sn)
    for path in ("plain", "adjoint"):
        e, f = ipg.snap.energy_and_force(s, lbp, path=path)
        assert abs(e - eo) <= 1e-10 * abs(eo), path
        assert rel(f, fo) < 1e-10, path


def test_mixed_pairs_256k_vs_oracle(ipg, orc):
    """The 256,000-atom LJ + Morse + ZBL workload: FP64 energy, forces and
    virial within 1e-12 relative of the oracle, which evaluates the three
    potentials one by one with a cell list each, as
    LinearCombinationPotential does."""
    from test_gpu_pair import close, oracle_eval

    from ipgpu import workloads

    s, pot = workloads.mixed_pairs()
    e, f = ipg.energy_and_force(s, pot)
    vir = ipg.virial_stress(s, pot)
    eo, fo, vo = oracle_eval(orc, s, pot)
    close(e, eo)
    close(f, fo)
    close(vir, vo)


def test_batch_512_configs_vs_oracle(ipg, orc):
    """Four of the 512 batch configurations (the largest, the smallest and
    two others) against the oracle's compute_all_descriptors: FP64 batch
    within 1e-10 relative, FP32 batch within 2e-5."""
    from ipgpu import workloads

    names = ["W", "Mo"]
    systems = workloads.snap_batch(nconfig=512, nmin=50, nmax=2000, rcut=5.0, seed=4)
    sn = ipg.snap.SNAP(100, 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                       list_mode="cutoff")
    o = orc.Snap(8, 2, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True)
    sizes = np.array([len(x) for x in systems])
    pick = [int(np.argmax(sizes)), int(np.argmin(sizes)), 100, 400]
    r64 = ipg.snap.compute_batch_descriptors(systems, sn, precision="fp64")
    r32 = ipg.snap.compute_batch_descriptors(systems, sn, precision="fp32")
    for c in pick:
        sc = systems[c]
        sp = np.array([names.index(x) for x in sc.symbols], np.int32)
        off, nbr, R, rv = o.neighborlist(sc.positions, sc.box, sp, mode=orc.SNAP_LIST_CUTOFF)
        Bo, dBo, Wo = o.all_descriptors(sc.positions, sp, off, nbr, rv)
        a, b = r64.start[c], r64.start[c + 1]
        assert rel(r64.B[a:b], Bo) < 1e-10, c
        assert rel(r64.dB[a:b], dBo.transpose(0, 2, 1)) < 1e-10, c
        assert rel(r64.W[c], Wo.T) < 1e-10, c
        assert rel(r32.B[a:b], Bo) < 2e-5 and rel(r32.dB[a:b], dBo.transpose(0, 2, 1)) < 2e-5, c
