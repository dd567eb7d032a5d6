"""Benchmark: SNAP 2J=8 energy+force atoms/s (FP32 adjoint vs FP64 plain) and
LJ/Morse/ZBL pairs/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-lj]

Step (value): one SNAP 2J=8 energy+force evaluation of an 8192-atom
two-species bcc supercell (a = 3.16 A, seeded jitter 0.05 A, rcut = 5 A,
switching on, beta ~ N(0, 0.1) of shape 2x55) through the FP32 adjoint path,
inputs resident in HBM: SNAP neighbor list, U, Z/B/Y, per-pair dE/dr and
the force gather.  The timed region replays the step from its captured HIP
graph K times, as an MD loop does.  The dominant kernel (found first from a
kernel-by-kernel pass with events around every launch, reported as the
per-kernel breakdown) is then timed by HIP events around it in K more
replays of a graph that holds those events.  Also reported: the
FP64 plain and adjoint paths, the LJ/Morse/ZBL pair kernel on 256k atoms,
batched descriptors of 512 configurations (--batch-only runs just that),
and the oracle on the host CPU.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "interatomicpotentials.jl_amd"))

WORKLOAD = "snap_2J8_bcc16_8192atoms_2species_rcut5_adjoint_fp32"
PEAK_FP32_TFLOPS = 157.3  # MI355X published dense FP32
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def snap_setup():
    import ipgpu
    from ipgpu import workloads
    from ipgpu.device import DeviceArray, dev_system

    s = workloads.snap_bcc(ncell=16, a=3.16, jitter=0.05, nspecies=2, seed=5)
    names = ["W", "Mo"]
    sn = ipgpu.snap.SNAP(len(s), 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                         list_mode="cutoff")
    beta = np.random.default_rng(55).normal(scale=0.1, size=2 * sn.ncoeff)
    sp = np.array([names.index(x) for x in s.symbols], np.int32)
    d = {
        "pos": DeviceArray.from_host(s.positions),
        "sp": DeviceArray.from_host(sp),
        "beta": DeviceArray.from_host(beta),
        "e": DeviceArray((1,), np.float64),
        "f": DeviceArray((len(s), 3), np.float64),
    }
    dsys = dev_system(d["pos"], d["sp"], len(s), s.box, s.periodic)
    return s, sn, beta, sp, d, dsys


def snap_step(sn, dsys, d, path):
    import ipgpu
    from ipgpu._lib import check

    check(ipgpu.load().ipg_snap_energy_forces_dev(sn._h, C.byref(dsys), 1, C.c_void_p(d["beta"].ptr), 0.0, path,
                                                  C.c_void_p(d["e"].ptr), C.c_void_p(d["f"].ptr)))


def time_steps(fn, steps, warmup):
    from ipgpu.device import synchronize

    for _ in range(warmup):
        fn()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    synchronize()
    return (time.perf_counter() - t0) / steps


def kernel_avg(names):
    from ipgpu.device import prof_query

    out = {}
    for nm in names:
        n, ms = prof_query(nm)
        if n:
            out[nm] = (n, ms / n)
    return out


CPU_RUNS = 10  # timed runs per CPU figure (after one untimed warm-up); the median is reported


def host_info():
    """The host the CPU baselines ran on: logical CPUs, the CPUs this process
    may run on, the OpenMP thread count used for "all cores" and the CPU
    model (lscpu's "Model name", read from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    threads = max(1, min(threads, avail))
    return {"nproc": os.cpu_count(), "cpus_available": avail, "threads_all_cores": threads, "model": model}


def median_time(fn, runs=CPU_RUNS):
    """Median wall time of `runs` calls of fn after one untimed warm-up."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), len(ts)


def oracle_module():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    return O


def cpu_baseline_snap(s, beta, sp, sample1, hi):
    """The oracle (C restatement of the reference's plain SNAP route: per
    atom B, per pair dB/dr contracted with beta, forces scattered to both
    atoms), timed on this host: all cores (OpenMP) over the whole workload,
    and one thread over the first `sample1` atoms; median of CPU_RUNS."""
    O = oracle_module()
    o = O.Snap(8, 2, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True)
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, mode=O.SNAP_LIST_CUTOFF)
    n = len(s)
    nth = hi["threads_all_cores"]
    tn, runs = median_time(lambda: o.plain_energy_forces(sp, off, nbr, r, beta, nthreads=nth))
    k = min(sample1, n)
    t1, _ = median_time(lambda: o.plain_energy_forces(sp, off, nbr, r, beta, nthreads=1, count=k))
    return {"threads": nth, "rate": n / tn, "seconds": tn, "rate1": k / t1, "seconds1": t1, "sample1": k,
            "runs": runs}


def cpu_baseline_pairs(hi):
    """The oracle's pair path (every potential of the sum with its own cell
    list, as LinearCombinationPotential evaluates them; OpenMP over atoms
    with per-thread sums, as the reference threads _perform_pairwise) on the
    LJ/Morse/ZBL workload at its density: a 64,000-atom box on all cores and
    an 8,000-atom box on one thread; and the two 1000-atom LJ workloads in
    full.  Rates are half pairs per second; median of CPU_RUNS each."""
    import ipgpu
    from ipgpu import workloads
    from ipgpu.empirical import _terms_of

    O = oracle_module()

    def case(s, pot):
        table = s.unique_symbols()
        sp = np.array([table.index(x) for x in s.symbols], np.int32)
        L = [s.austrip_box()[d] if s.periodic[d] else np.inf for d in range(3)]
        mask = lambda p: O.ALL_SPECIES if p.species is None else sum(1 << k for k, x in enumerate(table)
                                                                      if x in p.species)
        terms = [dict(kind=p.kind, coeff=c, rcut=p.rcutoff, p=list(p.params()), species_mask=mask(p))
                 for p, c in _terms_of(pot)]
        X = s.austrip_positions()
        nl = ipgpu.nnlist.neighborlist_arrays(X, s.austrip_box(), s.periodic, pot.get_rcutoff(), want_disp=False)
        return X, L, sp, terms, int(len(nl[1]))

    out = {}
    nth = hi["threads_all_cores"]
    for label, n_side, threads in (("mixed_lj_morse_zbl_64000_all_cores", (40, 40, 40), nth),
                                   ("mixed_lj_morse_zbl_8000_1_thread", (20, 20, 20), 1)):
        s, pot = workloads.mixed_pairs(n_side=n_side)
        X, L, sp, terms, half = case(s, pot)
        t, runs = median_time(lambda: O.pair_energy_force_virial(X, L, sp, terms, cells=True, nthreads=threads))
        out[label] = {"natoms": len(s), "half_pairs": half, "threads": threads, "seconds": t,
                      "pairs_per_s": half / t, "runs": runs}
    s, pot = workloads.lj_argon_fcc()
    cl_X = np.loadtxt(os.path.join(REPO, "tests", "golden", "lj_1000.xyz"))
    cl = ipgpu.AtomsSystem(cl_X, [8.0, 8.0, 8.0], (False,) * 3, ["Ar"] * len(cl_X), unit="angstrom")
    cl_pot = ipgpu.LennardJones(1.0 * ipgpu.eV, 1.0 * ipgpu.angstrom, 100.0 * ipgpu.angstrom, ["Ar"])
    for label, (sx, px) in (("lj_argon_fcc_1000", (s, pot)), ("lj_cluster_1000_rcut100", (cl, cl_pot))):
        X, L, sp, terms, half = case(sx, px)
        row = {"natoms": len(sx), "half_pairs": half}
        for tag, threads in (("all_cores", nth), ("1_thread", 1)):
            t, runs = median_time(lambda: O.pair_energy_force_virial(X, L, sp, terms, cells=True, nthreads=threads))
            row[tag] = {"threads": threads, "seconds": t, "pairs_per_s": half / t, "runs": runs}
        out[label] = row
    return out


def batch_bench(steps, warmup, nconfig, cpu_configs):
    """Batched SNAP 2J=8 descriptors (B, dB and per-configuration energy and
    virial rows) of `nconfig` periodic configurations with 50..2000 atoms each, in
    one call per step: FP32 (FP64 sums) and FP64, each replayed from its
    captured HIP graph, plus one uncaptured FP32 pass with per-kernel
    events.  CPU: the oracle's descriptors of the first `cpu_configs`
    configurations on one thread."""
    import ipgpu
    from ipgpu import workloads
    from ipgpu._lib import IPG_FP32, IPG_FP64, check
    from ipgpu.device import DeviceArray, prof_enable, prof_reset, set_graphs, synchronize

    names = ["W", "Mo"]
    systems = workloads.snap_batch(nconfig=nconfig, nmin=50, nmax=2000, rcut=5.0, seed=4)
    sn = ipgpu.snap.SNAP(100, 8, names, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True,
                         list_mode="cutoff")
    natoms = np.array([len(s) for s in systems], dtype=np.int64)
    n = int(natoms.sum())
    M = 2 * sn.ncoeff
    box = np.array([s.box for s in systems], dtype=np.float64)
    per = np.ones((nconfig, 3), dtype=np.int32)
    pos = DeviceArray.from_host(np.concatenate([s.positions for s in systems]))
    sp = DeviceArray.from_host(np.concatenate([[names.index(x) for x in s.symbols] for s in systems]).astype(np.int32))
    B = DeviceArray((n, sn.ncoeff), np.float64)
    dB = DeviceArray((n, M, 3), np.float64)
    er = DeviceArray((nconfig, M), np.float64)
    W = DeviceArray((nconfig, M, 6), np.float64)
    L = ipgpu.load()

    def step(prec):
        check(L.ipg_snap_batch_descriptors_dev(sn._h, nconfig, natoms.ctypes.data, box.ctypes.data, per.ctypes.data,
                                               C.c_void_p(pos.ptr), C.c_void_p(sp.ptr), prec, C.c_void_p(B.ptr),
                                               C.c_void_p(dB.ptr), C.c_void_p(er.ptr), C.c_void_p(W.ptr)))

    res = {"configs": nconfig, "natoms": n, "twojmax": 8, "outputs": "B, dB, energy rows, virial rows"}
    for name, prec in (("fp32", IPG_FP32), ("fp64", IPG_FP64)):
        t0 = time.perf_counter()
        step(prec)  # first call: autotuning, exact run, capture
        synchronize()
        first = time.perf_counter() - t0
        t = time_steps(lambda: step(prec), steps, warmup)
        res[name] = {"ms_per_batch": t * 1e3, "atoms_per_s": n / t, "configs_per_s": nconfig / t,
                     "first_call_s": first, "tuning": ipgpu.snap.batch_info(sn, name)}
        log("batch %s: %.2f ms per batch of %d configs (%d atoms)" % (name, t * 1e3, nconfig, n))
        if name == "fp32":
            B32, er32 = B.download(), er.download()
            dB32 = dB.download()
    B64, er64, dB64 = B.download(), er.download(), dB.download()
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    rms = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2)))
    res["fp32_vs_fp64"] = {"B": rel(B32, B64), "dB": rel(dB32, dB64), "erow": rel(er32, er64),
                           "B_rms": rms(B32, B64), "dB_rms": rms(dB32, dB64), "erow_rms": rms(er32, er64),
                           "B_scale_rms": float(np.sqrt(np.mean(B64 ** 2))),
                           "dB_scale_rms": float(np.sqrt(np.mean(dB64 ** 2))),
                           "note": "B, dB, erow: max |fp32 - fp64| / max |fp64|; *_rms: RMS of fp32 - fp64; "
                                   "*_scale_rms: RMS of the fp64 values"}
    del dB32, dB64
    # per-kernel times of the FP32 pipeline, launched kernel by kernel
    set_graphs(False)
    prof_reset()
    prof_enable(True)
    step(IPG_FP32)
    synchronize()
    prof_enable(False)
    set_graphs(True)
    res["kernels_fp32_ms"] = {k: v[1] for k, v in kernel_avg(
        ["batch_scan", "batch_list_finish", "batch_list_rev", "snap_ui_fp32", "snap_zb_fp32", "snap_dbdr_fp32", "snap_db_gather",
         "batch_erow", "batch_w_atom", "batch_wrow", "scan", "fill_zero"]).items()}
    from ipgpu import perfmodel

    npairs = res["fp32"]["tuning"]["npairs"]
    res["npairs"] = npairs
    if "snap_dbdr_fp32" in res["kernels_fp32_ms"]:
        ms = res["kernels_fp32_ms"]["snap_dbdr_fp32"]
        fl = perfmodel.flops_dbdr_plain_per_pair(8) * npairs
        res["dominant_kernel"] = {"kernel": "snap_dbdr_fp32", "ms": ms, "algorithmic_flops": fl,
                                  "tflops": fl / (ms * 1e-3) / 1e12,
                                  "frac_of_fp32_peak": fl / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS}
    if cpu_configs > 0:
        O = oracle_module()
        o = O.Snap(8, 2, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True)
        prep = []
        for s in systems[:cpu_configs]:
            spc = np.array([names.index(x) for x in s.symbols], np.int32)
            off, nbr, R, r = o.neighborlist(s.positions, s.box, spc, mode=O.SNAP_LIST_CUTOFF)
            prep.append((s.positions, spc, off, nbr, r))
        na = sum(len(p[1]) for p in prep)
        tc, runs = median_time(lambda: [o.all_descriptors(*p) for p in prep])
        res["cpu_baseline"] = {"atoms_per_s": na / tc, "cores": 1, "kind": "port", "runs": runs,
                               "sample": "oracle compute_all_descriptors (B, dB, W) of the first %d "
                                         "configuration(s) (%d atoms) over their prebuilt lists, 1 thread "
                                         "(the descriptor routine is serial), median of %d runs: %.3f s"
                                         % (cpu_configs, na, runs, tc)}
    return res


def small_configs_bench(steps):
    """The remaining reference workloads, device-resident, one call per
    step: LJ argon (1000-atom fcc, periodic, rcut 8.5 A) and the reference's
    1000-atom LJ cluster (non-periodic, rcut 100 A: every pair), energy +
    forces + virial with the list built in the call; SNAP 2J=6 on a
    2000-atom bcc crystal (a = 3.32 A, rcut 4.67 A): plain FP64 energy +
    forces in both list modes, and B + dB."""
    import ipgpu
    from ipgpu import workloads
    from ipgpu._lib import IPG_FP64, IPG_SNAP_PLAIN_FP64, check
    from ipgpu.device import DeviceArray, dev_system
    from ipgpu.empirical import _terms_of

    L = ipgpu.load()
    ctx = ipgpu._lib.default_context()
    res = {}

    def pair_case(s, pot, name):
        table = s.unique_symbols()
        sp = np.array([table.index(x) for x in s.symbols], np.int32)
        pos = DeviceArray.from_host(s.austrip_positions())
        dsp = DeviceArray.from_host(sp)
        dsys = dev_system(pos, dsp, len(s), s.austrip_box(), s.periodic)
        terms = _terms_of(pot)
        arr = (ipgpu._lib.PairTerm * len(terms))()
        for k, (p, c) in enumerate(terms):
            arr[k] = p._term(c, table)
        e = DeviceArray((1,), np.float64)
        f = DeviceArray((len(s), 3), np.float64)
        v = DeviceArray((6,), np.float64)
        fn = lambda: check(L.ipg_pair_energy_forces_dev(ctx, C.byref(dsys), C.cast(arr, C.c_void_p), len(terms),
                                                        IPG_FP64, C.c_void_p(e.ptr), C.c_void_p(f.ptr),
                                                        C.c_void_p(v.ptr)))
        t = time_steps(fn, steps, 2)
        nl = ipgpu.nnlist.neighborlist_arrays(s.austrip_positions(), s.austrip_box(), s.periodic, pot.get_rcutoff(),
                                             want_disp=False)
        res[name] = {"natoms": len(s), "half_pairs": int(len(nl[1])), "ms_per_call": t * 1e3,
                     "pairs_per_s": len(nl[1]) / t, "energy_hartree": float(e.download()[0])}

    s, pot = workloads.lj_argon_fcc()
    pair_case(s, pot, "lj_argon_fcc_1000")
    X = np.loadtxt(os.path.join(REPO, "tests", "golden", "lj_1000.xyz"))
    cl = ipgpu.AtomsSystem(X, [8.0, 8.0, 8.0], (False,) * 3, ["Ar"] * len(X), unit="angstrom")
    pair_case(cl, ipgpu.LennardJones(1.0 * ipgpu.eV, 1.0 * ipgpu.angstrom, 100.0 * ipgpu.angstrom, ["Ar"]),
              "lj_cluster_1000_rcut100")

    s = workloads.snap_bcc(ncell=10, a=3.32, jitter=0.03, nspecies=1, seed=3)
    sn = ipgpu.snap.SNAP(len(s), 6, ["W"], 4.67, 0.0, 0.99363, [0.5], [1.0])
    beta = DeviceArray.from_host(np.random.default_rng(33).normal(scale=0.1, size=sn.ncoeff))
    pos = DeviceArray.from_host(s.positions)
    dsys = dev_system(pos, None, len(s), s.box, s.periodic)
    e = DeviceArray((1,), np.float64)
    f = DeviceArray((len(s), 3), np.float64)
    out = {"natoms": len(s), "twojmax": 6}
    for mode, nm in ((1, "cutoff"), (0, "reference")):
        fn = lambda: check(L.ipg_snap_energy_forces_dev(sn._h, C.byref(dsys), mode, C.c_void_p(beta.ptr), 0.0,
                                                        IPG_SNAP_PLAIN_FP64, C.c_void_p(e.ptr), C.c_void_p(f.ptr)))
        t = time_steps(fn, steps, 2)
        out["plain_fp64_ms_%s" % nm] = t * 1e3
        out["plain_fp64_atoms_per_s_%s" % nm] = len(s) / t
    natoms = np.array([len(s)], np.int64)
    box = np.array([s.box], np.float64)
    per = np.ones((1, 3), np.int32)
    B = DeviceArray((len(s), sn.ncoeff), np.float64)
    dB = DeviceArray((len(s), sn.ncoeff, 3), np.float64)
    fn = lambda: check(L.ipg_snap_batch_descriptors_dev(sn._h, 1, natoms.ctypes.data, box.ctypes.data,
                                                        per.ctypes.data, C.c_void_p(pos.ptr), None, IPG_FP64,
                                                        C.c_void_p(B.ptr), C.c_void_p(dB.ptr), None, None))
    t = time_steps(fn, steps, 2)
    out["B_dB_fp64_ms_cutoff"] = t * 1e3
    res["snap_J6_bcc10_2000"] = out
    return res


def lj_setup():
    """The mixed LJ/Morse/ZBL 256k-atom workload on the device with its
    neighbor list built once, and a call of the pair kernel over that list."""
    import ipgpu
    from ipgpu import workloads
    from ipgpu._lib import check
    from ipgpu.device import DeviceArray, dev_system

    s, pot = workloads.mixed_pairs()
    table = s.unique_symbols()
    sp = np.array([table.index(x) for x in s.symbols], np.int32)
    pos = DeviceArray.from_host(s.austrip_positions())
    dsp = DeviceArray.from_host(sp)
    dsys = dev_system(pos, dsp, len(s), s.austrip_box(), s.periodic)
    from ipgpu.empirical import _terms_of

    terms = _terms_of(pot)
    arr = (ipgpu._lib.PairTerm * len(terms))()
    for k, (p, c) in enumerate(terms):
        arr[k] = p._term(c, table)
    L = ipgpu.load()
    ctx = ipgpu._lib.default_context()
    lst = C.c_void_p()
    check(L.ipg_nlist_build_dev(ctx, C.byref(dsys), pot.get_rcutoff(), 1, C.byref(lst)))
    npairs = C.c_int64()
    check(L.ipg_nlist_info(lst, None, C.byref(npairs), None, None))
    half = npairs.value // 2
    e = DeviceArray((1,), np.float64)
    f = DeviceArray((len(s), 3), np.float64)
    v = DeviceArray((6,), np.float64)

    def kern(prec):
        check(L.ipg_pair_eval_dev(ctx, lst, C.byref(dsys), C.cast(arr, C.c_void_p), len(terms), prec,
                                  C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr)))

    keep = (pos, dsp, e, f, v, arr)  # device buffers the calls use
    return s, pot, L, ctx, lst, dsys, arr, terms, half, kern, keep


def lj_bench(steps, warmup):
    import ipgpu
    from ipgpu import perfmodel
    from ipgpu._lib import check
    from ipgpu.device import prof_enable, prof_reset

    s, pot, L, ctx, lst, dsys, arr, terms, half, kern_p, keep = lj_setup()
    pos, dsp, e, f, v, _ = keep
    res = {"natoms": len(s), "half_pairs": half}
    outs = {}
    for prec, nm in ((1, "fp32"), (0, "fp64")):
        kern = lambda: kern_p(prec)

        def e2e():
            check(L.ipg_pair_energy_forces_dev(ctx, C.byref(dsys), C.cast(arr, C.c_void_p), len(terms), prec,
                                               C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr)))

        prof_reset()
        prof_enable(True)
        t_k = time_steps(kern, steps, warmup)
        prof_enable(False)
        kn = "pair_" + nm
        ka = kernel_avg([kn])
        kms = ka[kn][1] if kn in ka else t_k * 1e3
        t_e = time_steps(e2e, max(3, steps // 2), 1)
        outs[nm] = (float(e.download()[0]), f.download(), v.download())
        b = perfmodel.pair_bytes(len(s), half)
        res[nm] = {
            "pairs_per_s_kernel": half / (kms * 1e-3),
            "pairs_per_s_end_to_end": half / t_e,
            "kernel_ms": kms,
            "end_to_end_ms": t_e * 1e3,
            "roofline": {"bound": "hbm", "achieved": b / (kms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": b / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, "traffic": read_traffic(kn)[0],
                         "traffic_source": read_traffic(kn)[1], "algorithmic_bytes": b},
        }
    (e32, f32, v32), (e64, f64, v64) = outs["fp32"], outs["fp64"]
    res["fp32_vs_fp64_error"] = {
        "force_max": float(np.abs(f32 - f64).max()), "force_rms": float(np.sqrt(np.mean((f32 - f64) ** 2))),
        "force_scale_max": float(np.abs(f64).max()), "force_scale_rms": float(np.sqrt(np.mean(f64 ** 2))),
        "energy_abs": abs(e32 - e64), "energy_rel": abs(e32 - e64) / abs(e64),
        "virial_max": float(np.abs(v32 - v64).max()), "virial_scale_max": float(np.abs(v64).max()),
        "units": "hartree, hartree/bohr"}
    check(L.ipg_nlist_destroy(lst))
    # MD-style step over a Verlet-skin list (skin 0.5 A, positions unchanged,
    # so no rebuild): displacement check + FP64 evaluation, against the
    # reference's rebuild on every call (end_to_end above)
    skin = C.c_void_p()
    check(L.ipg_nlist_build_skin_dev(ctx, C.byref(dsys), pot.get_rcutoff(), 0.5 * ipgpu.angstrom, 1,
                                     C.byref(skin)))

    def md_step():
        check(L.ipg_nlist_update_dev(skin, C.byref(dsys), None))
        check(L.ipg_pair_eval_dev(ctx, skin, C.byref(dsys), C.cast(arr, C.c_void_p), len(terms), 0,
                                  C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr)))

    t_md = time_steps(md_step, max(3, steps // 2), 1)
    res["fp64"]["md_step_ms_skin_list"] = t_md * 1e3
    res["fp64"]["pairs_per_s_md_step_skin_list"] = half / t_md
    check(L.ipg_nlist_destroy(skin))
    return res


COPY_BYTES = 256 << 20


def probe():
    """Counter-pass workload: the copy kernel's bytes are known (it reads and
    writes COPY_BYTES), so FETCH_SIZE / WRITE_SIZE of the SNAP kernels can be
    scaled by what the counters read for it."""
    import ipgpu
    from ipgpu._lib import IPG_SNAP_ADJOINT_FP32, IPG_SNAP_ADJOINT_FP64, IPG_SNAP_PLAIN_FP64, check, default_context
    from ipgpu.device import DeviceArray, set_graphs, synchronize

    src = DeviceArray((COPY_BYTES // 8,), np.float64)
    dst = DeviceArray((COPY_BYTES // 8,), np.float64)
    L = ipgpu.load()
    for _ in range(3):
        check(L.ipg_device_copy(default_context(), C.c_void_p(dst.ptr), C.c_void_p(src.ptr), COPY_BYTES))
    synchronize()
    s, sn, beta, sp, d, dsys = snap_setup()
    set_graphs(False)
    for _ in range(3):
        snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP32)
    snap_step(sn, dsys, d, IPG_SNAP_PLAIN_FP64)
    snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP64)
    synchronize()
    # one FP64 and one FP32 force pullback on the same workload
    lbp = ipgpu.snap.LBasisPotential(beta, 0.0, sn)
    gF = np.random.default_rng(57).normal(size=(len(s), 3))
    for prec in ("fp64", "fp32"):
        ipgpu.snap.force_vjp(s, lbp, gF, precision=prec)
    # one uncaptured FP32 batched-descriptor call over 32 configurations
    from ipgpu import workloads

    systems = workloads.snap_batch(nconfig=32, nmin=50, nmax=2000, rcut=5.0, seed=4)
    ipgpu.snap.compute_batch_descriptors(systems, sn, precision="fp32")
    # three FP64 and three FP32 pair-kernel calls over the prebuilt list of
    # the mixed 256k-atom workload
    s2, _, L2, _, lst, _, _, _, _, kern, keep = lj_setup()
    for prec in (0, 0, 0, 1, 1, 1):
        kern(prec)
    synchronize()
    check(L2.ipg_nlist_destroy(lst))
    print(json.dumps({"probe": True, "copy_bytes": COPY_BYTES, "natoms": len(s), "pair_natoms": len(s2)}))


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/*_traffic.txt, lines 'kernel: NAME',
    'traffic_bytes_per_launch: X' and 'measured: DATE'), and a label naming
    that file and its date: the figure comes from a separate counter run of
    bench.py --probe, not from this run.  (None, None) when absent."""
    import glob

    for fn in sorted(glob.glob(os.path.join(REPO, "profiles", "*_traffic.txt"))):
        kv = {}
        with open(fn) as fh:
            for line in fh:
                if ":" in line:
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if kv.get("kernel") == kernel and "traffic_bytes_per_launch" in kv:
            src = "profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of bench.py --probe, measured %s)" % (
                os.path.basename(fn), kv.get("measured", "date not recorded"))
            return float(kv["traffic_bytes_per_launch"]), src
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lj", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-configs", type=int, default=512, help="configurations of the batched descriptor run")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched descriptor section")
    ap.add_argument("--lj-only", action="store_true", help="run only the pair-potential sections")
    ap.add_argument("--cpu-sample", type=int, default=2048, help="atoms of the 1-thread CPU baseline")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware-counter passes: 3 calibration copies of 256 MiB, 3 uncaptured "
                         "SNAP FP32 steps, one FP64 plain and one FP64 adjoint step, one FP64 and one FP32 force "
                         "pullback, one FP32 batched-descriptor call (32 configurations), six pair-kernel calls")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this engine has no multi-GPU path "
              "(--gpus %d is not supported)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import ipgpu
    from ipgpu import perfmodel
    from ipgpu._lib import IPG_SNAP_ADJOINT_FP32, IPG_SNAP_PLAIN_FP64, IPG_SNAP_ADJOINT_FP64
    from ipgpu.device import prof_enable, prof_reset, prof_select, set_graphs

    ipgpu.load()
    if a.probe:
        probe()
        return
    if a.batch_only:
        print(json.dumps({"batch": batch_bench(max(3, a.steps // 4), 1, a.batch_configs, 0)}))
        return
    if a.lj_only:
        print(json.dumps({"lj": lj_bench(max(5, a.steps), a.warmup),
                          "other_workloads": small_configs_bench(max(5, a.steps // 2))}))
        return
    log("setup SNAP workload")
    s, sn, beta, sp, d, dsys = snap_setup()
    n = len(s)

    # FP32 adjoint.  First a per-kernel breakdown: the step launched kernel
    # by kernel (graphs off) with HIP events around every launch.
    step32 = lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP32)
    names32 = ["snap_ui_fp32", "snap_zy_fp32", "snap_dedr_fp32", "snap_force_gather", "snap_list_finish", "snap_list_rev",
               "snap_pairs", "nlist_rows", "cell_pack", "nlist_count", "nlist_fill", "nlist_sort", "cell_assign",
               "cell_place", "cell_rank", "fill_zero", "reduce_rows", "scan"]
    set_graphs(False)
    for _ in range(a.warmup):
        step32()
    prof_select(None)
    prof_reset()
    prof_enable(True)
    t32_unc = time_steps(step32, a.steps, 0)
    prof_enable(False)
    k32 = kernel_avg(names32)
    log("fp32 adjoint uncaptured with events on every kernel: %.3f ms/step" % (t32_unc * 1e3))
    nl = ipgpu.snap.snap_neighborlist(s, sn, list_mode="cutoff")
    npairs = len(nl.j)
    flops = {
        "snap_ui_fp32": perfmodel.flops_ui_per_pair(8) * npairs,
        "snap_zy_fp32": perfmodel.flops_zby_per_atom(8) * n,
        "snap_dedr_fp32": perfmodel.flops_dedr_per_pair(8) * npairs,
    }
    dom = max((k for k in flops if k in k32), key=lambda k: k32[k][1])
    # Timed region (value): K steps replayed from the captured HIP graph, as
    # an MD loop runs them, nothing recorded inside.  Then the same K steps
    # from a graph with HIP events (graph nodes on the context stream) around
    # the dominant kernel only, whose average duration gives the roofline
    # figure; an event node costs a few microseconds of idle GPU, so that
    # pass's step time is reported beside the value, not as it.
    set_graphs(True)
    prof_enable(False)
    for _ in range(max(2, a.warmup)):
        step32()
    t32 = time_steps(step32, a.steps, 0)
    f32 = d["f"].download()
    e32 = float(d["e"].download()[0])
    prof_select([dom])
    prof_reset()
    prof_enable(True)
    for _ in range(2):
        step32()
    prof_reset()
    t32_ev = time_steps(step32, a.steps, 0)
    prof_enable(False)
    kdom = kernel_avg([dom])
    prof_select(None)
    dom_ms = kdom[dom][1]
    ach = flops[dom] / (dom_ms * 1e-3) / 1e12
    ginfo = sn.graph_info()
    log("fp32 adjoint graph replay: %.3f ms/step (%.3f ms with events around %s, which took %.1f us) (%s)" %
        (t32 * 1e3, t32_ev * 1e3, dom, dom_ms * 1e3, ginfo))
    # FP64 plain (the reference's route) and FP64 adjoint, graph replay
    steps64 = max(3, a.steps // 4)
    t64 = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_PLAIN_FP64), steps64, 2)
    f64 = d["f"].download()
    e64 = float(d["e"].download()[0])
    t64a = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP64), steps64, 2)
    f64a = d["f"].download()
    log("fp64 plain: %.3f ms/step, fp64 adjoint %.3f ms/step" % (t64 * 1e3, t64a * 1e3))
    # Backward of a force-matching fit: the pullback of F(x, beta) for a
    # seeded gF, through the host-buffer entry point (positions and gF
    # uploaded, both gradients downloaded inside the time)
    lbp = ipgpu.snap.LBasisPotential(beta, 0.0, sn)
    gF = np.random.default_rng(57).normal(size=(n, 3))
    backward = {"entry": "ipg_snap_force_vjp (host buffers)"}
    for prec in ("fp64", "fp32"):
        tb = time_steps(lambda: ipgpu.snap.force_vjp(s, lbp, gF, precision=prec), steps64, 1)
        backward[prec + "_ms"] = tb * 1e3
        backward[prec + "_atoms_per_s"] = n / tb
    log("force pullback: fp64 %.3f ms, fp32 %.3f ms" % (backward["fp64_ms"], backward["fp32_ms"]))

    # FP32 vs FP64 force errors (SNAP energies/forces are in the units of beta.B)
    df = f32 - f64
    err = {"force_max": float(np.abs(df).max()), "force_rms": float(np.sqrt(np.mean(df ** 2))),
           "energy_abs": abs(e32 - e64), "force_scale_max": float(np.abs(f64).max()),
           "adjoint_fp64_vs_plain_fp64_force_max": float(np.abs(f64a - f64).max())}

    out = {
        "metric": "SNAP 2J=8 atoms/s energy+force, FP32 adjoint vs FP64 plain; LJ pairs/s",
        "value": n / t32,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded bcc supercell, seeded beta)",
        "config": {"workload": WORKLOAD, "natoms": n, "twojmax": 8, "rcut": 5.0, "species": 2,
                   "pairs": npairs, "switch_flag": True, "list_mode": "cutoff"},
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_TFLOPS, "traffic": read_traffic(dom)[0],
                     "traffic_source": read_traffic(dom)[1], "kernel": dom,
                     "kernel_ms": dom_ms, "algorithmic_flops": flops[dom],
                     "note": "FP32 vector (VALU) kernel, no MFMA, priced against the published FP32 peak "
                             "(157.3 TFLOP/s, which counts packed FP32); the code is built without packed "
                             "FP32, whose ceiling is 78.6 TFLOP/s"},
        "snap": {
            "fp32_adjoint_atoms_per_s": n / t32,
            "fp64_plain_atoms_per_s": n / t64,
            "fp64_adjoint_atoms_per_s": n / t64a,
            "speedup_fp32_adjoint_vs_fp64_plain": t64 / t32,
            "kernels_fp32_ms_uncaptured": {k: v[1] for k, v in k32.items()},
            "kernel_launches_per_step": {k: v[0] / a.steps for k, v in k32.items()},
            "fp32_adjoint_ms_uncaptured_all_events": t32_unc * 1e3,
            "timed_region": "graph replay, no events; %s timed by HIP events in %d further graph replays" % (dom,
                                                                                                   a.steps),
            "fp32_adjoint_ms_graph_with_events": t32_ev * 1e3,
            "graph": ginfo,
            "fp32_vs_fp64_error": err,
            "force_matching_backward": backward,
        },
    }
    if not a.no_lj:
        log("LJ/Morse/ZBL 256k pair kernel")
        out["lj"] = lj_bench(max(5, a.steps), a.warmup)
    if not a.no_lj:
        log("remaining workloads: LJ argon 1000, LJ cluster 1000, SNAP 2J=6 2000 atoms")
        out["other_workloads"] = small_configs_bench(max(5, a.steps // 2))
    if not a.no_batch:
        log("batched SNAP descriptors")
        out["batch"] = batch_bench(max(3, a.steps // 4), 1, a.batch_configs, 0 if a.no_cpu else 1)
    if not a.no_cpu:
        log("CPU baseline (oracle)")
        hi = host_info()
        cb = cpu_baseline_snap(s, beta, sp, a.cpu_sample, hi)
        out["cpu_baseline"] = {"value": cb["rate"], "unit": "atoms/s", "cores": cb["threads"], "kind": "port",
                               "sample": "oracle (C restatement of the reference's plain route: B, per-pair dB, "
                                         "beta . dB) SNAP 2J=8 E+F over all %d atoms on %d OpenMP threads, median "
                                         "of %d runs after a warm-up: %.3f s" % (n, cb["threads"], cb["runs"],
                                                                                  cb["seconds"]),
                               "value_1_thread": cb["rate1"],
                               "sample_1_thread": "first %d atoms, 1 thread, median of %d runs: %.3f s"
                                                  % (cb["sample1"], cb["runs"], cb["seconds1"]),
                               "host": hi}
        if not a.no_lj:
            out["cpu_baseline_pairs"] = cpu_baseline_pairs(hi)
            out["cpu_baseline_pairs"]["host"] = hi
            out["cpu_baseline_pairs"]["kind"] = "port"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
