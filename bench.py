"""Benchmark: SNAP 2J=8 energy+force atoms/s (FP32 adjoint vs FP64 plain) and
LJ/Morse/ZBL pairs/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-lj]

Step (value): one SNAP 2J=8 energy+force evaluation of an 8192-atom
two-species bcc supercell (a = 3.16 A, seeded jitter 0.05 A, rcut = 5 A,
switching on, beta ~ N(0, 0.1) of shape 2x55) through the FP32 adjoint path,
inputs resident in HBM: SNAP neighbor list, U, Z/B/Y, per-pair dE/dr and
the force gather.  Prints one JSON line.
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


def cpu_baseline_snap(s, sn_sizes, beta, sp, sample):
    """Oracle (C restatement of the reference's plain SNAP path: B then
    per-pair dB then dB . beta), one thread, on the first `sample` atoms."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    o = O.Snap(8, 2, 5.0, 0.0, 0.99363, [0.5, 0.5], [1.0, 0.9], switch_flag=True)
    off, nbr, R, r = o.neighborlist(s.positions, s.box, sp, mode=O.SNAP_LIST_CUTOFF)
    n = len(s)
    M = o.ncoeff * 2
    B = np.zeros((n, o.ncoeff))
    dB = np.zeros((n, M, 3))
    L = O.lib()
    spc = np.ascontiguousarray(sp, np.int32)
    t0 = time.perf_counter()
    L.orc_snap_all_descriptors(C.byref(o.p), o.t, sample, s.positions.ctypes.data, spc.ctypes.data,
                               off.ctypes.data, nbr.ctypes.data, np.ascontiguousarray(r).ctypes.data,
                               B.ctypes.data, dB.ctypes.data, None)
    F = np.einsum("nmk,m->nk", dB, beta)
    dt = time.perf_counter() - t0
    del F
    return sample / dt, dt


def lj_bench(steps, warmup):
    import ipgpu
    from ipgpu import workloads, perfmodel
    from ipgpu._lib import check
    from ipgpu.device import DeviceArray, dev_system, prof_enable, prof_reset

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
    res = {"natoms": len(s), "half_pairs": half}
    for prec, nm in ((1, "fp32"), (0, "fp64")):
        def kern():
            check(L.ipg_pair_eval_dev(ctx, lst, C.byref(dsys), C.cast(arr, C.c_void_p), len(terms), prec,
                                      C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr)))

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
        b = perfmodel.pair_bytes(len(s), half)
        res[nm] = {
            "pairs_per_s_kernel": half / (kms * 1e-3),
            "pairs_per_s_end_to_end": half / t_e,
            "kernel_ms": kms,
            "end_to_end_ms": t_e * 1e3,
            "roofline": {"bound": "hbm", "achieved": b / (kms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": b / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, "traffic": None,
                         "algorithmic_bytes": b},
        }
    check(L.ipg_nlist_destroy(lst))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lj", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=8192)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this engine has no multi-GPU path "
              "(--gpus %d is not supported)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import ipgpu
    from ipgpu import perfmodel
    from ipgpu._lib import IPG_SNAP_ADJOINT_FP32, IPG_SNAP_PLAIN_FP64, IPG_SNAP_ADJOINT_FP64
    from ipgpu.device import prof_enable, prof_reset, set_graphs

    ipgpu.load()
    log("setup SNAP workload")
    s, sn, beta, sp, d, dsys = snap_setup()
    n = len(s)

    # FP32 adjoint (value), kernels timed with HIP events over the timed region
    prof_reset()
    prof_enable(True)
    t32 = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP32), a.steps, a.warmup)
    prof_enable(False)
    k32 = kernel_avg(["snap_ui_fp32", "snap_zy_fp32", "snap_dedr_fp32", "snap_force_gather", "snap_list_count",
                      "snap_list_fill", "nlist_count", "nlist_fill", "snap_ycoef", "reduce_rows", "scan",
                      "rev_sort", "cell_sort"])
    f32 = d["f"].download()
    e32 = float(d["e"].download()[0])
    ginfo = sn.graph_info()
    log("fp32 adjoint: %.3f ms/step (graph %s)" % (t32 * 1e3, ginfo))
    # the same step without profiling events, and without graph replay
    t32_noev = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP32), a.steps, 1)
    set_graphs(False)
    t32_eager = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP32), a.steps, 1)
    set_graphs(True)
    log("fp32 adjoint without events %.3f ms, uncaptured %.3f ms" % (t32_noev * 1e3, t32_eager * 1e3))
    # FP64 plain (the reference's route) and FP64 adjoint
    steps64 = max(3, a.steps // 4)
    t64 = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_PLAIN_FP64), steps64, 1)
    f64 = d["f"].download()
    e64 = float(d["e"].download()[0])
    t64a = time_steps(lambda: snap_step(sn, dsys, d, IPG_SNAP_ADJOINT_FP64), steps64, 1)
    f64a = d["f"].download()

    log("fp64 plain: %.3f ms/step, fp64 adjoint %.3f ms/step" % (t64 * 1e3, t64a * 1e3))
    # roofline of the dominant FP32 kernel
    nl = ipgpu.snap.snap_neighborlist(s, sn, list_mode="cutoff")
    npairs = len(nl.j)
    flops = {
        "snap_ui_fp32": perfmodel.flops_ui_per_pair(8) * npairs,
        "snap_zy_fp32": perfmodel.flops_zby_per_atom(8) * n,
        "snap_dedr_fp32": perfmodel.flops_dedr_per_pair(8) * npairs,
    }
    dom = max((k for k in flops if k in k32), key=lambda k: k32[k][1])
    dom_ms = k32[dom][1]
    ach = flops[dom] / (dom_ms * 1e-3) / 1e12

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
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_TFLOPS, "traffic": None, "kernel": dom,
                     "kernel_ms": dom_ms, "algorithmic_flops": flops[dom],
                     "note": "VALU-bound FP32 kernel priced against the published dense FP32 peak"},
        "snap": {
            "fp32_adjoint_atoms_per_s": n / t32,
            "fp64_plain_atoms_per_s": n / t64,
            "fp64_adjoint_atoms_per_s": n / t64a,
            "speedup_fp32_adjoint_vs_fp64_plain": t64 / t32,
            "kernels_fp32_ms": {k: v[1] for k, v in k32.items()},
            "fp32_adjoint_ms_without_events": t32_noev * 1e3,
            "fp32_adjoint_ms_uncaptured": t32_eager * 1e3,
            "graph": ginfo,
            "fp32_vs_fp64_error": err,
        },
    }
    if not a.no_lj:
        log("LJ/Morse/ZBL 256k pair kernel")
        out["lj"] = lj_bench(max(5, a.steps), a.warmup)
    if not a.no_cpu:
        log("CPU baseline (oracle)")
        sample = min(a.cpu_sample, n)
        rate, dt = cpu_baseline_snap(s, None, beta, sp, sample)
        out["cpu_baseline"] = {"value": rate, "unit": "atoms/s", "cores": 1, "kind": "port",
                               "sample": "oracle plain SNAP 2J=8 E+F (B, per-pair dB, dB.beta) over the first "
                                         "%d of the 8192 atoms, %.1f s, 1 thread" % (sample, dt)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
