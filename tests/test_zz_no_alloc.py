"""The allocation-free pair evaluation (ipg_pair_eval_noalloc_dev) against
ipg_pair_eval_dev through the C ABI: bit-identical energy, forces and
virial on two calls in a row, the second on a smaller system (the static
scratch keeps the first call's contents), in FP64 and FP32; and the size
limit of its static scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed(ipg, n_side, seed):
    from ipgpu import workloads

    return workloads.mixed_pairs(n_side=n_side, seed=seed)


def device_case(ipg, s, pot):
    from ipgpu import _lib
    from ipgpu.device import DeviceArray, dev_system
    from ipgpu.empirical import _terms_of

    table = s.unique_symbols()
    sp = np.array([table.index(x) for x in s.symbols], np.int32)
    d = {"pos": DeviceArray.from_host(s.austrip_positions()), "sp": DeviceArray.from_host(sp)}
    dsys = dev_system(d["pos"], d["sp"], len(s), s.austrip_box(), s.periodic)
    flat = _terms_of(pot)
    terms = (_lib.PairTerm * len(flat))()
    for k, (p, c) in enumerate(flat):
        terms[k] = p._term(c, table)
    lst = C.c_void_p()
    _lib.check(_lib.load().ipg_nlist_build_dev(_lib.default_context(), C.byref(dsys), pot.get_rcutoff(), 1,
                                               C.byref(lst)))
    return d, dsys, terms, lst


def evaluate(ipg, fn, d, dsys, terms, lst, n, precision):
    from ipgpu import _lib
    from ipgpu.device import DeviceArray

    e, f, v = DeviceArray((1,), np.float64), DeviceArray((n, 3), np.float64), DeviceArray((6,), np.float64)
    L = _lib.load()
    _lib.check(getattr(L, fn)(_lib.default_context(), lst, C.byref(dsys), terms, len(terms), precision,
                              C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr)))
    _lib.check(L.ipg_ctx_synchronize(_lib.default_context()))
    return e.download()[0], f.download(), v.download()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_eval_dev_twice(ipg, precision):
    from ipgpu import _lib

    cases = [mixed(ipg, (40, 40, 20), 3), mixed(ipg, (16, 16, 12), 4)]  # 32000 atoms, then 3072
    for s, pot in cases:
        d, dsys, terms, lst = device_case(ipg, s, pot)
        try:
            want = evaluate(ipg, "ipg_pair_eval_dev", d, dsys, terms, lst, len(s), precision)
            got = evaluate(ipg, "ipg_pair_eval_noalloc_dev", d, dsys, terms, lst, len(s), precision)
        finally:
            _lib.load().ipg_nlist_destroy(lst)
        assert np.isfinite(want[0]) and want[0] != 0.0
        assert got[0] == want[0]
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[2], want[2])


def test_noalloc_rejects_systems_beyond_its_scratch(ipg):
    from ipgpu import _lib

    s, pot = mixed(ipg, (72, 72, 60), 5)  # 311,040 atoms at about 76 neighbors: 19,440 blocks > 18,715
    d, dsys, terms, lst = device_case(ipg, s, pot)
    try:
        from ipgpu.device import DeviceArray

        e, f, v = DeviceArray((1,), np.float64), DeviceArray((len(s), 3), np.float64), DeviceArray((6,), np.float64)
        L = _lib.load()
        rc = L.ipg_pair_eval_noalloc_dev(_lib.default_context(), lst, C.byref(dsys), terms, len(terms), 0,
                                         C.c_void_p(e.ptr), C.c_void_p(f.ptr), C.c_void_p(v.ptr))
        assert rc == _lib.IPG_ERR_INVALID
        assert b"static scratch" in L.ipg_last_error()
    finally:
        _lib.load().ipg_nlist_destroy(lst)
