"""Device-resident arrays through the C ABI (ipg_device_malloc / memcpy), for
callers that keep inputs in HBM across calls (the *_dev entry points)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


class DeviceArray:
    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ctx = _lib.default_context()
        p = C.c_void_p()
        check(_lib.load().ipg_device_malloc(self.ctx, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(_lib.load().ipg_memcpy_h2d(self.ctx, C.c_void_p(self.ptr), a.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(_lib.load().ipg_memcpy_d2h(self.ctx, out.ctypes.data, C.c_void_p(self.ptr), self.nbytes))
        return out

    def free(self):
        if self.ptr:
            _lib.load().ipg_device_free(self.ctx, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dev_system(pos_dev, species_dev, natoms, box, periodic):
    s = _lib.System()
    s.natoms = int(natoms)
    s.positions = pos_dev.ptr
    s.species = species_dev.ptr if species_dev is not None else None
    for d in range(3):
        s.box[d] = float(box[d]) if periodic[d] else 0.0
        s.periodic[d] = 1 if periodic[d] else 0
    return s


def synchronize():
    check(_lib.load().ipg_ctx_synchronize(_lib.default_context()))


def prof_enable(on):
    check(_lib.load().ipg_prof_enable(_lib.default_context(), int(on)))


def set_graphs(on):
    """Replay captured HIP graphs in the SNAP energy/force calls (default on)."""
    check(_lib.load().ipg_ctx_set_graphs(_lib.default_context(), int(on)))


def prof_reset():
    check(_lib.load().ipg_prof_reset(_lib.default_context()))


def prof_query(name):
    n = C.c_int64()
    ms = C.c_double()
    check(_lib.load().ipg_prof_query(_lib.default_context(), name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def prof_select(names=None):
    """Time only the named kernels (iterable of names; None: every kernel)."""
    arg = ",".join(names).encode() if names else None
    check(_lib.load().ipg_prof_select(_lib.default_context(), arg))
