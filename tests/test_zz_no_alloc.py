"""The allocation-free transport entry point against the allocating one it
varies: device arrays in and out (allocated here through the HIP runtime the
library is bound to), the library's static per-GPU scratch inside; bitwise
the same results on two calls in a row, the second on a smaller grid, in both
mass-flux modes; a grid above the scratch maximum is refused and leaves the
outputs untouched."""
import ctypes

import numpy as np
import pytest

from nimbus_amd import lib as L
from tests.helpers import small_case

pytestmark = pytest.mark.gpu

F32 = np.float32


def fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


class DeviceArray:
    """A device copy of a float32 array (the test's own allocation: the entry
    point under test allocates nothing). It comes from the HIP runtime that
    libnimbus_hip.so is bound to, found by that runtime's soname: a second
    HIP runtime in the process (as a torch wheel brings) would hand out
    pointers the library's runtime does not own."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        assert self.hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.shape, F32)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)


def dptr(a):
    return ctypes.cast(a.ptr, ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _case(nx, ny, nz, seed):
    g, base, st = small_case(nx=nx, ny=ny, nz=nz, cloudy=True, seed=seed)
    rng = np.random.default_rng(seed)
    shape = (nx, ny, nz)
    rho = st["rho"].astype(F32)
    u = (st["u"] + 5.0 * rng.standard_normal(shape)).astype(F32)
    v = (st["v"] + 5.0 * rng.standard_normal(shape)).astype(F32)
    w = (2.0 * rng.standard_normal(shape)).astype(F32)
    w[:, :, -1] = 0.0
    rb = rho.astype(np.float64)
    mfx = (0.5 * (rb + np.roll(rb, -1, 0)) * u).astype(F32)
    mfy = (0.5 * (rb + np.roll(rb, -1, 1)) * v).astype(F32)
    mfz = (0.5 * (rb + np.roll(rb, -1, 2)) * w).astype(F32)
    rho_e = (rb * (1.0 + 1e-3 * rng.standard_normal(shape))).astype(F32)
    qs = [np.ascontiguousarray(st[k], F32) for k in ("theta", "qv", "qc", "qr")]
    return g, dict(rho=rho, u=u, v=v, w=w, mfx=mfx, mfy=mfy, mfz=mfz, rho_e=rho_e), qs


def _dispatch(g, arrays, qin, qout, supplied, ptr):
    pd = (ctypes.c_int * 4)(0, 1, 1, 1)
    kap = np.zeros(4, F32)
    qi = (ctypes.POINTER(ctypes.c_float) * 4)(*[ptr(q) for q in qin])
    qo = (ctypes.POINTER(ctypes.c_float) * 4)(*[ptr(q) for q in qout])
    a = arrays
    d = L.ConservativeTransportDispatch(
        None if supplied else ptr(a["u"]), None if supplied else ptr(a["v"]), None if supplied else ptr(a["w"]),
        ptr(a["rho_e"] if supplied else a["rho"]), qi, qo, fp(kap), pd, 4, g.nx, g.ny, g.nz, 1.0 / g.dx, 1.0 / g.dy,
        1.0 / g.dz, 0.0, 0.0, 0.0, 4.0, 0.05, ptr(a["rho"]) if supplied else None,
        ptr(a["mfx"]) if supplied else None, ptr(a["mfy"]) if supplied else None,
        ptr(a["mfz"]) if supplied else None, int(supplied), 0.0, 0.0, 1, 0)
    return d, (qi, qo, pd, kap)


@pytest.mark.parametrize("supplied", [True, False])
def test_no_alloc_transport_equals_the_allocating_entry_point(hip_ctx, supplied):
    for nx, ny, nz, seed in ((48, 40, 30, 1), (20, 18, 14, 2)):  # the second call is smaller
        print(f"case: no-alloc transport supplied={supplied} grid {nx}x{ny}x{nz}", flush=True)
        g, arrays, qs = _case(nx, ny, nz, seed)
        host_out = [np.zeros_like(q) for q in qs]
        d, keep = _dispatch(g, arrays, qs, host_out, supplied, fp)
        L.check(hip_ctx.lib.nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid(hip_ctx.handle,
                                                                                           ctypes.byref(d)), "ref")
        dev = {k: DeviceArray(v) for k, v in arrays.items()}
        dq = [DeviceArray(q) for q in qs]
        dout = [DeviceArray(np.full(q.shape, np.nan, F32)) for q in qs]
        d2, keep2 = _dispatch(g, dev, dq, dout, supplied, dptr)
        rc = hip_ctx.lib.nimbus_hip_advect_scalars_conservative_noalloc(hip_ctx.handle, ctypes.byref(d2))
        print(f"  rc {rc}: {hip_ctx.lib.nimbus_hip_last_error().decode() if rc else 'ok'}", flush=True)
        assert rc == 0
        for q in range(4):
            got = dout[q].get()
            print(f"  scalar {q}: max diff {float(np.nanmax(np.abs(got.astype(np.float64) - host_out[q]))):.3e}, "
                  f"nan {int(np.isnan(got).sum())}", flush=True)
            assert np.array_equal(got, host_out[q]), (nx, q)


def test_no_alloc_transport_refuses_grids_above_the_scratch_maximum(hip_ctx):
    nx, ny, nz = 130, 130, 64  # 1,081,600 cells > 2^20
    g = L.make_grid(nx, ny, nz, 1000.0, 1000.0, 500.0)
    ones = np.ones((nx, ny, nz), F32)
    arrays = {k: DeviceArray(ones) for k in ("rho", "u", "v", "w", "mfx", "mfy", "mfz", "rho_e")}
    qin = [DeviceArray(ones) for _ in range(4)]
    qout = [DeviceArray(np.full((nx, ny, nz), 7.0, F32)) for _ in range(4)]
    d, keep = _dispatch(g, arrays, qin, qout, True, dptr)
    rc = hip_ctx.lib.nimbus_hip_advect_scalars_conservative_noalloc(hip_ctx.handle, ctypes.byref(d))
    assert rc == -1
    assert "scratch maximum" in hip_ctx.lib.nimbus_hip_last_error().decode()
    assert all(np.all(o.get() == 7.0) for o in qout)
