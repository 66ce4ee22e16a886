"""sf_norm_noalloc (the stability norm without any allocation) against the reference-ABI norms.

get_norm_elastic_from_device_ / get_norm_acoustic_from_device_ take the mesh's own fields; the
same arrays are copied into buffers this test allocates, and sf_norm_noalloc must return the same
bits (a maximum is exact, and the per-point norm is the same arithmetic). Two calls in a row, the
second on a smaller field whose values are all smaller than the first's: a reduction that read
partial maxima left in the shared scratch array by the first call would return the first maximum.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import small_coupled_box, small_elastic_box

pytestmark = pytest.mark.gpu


class Hip:
    """The HIP runtime the library itself links (test-side allocations only)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)

    def check(self, e, what):
        assert e == 0, f"{what}: hipError {e}"

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))), "hipMalloc")
        if a.nbytes:
            self.check(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")
        return p

    def download_scalar(self, p):
        out = np.zeros(1, np.float32)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(4), 2), "hipMemcpy D2H")
        return out[0]

    def sync(self):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.check(self.h.hipFree(p), "hipFree")


def _lib():
    from specfem3d_amd import abi
    L = abi.load()
    L.sf_norm_noalloc.restype = C.c_int
    L.sf_norm_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    return L


def _abi_norm_elastic(db, src, rec, displ):
    from specfem3d_amd import abi
    from specfem3d_amd.solver import GPUSolver
    g = GPUSolver(db, src, rec, nstep=1)
    z = np.zeros_like(displ)
    g.L.transfer_fields_el_to_device_(C.byref(C.c_int(displ.size)), abi.ptr(displ), abi.ptr(z), abi.ptr(z), g.MP)
    out = C.c_float(0)
    g.L.get_norm_elastic_from_device_(C.byref(out), g.MP, C.byref(C.c_int(1)))
    g.close()
    return np.float32(out.value)


def _abi_norm_acoustic(db, src, rec, pot):
    from specfem3d_amd import abi
    from specfem3d_amd.solver import GPUSolver
    g = GPUSolver(db, src, rec, nstep=1)
    z = np.zeros_like(pot)
    g.L.transfer_fields_ac_to_device_(C.byref(C.c_int(pot.size)), abi.ptr(pot), abi.ptr(z), abi.ptr(z), g.MP)
    out = C.c_float(0)
    g.L.get_norm_acoustic_from_device_(C.byref(out), g.MP, C.byref(C.c_int(1)))
    g.close()
    return np.float32(out.value)


def _noalloc(L, hip, field, ncomp):
    d = hip.upload(field)
    o = hip.upload(np.full(1, -1.0, np.float32))
    rc = L.sf_norm_noalloc(d, field.size // ncomp, ncomp, o, None)
    assert rc == 0, L.sf_last_error()
    hip.sync()
    v = hip.download_scalar(o)
    hip.free(d)
    hip.free(o)
    return v


def test_norm_noalloc_matches_abi_norms_two_calls():
    L, hip = _lib(), Hip()
    rng = np.random.default_rng(7)
    big = small_elastic_box(n=6)
    small = small_elastic_box(n=3)
    d_big = (rng.standard_normal(3 * big[0].nglob) * 10.0).astype(np.float32)
    d_small = (rng.standard_normal(3 * small[0].nglob) * 0.1).astype(np.float32)
    ref_big = _abi_norm_elastic(*big, d_big)
    ref_small = _abi_norm_elastic(*small, d_small)
    print(f"[gpu] noalloc elastic: nglob {big[0].nglob} then {small[0].nglob}", flush=True)
    got_big = _noalloc(L, hip, d_big, 3)
    got_small = _noalloc(L, hip, d_small, 3)
    assert got_big == ref_big and got_small == ref_small, (got_big, ref_big, got_small, ref_small)
    assert got_small < got_big

    cb = small_coupled_box(n=6)
    cs = small_coupled_box(n=3)
    p_big = (rng.standard_normal(cb[0].nglob) * 5.0).astype(np.float32)
    p_small = (rng.standard_normal(cs[0].nglob) * 0.5).astype(np.float32)
    ref_pb = _abi_norm_acoustic(*cb, p_big)
    ref_ps = _abi_norm_acoustic(*cs, p_small)
    print(f"[gpu] noalloc acoustic: nglob {cb[0].nglob} then {cs[0].nglob}", flush=True)
    assert _noalloc(L, hip, p_big, 1) == ref_pb
    assert _noalloc(L, hip, p_small, 1) == ref_ps


def test_norm_noalloc_nan_and_empty():
    L, hip = _lib(), Hip()
    f = np.ones(3 * 5000, np.float32)
    f[3 * 4321 + 1] = np.nan
    assert _noalloc(L, hip, f, 3) == np.inf
    assert _noalloc(L, hip, np.zeros(0, np.float32), 3) == 0.0


def test_norm_noalloc_rejects_bad_sizes():
    L = _lib()
    dummy = C.c_void_p(16)  # never dereferenced: the calls are refused before any launch
    assert L.sf_norm_noalloc(dummy, 268435457, 3, dummy, None) == -1
    assert b"npts" in L.sf_last_error()
    assert L.sf_norm_noalloc(dummy, 100, 2, dummy, None) == -1
    assert b"ncomp" in L.sf_last_error()
    assert L.sf_norm_noalloc(None, 100, 1, dummy, None) == -1
