"""GPU: cylmhd_setdt_noalloc -- Setdt on a caller's device arrays, with its
block partials in the library's static device array instead of memory it
allocates -- against cylmhd_setdt on a context holding the same state: dt,
ch and the limiting cell equal, on two calls in a row, the second on a
smaller disc, on the null stream and on a stream of the caller's; sizes it
cannot take are rejected with an error.  The caller's device buffers come
from the HIP runtime directly (ctypes)."""
import ctypes as C
import os

import numpy as np
import pytest

import cylmhd as cm
from cases import disc_case
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("libamdhip64.so not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _to_device(h, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0     # host -> device
    return p


def test_setdt_noalloc_matches_setdt(gpu):
    h = _hip()
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    bufs = []
    try:
        for size in ((48, 40, 24), (24, 16, 12)):          # the second smaller
            p, Ui, Vi, _, _ = disc_case(3, *size)
            p.path = cm.PATH_PLAIN64
            with cm.Context(p) as g:
                g.set_initial(Ui, Vi)
                ref = g.setdt()
                _, V = g.download()
            o = Oracle(p)
            minlength = o.geometry()[5]                    # [NX][NZ] (cylindricalgrid.h:146)
            mml = o.grid()["minminlength"]
            o.close()
            Vd, md = _to_device(h, V), _to_device(h, minlength)
            bufs += [Vd, md]
            got = cm.setdt_noalloc(p, Vd.value, md.value, mml)
            got_s = cm.setdt_noalloc(p, Vd.value, md.value, mml, stream=stream.value)
            print(size, ref, got, got_s)
            assert got == ref and got_s == ref, (size, got, got_s, ref)
        q = p.copy()
        q.ix = 8                                            # below what Setdt's r-range needs
        with pytest.raises(cm.CylmhdError) as e:
            cm.setdt_noalloc(q, Vd.value, md.value, mml)
        assert e.value.code == cm.EINVAL
    finally:
        for b in bufs:
            h.hipFree(b)
        h.hipStreamDestroy(stream)
