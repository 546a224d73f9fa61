"""etd_diag_noalloc, the allocation-free variant of etd_diag: the same
check_soln_new statistics, written to a caller's device buffer, with the
partial sums in a static device array instead of the context's buffer.
Checked bit for bit against etd_diag on two calls in a row, the second on a
smaller context (the array then holds the first call's larger contents), in
FP64 and in the mixed path, and after stepping; a context whose partial sums
exceed the array is refused with a message."""
import ctypes as C
import sys

import numpy as np
import pytest

import pyetd
from conftest import noise_fields, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hip():
    """The HIP runtime libetd itself is linked against (already loaded by
    pyetd.lib(), so dlopen returns that copy), for the caller's device buffer
    of the statistics; no second runtime in the process."""
    pyetd.lib()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.restype = C.c_int
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.restype = C.c_int
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.restype = C.c_int
    H.hipFree.argtypes = [C.c_void_p]
    H.hipDeviceSynchronize.restype = C.c_int
    return H


def noalloc_stats(H, g):
    L = pyetd.lib()
    L.etd_diag_noalloc.restype = C.c_int
    L.etd_diag_noalloc.argtypes = [C.c_void_p, C.c_void_p]
    cnt = 4 * g.V + 1
    buf = C.c_void_p()
    assert H.hipMalloc(C.byref(buf), 8 * cnt) == 0
    host = np.full(cnt, np.nan)
    assert H.hipMemcpy(buf, host.ctypes.data, 8 * cnt, 1) == 0  # host -> device: every slot starts as NaN
    rc = L.etd_diag_noalloc(g.ctx, buf)
    assert rc == 0, L.etd_last_error(g.ctx).decode()
    assert H.hipDeviceSynchronize() == 0  # a device-wide wait also covers the context's own stream
    assert H.hipMemcpy(host.ctypes.data, buf, 8 * cnt, 2) == 0  # device -> host
    assert H.hipFree(buf) == 0
    return host


def context(prm, seed):
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(noise_fields(prm, seed=seed, amp=0.4)):
        g.set_field(v, f)
    return g


def test_diag_noalloc_equals_diag_on_two_calls_the_second_smaller():
    H = hip()
    cases = [
        dict(small_params(n=40, nac=4, nch=1, aniso=1, h=(0.9, 1.0, 1.1)), precision=0),   # 36^3 x 5, FP64
        dict(small_params(n=20, nac=2, nch=1), precision=2),                               # 16^3 x 3, mixed
        dict(small_params(n=16, nac=3, nch=0, elastic=1), precision=0),                    # 12^3 x 3, no CH
    ]
    for i, prm in enumerate(cases):
        say(f"diag_noalloc case {i}: n={prm['n']} nac={prm['nac']} nch={prm['nch']} precision={prm['precision']}")
        g = context(prm, 30 + i)
        for rep in range(2):
            got = noalloc_stats(H, g)
            st, cp = g.diag()
            ref = np.concatenate([st.reshape(-1), [cp]])
            assert np.array_equal(got, ref), (i, rep, np.abs(got - ref).max())
            g.step(1)
        g.close()


def test_diag_noalloc_refuses_contexts_beyond_its_buffer():
    # (3 V + 1 + nch + 1)(N - 1) 8 = (192 + 65) 95 8 = 195320 > 65536 doubles
    prm = dict(small_params(n=100, nac=1, nch=63), precision=2)
    g = pyetd.Etd.from_dict(prm)
    L = pyetd.lib()
    L.etd_diag_noalloc.restype = C.c_int
    L.etd_diag_noalloc.argtypes = [C.c_void_p, C.c_void_p]
    rc = L.etd_diag_noalloc(g.ctx, C.c_void_p(16))  # refused before any use of the pointer
    assert rc == pyetd.ERR_UNSUPPORTED
    assert "195320 doubles" in L.etd_last_error(g.ctx).decode()
    g.close()
