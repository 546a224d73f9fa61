"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, to check results.  The product (libetd) never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("g", C.c_int), ("nac", C.c_int), ("nch", C.c_int),
        ("etd2", C.c_int), ("elastic", C.c_int), ("aniso", C.c_int),
        ("dt", C.c_double), ("epn2", C.c_double),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("kk", C.c_double * 3),
        ("LE", C.c_double), ("KE", C.c_double),
        ("a2", C.c_double), ("a3", C.c_double), ("a4", C.c_double),
        ("LCI", C.c_double), ("KCI", C.c_double),
        ("elastic_scale", C.c_double),
        ("C", C.c_double * 36),
        ("eps0", _dp), ("lam", _dp), ("V", _dp), ("Vi", _dp), ("D", _dp),
        ("free_energy", C.c_int), ("mask", C.POINTER(C.c_int)), ("varmob", C.c_int),
        ("physical", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_field.restype = _dp
        L.orc_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_felas.restype = _dp
        L.orc_felas.argtypes = [C.c_void_p, C.c_int]
        L.orc_ghost_fill.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_transfer.argtypes = [C.c_void_p]
        L.orc_elastic.argtypes = [C.c_void_p]
        L.orc_nonlinear.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_modeprod.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int]
        L.orc_phi_update.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.orc_step_pair.argtypes = [C.c_void_p, C.c_double, _dp]
        L.orc_diag.argtypes = [C.c_void_p, _dp]
        L.orc_eigbasis.restype = C.c_int
        L.orc_eigbasis.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_fft3d.restype = C.c_int
        L.orc_fft3d.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_bpq.restype = C.c_double
        L.orc_bpq.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_calphad.argtypes = [_dp, _dp]
        L.orc_calphad_terms.argtypes = [_dp]
        L.orc_div_m_grad.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_group_create.restype = C.c_void_p
        L.orc_group_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]
        L.orc_group_destroy.argtypes = [C.c_void_p]
        L.orc_group_transfer.argtypes = [C.c_void_p]
        L.orc_group_elastic.argtypes = [C.c_void_p]
        L.orc_group_step.argtypes = [C.c_void_p]
        L.orc_group_diag.argtypes = [C.c_void_p, _dp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(_dp)


def eigbasis(n, bcid=3):
    """Closed-form V, V^-1, D (column-major n x n as numpy arrays A[r, c])."""
    V = np.zeros(n * n)
    Vi = np.zeros(n * n)
    D = np.zeros(n)
    rc = lib().orc_eigbasis(n, bcid, _ptr(V), _ptr(Vi), _ptr(D))
    if rc != 0:
        raise ValueError("orc_eigbasis failed")
    # column-major storage -> A[r, c] = buf[r + n c]
    return V.reshape(n, n).T.copy(), Vi.reshape(n, n).T.copy(), D


def fft3d(x, direction):
    """Oracle 3D complex FFT of an N^3 complex array (x fastest = last axis)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    N = a.shape[0]
    buf = a.view(np.float64).reshape(-1)
    lib().orc_fft3d(N, _ptr(buf), int(direction))
    return a


def calphad():
    """(G_bcc, G_hcp) at the alpha driver's fixed compositions."""
    a, b = C.c_double(), C.c_double()
    lib().orc_calphad(C.byref(a), C.byref(b))
    return a.value, b.value


def calphad_terms():
    t = np.zeros(13)
    lib().orc_calphad_terms(_ptr(t))
    return t


def div_m_grad(M, a, h=(1.0, 1.0, 1.0)):
    """div(M grad a) as the variable-mobility CH forms it (n^3 arrays)."""
    n = M.shape[0]
    Mf = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
    af = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    out = np.zeros(n ** 3)
    lib().orc_div_m_grad(n, h[0], h[1], h[2], _ptr(Mf), _ptr(af), _ptr(out))
    return out.reshape(n, n, n)


def modeprod(A, u, rotate=True):
    """out[(b + n c) + n^2 m] = sum_a A[m, a] u[a + n b + n^2 c] (flat arrays)."""
    n = A.shape[0]
    Acm = np.ascontiguousarray(A.T).reshape(-1)  # column-major
    inp = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    out = np.zeros(n ** 3)
    lib().orc_modeprod(n, _ptr(Acm), _ptr(inp), _ptr(out), 1 if rotate else 0)
    return out


class Oracle:
    """Owns an orc_ctx built from a parameter dict (see etd_params in tests)."""

    def __init__(self, prm):
        self.prm = prm
        n = prm["n"]
        self.n = n
        self.nac = prm["nac"]
        self.nch = prm["nch"]
        self.V = self.nac + self.nch
        p = OrcParams()
        p.n = n
        p.g = prm.get("g", 2)
        p.nac = prm["nac"]
        p.nch = prm["nch"]
        p.etd2 = prm.get("etd2", 0)
        p.elastic = prm.get("elastic", 0)
        p.aniso = prm.get("aniso", 0)
        p.dt = prm["dt"]
        p.epn2 = prm["epn2"]
        h = prm.get("h", (1.0, 1.0, 1.0))
        p.hx, p.hy, p.hz = h
        for i in range(3):
            p.kk[i] = prm.get("kk", (1.0, 1.0, 1.0))[i]
        p.LE = prm.get("LE", 1.0)
        p.KE = prm.get("KE", 42.0)
        p.a2, p.a3, p.a4 = prm.get("landau", (2376.0 / 2372.0, (7128.0 + 12.0 * 2372.0) / 2372.0,
                                               (4752.0 + 12.0 * 2372.0) / 2372.0))
        p.LCI = prm.get("LCI", 1.0)
        p.KCI = prm.get("KCI", 2.0)
        p.elastic_scale = prm.get("elastic_scale", 1.0)
        Cv = np.asarray(prm.get("C", np.zeros(36)), dtype=np.float64).reshape(36)
        for i in range(36):
            p.C[i] = Cv[i]
        nac = max(self.nac, 1)
        self._eps = np.ascontiguousarray(np.asarray(prm.get("eps0", np.zeros((nac, 3, 3))), dtype=np.float64).reshape(-1))
        if self._eps.size < 9 * nac:
            self._eps = np.zeros(9 * nac)
        lam = prm.get("lam")
        if lam is None:
            kk = prm.get("kk", (1.0, 1.0, 1.0))
            lam = np.array([np.diag(kk)] * nac)
        self._lam = np.ascontiguousarray(np.asarray(lam, dtype=np.float64).reshape(-1))
        Vm = prm.get("Vmat")
        if Vm is None:
            # boundary id 3 (neighbours on both sides) for a periodic box, 0 (physical both sides) otherwise
            Vm, Vim, D = eigbasis(n, 3 if prm.get("periodic", 1) else 0)
        else:
            Vim = prm["Vinv"]
            D = prm["D"]
        self._V = np.ascontiguousarray(np.asarray(Vm).T).reshape(-1)
        self._Vi = np.ascontiguousarray(np.asarray(Vim).T).reshape(-1)
        self._D = np.ascontiguousarray(np.asarray(D, dtype=np.float64))
        p.eps0 = _ptr(self._eps)
        p.lam = _ptr(self._lam)
        p.V = _ptr(self._V)
        p.Vi = _ptr(self._Vi)
        p.D = _ptr(self._D)
        p.free_energy = prm.get("free_energy", 0)
        p.varmob = prm.get("varmob", 0)
        p.physical = 0 if prm.get("periodic", 1) else 1
        mask = prm.get("mask")
        if mask is not None:
            self._mask = np.ascontiguousarray(np.asarray(mask, dtype=np.int32).reshape(-1))
            p.mask = self._mask.ctypes.data_as(C.POINTER(C.c_int))
        self._p = p
        self.ctx = lib().orc_create(C.byref(p))
        if not self.ctx:
            raise ValueError("orc_create rejected the parameters")

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().orc_destroy(self.ctx)
            self.ctx = None

    def _arr(self, ptr):
        return np.ctypeslib.as_array(ptr, shape=(self.n ** 3,))

    def field(self, v):
        return self._arr(lib().orc_field(self.ctx, v))

    def set_field(self, v, data):
        self.field(v)[:] = np.asarray(data, dtype=np.float64).reshape(-1)

    def get_field(self, v):
        return self.field(v).copy().reshape(self.n, self.n, self.n)

    def felas(self, p):
        return self._arr(lib().orc_felas(self.ctx, p)).copy().reshape(self.n, self.n, self.n)

    def transfer(self):
        lib().orc_transfer(self.ctx)

    def elastic(self):
        lib().orc_elastic(self.ctx)

    def nonlinear(self, v):
        out = np.zeros(self.n ** 3)
        lib().orc_nonlinear(self.ctx, v, _ptr(out))
        return out.reshape(self.n, self.n, self.n)

    def step(self, nsteps=1):
        for _ in range(nsteps):
            lib().orc_step(self.ctx)

    def set_dt(self, dt):
        lib().orc_set_dt(self.ctx, dt)

    def step_pair(self, dt):
        """Shared-N step-doubling pair: returns u_n + dt (V x n^3) and leaves
        the fields at u_n + dt/2, dt at dt/2."""
        full = np.empty(self.V * self.n ** 3)
        lib().orc_step_pair(self.ctx, dt, _ptr(full))
        return full.reshape(self.V, self.n ** 3)

    def diag(self):
        st = np.zeros(4 * self.V + 1)
        lib().orc_diag(self.ctx, _ptr(st))
        return st[:-1].reshape(self.V, 4), st[-1]

    def bpq(self, N, gx, gy, gz, p, q):
        return lib().orc_bpq(C.byref(self._p), N, gx, gy, gz, p, q)


class OracleGroup:
    """A decomposed run: px x py x pz boxes (omega's procs line, periodic),
    box b at (b % px, (b // px) % py, b // (px py)), each an Oracle with the
    same parameters (h = L / (p N) in prm["h"])."""

    def __init__(self, prm, procs):
        self.px, self.py, self.pz = procs
        self.P = self.px * self.py * self.pz
        self.boxes = [Oracle(prm) for _ in range(self.P)]
        self.n = prm["n"]
        self.V = self.boxes[0].V
        arr = (C.c_void_p * self.P)(*[b.ctx for b in self.boxes])
        self.grp = lib().orc_group_create(arr, self.px, self.py, self.pz)
        if not self.grp:
            raise ValueError("orc_group_create rejected the boxes")

    def __del__(self):
        if getattr(self, "grp", None):
            lib().orc_group_destroy(self.grp)
            self.grp = None

    def box(self, b):
        return self.boxes[b]

    def coords(self, b):
        return b % self.px, (b // self.px) % self.py, b // (self.px * self.py)

    def transfer(self):
        lib().orc_group_transfer(self.grp)

    def elastic(self):
        lib().orc_group_elastic(self.grp)

    def step(self, nsteps=1):
        for _ in range(nsteps):
            lib().orc_group_step(self.grp)

    def diag(self):
        st = np.zeros(4 * self.V + 1)
        lib().orc_group_diag(self.grp, _ptr(st))
        return st[:-1].reshape(self.V, 4), st[-1]


def ghost_fill(u, g=2):
    a = np.ascontiguousarray(u, dtype=np.float64).copy()
    n = a.shape[0]
    lib().orc_ghost_fill(n, g, _ptr(a.reshape(-1)))
    return a


def threads():
    return lib().orc_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def set_mutation(m):
    """Plant a known mistake in the oracle (tests of the tests): 1 flips the
    sign of CH_f, 2 uses lambda for lambda^2 in the CH operator; 0 undoes."""
    lib().orc_set_mutation(int(m))
