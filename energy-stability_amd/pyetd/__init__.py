"""Python host binding of libetd (ctypes over the C ABI in include/etd.h).

Mirrors the reference's step boundary: an Etd context owns the device state
of one periodic subdomain; step() is the body of the time loop of
source_beta_to_omega/src/ScLETD.c:84-191 and diag() is check_soln_new
(source_beta_to_omega/src/field.c:407-678).  Errors raise EtdError with the
library's message (the reference prints and exits).

The binding never falls back to Python or CPU arithmetic: if libetd.so is
missing the import-time load raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libetd.so")
MAX_VAR = 64

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class EtdError(RuntimeError):
    """A negative return code of libetd: .code is the code (ETD_ERR_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


ERR_ARG, ERR_IO, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NOMEM, ERR_NONFINITE = -1, -2, -3, -4, -5, -6


class EtdParams(C.Structure):
    """Mirror of etd_params (include/etd.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("px", C.c_int), ("py", C.c_int), ("pz", C.c_int),
        ("nghost", C.c_int),
        ("work_dir", C.c_char * 1024), ("data_dir", C.c_char * 1024),
        ("dt", C.c_double), ("t_total", C.c_double), ("epn2", C.c_double),
        ("periodic", C.c_int), ("nac", C.c_int), ("nch", C.c_int),
        ("restart", C.c_int), ("restart_iter", C.c_int), ("restart_dir", C.c_char * 1024),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmin", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("etd2", C.c_int), ("elastic", C.c_int), ("anisotropic", C.c_int),
        ("kkx", C.c_double), ("kky", C.c_double), ("kkz", C.c_double),
        ("variable_mobility", C.c_int),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("LE", C.c_double), ("KE", C.c_double),
        ("a2", C.c_double), ("a3", C.c_double), ("a4", C.c_double),
        ("LCI", C.c_double), ("KCI", C.c_double),
        ("elastic_scale", C.c_double),
        ("C", C.c_double * 36),
        ("eps0", C.c_double * (MAX_VAR * 9)),
        ("lam", C.c_double * (MAX_VAR * 9)),
        ("precision", C.c_int),
        ("checkpoint", C.c_int), ("nchk", C.c_int), ("chk", C.c_int), ("nout", C.c_int), ("approx", C.c_int),
        ("free_energy", C.c_int),
    ]


NFEAT = 16
LOG_COLS = 4 + NFEAT
MLP_PARAMS = 64 * NFEAT + 64 + 64 * 64 + 64 + 64 + 1
CTRL_PLAIN, CTRL_LEARNED = 0, 1


class AdaptOpts(C.Structure):
    """Mirror of etd_adapt_opts."""

    _fields_ = [("t_end", C.c_double), ("dt0", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("tol", C.c_double), ("safety", C.c_double), ("grow", C.c_double), ("shrink", C.c_double),
                ("learned_safety", C.c_double), ("controller", C.c_int), ("extrapolate", C.c_int), ("energy_test", C.c_int),
                ("energy_elastic", C.c_int), ("energy_kind", C.c_int), ("max_attempts", C.c_int),
                ("learned_online", C.c_int)]


class AdaptStats(C.Structure):
    """Mirror of etd_adapt_stats."""

    _fields_ = [("t", C.c_double), ("dt_last", C.c_double), ("energy", C.c_double),
                ("accepted", C.c_int), ("rejected_err", C.c_int), ("rejected_energy", C.c_int),
                ("learned_proposals", C.c_int), ("learned_fallbacks", C.c_int), ("step_evals", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    """Mirror of etd_tuning."""

    _fields_ = [("mp_variant", C.c_int), ("split_waves", C.c_int), ("fft_variant", C.c_int),
                ("ms_per_step", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise EtdError(f"{LIB_PATH} is not built (run `make -C energy-stability_amd`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(EtdParams)
        ctxp = C.c_void_p
        sig = {
            "etd_default_params": (None, [P]),
            "etd_parse_ini": (C.c_int, [C.c_char_p, P]),
            "etd_parse_ini_alpha": (C.c_int, [C.c_char_p, P]),
            "etd_parse_elastic": (C.c_int, [C.c_char_p, P]),
            "etd_parse_anisotropic": (C.c_int, [C.c_char_p, P]),
            "etd_parse_landau": (C.c_int, [C.c_char_p, P]),
            "etd_read_matrices": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp]),
            "etd_eigbasis": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp]),
            "etd_create": (C.c_int, [P, _dp, _dp, _dp, C.POINTER(ctxp)]),
            "etd_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(ctxp)]),
            "etd_close": (None, [ctxp]),
            "etd_last_error": (C.c_char_p, [ctxp]),
            "etd_error_string": (C.c_char_p, [C.c_int]),
            "etd_get_params": (C.c_int, [ctxp, P]),
            "etd_set_field": (C.c_int, [ctxp, C.c_int, _dp]),
            "etd_set_mask": (C.c_int, [ctxp, C.POINTER(C.c_int)]),
            "etd_get_field": (C.c_int, [ctxp, C.c_int, _dp]),
            "etd_init_noise": (C.c_int, [ctxp, C.c_int, C.c_ulonglong, C.c_double, C.c_double]),
            "etd_step": (C.c_int, [ctxp, C.c_int]),
            "etd_diag": (C.c_int, [ctxp, _dp]),
            "etd_write_csv": (C.c_int, [ctxp, C.c_int, _dp]),
            "etd_write_field": (C.c_int, [ctxp, C.c_int, C.c_int]),
            "etd_k_transfer": (C.c_int, [ctxp]),
            "etd_k_elastic": (C.c_int, [ctxp, _dp]),
            "etd_k_nonlinear": (C.c_int, [ctxp, C.c_int, _dp]),
            "etd_k_fft3d": (C.c_int, [ctxp, _dp, C.c_int]),
            "etd_k_modeprod": (C.c_int, [ctxp, C.c_int, _dp, _dp, _dp, C.c_int]),
            "etd_k_phiscale": (C.c_int, [ctxp, C.c_int, _dp, _dp]),
            "etd_time_steps": (C.c_int, [ctxp, C.c_int, C.c_int, _dp, C.POINTER(C.c_int)]),
            "etd_calib_copy": (C.c_int, [ctxp, C.c_size_t]),
            "etd_write_checkpoint": (C.c_int, [ctxp, C.c_int, C.c_int]),
            "etd_read_checkpoint": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_int)]),
            "etd_write_field2b": (C.c_int, [ctxp, C.c_int]),
            "etd_detector_feed": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_ubyte), _fp]),
            "etd_adapt_defaults": (None, [C.POINTER(AdaptOpts)]),
            "etd_run_adaptive": (C.c_int, [ctxp, C.POINTER(AdaptOpts), C.POINTER(AdaptStats), _dp, C.c_int,
                                           C.POINTER(C.c_int)]),
            "etd_get_time": (C.c_double, [ctxp]),
            "etd_set_time": (None, [ctxp, C.c_double]),
            "etd_mlp_create": (C.c_int, [C.c_int, C.POINTER(ctxp)]),
            "etd_mlp_destroy": (None, [ctxp]),
            "etd_mlp_init": (C.c_int, [ctxp, C.c_ulonglong]),
            "etd_mlp_set": (C.c_int, [ctxp, _fp, _fp, _fp, C.c_float, C.c_float]),
            "etd_mlp_get": (C.c_int, [ctxp, _fp, _fp, _fp, _fp, _fp]),
            "etd_mlp_forward": (C.c_int, [ctxp, _fp, C.c_int, _fp]),
            "etd_mlp_loss_grad": (C.c_int, [ctxp, _fp, _fp, C.c_int, _fp, _fp]),
            "etd_mlp_train": (C.c_int, [ctxp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _fp]),
            "etd_mlp_predict": (C.c_int, [ctxp, _dp, _dp]),
            "etd_set_controller": (C.c_int, [ctxp, ctxp]),
            "etd_ensemble_create": (C.c_int, [C.POINTER(ctxp), C.c_int, C.c_int, C.POINTER(ctxp)]),
            "etd_ensemble_run": (C.c_int, [ctxp, C.c_int]),
            "etd_ensemble_destroy": (None, [ctxp]),
            "etd_ensemble_error": (C.c_char_p, [ctxp]),
            "etd_autotune": (C.c_int, [P, _dp, _dp, _dp, C.c_int, C.POINTER(Tuning)]),
            "etd_set_tuning": (C.c_int, [ctxp, C.POINTER(Tuning)]),
            "etd_get_tuning": (C.c_int, [ctxp, C.POINTER(Tuning)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _fptr(a):
    return a.ctypes.data_as(_fp)


def adapt_opts(**kw):
    """etd_adapt_opts with the library's defaults, overridden by keywords."""
    o = AdaptOpts()
    lib().etd_adapt_defaults(C.byref(o))
    for k, v in kw.items():
        if k == "controller" and isinstance(v, str):
            v = {"plain": CTRL_PLAIN, "learned": CTRL_LEARNED}[v]
        setattr(o, k, v)
    return o


def autotune(prm, steps=3):
    """etd_autotune for a parameter dict: the fastest kernel choice (Tuning)."""
    t = Tuning()
    _check(lib().etd_autotune(C.byref(params_from_dict(prm)), None, None, None, steps, C.byref(t)), what="autotune")
    return t


class Ensemble:
    """One hipGraph stepping many contexts (etd_ensemble)."""

    def __init__(self, members, steps_per_launch):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.h = C.c_void_p()
        rc = lib().etd_ensemble_create(arr, len(self.members), steps_per_launch, C.byref(self.h))
        if rc != 0:
            msg = lib().etd_ensemble_error(self.h).decode() if self.h else ""
            if self.h:
                lib().etd_ensemble_destroy(self.h)
            self.h = C.c_void_p()
            raise EtdError(f"ensemble_create: {lib().etd_error_string(rc).decode()}: {msg}")

    def run(self, launches=1):
        rc = lib().etd_ensemble_run(self.h, launches)
        if rc != 0:
            raise EtdError(f"ensemble_run: {lib().etd_ensemble_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().etd_ensemble_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned dt controller's network (etd_mlp) on one GPU."""

    def __init__(self, device=0, seed=1):
        self.h = C.c_void_p()
        _check(lib().etd_mlp_create(device, C.byref(self.h)), what="mlp_create")
        if seed != 1:
            _check(lib().etd_mlp_init(self.h, seed), what="mlp_init")

    def close(self):
        if self.h:
            lib().etd_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, params, fmean, fstd, ymean, ystd):
        p = np.ascontiguousarray(params, dtype=np.float32)
        fm = np.ascontiguousarray(fmean, dtype=np.float32)
        fs = np.ascontiguousarray(fstd, dtype=np.float32)
        assert p.size == MLP_PARAMS and fm.size == NFEAT and fs.size == NFEAT
        _check(lib().etd_mlp_set(self.h, _fptr(p), _fptr(fm), _fptr(fs), float(ymean), float(ystd)), what="mlp_set")

    def get(self):
        p = np.empty(MLP_PARAMS, np.float32)
        fm = np.empty(NFEAT, np.float32)
        fs = np.empty(NFEAT, np.float32)
        ym = np.zeros(1, np.float32)
        ys = np.zeros(1, np.float32)
        _check(lib().etd_mlp_get(self.h, _fptr(p), _fptr(fm), _fptr(fs), _fptr(ym), _fptr(ys)), what="mlp_get")
        return p, fm, fs, float(ym[0]), float(ys[0])

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEAT)
        y = np.empty(x.shape[0], np.float32)
        _check(lib().etd_mlp_forward(self.h, _fptr(x), x.shape[0], _fptr(y)), what="mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEAT)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        loss = np.zeros(1, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        _check(lib().etd_mlp_loss_grad(self.h, _fptr(x), _fptr(t), x.shape[0], _fptr(loss), _fptr(g)),
               what="mlp_loss_grad")
        return float(loss[0]), g

    def train(self, x, t, epochs, batch, lr, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEAT)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        el = np.zeros(epochs, np.float32)
        _check(lib().etd_mlp_train(self.h, _fptr(x), _fptr(t), x.shape[0], epochs, batch, float(lr), seed, _fptr(el)),
               what="mlp_train")
        return el

    def predict(self, features):
        f = np.ascontiguousarray(features, dtype=np.float64).reshape(NFEAT)
        d = np.zeros(1)
        _check(lib().etd_mlp_predict(self.h, _ptr(f), _ptr(d)), what="mlp_predict")
        return float(d[0])


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = lib().etd_error_string(rc).decode()
        if ctx:
            detail = lib().etd_last_error(ctx)
            if detail:
                msg += ": " + detail.decode()
        raise EtdError(f"{what}: {msg}", rc)


def default_params():
    p = EtdParams()
    lib().etd_default_params(C.byref(p))
    return p


def parse_inputs(directory, suffix):
    """input<suffix>.ini + elastic_input.txt + anisotropic_input.txt (+ landau_input.txt)."""
    p = default_params()
    d = os.fspath(directory)
    _check(lib().etd_parse_ini(os.path.join(d, f"input{suffix}.ini").encode(), C.byref(p)), what="parse_ini")
    if p.elastic:
        _check(lib().etd_parse_elastic(os.path.join(d, "elastic_input.txt").encode(), C.byref(p)), what="elastic_input")
    if p.anisotropic == 1:
        _check(lib().etd_parse_anisotropic(os.path.join(d, "anisotropic_input.txt").encode(), C.byref(p)),
               what="anisotropic_input")
    lp = os.path.join(d, "landau_input.txt")
    if os.path.exists(lp):
        _check(lib().etd_parse_landau(lp.encode(), C.byref(p)), what="landau_input")
    return p


def eigbasis(n, bcid=3):
    V = np.zeros(n * n)
    Vi = np.zeros(n * n)
    D = np.zeros(n)
    _check(lib().etd_eigbasis(n, bcid, _ptr(V), _ptr(Vi), _ptr(D)), what="eigbasis")
    return V.reshape(n, n).T.copy(), Vi.reshape(n, n).T.copy(), D


def params_from_dict(prm):
    """Build etd_params from the test/bench dictionary form (same keys as the oracle)."""
    p = default_params()
    n = prm["n"]
    p.nx = p.ny = p.nz = n
    p.nghost = prm.get("g", 2)
    p.dt = prm["dt"]
    p.t_total = prm.get("t_total", 1.0)
    p.epn2 = prm["epn2"]
    p.nac = prm["nac"]
    p.nch = prm["nch"]
    p.etd2 = prm.get("etd2", 0)
    p.elastic = prm.get("elastic", 0)
    p.anisotropic = prm.get("aniso", 0)
    kk = prm.get("kk", (1.0, 1.0, 1.0))
    p.kkx, p.kky, p.kkz = kk
    h = prm.get("h", (1.0, 1.0, 1.0))
    N = n - 2 * p.nghost
    procs = prm.get("procs", (1, 1, 1))  # a decomposed run: boxes per axis
    p.px, p.py, p.pz = procs
    p.xmin = p.ymin = p.zmin = 0.0
    p.xmax, p.ymax, p.zmax = h[0] * N * procs[0], h[1] * N * procs[1], h[2] * N * procs[2]
    p.hx, p.hy, p.hz = h
    p.LE = prm.get("LE", 1.0)
    p.KE = prm.get("KE", 42.0)
    if "landau" in prm:
        p.a2, p.a3, p.a4 = prm["landau"]
    p.LCI = prm.get("LCI", 1.0)
    p.KCI = prm.get("KCI", 2.0)
    p.elastic_scale = prm.get("elastic_scale", 1.0)
    Cv = np.asarray(prm.get("C", np.zeros(36)), dtype=np.float64).reshape(-1)
    for i in range(36):
        p.C[i] = Cv[i]
    nac = p.nac
    nel = p.nch if p.elastic == 2 else nac  # the variants the elastic force acts on
    if nel:
        eps = np.asarray(prm.get("eps0", np.zeros((nel, 3, 3))), dtype=np.float64).reshape(-1)[: 9 * nel]
        for i, x in enumerate(eps):
            p.eps0[i] = x
    if nac:
        lam = prm.get("lam")
        if lam is None:
            lam = np.array([np.diag(kk)] * nac)
        lam = np.asarray(lam, dtype=np.float64).reshape(-1)[: 9 * nac]
        for i, x in enumerate(lam):
            p.lam[i] = x
    p.precision = prm.get("precision", 0)
    p.free_energy = prm.get("free_energy", 0)
    p.variable_mobility = prm.get("varmob", 0)
    p.periodic = prm.get("periodic", 1)
    return p


class Etd:
    """One subdomain on the GPU, or a decomposed run (px x py x pz boxes):
    then fields, nonlinear terms and forces come per box, box b at index b
    of the leading axis (b = (bz py + by) px + bx)."""

    def __init__(self, params, V=None, Vi=None, D=None, carry=True):
        """carry=False: every step transforms u itself (ETD_CARRY=0) instead of
        carrying the AC spectrum from the step before."""
        if not carry:
            os.environ["ETD_CARRY"] = "0"
        try:
            self._create(params, V, Vi, D)
        finally:
            if not carry:
                os.environ.pop("ETD_CARRY", None)

    def _create(self, params, V, Vi, D):
        self.ctx = C.c_void_p()
        self.p = params
        self.n = params.nx
        self.V = params.nac + params.nch
        self.P = params.px * params.py * params.pz
        if V is not None:
            Vc = np.ascontiguousarray(np.asarray(V).T).reshape(-1)
            Vic = np.ascontiguousarray(np.asarray(Vi).T).reshape(-1)
            Dc = np.ascontiguousarray(np.asarray(D, dtype=np.float64))
            rc = lib().etd_create(C.byref(params), _ptr(Vc), _ptr(Vic), _ptr(Dc), C.byref(self.ctx))
        else:
            rc = lib().etd_create(C.byref(params), None, None, None, C.byref(self.ctx))
        if rc != 0:
            ctx = self.ctx if self.ctx else None
            err = lib().etd_last_error(ctx).decode() if ctx else ""
            if ctx:
                lib().etd_close(ctx)
            self.ctx = None
            raise EtdError(f"etd_create: {lib().etd_error_string(rc).decode()} {err}")

    @classmethod
    def from_dict(cls, prm, carry=True):
        return cls(params_from_dict(prm), carry=carry)

    @classmethod
    def open(cls, directory, suffix):
        obj = cls.__new__(cls)
        obj.ctx = C.c_void_p()
        rc = lib().etd_open(os.fspath(directory).encode(), str(suffix).encode(), C.byref(obj.ctx))
        if rc != 0:
            raise EtdError(f"etd_open: {lib().etd_error_string(rc).decode()}")
        obj.p = EtdParams()
        lib().etd_get_params(obj.ctx, C.byref(obj.p))
        obj.n = obj.p.nx
        obj.V = obj.p.nac + obj.p.nch
        obj.P = obj.p.px * obj.p.py * obj.p.pz
        return obj

    def close(self):
        if getattr(self, "ctx", None):
            lib().etd_close(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # fields
    def set_mask(self, mask):
        """Grain mask (n^3 ints, 0 outside the grain) or None."""
        L = lib()
        if mask is None:
            _check(L.etd_set_mask(self.ctx, None), self.ctx, "set_mask")
            return
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32).reshape(-1))
        assert m.size == self.n ** 3
        _check(L.etd_set_mask(self.ctx, m.ctypes.data_as(C.POINTER(C.c_int))), self.ctx, "set_mask")

    def _boxes(self, a):
        n = self.n
        return a.reshape(n, n, n) if self.P == 1 else a.reshape(self.P, n, n, n)

    def set_field(self, var, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
        if a.size != self.P * self.n ** 3:
            raise EtdError("set_field: wrong size")
        _check(lib().etd_set_field(self.ctx, var, _ptr(a)), self.ctx, "set_field")

    def get_field(self, var):
        a = np.empty(self.P * self.n ** 3)
        _check(lib().etd_get_field(self.ctx, var, _ptr(a)), self.ctx, "get_field")
        return self._boxes(a)

    def init_noise(self, var, seed, mean, amp):
        _check(lib().etd_init_noise(self.ctx, var, seed, mean, amp), self.ctx, "init_noise")

    # the step and its parts
    def step(self, nsteps=1):
        _check(lib().etd_step(self.ctx, nsteps), self.ctx, "step")

    def diag(self):
        st = np.zeros(4 * self.V + 1)
        _check(lib().etd_diag(self.ctx, _ptr(st)), self.ctx, "diag")
        return st[:-1].reshape(self.V, 4), st[-1]

    def transfer(self):
        _check(lib().etd_k_transfer(self.ctx), self.ctx, "transfer")

    def elastic(self, want=True):
        """etd_k_elastic; the force on the AC fields (ELASTIC = 1) or the CH fields (2)."""
        if want:
            nel = self.p.nac if self.p.elastic == 1 else self.p.nch if self.p.elastic == 2 else 0
            out = np.empty(max(nel, 1) * self.P * self.n ** 3)
            _check(lib().etd_k_elastic(self.ctx, _ptr(out)), self.ctx, "elastic")
            n = self.n
            out = out[: nel * self.P * n ** 3]
            return out.reshape(nel, n, n, n) if self.P == 1 else out.reshape(nel, self.P, n, n, n)
        _check(lib().etd_k_elastic(self.ctx, None), self.ctx, "elastic")
        return None

    def fft3d(self, x, direction):
        """In-place 3D FFT of an N^3 complex array on the GPU (unscaled)."""
        a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
        buf = a.view(np.float64).reshape(-1)
        _check(lib().etd_k_fft3d(self.ctx, _ptr(buf), int(direction)), self.ctx, "fft3d")
        return a

    def nonlinear(self, var):
        out = np.empty(self.P * self.n ** 3)
        _check(lib().etd_k_nonlinear(self.ctx, var, _ptr(out)), self.ctx, "nonlinear")
        return self._boxes(out)

    def modeprod(self, A, u, rotate=True, slowest=False):
        """One mode product by A (n x n): over the fastest axis (PUX,
        out[(b + n c) + n^2 m] rotated, out[m + n (b + n c)] not), or with
        slowest=True over the slowest axis (FP16 split-2 contexts:
        out[m + n col] rotated, out[col + n^2 m] not, col = a + n b)."""
        n = self.n
        Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T).reshape(-1)
        inp = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
        out = np.empty(n ** 3)
        code = (2 if rotate else 3) if slowest else (1 if rotate else 0)
        _check(lib().etd_k_modeprod(self.ctx, n, _ptr(Acm), _ptr(inp), _ptr(out), code), self.ctx, "modeprod")
        return out

    def phiscale(self, var, uhat, nhat):
        u = np.ascontiguousarray(np.asarray(uhat, dtype=np.float64).reshape(-1)).copy()
        nh = np.ascontiguousarray(np.asarray(nhat, dtype=np.float64).reshape(-1))
        _check(lib().etd_k_phiscale(self.ctx, var, _ptr(u), _ptr(nh)), self.ctx, "phiscale")
        return u

    # the beta->alpha output and restart formats
    def write_checkpoint(self, irun, chk):
        _check(lib().etd_write_checkpoint(self.ctx, irun, chk), self.ctx, "write_checkpoint")

    def read_checkpoint(self, chk):
        irun = C.c_int(0)
        _check(lib().etd_read_checkpoint(self.ctx, chk, C.byref(irun)), self.ctx, "read_checkpoint")
        return irun.value

    def write_field2b(self, irun):
        _check(lib().etd_write_field2b(self.ctx, irun), self.ctx, "write_field2b")

    def detector_feed(self, var, want_binary=True):
        """(bits (N/2)^3 / 8 uint8, binary (N/2)^3 float32 or None) of AC field var."""
        d = (self.n - 2 * self.p.nghost) // 2
        bits = np.zeros(d ** 3 // 8, np.uint8)
        binary = np.zeros(d ** 3, np.float32) if want_binary else None
        _check(lib().etd_detector_feed(self.ctx, var, bits.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                       _fptr(binary) if want_binary else None), self.ctx, "detector_feed")
        return bits, binary

    def calib_copy(self, nbytes):
        """One plain device copy of nbytes (HBM counter calibration)."""
        _check(lib().etd_calib_copy(self.ctx, nbytes), self.ctx, "calib_copy")

    @property
    def time(self):
        return lib().etd_get_time(self.ctx)

    @time.setter
    def time(self, t):
        lib().etd_set_time(self.ctx, t)

    def set_tuning(self, t):
        _check(lib().etd_set_tuning(self.ctx, C.byref(t)), self.ctx, "set_tuning")

    def get_tuning(self):
        t = Tuning()
        _check(lib().etd_get_tuning(self.ctx, C.byref(t)), self.ctx, "get_tuning")
        return t

    def set_controller(self, mlp):
        _check(lib().etd_set_controller(self.ctx, mlp.h if mlp is not None else None), self.ctx, "set_controller")
        self._mlp = mlp

    def run_adaptive(self, log_cap=0, **kw):
        """etd_run_adaptive; returns (stats dict, log rows [t, dt, err, E, features...])."""
        o = adapt_opts(**kw)
        st = AdaptStats()
        log = np.zeros((max(log_cap, 1), LOG_COLS))
        rows = C.c_int(0)
        rc = lib().etd_run_adaptive(self.ctx, C.byref(o), C.byref(st), _ptr(log) if log_cap else None, log_cap,
                                    C.byref(rows))
        try:
            _check(rc, self.ctx, "run_adaptive")
        except EtdError as e:  # the run's statistics up to the failure
            e.stats, e.log = st.as_dict(), log[: rows.value].copy()
            raise
        return st.as_dict(), log[: rows.value].copy()

    def time_steps(self, nsteps, with_diag=True):
        """etd_time_steps: (ms[10], launches[8]); ms[9] is the count of
        field passes the mode-product launches made."""
        ms = np.zeros(10)
        launches = (C.c_int * 8)()
        _check(lib().etd_time_steps(self.ctx, nsteps, 1 if with_diag else 0, _ptr(ms), launches), self.ctx,
               "time_steps")
        return ms, list(launches)
