"""ctypes binding of libelastoplasm_hip.so (include/elastoplasm.h).

This is the same binding a Julia ``ccall`` shim makes: raw
pointers into the column-major host arrays of ``Mesh`` / ``Point``.  There is
no CPU fallback: if the library is missing or no GPU is visible, the calls
fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .types import Mesh, Point

LIBDIR = Path(__file__).resolve().parent.parent / "lib"
LIBNAME = "libelastoplasm_hip.so"

EP_OK = 0
ERRORS = {-1: "EP_E_ARG", -2: "EP_E_CONFIG", -3: "EP_E_DOMAIN", -4: "EP_E_NONFINITE",
          -5: "EP_E_HIP", -6: "EP_E_STATE"}

_dp = C.POINTER(C.c_double)


class EpConfig(C.Structure):
    _fields_ = [("dim", C.c_int32), ("basis", C.c_int32), ("transfer", C.c_int32), ("musl", C.c_int32),
                ("deform", C.c_int32), ("plast", C.c_int32), ("locking", C.c_int32), ("precision", C.c_int32),
                ("nmp", C.c_int64), ("nel", C.c_int64 * 3), ("nno", C.c_int64 * 3),
                ("h", C.c_double * 3), ("x0", C.c_double * 3), ("xB", (C.c_double * 2) * 3),
                ("C_pf", C.c_double), ("damping", C.c_double), ("Kc", C.c_double), ("Gc", C.c_double),
                ("Hp", C.c_double), ("c", C.c_double), ("Del", C.c_double * 36),
                ("surrogate_tol", C.c_double), ("domain", C.c_int32), ("nonloc", C.c_int32),
                ("reserved", C.c_int32 * 6), ("nonloc_ls", C.c_double)]


class EpGridView(C.Structure):
    _fields_ = [("x", _dp), ("bcs", C.POINTER(C.c_uint8))]


_PFIELDS = ["x", "v", "l0", "l", "F", "b", "gradv", "eps", "sig", "tau", "epII", "rho0", "rho", "Om0", "Om",
            "n0", "n", "J", "dJ", "c0", "cr", "phi", "dlam", "vmax", "Bij"]


class EpParticlesView(C.Structure):
    _fields_ = [(f, _dp) for f in _PFIELDS]


class EpGridFields(C.Structure):
    _fields_ = [("m", _dp), ("p", _dp), ("oobf", _dp), ("a", _dp), ("v", _dp)]


class EpStepParams(C.Structure):
    _fields_ = [("T", C.c_double), ("G", C.c_double), ("tg", C.c_double), ("ramp", C.c_int32),
                ("with_plast", C.c_int32), ("dt_mode", C.c_int32), ("record_dt", C.c_int32),
                ("dt_fixed", C.c_double)]


class EpStepResult(C.Structure):
    _fields_ = [("t", C.c_double), ("last_dt", C.c_double), ("steps", C.c_int64), ("it", C.c_int64)]


class EpTuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("brick", C.c_int32 * 3), ("ms_per_step", C.c_double), ("ntried", C.c_int32),
                ("tried_block", C.c_int32 * 16), ("tried_brick", (C.c_int32 * 3) * 16), ("tried_ms", C.c_double * 16)]


class EpStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("steps", C.c_int64), ("fallbacks", C.c_int64),
                ("predictions", C.c_int64), ("j2_maxit", C.c_int64), ("max_residual", C.c_double),
                ("resorts", C.c_int64)]


EXPORTS = ["ep_version", "ep_last_error", "ep_create", "ep_destroy", "ep_upload_grid", "ep_upload_particles",
           "ep_download_particles", "ep_download_grid", "ep_set_time", "ep_set_resort", "ep_set_layout", "ep_step", "ep_get_dts", "ep_profile", "ep_profile_kernel",
           "ep_kernel_times", "ep_get_stats", "ep_surrogate_create", "ep_surrogate_destroy",
           "ep_surrogate_get_params", "ep_surrogate_set_params", "ep_surrogate_sample", "ep_surrogate_get_samples",
           "ep_surrogate_train", "ep_surrogate_grad", "ep_surrogate_predict", "ep_surrogate_harvest",
           "ep_attach_surrogate",
           "ep_ensemble_create", "ep_ensemble_destroy", "ep_ensemble_step", "ep_ensemble_tune",
           "ep_ensemble_get_stats", "ep_grf_gauss", "ep_grf_gauss_noalloc", "ep_slump_points"]

NPARAM = 34692
_fp = C.POINTER(C.c_float)

_lib = None


def lib_path() -> Path:
    return Path(os.environ.get("EP_HIP_LIB", LIBDIR / LIBNAME))


def load() -> C.CDLL:
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not p.exists():
        raise RuntimeError(f"{p} is missing: run __graft_entry__.build() (make -C elastoplasm.jl_amd)")
    L = C.CDLL(str(p))
    L.ep_version.restype = C.c_int
    L.ep_last_error.restype = C.c_char_p
    L.ep_create.argtypes = [C.POINTER(EpConfig), C.POINTER(C.c_void_p)]
    L.ep_destroy.argtypes = [C.c_void_p]
    L.ep_destroy.restype = None
    L.ep_upload_grid.argtypes = [C.c_void_p, C.POINTER(EpGridView)]
    L.ep_upload_particles.argtypes = [C.c_void_p, C.POINTER(EpParticlesView)]
    L.ep_download_particles.argtypes = [C.c_void_p, C.POINTER(EpParticlesView)]
    L.ep_download_grid.argtypes = [C.c_void_p, C.POINTER(EpGridFields)]
    L.ep_set_time.argtypes = [C.c_void_p, C.c_double]
    L.ep_set_resort.argtypes = [C.c_void_p, C.c_int32]
    L.ep_set_layout.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.ep_step.argtypes = [C.c_void_p, C.c_int32, C.POINTER(EpStepParams), C.POINTER(EpStepResult)]
    L.ep_get_dts.argtypes = [C.c_void_p, _dp, C.c_int64, C.POINTER(C.c_int64)]
    L.ep_profile.argtypes = [C.c_void_p, C.c_int32]
    L.ep_profile_kernel.argtypes = [C.c_void_p, C.c_char_p]
    L.ep_kernel_times.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, _dp,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.ep_get_stats.argtypes = [C.c_void_p, C.POINTER(EpStats)]
    L.ep_surrogate_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
    L.ep_surrogate_destroy.argtypes = [C.c_void_p]
    L.ep_surrogate_destroy.restype = None
    L.ep_surrogate_get_params.argtypes = [C.c_void_p, _fp]
    L.ep_surrogate_set_params.argtypes = [C.c_void_p, _fp]
    L.ep_surrogate_sample.argtypes = [C.c_void_p, C.c_int64, C.c_uint64]
    L.ep_surrogate_get_samples.argtypes = [C.c_void_p, C.c_int64, C.c_int64, _fp, _fp, _dp]
    L.ep_surrogate_train.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double, _fp]
    L.ep_surrogate_grad.argtypes = [C.c_void_p, _fp, _fp, C.c_int64, _fp, _fp]
    L.ep_surrogate_predict.argtypes = [C.c_void_p, _fp, C.c_int64, _fp]
    L.ep_surrogate_harvest.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
    L.ep_attach_surrogate.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    L.ep_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
    L.ep_ensemble_destroy.argtypes = [C.c_void_p]
    L.ep_ensemble_destroy.restype = None
    L.ep_ensemble_step.argtypes = [C.c_void_p, C.c_int32, C.POINTER(EpStepParams), C.POINTER(EpStepResult)]
    L.ep_ensemble_tune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(EpStepParams), C.POINTER(EpTuneResult)]
    L.ep_ensemble_get_stats.argtypes = [C.c_void_p, C.POINTER(EpStats)]
    _lib = L
    return L


class EpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _check(r):
    if r != EP_OK:
        raise EpError(r, load().ep_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    if not (a.flags["F_CONTIGUOUS"] and a.dtype == np.float64):
        raise ValueError("arrays passed to the C ABI must be Float64 column-major")
    return a.ctypes.data_as(_dp)


def make_config(mesh: Mesh, mpts: Point, cmpr, instr) -> EpConfig:
    ops = instr["ops"]
    cfg = EpConfig()
    d = mesh.dim
    cfg.dim = d
    cfg.basis, cfg.transfer, cfg.musl = ops["basis"], ops["transfer"], ops["musl"]
    cfg.deform, cfg.plast, cfg.locking, cfg.precision = ops["deform"], ops["plast"], ops["locking"], ops["precision"]
    cfg.nmp = mpts.nmp
    for i in range(d):
        cfg.nel[i] = int(mesh.nel[i])
        cfg.nno[i] = int(mesh.nno[i])
        cfg.h[i] = float(mesh.h[i])
        cfg.x0[i] = float(mesh.x0[i])
        cfg.xB[i][0] = float(mesh.xB[i, 0])
        cfg.xB[i][1] = float(mesh.xB[i, 1])
    cfg.C_pf, cfg.damping = ops["C_pf"], ops["damping"]
    cfg.Kc, cfg.Gc, cfg.Hp, cfg.c = cmpr.Kc, cmpr.Gc, cmpr.Hp, cmpr.c
    Del = np.asarray(cmpr.Del, dtype=np.float64).ravel(order="F")
    for k, v in enumerate(Del):
        cfg.Del[k] = v
    cfg.surrogate_tol = float(instr.get("surrogate_tol", 1e-3))
    cfg.domain = ops.get("domain", 0)
    nl = instr.get("nonloc", {})
    cfg.nonloc = 1 if nl.get("status", False) else 0   # ep_config.nonlocal
    cfg.nonloc_ls = float(np.max(nl.get("ls", 0.5)))    # ep_config.nonlocal_ls
    return cfg


def particles_view(mpts: Point) -> EpParticlesView:
    s = mpts.s
    v = EpParticlesView()
    m = {"x": mpts.x, "v": s.v, "l0": mpts.l0, "l": mpts.l, "F": s.F, "b": s.b, "gradv": s.gradv, "eps": s.eps,
         "sig": s.sig, "tau": s.tau, "epII": s.epII, "rho0": s.rho0, "rho": s.rho, "Om0": mpts.Om0, "Om": mpts.Om,
         "n0": mpts.n0, "n": mpts.n, "J": mpts.J, "dJ": mpts.dJ, "c0": s.c0, "cr": s.cr, "phi": s.phi,
         "dlam": s.dlam, "vmax": mpts.vmax, "Bij": mpts.Bij}
    for k, a in m.items():
        setattr(v, k, _ptr(a))
    return v


def step_params(T: float, *, G: float = 9.81, tg: float = 0.0, ramp: bool = False, plast: bool = True,
                dt: float | None = None, record_dt: bool = False) -> EpStepParams:
    return EpStepParams(T=float(T), G=float(G), tg=float(tg), ramp=1 if ramp else 0,
                        with_plast=1 if plast else 0, dt_mode=0 if dt is None else 1,
                        record_dt=1 if record_dt else 0, dt_fixed=0.0 if dt is None else float(dt))


class Device:
    """One device context: a Mesh + Point pair resident in HBM."""

    def __init__(self, mesh: Mesh, mpts: Point, cmpr, instr):
        self.L = load()
        self.mesh, self.mpts, self.cmpr, self.instr = mesh, mpts, cmpr, instr
        self.cfg = make_config(mesh, mpts, cmpr, instr)
        h = C.c_void_p()
        _check(self.L.ep_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self._bcs = np.asfortranarray(mesh.bcs.status.astype(np.uint8))
        gv = EpGridView(_ptr(mesh.x), self._bcs.ctypes.data_as(C.POINTER(C.c_uint8)))
        _check(self.L.ep_upload_grid(self.h, C.byref(gv)))
        self.upload()

    def upload(self, t: float | None = None):
        _check(self.L.ep_upload_particles(self.h, C.byref(particles_view(self.mpts))))
        if t is not None:
            _check(self.L.ep_set_time(self.h, float(t)))

    def set_resort(self, every: int):
        """Re-sort the particles into brick-major cell order every `every`
        steps; every < 0 (the library's default, -5): adaptively, checking
        the growth of the P2G node boxes and their overflow count every
        -every steps; 0: never."""
        _check(self.L.ep_set_resort(self.h, int(every)))

    def set_layout(self, block: int, brick):
        """Threads per block and cells per sort brick (ep_set_layout)."""
        b = (C.c_int32 * 3)(*[int(v) for v in brick] + [1] * (3 - len(brick)))
        _check(self.L.ep_set_layout(self.h, int(block), b))

    def set_time(self, t: float):
        _check(self.L.ep_set_time(self.h, float(t)))

    def download(self):
        _check(self.L.ep_download_particles(self.h, C.byref(particles_view(self.mpts))))
        return self.mpts

    def download_grid(self):
        me = self.mesh
        f = EpGridFields(_ptr(me.m if me.m.flags["F_CONTIGUOUS"] else None), _ptr(me.p), _ptr(me.oobf),
                         _ptr(me.a), _ptr(me.v))
        _check(self.L.ep_download_grid(self.h, C.byref(f)))
        return me

    def step(self, nsteps: int, T: float, *, G: float = 9.81, tg: float = 0.0, ramp: bool = False,
             plast: bool = True, dt: float | None = None, record_dt: bool = False):
        p = step_params(T, G=G, tg=tg, ramp=ramp, plast=plast, dt=dt, record_dt=record_dt)
        r = EpStepResult()
        _check(self.L.ep_step(self.h, int(nsteps), C.byref(p), C.byref(r)))
        return r

    def dts(self, cap: int) -> np.ndarray:
        out = np.zeros(cap)
        n = C.c_int64()
        _check(self.L.ep_get_dts(self.h, out.ctypes.data_as(_dp), cap, C.byref(n)))
        return out[: n.value]

    def profile(self, on: bool = True):
        _check(self.L.ep_profile(self.h, 1 if on else 0))

    def profile_kernel(self, name: str):
        """Time only the launches of kernel `name` (a key of kernel_times)."""
        _check(self.L.ep_profile_kernel(self.h, name.encode()))

    def kernel_times(self) -> dict:
        nmax, nl = 32, 32
        names = C.create_string_buffer(nmax * nl)
        ms = (C.c_double * nmax)()
        la = (C.c_int64 * nmax)()
        cnt = C.c_int32()
        _check(self.L.ep_kernel_times(self.h, nmax, names, nl, ms, la, C.byref(cnt)))
        out = {}
        for k in range(cnt.value):
            nm = names.raw[k * nl:(k + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (ms[k], la[k])
        return out

    def attach_surrogate(self, net: "Surrogate", tol: float = 0.0):
        """Network for a precision EP_FP32_SURROGATE context (kept alive here)."""
        self._net = net
        _check(self.L.ep_attach_surrogate(self.h, net.h, float(tol)))

    def stats(self) -> EpStats:
        st = EpStats()
        _check(self.L.ep_get_stats(self.h, C.byref(st)))
        return st

    def close(self):
        if getattr(self, "h", None):
            self.L.ep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class Surrogate:
    """The learned Drucker-Prager return (ep_surrogate_*): an MLP
    8-128-128-128-4 trained and evaluated on the device with bf16 MFMA."""

    def __init__(self, seed: int = 0):
        self.L = load()
        h = C.c_void_p()
        _check(self.L.ep_surrogate_create(int(seed), C.byref(h)))
        self.h = h

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(NPARAM, np.float32)
        _check(self.L.ep_surrogate_get_params(self.h, out.ctypes.data_as(_fp)))
        return out

    @params.setter
    def params(self, w):
        w = _f32(w, (NPARAM,))
        _check(self.L.ep_surrogate_set_params(self.h, w.ctypes.data_as(_fp)))

    def sample(self, n: int, seed: int):
        _check(self.L.ep_surrogate_sample(self.h, int(n), int(seed)))

    def samples(self, first: int, n: int):
        x = np.zeros((n, 8), np.float32)
        y = np.zeros((n, 4), np.float32)
        raw = np.zeros((n, 6), np.float64)
        _check(self.L.ep_surrogate_get_samples(self.h, int(first), int(n), x.ctypes.data_as(_fp),
                                                y.ctypes.data_as(_fp), raw.ctypes.data_as(_dp)))
        return x, y, raw

    def harvest(self, dev: "Device", append: bool = True) -> int:
        """Add the plastic trial states of dev's last step, with their FP64
        returns, to the training set (ep_surrogate_harvest); returns the
        number of samples added."""
        n = C.c_int64(0)
        _check(self.L.ep_surrogate_harvest(self.h, dev.h, 1 if append else 0, C.byref(n)))
        return int(n.value)

    def train(self, steps: int, batch: int, lr0: float, lr1: float) -> np.ndarray:
        loss = np.zeros(max(int(steps), 1), np.float32)
        _check(self.L.ep_surrogate_train(self.h, int(steps), int(batch), float(lr0), float(lr1),
                                         loss.ctypes.data_as(_fp)))
        return loss[: int(steps)]

    def grad(self, x, y):
        n = len(x)
        x = _f32(x, (n, 8))
        y = _f32(y, (n, 4))
        g = np.zeros(NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        _check(self.L.ep_surrogate_grad(self.h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp), n,
                                        g.ctypes.data_as(_fp), loss.ctypes.data_as(_fp)))
        return g, float(loss[0])

    def predict(self, x) -> np.ndarray:
        n = len(x)
        x = _f32(x, (n, 8))
        y = np.zeros((n, 4), np.float32)
        _check(self.L.ep_surrogate_predict(self.h, x.ctypes.data_as(_fp), n, y.ctypes.data_as(_fp)))
        return y

    def close(self):
        if getattr(self, "h", None):
            self.L.ep_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Many uploaded Devices advanced together, one captured graph replay per
    step (ep_ensemble_*).  The Devices stay usable for download."""

    def __init__(self, devices):
        self.L = load()
        self.devices = list(devices)
        hs = (C.c_void_p * len(self.devices))(*[d.h.value for d in self.devices])
        h = C.c_void_p()
        _check(self.L.ep_ensemble_create(hs, len(self.devices), C.byref(h)))
        self.h = h

    def step(self, nsteps: int, T: float, **kw):
        p = step_params(T, **kw)
        res = (EpStepResult * len(self.devices))()
        _check(self.L.ep_ensemble_step(self.h, int(nsteps), C.byref(p), res))
        return [(r.t, r.last_dt, r.steps) for r in res]

    def tune(self, steps: int, T: float, **kw) -> dict:
        p = step_params(T, **kw)
        r = EpTuneResult()
        _check(self.L.ep_ensemble_tune(self.h, int(steps), C.byref(p), C.byref(r)))
        tried = [{"block": r.tried_block[k], "brick": list(r.tried_brick[k]), "ms_per_step": r.tried_ms[k]}
                 for k in range(min(r.ntried, 16))]
        return {"block": r.block, "brick": list(r.brick), "ms_per_step": r.ms_per_step, "tried": tried}

    def stats(self) -> EpStats:
        st = EpStats()
        _check(self.L.ep_ensemble_get_stats(self.h, C.byref(st)))
        return st

    def download(self):
        for d in self.devices:
            d.download()

    def close(self):
        if getattr(self, "h", None):
            self.L.ep_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
