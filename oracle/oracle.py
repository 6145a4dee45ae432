"""ctypes wrapper of the CPU oracle (libep_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
It steps the host-side Mesh/Point mirrors in place, in the reference's
column-major layout, exactly as the Julia kernels would
(src/home/program/workflow/worflow.jl:23-93).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libep_oracle.so"
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class OrMesh(C.Structure):
    _fields_ = [("dim", C.c_int32), ("nn", C.c_int32), ("nel", C.c_int64 * 4), ("nno", C.c_int64 * 4),
                ("h", C.c_double * 3), ("x0", C.c_double * 3), ("xB", C.c_double * 6),
                ("x", _dp), ("m", _dp), ("oobf", _dp), ("a", _dp), ("p", _dp), ("v", _dp), ("dJ", _dp),
                ("e2n", _ip), ("bcs", C.POINTER(C.c_uint8))]


class OrPoint(C.Structure):
    _fields_ = [("nmp", C.c_int64), ("dim", C.c_int32), ("pad0", C.c_int32), ("vmax", C.c_double * 3),
                ("phi", _dp), ("dnp", _dp), ("p2e", _ip), ("p2n", _ip),
                ("x", _dp), ("l0", _dp), ("l", _dp),
                ("n0", _dp), ("n", _dp), ("Om0", _dp), ("Om", _dp), ("dJ", _dp), ("J", _dp),
                ("v", _dp), ("rho0", _dp), ("rho", _dp), ("c0", _dp), ("cr", _dp), ("phif", _dp), ("dlam", _dp),
                ("epII", _dp), ("sig", _dp), ("tau", _dp),
                ("gradv", _dp), ("dF", _dp), ("F", _dp), ("b", _dp), ("eps", _dp), ("omg", _dp),
                ("Bij", _dp), ("Dij", _dp)]


class OrInstr(C.Structure):
    _fields_ = [("basis", C.c_int32), ("trsfr", C.c_int32), ("musl", C.c_int32), ("deform", C.c_int32),
                ("plast", C.c_int32), ("locking", C.c_int32), ("nthreads", C.c_int32), ("domain", C.c_int32),
                ("C_pf", C.c_double), ("damping", C.c_double), ("Kc", C.c_double), ("Gc", C.c_double),
                ("Hp", C.c_double), ("c", C.c_double), ("Del", C.c_double * 36),
                ("nonloc", C.c_int32), ("pad1", C.c_int32), ("ls", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.or_get_dt.restype = C.c_double
        L.or_det.restype = C.c_double
        L.or_det.argtypes = [C.c_int, _dp]
        L.or_inv.argtypes = [C.c_int, _dp, _dp]
        L.or_get_dt.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.c_double, C.c_double, C.c_double]
        L.or_step.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.POINTER(OrInstr), _dp, C.c_double, C.c_int]
        L.or_shpfun.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.POINTER(OrInstr)]
        L.or_p2n.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), _dp, C.POINTER(OrInstr)]
        L.or_solve.argtypes = [C.POINTER(OrMesh), C.c_double, C.POINTER(OrInstr)]
        L.or_n2p.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.c_double, C.POINTER(OrInstr)]
        L.or_update.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.c_double, C.POINTER(OrInstr)]
        L.or_elast.argtypes = [C.POINTER(OrPoint), C.POINTER(OrInstr)]
        L.or_plast.argtypes = [C.POINTER(OrPoint), C.POINTER(OrInstr)]
        L.or_nonlocal.argtypes = [C.POINTER(OrPoint), C.POINTER(OrMesh), C.POINTER(OrInstr)]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_bsmpm_1d.argtypes = [C.c_double, C.c_int, _dp, _dp]
        L.or_gimp_1d.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.or_smpm_1d.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.or_sym_eig.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_det.argtypes = [C.c_int, _dp]
        L.or_det.restype = C.c_double
        L.or_lu_solve.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_dp_return.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, _dp, _dp]
        L.or_j2_return.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, _dp, C.c_double, _dp,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["F_CONTIGUOUS"], "oracle arrays must be column-major"
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


class Oracle:
    """Steps (mesh, mpts) in place with the restated reference kernels."""

    def __init__(self, mesh, mpts, cmpr, instr, threads: int = 1):
        from elastoplasm.ic import mesh_e2n  # host-side setup mirror
        self.L = load()
        self.mesh, self.mpts, self.cmpr, self.instr = mesh, mpts, cmpr, instr
        ops = instr["ops"]
        d = mesh.dim
        mpts.ensure_stencil(mesh.nn)
        self._e2n = mesh_e2n(mesh)
        self._bcs = np.asfortranarray(mesh.bcs.status.astype(np.uint8))
        m = OrMesh()
        m.dim, m.nn = d, mesh.nn
        for i in range(d + 1):
            m.nel[i] = int(mesh.nel[i])
            m.nno[i] = int(mesh.nno[i])
        for i in range(d):
            m.h[i], m.x0[i] = mesh.h[i], mesh.x0[i]
        xB = mesh.xB.ravel(order="F")
        for k in range(2 * d):
            m.xB[k] = xB[k]
        m.x, m.m, m.oobf, m.a, m.p, m.v, m.dJ = (_p(mesh.x), _p(mesh.m), _p(mesh.oobf), _p(mesh.a), _p(mesh.p),
                                                  _p(mesh.v), _p(mesh.dJ))
        m.e2n = _p(self._e2n)
        m.bcs = self._bcs.ctypes.data_as(C.POINTER(C.c_uint8))
        self.m = m
        s = mpts.s
        pt = OrPoint()
        pt.nmp, pt.dim = mpts.nmp, d
        for i in range(d):
            pt.vmax[i] = mpts.vmax[i]
        pt.phi, pt.dnp, pt.p2e, pt.p2n = _p(mpts.phi), _p(mpts.dnp), _p(mpts.p2e), _p(mpts.p2n)
        pt.x, pt.l0, pt.l = _p(mpts.x), _p(mpts.l0), _p(mpts.l)
        pt.n0, pt.n, pt.Om0, pt.Om, pt.dJ, pt.J = (_p(mpts.n0), _p(mpts.n), _p(mpts.Om0), _p(mpts.Om), _p(mpts.dJ),
                                                   _p(mpts.J))
        pt.v, pt.rho0, pt.rho, pt.c0, pt.cr, pt.phif, pt.dlam = (_p(s.v), _p(s.rho0), _p(s.rho), _p(s.c0), _p(s.cr),
                                                                 _p(s.phi), _p(s.dlam))
        pt.epII, pt.sig, pt.tau = _p(s.epII), _p(s.sig), _p(s.tau)
        pt.gradv, pt.dF, pt.F, pt.b, pt.eps, pt.omg = (_p(s.gradv), _p(s.dF), _p(s.F), _p(s.b), _p(s.eps),
                                                       _p(s.omg))
        pt.Bij, pt.Dij = _p(mpts.Bij), _p(mpts.Dij)
        self.pt = pt
        ins = OrInstr()
        ins.basis, ins.trsfr, ins.musl, ins.deform = ops["basis"], ops["transfer"], ops["musl"], ops["deform"]
        ins.plast, ins.locking, ins.nthreads = ops["plast"], ops["locking"], threads
        ins.domain = ops.get("domain", 0)
        ins.C_pf, ins.damping = ops["C_pf"], ops["damping"]
        nl = instr.get("nonloc", {})
        ins.nonloc = 1 if nl.get("status", False) else 0
        ins.ls = float(np.max(nl.get("ls", 0.5)))
        ins.Kc, ins.Gc, ins.Hp, ins.c = cmpr.Kc, cmpr.Gc, cmpr.Hp, cmpr.c
        for k, v in enumerate(np.asarray(cmpr.Del).ravel(order="F")):
            ins.Del[k] = v
        self.ins = ins
        self.L.or_set_threads(threads)
        self.t = 0.0
        self.it = 0

    def sync_vmax(self):
        self.mpts.vmax[:] = [self.pt.vmax[i] for i in range(self.mesh.dim)]

    def get_dt(self, T: float) -> float:
        return self.L.or_get_dt(C.byref(self.pt), C.byref(self.m), self.cmpr.c, self.t, T)

    def step(self, g, dt: float, with_plast: bool = True) -> int:
        gg = np.zeros(3)
        gg[: len(g)] = g
        r = self.L.or_step(C.byref(self.pt), C.byref(self.m), C.byref(self.ins), gg.ctypes.data_as(_dp), dt,
                           1 if with_plast else 0)
        if r != 0:
            raise RuntimeError(f"Invalid position (oracle error {r})")
        return r

    def gravity(self, ramp: bool, tg: float, G: float = 9.81):
        d = self.mesh.dim
        if ramp:
            G = (G * self.t / tg) if self.t <= tg else G
        g = np.zeros(d)
        g[-1] = -G
        return g

    def run(self, nsteps: int, T: float, *, ramp: bool = False, tg: float = 0.0, plast: bool = True,
            dt: float | None = None) -> list:
        """Up to nsteps iterations of `while T > t` (worflow.jl:28-39, 74-85)."""
        dts = []
        for _ in range(nsteps):
            if not (T > self.t):
                break
            g = self.gravity(ramp, tg)
            h = self.get_dt(T) if dt is None else dt
            if dt is not None:
                for i in range(self.mesh.dim):
                    self.pt.vmax[i] = 0.0
            self.step(g, h, plast)
            self.t = self.t + h
            self.it += 1
            dts.append(h)
        self.sync_vmax()
        return dts
