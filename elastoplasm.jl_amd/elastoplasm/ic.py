"""Problem setup: mesh, materials, particles and time (the reference's
src/home/program/utils/setup/).  One-time host work, restated so that the
device path and the CPU oracle see identical inputs.
"""
from __future__ import annotations

import math

import numpy as np

from .jrange import frange, linrange
from .types import Boundary, Mesh, Point, Solid, fzeros


class NT(dict):
    """NamedTuple stand-in: ``ic.mesh`` and ``ic['mesh']`` both work."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v


# --------------------------------------------------------------------------
# mesh (get_geom.jl:21-30, get_coords.jl:22-55, get_bc.jl:14-57, e2n.jl:20-78,
#       setup_mesh.jl:26-99)
# --------------------------------------------------------------------------
def get_geom(nel, L, instr) -> NT:
    L = np.asarray(L, dtype=np.float64)
    nel = np.asarray(nel, dtype=np.int64)
    ndim = len(L)
    if instr["basis"]["which"] == "bsmpm":
        h = np.minimum(L / nel, L / 4.0)
    else:
        h = L / nel
    return NT(ndim=ndim, nn=4 ** ndim, h=h.astype(np.float64), nel=nel, L=L)


def get_coords(ndim, L, h, ghosts):
    if ndim == 2:
        xs = frange(0.0 - ghosts[0], h[0], L[0] + ghosts[0])
        zs = frange(0.0 - ghosts[1], h[1], L[1] + ghosts[1])
        nno = np.array([len(xs), len(zs), len(xs) * len(zs)], dtype=np.int64)
        nel = np.array([nno[0] - 1, nno[1] - 1, (nno[0] - 1) * (nno[1] - 1)], dtype=np.int64)
        x = fzeros((2, nno[2]))
        # node no = iz + nno_z*ix (vec of an nno_z x nno_x matrix)
        x[0, :] = np.repeat(xs, len(zs))
        x[1, :] = np.tile(zs, len(xs))
        axes = (xs, zs)
    else:
        xs = frange(0.0 - ghosts[0], h[0], L[0] + ghosts[0])
        ys = frange(0.0 - ghosts[1], h[1], L[1] + ghosts[1])
        zs = frange(0.0 - ghosts[2], h[2], L[2] + 2.0 * h[2] + ghosts[2])
        nno = np.array([len(xs), len(ys), len(zs), len(xs) * len(ys) * len(zs)], dtype=np.int64)
        nel = np.array([nno[0] - 1, nno[1] - 1, nno[2] - 1, (nno[0] - 1) * (nno[1] - 1) * (nno[2] - 1)],
                       dtype=np.int64)
        x = fzeros((3, nno[3]))
        # node no = iz + nno_z*(ix + nno_x*iy)
        iz, ix, iy = np.meshgrid(np.arange(len(zs)), np.arange(len(xs)), np.arange(len(ys)), indexing="ij")
        iz, ix, iy = iz.ravel(order="F"), ix.ravel(order="F"), iy.ravel(order="F")
        x[0, :] = xs[ix]
        x[1, :] = ys[iy]
        x[2, :] = zs[iz]
        axes = (xs, ys, zs)
    return x, nel, nno, axes


def get_bc(xn, instr, ghosts):
    ndim, nno = xn.shape
    xB = np.asfortranarray(np.stack([xn.min(axis=1) + ghosts, xn.max(axis=1) - ghosts], axis=1))
    bc = np.zeros((ndim, nno), dtype=np.bool_, order="F")
    dirichlet = instr["bcs"]["dirichlet"]
    for d in range(ndim):
        for k in range(2):
            ids = xn[d, :] == xB[d, k]
            kind = dirichlet[d][k]
            if kind in ("roller", ":roller"):
                bc[d, ids] = True
            elif kind in ("fixed", ":fixed"):
                bc[:, ids] = True
    return bc, xB


def e2n(ndim, nno, nel, nn):
    """Element-to-node table, 1-based with 0 for absent (e2n.jl:20-78)."""
    out = np.zeros((nn, int(nel[-1])), dtype=np.int64, order="F")
    if ndim == 2:
        nx, nz = int(nno[0]), int(nno[1])
        ex, ez = np.meshgrid(np.arange(nel[0]), np.arange(nel[1]), indexing="ij")
        ex, ez = ex.ravel(), ez.ravel()                    # element id = ez + nel_z*ex
        el = ez + nel[1] * ex
        k = 0
        for i in range(-1, 3):
            for j in range(-1, 3):
                ix, iz = ex + i, ez + j
                ok = (ix >= 0) & (ix < nx) & (iz >= 0) & (iz < nz)
                out[k, el] = np.where(ok, 1 + iz + nz * ix, 0)
                k += 1
    else:
        nx, ny, nz = int(nno[0]), int(nno[1]), int(nno[2])
        ey, ex, ez = np.meshgrid(np.arange(nel[1]), np.arange(nel[0]), np.arange(nel[2]), indexing="ij")
        ey, ex, ez = ey.ravel(), ex.ravel(), ez.ravel()
        el = ez + nel[2] * ex + nel[2] * nel[0] * ey
        k = 0
        for kk in range(-1, 3):
            for i in range(-1, 3):
                for j in range(-1, 3):
                    ix, iy, iz = ex + i, ey + kk, ez + j
                    ok = (ix >= 0) & (ix < nx) & (iy >= 0) & (iy < ny) & (iz >= 0) & (iz < nz)
                    out[k, el] = np.where(ok, 1 + iz + nz * (ix + nx * iy), 0)
                    k += 1
    return out


def setup_mesh(instr, geom) -> Mesh:
    buffer = 2.0 if instr["basis"]["ghost"] else 0.0
    ndim = int(geom.ndim)
    L, nn, h = geom.L, int(geom.nn), geom.h
    ghosts = buffer * h
    xn, nel, nno, axes = get_coords(ndim, L, h, ghosts)
    status, xB = get_bc(xn, instr, ghosts)
    nt = int(nno[-1])
    mesh = Mesh(
        dim=ndim, nel=nel, nno=nno, nn=nn, L=L.copy(), h=h.copy(),
        x0=xn.min(axis=1).copy(), x=xn,
        m=np.zeros(nt), oobf=fzeros((ndim, nt)), a=fzeros((ndim, nt)), p=fzeros((ndim, nt)),
        v=fzeros((ndim, nt)), dJ=fzeros((ndim, nt)),
        e2n=None, xB=xB, bcs=Boundary(status),
    )
    mesh.axes = axes
    return mesh


def mesh_e2n(mesh: Mesh) -> np.ndarray:
    if mesh.e2n is None:
        mesh.e2n = e2n(mesh.dim, mesh.nno, mesh.nel, mesh.nn)
    return mesh.e2n


# --------------------------------------------------------------------------
# materials (setup_cmpr.jl:21-48, 120-143)
# --------------------------------------------------------------------------
def get_elastic_stiffness(E, nu, ndim):
    Kc, Gc = E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
    i, I = 1.0 - nu, (1.0 - 2.0 * nu) / 2.0
    if ndim == 2:
        D = np.array([[i, nu, 0.0], [nu, i, 0.0], [0.0, 0.0, I]])
    else:
        D = np.array([[i, nu, nu, 0, 0, 0], [nu, i, nu, 0, 0, 0], [nu, nu, i, 0, 0, 0],
                      [0, 0, 0, I, 0, 0], [0, 0, 0, 0, I, 0], [0, 0, 0, 0, 0, I]], dtype=np.float64)
    Del = (E / ((1.0 + nu) * (1.0 - 2.0 * nu))) * D
    return Kc, Gc, np.asfortranarray(Del)


def setup_cmpr(mesh, E=1.0e6, nu=0.3, rho0=2700.0, **over) -> NT:
    """``setup_cmpr`` with the reference defaults.  The keyword overrides
    c0, cr, phi0, phir, Hp (hard-coded in the reference) are additions for
    the cohesive column workloads."""
    K, G, Del = get_elastic_stiffness(E, nu, mesh.dim)
    c = math.sqrt((K + 4.0 / 3.0 * G) / rho0)
    c0, cr = 20.0e3, 4.0e3
    phi0, phir = 20.0 * math.pi / 180, 7.5 * math.pi / 180
    Hp = -30.0e3
    out = NT(E=float(E), nu=float(nu), Kc=K, Gc=G, Del=Del, Hp=Hp, c0=c0, cr=cr,
             phi0=phi0, phir=phir, rho0=float(rho0), c=c)
    for k, v in over.items():
        if k not in ("c0", "cr", "phi0", "phir", "Hp"):
            raise KeyError(f"unknown material override {k}")
        out[k] = float(v)
    return out


# --------------------------------------------------------------------------
# particles (mpts_populate.jl:26-56, get_slump.jl:18-95, get_collapse.jl:17-48,
#            setup_mpts.jl:26-163)
# --------------------------------------------------------------------------
def _lattice(mesh, ni, top=None):
    d = mesh.dim
    axes = []
    for i in range(d):
        hi = mesh.xB[i, 1] if (top is None or i < d - 1) else top
        axes.append(frange(mesh.xB[i, 0] + (0.5 * mesh.h[i] / ni), mesh.h[i] / ni, hi))
    if d == 2:
        xs, zs = axes
        xp = np.repeat(xs, len(zs))            # vec of an nz x nx matrix: z fastest
        zp = np.tile(zs, len(xs))
        return np.stack([xp, zp])
    xs, ys, zs = axes
    iz, ix, iy = np.meshgrid(np.arange(len(zs)), np.arange(len(xs)), np.arange(len(ys)), indexing="ij")
    iz, ix, iy = iz.ravel(order="F"), ix.ravel(order="F"), iy.ravel(order="F")
    return np.stack([xs[ix], ys[iy], zs[iz]])


def mpts_populate(mesh, cmpr, instr, ni=2):
    x = _lattice(mesh, ni)
    if instr["grf"]["status"]:
        raise ValueError("the reference's unseeded spectral GRF is not restated; "
                         "use cohesion_field=... with a seeded field instead")
    c0 = np.full(x.shape[1], cmpr.c0)
    return NT(ni=ni, x=x, c0=c0)


def _axes(mesh, ni):
    return [frange(mesh.xB[i, 0] + (0.5 * mesh.h[i] / ni), mesh.h[i] / ni, mesh.xB[i, 1]) for i in range(mesh.dim)]


def _device_points(mesh, ni, geometry, par):
    """The lattice and its filter evaluated by ep_slump_points (host ranges,
    device lattice, filter and compaction)."""
    import ctypes as C
    from . import hip
    L = hip.load()
    dp = C.POINTER(C.c_double)
    L.ep_slump_points.argtypes = [C.c_int32, C.POINTER(dp), C.POINTER(C.c_int64), C.c_int32, dp, C.c_int64, dp,
                                  C.POINTER(C.c_int64)]
    d = mesh.dim
    axes = [np.ascontiguousarray(a, dtype=np.float64) for a in _axes(mesh, ni)]
    ap = (dp * 3)(*[a.ctypes.data_as(dp) for a in axes] + [None] * (3 - d))
    na = (C.c_int64 * 3)(*[len(a) for a in axes] + [1] * (3 - d))
    pr = np.zeros(5)
    pr[:len(par)] = par
    cap = int(np.prod([len(a) for a in axes]))
    x = np.zeros((d, cap), order="F")
    n = C.c_int64()
    hip._check(L.ep_slump_points(d, ap, na, geometry, pr.ctypes.data_as(dp), cap, x.ctypes.data_as(dp), C.byref(n)))
    return np.asfortranarray(x[:, :n.value])


def get_slump(mesh, cmpr, instr, ni=2, lz=12.80, device=False):
    """get_slump.jl:18-95.  device=True evaluates the lattice and the filter
    on the GPU (ep_slump_points); the result is the same array."""
    wl = 0.15 * lz
    if device:
        if instr["grf"]["status"]:
            raise ValueError("the reference's unseeded spectral GRF is not restated; "
                             "use cohesion_field=... with a seeded field instead")
        xs = _axes(mesh, ni)[0]
        a = -1.25
        xmax = float(xs.max())      # the filter line's last point: max x of the kept lattice (its lowest z row)
        xp = _device_points(mesh, ni, 0, [lz - (0.5 * mesh.h[-1] / ni), wl, a, xmax + 0.5 * mesh.L[0], a * xmax])
        nmp = xp.shape[1]
        phi = np.full(nmp, cmpr.phi0)
        phi[xp[-1, :] <= 2 * wl] = cmpr.phir
        return NT(xp=xp, coh0=np.full(nmp, cmpr.c0), cohr=np.full(nmp, cmpr.cr), phi=phi, ni=ni, nmp=nmp)
    out = mpts_populate(mesh, cmpr, instr, ni=ni)
    keep = out.x[-1, :] <= lz - (0.5 * mesh.h[-1] / ni)
    xp = out.x[:, keep]
    c = out.c0[keep]
    xline = linrange(xp[0].min(), xp[0].max(), 200)
    a = -1.25
    X, Z = xline[-1] + 0.5 * mesh.L[0], a * xline[-1]   # only the last line point decides
    dx, dz = xp[0] - X, xp[-1] - Z
    s = dx * a + dz * (-1.0)
    pos = (s > 0) | (xp[-1] < wl)
    xp = np.asfortranarray(xp[:, pos])
    nmp = xp.shape[1]
    coh0 = c[pos].copy()
    cohr = np.full(nmp, cmpr.cr)
    phi = np.full(nmp, cmpr.phi0)
    phi[xp[-1, :] <= 2 * wl] = cmpr.phir
    return NT(xp=xp, coh0=coh0, cohr=cohr, phi=phi, ni=ni, nmp=nmp)


def get_column(mesh, cmpr, instr, ni=2, width=0.25, height=None, device=False):
    """Rectangular column x <= width*L_x, z <= height (default the domain
    height L_z, i.e. below the 3D mesh's 2h headroom): the column geometry
    of the 3D slump workloads, an addition to the reference's wedge.
    device=True evaluates it on the GPU (ep_slump_points)."""
    zmax = mesh.L[-1] if height is None else height
    if device:
        xp = _device_points(mesh, ni, 1, [width * mesh.L[0], zmax])
    else:
        x = _lattice(mesh, ni)
        keep = x[0, :] <= width * mesh.L[0]
        keep &= x[-1, :] <= zmax
        xp = np.asfortranarray(x[:, keep])
    nmp = xp.shape[1]
    return NT(xp=xp, coh0=np.full(nmp, cmpr.c0), cohr=np.full(nmp, cmpr.cr),
              phi=np.full(nmp, cmpr.phi0), ni=ni, nmp=nmp)


def get_collapse(mesh, cmpr, ni, l0=0.0):
    if mesh.dim != 2:
        raise ValueError("get_collapse is 2D in the reference (get_collapse.jl:26-36)")
    xs = frange(mesh.xB[0, 0] + (0.5 * mesh.h[0] / ni), mesh.h[0] / ni, mesh.xB[0, 1])
    zs = frange(mesh.xB[1, 0] + (0.5 * mesh.h[1] / ni), mesh.h[1] / ni, l0)
    xp = np.asfortranarray(np.stack([np.repeat(xs, len(zs)), np.tile(zs, len(xs))]))
    nmp = xp.shape[1]
    return NT(xp=xp, coh0=np.full(nmp, cmpr.c0), cohr=np.full(nmp, cmpr.cr),
              phi=np.full(nmp, cmpr.phi0), ni=ni, nmp=nmp)


def setup_mpts(mesh, cmpr, geom) -> Point:
    d = mesh.dim
    nstr = 3 if d == 2 else 6
    ni, nmp, xp = geom.ni, int(geom.nmp), geom.xp
    n0 = 0.1 * np.ones(nmp)
    l0 = np.asfortranarray(np.ones((d, nmp)) * 0.5 * (mesh.h / ni)[:, None])
    v0 = np.prod(2 * l0, axis=0)
    eye = np.asfortranarray(np.repeat(np.eye(d)[:, :, None], nmp, axis=2))
    s = Solid(
        v=fzeros((d, nmp)),
        rho0=np.full(nmp, cmpr.rho0), rho=np.full(nmp, cmpr.rho0),
        c0=np.asarray(geom.coh0, dtype=np.float64).copy(), cr=np.asarray(geom.cohr, dtype=np.float64).copy(),
        phi=np.asarray(geom.phi, dtype=np.float64).copy(), dlam=np.zeros(nmp),
        epII=fzeros((2, nmp)), sig=fzeros((nstr, nmp)), tau=fzeros((nstr, nmp)),
        gradv=fzeros((d, d, nmp)), dF=fzeros((d, d, nmp)), F=eye.copy(order="F"),
        b=eye.copy(order="F"), eps=fzeros((d, d, nmp)), omg=fzeros((d, d, nmp)),
    )
    return Point(ndim=d, nmp=nmp, vmax=np.zeros(d), x=np.asfortranarray(xp.copy()), l0=l0,
                 l=l0.copy(order="F"), n0=n0, n=n0.copy(), Om0=v0.copy(), Om=v0.copy(),
                 dJ=np.ones(nmp), J=np.ones(nmp), s=s, z0=xp[-1, :].copy(), Bij=fzeros((d, d, nmp)))


def setup_time(instr, te=0.0, tg=0.0, tep=0.0) -> NT:
    return NT(t=np.array([0.0, te + tep]), te=float(te), tg=float(te if tg > te else tg), tep=float(tep))
