"""Time loop and user entry points on the device path.

``elastodynamic_`` / ``elastoplastic_`` / ``elastoplasm`` restate
src/home/program/workflow/worflow.jl:23-141; ``ic_slump`` / ``slump`` /
``slump_`` restate src/home/script/example/slump.jl:22-109 and
``ic_collapse`` / ``collapse`` restate collapse.jl:25-115.  Each inner
``while T > t`` loop runs on the device in chunks of steps; the host only
reads the clock back once per chunk.  Plotting and file output
(savlot, get_plot_field) are not part of this path.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import hip
from .config import kwargser
from .ic import (NT, get_collapse, get_column, get_geom, get_slump, setup_cmpr, setup_mesh, setup_mpts,
                 setup_time)
from .jrange import frange

CHUNK = 256


def _run_window(dev: hip.Device, time, T: float, *, ramp: bool, plast: bool, G: float = 9.81) -> int:
    steps = 0
    while T > time.t[0]:
        r = dev.step(CHUNK, T, G=G, tg=time.tg, ramp=ramp, plast=plast)
        time.t[0] = r.t
        steps += r.steps
        if r.steps < CHUNK:
            break
    return steps


def elastodynamic_(mpts, mesh, cmpr, time, instr, dev: hip.Device | None = None) -> int:
    """worflow.jl:23-47: gravity ramp, elastic update only."""
    own = dev is None
    dev = dev or hip.Device(mesh, mpts, cmpr, instr)
    dev.set_time(time.t[0])
    it = 0
    for T in np.sort(frange(time.t[0], instr["plot"]["freq"], time.te)):
        it += _run_window(dev, time, float(T), ramp=True, plast=False)
    if own:
        dev.download()
        dev.close()
    return it


def elastoplastic_(mpts, mesh, cmpr, time, instr, dev: hip.Device | None = None) -> int:
    """worflow.jl:68-93: full gravity, elastoplastic update."""
    own = dev is None
    dev = dev or hip.Device(mesh, mpts, cmpr, instr)
    dev.set_time(time.t[0])
    it = 0
    for T in np.sort(frange(time.t[0], instr["plot"]["freq"], time.t[1])):
        it += _run_window(dev, time, float(T), ramp=False, plast=True)
    if own:
        dev.download()
        dev.close()
    return it


def elastoplasm(ic, cfg, mode: str = "elastodynamic"):
    """worflow.jl:113-141."""
    mesh, mpts, cmpr, time = ic["mesh"], ic["mpts"], ic["cmpr"], ic["time"]
    instr = cfg["instr"]
    if mode not in ("elastodynamic", "elastoplastic", "all-in-one"):
        raise ValueError(f"Invalid workflow: {mode}. Choose 'elastodynamic', 'elastoplastic' or 'all-in-one'.")
    dev = hip.Device(mesh, mpts, cmpr, instr)
    try:
        if mode in ("elastodynamic", "all-in-one"):
            elastodynamic_(mpts, mesh, cmpr, time, instr, dev)
        if mode in ("elastoplastic", "all-in-one"):
            elastoplastic_(mpts, mesh, cmpr, time, instr, dev)
        dev.download()
    finally:
        dev.close()
    return NT(ic=ic, cfg=cfg)


def ic_slump(L, nel, fid: str = "slump", cmpr=None, geometry: str = "wedge", column_width: float = 0.25,
             device: bool = False, **kwargs):
    """slump.jl:22-49.  ``cmpr`` (material overrides for setup_cmpr) and
    ``geometry`` ('wedge' = the reference's get_slump, 'column' = a
    quarter-width column) are additions for the cohesive column workloads;
    ``device`` builds the particle lattice and its geometry filter on the GPU
    (the same arrays)."""
    instr = kwargser("instr", kwargs, dim=len(L))
    geom = get_geom(nel, L, instr)
    mesh = setup_mesh(instr, geom)
    cm = dict(cmpr or {})
    base = {k: cm.pop(k) for k in ("E", "nu", "rho0") if k in cm}
    cmp = setup_cmpr(mesh, **base, **cm)
    if geometry == "wedge":
        g = get_slump(mesh, cmp, instr, device=device)
    elif geometry == "column":
        g = get_column(mesh, cmp, instr, width=column_width, device=device)
    else:
        raise ValueError(f"unknown geometry {geometry}")
    mpts = setup_mpts(mesh, cmp, g)
    time = setup_time(instr, te=10.0, tg=10.0, tep=5.0)
    misc = NT(file=f"{mesh.dim}d_{instr['fwrk']['trsfr']}")
    return NT(mesh=mesh, mpts=mpts, cmpr=cmp, time=time), NT(instr=instr, paths=NT(plot=fid), misc=misc)


ENSEMBLE_L = [64.1584, 16.0396]


def ic_ensemble(n: int = 256, seed: int = 43, *, scales=(1, 2, 3, 4, 5, 6, 7, 8), power: float = 2.0,
                trsfr: str = "tpic", musl: bool = True, precision: str = "fp32", geometry: str = "column",
                **kwargs):
    """An ensemble of n independent 2D slumps (2x2 particles per cell) of
    uneven size: instance i has nel = (40 s, 10 s) with s drawn from `scales`
    with probability ∝ s^-power, cohesion c log-uniform in [5, 50] kPa and
    friction angle φ uniform in [20°, 40°] (no softening), ρ 2650, E 1e6,
    ν 0.3.  Defaults: the quarter-width full-height column (about 0.54M
    particles over 256 instances at power 2) and TPIC with the mUSL
    re-projection (trsfr="tpic", musl=true, the TPIC form the reference
    constructs, mapsto.jl:44-53).  Without mUSL (TPIC-USL) the finest
    instances dilate without bound at free-surface particles, in the FP64
    CPU oracle as on the device.  geometry="wedge" gives ic_slump's
    get_slump geometry.  Returns the list of (ic, cfg) pairs and the drawn
    (s, c, φ) per instance."""
    rng = np.random.default_rng(seed)
    sc = np.asarray(scales, dtype=np.int64)
    pr = sc.astype(np.float64) ** (-power)
    s = rng.choice(sc, size=n, p=pr / pr.sum())
    c = 10.0 ** rng.uniform(math.log10(5.0e3), math.log10(50.0e3), size=n)
    phi = np.radians(rng.uniform(20.0, 40.0, size=n))
    out = []
    for i in range(n):
        mat = dict(rho0=2650.0, E=1.0e6, nu=0.3, c0=float(c[i]), cr=float(c[i]), phi0=float(phi[i]),
                   phir=float(phi[i]), Hp=0.0)
        kw = dict(basis={"which": "bsmpm", "how": None, "ghost": False},
                  fwrk=dict(deform="finite", trsfr=trsfr, musl=musl, C_pf=1.0, locking=True, damping=0.1),
                  nonloc=dict(status=False, ls=0.5), plast=dict(status=True, constitutive="DP"))
        kw.update(kwargs)
        ic, cfg = ic_slump(ENSEMBLE_L, [40 * int(s[i]), 10 * int(s[i])], cmpr=mat, geometry=geometry, **kw)
        cfg.instr["precision"] = precision
        cfg.instr["ops"]["precision"] = {"fp64": 0, "fp32": 1}[precision]
        out.append((ic, cfg))
    return out, dict(s=s, c=c, phi=phi)


def slump(ic, cfg, workflow: str = "elastodynamic"):
    """slump.jl:73-79: runs on deep copies."""
    out = elastoplasm(copy.deepcopy(ic), copy.deepcopy(cfg), mode=workflow)
    out["success"] = True
    return out


def slump_(ic, cfg, workflow: str = "elastodynamic"):
    """slump!: slump.jl:103-109, mutates ic."""
    out = elastoplasm(ic, cfg, mode=workflow)
    out["success"] = True
    return out


def ic_collapse(nel, nu, E, rho0, l0, fid: str = "collapse", **kwargs):
    """collapse.jl:25-57."""
    dim = len(nel)
    if dim != 2:
        raise ValueError("ic_collapse is 2D in the reference")
    L = [10.0, 1.25 * l0]
    instr = kwargser("instr", kwargs, dim=dim)
    h = np.array([L[-1] / nel[-1], L[-1] / nel[-1]])
    nel2 = np.array([5, nel[-1]], dtype=np.int64)
    L2 = np.array([5 * h[-1], L[-1]])
    geom = NT(ndim=2, nn=16, h=h, nel=nel2, L=L2)
    mesh = setup_mesh(instr, geom)
    cmp = setup_cmpr(mesh, E=E, nu=nu, rho0=rho0)
    mpts = setup_mpts(mesh, cmp, get_collapse(mesh, cmp, 2, l0=l0))
    tg = math.ceil((1.0 / cmp.c) * (2.0 * l0) * 40.0)
    time = setup_time(instr, te=tg, tg=tg)
    misc = NT(file=f"{mesh.dim}d_{instr['fwrk']['trsfr']}")
    return NT(mesh=mesh, mpts=mpts, cmpr=cmp, time=time), NT(instr=instr, paths=NT(plot=fid), misc=misc)


def collapse(ic, cfg):
    """collapse.jl:80-86."""
    out = elastoplasm(copy.deepcopy(ic), copy.deepcopy(cfg))
    out["success"] = True
    return out
