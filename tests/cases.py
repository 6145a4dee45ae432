"""Shared problem builders and comparisons for the parity tests."""
from __future__ import annotations

import copy

import numpy as np

import elastoplasm as ep

L2 = [64.1584, 64.1584 / 4.0]          # test/testset/test_slump.jl:63
L3 = [64.1584, 64.1584 / 4.0, 64.1584 / 4.0]  # test_slump.jl:69


def basis_kw(which: str, how: str = "undeformed") -> dict:
    if which == "bsmpm":
        return {"which": "bsmpm", "how": None, "ghost": False}
    if which == "gimpm":
        return {"which": "gimpm", "how": how, "ghost": True}
    return {"which": "smpm", "how": None, "ghost": True}


def slump(dim=2, *, deform="finite", trsfr="std", musl=True, locking=True, C_pf=0.99, basis="bsmpm",
          plast="DP", nel=None, cmpr=None, geometry="wedge", damping=0.1, how="undeformed"):
    L = L2 if dim == 2 else L3
    nel = nel or ([40, 10] if dim == 2 else [40, 10, 10])
    kw = dict(basis=basis_kw(basis, how),
              fwrk=dict(deform=deform, trsfr=trsfr, musl=musl, C_pf=C_pf, locking=locking, damping=damping),
              nonloc=dict(status=False, ls=0.5), plast=dict(status=True, constitutive=plast))
    return ep.ic_slump(L, nel, cmpr=cmpr, geometry=geometry, **kw)


def clone(ic):
    return copy.deepcopy(ic)


FIELDS = ["x", "v", "F", "b", "tau", "sig", "epII", "J", "dlam"]


def field(mpts, name):
    s = mpts.s
    return {"x": mpts.x, "v": s.v, "F": s.F, "b": s.b, "tau": s.tau, "sig": s.sig, "epII": s.epII[0],
            "J": mpts.J, "dlam": s.dlam, "gradv": s.gradv, "Bij": mpts.Bij, "l": mpts.l}[name]


def relerr(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0.0:
        return float(d)
    return float(d / nb)


def compare(mp_dev, mp_ref, fields) -> dict:
    return {f: relerr(field(mp_dev, f), field(mp_ref, f)) for f in fields}
