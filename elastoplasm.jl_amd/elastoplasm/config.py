"""Instruction tree: ``require`` and ``kwargser`` (src/home/api/require.jl:34-88,
src/home/api/kwargser.jl:31-62).

The reference stores kernels in ``instr.cairn``; here ``kwargser`` validates the
same string switches and maps them onto the integer enums of the C ABI
(``include/elastoplasm.h``), raising the reference's errors for unsupported
choices (shpfun.jl:50, mapsto.jl:44, plast.jl:31).
"""
from __future__ import annotations

import copy
import warnings

BASIS = {"bsmpm": 0, "gimpm": 1, "smpm": 2}
TRANSFER = {"std": 0, "tpic": 1, "apic": 2}
# GIMP domain updates (domain.jl:1-52) by framework (update.jl:7-29)
DOMAIN_FINITE = {"undeformed": 0, "detFij": 1, "Fii": 2, "Uii": 3}
DOMAIN_INFINITESIMAL = {"undeformed": 0, "detΔFij": 4, "ΔFii": 5, "ΔUii": 6}
DEFORM = {"finite": 0, "infinitesimal": 1}
PLAST = {"DP": 1, "J2": 2}
PRECISION = {"fp64": 0, "fp32": 1, "fp32_surrogate": 2}


def require(what: str = "instr") -> dict:
    """Default instruction tree (require.jl:34-88)."""
    if what not in ("instr", ":instr"):
        raise ValueError(f"UnsupportedSymbol: {what}")
    return {
        "dtype": 64,
        "basis": {"which": "bsmpm", "how": None, "ghost": False},
        "fwrk": {"deform": "finite", "trsfr": "std", "C_pf": 1.0, "musl": True,
                 "locking": True, "damping": 0.1},
        "bcs": {"dirichlet": [["roller", "roller"], ["roller", "roller"], ["roller", "roller"]]},
        "grf": {"status": False, "covariance": "gaussian",
                "param": {"Ix": [2.5, 2.5, 2.5], "Nh": 5000, "km": 100}},
        "plast": {"status": False, "constitutive": "DP"},
        "nonloc": {"status": True, "ls": 0.5},
        "plot": {"status": True, "freq": 1.0, "what": ["epII"], "dims": (500.0, 250.0)},
        "perf": {"status": False},
        # addition of this implementation: arithmetic of the device path
        "precision": "fp64",
    }


def kwargser(what: str, kwargs: dict | None = None, dim: int = 2) -> dict:
    """Merge top-level user keys into ``require()`` (kwargser.jl:31-62).

    As in the reference, a nested key replaces the whole subtree.  The result
    carries ``instr['ops']``: the integer switches the device path uses in
    place of the reference's ``cairn`` kernel table.
    """
    ref = require(what)
    kwargs = dict(kwargs or {})
    unused = [k for k in kwargs if k not in ref]
    if unused:
        warnings.warn("miscellaneous kwargs for require():\n\t- " + "\n\t- ".join(map(str, unused)))
    instr = copy.deepcopy(ref)
    for k, v in kwargs.items():
        if k in ref:
            instr[k] = copy.deepcopy(v)
    if instr["dtype"] not in (32, 64):
        raise ValueError("dtype must be 32 or 64")
    instr["ops"] = init_ops(dim, instr)
    return instr


def init_ops(dim: int, instr: dict) -> dict:
    """Integer form of the cairn (shpfun.jl:14-60, mapsto.jl:13-54,
    update.jl:1-42, elast.jl:93-105, plast.jl:12-34)."""
    if dim not in (2, 3):
        raise ValueError("the device path covers 2D and 3D")
    basis = instr["basis"]["which"]
    if basis not in BASIS:
        raise ValueError(f"{basis} is not a supported shape function basis")
    fw = instr["fwrk"]
    if fw["trsfr"] not in TRANSFER:
        raise ValueError(f"{fw['trsfr']} is an unsupported transfer scheme")
    if fw["deform"] not in DEFORM:
        raise ValueError(f"{fw['deform']} is an unsupported deformation framework")
    domain = 0
    if basis == "gimpm":
        # update.jl:7-28: the `how` kernels of each framework; any other value
        # leaves domain! undefined (an UndefVarError at setup)
        how = instr["basis"].get("how")
        table = DOMAIN_FINITE if fw["deform"] == "finite" else DOMAIN_INFINITESIMAL
        if how not in table:
            raise ValueError(f"gimpm: basis.how = {how!r} is not a {fw['deform']} domain update")
        domain = table[how]
    cons = instr["plast"]["constitutive"]
    if cons not in PLAST:
        raise ValueError(f"InvalidReturnMapping: {cons}")
    prec = instr.get("precision", "fp64")
    if prec not in PRECISION:
        raise ValueError(f"unknown precision {prec}")
    return {
        "dim": dim,
        "basis": BASIS[basis],
        "transfer": TRANSFER[fw["trsfr"]],
        "musl": 1 if fw["musl"] else 0,
        "deform": DEFORM[fw["deform"]],
        "plast": PLAST[cons],
        "locking": 1 if fw["locking"] else 0,
        "C_pf": float(fw["C_pf"]),
        "damping": float(fw["damping"]),
        "precision": PRECISION[prec],
        "domain": domain,
    }
