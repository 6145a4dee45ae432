"""Periodic device re-sort (ep_set_resort): re-sorting the particles into
brick-major cell order between steps changes nothing but the summation order
of the scatters.  Every configuration is stepped with and without re-sorts
and against the oracle, with downloads between ep_step calls (the slot list
is read back from the device)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from cases import clone, compare, relerr, slump
from elastoplasm import hip
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

EMU = "cpucheck" in os.environ.get("EP_HIP_LIB", "")

CASES = [
    dict(dim=3),
    dict(dim=2, trsfr="tpic", musl=False),
    dict(dim=2, trsfr="apic"),
    dict(dim=2, basis="gimpm", how="Uii"),
    dict(dim=2, basis="gimpm", deform="infinitesimal", how="ΔFii"),
    dict(dim=3, precision="fp32"),
]


def _run(ic, cfg, chunks, every):
    st = clone(ic)
    dev = hip.Device(st.mesh, st.mpts, st.cmpr, cfg.instr)
    dev.set_resort(every)
    for n in chunks:
        assert dev.step(n, 100.0).steps == n
        dev.download()
    dev.close()
    return st


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_resort_changes_only_rounding(case):
    case = dict(case)
    prec = case.pop("precision", "fp64")
    ic, cfg = slump(C_pf=1.0, **case)
    cfg = clone(cfg)
    cfg.instr["precision"] = prec
    cfg.instr["ops"]["precision"] = {"fp64": 0, "fp32": 1}[prec]
    chunks = (4, 3) if EMU else (13, 17)
    a = _run(ic, cfg, chunks, 0)
    b = _run(ic, cfg, chunks, 3)
    fields = ["x", "v", "F", "J", "epII"] + (["tau", "b"] if case.get("deform", "finite") == "finite" else ["sig"])
    if case.get("trsfr") == "tpic":
        fields.append("gradv")
    if case.get("trsfr") == "apic":
        fields.append("Bij")
    if case.get("basis") == "gimpm":
        fields.append("l")
    tol = 1e-11 if prec == "fp64" else 1e-4
    err = compare(b.mpts, a.mpts, fields)
    print("re-sorted vs not", err)
    for f, e in err.items():
        assert e <= tol, (f, e)
    if prec == "fp64":
        ref = clone(ic)
        Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(sum(chunks), 100.0)
        x0 = ic.mpts.x
        assert relerr(b.mpts.x - x0, ref.mpts.x - x0) <= 1e-9
        assert max(compare(b.mpts, ref.mpts, fields).values()) <= 1e-9


def test_adaptive_resort_and_ensemble_members():
    """every < 0: the overflow count is read every -every steps (here every
    step); the result is the never-re-sorted one up to rounding.  Members of
    an ensemble are not re-sorted."""
    ic, cfg = slump(2, C_pf=1.0)
    a = _run(ic, cfg, (5, 5), 0)
    st = clone(ic)
    dev = hip.Device(st.mesh, st.mpts, st.cmpr, cfg.instr)
    dev.set_resort(-1)
    assert dev.step(10, 100.0).steps == 10
    dev.download()
    assert dev.stats().resorts >= 0
    dev.close()
    for f, e in compare(st.mpts, a.mpts, ["x", "v", "F", "tau", "b"]).items():
        assert e <= 1e-11, (f, e)
    import elastoplasm as ep
    lst, _ = ep.ic_ensemble(2, 7, scales=(1,), power=1.0, precision="fp64")
    devs = []
    for ic_i, cfg_i in lst:
        st = clone(ic_i)
        d = hip.Device(st.mesh, st.mpts, st.cmpr, cfg_i.instr)
        d.set_resort(1)
        devs.append((st, d))
    ens = hip.Ensemble([d for _, d in devs])
    ens.step(3, 100.0)
    ens.download()
    for (ic_i, cfg_i), (st, _) in zip(lst, devs):
        alone = clone(ic_i)
        d = hip.Device(alone.mesh, alone.mpts, alone.cmpr, cfg_i.instr)
        d.step(3, 100.0)
        d.download()
        d.close()
        assert relerr(st.mpts.s.tau, alone.mpts.s.tau) <= 1e-12
    ens.close()
    for _, d in devs:
        d.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_spread_particles_take_the_box_fallbacks(dim):
    """Particles scattered at random over the whole mesh, a few per hundred
    cells: a block's particles then span far more nodes than its LDS box
    holds, so P2G takes its per-wave boxes and, where a wave is spread too,
    global atomics; the gathers read global memory.  Two steps still match
    the oracle."""
    ic, cfg = slump(dim, C_pf=1.0, nel=[80, 20] if dim == 2 else [40, 10, 10])
    rng = np.random.default_rng(31)
    lo, hi = ic.mesh.xB[:, 0], ic.mesh.xB[:, 1]
    h = ic.mesh.h
    n = ic.mpts.nmp
    ic.mpts.x[:] = lo[:, None] + 2 * h[:, None] + rng.random((dim, n)) * (hi - lo - 4 * h)[:, None]
    ref, dev = clone(ic), clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(2, 100.0)
    d = hip.Device(dev.mesh, dev.mpts, dev.cmpr, cfg.instr)
    d.set_resort(0)
    d.set_layout(256, (8, 8) if dim == 2 else (4, 2, 4))    # four waves per block
    assert d.step(2, 100.0).steps == 2
    d.download()
    d.close()
    err = compare(dev.mpts, ref.mpts, ["x", "v", "F", "J", "b"])
    # isolated particles in free fall carry no stress: tau is rounding noise
    # of Del * log(b) with b = I to 1e-16, so it is compared in pascals
    dtau = float(np.abs(dev.mpts.s.tau - ref.mpts.s.tau).max())
    print("spread particles", err, "max |dtau| [Pa]", dtau)
    assert max(err.values()) <= 1e-10 and dtau <= 1e-6
