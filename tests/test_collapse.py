"""Elastic column compaction against its closed form σ_yy = -ρ0 g (l0 - z0)
(test/testset/test_collapse.jl:1-108).  The reference's own test passes
trsfr="musl", which init_mapsto rejects (mapsto.jl:43-44); here the intended
std + mUSL transfer is used.  The error must fall under mesh refinement."""
from __future__ import annotations

import numpy as np
import pytest

import elastoplasm as ep
from oracle.oracle import Oracle

L0 = 50.0
FWRK = dict(deform="finite", trsfr="std", musl=True, C_pf=1.0, locking=False, damping=0.1)


def collapse_error(mpts, cmpr, z0, current=False):
    # test_collapse.jl:9-14; σ is the transform of the last step's start
    # (p2n.jl:12-21); current=True uses τ/J of the final state instead
    xnum = np.abs(mpts.s.tau[1, :] / mpts.J if current else mpts.s.sig[1, :])
    x = np.abs(cmpr.rho0 * 9.81 * (L0 - z0))
    return float(np.sum(np.abs(xnum - x) * mpts.Om0) / (9.81 * cmpr.rho0 * L0 * np.sum(mpts.Om0)))


def run_oracle(nel):
    ic, cfg = ep.ic_collapse(nel, 0.0, 1.0e4, 80.0, L0, fwrk=FWRK, nonloc=dict(status=False, ls=0.5))
    z0 = ic.mpts.x[-1].copy()
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    t = ic.time
    # elastodynamic! over every plot checkpoint (worflow.jl:23-47)
    for T in ep.jrange.frange(t.t[0], 1.0, t.te):
        while T > o.t:
            o.run(1, float(T), ramp=True, tg=t.tg, plast=False)
    return collapse_error(ic.mpts, ic.cmpr, z0), ic


def test_compaction_error_decreases_under_refinement_oracle():
    errs = [run_oracle([5, n])[0] for n in (5, 10, 20)]
    print("compaction errors", errs)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.05


@pytest.mark.gpu
def test_compaction_through_device_workflow():
    """collapse() runs elastodynamic! on the device through the C ABI."""
    errs = []
    for n in (5, 10, 20):
        ic, cfg = ep.ic_collapse([5, n], 0.0, 1.0e4, 80.0, L0, fwrk=FWRK, nonloc=dict(status=False, ls=0.5))
        z0 = ic.mpts.x[-1].copy()
        out = ep.collapse(ic, cfg)
        assert out.success
        errs.append(collapse_error(out.ic.mpts, out.ic.cmpr, z0, current=True))
    print("device compaction errors", errs)
    assert errs[0] > errs[1] > errs[2]
    _, ico = run_oracle([5, 20])
    eo = collapse_error(ico.mpts, ico.cmpr, z0, current=True)
    assert abs(errs[2] - eo) <= 1e-6 * eo
