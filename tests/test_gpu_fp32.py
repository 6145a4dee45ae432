"""FP32 path (FP32 particle state and arithmetic, FP64 grid and
near-degenerate eigen-decompositions) against the FP64 path and the oracle.
Bar: relative stress error <= 1e-3."""
from __future__ import annotations

import math

import pytest

from cases import clone, relerr, slump
from elastoplasm import hip
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-3
C1 = dict(rho0=2650.0, E=1.0e6, nu=0.3, c0=20.0e3, cr=20.0e3, phi0=math.radians(30.0),
          phir=math.radians(30.0), Hp=0.0)


def run(ic, cfg, nsteps, precision):
    cfg = clone(cfg)
    cfg.instr["precision"] = precision
    cfg.instr["ops"]["precision"] = {"fp64": 0, "fp32": 1, "fp32_surrogate": 2}[precision]
    out = clone(ic)
    dev = hip.Device(out.mesh, out.mpts, out.cmpr, cfg.instr)
    try:
        r = dev.step(nsteps, 100.0)
        dev.download()
    finally:
        dev.close()
    assert r.steps == nsteps
    return out


@pytest.mark.parametrize("dim,n", [(2, 100), (3, 20)])
def test_fp32_stress_within_1e3_of_fp64(dim, n):
    ic, cfg = slump(dim, C_pf=1.0, cmpr=C1)
    ref = clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(n, 100.0)
    d64 = run(ic, cfg, n, "fp64")
    d32 = run(ic, cfg, n, "fp32")
    e32 = relerr(d32.mpts.s.tau, ref.mpts.s.tau)
    e64 = relerr(d64.mpts.s.tau, ref.mpts.s.tau)
    x0 = ic.mpts.x
    eu = relerr(d32.mpts.x - x0, ref.mpts.x - x0)
    print(f"{dim}D {n} steps: stress err fp32 {e32:.2e} fp64 {e64:.2e}; displacement err fp32 {eu:.2e}")
    assert e64 <= 1e-6
    assert e32 <= TOL and eu <= TOL
    assert (ref.mpts.s.epII[0] > 0).sum() > 0
