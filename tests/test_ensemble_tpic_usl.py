"""Why the 256-slump ensemble runs TPIC with the mUSL re-projection.

The first ensemble mix (seed 43, scales ∝ s^-2.5, ic_slump wedge, TPIC
without mUSL, FP32) went non-finite on the device in one instance.  The
cause is the scheme, not the device: in the FP64 oracle the finest instance
of that mix (nel 320x80, c = 9.68 kPa, φ = 23.7°) dilates without bound at
free-surface particles from about step 80 on (max J ≈ 9 at step 80, 1.7e3 at
100, 1e146 at 160), while the same instance with mUSL stays bounded (max J below 25 over 1000
steps).  TPIC
without mUSL is a combination the reference does not construct: init_mapsto
only defines the `augm` table when `musl` is set (mapsto.jl:49-54), so
`(; mapsto[:augm]...)` fails for musl = false.  CPU only (oracle)."""
from __future__ import annotations

import numpy as np

import elastoplasm as ep
from elastoplasm.workflow import ENSEMBLE_L
from oracle.oracle import Oracle

# instance 24 of ic_ensemble(256, 43, power=2.5, geometry="wedge"): the only
# scale-8 member of that draw
C_24 = 9678.83467657739
PHI_24 = 0.4139595366119195


def _max_j(musl: bool, nsteps: int) -> list[float]:
    mat = dict(rho0=2650.0, E=1.0e6, nu=0.3, c0=C_24, cr=C_24, phi0=PHI_24, phir=PHI_24, Hp=0.0)
    ic, cfg = ep.ic_slump(ENSEMBLE_L, [320, 80], cmpr=mat, geometry="wedge",
                          basis={"which": "bsmpm", "how": None, "ghost": False},
                          fwrk=dict(deform="finite", trsfr="tpic", musl=musl, C_pf=1.0, locking=True, damping=0.1),
                          nonloc=dict(status=False, ls=0.5), plast=dict(status=True, constitutive="DP"))
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr, threads=8)
    out = []
    for _ in range(0, nsteps, 20):
        o.run(20, 1e9)
        assert np.isfinite(ic.mpts.J).all()
        out.append(float(ic.mpts.J.max()))
    return out


def test_tpic_without_musl_dilates_without_bound_in_the_oracle():
    usl = _max_j(False, 100)
    musl = _max_j(True, 100)
    print("max J per 20 steps, TPIC-USL:", usl, "TPIC + mUSL:", musl)
    # TPIC-USL: J grows by two orders of magnitude between steps 80 and 100
    assert usl[-1] > 100.0 and usl[-1] > 100.0 * usl[-2]
    # with the re-projection the same free-surface particles dilate too, but
    # boundedly (max J 18-24 over 1000 steps)
    assert max(musl) < 10.0 and musl[-1] < 5.0 * musl[-2]
