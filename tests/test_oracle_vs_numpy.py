"""Cross-check of the C oracle against the independent numpy restatement
(tests/numpy_step.py) over several whole steps on tiny slumps."""
from __future__ import annotations

import copy
import math

import numpy as np
import pytest

import numpy_step
from cases import relerr, slump
from oracle.oracle import Oracle


@pytest.mark.parametrize("dim,nel,locking,musl", [(2, [8, 2], True, True), (2, [8, 2], False, False),
                                                   (3, [8, 2, 2], True, True)])
def test_oracle_matches_numpy_restatement(dim, nel, locking, musl):
    cm = dict(c0=2.0e3, cr=1.0e3, Hp=-1e4, phi0=math.radians(30), phir=math.radians(20))
    ic, cfg = slump(dim, nel=nel, locking=locking, musl=musl, C_pf=0.95, cmpr=cm)
    assert 0 < ic.mpts.nmp < 400
    ref = copy.deepcopy(ic)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    g = np.zeros(dim)
    g[-1] = -9.81
    plastic = 0
    for k in range(8):
        dt = 0.02
        o.step(g, dt, True)
        numpy_step.step(ref.mesh, ref.mpts, ref.cmpr, cfg.instr["ops"], g, dt, True)
        for name, a, b in [("x", ic.mpts.x, ref.mpts.x), ("v", ic.mpts.s.v, ref.mpts.s.v),
                           ("F", ic.mpts.s.F, ref.mpts.s.F), ("b", ic.mpts.s.b, ref.mpts.s.b),
                           ("tau", ic.mpts.s.tau, ref.mpts.s.tau), ("epII", ic.mpts.s.epII[0], ref.mpts.s.epII[0]),
                           ("J", ic.mpts.J, ref.mpts.J)]:
            assert relerr(a, b) < 1e-10, (k, name, relerr(a, b))
        plastic = max(plastic, int((ic.mpts.s.epII[0] > 0).sum()))
    assert plastic > 0
