"""Per-kernel timing with HIP events (ep_profile, ep_profile_kernel,
ep_kernel_times): every kernel or the launches of one kernel are timed, and
timing changes nothing in the state.  Also ep_create's check that finite
strain gets an isotropic Del."""
from __future__ import annotations

import pytest

from cases import clone, relerr, slump
from elastoplasm import hip

pytestmark = pytest.mark.gpu


def test_profile_every_kernel_then_one_kernel():
    ic, cfg = slump(2)
    st, plain = clone(ic), clone(ic)
    dev = hip.Device(st.mesh, st.mpts, st.cmpr, cfg.instr)
    dev.profile(True)
    assert dev.step(3, 100.0).steps == 3
    kt = dev.kernel_times()
    for k in ("begin", "zero_grid", "p2g", "euler", "g2p", "augv", "deform", "fbar", "end"):
        ms, n = kt[k]
        assert n == 3 and ms > 0.0, (k, kt[k])
    dev.profile_kernel("fbar")
    assert dev.step(4, 100.0).steps == 4
    kt = dev.kernel_times()
    assert list(kt) == ["fbar"] and kt["fbar"][1] == 4 and kt["fbar"][0] > 0.0, kt
    with pytest.raises(hip.EpError):
        dev.profile_kernel("no_such_kernel")
    dev.profile(False)
    assert dev.step(2, 100.0).steps == 2
    assert dev.kernel_times() == {} or all(ms == 0.0 for ms, _ in dev.kernel_times().values())
    dev.download()
    dev.close()

    ref = hip.Device(plain.mesh, plain.mpts, plain.cmpr, cfg.instr)
    assert ref.step(9, 100.0).steps == 9
    ref.download()
    ref.close()
    # events on the stream do not change the arithmetic (runs differ only in
    # the order of the atomic sums)
    assert relerr(st.mpts.x, plain.mpts.x) < 1e-13
    assert relerr(st.mpts.s.tau, plain.mpts.s.tau) < 1e-10


def test_finite_strain_rejects_an_anisotropic_del():
    """The finite-strain update runs in the principal frame of the trial b,
    which needs the reference's isotropic Del; ep_create says so for any
    other (EP_E_CONFIG) and accepts the isotropic one."""
    import numpy as np
    ic, cfg = slump(3)
    cm = clone(ic.cmpr)
    Del = np.array(cm.Del, dtype=np.float64, order="F")
    Del[3, 3] *= 1.5                        # shear modulus off (a - b)/2
    cm.Del = Del
    with pytest.raises(hip.EpError) as e:
        hip.Device(ic.mesh, ic.mpts, cm, cfg.instr)
    assert e.value.code == -2 and "isotropic" in str(e.value)
    dev = hip.Device(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    dev.close()
