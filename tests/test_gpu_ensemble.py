"""Ensemble of uneven 2D slumps advanced by one captured graph per step
(ep_ensemble_*): every instance against the same instance stepped alone
through ep_step and against the oracle; the autotuner keeps the state."""
from __future__ import annotations

import os

import numpy as np
import pytest

import elastoplasm as ep
from cases import clone, relerr
from elastoplasm import hip
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

EMU = "cpucheck" in os.environ.get("EP_HIP_LIB", "")


def _build(n, seed, precision, scales=(1, 2), **kw):
    lst, draw = ep.ic_ensemble(n, seed, scales=scales, power=1.0, precision=precision, **kw)
    return lst, draw


def _devices(lst):
    out = []
    for ic, cfg in lst:
        st = clone(ic)
        out.append((st, hip.Device(st.mesh, st.mpts, st.cmpr, cfg.instr)))
    return out


@pytest.mark.parametrize("trsfr,musl", [("tpic", False), ("apic", True)])
def test_ensemble_matches_single_contexts_and_oracle(trsfr, musl):
    lst, _ = _build(3 if EMU else 8, 1, "fp64", trsfr=trsfr, musl=musl)
    nsteps = (2, 1) if EMU else (6, 4)
    devs = _devices(lst)
    ens = hip.Ensemble([d for _, d in devs])
    res = [ens.step(k, 100.0) for k in nsteps]
    ens.download()
    for i, (ic, cfg) in enumerate(lst):
        alone = clone(ic)
        dev = hip.Device(alone.mesh, alone.mpts, alone.cmpr, cfg.instr)
        for k in nsteps:
            dev.step(k, 100.0)
        dev.download()
        dev.close()
        ref = clone(ic)
        Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(sum(nsteps), 100.0)
        st = devs[i][0]
        assert res[-1][i][2] == nsteps[-1]
        e1 = relerr(st.mpts.s.tau, alone.mpts.s.tau)
        e2 = relerr(st.mpts.s.tau, ref.mpts.s.tau)
        x0 = ic.mpts.x
        e3 = relerr(st.mpts.x - x0, ref.mpts.x - x0)
        print(f"instance {i} nmp {ic.mpts.nmp}: vs alone {e1:.1e}, vs oracle stress {e2:.1e} displacement {e3:.1e}")
        assert e1 <= 1e-12 and e2 <= 1e-9 and e3 <= 1e-9
        if trsfr == "apic":
            assert relerr(st.mpts.Bij, ref.mpts.Bij) <= 1e-9
    ens.close()
    for _, d in devs:
        d.close()


def test_ensemble_fp32_tpic_musl_within_1e3():
    lst, draw = _build(2 if EMU else 6, 2, "fp32")
    n = 3 if EMU else 40
    devs = _devices(lst)
    ens = hip.Ensemble([d for _, d in devs])
    ens.step(n, 100.0)
    ens.download()
    for i, (ic, cfg) in enumerate(lst):
        ref = clone(ic)
        Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(n, 100.0)
        e = relerr(devs[i][0].mpts.s.tau, ref.mpts.s.tau)
        print(f"fp32 instance {i} (s={draw['s'][i]}, c={draw['c'][i]:.0f}): stress err {e:.2e}")
        assert e <= 1e-3
    ens.close()
    for _, d in devs:
        d.close()


def test_bench_mix_1000_steps():
    """The 256-slump mix the benchmark runs (seed 43, nel 40x10 .. 320x80,
    FP32, TPIC + mUSL), all 1000 steps: the smallest and the largest
    instance against the FP64 oracle after 40 steps, then every instance
    finite, inside its domain and with bounded dilation at step 1000."""
    if EMU:
        pytest.skip("1000 steps of 0.54M particles: GPU only")
    lst, draw = ep.ic_ensemble(256, 43)
    devs = _devices(lst)
    ens = hip.Ensemble([d for _, d in devs])
    ens.tune(5, 1e9)
    r = ens.step(40, 1e9)
    assert all(x[2] == 40 for x in r)
    ens.download()
    sizes = [ic.mpts.nmp for ic, _ in lst]
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        ic, cfg = lst[i]
        ref = clone(ic)
        ref_cfg = clone(cfg)
        ref_cfg.instr["ops"]["precision"] = 0
        Oracle(ref.mesh, ref.mpts, ref.cmpr, ref_cfg.instr, threads=8).run(40, 1e9)
        e = relerr(devs[i][0].mpts.s.tau, ref.mpts.s.tau)
        print(f"instance {i} (nmp {sizes[i]}, s={draw['s'][i]}): fp32 stress err vs fp64 oracle at step 40 {e:.2e}")
        assert e <= 1e-3
    r = ens.step(960, 1e9)
    assert all(x[2] == 960 for x in r)
    ens.download()
    for i, ((st, _), (ic, _)) in enumerate(zip(devs, lst)):
        m = st.mpts
        assert np.isfinite(m.s.tau).all() and np.isfinite(m.x).all(), i
        assert ((m.x >= ic.mesh.xB[:, :1]) & (m.x <= ic.mesh.xB[:, 1:])).all(), i
        assert 0.1 < m.J.min() and m.J.max() < 30.0, (i, m.J.min(), m.J.max())
    ens.close()
    for _, d in devs:
        d.close()


def test_window_stops_each_instance_at_T():
    lst, _ = _build(3, 3, "fp64")
    devs = _devices(lst)
    ens = hip.Ensemble([d for _, d in devs])
    r = ens.step(50, 0.05)
    steps = [x[2] for x in r]
    for i, (ic, cfg) in enumerate(lst):
        alone = clone(ic)
        dev = hip.Device(alone.mesh, alone.mpts, alone.cmpr, cfg.instr)
        ra = dev.step(50, 0.05)
        dev.close()
        assert steps[i] == ra.steps and abs(r[i][0] - ra.t) <= 1e-15 * ra.t
        assert steps[i] < 50 and r[i][0] == pytest.approx(0.05, rel=1e-12)
    ens.close()
    for _, d in devs:
        d.close()


def test_autotuner_keeps_state():
    lst, _ = _build(2 if EMU else 6, 4, "fp32")
    devs = _devices(lst)
    ens = hip.Ensemble([d for _, d in devs])
    ens.step(2, 100.0)
    before = []
    for st, d in devs:
        d.download()
        before.append(clone(st.mpts))
    tune = ens.tune(1 if EMU else 3, 100.0)
    print("autotuner:", tune)
    assert tune["block"] in (64, 128, 256) and tune["ms_per_step"] > 0
    assert len(tune["tried"]) == (3 if EMU else 9)
    for (st, d), b in zip(devs, before):
        d.download()
        for f in ("x", "v", "F", "b", "tau", "epII", "gradv"):
            a = getattr(st.mpts, f) if hasattr(st.mpts, f) else getattr(st.mpts.s, f)
            c = getattr(b, f) if hasattr(b, f) else getattr(b.s, f)
            np.testing.assert_array_equal(a, c)
    # and the tuned layout steps on from there as the untuned one would
    ens.step(2, 100.0)
    ens.download()
    for i, (ic, cfg) in enumerate(lst):
        alone = clone(ic)
        dev = hip.Device(alone.mesh, alone.mpts, alone.cmpr, cfg.instr)
        dev.step(2, 100.0)
        dev.step(2, 100.0)
        dev.download()
        dev.close()
        assert relerr(devs[i][0].mpts.s.tau, alone.mpts.s.tau) <= 1e-5
    ens.close()
    for _, d in devs:
        d.close()


def test_ensemble_rejects_mixed_configs():
    lst, _ = _build(2, 5, "fp64")
    a = hip.Device(*[getattr(lst[0][0], k) for k in ("mesh", "mpts", "cmpr")], lst[0][1].instr)
    cfg = clone(lst[1][1])
    cfg.instr["ops"]["precision"] = 1
    b = hip.Device(*[getattr(lst[1][0], k) for k in ("mesh", "mpts", "cmpr")], cfg.instr)
    with pytest.raises(hip.EpError, match="EP_E_CONFIG"):
        hip.Ensemble([a, b])
    a.close()
    b.close()
