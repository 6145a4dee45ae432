"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Bar: plain FP64 matches the reference's CPU kernels to
<= 1e-10 relative per step, and <= 1e-6 relative in displacement and stress
after 100 steps.  Errors are norm-wise over all particles,
||dev - ref|| / ||ref||.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from cases import clone, compare, field, relerr, slump
from elastoplasm import hip
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-10
RUN_TOL = 1e-6

# C1 materials: the 2D slump with c = 20 kPa, φ = 30°
C1 = dict(rho0=2650.0, E=1.0e6, nu=0.3, c0=20.0e3, cr=20.0e3, phi0=math.radians(30.0),
          phir=math.radians(30.0), Hp=0.0)


def _device_run(ic, cfg, nsteps, T, *, t0=0.0, plast=True, ramp=False, tg=0.0, record=False, resort=None):
    dev = hip.Device(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    try:
        if resort is not None:
            dev.set_resort(resort)
        dev.set_time(t0)
        r = dev.step(nsteps, T, plast=plast, ramp=ramp, tg=tg, record_dt=record)
        dts = dev.dts(nsteps) if record else None
        dev.download()
    finally:
        dev.close()
    return r, dts


def _warm_state(ic, cfg, nwarm, T=100.0):
    """Advance with the oracle to a state with motion and plasticity."""
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    o.run(nwarm, T, plast=True)
    return o.t


def _one_step(ic, cfg, nwarm, fields, tol=STEP_TOL, T=100.0):
    t0 = _warm_state(ic, cfg, nwarm, T) if nwarm else 0.0
    ref, dev = clone(ic), clone(ic)
    o = Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr)
    o.t = t0
    o.run(1, T, plast=True)
    r, _ = _device_run(dev, cfg, 1, T, t0=t0)
    assert r.steps == 1
    err = compare(dev.mpts, ref.mpts, fields)
    print("one-step", err)
    for f, e in err.items():
        assert e <= tol, (f, e)
    return err


CASES_2D = [
    dict(deform="finite", trsfr="std", musl=True, locking=True),
    dict(deform="finite", trsfr="std", musl=True, locking=False),
    dict(deform="finite", trsfr="tpic", musl=True, locking=True),
    dict(deform="finite", trsfr="tpic", musl=False, locking=True),
    dict(deform="infinitesimal", trsfr="std", musl=True, locking=True),
    dict(deform="infinitesimal", trsfr="std", musl=True, locking=False),
    dict(deform="infinitesimal", trsfr="tpic", musl=False, locking=False),
    # the APIC rows of the reference's slump test matrix (test_slump.jl:8-15)
    dict(deform="finite", trsfr="apic", musl=True, locking=True),
    dict(deform="finite", trsfr="apic", musl=True, locking=False),
    dict(deform="infinitesimal", trsfr="apic", musl=True, locking=True),
    dict(deform="infinitesimal", trsfr="apic", musl=True, locking=False),
    dict(deform="finite", trsfr="apic", musl=False, locking=True),
]


@pytest.mark.parametrize("case", CASES_2D, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_one_step_2d_bsmpm(case):
    ic, cfg = slump(2, **case)
    fields = ["x", "v", "F", "J", "epII", "dlam"] + (["b", "tau"] if case["deform"] == "finite" else ["sig"])
    if case["trsfr"] == "tpic":
        fields.append("gradv")
    if case["trsfr"] == "apic":
        fields.append("Bij")
    _one_step(ic, cfg, 40, fields)


@pytest.mark.parametrize("basis", ["gimpm", "smpm"])
def test_one_step_2d_other_bases(basis):
    ic, cfg = slump(2, basis=basis)
    _one_step(ic, cfg, 30, ["x", "v", "F", "J", "b", "tau", "epII"])


DOMAINS = [("finite", "detFij"), ("finite", "Fii"), ("finite", "Uii"), ("infinitesimal", "detΔFij"),
           ("infinitesimal", "ΔFii"), ("infinitesimal", "ΔUii")]


@pytest.mark.parametrize("deform,how", DOMAINS)
def test_one_step_2d_gimp_domain(deform, how):
    """GIMP with each domain update of domain.jl (basis.how)."""
    ic, cfg = slump(2, basis="gimpm", deform=deform, how=how)
    stress = ["b", "tau"] if deform == "finite" else ["sig"]
    _one_step(ic, cfg, 30, ["x", "v", "F", "J", "epII", "l"] + stress)


@pytest.mark.parametrize("deform,how", [("finite", "Uii"), ("infinitesimal", "ΔUii")])
def test_gimp_domain_3d_and_60_steps(deform, how):
    """3D GIMP one step, and 60 2D steps from rest, with the stretch-tensor
    domain updates."""
    ic, cfg = slump(3, basis="gimpm", deform=deform, how=how)
    stress = ["b", "tau"] if deform == "finite" else ["sig"]
    _one_step(ic, cfg, 5, ["x", "v", "F", "J", "l"] + stress)
    ic, cfg = slump(2, basis="gimpm", deform=deform, how=how)
    x0 = ic.mpts.x.copy(order="F")
    ref, dev = clone(ic), clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(60, 100.0, plast=True)
    r, _ = _device_run(dev, cfg, 60, 100.0)
    assert r.steps == 60
    du = relerr(dev.mpts.x - x0, ref.mpts.x - x0)
    dl = relerr(dev.mpts.l - ic.mpts.l, ref.mpts.l - ic.mpts.l)
    print(f"GIMP {how} 60 steps: displacement {du:.2e}, change of l {dl:.2e}")
    assert np.abs(ref.mpts.l - ic.mpts.l).max() > 0
    assert du <= RUN_TOL and dl <= RUN_TOL


def test_one_step_2d_j2():
    ic, cfg = slump(2, plast="J2", cmpr=dict(c0=8.0e3, cr=4.0e3, Hp=1.0e4))
    _one_step(ic, cfg, 40, ["x", "v", "F", "J", "b", "tau", "epII"])


# the reference's 3D slump test matrix (test_slump.jl:3-16, 66-73): bsmpm,
# musl = true, C_pf = 0.99, every deformation x transfer x locking row
CASES_3D = [dict(deform=d, trsfr=t, locking=lk) for d in ("finite", "infinitesimal") for t in ("std", "tpic", "apic")
            for lk in (True, False)]


@pytest.mark.parametrize("case", CASES_3D, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_one_step_3d(case):
    ic, cfg = slump(3, musl=True, **case)
    fields = ["x", "v", "F", "J", "epII", "dlam"] + (["b", "tau"] if case["deform"] == "finite" else ["sig"])
    if case["trsfr"] == "tpic":
        fields.append("gradv")
    if case["trsfr"] == "apic":
        fields.append("Bij")
    _one_step(ic, cfg, 10, fields)


@pytest.mark.parametrize("basis", ["gimpm", "smpm"])
def test_one_step_3d_other_bases(basis):
    """3D GIMP and linear (smpm_3d, smpm.jl:46-66) bases."""
    ic, cfg = slump(3, basis=basis)
    _one_step(ic, cfg, 10, ["x", "v", "F", "J", "b", "tau", "epII"])


def test_3d_tpic_usl_20_steps():
    """3D TPIC without the mUSL re-projection (tpic_3d_p2n, p2n.jl:164-191)
    over 20 steps."""
    ic, cfg = slump(3, C_pf=1.0, cmpr=C1, trsfr="tpic", musl=False)
    x0 = ic.mpts.x.copy(order="F")
    ref, dev = clone(ic), clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(20, 100.0, plast=True)
    r, _ = _device_run(dev, cfg, 20, 100.0)
    assert r.steps == 20
    err = compare(dev.mpts, ref.mpts, ["v", "F", "tau", "gradv"])
    du = relerr(dev.mpts.x - x0, ref.mpts.x - x0)
    print("3D TPIC-USL 20 steps", err, du)
    assert du <= RUN_TOL and max(err.values()) <= RUN_TOL


def test_one_step_3d_gimp_infinitesimal_j2():
    ic, cfg = slump(3, basis="gimpm", deform="infinitesimal", plast="J2",
                    cmpr=dict(c0=8.0e3, cr=4.0e3, Hp=1.0e4))
    _one_step(ic, cfg, 10, ["x", "v", "F", "J", "sig", "epII"])


@pytest.mark.parametrize("trsfr", ["std", "apic"])
def test_c1_100_steps(trsfr):
    """C1: 2D slump, DP c=20 kPa, phi=30, rho 2650, bsmpm, mUSL-FLIP (or
    APIC), finite, CFL 0.5, 100 steps from t=0."""
    ic, cfg = slump(2, C_pf=1.0, cmpr=C1, trsfr=trsfr)
    x0 = ic.mpts.x.copy(order="F")
    ref, dev = clone(ic), clone(ic)
    o = Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr)
    dts_ref = o.run(100, 100.0, plast=True)
    r, dts = _device_run(dev, cfg, 100, 100.0, record=True)
    assert r.steps == 100
    assert relerr(dts, dts_ref) <= 1e-9
    assert abs(r.t - o.t) <= 1e-9 * o.t
    du = relerr(dev.mpts.x - x0, ref.mpts.x - x0)
    dtau = relerr(dev.mpts.s.tau, ref.mpts.s.tau)
    print("C1 100 steps: displacement", du, "stress", dtau, "plastic particles",
          int((ref.mpts.s.epII[0] > 0).sum()))
    assert (ref.mpts.s.epII[0] > 0).sum() > 0
    assert du <= RUN_TOL and dtau <= RUN_TOL


@pytest.mark.parametrize("trsfr", ["std", "apic"])
def test_3d_20_steps(trsfr):
    ic, cfg = slump(3, C_pf=1.0, cmpr=C1, trsfr=trsfr)
    x0 = ic.mpts.x.copy(order="F")
    ref, dev = clone(ic), clone(ic)
    o = Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr)
    o.run(20, 100.0, plast=True)
    r, _ = _device_run(dev, cfg, 20, 100.0)
    du = relerr(dev.mpts.x - x0, ref.mpts.x - x0)
    dtau = relerr(dev.mpts.s.tau, ref.mpts.s.tau)
    print("3D 20 steps: displacement", du, "stress", dtau)
    assert du <= RUN_TOL and dtau <= RUN_TOL


def test_window_stops_at_T():
    """Device-side `while T > t` guard: steps beyond the window are skipped
    and the last dt is clipped to T - t (physics.jl:24)."""
    ic, cfg = slump(2)
    ref, dev = clone(ic), clone(ic)
    o = Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr)
    T = 0.5
    dts_ref = o.run(1000, T)
    r, dts = _device_run(dev, cfg, 64, T, record=True)
    assert r.steps == len(dts_ref) < 64
    assert r.t == pytest.approx(T, rel=0, abs=1e-12)
    assert relerr(dts, dts_ref) <= 1e-9
    assert relerr(dev.mpts.x, ref.mpts.x) <= 1e-10


def test_elastodynamic_ramp_parity():
    """elastodynamic! with the gravity ramp of get_spacetime."""
    ic, cfg = slump(2)
    ref, dev = clone(ic), clone(ic)
    o = Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr)
    o.run(50, 10.0, ramp=True, tg=10.0, plast=False)
    r, _ = _device_run(dev, cfg, 50, 10.0, ramp=True, tg=10.0, plast=False)
    assert r.steps == 50
    err = compare(dev.mpts, ref.mpts, ["x", "v", "F", "b", "tau"])
    print("elastodynamic", err)
    assert max(err.values()) <= RUN_TOL


@pytest.mark.parametrize("dim", [2, 3])
def test_resume_through_host_state(dim):
    """slump! resumes from the mutated ic (slump.jl:103-109): here the state
    goes to the host (ep_download_particles, vmax included), into a new
    context (ep_upload_particles, which sorts by the new positions) and on
    from time t (ep_set_time).  30 steps in one go and 15 + 15 across the
    round trip agree to rounding, and both match the oracle."""
    ic, cfg = slump(dim, C_pf=1.0, cmpr=C1)
    one, _ = clone(ic), None
    r1, _ = _device_run(one, cfg, 30, 100.0)
    two = clone(ic)
    ra, _ = _device_run(two, cfg, 15, 100.0)
    rb, _ = _device_run(two, cfg, 15, 100.0, t0=ra.t)
    assert rb.t == pytest.approx(r1.t, rel=1e-12)
    fields = ["x", "v", "F", "J", "b", "tau", "epII"]
    err = compare(two.mpts, one.mpts, fields)
    print("resume vs one run", err)
    assert max(err.values()) <= 1e-11
    ref = clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(30, 100.0)
    assert max(compare(two.mpts, ref.mpts, fields).values()) <= RUN_TOL


@pytest.mark.parametrize("dim,nsteps,ls,resort", [(2, 80, 2.0, None), (2, 40, 5.0, None), (3, 25, 2.0, None),
                                                    (2, 40, 2.0, 3)])
def test_nonlocal_regularisation(dim, nsteps, ls, resort):
    """Non-local plastic strain (nonlocal.jl:18-47, plast.jl:49-60) with the
    reference's softening material (c0 20 kPa -> cr 4 kPa, Hp -30 kPa, weak
    layer φr 7.5°), so that ϵpII[2] enters the cohesion of the return.  The
    last case re-sorts every 3 steps: the pass excludes pairs by the
    particles' original order, which must follow the particles."""
    ic, cfg = slump(dim, C_pf=1.0)
    cfg = clone(cfg)
    cfg.instr["nonloc"] = {"status": True, "ls": ls}
    x0 = ic.mpts.x.copy(order="F")
    ref, dev = clone(ic), clone(ic)
    Oracle(ref.mesh, ref.mpts, ref.cmpr, cfg.instr).run(nsteps, 100.0)
    r, _ = _device_run(dev, cfg, nsteps, 100.0, resort=resort)
    assert r.steps == nsteps
    a, b = dev.mpts, ref.mpts
    npl = int((b.s.dlam != 0).sum())
    e1, e2 = relerr(a.s.epII[0], b.s.epII[0]), relerr(a.s.epII[1], b.s.epII[1])
    du = relerr(a.x - x0, b.x - x0)
    dtau = relerr(a.s.tau, b.s.tau)
    print(f"non-local {dim}D ls={ls} {nsteps} steps: plastic {npl}, epII1 {e1:.2e}, epII2 {e2:.2e}, "
          f"displacement {du:.2e}, stress {dtau:.2e}")
    assert npl > 0 and np.abs(b.s.epII[1]).max() > 0
    assert max(e1, e2, du, dtau) <= RUN_TOL
