#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X, in particle-updates/s.

Workload (default): the large 3D slump -- domain 256 x 64 x 128 m, 1 m cells,
2x2x2 particles per cell in the left 40 % column (6.71M particles), cubic
B-spline, mUSL-FLIP (C_pf = 1), log-strain finite deformation, F-bar on,
Drucker-Prager with cohesion from a seeded log-normal random field (mean
20 kPa, CoV 0.3, correlation length 5 m), phi = 30 deg, rho = 2650, E = 1e6,
nu = 0.3, CFL 0.5.
A step is one pass of shpfun + mapsto + elastoplast over all particles, with
every particle and grid array already resident in HBM.

Prints one JSON line (value = whole-job particle-updates/s, roofline of the
dominant kernel from HIP events on the library's stream, the CPU oracle timed
on the host cores over a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import copy
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "elastoplasm.jl_amd"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import elastoplasm as ep  # noqa: E402
from elastoplasm import hip  # noqa: E402

HBM_PEAK_GBS = 8000.0
SURR_TOL = 3e-5      # scaled residual a learned-return prediction must meet
RESORT = -5          # device re-sort policy: adaptive, checked every 5 steps (ep_set_resort)
MAT = dict(rho0=2650.0, E=1.0e6, nu=0.3, c0=20.0e3, cr=20.0e3, phi0=math.radians(30.0), phir=math.radians(30.0),
           Hp=0.0)

WORKLOADS = {
    # name: (L, nel, basis, column width)
    "slump3d_256x64x128": ([256.0, 64.0, 128.0], [256, 64, 128], "bsmpm", 0.4),
    # the same domain filled completely: 16.8M particles
    "block3d_256x64x128_full": ([256.0, 64.0, 128.0], [256, 64, 128], "bsmpm", 1.0),
    "slump3d_128x32x32_gimp": ([128.0, 32.0, 32.0], [128, 32, 32], "gimpm", 0.4),
    "slump2d_40x10": ([64.1584, 16.0396], [40, 10], "bsmpm", None),
}

# Algorithmic bytes per particle for each particle kernel of the 3D finite /
# mUSL / F-bar / Drucker-Prager step: the particle values the kernel must read
# and write, 8 B each in FP64; in FP32 4 B each except the FP64 positions.
# Grid traffic is excluded.  The whole step's minimum (each persistent value
# read and written once per step) is B_ALG.
KERNEL_BYTES_3D = {
    "fp64": {
        "p2g": 8 * (3 + 3 + 6 + 1 + 1 + 1),                  # x v tau J rho0 Om0
        "g2p": 8 * ((3 + 3 + 1 + 1 + 1) + (3 + 3)),          # reads x v J rho0 Om0, writes x' v
        "deform": 8 * (3 + 1 + 1 + 1),                       # reads x J rho0 Om0 (scatters m*dJ to the grid)
        "fbar": 8 * ((3 + 9 + 6 + 1 + 3) + (9 + 1 + 6 + 6 + 1 + 1)),  # reads x F b epII c0 cr phi,
    },                                                        # writes F J b tau epII dlam
    "fp32": {
        "p2g": 8 * 3 + 4 * (3 + 6 + 1 + 1 + 1),
        "g2p": 8 * (3 + 3) + 4 * ((3 + 1 + 1 + 1) + 3),
        "deform": 8 * 3 + 4 * (1 + 1 + 1),
        "fbar": 8 * 3 + 4 * ((9 + 6 + 1 + 3) + (9 + 1 + 6 + 6 + 1 + 1)),
    },
}
B_ALG = {3: 496, 2: 288}


# heterogeneous cohesion of the large slump: log-normal, mean 20 kPa, CoV 0.3,
# squared-exponential correlation length 5 m, seed 42
GRF = dict(mean=20.0e3, cov=0.3, ell=5.0, seed=42)


def build_workload(name: str, basis_override=None):
    L, nel, basis, width = WORKLOADS[name]
    basis = basis_override or basis
    bkw = {"which": basis, "how": "undeformed" if basis == "gimpm" else None, "ghost": basis != "bsmpm"}
    kw = dict(basis=bkw, fwrk=dict(deform="finite", trsfr="std", musl=True, C_pf=1.0, locking=True, damping=0.1),
              nonloc=dict(status=False, ls=0.5), plast=dict(status=True, constitutive="DP"))
    if width is None:
        return ep.ic_slump(L, nel, cmpr=MAT, **kw)
    # lattice and column filter built on the device (ep_slump_points)
    ic, cfg = ep.ic_slump(L, nel, cmpr=MAT, geometry="column", column_width=width, device=True, **kw)
    if len(L) == 3:
        ep.grf.set_cohesion_field(ic, **GRF)
    return ic, cfg


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name: str, budget_s: float, threads: int):
    """The CPU oracle (a restatement of the reference's KernelAbstractions CPU
    kernels, OpenMP over particles with atomic scatters) on a y-slab of the
    same workload: same x/z extent, resolution and column, 8 m wide."""
    from oracle.oracle import Oracle
    L, nel, basis, width = WORKLOADS[name]
    if len(L) == 3:
        L = [L[0], 8.0, L[2]]
        nel = [nel[0], 8, nel[2]]
    bkw = {"which": basis, "how": "undeformed" if basis == "gimpm" else None, "ghost": basis != "bsmpm"}
    kw = dict(basis=bkw, fwrk=dict(deform="finite", trsfr="std", musl=True, C_pf=1.0, locking=True, damping=0.1),
              nonloc=dict(status=False, ls=0.5), plast=dict(status=True, constitutive="DP"))
    if width is None:
        ic, cfg = ep.ic_slump(L, nel, cmpr=MAT, **kw)
    else:
        ic, cfg = ep.ic_slump(L, nel, cmpr=MAT, geometry="column", column_width=width, **kw)
        if len(L) == 3:
            ep.grf.set_cohesion_field(ic, **GRF)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr, threads=threads)
    o.run(1, 1e9)   # warm-up step (first touch of the stencil arrays)
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.run(1, 1e9)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    nmp = ic.mpts.nmp
    return {"value": nmp * steps / el, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{'x'.join(str(int(v)) for v in nel)} cells ({'x'.join(str(v) for v in L)} m) slab of the "
                      f"same column, {nmp} particles, {steps} steps in {el:.1f} s, OpenMP {threads} threads"}


def ensemble_leg(steps: int) -> dict:
    """256 independent 2D slumps of uneven size (nel 40x10 .. 320x80, seed
    43; quarter-width columns), FP32, cubic B-spline, TPIC with the mUSL
    re-projection: autotuned layout, then `steps` graph replays (one per step
    of every instance).  A failure is reported in the leg and marks the line
    non-finite."""
    t0 = time.perf_counter()
    lst, draw = ep.ic_ensemble(256, 43)
    devs = []
    for ic, cfg in lst:
        devs.append(hip.Device(ic.mesh, ic.mpts, ic.cmpr, cfg.instr))
    e = hip.Ensemble(devs)
    nmp = sum(ic.mpts.nmp for ic, _ in lst)
    t_setup = time.perf_counter() - t0
    scheme = "cubic B-spline, TPIC + mUSL, finite, F-bar, Drucker-Prager, quarter-width columns"
    head = {"instances": 256, "nmp_total": nmp, "nmp_min": int(min(ic.mpts.nmp for ic, _ in lst)),
            "nmp_max": int(max(ic.mpts.nmp for ic, _ in lst)), "steps": steps, "unit": "particle-updates/s",
            "dtype": "f32", "scheme": scheme}
    try:
        t0 = time.perf_counter()
        tune = e.tune(20, 1e9)
        t_tune = time.perf_counter() - t0
        e.step(10, 1e9)
        t0 = time.perf_counter()
        r = e.step(steps, 1e9)
        el = time.perf_counter() - t0
        if not all(x[2] == steps for x in r):
            raise RuntimeError("instances skipped steps")
        e.download()
        finite = all(bool(np.isfinite(ic.mpts.s.tau).all()) for ic, _ in lst)
    except (hip.EpError, RuntimeError) as err:
        print(f"[bench] ensemble FAILED: {err}", file=sys.stderr, flush=True)
        return dict(head, value=None, error=str(err), state_finite=False)
    finally:
        e.close()
        for d in devs:
            d.close()
    value = nmp * steps / el
    print(f"[bench] ensemble: 256 slumps, {nmp} particles, {steps} steps in {el:.3f}s = {value:.4g} pu/s "
          f"({1000 * el / steps:.3f} ms/step); tuned block {tune['block']} brick {tune['brick']} "
          f"(setup {t_setup:.1f}s, tuning {t_tune:.1f}s)", file=sys.stderr, flush=True)
    return dict(head, value=value, ms_per_step=1000 * el / steps, autotune=tune, state_finite=finite)


def measured_traffic(workload: str, precision: str, kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/hbm_traffic_<workload>.md
    (counters already scaled by the copy-kernel calibration); None when
    that workload has not been profiled."""
    f = ROOT / "profiles" / f"hbm_traffic_{workload}.md"
    if not f.exists():
        return None
    txt = f.read_text()
    try:
        blob = txt[txt.index("```json") + 7:]
        blob = blob[:blob.index("```")]
        return json.loads(blob)[precision][kernel]["traffic_bytes"]
    except (ValueError, KeyError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="slump3d_256x64x128", choices=list(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="all", choices=["all", "fp64", "fp32", "fp32_surrogate"],
                    help="fp64 alone gives the headline line without the comparisons (used under profilers)")
    ap.add_argument("--surr-samples", type=int, default=10_000_000, help="training samples of the learned return")
    ap.add_argument("--surr-steps", type=int, default=6000, help="Adam steps (batch 65536)")
    ap.add_argument("--ensemble-steps", type=int, default=1000, help="steps of the 256-slump ensemble (0: skip)")
    ap.add_argument("--long-steps", type=int, default=2000,
                    help="whole-trajectory run of the workload, FP64 and FP32 + learned return (0: skip)")
    ap.add_argument("--full-fill-steps", type=int, default=20,
                    help="steps of the 16.8M-particle full-fill line (0: skip)")
    ap.add_argument("--surr-tol", type=float, default=SURR_TOL, help="residual tolerance of the learned return")
    ap.add_argument("--finetune-steps", type=int, default=2000,
                    help="Adam steps of fine-tuning on harvested in-step trial states (0: the C3 network as trained)")
    ap.add_argument("--breakdown-steps", type=int, default=10,
                    help="steps after the timed ones that time every kernel (per-kernel breakdown)")
    ap.add_argument("--resort", type=int, default=RESORT,
                    help="re-sort policy (ep_set_resort): K > 0 every K steps, -K adaptive every K steps, 0 never")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the step would shard by spatial "
              "slabs with halo exchange).", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    ic, cfg = build_workload(args.workload)
    nmp = ic.mpts.nmp
    dim = ic.mesh.dim
    print(f"[bench] {args.workload}: nmp={nmp} nno={int(ic.mesh.nno[-1])} setup {time.perf_counter()-t_setup:.1f}s",
          file=sys.stderr, flush=True)

    net = None
    surr = None
    if args.precision in ("all", "fp32_surrogate") and dim == 3:
        # learned return: training samples and Adam training on the device
        net = hip.Surrogate(seed=45)
        t0 = time.perf_counter()
        net.sample(args.surr_samples, 44)
        t_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss = net.train(args.surr_steps, 65536, 2e-3, 2e-5)
        t_t = time.perf_counter() - t0
        print(f"[bench] surrogate: {args.surr_samples} samples in {t_s:.2f}s, {args.surr_steps} Adam steps "
              f"(batch 65536) in {t_t:.2f}s, loss {loss[:10].mean():.3e} -> {loss[-50:].mean():.3e}",
              file=sys.stderr, flush=True)
        surr = {"train_samples": args.surr_samples, "train_steps": args.surr_steps, "batch": 65536,
                "sample_s": t_s, "train_s": t_t, "train_samples_per_s": args.surr_steps * 65536 / t_t,
                "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-50:].mean())}

    def device(st, precision):
        instr = copy.deepcopy(cfg.instr)
        instr["ops"]["precision"] = {"fp64": 0, "fp32": 1, "fp32_surrogate": 2}[precision]
        dev = hip.Device(st.mesh, st.mpts, st.cmpr, instr)
        dev.set_resort(args.resort)
        if precision == "fp32_surrogate":
            dev.attach_surrogate(net, args.surr_tol)
        return dev

    def timed_run(precision):
        """warmup + timed steps from the initial state, then `args.breakdown_steps`
        more.  The warmup times every kernel with HIP events to find the
        dominant one; the timed steps time only that kernel (its average launch
        duration is the roofline's), so the events of the other kernels stay
        out of the wall time; the steps after them time every kernel for the
        per-kernel breakdown.  Returns (pu/s, s, dominant kernel, its ms per
        launch, {kernel: ms per launch}, {kernel: ms per step}, state, stats)."""
        st = copy.deepcopy(ic)
        dev = device(st, precision)
        kbytes = KERNEL_BYTES_3D["fp64" if precision == "fp64" else "fp32"]
        if args.warmup:
            dev.profile(True)
            dev.step(args.warmup, 1e9)
            kw = dev.kernel_times()
            dom = max((k for k in kw if k in kbytes), key=lambda k: kw[k][0] / kw[k][1])
            dev.profile_kernel(dom)
        else:
            dom = None
            dev.profile(True)
        t0 = time.perf_counter()
        r = dev.step(args.steps, 1e9)
        el = time.perf_counter() - t0
        kt = dev.kernel_times()
        assert r.steps == args.steps, "steps were skipped"
        if dom is None:
            dom = max((k for k in kt if k in kbytes), key=lambda k: kt[k][0] / kt[k][1])
        dom_ms = kt[dom][0] / kt[dom][1]
        per_launch, per_step = {}, {}
        if args.breakdown_steps:
            dev.profile(True)
            dev.step(args.breakdown_steps, 1e9)
            kb = dev.kernel_times()
            per_launch = {k: ms / n for k, (ms, n) in kb.items()}
            per_step = {k: ms / args.breakdown_steps for k, (ms, n) in kb.items()}
            top = max((k for k in per_launch if k in kbytes), key=lambda k: per_launch[k])
            if top != dom:
                print(f"[bench] {precision}: {top} ({per_launch[top]*1e3:.0f} us) outran {dom} "
                      f"({per_launch[dom]*1e3:.0f} us) after the timed steps", file=sys.stderr, flush=True)
        dev.profile(False)
        dev.download()
        stats = dev.stats()
        dev.close()
        return nmp * args.steps / el, el, dom, dom_ms, per_launch, per_step, st, stats

    def long_run(precision, steps, tail=100):
        tail = max(1, min(tail, steps // 2))
        """The whole trajectory from rest (C5 runs 2000 steps): wall time of
        every step, per-kernel times of the last `tail` steps, re-sorts."""
        st = copy.deepcopy(ic)
        dev = device(st, precision)
        t0 = time.perf_counter()
        try:
            dev.step(steps - tail, 1e9)
            dev.profile(True)
            r = dev.step(tail, 1e9)
            el = time.perf_counter() - t0
            kt = dev.kernel_times()
            dev.download()
            stats = dev.stats()
        except hip.EpError as err:
            print(f"[bench] long run {precision} FAILED: {err}", file=sys.stderr, flush=True)
            return {"precision": precision, "steps": steps, "value": None, "error": str(err), "state_finite": False}
        finally:
            dev.close()
        finite = bool(np.isfinite(st.mpts.s.tau).all() and np.isfinite(st.mpts.x).all())
        inside = bool(((st.mpts.x >= ic.mesh.xB[:, :1]) & (st.mpts.x <= ic.mesh.xB[:, 1:])).all())
        out = {"precision": precision, "steps": steps, "value": nmp * steps / el, "ms_per_step": 1000 * el / steps,
               "t_end": r.t, "resorts": int(stats.resorts), "state_finite": finite, "particles_inside_domain": inside,
               "last_steps": tail, "kernels_us_last_steps": {k: 1000 * ms / n for k, (ms, n) in kt.items()}}
        if precision == "fp32_surrogate":
            out.update(predictions=int(stats.predictions), accepted=int(stats.predictions - stats.fallbacks))
        print(f"[bench] long run {precision}: {steps} steps in {el:.2f}s = {out['value']:.4g} pu/s "
              f"({out['ms_per_step']:.3f} ms/step), {out['resorts']} re-sorts, p2g at the end "
              f"{out['kernels_us_last_steps'].get('p2g', float('nan')):.0f} us", file=sys.stderr, flush=True)
        return out

    value, el, dom, dom_ms, per_kernel, per_step, st64, _ = timed_run("fp64")
    ms_step = 1000.0 * el / args.steps
    print(f"[bench] {args.steps} steps in {el:.3f}s: {ms_step:.3f} ms/step, {value:.4g} pu/s; {dom} "
          f"{dom_ms*1000:.1f} us/launch over the timed steps", file=sys.stderr, flush=True)
    for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1]):
        print(f"[bench]   {k:10s} {v*1000:9.1f} us/launch", file=sys.stderr)
    kb = KERNEL_BYTES_3D["fp64"][dom] * (1 if dim == 3 else 0.5)
    achieved = kb * nmp / (dom_ms * 1e-3) / 1e9
    step_ms = sum(per_step.values())
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, "fp64", dom),
            "kernel": dom, "kernel_ms": dom_ms, "bytes_per_particle": kb,
            "algorithmic_bytes": kb * nmp}
    step_roof = {"bytes_per_particle": B_ALG[dim], "achieved": B_ALG[dim] * value / 1e9,
                 "frac": B_ALG[dim] * value / 1e9 / HBM_PEAK_GBS, "kernel_ms_per_step": step_ms}
    finite = {"fp64": bool(np.isfinite(st64.mpts.s.tau).all())}
    fp32 = None
    if args.precision in ("all", "fp32"):
        v32, el32, d32, d32_ms, pk32, _, st32, _ = timed_run("fp32")
        finite["fp32"] = bool(np.isfinite(st32.mpts.s.tau).all())
        err32 = float(np.linalg.norm(st32.mpts.s.tau - st64.mpts.s.tau) / np.linalg.norm(st64.mpts.s.tau))
        x0 = ic.mpts.x
        errx32 = float(np.linalg.norm(st32.mpts.x - st64.mpts.x) / np.linalg.norm(st64.mpts.x - x0))
        print(f"[bench] fp32: {1000*el32/args.steps:.3f} ms/step, {v32:.4g} pu/s, stress err vs fp64 {err32:.2e}",
              file=sys.stderr, flush=True)
        for k, v in sorted(pk32.items(), key=lambda kv: -kv[1]):
            print(f"[bench]   fp32 {k:10s} {v*1000:9.1f} us/launch", file=sys.stderr)
        kb32 = KERNEL_BYTES_3D["fp32"][d32] * (1 if dim == 3 else 0.5)
        a32 = kb32 * nmp / (d32_ms * 1e-3) / 1e9
        fp32 = {"value": v32, "ms_per_step": 1000.0 * el32 / args.steps, "speedup_vs_fp64": v32 / value,
                "roofline": {"bound": "hbm", "achieved": a32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": a32 / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, "fp32", d32),
                             "kernel": d32, "kernel_ms": d32_ms, "bytes_per_particle": kb32},
                "rel_stress_err_vs_fp64": err32, "rel_displacement_err_vs_fp64": errx32,
                "kernels_us": {k: v * 1000 for k, v in pk32.items()},
                "dtype": "f32 particles, f64 grid and near-degenerate eigen-decompositions"}
        del st32
    if net is not None and args.finetune_steps > 0:
        # fine-tuning on the step's own plastic trial states, harvested
        # (ep_surrogate_harvest) every 5 steps of the same slump under another
        # random cohesion field: the network never sees the timed trajectory
        t0 = time.perf_counter()
        st_h = copy.deepcopy(ic)
        ep.grf.set_cohesion_field(st_h, **dict(GRF, seed=GRF["seed"] + 1))
        dev_h = device(st_h, "fp32_surrogate")
        harvested = 0
        for k in range(args.warmup + args.steps):
            dev_h.step(1, 1e9)
            if k % 5 == 4:
                harvested += net.harvest(dev_h, append=harvested > 0)
        sth = dev_h.stats()
        dev_h.close()
        del st_h
        t_h = time.perf_counter() - t0
        acc0 = (int(sth.predictions) - int(sth.fallbacks)) / max(int(sth.predictions), 1)
        t0 = time.perf_counter()
        lft = net.train(args.finetune_steps, min(65536, harvested // 32 * 32), 1e-3, 1e-5)
        t_f = time.perf_counter() - t0
        print(f"[bench] surrogate fine-tuning: {harvested} trial states harvested from {args.warmup + args.steps} "
              f"steps (cohesion field seed {GRF['seed'] + 1}; acceptance there with the C3 network {acc0:.3f}) in "
              f"{t_h:.2f}s, {args.finetune_steps} Adam steps in {t_f:.2f}s, loss {lft[:5].mean():.3e} -> "
              f"{lft[-20:].mean():.3e}", file=sys.stderr, flush=True)
        surr["finetune"] = {"harvest_field_seed": GRF["seed"] + 1, "harvest_steps": args.warmup + args.steps,
                            "harvest_every": 5, "samples": harvested, "steps": args.finetune_steps,
                            "lr": [1e-3, 1e-5], "harvest_s": t_h, "train_s": t_f,
                            "loss_first": float(lft[:5].mean()), "loss_last": float(lft[-20:].mean()),
                            "acceptance_before_on_harvest_run": acc0}
    if net is not None:
        vs, els, _, _, pks, _, sts, stt = timed_run("fp32_surrogate")
        finite["fp32_surrogate"] = bool(np.isfinite(sts.mpts.s.tau).all())
        errs = float(np.linalg.norm(sts.mpts.s.tau - st64.mpts.s.tau) / np.linalg.norm(st64.mpts.s.tau))
        x0 = ic.mpts.x
        errxs = float(np.linalg.norm(sts.mpts.x - st64.mpts.x) / np.linalg.norm(st64.mpts.x - x0))
        pred = int(stt.predictions)
        print(f"[bench] fp32+surrogate: {1000*els/args.steps:.3f} ms/step, {vs:.4g} pu/s, stress err vs fp64 "
              f"{errs:.2e}, predictions {pred}, fallbacks {int(stt.fallbacks)}", file=sys.stderr, flush=True)
        for k, v in sorted(pks.items(), key=lambda kv: -kv[1]):
            print(f"[bench]   surr {k:10s} {v*1000:9.1f} us/launch", file=sys.stderr)
        surr.update({"value": vs, "ms_per_step": 1000.0 * els / args.steps, "speedup_vs_fp64": vs / value,
                     "rel_stress_err_vs_fp64": errs, "rel_displacement_err_vs_fp64": errxs,
                     "predictions": pred, "accepted": pred - int(stt.fallbacks),
                     "acceptance": ((pred - int(stt.fallbacks)) / pred) if pred else None,
                     "fallback_fraction": (int(stt.fallbacks) / pred) if pred else None,
                     "max_scaled_residual": float(stt.max_residual), "tol": args.surr_tol,
                     "kernels_us": {k: v * 1000 for k, v in pks.items()}})
        del sts
    del st64
    longr = None
    if args.long_steps > 0 and args.precision == "all":
        longr = {"fp64": long_run("fp64", args.long_steps)}
        if net is not None:
            longr["fp32_surrogate"] = long_run("fp32_surrogate", args.long_steps)
        for k, v in longr.items():
            finite[f"long_{k}"] = v["state_finite"]
    if net is not None:
        net.close()
    ens = None
    if args.ensemble_steps > 0 and args.precision == "all":
        ens = ensemble_leg(args.ensemble_steps)
        finite["ensemble"] = ens["state_finite"]
    full = None
    if args.full_fill_steps > 0 and args.precision == "all" and args.workload == "slump3d_256x64x128":
        full = full_fill_leg(args.full_fill_steps, args.warmup, args.resort)
        finite["full_fill"] = full["state_finite"]
    cpu = None
    if not args.no_cpu_baseline:
        thr = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.workload, args.cpu_seconds, thr)
        one = cpu_baseline(args.workload, args.cpu_seconds, 1)
        cpu.update({"one_core": {"value": one["value"], "cores": 1, "sample": one["sample"]},
                    "nproc": os.cpu_count(), "cpu_model": _cpu_model()})
        print(f"[bench] cpu oracle: {cpu['value']:.4g} pu/s on {thr} threads, {one['value']:.4g} on 1 "
              f"({cpu['cpu_model']}, nproc {cpu['nproc']})", file=sys.stderr, flush=True)
    out = {
        "metric": "particle-updates/s on 1 MI355X (plain FP64 explicit MPM step)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (ic_slump column geometry, seeded deterministic lattice)",
        "config": {"workload": args.workload, "nmp": nmp, "nno": int(ic.mesh.nno[-1]), "basis": "bsmpm",
                   "transfer": "mUSL-FLIP", "deform": "finite", "locking": True, "plast": "DP",
                   "cohesion": (f"log-normal random field, mean {GRF['mean']:g} Pa, CoV {GRF['cov']}, "
                                f"correlation length {GRF['ell']} m, seed {GRF['seed']}" if dim == 3 else "uniform"),
                   "precision": "fp64", "resort": args.resort, "parallelism": "single-gpu"},
        "roofline": roof,
        "step_roofline": step_roof,
        "kernels_us": {k: v * 1000 for k, v in per_kernel.items()},
        "fp32": fp32,
        "fp32_surrogate": surr,
        "long_run": longr,
        "ensemble": ens,
        "full_fill": full,
        "state_finite": all(finite.values()),
        "state_finite_legs": finite,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def full_fill_leg(steps: int, warmup: int, resort: int) -> dict:
    """The 256 x 64 x 128 m domain filled with 2x2x2 particles per cell
    (16.8M particles), the same scheme and materials, FP64 and FP32: the
    particle count of the 16M-particle target on one GPU."""
    t0 = time.perf_counter()
    ic, cfg = build_workload("block3d_256x64x128_full")
    nmp = ic.mpts.nmp
    t_setup = time.perf_counter() - t0
    out = {"workload": "block3d_256x64x128_full", "nmp": nmp, "steps": steps, "warmup": warmup,
           "unit": "particle-updates/s", "setup_s": t_setup}
    finite = True
    for prec in ("fp64", "fp32"):
        instr = copy.deepcopy(cfg.instr)
        instr["ops"]["precision"] = {"fp64": 0, "fp32": 1}[prec]
        dev = hip.Device(ic.mesh, ic.mpts, ic.cmpr, instr)   # uploads the initial state; never downloaded
        dev.set_resort(resort)
        try:
            dev.step(warmup, 1e9)
            dev.profile(True)
            t0 = time.perf_counter()
            dev.step(steps, 1e9)
            el = time.perf_counter() - t0
            kt = dev.kernel_times()
            out[prec] = {"value": nmp * steps / el, "ms_per_step": 1000 * el / steps,
                         "kernels_us": {k: 1000 * ms / n for k, (ms, n) in kt.items()}}
        except hip.EpError as err:
            out[prec] = {"value": None, "error": str(err)}
            finite = False
        finally:
            dev.close()
        print(f"[bench] full fill {prec}: {nmp} particles, {out[prec].get('ms_per_step', float('nan')):.3f} ms/step, "
              f"{out[prec]['value'] or 0:.4g} pu/s", file=sys.stderr, flush=True)
    out["state_finite"] = finite
    return out


if __name__ == "__main__":
    main()
