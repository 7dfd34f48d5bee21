#!/usr/bin/env python3
"""Benchmark of the MI355X structured shallow-water step.

Workload: forward SSP-RK2 steps of a 4096 x 4096 FP64 tidal basin (dx = dy =
20 m, mean depth 30 m with 64 seeded Gaussian bumps, west edge tidal
eta = 1.0 sin(2 pi t / 44712 s), other edges walls, Manning n = 0.025, HLLC,
CFL 0.45), state resident in HBM.  A step is one full SSP-RK2 step (two
stage kernels) of the whole grid; the metric is cell-updates per second
(interior cells x steps / wall time).

Prints one JSON line with the throughput, the HBM roofline of the dominant
stage kernel (kernel time from HIP events on the launching stream) and a
CPU baseline (the oracle, OpenMP, on a bounded sample of the same run).

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20]

An "ensemble" object adds the batched-vs-loop comparison on 64 uneven
tidal basins (500 steps each); an "adjoint" object the cost of the gauge-
misfit gradient (w.r.t. bed and per-cell Manning n) on a 1024 x 1024 basin
over 500 steps against the forward run, and its central-difference check;
a "batched_adjoint" object the gradients of 32 uneven basins (256..1024
cells a side, 300 steps) in one batch against a loop over the members.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FP64 cell-updates/s, %HBM roofline @4096²; adjoint/fwd cost; batched vs loop"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per interior cell and launch (FP64 words x 8):
# stage 1 reads h, hu, hv, b and writes U* (7 words); stage 2 reads U*, b and
# U^n and writes U^{n+1} (10 words); 17 words = 136 B per cell-update.
STAGE_BYTES = {1: 56, 2: 80}


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=4096)
    ap.add_argument("--kernel-steps", type=int, default=50,
                    help="steps of the per-kernel timing pass")
    ap.add_argument("--tile-rows", type=int, default=0, help="0 = library choice by grid size")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU work of the baseline sample (0 disables)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = the CPUs of this process, up to OMP_NUM_THREADS")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the plain-copy bandwidth reference (0 disables)")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the uneven-ensemble comparison (0 disables)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--adjoint-n", type=int, default=1024, help="adjoint basin size (0 disables)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--adjoint-fd", type=int, default=32, help="finite-difference directions")
    ap.add_argument("--batch-members", type=int, default=32,
                    help="members of the batched-adjoint run (0 disables)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-fd", type=int, default=8, help="finite-difference directions per member")
    ap.add_argument("--feature-steps", type=int, default=50,
                    help="steps per optional-source-term variant of the 4096^2 step (0 disables)")
    return ap.parse_args()


def main():
    import faulthandler
    faulthandler.enable()
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step is measured "
              "on one GPU's grid; sharding over several GPUs is designed but not built or measured.",
              file=sys.stderr)
        return 2

    from rakali_amd import Grid
    from rakali_amd import scenarios as S

    scn = S.tidal_bumps(nx=args.nx, ny=args.ny, steps=args.steps)
    g = Grid(scn.nx, scn.ny, scn.dx, scn.dy, **scn.grid_kw)
    if args.tile_rows:
        g.tile_rows = args.tile_rows
    tile_rows = g.tile_rows
    for k in ("b", "h", "hu", "hv"):
        g.set_field(k, getattr(scn, k))
    cells = scn.nx * scn.ny

    g.step(args.warmup)
    g.sync()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = cells * args.steps / wall

    # per-kernel timing pass (events on the grid's stream around each launch)
    g.profile(True)
    g.step(args.kernel_steps)
    ms1, ms2, n = g.profile_read()
    g.profile(False)
    avg = {1: ms1 / n, 2: ms2 / n}
    dom = max(avg, key=avg.get)
    bytes_launch = STAGE_BYTES[dom] * cells
    achieved = bytes_launch / (avg[dom] * 1e-3) / 1e9
    per_step_ms = avg[1] + avg[2]
    step_gbs = (STAGE_BYTES[1] + STAGE_BYTES[2]) * cells / (per_step_ms * 1e-3) / 1e9
    t_sim, dt, nsteps = g.clock
    mass = g.total_mass()
    g.close()
    copy_gbs = copy_reference(args.copy_mib)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "forward SSP-RK2, 4096x4096 tidal basin (west tide A=1 m, T=44712 s; "
                        "walls elsewhere), dx=dy=20 m, 64 seeded bumps, Manning n=0.025, HLLC, CFL 0.45",
            "nx": scn.nx, "ny": scn.ny, "dx": scn.dx,
            "dry_fraction_initial": round(scn.dry_fraction(), 4),
            "tile_rows": tile_rows,
            "simulated_time_s": t_sim, "last_dt_s": dt, "total_steps": nsteps, "mass": mass,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_stage<{dom}> (stage {dom})",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom),
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": avg[dom],
            "stage_avg_ms": {"1": avg[1], "2": avg[2]},
            "step_algorithmic_GBs": step_gbs,
            "step_frac": step_gbs / HBM_PEAK_GBS,
            "copy_kernel_GBs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs if copy_gbs else None,
        },
    }
    if args.feature_steps > 0:
        out["source_terms"] = run_source_terms(scn, args.feature_steps)
    if args.ensemble_members > 0:
        out["ensemble"] = run_ensemble(args.ensemble_members, args.ensemble_steps)
    if args.adjoint_n > 0:
        out["adjoint"] = run_adjoint(args.adjoint_n, args.adjoint_steps, args.adjoint_fd)
    if args.batch_members > 0:
        out["batched_adjoint"] = run_batched_adjoint(args.batch_members, args.batch_steps, 30, args.batch_fd)
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = run_cpu_baseline(scn, args.cpu_seconds, args.cpu_threads)
    print(json.dumps(out), flush=True)
    return 0


def measured_traffic(stage: int):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3
    summary (FETCH_SIZE and WRITE_SIZE scaled by the same counters on a copy
    kernel of known size); None when no summary is present."""
    import re
    p = ROOT / "profiles" / "stage_kernels_tidal4096.md"
    if not p.exists():
        return None
    m = re.search(r"k_stage<%d>=(\d+)" % stage, p.read_text())
    return int(m.group(1)) if m else None


def run_source_terms(scn, steps: int, warmup: int = 5):
    """Stage-kernel time of the same 4096^2 step with the optional source
    terms on: surface forcing (wind, pressure, rain), a 64-cell sponge band
    on the east edge, a Chapman east edge, and all three.  Each set runs its
    own kernel instantiation; the plain kernels do not change."""
    import numpy as np
    from rakali_amd import Grid
    x = (np.arange(scn.nx) + 0.5)[:, None] * scn.dx
    y = (np.arange(scn.ny) + 0.5)[None, :] * scn.dy
    forcing = {
        "wind_x": np.asfortranarray(np.full((scn.nx, scn.ny), 0.1)),
        "wind_y": np.asfortranarray(np.full((scn.nx, scn.ny), -0.05)),
        "atm_pressure": np.asfortranarray(101325.0 - 0.01 * x - 0.005 * y),
        "precipitation": np.asfortranarray(np.full((scn.nx, scn.ny), 1e-6)),
    }
    variants = {
        "plain": dict(),
        "forcing": dict(forcing=True),
        "sponge": dict(bc=("tidal", "sponge", "wall", "wall"), sponge=(64, 1e-3, 30.0)),
        "chapman": dict(bc=("tidal", "chapman", "wall", "wall")),
        "all": dict(bc=("tidal", "chapman", "sponge", "wall"), sponge=(64, 1e-3, 30.0), forcing=True),
    }
    res = {}
    for name, v in variants.items():
        kw = dict(scn.grid_kw)
        if "bc" in v:
            kw["bc"] = v["bc"]
        g = Grid(scn.nx, scn.ny, scn.dx, scn.dy, **kw)
        for k in ("b", "h", "hu", "hv"):
            g.set_field(k, getattr(scn, k))
        if "sponge" in v:
            g.sponge = v["sponge"]
        if v.get("forcing"):
            g.forcing = True
            for k, a in forcing.items():
                g.set_field(k, a)
        g.step(warmup)
        g.profile(True)
        g.step(steps)
        ms1, ms2, n = g.profile_read()
        g.close()
        res[name] = {"stage1_ms": ms1 / n, "stage2_ms": ms2 / n, "step_ms": (ms1 + ms2) / n}
    base = res["plain"]["step_ms"]
    for name in res:
        res[name]["vs_plain"] = res[name]["step_ms"] / base
    return {"workload": f"the 4096x4096 step above, {steps} steps per variant (stage kernels, HIP events)",
            "variants": res}


def run_ensemble(n_members: int, steps: int, warmup: int = 2):
    """Batched vs looped: `n_members` uneven tidal basins (nx, ny log-uniform
    in [128, 2048], dx = 25 m, own bed, tide and Manning n), `steps` steps of
    every member.  Batched = one launch per RK stage for the whole ensemble
    (schedule from the autotuner, steps replayed from a hipGraph).  Loop =
    one grid per member, all on one stream, each member's steps launched
    back to back.  Both start from the same state and take the same steps;
    the final depths must agree bit for bit."""
    import ctypes
    import zlib

    import numpy as np

    from rakali_amd import Grid, _lib
    from rakali_amd import scenarios as S
    from rakali_amd.ensemble import from_scenarios

    scns = S.tidal_ensemble(n=n_members, seed=4, size=(128, 2048), dx=25.0, steps=steps)
    cells = sum(s.nx * s.ny for s in scns)

    e = from_scenarios(scns)
    tune_ms = e.autotune(steps=3)
    sched = e.schedule
    e.set_schedule(sched["schedule"], use_graph=True)
    e.step(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    batched_s = time.perf_counter() - t0
    crc_batched = 0
    for m in range(len(scns)):
        crc_batched = zlib.crc32(e.get_field(m, "h").tobytes(), crc_batched)
    e.close()

    L = _lib.load()
    grids = []
    stream = None
    for s in scns:
        g = Grid(s.nx, s.ny, s.dx, s.dy, stream=stream, **s.grid_kw)
        for k in ("b", "h", "hu", "hv"):
            g.set_field(k, getattr(s, k))
        if stream is None:
            p = ctypes.c_void_p()
            _lib.check("rkg_get_stream", L.rkg_get_stream(g.handle, ctypes.byref(p)))
            stream = p.value
        grids.append(g)
    for g in grids:
        g.step(warmup)
    grids[0].sync()
    t0 = time.perf_counter()
    for g in grids:
        g.step(steps)
    grids[0].sync()
    loop_s = time.perf_counter() - t0
    crc_loop = 0
    for g in grids:
        crc_loop = zlib.crc32(g.get_field("h").tobytes(), crc_loop)
    for g in grids:   # the library destroys the shared stream with the last grid using it
        g.close()

    return {
        "workload": f"{n_members} tidal basins, nx, ny log-uniform in [128, 2048], dx=25 m, own bed, "
                    f"tide A~U[0.5,2] m, Manning n~U[0.015,0.04], HLLC, {steps} steps each",
        "members": n_members, "cells": cells, "steps": steps,
        "batched": {"value": cells * steps / batched_s, "unit": "cell-updates/s",
                    "ms_per_step": batched_s / steps * 1e3, "schedule": sched,
                    "autotune_ms_per_step": tune_ms, "graph": True},
        "loop": {"value": cells * steps / loop_s, "unit": "cell-updates/s",
                 "ms_per_step": loop_s / steps * 1e3, "launches_per_step": 2 * n_members},
        "speedup": loop_s / batched_s,
        "identical_results": crc_batched == crc_loop,
    }


def run_adjoint(n: int, steps: int, fd_dirs: int, checkpoint: int = 50, n_gauges: int = 64):
    """Adjoint of an n x n tidal basin (dx = 50 m, the tidal-basin bed
    recipe and tide, still water, per-cell Manning n = 0.025, HLLC, CFL
    0.45), `steps` steps with a checkpoint every `checkpoint`.  Loss: gauge
    misfit at 64 seeded gauge cells over all steps, against the gauge series
    of a run with the bed plus 0.5 m Gaussian noise and n = 0.03.  Cost of
    the reverse sweep (segment replays included) against the plain forward
    run, and central differences of the forward run along `fd_dirs` seeded
    random directions in (b, n), with step 1e-6 and, because a step that
    large crosses the scheme's minmod and wave-speed kinks somewhere in a
    million cells over hundreds of steps, also with step 1e-8."""
    import numpy as np

    from rakali_amd import Grid
    from rakali_amd import scenarios as S
    from rakali_amd.adjoint import GaugeMisfit, start_still_water

    scn = S.tidal_bumps(nx=n, ny=n, dx=50.0, steps=steps)
    rng = np.random.default_rng(3)
    b = scn.b
    nf = np.full((n, n), 0.025)
    gauges = [(int(rng.integers(0, n)), int(rng.integers(0, n))) for _ in range(n_gauges)]
    kw = dict(scn.grid_kw)
    kw.pop("manning_n")
    g = Grid(n, n, scn.dx, scn.dy, manning_n=0.025, **kw)
    sess = GaugeMisfit(g, steps, checkpoint, gauges)
    # target: perturbed bed, n = 0.03
    start_still_water(g, b + 0.5 * rng.normal(size=b.shape), np.full((n, n), 0.03))
    sess.forward()
    sess.set_targets(sess.gauge_eta())
    # plain forward cost (same kernels: per-cell n)
    start_still_water(g, b, nf)
    g.step(2)
    g.sync()
    start_still_water(g, b, nf)
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    fwd_s = time.perf_counter() - t0
    # loss and gradients
    start_still_water(g, b, nf)
    t0 = time.perf_counter()
    loss = sess.forward()
    sess.backward()
    sess.fold_still_water(0.0)
    adj_s = time.perf_counter() - t0
    f_ms, b_ms = sess.timing()
    gb, gn = sess.gradient("b"), sess.gradient("n")
    fd_steps = (1e-6, 1e-8)
    errs = {e: [] for e in fd_steps}
    drng = np.random.default_rng(11)
    fd_t0 = time.perf_counter()
    for _ in range(fd_dirs):
        db = drng.normal(size=b.shape)
        dn = drng.normal(size=b.shape) * 1e-3
        ad = float(np.sum(gb * db) + np.sum(gn * dn))
        for eps in fd_steps:
            vals = []
            for sgn in (1.0, -1.0):
                start_still_water(g, b + sgn * eps * db, nf + sgn * eps * dn)
                vals.append(sess.forward())
            fd = (vals[0] - vals[1]) / (2 * eps)
            errs[eps].append(abs(fd - ad) / max(abs(fd), 1e-300))
    fd_s = time.perf_counter() - fd_t0
    sess.close()
    g.close()
    return {
        "workload": f"{n}x{n} tidal basin, dx=50 m, {steps} steps, checkpoint every {checkpoint}, "
                    f"{n_gauges} gauges; target from bed + 0.5 m noise and n = 0.03; gradient w.r.t. "
                    f"bed and per-cell Manning n",
        "loss": loss,
        "forward_ms_per_step": fwd_s / steps * 1e3,
        "loss_forward_ms_per_step": f_ms / steps,
        "reverse_ms_per_step": b_ms / steps,
        "adjoint_over_forward": b_ms / (fwd_s * 1e3),
        "forward_plus_adjoint_over_forward": adj_s / fwd_s,
        "fd_directions": fd_dirs,
        "fd_max_rel_err": max(errs[1e-6]) if fd_dirs else None,
        "fd_median_rel_err": float(np.median(errs[1e-6])) if fd_dirs else None,
        "fd_max_rel_err_step_1e-8": max(errs[1e-8]) if fd_dirs else None,
        "fd_median_rel_err_step_1e-8": float(np.median(errs[1e-8])) if fd_dirs else None,
        "fd_seconds": fd_s,
        "max_abs_grad_b": float(np.abs(gb).max()),
        "max_abs_grad_n": float(np.abs(gn).max()),
    }


def run_batched_adjoint(n_members: int, steps: int, checkpoint: int, fd_dirs: int, n_gauges: int = 64):
    """Gradients of 32 uneven tidal basins in one batch: nx, ny log-uniform
    in [256, 1024], dx = 50 m, the tidal-basin recipe (own bed, tide and
    Manning field per member), still water, `steps` steps with a checkpoint
    every `checkpoint`; loss per member as in the adjoint workload (64
    seeded gauges, target from the bed plus 0.5 m noise and n = 0.03).
    Batched: every member's forward run, then every member's reverse sweep,
    with the members' launches interleaved step by step on their own
    streams.  Loop: one member after the other.  The gradients of the two
    must agree bit for bit; central differences (steps 1e-6 and 1e-8) along
    `fd_dirs` seeded random directions in (b, n) check each member.

    Each bed carries 1 cm of seeded roughness.  The members' narrow bumps
    leave most of the -30 m base exactly flat, and two neighbours with the
    same bed level put the hydrostatic reconstruction on the kink of
    max(b_L, b_R): along a random direction the loss then changes by
    max(db_L, db_R), which is not linear in db, and central differences
    measure no derivative to compare with."""
    import numpy as np

    from rakali_amd import Grid
    from rakali_amd import scenarios as S
    from rakali_amd.adjoint import GaugeMisfit, backward_many, forward_many, start_still_water

    scns = S.tidal_ensemble(n_members, seed=5, size=(256, 1024), dx=50.0, steps=steps)
    grids, sess, beds, nfs = [], [], [], []
    for m, scn in enumerate(scns):
        rng = np.random.default_rng(100 + m)
        kw = dict(scn.grid_kw)
        n0 = kw.pop("manning_n")
        g = Grid(scn.nx, scn.ny, scn.dx, scn.dy, manning_n=n0, **kw)
        gauges = [(int(rng.integers(0, scn.nx)), int(rng.integers(0, scn.ny))) for _ in range(n_gauges)]
        sess.append(GaugeMisfit(g, steps, checkpoint, gauges))
        bed = scn.b + 0.01 * np.random.default_rng(500 + m).normal(size=scn.b.shape)
        start_still_water(g, bed + 0.5 * rng.normal(size=scn.b.shape), np.full(scn.b.shape, 0.03))
        grids.append(g)
        beds.append(bed)
        nfs.append(np.full(scn.b.shape, n0))
    forward_many(sess)
    for s in sess:
        s.set_targets(s.gauge_eta())

    def start(db=None, dn=None, eps=0.0):
        for m, g in enumerate(grids):
            b = beds[m] if db is None else beds[m] + eps * db[m]
            nf = nfs[m] if dn is None else nfs[m] + eps * dn[m]
            start_still_water(g, b, nf)

    cells = sum(scn.nx * scn.ny for scn in scns)
    start()
    forward_many(sess)                        # warm-up of both passes
    backward_many(sess)
    start()
    t0 = time.perf_counter()
    losses = forward_many(sess)
    t1 = time.perf_counter()
    backward_many(sess)
    for s in sess:
        s.fold_still_water(0.0)
    t2 = time.perf_counter()
    grads = [(s.gradient("b"), s.gradient("n")) for s in sess]
    # loop: one member at a time
    t3 = time.perf_counter()
    same = True
    for m, (g, s) in enumerate(zip(grids, sess)):
        start_still_water(g, beds[m], nfs[m])
        s.forward()
        s.backward()
        s.fold_still_water(0.0)
        same = same and np.array_equal(s.gradient("b"), grads[m][0]) and np.array_equal(s.gradient("n"), grads[m][1])
    t4 = time.perf_counter()
    # central differences, all members per batched forward pair.  Besides
    # |fd - ad| / |fd|, the error is also given relative to sum |g_i d_i|,
    # the size of the terms the directional derivative sums: a direction
    # along which they cancel makes the plain relative error measure the
    # difference quotient's rounding (about 1e-16 L / step) instead.
    # Each direction also gets its two one-sided quotients (L(+eps) - L(0)) /
    # eps and (L(0) - L(-eps)) / eps.  Where they disagree by more than 1 %
    # (plus the rounding of the loss) the segment [-eps, eps] crosses a kink
    # of the piecewise-smooth scheme (a wet/dry switch, a minmod or max
    # branch): the central difference there is not the derivative of either
    # side, and such directions are counted apart.
    fd_res = {}
    for eps in (1e-6, 1e-8):
        errs, scaled, worst, kinks, smooth = [], [], None, 0, []
        drng = np.random.default_rng(17)
        for _ in range(fd_dirs):
            db = [drng.normal(size=b.shape) for b in beds]
            dn = [drng.normal(size=b.shape) * 1e-3 for b in beds]
            start(db, dn, eps)
            lp = forward_many(sess)
            start(db, dn, -eps)
            lm = forward_many(sess)
            for m in range(n_members):
                fd = (lp[m] - lm[m]) / (2 * eps)
                qp = (lp[m] - losses[m]) / eps
                qm = (losses[m] - lm[m]) / eps
                kink = abs(qp - qm) > 1e-2 * max(abs(qp), abs(qm)) + 1e-15 * abs(losses[m]) / eps
                terms = np.concatenate([(grads[m][0] * db[m]).ravel(), (grads[m][1] * dn[m]).ravel()])
                ad = float(np.sum(terms))
                e = abs(fd - ad) / max(abs(fd), 1e-300)
                errs.append(e)
                kinks += int(kink)
                if not kink:
                    smooth.append(e)
                scaled.append(abs(fd - ad) / float(np.sum(np.abs(terms))))
                if worst is None or e > worst[0]:
                    worst = (e, m, fd, ad, float(lp[m]), qp, qm, kink)
        fd_res[eps] = {"max_rel_err": max(errs) if errs else None,
                       "median_rel_err": float(np.median(errs)) if errs else None,
                       "max_err_over_sum_abs_terms": max(scaled) if scaled else None,
                       "kink_directions": kinks,
                       "smooth_directions": len(smooth),
                       "smooth_max_rel_err": max(smooth) if smooth else None,
                       "smooth_median_rel_err": float(np.median(smooth)) if smooth else None,
                       "worst": None if worst is None else
                       {"member": worst[1], "fd": worst[2], "adjoint": worst[3], "loss": worst[4],
                        "one_sided_plus": worst[5], "one_sided_minus": worst[6], "kink": bool(worst[7])}}
    for s in sess:
        s.close()
    for g in grids:
        g.close()
    fwd_s, rev_s, loop_s = t1 - t0, t2 - t1, t4 - t3
    return {
        "workload": f"{n_members} tidal basins, nx, ny log-uniform in [256, 1024], dx=50 m, beds with 1 cm "
                    f"roughness, {steps} steps, checkpoint every {checkpoint}, {n_gauges} gauges each; gradient "
                    f"w.r.t. each member's bed and per-cell Manning n",
        "members": n_members,
        "cells": cells,
        "batched_forward_s": fwd_s,
        "batched_reverse_s": rev_s,
        "batched_reverse_over_forward": rev_s / fwd_s,
        "batched_cell_steps_per_s": cells * steps / (fwd_s + rev_s),
        "loop_forward_plus_reverse_s": loop_s,
        "speedup_vs_loop": loop_s / (fwd_s + rev_s),
        "identical_gradients": bool(same),
        "fd_directions_per_member": fd_dirs,
        "fd_step_1e-6": fd_res[1e-6],
        "fd_step_1e-8": fd_res[1e-8],
        "total_loss": float(np.sum(losses)),
    }


def copy_reference(mib: int):
    """Achievable HBM bandwidth of a plain copy kernel (read + write bytes)."""
    if mib <= 0:
        return None
    import ctypes
    from rakali_amd import _lib
    L = _lib.load()
    gbs, ms = ctypes.c_double(), ctypes.c_double()
    _lib.check("rkg_copy_bandwidth", L.rkg_copy_bandwidth(mib << 20, 20, ctypes.byref(gbs), ctypes.byref(ms)))
    return gbs.value


def run_cpu_baseline(scn, seconds: float, threads: int):
    """The oracle (oracle/rk_oracle.c, OpenMP over rows) timed on this host on
    a bounded sample of the same run: the same bed and initial state, as many
    steps as fit in about `seconds` (at least one)."""
    import oracle as O

    # every CPU this process may run on, up to the OpenMP thread count the
    # host sets for one GPU's share of the machine (OMP_NUM_THREADS)
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS") or affinity)
    nthreads = threads or max(1, min(affinity, share))
    O.set_threads(nthreads)
    kw = dict(scn.grid_kw)
    bc = kw.pop("bc")
    p = O.make_params(scn.nx, scn.ny, ng=2, dx=scn.dx, dy=scn.dy, bc=bc,
                      use_hllc=int(kw.pop("use_hllc")), **kw)
    sim = O.Sim(p)
    for k in ("b", "h", "hu", "hv"):
        sim.set_interior(k, getattr(scn, k))
    sim.finalize()
    t0 = time.perf_counter()
    sim.step(1)
    one = time.perf_counter() - t0
    more = max(0, int(seconds / max(one, 1e-9)) - 1)
    t0 = time.perf_counter()
    if more:
        sim.step(more)
    rest = time.perf_counter() - t0
    steps = 1 + more
    wall = one + rest
    return {
        "value": scn.nx * scn.ny * steps / wall,
        "unit": "cell-updates/s",
        "cores": nthreads,
        "host_cpus": os.cpu_count(),
        "affinity_cpus": affinity,
        "kind": "port",
        "sample": f"{steps} SSP-RK2 steps of the same {scn.nx}x{scn.ny} run from the same initial "
                  f"state ({wall:.1f} s of CPU work, {nthreads} OpenMP threads on a host of "
                  f"{os.cpu_count()} CPUs, {affinity} in this process's affinity mask, "
                  f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')})",
    }


if __name__ == "__main__":
    sys.exit(main())
