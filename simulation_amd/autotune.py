"""Autotuner for the temporally blocked launches.

Chooses the temporal depth (leapfrog steps fused per launch), rows per block
and columns per lane -- plus, for batches, the number of resident waves --
by timing candidate launch shapes on the device, and caches the winner in a
JSON file keyed by device name, problem shape and form.  Every candidate
computes the same values bit for bit (the shape only changes which cells are
computed when), so the choice is purely a speed decision.  The reference has
no counterpart: it launches fixed 16x16 blocks, one step per launch
(fd2d/pycuda/fd2d_3_4.py:345-351).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import torch

from . import fd2d

CACHE = os.environ.get("FDTD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "fdtd_mi355x_tune.json"))


def candidates(nx: int):
    """(depth, rows, cols, form) shapes tried for an nx-row grid: single-wave
    strips at depth 10 (and from 4096 rows on depth 11) with row blocks of
    nx/170, nx/128 and nx/96 rows and at depth 8 with nx/128 (MI355X sweeps
    put the optimum at depth 10 at 2048^2 and depth 11 from 4096^2 on, rows
    ~nx/170 at 2048^2 and ~nx/128 from 4096^2 on), two columns per lane at
    depth 8, and the multi-wave form with short blocks."""
    def r(d):
        return min(512, max(8, nx // d))
    for depth in ((10, 11) if nx >= 4096 else (10,)):
        for rows in sorted({r(170), r(128), r(96)}):
            yield depth, rows, 1, 1
    yield 8, r(128), 1, 1
    yield 8, r(128), 2, 1
    yield 8, 32, 1, 2


def batch_candidates():
    """(depth, rows, cols, waves) shapes tried for a batch: depth 8 with
    96-160 rows and depth 10 with 128 (on MI355X the 256-grid batch of sides
    128-1024 peaks at depth 8, 128 rows), and two columns per lane."""
    for rows in (96, 128, 160):
        yield 8, rows, 1, 0
    yield 10, 128, 1, 0
    yield 8, 128, 2, 0


def _device():
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def _time(fn, reps: int = 2) -> float:
    """Best of `reps` timed calls after one untimed call."""
    fn()                                   # warm-up: first touch of the shape
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(solver: fd2d.TMSolver, steps: int = 80, use_cache: bool = True, configs=None,
         verbose: bool = False) -> dict:
    """Time (depth, rows, cols[, form]) candidates on this solver's grid and
    set the fastest on it.  The solver's state, probe traces and step are
    restored."""
    key = (f"tm2|{_device()}|{solver.nx}x{solver.ny}|lossy={int(solver.lossy)}"
           f"|pml={int(solver.pml is not None)}")
    db = _load() if use_cache and configs is None else {}
    if key in db:
        best = db[key]
    else:
        saved = {k: getattr(solver, k).clone() for k in fd2d.STATE if getattr(solver, k) is not None}
        probe_saved = None if solver.probe_out is None else solver.probe_out.clone()
        t_saved = solver.t
        base = steps if not solver.probes else min(steps, solver.probe_len)
        results = []
        for cand in (configs or candidates(solver.nx)):
            depth, rows, cols = cand[:3]
            form = cand[3] if len(cand) > 3 else 0
            solver.depth, solver.rows, solver.cols, solver.form = depth, rows, cols, form
            # whole pairs of full-depth launches where the run is long enough
            # (a long run's launches are nearly all full depth), else the
            # run's own length
            n = base - base % (2 * depth) if base >= 2 * depth else base - base % 2

            def go(n=n):
                solver.t = 0
                solver.run(n)
            dt = _time(go)
            rate = solver.nx * solver.ny * n / dt
            results.append((rate, depth, rows, cols, form))
            if verbose:
                print(f"  tm depth={depth} rows={rows} cols={cols} form={form}: "
                      f"{rate / 1e9:.1f} G/s", flush=True)
        for k, v in saved.items():
            getattr(solver, k).copy_(v)
        if probe_saved is not None:
            solver.probe_out.copy_(probe_saved)
        solver.t = t_saved
        rate, depth, rows, cols, form = max(results)
        best = {"depth": depth, "rows": rows, "cols": cols, "form": form,
                "gcups": round(rate / 1e9, 2)}
        if use_cache and configs is None:
            db = _load()
            db[key] = best
            _save(db)
    solver.depth, solver.rows, solver.cols = best["depth"], best["rows"], best["cols"]
    solver.form = best.get("form", 0)
    return best


def tune_batch(b, steps: int = 48, use_cache: bool = True, configs=None,
               verbose: bool = False) -> dict:
    """Time (depth, rows, cols, waves) candidates on a BatchSolver before its
    first step and plan the fastest; the batch is reset to step 0 afterwards."""
    if b.t != 0:
        raise ValueError("tune_batch runs before the first step")
    sig = hashlib.sha1(repr([(s.nx, s.ny, s.npml) for s in b.specs]).encode()).hexdigest()[:16]
    key = f"batch|{_device()}|{len(b.specs)} grids|{sig}"
    db = _load() if use_cache and configs is None else {}
    if key in db:
        best = db[key]
    else:
        results = []
        for depth, rows, cols, waves in (configs or batch_candidates()):
            n = 2 * depth * max(1, steps // (2 * depth))
            if n > b.nsteps_max:
                continue
            b.plan(depth, rows, cols, waves)

            def go():
                b.t = 0
                b.run(n)
            dt = _time(go)
            rate = b.cells * n / dt
            results.append((rate, depth, rows, cols, waves))
            if verbose:
                print(f"  batch depth={depth} rows={rows} cols={cols}: {rate / 1e9:.1f} G/s",
                      flush=True)
        rate, depth, rows, cols, waves = max(results)
        best = {"depth": depth, "rows": rows, "cols": cols, "waves": waves,
                "gcups": round(rate / 1e9, 2)}
        if use_cache and configs is None:
            db = _load()
            db[key] = best
            _save(db)
    b.plan(best["depth"], best["rows"], best["cols"], best["waves"])
    b.reset()
    return best
