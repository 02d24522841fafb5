"""On-disk formats: field snapshots, probe traces, checkpoint/resume.

The reference writes only plots and animations (PNG/MP4; its animation
scripts keep every frame in memory, fd2d/animation/fd2d_3_4.py).  Here:

* snapshots: one ``.npy`` per field and step, ``<field>_t<step:06d>.npy``,
  (nx, ny) or (nx,) float arrays exactly as on the device;
* probe traces: CSV, one row per step ``t, probe_0, probe_1, ...``, header
  ``t,ez[i;j],...``; values written with ``repr`` of the float32 (round-trip
  exact);
* checkpoint: one ``.npz`` (no pickled objects) holding the solver's whole
  state set and its step count.  The scheme is deterministic, so a solver
  restored from a checkpoint at step t and advanced by n steps holds the same
  bits as the original advanced from t by n steps.

Checkpoints cover the solvers of this package: fd2d.TMSolver (dz iz hx hy
ihx ihy ez, probe traces), learned.LearnedTMSolver (+ the boundary-sample
history), fd2d.PlaneWaveSolver (+ incident grid, its boundary state and the
DFT planes) and fd1d.DFormSolver (dx ix sx ex hy bc, DFT planes, probes).
"""
from __future__ import annotations

import csv
import json
import os

import numpy as np
import torch

FORMAT = 1


def snapshot_path(out_dir: str, field: str, t: int) -> str:
    return os.path.join(out_dir, f"{field}_t{int(t):06d}.npy")


def write_snapshot(out_dir: str, field: str, t: int, value) -> str:
    """Save one field (a device tensor or array) at step t as .npy."""
    os.makedirs(out_dir, exist_ok=True)
    a = value.detach().cpu().numpy() if isinstance(value, torch.Tensor) else np.asarray(value)
    p = snapshot_path(out_dir, field, t)
    np.save(p, a, allow_pickle=False)
    return p


def read_snapshot(out_dir: str, field: str, t: int) -> np.ndarray:
    return np.load(snapshot_path(out_dir, field, t), allow_pickle=False)


def write_probes_csv(path: str, traces, names, t0: int = 0) -> None:
    """traces (nprobe, nsteps) of steps t0+1 .. t0+nsteps; names: one label
    per probe (e.g. "ez[40;60]")."""
    tr = np.asarray(traces.detach().cpu().numpy() if isinstance(traces, torch.Tensor) else traces)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["t"] + list(names))
        for s in range(tr.shape[1]):
            w.writerow([t0 + s + 1] + [repr(float(v)) for v in tr[:, s]])


def read_probes_csv(path: str, dtype=np.float32):
    """(t array, names, traces (nprobe, nsteps))."""
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    names = rows[0][1:]
    t = np.array([int(r[0]) for r in rows[1:]], np.int64)
    tr = np.array([[float(v) for v in r[1:]] for r in rows[1:]], dtype).T.reshape(len(names), -1)
    return t, names, tr


# ------------------------------------------------ checkpoint / resume

def _fields(solver) -> tuple[str, dict]:
    """(kind, {name: tensor or None}) of a solver's state set."""
    from . import fd1d, fd2d
    kind = type(solver).__name__
    if isinstance(solver, fd2d.TMSolver):
        st = {k: getattr(solver, k) for k in ("dz", "iz", "hx", "hy", "ihx", "ihy", "ez")}
        st["probe_out"] = solver.probe_out
        return kind, st
    if isinstance(solver, fd2d.PlaneWaveSolver):
        _, st = _fields(solver.tm)
        st.update({k: getattr(solver, k) for k in ("ezi", "hxi", "bc")})
        if solver.ft is not None:
            st.update({"ft_" + k: getattr(solver.ft, k) for k in solver.ft._fields})
        return kind, st
    if isinstance(solver, fd1d.DFormSolver):
        st = {k: getattr(solver, k) for k in ("dx", "ix", "sx", "ex", "hy", "bc", "probe_out")}
        if solver.ft is not None:
            st.update({"ft_" + k: getattr(solver.ft, k) for k in solver.ft._fields})
        return kind, st
    from . import learned
    if isinstance(solver, learned.LearnedTMSolver):
        _, st = _fields(solver.tm)
        st["hist"] = solver.hist
        return kind, st
    raise TypeError(f"no checkpoint format for {kind}")


def _step_owner(solver):
    from . import fd1d
    return solver if isinstance(solver, fd1d.DFormSolver) else getattr(solver, "tm", solver)


def save_checkpoint(path: str, solver) -> None:
    """Write the solver's state set and step count to path (.npz)."""
    kind, st = _fields(solver)
    torch.cuda.synchronize()
    arrays = {k: v.detach().cpu().numpy() for k, v in st.items() if v is not None}
    meta = {"format": FORMAT, "kind": kind, "t": int(_step_owner(solver).t),
            "shapes": {k: list(a.shape) for k, a in arrays.items()}}
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez(path, meta=np.array(json.dumps(meta)), **arrays)


def load_checkpoint(path: str, solver) -> int:
    """Restore a checkpoint into a solver built with the same configuration
    (same kind, shapes and optional fields); returns the restored step."""
    kind, st = _fields(solver)
    with np.load(path, allow_pickle=False) as d:
        meta = json.loads(str(d["meta"]))
        if meta.get("format") != FORMAT or meta.get("kind") != kind:
            raise ValueError(f"{path}: a {meta.get('kind')} checkpoint, not {kind}")
        for k, v in st.items():
            if v is None:
                if k in d.files:
                    raise ValueError(f"{path}: holds {k}, which this solver does not have")
                continue
            if k not in d.files:
                raise ValueError(f"{path}: {k} missing")
            a = d[k]
            if tuple(a.shape) != tuple(v.shape):
                raise ValueError(f"{path}: {k} has shape {a.shape}, the solver {tuple(v.shape)}")
            v.copy_(torch.from_numpy(a).to(v.device))
    # a restored state may hold another run's iz: free-space flags are
    # recorded afresh by the next run
    tm = getattr(solver, "tm", solver)
    if hasattr(tm, "forget_tiles"):
        tm.forget_tiles()
    _step_owner(solver).t = int(meta["t"])
    torch.cuda.synchronize()
    return int(meta["t"])
