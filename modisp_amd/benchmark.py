"""Backend benchmark with ROCm rows: the reference's timing protocol on R64 / R32.

The reference times its backends with tools/benchmark_backends.py: paths
C64 / G64 / G32 (CPU FP64, CUDA FP64, CUDA FP32; PATHS and ORDERS at
:35-41), one first call, three warm-up calls and twelve steady calls per path
in rotating order, medians and quartiles of the steady calls, valid roots/s =
status-0 cells / median call time (:556-559), an NVML monitor (:259-345) and
an environment capture (:68-84). This module runs that protocol on this
plugin's two front-end paths:

* R64: ``Solver(backend="rocm", precision="float64")`` -- the plain FP64 grid
  scan, FP64 refinement, float64 velocities;
* R32: ``Solver(backend="rocm", precision="float32")`` -- FP32 grid scan with
  FP64 recheck and refinement, float32 velocities;

in one process with one solver per path, rotating the path order every
round. The device monitor samples ``amd-smi metric`` (or ``rocm-smi`` when
amd-smi is absent) in a background thread; the environment records
``hipcc --version``, ``amd-smi static``, ``lscpu``, CPU affinity and the
library digest. Each call's result arrays are hashed, so a changed repeat
is reported.

    python -m modisp_amd.benchmark --case plain --output bench-plain.json
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import platform
import shutil
import subprocess
import sys
import threading
import time
from pathlib import Path

import numpy as np

PATHS = {"R64": dict(backend="rocm", precision="float64"), "R32": dict(backend="rocm", precision="float32")}
ORDERS = (("R64", "R32"), ("R32", "R64"))
SIGNED = ("phase_velocity", "mode_mask", "status", "trace_flags", "search_level", "root_count", "search_bounds")
CASES = {  # name: (workload factory, period slice, wave, modes)
    "plain": ("plain_batch", slice(None), "rayleigh", 3),
    "precision_rayleigh": ("precision_batch", slice(None), "rayleigh", 5),
    "precision_love": ("precision_batch", slice(None), "love", 5),
    "gradient": ("gradient_batch", slice(None), "rayleigh", 3),
    "sweep_rayleigh": ("sweep_batch", slice(None), "rayleigh", 4),
}


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def run(command, timeout=20):
    """A tool's output, or the reason it could not run."""
    exe = shutil.which(command[0]) or (f"/opt/rocm/bin/{command[0]}" if Path(f"/opt/rocm/bin/{command[0]}").exists()
                                      else None)
    if exe is None:
        return dict(command=command, error="not installed")
    try:
        p = subprocess.run([exe, *command[1:]], capture_output=True, text=True, timeout=timeout)
        return dict(command=command, returncode=p.returncode, stdout=p.stdout[-20000:], stderr=p.stderr[-4000:])
    except Exception as exc:  # noqa: BLE001 - recorded, never fatal
        return dict(command=command, error=f"{type(exc).__name__}: {exc}")


def environment():
    from . import _native, VERSION
    return dict(plugin_version=VERSION, library=str(_native.LIBRARY), library_sha256=sha256(_native.LIBRARY),
                python=platform.python_version(), numpy=np.__version__, platform=platform.platform(),
                affinity=sorted(os.sched_getaffinity(0)), load_average=list(os.getloadavg()),
                env={k: os.environ.get(k) for k in ("HIP_VISIBLE_DEVICES", "ROCR_VISIBLE_DEVICES", "OMP_NUM_THREADS",
                                                    "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "GPU_MAX_HW_QUEUES")},
                hipcc=run(["hipcc", "--version"]), gpu=run(["amd-smi", "static", "--json"]),
                lscpu=run(["lscpu"]))


def parse_amd_smi(text):
    """GPU use (%), memory in use (bytes), power (W) and junction temperature
    (C) from ``amd-smi metric --json`` output (one device); missing fields
    are None."""
    try:
        data = json.loads(text)
    except json.JSONDecodeError:
        return {}
    if isinstance(data, dict) and "gpu_data" in data:
        data = data["gpu_data"]
    dev = data[0] if isinstance(data, list) and data else data if isinstance(data, dict) else {}

    def value(*keys):
        x = dev
        for k in keys:
            if not isinstance(x, dict) or k not in x:
                return None
            x = x[k]
        if isinstance(x, dict):
            x = x.get("value")
        try:
            return float(x)
        except (TypeError, ValueError):
            return None
    used = value("mem_usage", "used_vram")
    return dict(gpu_util=value("usage", "gfx_activity"), memory_used_bytes=None if used is None else used * 2 ** 20,
                power_w=value("power", "socket_power"), temperature_c=value("temperature", "hotspot"))


class Monitor:
    """Device samples every `interval` seconds from amd-smi in a thread."""

    def __init__(self, interval=.5):
        self.interval = interval
        self.samples, self.errors = [], []
        self._stop = threading.Event()
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self.tool = "amd-smi" if (shutil.which("amd-smi") or Path("/opt/rocm/bin/amd-smi").exists()) else None

    def _loop(self):
        while not self._stop.is_set():
            t0 = time.monotonic()
            if self.tool is None:
                if "amd-smi not installed" not in self.errors:
                    self.errors.append("amd-smi not installed")
            else:
                out = run(["amd-smi", "metric", "--json"], timeout=10)
                if "stdout" in out and out.get("returncode") == 0:
                    self.samples.append(dict(time_ns=time.time_ns(), load_average=os.getloadavg()[0],
                                             **parse_amd_smi(out["stdout"])))
                else:
                    label = out.get("error") or (out.get("stderr") or "")[:200]
                    if label not in self.errors:
                        self.errors.append(label)
            self._stop.wait(max(0., self.interval - (time.monotonic() - t0)))

    def __enter__(self):
        self._thread.start()
        return self

    def __exit__(self, *exc):
        self._stop.set()
        self._thread.join(timeout=30)

    def summary(self):
        def span(key):
            v = [s[key] for s in self.samples if s.get(key) is not None]
            return [min(v), max(v)] if v else None
        return dict(tool=self.tool, requested_interval_s=self.interval, sample_count=len(self.samples),
                    gpu_util_range=span("gpu_util"), memory_used_bytes_range=span("memory_used_bytes"),
                    power_w_range=span("power_w"), temperature_c_range=span("temperature_c"),
                    load_average_range=span("load_average"), errors=self.errors)


def signature(result):
    h = hashlib.sha256()
    for key in SIGNED:
        a = np.ascontiguousarray(getattr(result, key))
        if a.dtype.kind == "f":
            a = np.where(np.isnan(a), np.nan, a)   # one NaN payload
        h.update(key.encode())
        h.update(a.dtype.str.encode())
        h.update(a.tobytes())
    return h.hexdigest()


def protocol(calls, warmups=3, rounds=12):
    """Time callables {path: f() -> result} by the reference protocol: one
    first call per path, `warmups` warm-up rounds and `rounds` steady rounds,
    the path order rotating through ORDERS. Returns per-path records."""
    records = {p: dict(first=None, warmup=[], steady=[]) for p in calls}
    orders = [o for o in ORDERS if set(o) == set(calls)] or [tuple(calls)]

    def one(path, tag, keep):
        t0 = time.perf_counter_ns()
        res = calls[path]()
        seconds = (time.perf_counter_ns() - t0) * 1e-9
        valid = int(np.count_nonzero(res.status == 0))
        entry = dict(tag=tag, seconds=seconds, valid_roots=valid, signature=signature(res))
        if keep:
            entry["status_counts"] = {str(s): int(np.count_nonzero(res.status == s)) for s in range(4)}
        return entry

    for path in calls:
        records[path]["first"] = one(path, "first", True)
    for w in range(warmups):
        for path in orders[w % len(orders)]:
            records[path]["warmup"].append(one(path, f"warm{w}", False))
    for r in range(rounds):
        for pos, path in enumerate(orders[r % len(orders)]):
            e = one(path, f"steady{r:02}", r in (0, rounds - 1))
            e.update(round=r, position=pos)
            records[path]["steady"].append(e)
    for path, rec in records.items():
        t = np.array([e["seconds"] for e in rec["steady"]])
        roots = [e["valid_roots"] for e in rec["steady"]]
        first_sig = rec["first"]["signature"]
        changed = sum(e["signature"] != first_sig for e in rec["warmup"] + rec["steady"])
        med = float(np.median(t)) if t.size else None
        rec["summary"] = dict(steady_calls=int(t.size), median_s=med,
                              q25_s=float(np.quantile(t, .25)) if t.size else None,
                              q75_s=float(np.quantile(t, .75)) if t.size else None,
                              min_s=float(t.min()) if t.size else None, max_s=float(t.max()) if t.size else None,
                              valid_roots=roots[-1] if roots else None,
                              valid_roots_per_s=(roots[-1] / med) if t.size and med else None,
                              repeated_changes=int(changed))
    return records


def case_inputs(name, models=None):
    from . import workloads
    factory, sl, wave, modes = CASES[name]
    kw = {} if models is None else dict(B=int(models))
    wl = getattr(workloads, factory)(**kw)
    return wl, np.asarray(wl.periods)[sl], wave, modes


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m modisp_amd.benchmark")
    ap.add_argument("--case", choices=sorted(CASES), default="plain")
    ap.add_argument("--models", type=int, help="models of the case (default: the workload's full size)")
    ap.add_argument("--accuracy", choices=["auto", "high", "balanced", "fast"], default="high")
    ap.add_argument("--c-tol-rel", type=float, default=1e-12)
    ap.add_argument("--group-velocity", action="store_true")
    ap.add_argument("--warmups", type=int, default=3)
    ap.add_argument("--rounds", type=int, default=12)
    ap.add_argument("--workspace-bytes", type=int, default=8 << 30)
    ap.add_argument("--monitor-interval", type=float, default=.5)
    ap.add_argument("--output", type=Path, required=True)
    args = ap.parse_args(argv)
    from .api import Solver
    wl, periods, wave, modes = case_inputs(args.case, args.models)
    options = dict(accuracy=args.accuracy, c_tol_rel=args.c_tol_rel, search_strategy="shape_adaptive",
                   refinement="brent", love_strategy="phase", close_root_rtol=.005, f_accept=1.,
                   group_velocity=args.group_velocity, workspace_bytes=args.workspace_bytes)
    solvers = {p: Solver(**options, **PATHS[p]) for p in PATHS}
    model = dict(thickness=wl.thickness, vp=wl.vp, vs=wl.vs, rho=wl.rho)
    calls = {p: (lambda s=s: s(**model, periods=periods, nmodes=modes, wave=wave)) for p, s in solvers.items()}
    started = time.time()
    try:
        with Monitor(args.monitor_interval) as mon:
            records = protocol(calls, args.warmups, args.rounds)
    finally:
        for s in solvers.values():
            s.close()
    r64, r32 = records["R64"]["summary"], records["R32"]["summary"]
    doc = dict(case=args.case, models=int(wl.B), periods=int(np.size(periods)), wave=wave, nmodes=modes,
               options=options, paths=PATHS, orders=ORDERS, started_unix=started, finished_unix=time.time(),
               records=records, monitor=mon.summary(), environment=environment(),
               ratio_r32_over_r64=(r32["valid_roots_per_s"] / r64["valid_roots_per_s"]
                                   if r32["valid_roots_per_s"] and r64["valid_roots_per_s"] else None))
    args.output.parent.mkdir(parents=True, exist_ok=True)
    args.output.write_text(json.dumps(doc, indent=1, allow_nan=False) + "\n")
    print(json.dumps({p: records[p]["summary"] for p in records}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
