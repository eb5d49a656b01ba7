"""Benchmark of the modisp-rocm hot path on one MI355X.

Workload: the precision batch -- 16384 seeded monotone models of 12 layers
plus half-space, 100 log-spaced periods 0.5-60 s, Rayleigh and Love, 5 modes
each, phase window [0.8 min Vs, Vs_halfspace), FP64 refinement to 1e-12
relative, analytic FP64 group velocity.

A step is one pass of the multimode root pipeline over the whole batch for
both wave types (scan -> FP64 recheck / pair probes -> brackets -> FP64
refinement -> group velocity). Inputs are resident in HBM before the timed
region and outputs stay there. `value` is valid roots (finite phase cells)
per second of the FP32-scan path. The same steps are timed for the plain
FP64 path (identical grid and pipeline, every scan evaluation in FP64) and,
for reference, for the literal 512-sample uniform grid in both precisions.

`graph` is the same device-resident step replayed from one captured hipGraph
per wave type. HIP runs the replays of both solvers' graphs on one internal
queue, so the two wave types' replays run one after the other where their eager
launches overlap on two streams; the leg therefore also reports each wave
type's replay device time beside its eager pass. `end_to_end` follows the reference's own timing protocol:
Solver calls on host arrays (copies and host result assembly included), median
of the steady calls, valid roots = status-0 cells.

`fwd_adjoint` is the second half of the metric: models/s of forward root
solve + misfit loss + adjoint backward on the differentiation workload
(4096 models x 21 rows, 50 periods, Rayleigh, 3 modes) through the
torch.autograd binding.

`uneven` times the ragged uneven batch (65536 models, per-model rows,
periods and modes) resident on the device: eagerly with the default launch
shape, eagerly with the autotuned shape, and replayed from one captured
hipGraph (`value` of the leg); anchor continuation is timed beside it with its
acceptance, fallback reasons and FP64 evaluations per target.

`sweep` is the inversion-scale batch (262144 models x 31 rows, 200 periods,
Rayleigh and Love, 4 modes, group velocity): forward roots/s, forward +
adjoint gradient models/s and the device memory in use.

`roofline` is for the dominant kernel, the FP32 Rayleigh grid scan: its
algorithmic FLOP (secular evaluations x layer steps x 160 FLOP per Dunkin
layer step, counted on the device) divided by its HIP-event time on the
solver's stream. `cpu_baseline` is the oracle (C++17 FP64 restatement of the
reference CPU backend) on the host's physical cores (capped by the job's CPU
share): a 32-model sample of the same workload per call, timed by the
reference's backend protocol (1 first + 3 warm-up + 12 steady calls, median
and quartiles), and the plain FP64 configuration (1024 models x 7 rows) in
full beside the device's plain-FP64 and FP32-scan paths on it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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

METRIC = "roots/s at ≤1e-9 rel err: FP32-scan+FP64 vs plain FP64; fwd+adjoint models/s"
FLOP_PER_LAYER = {0: 160.0, 1: 20.0}   # algorithmic FLOP per layer step: Dunkin (Rayleigh) / SH (Love)
PEAK_FP32 = 157.3                      # MI355X published FP32 vector peak, TFLOP/s (packed dual issue)
PEAK_FP32_UNPACKED = 78.6              # the same without packed FP32 (not used by these kernels)
TRAFFIC_FILE = ROOT / "profiles" / "scan_kernel_precision_traffic.json"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class DeviceCase:
    """One wave type of the workload resident on the device, with its solver."""

    def __init__(self, modisp_amd, torch, wl, window, wave, scan, grid="adaptive", group=True, graph=False):
        dev = torch.device("cuda", 0)
        self.B, self.L = wl.thickness.shape
        self.T = wl.periods.size
        self.M = int(wl.nmodes)
        self.wave = wave
        t = lambda a: torch.from_numpy(a).to(dev).contiguous()
        self.h, self.vp, self.vs, self.rho = t(wl.thickness), t(wl.vp), t(wl.vs), t(wl.rho)
        self.periods, self.bounds = t(wl.periods), t(window)
        n = self.B * self.M * self.T
        self.velocity = torch.empty(n, dtype=torch.float64, device=dev)
        self.flags = torch.empty(self.B * self.T, dtype=torch.int32, device=dev)
        self.levels = torch.empty(self.B * self.T, dtype=torch.int32, device=dev)
        self.counts = torch.empty(self.B * self.T, dtype=torch.int32, device=dev)
        self.group = torch.empty(n, dtype=torch.float64, device=dev) if group else None
        self.gflags = torch.empty(n, dtype=torch.int32, device=dev) if group else None
        self.solver = modisp_amd.QeSolver(self.B, self.L, self.T, self.M, group_velocity=group, scan=scan, grid=grid,
                                          nc=512, graph=graph)
        self.ptrs = dict(thickness=self.h.data_ptr(), vp=self.vp.data_ptr(), vs=self.vs.data_ptr(),
                         rho=self.rho.data_ptr(), periods=self.periods.data_ptr(), bounds=self.bounds.data_ptr(),
                         velocity=self.velocity.data_ptr(), flags=self.flags.data_ptr(),
                         levels=self.levels.data_ptr(), counts=self.counts.data_ptr(),
                         group=self.group.data_ptr() if group else None,
                         group_flags=self.gflags.data_ptr() if group else None)

    def launch(self):
        self.solver.run_device_batch(self.B, self.ptrs, 1e-12, self.wave)

    def stages(self):
        """Per-stage device milliseconds of the last run (synchronizes its stream)."""
        self.solver.synchronize()
        return self.solver.last_timing()

    def valid_roots(self):
        return int(self.velocity.isfinite().sum().item())

    def close(self):
        self.solver.close()


def timed_steps(torch, cases, steps, warmup):
    """W untimed steps, then exactly K timed steps bracketed by device syncs.
    Returns wall seconds and the mean per-step stage times of every case."""
    for _ in range(warmup):
        for c in cases:
            c.launch()
        for c in cases:
            c.stages()
    torch.cuda.synchronize()
    stage = [None] * len(cases)
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in cases:
            c.launch()
        for i, c in enumerate(cases):
            ms = c.stages()
            stage[i] = ms if stage[i] is None else stage[i] + ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [s / steps for s in stage]


def cpu_threads():
    """Threads for the CPU baseline: the physical cores among the CPUs this
    process may run on, capped by the job's CPU share where the launcher
    states one (OMP_NUM_THREADS). Returns (threads, physical, logical, share)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            cores.add((Path(base + "physical_package_id").read_text().strip(),
                       Path(base + "core_id").read_text().strip()))
        except OSError:
            cores.add(("cpu", c))
    share = os.environ.get("OMP_NUM_THREADS")
    share = int(share) if share and share.isdigit() and int(share) > 0 else None
    threads = min(len(cores), share) if share else len(cores)
    return max(1, threads), len(cores), len(cpus), share


def timing_protocol(call, warmups=3, rounds=12):
    """One first call, `warmups` warm-up calls and `rounds` steady calls of
    call() -> roots; median and quartiles of the steady calls (the
    reference's backend-benchmark protocol, docs/BACKEND_BENCHMARK_0_4_0.md:19)."""
    import numpy as np
    samples = {"first": None, "warmup": [], "steady": []}
    roots = 0
    for phase, n in (("first", 1), ("warmup", warmups), ("steady", rounds)):
        for _ in range(n):
            t0 = time.perf_counter()
            roots = call()
            dt = time.perf_counter() - t0
            if phase == "first":
                samples["first"] = dt
            else:
                samples[phase].append(dt)
    t = np.array(samples["steady"])
    med = float(np.median(t))
    return {"roots_per_call": int(roots), "median_s": med, "q25_s": float(np.quantile(t, .25)),
            "q75_s": float(np.quantile(t, .75)), "value": roots / med, "samples_s": samples}


def cpu_baseline(wl, window, gpu_cases, sample, warmups, rounds, plain_gpu, plain_rounds):
    """The oracle (C++17 FP64 restatement of the reference CPU backend:
    shape-free qe_adaptive ladder, Brent to 1e-12, Love phase counting,
    analytic group velocity) on the host cores:

    * precision sample: the first `sample` models of the workload, both waves,
      all periods, timed by the protocol (1 + warmups + rounds calls, median);
      the device roots of those models are checked against it;
    * cfg0 in full: the plain FP64 batch (1024 models x 7 rows, 60 periods,
      Rayleigh, 3 modes, phase only), 1 + 0 + plain_rounds calls, beside the
      device plain-FP64 and FP32-scan paths on the same batch."""
    import numpy as np
    import oracle
    oracle.build(force=False)
    threads, physical, logical, share = cpu_threads()
    sub = wl.subset(np.arange(sample))
    refs = {}

    def precision_call():
        n = 0
        for case in gpu_cases:
            ref = oracle.solve(sub.thickness, sub.vp, sub.vs, sub.rho, sub.periods, window[:sample], case.M,
                               case.wave, tol=1e-12, group=2, threads=threads)
            refs[case.wave] = ref
            n += int(np.isfinite(ref["phase_velocity"]).sum())
        return n

    prec = timing_protocol(precision_call, warmups, rounds)
    agree = {"count_match": 1.0, "mask_match": 1.0, "max_rel_phase": 0.0, "max_rel_group": 0.0}
    for case in gpu_cases:
        ref = refs[case.wave]
        rv = ref["phase_velocity"]
        B, M, T = case.B, case.M, case.T
        gv = case.velocity.view(B, M, T)[:sample].cpu().numpy()
        gc = case.counts.view(B, T)[:sample].cpu().numpy()
        gg = case.group.view(B, M, T)[:sample].cpu().numpy()
        resolved = ref["root_count"] >= 0
        agree["count_match"] = min(agree["count_match"], float(np.mean((gc == ref["root_count"])[resolved])))
        agree["mask_match"] = min(agree["mask_match"], float(np.mean(np.isfinite(gv) == np.isfinite(rv))))
        both = np.isfinite(gv) & np.isfinite(rv)
        if both.any():
            agree["max_rel_phase"] = max(agree["max_rel_phase"], float(np.max(np.abs(gv[both] - rv[both]) / rv[both])))
        rg = ref["group_velocity"]
        gb = np.isfinite(gg) & np.isfinite(rg)
        if gb.any():
            agree["max_rel_group"] = max(agree["max_rel_group"],
                                         float(np.max(np.abs(gg[gb] - rg[gb]) / np.abs(rg[gb]))))
    pw, pwin = plain_gpu["workload"], plain_gpu["window"]
    plain_ref = {}

    def plain_call():
        r = oracle.solve(pw.thickness, pw.vp, pw.vs, pw.rho, pw.periods, pwin, 3, 0, tol=1e-12, group=0,
                         threads=threads)
        plain_ref["r"] = r
        return int(np.isfinite(r["phase_velocity"]).sum())

    plain = timing_protocol(plain_call, 0, plain_rounds)
    r = plain_ref["r"]
    gv = plain_gpu["velocity"]
    plain["gpu_count_match"] = float(np.mean(plain_gpu["counts"] == r["root_count"]))
    plain["gpu_mask_match"] = float(np.mean(np.isfinite(gv) == np.isfinite(r["phase_velocity"])))
    both = np.isfinite(gv) & np.isfinite(r["phase_velocity"])
    plain["gpu_max_rel_phase"] = float(np.max(np.abs(gv[both] - r["phase_velocity"][both]) / r["phase_velocity"][both]))
    plain["gpu"] = plain_gpu["rates"]
    where = f"{threads} threads ({physical} physical cores in the affinity mask of {logical} CPUs" + (
        f", job CPU share {share})" if share else ")")
    return {"value": prec["value"], "unit": "roots/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} of {wl.B} models of the workload, Rayleigh+Love, {gpu_cases[0].M} modes, "
                      f"{gpu_cases[0].T} periods, analytic group velocity, {prec['roots_per_call']} roots per call; "
                      f"median of {rounds} steady calls after 1 first + {warmups} warm-up calls on {where}",
            "protocol": prec, "physical_cores": physical, "logical_cpus": logical, "cpu_share": share,
            "gpu_parity_on_sample": agree,
            "plain_fp64_full": dict(plain, workload="plain: 1024 models x 7 rows, 60 periods 1-50 s, Rayleigh, 3 "
                                                    "modes, phase only (the CPU-runnable configuration, in full)")}


def plain_gpu_leg(torch, modisp_amd, workloads, steps, warmup):
    """The plain FP64 batch on the device: plain-FP64 and FP32-scan paths
    (device-resident roots/s), and the FP64 path's results for the CPU check."""
    wl = workloads.plain_batch()
    window = workloads.scan_window(wl.vs)
    rates = {}
    keep = None
    for scan in ("float64", "float32"):
        case = DeviceCase(modisp_amd, torch, wl, window, 0, scan, group=False)
        wall, _ = timed_steps(torch, [case], steps, warmup)
        roots = case.valid_roots()
        rates[scan] = {"value": roots * steps / wall, "unit": "roots/s", "ms_per_step": wall / steps * 1e3,
                       "valid_roots": roots}
        if scan == "float64":
            keep = (case.velocity.view(case.B, case.M, case.T).cpu().numpy(),
                    case.counts.view(case.B, case.T).cpu().numpy())
        case.close()
    return {"workload": wl, "window": window, "velocity": keep[0], "counts": keep[1], "rates": rates}


def end_to_end_leg(wl, steps):
    """The reference's own timing protocol (tools/benchmark_backends.py:140-200,
    556-559) on the same workload: Solver(precision="float32", ...) called
    with host arrays per wave type, one first call, then `steps` steady calls;
    valid roots are status-0 cells and roots/s = the two waves' valid roots /
    the sum of their median call times. Includes the host-to-device copies of
    the models, the device-to-host copies of every result array and the host
    result assembly (status, masks, flags), so it is the PCIe-inclusive rate,
    not `value`. The tolerance is the device leg's (1e-12 relative) and the
    window the front end's own (reference bounds)."""
    import numpy as np
    from modisp_amd import api
    opts = dict(backend="rocm", precision="float32", secular="fast", accuracy="high", c_tol_rel=1e-12,
                search_strategy="shape_adaptive", refinement="brent", love_strategy="phase", close_root_rtol=.005,
                f_accept=1., group_velocity=True, workspace_bytes=8 << 30)
    out, roots, seconds = {}, 0, 0.0
    for wave in ("rayleigh", "love"):
        with api.Solver(**opts) as solver:
            call = lambda: solver(wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods, nmodes=int(wl.nmodes), wave=wave)
            t0 = time.perf_counter()
            res = call()
            first = time.perf_counter() - t0
            times = []
            for _ in range(steps):
                t0 = time.perf_counter()
                res = call()
                times.append(time.perf_counter() - t0)
            valid = int(np.count_nonzero(res.status == 0))
            med = float(np.median(times))
            out[wave] = {"valid_roots": valid, "median_s": med, "first_s": first,
                         "iqr_s": float(np.subtract(*np.percentile(times, [75, 25])))}
            roots += valid
            seconds += med
    out.update({"value": roots / seconds, "unit": "roots/s", "steady_calls": steps, "options": opts,
                "note": "workspace_bytes holds the whole batch in one device chunk"})
    return out


def gradient_leg(torch, modisp_amd, workloads, steps, warmup):
    """Forward + adjoint on the differentiation workload (4096 models of 20
    layers + half-space, 50 periods 2-40 s, Rayleigh, 3 modes): a step is the
    device root solve, the misfit loss sum (c - c_target)^2 over valid cells
    and its backward pass (the adjoint kernel), targets precomputed."""
    from modisp_amd.autograd import phase_velocity
    wl = workloads.gradient_batch()
    dev = torch.device("cuda", 0)
    B, L = wl.thickness.shape
    window = workloads.scan_window(wl.vs)
    solver = modisp_amd.QeSolver(B, L, wl.periods.size, 3, scan="float32")
    per = torch.tensor(wl.periods, device=dev)
    win = torch.tensor(window, device=dev)
    with torch.no_grad():
        tgt = phase_velocity(*(torch.tensor(x, device=dev) for x in (wl.thickness, wl.vp, wl.extra["target_vs"], wl.rho)),
                             per, win, nmodes=3)
    params = [torch.tensor(getattr(wl, k), device=dev, requires_grad=True) for k in ("thickness", "vp", "vs", "rho")]

    def step():
        for p in params:
            p.grad = None
        c = phase_velocity(*params, per, win, nmodes=3, solver=solver)
        ok = torch.isfinite(c) & torch.isfinite(tgt)
        loss = ((torch.where(ok, c, 0.) - torch.where(ok, tgt, 0.)) ** 2).sum()
        loss.backward()
        return c

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        c = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    roots = int(torch.isfinite(c).sum().item())
    # adjoint alone: the last backward's share, timed with events on torch's stream
    c = phase_velocity(*params, per, win, nmodes=3, solver=solver)
    ok = torch.isfinite(c) & torch.isfinite(tgt)
    loss = ((torch.where(ok, c, 0.) - torch.where(ok, tgt, 0.)) ** 2).sum()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    loss.backward()
    e1.record()
    torch.cuda.synchronize()
    solver.close()
    return {"value": B * steps / wall, "unit": "models/s", "ms_per_step": wall / steps * 1e3,
            "adjoint_ms": e0.elapsed_time(e1), "roots_per_step": roots,
            "config": {"workload": "differentiation: 4096 models x 21 rows, 50 periods 2-40 s, Rayleigh, 3 modes, "
                                   "loss sum (c - c_target)^2, targets from Vs + N(0, 0.05)",
                       "parameters": "h, vp, vs, rho of every row", "dtype": "f32 scan, f64 refine and adjoint"}}


def uneven_leg(torch, modisp_amd, workloads, models, steps):
    """The uneven batch (65536 models of 4..41 rows, 20..200 periods 0.2-100 s,
    1..10 Rayleigh modes each) resident on the device as one ragged batch
    (md_rocm_run_ragged_device: preparation, scan, rechecks and refinement on
    the device). Timed eagerly with the default launch shape, then the
    autotuner picks the launch shape, then the whole sequence is captured in
    one hipGraph and replayed: HIP-event milliseconds per call (median of
    `steps`) and wall milliseconds per call for each. Anchor continuation runs
    through the host ragged entry (its anchors are planned from host periods):
    device milliseconds and acceptance."""
    import numpy as np
    wl = workloads.uneven_batch(B=models)
    packed = workloads.pack_ragged(wl)
    geom, cells = modisp_amd.ragged_geometry(packed["row_offsets"], packed["period_offsets"], packed["modes"])
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    bufs = {k: t(packed[k]) for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets",
                                      "modes", "bounds")}
    bufs["cell_offsets"] = t(cells)
    bufs.update(velocity=torch.empty(geom["ncells"], dtype=torch.float64, device=dev),
                flags=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev),
                levels=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev),
                counts=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev))
    ptr = {k: v.data_ptr() for k, v in bufs.items()}

    def timed(solver):
        ev, wall = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            solver.run_ragged_device_batch(geom, ptr, 1e-12, 0)
            solver.synchronize()
            wall.append(time.perf_counter() - t0)
            ev.append(float(solver.last_timing()[0]))
        return {"device_ms": float(np.median(ev)), "wall_ms": float(np.median(wall)) * 1e3,
                "device_ms_samples": ev}

    out = {}
    B, R = geom["B"], geom["max_rows"]
    with modisp_amd.QeSolver(B, R, 1, 1, scan="float32") as s:
        s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
        out["eager_default"] = timed(s)
        out["eager_default"]["tiles"] = {"scan_waves": 4, "consumer_blocks": 8}
        w, b = s.autotune_ragged_device_batch(geom, ptr, 1e-12, 0)
        out["eager_tuned"] = timed(s)
        out["eager_tuned"]["tiles"] = {"scan_waves": w, "consumer_blocks": b}
        ref = bufs["velocity"].cpu().numpy().copy()
    with modisp_amd.QeSolver(B, R, 1, 1, scan="float32", graph=True, scan_waves=w, consumer_blocks=b) as g:
        g.run_ragged_device_batch(geom, ptr, 1e-12, 0)      # eager run, then capture
        g.synchronize()
        out["graph_replay"] = timed(g)
        out["graph_replay"]["tiles"] = {"scan_waves": w, "consumer_blocks": b}
        got = bufs["velocity"].cpu().numpy()
        out["graph_replay"]["equals_eager"] = bool(np.array_equal(got.view(np.uint64), ref.view(np.uint64)))
    roots = int(np.isfinite(ref).sum())
    out["roots"] = roots
    out["value"] = roots / (out["graph_replay"]["device_ms"] * 1e-3)
    out["unit"] = "roots/s"
    out["graph_over_eager"] = out["eager_tuned"]["device_ms"] / out["graph_replay"]["device_ms"]
    args = [packed[k] for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets", "modes",
                                "bounds")]
    with modisp_amd.QeSolver(B, R, 1, 1, continuation=True, scan="float32") as s:
        s.run_ragged(*args, 1e-12, 0)      # warm-up
        t0 = time.perf_counter()
        r = s.run_ragged(*args, 1e-12, 0)
        wall = time.perf_counter() - t0
        pred = r["prediction_flags"]
        st = s.last_stats()
        fell = (pred & 4) != 0
        reasons = {name: int(np.count_nonzero(fell & ((pred & bit) != 0)))
                   for name, bit in (("source_uncertain", 64), ("root_count", 128), ("anchor_unsafe", 256),
                                     ("prediction_window", 512), ("bracket", 1024), ("correction", 2048),
                                     ("audit", 32))}
        out["continuation"] = {"device_ms": float(s.last_timing()[0]), "wall_ms": wall * 1e3,
                               "fp64_evaluations": st["continuation_evaluations"],
                               "fp64_evaluations_per_target": st["continuation_evaluations"] / max(1, st[
                                   "continuation_targets"]),
                               "fallback_reasons": reasons,
                               "roots": int(np.isfinite(r["velocity"]).sum()),
                               "targets": int(np.count_nonzero((pred & 1) == 0)),
                               "accepted": int(np.count_nonzero(pred & 2)),
                               "fallbacks": int(np.count_nonzero(pred & 4)),
                               "entry_point": "md_rocm_run_ragged (host arrays; PCIe copies inside the events)"}
    out["config"] = {"workload": f"uneven: {B} models, {geom['ntraces']} traces, rows 4..41, periods 20..200, "
                                 "modes 1..10, Rayleigh, 20% low-velocity zones",
                     "entry_point": "md_rocm_run_ragged_device + md_rocm_autotune_ragged_device, hipGraph replay"}
    return out


def uneven_profile(torch, modisp_amd, workloads, models, steps):
    """Profiling mode: the uneven ragged batch resident on the device, one
    warm-up run and `steps` eager runs with the default launch shape; per-run
    stage milliseconds and the work counters of the last run."""
    import numpy as np
    wl = workloads.uneven_batch(B=models)
    packed = workloads.pack_ragged(wl)
    geom, cells = modisp_amd.ragged_geometry(packed["row_offsets"], packed["period_offsets"], packed["modes"])
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    bufs = {k: t(packed[k]) for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets",
                                      "modes", "bounds")}
    bufs["cell_offsets"] = t(cells)
    bufs.update(velocity=torch.empty(geom["ncells"], dtype=torch.float64, device=dev),
                flags=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev),
                levels=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev),
                counts=torch.empty(geom["ntraces"], dtype=torch.int32, device=dev))
    ptr = {k: v.data_ptr() for k, v in bufs.items()}
    runs = []
    with modisp_amd.QeSolver(geom["B"], geom["max_rows"], 1, 1, scan="float32") as s:
        s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
        s.synchronize()
        for _ in range(steps):
            s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
            s.synchronize()
            runs.append(dict(zip(s.STAGES, map(float, s.last_timing()))))
        stats = s.last_stats()
    return {"workload": f"uneven: {geom['B']} models, {geom['ntraces']} traces, Rayleigh, FP32 scan",
            "runs": runs, "work": stats, "roots": int(bufs["velocity"].isfinite().sum().item())}


def sweep_leg(torch, modisp_amd, workloads, steps):
    """The inversion-scale sweep (262144 models of 30 layers + half-space,
    200 periods 0.5-100 s, Rayleigh and Love, 4 modes, group velocity):
    forward roots/s of the FP32-scan path (both waves; timed like the main
    steps), then forward + adjoint of the misfit sum (c - c_target)^2 over both
    waves' phase velocities through torch.autograd (gradient models/s; targets
    from Vs + N(0, 0.05)), and the device memory in use after each part (this
    package's buffers persist in their solvers, so that is the peak)."""
    import numpy as np
    from modisp_amd.autograd import phase_velocity
    wl = workloads.sweep_batch()
    window = workloads.scan_window(wl.vs)
    B = wl.B
    free0, total = modisp_amd.memory_info()
    cases = [DeviceCase(modisp_amd, torch, wl, window, w, "float32") for w in (0, 1)]
    wall, _ = timed_steps(torch, cases, steps, 1)
    roots = sum(c.valid_roots() for c in cases)
    stats = [c.solver.last_stats() for c in cases]
    fwd_used = total - modisp_amd.memory_info()[0]
    for c in cases:
        c.close()
    del cases
    torch.cuda.empty_cache()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    target_vs = np.maximum(wl.vs + rng.normal(0., 0.05, wl.vs.shape), 0.2)
    per = torch.tensor(wl.periods, device=dev)
    win = torch.tensor(window, device=dev)
    L = wl.thickness.shape[1]
    solvers = [modisp_amd.QeSolver(B, L, wl.periods.size, 4, scan="float32") for _ in (0, 1)]
    with torch.no_grad():
        tm = [torch.tensor(x, device=dev) for x in (wl.thickness, wl.vp, target_vs, wl.rho)]
        tgt = [phase_velocity(*tm, per, win, nmodes=4, wave=w, solver=solvers[w]) for w in (0, 1)]
    params = [torch.tensor(getattr(wl, k), device=dev, requires_grad=True) for k in ("thickness", "vp", "vs", "rho")]

    def step():
        for p_ in params:
            p_.grad = None
        loss = 0.
        for w in (0, 1):
            c = phase_velocity(*params, per, win, nmodes=4, wave=w, solver=solvers[w])
            ok = torch.isfinite(c) & torch.isfinite(tgt[w])
            loss = loss + ((torch.where(ok, c, 0.) - torch.where(ok, tgt[w], 0.)) ** 2).sum()
        loss.backward()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    gwall = time.perf_counter() - t0
    grad_used = total - modisp_amd.memory_info()[0]
    finite = all(bool(torch.isfinite(p_.grad).all().item()) for p_ in params)
    for s_ in solvers:
        s_.close()
    return {"value": roots * steps / wall, "unit": "roots/s", "ms_per_step": wall / steps * 1e3, "roots_per_step": roots,
            "work": {"rayleigh": stats[0], "love": stats[1]},
            "fwd_adjoint": {"value": B * steps / gwall, "unit": "models/s", "ms_per_step": gwall / steps * 1e3,
                            "gradients_finite": finite},
            "hbm_in_use_bytes": {"forward": int(fwd_used), "forward_adjoint": int(grad_used),
                                 "before": int(total - free0), "total": int(total)},
            "config": {"workload": "sweep: 262144 models x 31 rows (Vs random walk), 200 periods 0.5-100 s, "
                                   "Rayleigh+Love, 4 modes, group velocity; gradient w.r.t. h, vp, vs, rho",
                       "steps": steps}}


def scan_traffic():
    """HBM bytes of the Rayleigh scan from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by a copy-kernel calibration), if present:
    (bytes per run -- one pass over the batch, two trace-chunk dispatches --,
    mean bytes per dispatch)."""
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
        return (float(d["fetch_bytes_per_run"] + d["write_bytes_per_run"]),
                float(d["fetch_bytes_per_dispatch"] + d["write_bytes_per_dispatch"]))
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--models", type=int, default=16384)
    ap.add_argument("--cpu-sample", type=int, default=32, help="models per call of the CPU baseline protocol")
    ap.add_argument("--cpu-plain-rounds", type=int, default=2,
                    help="steady calls of the CPU baseline on the full plain FP64 batch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-uniform", action="store_true", help="skip the 512-sample uniform-grid comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the ragged uneven-batch leg")
    ap.add_argument("--uneven-models", type=int, default=65536)
    ap.add_argument("--no-sweep", action="store_true", help="skip the inversion-scale sweep leg")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replay leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end Solver leg (host arrays)")
    ap.add_argument("--profile-uneven", action="store_true",
                    help="profiling mode: only `steps` eager runs of the uneven ragged batch resident on the device "
                         "(stage times and work counters as one JSON line), for a kernel trace of that workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: every trace is independent, "
              f"so the batch would be split by model ranges with no exchange, and this build runs on one GPU",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import modisp_amd
    from modisp_amd import workloads

    if not torch.cuda.is_available() or not modisp_amd.available(0):
        print("no gfx950 GPU visible", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    modisp_amd.set_device(0)
    if args.profile_uneven:
        print(json.dumps(uneven_profile(torch, modisp_amd, workloads, args.uneven_models, args.steps)))
        return
    wl = workloads.precision_batch(B=args.models)
    window = workloads.scan_window(wl.vs)
    L = wl.thickness.shape[1]
    log(f"workload {wl.name}: B={wl.B} L={L} T={wl.periods.size} modes={wl.nmodes}")

    def run_path(scan, grid, graph=False):
        cases = [DeviceCase(modisp_amd, torch, wl, window, w, scan, grid, graph=graph) for w in (0, 1)]
        wall, _ = timed_steps(torch, cases, args.steps, args.warmup)
        roots = sum(c.valid_roots() for c in cases)
        # Stage times and work counters from one more pass of each wave type on
        # its own (in the timed steps the two waves' streams overlap).
        stages, stats = [], []
        for c in cases:
            c.launch()
            stages.append(c.stages())
            stats.append(c.solver.last_stats())
        log(f"{scan}/{grid}: {wall / args.steps * 1e3:.2f} ms/step, {roots} roots; stages R "
            + " ".join(f"{n}={v:.2f}" for n, v in zip(cases[0].solver.STAGES, stages[0]))
            + " | L " + " ".join(f"{n}={v:.2f}" for n, v in zip(cases[1].solver.STAGES, stages[1])))
        return cases, wall, stages, roots, stats

    fast, wall32, st32, roots32, stats32 = run_path("float32", "adaptive")
    plain, wall64, st64, roots64, stats64 = run_path("float64", "adaptive")
    for c in plain:
        c.close()
    value = roots32 * args.steps / wall32
    value64 = roots64 * args.steps / wall64
    uniform = {}
    if not args.no_uniform:
        for scan in ("float32", "float64"):
            cases, wall, stages, roots, _ = run_path(scan, "uniform")
            uniform[scan] = {"value": roots * args.steps / wall, "unit": "roots/s",
                             "ms_per_step": wall / args.steps * 1e3}
            for c in cases:
                c.close()

    # Roofline of the dominant kernel: the FP32 Rayleigh scan.
    evals = stats32[0]["scan_nodes"]
    flop = evals * (L - 1) * FLOP_PER_LAYER[0]
    scan_ms = float(st32[0][1])
    achieved = flop / (scan_ms * 1e-3) / 1e12
    love_flop = stats32[1]["scan_nodes"] * (L - 1) * FLOP_PER_LAYER[1]
    traffic_run, traffic_dispatch = scan_traffic()
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32, "traffic": traffic_run,
                "traffic_scope": "HBM bytes per run of the kernel over the batch (the same scope as "
                                 "algorithmic_flop_per_run and kernel_ms_per_run)",
                "traffic_per_dispatch": traffic_dispatch,
                "kernel": "scan_kernel<float,64,4> (Rayleigh)",
                "algorithmic_flop_per_run": flop, "secular_evaluations_per_run": evals,
                "flop_per_layer_step": FLOP_PER_LAYER[0], "layer_steps_per_evaluation": L - 1,
                "kernel_ms_per_run": scan_ms, "launches_per_run": 2,
                "peak_without_packed_fp32": PEAK_FP32_UNPACKED,
                "frac_of_unpacked_peak": achieved / PEAK_FP32_UNPACKED,
                "love_scan_tflops": love_flop / (float(st32[1][1]) * 1e-3) / 1e12}
    result = {
        "metric": METRIC, "value": value, "unit": "roots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall32 / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 scan, f64 recheck/refine/group",
        "data": "synthetic (seeded monotone layered models, precision workload recipe)",
        "config": {"workload": "precision: 16384 models x 13 rows, 100 periods 0.5-60 s, Rayleigh+Love, 5 modes, "
                               "group velocity", "models": wl.B, "rows": int(L), "periods": int(wl.periods.size),
                   "waves": "rayleigh+love", "modes": int(wl.nmodes), "grid": "adaptive (8 nodes per WKB mode)",
                   "refine_tol_rel": 1e-12, "group_velocity": "analytic fp64"},
        "fp64_path": {"value": value64, "unit": "roots/s", "ms_per_step": wall64 / args.steps * 1e3,
                      "speedup_fp32_over_fp64": value / value64},
        "uniform512": {k: v for k, v in uniform.items()},
        "stages_ms": {"fp32_rayleigh": dict(zip(fast[0].solver.STAGES, map(float, st32[0]))),
                      "fp32_love": dict(zip(fast[1].solver.STAGES, map(float, st32[1]))),
                      "fp64_rayleigh": dict(zip(fast[0].solver.STAGES, map(float, st64[0]))),
                      "fp64_love": dict(zip(fast[1].solver.STAGES, map(float, st64[1])))},
        "work": {"fp32_rayleigh": stats32[0], "fp32_love": stats32[1]},
        "roofline": roofline,
    }
    if uniform:
        result["uniform512"]["speedup_fp32_over_fp64"] = uniform["float32"]["value"] / uniform["float64"]["value"]
    if not args.no_graph:
        # The same device-resident step replayed from one captured hipGraph per
        # wave type (captured on the first warmup call).
        cases, wall, gst, roots, _ = run_path("float32", "adaptive", graph=True)
        for c in cases:
            c.close()
        result["graph"] = {"value": roots * args.steps / wall, "unit": "roots/s",
                           "ms_per_step": wall / args.steps * 1e3, "valid_roots": roots,
                           "replay_device_ms": {"rayleigh": float(gst[0][0]), "love": float(gst[1][0])},
                           "eager_device_ms": {"rayleigh": float(st32[0][0]), "love": float(st32[1][0])},
                           "note": "both solvers' replays share one HIP-internal queue (no overlap of the wave "
                                   "types); per wave type, replay device time vs the eager pass"}
    if not args.no_e2e:
        result["end_to_end"] = end_to_end_leg(wl, args.steps)
        log(f"end to end: {result['end_to_end']['value']:.3g} roots/s (Solver calls on host arrays)")
    if not args.no_adjoint:
        result["fwd_adjoint"] = gradient_leg(torch, modisp_amd, workloads, args.steps, args.warmup)
        log(f"fwd+adjoint: {result['fwd_adjoint']['ms_per_step']:.2f} ms/step, "
            f"{result['fwd_adjoint']['value']:.0f} models/s")
    if not args.no_uneven:
        result["uneven"] = uneven_leg(torch, modisp_amd, workloads, args.uneven_models, args.steps)
        u = result["uneven"]
        log(f"uneven: eager {u['eager_default']['device_ms']:.1f} ms, tuned {u['eager_tuned']['device_ms']:.1f} ms "
            f"{u['eager_tuned']['tiles']}, graph replay {u['graph_replay']['device_ms']:.1f} ms; continuation "
            f"{u['continuation']['device_ms']:.1f} ms ({u['continuation']['accepted']} of "
            f"{u['continuation']['targets']} targets accepted)")
    if not args.no_sweep:
        result["sweep"] = sweep_leg(torch, modisp_amd, workloads, 2)
        sw = result["sweep"]
        log(f"sweep: {sw['ms_per_step']:.1f} ms/step, {sw['value']:.3g} roots/s; fwd+adjoint "
            f"{sw['fwd_adjoint']['ms_per_step']:.1f} ms/step, {sw['fwd_adjoint']['value']:.0f} models/s; "
            f"HBM in use {sw['hbm_in_use_bytes']['forward_adjoint'] / 2**30:.1f} GiB")
    if not args.no_cpu:
        plain_gpu = plain_gpu_leg(torch, modisp_amd, workloads, args.steps, args.warmup)
        result["cpu_baseline"] = cpu_baseline(wl, window, fast, min(args.cpu_sample, wl.B), 3, 12, plain_gpu,
                                              args.cpu_plain_rounds)
        cb = result["cpu_baseline"]
        log(f"cpu baseline: {cb['value']:.3g} roots/s on {cb['cores']} threads (median of 12); plain FP64 full "
            f"{cb['plain_fp64_full']['value']:.3g} roots/s")
    for c in fast:
        c.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
