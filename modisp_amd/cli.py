"""Command line of the ROCm backend: ``doctor``, ``forward`` and ``check-modes``.

The commands, options and the NPZ layout of ``forward``'s output are the
reference's (python/modisp/cli.py:9-117; NPZ keys and dtypes at 98-112), with
``--backend rocm`` (``auto`` resolves to it here). ``forward`` reads an NPZ
with ``thickness``, ``vp``, ``vs``, ``rho`` ([B, L] or [L]) and ``periods``
and exits 2 when any trace's search did not converge (TRACE_SEARCH_UNCONVERGED).

    python -m modisp_amd.cli doctor
    python -m modisp_amd.cli forward models.npz --output out.npz --wave love --group-velocity
    python -m modisp_amd.cli check-modes out.npz --output modes.npz
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import ABI_VERSION, SUPPORTS_AUTO_ACCURACY, SUPPORTS_FP32, VERSION, BackendUnavailableError
from . import available, device_name, memory_info, set_device
from .api import DEFAULT_ACCURACY, TRACE_SEARCH_UNCONVERGED, Solver
from .modes import check_mode_reliability

# Option table of `forward`: (flag, keyword arguments of add_argument).
_FORWARD = (
    ("input", {}),
    ("--output", dict(required=True)),
    ("--wave", dict(choices=["rayleigh", "love"], default="rayleigh")),
    ("--backend", dict(choices=["auto", "cpu", "cuda", "rocm"], default="auto")),
    ("--precision", dict(choices=["float64", "float32"], default="float64",
                         help="secular arithmetic precision of the grid scan (float32: FP32 scan, FP64 refinement)")),
    ("--nmodes", dict(type=int, default=3)),
    ("--threads", dict(type=int)),
    ("--device", dict(type=int, default=0)),
    ("--cmin", dict(type=float)),
    ("--cmax", dict(type=float)),
    ("--accuracy", dict(choices=["auto", "high", "balanced", "fast"], default=None,
                        help=f"search accuracy preset (default: {DEFAULT_ACCURACY})")),
    ("--search-strategy", dict(choices=["shape_adaptive", "qe_adaptive", "uniform"], default="shape_adaptive")),
    ("--sampling-rtol", dict(type=float)),
    ("--close-root-rtol", dict(type=float, default=.005)),
    ("--c-tol-rel", dict(type=float)),
    ("--secular", dict(choices=["dunkin", "fast"], default=None,
                       help="secular evaluator (this backend builds the contracted 'fast' form)")),
    ("--refinement", dict(choices=["legacy", "brent", "chandrupatla"], default=None)),
    ("--love-strategy", dict(choices=["phase", "scan"], default="phase")),
    ("--group-velocity", dict(action="store_true", help="compute FP64 implicit-derivative group velocities")),
    ("--group-rtol", dict(type=float, default=.001, help="group-velocity rounding-error tolerance")),
    ("--frequency-strategy", dict(choices=["independent", "anchor_continuation"], default="independent")),
    ("--anchor-stride", dict(type=int, default=4, help="maximum spacing of continuation anchor periods")),
    ("--no-continuity-check", dict(action="store_true",
                                   help="disable the bilateral phase-curve and FP32 duplicate checks")),
    ("--check-continuity", dict(action="store_true", help="save separate heuristic mode-reliability diagnostics")),
)
_CHECK = (
    ("input", {}),
    ("--output", dict(required=True)),
    ("--max-relative-step", dict(type=float, default=.1)),
    ("--prediction-rtol", dict(type=float, default=.05)),
    ("--max-log-period-gap", dict(type=float, default=.25)),
    ("--close-root-rtol", dict(type=float)),
)


def parser():
    p = argparse.ArgumentParser(prog="modisp-rocm")
    p.add_argument("--version", action="version", version=VERSION)
    sub = p.add_subparsers(dest="command", required=True)
    sub.add_parser("doctor", help="show the installed plugin and device status")
    for name, table, text in (("forward", _FORWARD, "compute phase velocities from NPZ model/period arrays"),
                              ("check-modes", _CHECK, "check an existing forward NPZ without solving again")):
        cmd = sub.add_parser(name, help=text)
        for flag, kw in table:
            cmd.add_argument(flag, **kw)
    return p


def doctor():
    """Plugin, defaults and device status as one JSON object."""
    from . import _native
    info = dict(version=VERSION, python=sys.executable, module=__package__, library=str(_native.LIBRARY),
                private_abi=ABI_VERSION, default_precision="float64", default_accuracy=DEFAULT_ACCURACY,
                accuracy_options=["auto", "high", "balanced", "fast"], default_refinement="brent",
                refinement_options=["legacy", "brent", "chandrupatla"], default_love_strategy="phase",
                default_search_strategy="shape_adaptive", default_secular_by_backend={"rocm": "fast"},
                secular_options=["fast"], default_continuity_check=True, default_group_velocity=False,
                default_group_rtol=.001, default_frequency_strategy="independent",
                frequency_strategies=["independent", "anchor_continuation"], default_anchor_stride=4,
                rocm_float32_available=False, rocm_auto_accuracy_available=bool(SUPPORTS_AUTO_ACCURACY))
    try:
        ok = bool(available(0))
        info.update(rocm_available=ok, auto_backend="rocm" if ok else "none")
        if ok:
            set_device(0)
            free, total = memory_info()
            info.update(rocm_device=device_name(0), free_device_bytes=int(free), total_device_bytes=int(total),
                        rocm_float32_available=bool(SUPPORTS_FP32))
        else:
            info["rocm_reason"] = "no gfx950 device is visible"
    except (BackendUnavailableError, RuntimeError) as exc:
        info.update(rocm_available=False, rocm_reason=str(exc), auto_backend="none")
    info["default_secular"] = info["default_secular_by_backend"].get(info["auto_backend"])
    return info


def _solve(args):
    with np.load(args.input, allow_pickle=False) as src:
        model = {k: src[k] for k in ("thickness", "vp", "vs", "rho")}
        periods = src["periods"]
    options = dict(backend=args.backend, threads=args.threads, device=args.device, cmin=args.cmin, cmax=args.cmax,
                   accuracy=args.accuracy, close_root_rtol=args.close_root_rtol, c_tol_rel=args.c_tol_rel,
                   search_strategy=args.search_strategy, sampling_rtol=args.sampling_rtol, secular=args.secular,
                   refinement=args.refinement, love_strategy=args.love_strategy, precision=args.precision,
                   group_velocity=args.group_velocity, group_rtol=args.group_rtol,
                   frequency_strategy=args.frequency_strategy, anchor_stride=args.anchor_stride,
                   continuity_check=not args.no_continuity_check)
    with Solver(**options) as solver:
        out = solver(**model, periods=periods, nmodes=args.nmodes, wave=args.wave)
        stats = solver.stats()
    return out, stats


def forward_arrays(out, stats, group_rtol, check_continuity=False):
    """The NPZ content of `forward` (reference key set and dtypes)."""
    text = lambda x: np.asarray("" if x is None else x)
    arrays = dict(refinement=text(out.refinement), love_strategy=text(out.love_strategy),
                  phase_velocity=out.phase_velocity, periods=out.period, mode_mask=out.mode_mask, status=out.status,
                  trace_flags=out.trace_flags, search_bounds=out.search_bounds, search_level=out.search_level,
                  root_count=out.root_count, backend=text(out.backend_used), wave=text(out.wave),
                  precision=text(out.precision), accuracy=text(out.accuracy), numerical_flags=out.numerical_flags,
                  approximate_search=out.approximate_search, close_root_mask=out.close_root_mask,
                  potentially_ambiguous=out.potentially_ambiguous, mode_index_uncertain=out.mode_index_uncertain,
                  close_root_risk=out.close_root_risk, close_root_rtol=np.asarray(out.close_root_rtol),
                  search_strategy=text(out.search_strategy),
                  sampling_rtol=np.asarray(np.nan if out.sampling_rtol is None else out.sampling_rtol),
                  secular=text(out.secular), group_rtol=np.asarray(group_rtol),
                  frequency_strategy=text(out.frequency_strategy),
                  frequency_strategy_used=text(out.frequency_strategy_used),
                  frequency_strategy_reason=text(out.frequency_strategy_reason),
                  anchor_stride=np.asarray(out.anchor_stride))
    if check_continuity:
        check = check_mode_reliability(out)
        arrays.update(mode_reliability_flags=check.flags, continuity_assessed=check.assessed,
                      continuity_previous_index=check.previous_index, continuity_sorted_order=check.sorted_order)
    if out.accuracy == "auto":
        arrays["auto_relaxed"] = out.auto_relaxed
        top = stats.get("sampling_rtol_max")
        arrays["sampling_rtol_max"] = np.asarray(np.nan if top is None else top)
    if out.continuity_flags is not None:
        arrays["continuity_flags"] = out.continuity_flags
    if out.group_velocity is not None:
        arrays.update(group_velocity=out.group_velocity, group_velocity_flags=out.group_velocity_flags,
                      group_velocity_mask=out.group_velocity_mask, group_velocity_valid=out.group_velocity_valid)
    if out.prediction_flags is not None:
        arrays["prediction_flags"] = out.prediction_flags
    return arrays


def check_modes(args):
    with np.load(args.input, allow_pickle=False) as z:
        close = (args.close_root_rtol if args.close_root_rtol is not None
                 else float(z["close_root_rtol"]) if "close_root_rtol" in z else .005)
        check = check_mode_reliability(z["phase_velocity"], z["periods"], trace_flags=z["trace_flags"],
                                       root_count=z["root_count"], close_root_rtol=close,
                                       max_relative_step=args.max_relative_step, prediction_rtol=args.prediction_rtol,
                                       max_log_period_gap=args.max_log_period_gap)
    np.savez_compressed(args.output, continuity_flags=check.flags, continuity_assessed=check.assessed,
                        continuity_previous_index=check.previous_index, continuity_sorted_order=check.sorted_order)
    return check.summary()


def main(argv=None):
    args = parser().parse_args(argv)
    if args.command == "doctor":
        print(json.dumps(doctor(), indent=2))
        return 0
    if args.command == "check-modes":
        print(json.dumps(check_modes(args), indent=2))
        return 0
    out, stats = _solve(args)
    np.savez_compressed(args.output, **forward_arrays(out, stats, args.group_rtol, args.check_continuity))
    return 2 if np.any(out.trace_flags & TRACE_SEARCH_UNCONVERGED) else 0


if __name__ == "__main__":
    raise SystemExit(main())
