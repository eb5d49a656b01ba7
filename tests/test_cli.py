"""Command line (modisp_amd.cli): the reference's commands and NPZ layout
(python/modisp/cli.py:9-117) with --backend rocm.

CPU: doctor without a device, check-modes on a stored result, argument
handling. GPU: forward writes the reference's key set with the reference's
dtypes, and its phase velocities equal the front end's."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from modisp_amd import cli, workloads

ROOT = Path(__file__).resolve().parents[1]
# Keys of the reference's forward NPZ (python/modisp/cli.py:98-112), always present.
FORWARD_KEYS = {"refinement", "love_strategy", "phase_velocity", "periods", "mode_mask", "status", "trace_flags",
                "search_bounds", "search_level", "root_count", "backend", "wave", "precision", "accuracy",
                "numerical_flags", "approximate_search", "close_root_mask", "potentially_ambiguous",
                "mode_index_uncertain", "close_root_risk", "close_root_rtol", "search_strategy", "sampling_rtol",
                "secular", "group_rtol", "frequency_strategy", "frequency_strategy_used",
                "frequency_strategy_reason", "anchor_stride"}


def run(*args):
    return subprocess.run([sys.executable, "-m", "modisp_amd.cli", *args], capture_output=True, text=True,
                          cwd=ROOT)


def test_doctor_reports_the_plugin():
    proc = run("doctor")
    assert proc.returncode == 0, proc.stderr
    info = json.loads(proc.stdout)
    assert info["version"] == "0.5.0" and info["private_abi"] == 9 and info["default_accuracy"] == "auto"
    assert "rocm_available" in info and info["library"].endswith("libmodisp_rocm.so")


def test_check_modes_on_a_stored_result(tmp_path):
    per = np.geomspace(1., 3., 12)
    v = np.broadcast_to(3. + .4 * np.arange(3)[None, :, None] + .05 * np.log(per), (2, 3, 12)).copy()
    v[0, 1, 6] *= 1.2
    src = tmp_path / "in.npz"
    np.savez(src, phase_velocity=v, periods=per, trace_flags=np.zeros((2, 12), np.uint32),
             root_count=np.full((2, 12), 3, np.int32), close_root_rtol=np.asarray(.005))
    proc = run("check-modes", str(src), "--output", str(tmp_path / "out.npz"))
    assert proc.returncode == 0, proc.stderr
    summary = json.loads(proc.stdout)
    assert summary["cells"] == 72 and summary["flags"]["STEP_JUMP"] >= 1
    with np.load(tmp_path / "out.npz") as z:
        assert set(z.files) == {"continuity_flags", "continuity_assessed", "continuity_previous_index",
                                "continuity_sorted_order"}
        assert z["continuity_flags"][0, 1, 6] & 32


def test_forward_options_are_the_reference_options():
    p = cli.parser()
    a = p.parse_args(["forward", "in.npz", "--output", "o.npz", "--backend", "rocm", "--group-velocity",
                      "--frequency-strategy", "anchor_continuation", "--check-continuity"])
    assert a.backend == "rocm" and a.group_velocity and a.check_continuity and a.nmodes == 3
    with pytest.raises(SystemExit):
        p.parse_args(["forward", "in.npz"])     # --output is required


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--group-velocity", "--check-continuity", "--accuracy", "balanced"]])
def test_forward_writes_the_reference_npz(tmp_path, extra):
    import modisp_amd
    from modisp_amd import api
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    wl = workloads.plain_batch(B=3, seed=14)
    src = tmp_path / "in.npz"
    np.savez(src, thickness=wl.thickness, vp=wl.vp, vs=wl.vs, rho=wl.rho, periods=wl.periods[::6])
    rc = cli.main(["forward", str(src), "--output", str(tmp_path / "out.npz"), "--backend", "rocm",
                   "--nmodes", "3", *extra])
    assert rc == 0
    with np.load(tmp_path / "out.npz") as z:
        assert FORWARD_KEYS <= set(z.files)
        assert z["phase_velocity"].dtype == np.float64 and z["trace_flags"].dtype == np.uint32
        assert z["status"].dtype == np.uint8 and z["root_count"].dtype == np.int32
        assert str(z["backend"]) == "rocm" and str(z["wave"]) == "rayleigh"
        if extra:
            assert {"group_velocity", "group_velocity_flags", "mode_reliability_flags", "continuity_flags"} <= set(
                z.files)
        accuracy = "balanced" if extra else None
        with api.Solver(backend="rocm", accuracy=accuracy, group_velocity=bool(extra)) as s:
            ref = s(wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods[::6], 3)
        np.testing.assert_array_equal(z["phase_velocity"], ref.phase_velocity)
