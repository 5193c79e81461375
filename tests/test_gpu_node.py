"""The Node drop-in (Sim / RigidSphere over the N-API addon) on the GPU: the
reference's own call sequence reproduces the reference's results, through the
native-sphere path and through the generic duck-typed solid path, and the
counterparts of the reference's scenario scripts run clean."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_io
from gpu_util import announce, interior_fluid, rel_err, same_values

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "tests", "js")
NODE = shutil.which("node")


def run_node(script, *args, timeout=240):
    if NODE is None:
        pytest.skip("node is not installed")
    announce("node", script, *args)
    r = subprocess.run([NODE, os.path.join(JS, script), *args], capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    print(r.stdout[-4000:])
    print(r.stderr[-2000:])
    return r


def read_out(path, nsteps, n):
    raw = open(path, "rb").read()
    off = 0
    sph = np.frombuffer(raw, np.float64, nsteps * 27, off).reshape(nsteps, 27)
    off += nsteps * 27 * 8
    fields = {}
    for k in ("phi", "rho", "ux", "uy", "uz"):
        fields[k] = np.frombuffer(raw, np.float32, n, off)
        off += n * 4
    fields["tag"] = np.frombuffer(raw, np.uint8, n, off)
    return sph, fields


@pytest.mark.parametrize("mode", ["passes", "generic"])
def test_dropin_config0_bit_exact(tmp_path, mode):
    out = str(tmp_path / "o.bin")
    r = run_node("parity_run.mjs", "config0", mode, out)
    assert r.returncode == 0
    meta, A = golden_io.load("config0")
    sph, fields = read_out(out, 200, 40 * 28 * 40)
    for t in range(200):
        assert list(sph[t]) == meta["sphere"][t], (mode, t)
    for k in ("phi", "rho", "ux", "uy", "uz", "tag"):
        assert same_values(fields[k], A[f"step200/{k}"]), k


def test_dropin_fused_config0_within_tolerance(tmp_path):
    out = str(tmp_path / "o.bin")
    r = run_node("parity_run.mjs", "config0", "fused", out)
    assert r.returncode == 0
    meta, A = golden_io.load("config0")
    sph, fields = read_out(out, 200, 40 * 28 * 40)
    for t in range(200):
        ref = meta["sphere"][t]
        assert np.max(np.abs(sph[t][3:6] - np.array(ref[3:6]))) <= 1e-4, t
    mask = interior_fluid(A["step200/tag"], 40, 28, 40)
    assert np.array_equal(fields["tag"], A["step200/tag"])
    for k in ("phi", "rho", "ux", "uy", "uz"):
        assert rel_err(fields[k], A[f"step200/{k}"], mask) <= 1e-5, k


def test_dropin_smoke12_bit_exact(tmp_path):
    out = str(tmp_path / "o.bin")
    r = run_node("parity_run.mjs", "smoke12", "passes", out)
    assert r.returncode == 0
    meta, A = golden_io.load("smoke12")
    sph, fields = read_out(out, 80, 12 ** 3)
    for t in range(80):
        assert list(sph[t]) == meta["sphere"][t], t
    for k in ("phi", "rho", "ux", "uy", "uz", "tag"):
        assert same_values(fields[k], A[f"step80/{k}"]), k


def test_dropin_scenarios():
    r = run_node("scenarios.mjs", timeout=400)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "ALL SCENARIOS OK" in r.stdout


def test_frame_delivery_end_to_end():
    """frame_server.js + remote_worker.js: the worker protocol over a WebSocket,
    frames equal to packVolume of the same scene stepped locally."""
    r = run_node("frame_delivery.mjs", timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert "OK frames=" in r.stdout


def test_dropin_run_async_and_check():
    """runAsync (worker thread, Promise) equals run; checkFields on the state."""
    r = run_node("async_run.mjs", timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert "ASYNC OK" in r.stdout
