"""The host C++ layer on the GPU: the C++ pipeline test program
(tests/cpp/test_cpp_pipeline.cpp) and the refactor / reconstruct apps on a
small field. The programs are built by `python -m prodm_amd.build`."""
import os
import subprocess
import sys

import numpy as np
import pytest

from fields import bumps_field

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "prodm_amd", "bin")


def run(args, **kw):
    print("run:", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=240, **kw)
    sys.stdout.write(r.stdout[-4000:])
    sys.stderr.write(r.stderr[-4000:])
    return r


def test_cpp_pipeline(gpu, tmp_path):
    r = run([os.path.join(BIN, "test_cpp_pipeline"), str(tmp_path)])
    assert r.returncode == 0, r.stdout[-2000:]
    assert "PASSED" in r.stdout


@pytest.mark.parametrize("mode", [[], ["--grouped"], ["--composed"], ["--squared"], ["--hierarchical"]])
def test_apps_refactor_reconstruct(gpu, tmp_path, mode):
    dims = [65, 65, 33]
    u = bumps_field(dims, seed=5)
    f = tmp_path / "field.f32"
    u.tofile(f)
    out = tmp_path / "refactored"
    r = run([os.path.join(BIN, "prodm_refactor"), str(f), "4", "32", "3", *map(str, dims), "--out", str(out), *mode])
    assert r.returncode == 0
    R = float(u.max() - u.min())
    tols = [1e-1 * R, 1e-3 * R, 1e-5 * R]
    extra = ["--l2"] if mode == ["--squared"] else (["--hierarchical"] if mode == ["--hierarchical"] else [])
    if extra == ["--l2"]:
        tols = [(t * float(np.linalg.norm(u))) ** 2 for t in (1e-2, 1e-4)]
    r = run([os.path.join(BIN, "prodm_reconstruct"), str(f), str(len(tols)), *map(repr, tols), "--in", str(out),
             *extra])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("tolerance")]
    assert len(lines) == len(tols)
    if extra != ["--l2"]:
        for t, line in zip(tols, lines):
            err = float(line.split("max error ")[1].split()[0])
            assert err <= t, line
