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


@pytest.mark.parametrize("mode", [[], ["--grouped"], ["--composed"], ["--squared"], ["--squared-encoder"],
                                  ["--squared-encoder", "--composed"], ["--hierarchical"]])
def test_apps_refactor_reconstruct(gpu, tmp_path, mode):
    """Refactor a 65x65x33 field into files and reconstruct it progressively.
    L-inf tolerances 1e-1, 1e-3, 1e-5 of the range are met; with the
    negabinary encoder's squared-error tables (--squared-encoder), S-norm
    retrievals at relative L2 1e-2 .. 1e-6 meet each target. (Files written
    with --squared carry SquaredErrorCollector's tables, which model
    sign-magnitude truncation; they are reconstructed to L-inf tolerances.)"""
    dims = [65, 65, 33]
    u = bumps_field(dims, seed=5)
    f = tmp_path / "field.f32"
    u.tofile(f)
    out = tmp_path / "refactored"
    r = run([os.path.join(BIN, "prodm_refactor"), str(f), "4", "32", "3", *map(str, dims), "--out", str(out), *mode])
    assert r.returncode == 0
    R = float(u.max() - u.min())
    norm2 = float(np.sum(u.astype(np.float64) ** 2))
    l2 = "--squared-encoder" in mode
    if l2:
        rels = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6]
        tols = [rel * rel * norm2 for rel in rels]  # S-norm estimate: squared error on the fine grid
        extra = ["--l2"]
    else:
        tols = [1e-1 * R, 1e-3 * R, 1e-5 * R]
        extra = ["--hierarchical"] if "--hierarchical" in mode else []
    r = run([os.path.join(BIN, "prodm_reconstruct"), str(f), str(len(tols)), *map(repr, tols), "--in", str(out),
             *extra])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("tolerance")]
    assert len(lines) == len(tols)
    if l2:
        for rel, line in zip(rels, lines):
            got = float(line.split("L2 rel ")[1].split(",")[0])
            assert got <= rel, line
    else:
        for t, line in zip(tols, lines):
            err = float(line.split("max error ")[1].split()[0])
            assert err <= t, line
