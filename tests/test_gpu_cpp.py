"""The host C++ layer on the GPU: the C++ pipeline test program
(tests/cpp/test_cpp_pipeline.cpp) and the refactor / reconstruct apps on a
small field; files written by the GPU refactor read, decoded and
repositioned by ProDM's own code (the compiled reference), and files
written by ProDM's writer reconstructed by the GPU app. The programs are
built by `python -m prodm_amd.build`."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
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


@pytest.mark.parametrize("mode", [[], ["--grouped"], ["--composed", "--grouped"]])
def test_gpu_files_read_by_reference(gpu, tmp_path, mode):
    """prodm_refactor (GpuRefactor, or the composed GPU components) writes
    the files; ProDM's code reads them (deserialize in load_metadata's order,
    ConcatLevelFileRetriever in two progressive steps,
    AdaptiveLevelCompressor), decodes every plane with ProDM's decoder
    (NegaBinaryBPEncoder / GroupedBPEncoder, exponent from the stored bound,
    ComposedReconstructor.hpp:202-204,240-242) and repositions with ProDM's
    DirectInterleaver; the oracle recomposes. The field comes back to the
    truncation of all 32 planes."""
    from test_host_format import paths, read
    if not ob.ref_available():
        pytest.skip("compiled reference not built (oracle/_ref)")
    dims, L, nb = [65, 65, 33], 4, 32
    u = bumps_field(dims, seed=5)
    f = tmp_path / "field.f32"
    u.tofile(f)
    out = tmp_path / "refactored"
    r = run([os.path.join(BIN, "prodm_refactor"), str(f), str(L), str(nb), "3", *map(str, dims), "--out", str(out),
             *mode])
    assert r.returncode == 0
    meta, lv = paths(str(out), L)
    cap = sum(os.path.getsize(x) for x in lv) * 4 + 8 * 4 * u.size
    got = read(ob.ref(), "ref_read_mdr", meta, lv, np.float32, 64, [(3 * l) % nb for l in range(L + 1)], cap)
    assert got["dims"] == dims and got["level_num"] == [1] * (L + 1)
    nega = "--grouped" not in mode
    assert got["negabinary"] == int(nega)
    ld = ob.level_dims(dims, L)
    data = np.zeros_like(u)
    dec = ob.GBDecoder(np.float32, lib="ref")
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
        n = int(np.prod(ld[l].astype(np.int64)) - np.prod(coarse.astype(np.int64)))
        bound = np.float32(got["bounds"][l])
        if nega:
            e = ob.frexp_exp(bound / np.float32(4), np.float32)
            planes = np.stack([np.frombuffer(p, np.uint32) for p in got["planes"][l]])
            vals = ob.nb_decode(planes, n, e, 0, nb, np.float32, lib="ref")
        else:
            e = ob.frexp_exp(bound, np.float32)
            vals = dec.decode(l, got["planes"][l], n, e, 0, nb)
        ob.reposition(vals, data, dims, ld[l], coarse, lib="ref")
    rec = ob.recompose(data, dims, L)
    R = float(u.max() - u.min())
    err = float(np.max(np.abs(rec.astype(np.float64) - u)))
    print(f"{' '.join(mode) or 'negabinary'}: reference reader + decoder, oracle recompose: max error {err / R:.3e} of range",
          flush=True)
    assert err <= 1e-5 * R


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reference_files_reconstructed_on_gpu(gpu, tmp_path, dtype):
    """ProDM's writer (AdaptiveLevelCompressor(64), ConcatLevelFileWriter,
    serialize in write_metadata's order) stores the oracle's plane streams of
    a field (NegaBinary for float32, GroupedBP for float64); prodm_reconstruct
    reads the files progressively and meets each L-inf tolerance."""
    from test_host_format import level_streams, write
    if not ob.ref_available():
        pytest.skip("compiled reference not built (oracle/_ref)")
    dims, L = [65, 65, 33], 4
    nb = 32 if dtype == np.float32 else 64
    levels = level_streams(dtype, dims, L, nb)
    u = bumps_field(dims, seed=7, dtype=dtype)
    f = tmp_path / "field.bin"
    u.tofile(f)
    d = tmp_path / "theirs"
    d.mkdir()
    write(ob.ref(), "ref_write_mdr", str(d), dims, levels, dtype, dtype == np.float32, 64)
    R = float(u.max() - u.min())
    tols = [1e-1 * R, 1e-3 * R, 1e-5 * R]
    extra = ["--f64"] if dtype == np.float64 else []
    r = run([os.path.join(BIN, "prodm_reconstruct"), str(f), str(len(tols)), *map(repr, tols), "--in", str(d), *extra])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("tolerance")]
    assert len(lines) == len(tols)
    for t, line in zip(tols, lines):
        err = float(line.split("max error ")[1].split()[0])
        assert err <= t, line
