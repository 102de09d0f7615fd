"""The CPU oracle against the golden fixtures written from the compiled ProDM
reference (tests/golden/gen_golden.py). Bit-exact for every integer, byte and
index result; the decoders' float outputs are compared bit for bit too.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from golden.gen_golden import GB_CASES, NB_CASES, splits

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def test_geometry(g):
    for i in range(int(g["geom_count"])):
        dims, L = g[f"geom{i}_dims"], int(g[f"geom{i}_L"])
        np.testing.assert_array_equal(ob.level_dims(dims, L), g[f"geom{i}_ld"])
        np.testing.assert_array_equal(ob.level_elements(dims, L), g[f"geom{i}_el"].astype(np.uint64))


def test_interleave(g):
    x, dims = g["ilv3_x"], g["ilv3_dims"]
    ld = ob.level_dims(dims, 2)
    parts = []
    for l in range(3):
        coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
        got = ob.interleave(x, dims, ld[l], coarse)
        np.testing.assert_array_equal(got, g[f"ilv3_l{l}"])
        parts.append(got)
    # reposition of every level rebuilds the array bit for bit
    y = np.zeros_like(x)
    for l in range(3):
        coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
        ob.reposition(parts[l], y, dims, ld[l], coarse)
    np.testing.assert_array_equal(y, x)
    got = ob.interleave(x, [9, 5, 17], [9, 5, 17], [5, 3, 9], strides=[9 * 33, 33, 1])
    np.testing.assert_array_equal(got, g["ilv3s_l1"])
    x1 = g["ilv1_x"]
    ld1 = ob.level_dims([65], 3)
    for l in range(4):
        coarse = ld1[l - 1] if l else np.zeros(1, np.uint32)
        np.testing.assert_array_equal(ob.interleave(x1, [65], ld1[l], coarse), g[f"ilv1_l{l}"])


@pytest.mark.parametrize("case", NB_CASES, ids=[c[0] for c in NB_CASES])
def test_negabinary(g, case):
    name, dt, n, nb, _ = case
    x, e = g[f"{name}_x"], int(g[f"{name}_exp"])
    planes = ob.nb_encode(x, e, nb)
    np.testing.assert_array_equal(planes, g[f"{name}_planes"])
    for j, (s, k) in enumerate(splits(nb)):
        got = ob.nb_decode(planes[s:s + k], n, e, s, k, dt)
        np.testing.assert_array_equal(got.view(np.uint8), g[f"{name}_dec{j}"].view(np.uint8))


@pytest.mark.parametrize("case", GB_CASES, ids=[c[0] for c in GB_CASES])
def test_grouped(g, case):
    name, dt, n, nb, _ = case
    x, e = g[f"{name}_x"], int(g[f"{name}_exp"])
    # the reference's (int64_t) conversion saturates on x86 for |v| >= 2^63
    compat = dt == np.float64 and nb == 64
    streams = ob.gb_encode(x, e, nb, compat=compat)
    sizes = g[f"{name}_sizes"]
    np.testing.assert_array_equal(np.array([len(s) for s in streams], np.uint32), sizes)
    assert b"".join(streams) == g[f"{name}_bytes"].tobytes()
    dec = ob.GBDecoder(dt)
    for j, (s, k) in enumerate(splits(nb)[1:]):
        got = dec.decode(0, streams[s:s + k], n, e, s, k)
        np.testing.assert_array_equal(got.view(np.uint8), g[f"{name}_dec{j}"].view(np.uint8))
    got = ob.GBDecoder(dt).decode(0, streams, n, e, 0, nb)
    np.testing.assert_array_equal(got.view(np.uint8), g[f"{name}_full"].view(np.uint8))


def test_perbit(g):
    """PerBit streams and progressive decodes reproduce the reference's bytes"""
    for i in range(int(g["pb_count"])):
        x, e, nb = g[f"pb{i}_x"], int(g[f"pb{i}_exp"]), int(g[f"pb{i}_nb"])
        sizes, raw = g[f"pb{i}_sizes"], g[f"pb{i}_bytes"].tobytes()
        want, off = [], 0
        for sz in sizes:
            want.append(raw[off:off + int(sz)])
            off += int(sz)
        got = ob.pb_encode(x, e, nb)
        assert got == want, i
        dec = ob.PBDecoder(x.dtype)
        for j, (s, k) in enumerate(splits(nb)[1:]):
            np.testing.assert_array_equal(dec.decode(0, got[s:s + k], x.size, e, s, k).view(np.uint8),
                                          g[f"pb{i}_dec{j}"].view(np.uint8))


def test_collect_max(g):
    for t, dt, nb in (("f32", np.float32, 32), ("f64", np.float64, 64)):
        for b, want in zip(g[f"max_{t}_bounds"], g[f"max_{t}"]):
            np.testing.assert_array_equal(ob.collect_max(dt(b), nb, dt), want)


def test_collect_squared_sequential_order(g):
    """The oracle's reference-order sum reproduces the reference bit for bit."""
    for i in range(int(g["sq_count"])):
        x, nb = g[f"sq{i}_x"], int(g[f"sq{i}_nb"])
        m = np.float32(np.max(np.abs(x)))
        np.testing.assert_array_equal(ob.collect_squared(x, nb, m, mode="seq"), g[f"sq{i}"])


def test_collect_squared_exact_vs_sequential(g):
    """The exactly rounded table differs from the sequential FP64 sum by at
    most the sequential sum's own rounding (n * 2^-52 relative)."""
    for i in range(int(g["sq_count"])):
        x, nb = g[f"sq{i}_x"], int(g[f"sq{i}_nb"])
        m = np.float32(np.max(np.abs(x)))
        ex = ob.collect_squared(x, nb, m, mode="exact")
        seq = g[f"sq{i}"]
        assert np.all(np.abs(ex - seq) <= x.size * 2.0 ** -52 * np.abs(seq) + 1e-300)


def test_fused_level_errors_sequential_order(g):
    """The oracle's reference-order sums reproduce the encoders' level_errors
    overloads (NegaBinaryBPEncoder.hpp:65-125, GroupedBPEncoder.hpp:76-145)
    bit for bit, including a case with elements far below 2^-nb of the max."""
    for i in range(int(g["fle_count"])):
        x, e, nb = g[f"fle{i}_x"], int(g[f"fle{i}_exp"]), int(g[f"fle{i}_nb"])
        for enc, key in (("nb", "negabinary"), ("gb", "grouped")):
            np.testing.assert_array_equal(ob.fused_level_errors(x, e, nb, enc, mode="seq"), g[f"fle{i}_{key}"])


def test_fused_level_errors_exact_vs_sequential(g):
    for i in range(int(g["fle_count"])):
        x, e, nb = g[f"fle{i}_x"], int(g[f"fle{i}_exp"]), int(g[f"fle{i}_nb"])
        for enc, key in (("nb", "negabinary"), ("gb", "grouped")):
            ex = ob.fused_level_errors(x, e, nb, enc, mode="exact")
            seq = g[f"fle{i}_{key}"]
            assert np.all(np.abs(ex - seq) <= x.size * 2.0 ** -52 * np.abs(seq) + 1e-300), (i, enc)


def test_max_estimator_constant(g):
    from prodm_amd import interpret
    np.testing.assert_allclose([interpret.max_estimator_ob_c(nd) for nd in (1, 2, 3)], g["maxest_c"], rtol=0,
                               atol=0)
