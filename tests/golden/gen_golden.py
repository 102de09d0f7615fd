"""Writes the golden fixtures under tests/golden/ from the compiled ProDM
reference checker (oracle/_ref/libref.so, built by `make -C oracle ref` from
the reference headers in place).

The fixtures are data only: seeded inputs and the outputs the reference
produced for them. They pin the oracle (tests/test_oracle_golden.py) without
needing the reference sources at test time.

    python tests/golden/gen_golden.py

Covered reference functions (paths relative to the ProDM repository):
  compute_level_dims / compute_level_elements   include/MDR/RefactorUtils.hpp:21-59
  compute_max_abs_value                          include/MDR/RefactorUtils.hpp:68-76
  DirectInterleaver interleave / reposition      include/MDR/Interleaver/DirectInterleaver.hpp:12-63
  NegaBinaryBPEncoder encode / progressive_decode include/MDR/BitplaneEncoder/NegaBinaryBPEncoder.hpp:18-62,132-192
  GroupedBPEncoder encode / progressive_decode   include/MDR/BitplaneEncoder/GroupedBPEncoder.hpp:18-73,210-297
  MaxErrorCollector                              include/MDR/ErrorCollector/MaxErrorCollector.hpp:14-25
  SquaredErrorCollector<float>                   include/MDR/ErrorCollector/SquaredErrorCollector.hpp:23-65
  MaxErrorEstimatorOB constant                   include/MDR/ErrorEstimator/MaxErrorEstimator.hpp:23-31
  size interpreters                              include/MDR/SizeInterpreter/GreedyBasedSizeInterpreter.hpp:23-250,
                                                 include/MDR/SizeInterpreter/BasicSizeInterpreter.hpp:11-112
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402


def codec_input(n, dtype, seed, tiny=False):
    """Mixed magnitudes and signs, exact zeros, one element at the maximum."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp2(rng.integers(-12, 1, n))
    x[rng.random(n) < 0.05] = 0.0
    x[rng.integers(0, n)] = -1.75
    if tiny:
        x = x * (1e-36 if dtype == np.float32 else 1e-300)
    return x.astype(dtype)


def level_exp(x):
    m = float(np.max(np.abs(x)))
    return math.frexp(m)[1]


def gen_geometry(out):
    cases = [([129, 129, 129], 4), ([513, 513, 513], 6), ([1025, 1025, 257], 7), ([65], 5), ([17, 9, 33], 2),
             ([100, 37, 8], 2), ([1, 33, 17], 3)]
    for i, (dims, L) in enumerate(cases):
        out[f"geom{i}_dims"] = np.array(dims, np.uint32)
        out[f"geom{i}_L"] = np.array(L)
        out[f"geom{i}_ld"] = ob.level_dims(dims, L, lib=ob.ref())
        el = np.zeros(L + 1, np.uint32)
        import ctypes as C
        ob.ref().ref_level_elements(np.array(dims, np.uint32).ctypes.data_as(C.c_void_p), len(dims), L,
                                   el.ctypes.data_as(C.c_void_p))
        out[f"geom{i}_el"] = el
    out["geom_count"] = np.array(len(cases))


def gen_interleave(out):
    rng = np.random.default_rng(11)
    # 3-D dense: 17 x 9 x 33, two levels
    dims, L = [17, 9, 33], 2
    x = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    ld = ob.level_dims(dims, L)
    out["ilv3_x"] = x
    out["ilv3_dims"] = np.array(dims, np.uint32)
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
        out[f"ilv3_l{l}"] = ob.interleave(x, dims, ld[l], coarse, lib="ref")
    # 3-D strided: a 9 x 5 x 17 box inside a 17 x 9 x 33 array (full strides)
    strides = [9 * 33, 33, 1]
    box = [9, 5, 17]
    out["ilv3s_l1"] = ob.interleave(x, box, box, [5, 3, 9], strides=strides, lib="ref")
    # 1-D fp64
    x1 = rng.standard_normal(65)
    out["ilv1_x"] = x1
    ld1 = ob.level_dims([65], 3)
    for l in range(4):
        coarse = ld1[l - 1] if l else np.zeros(1, np.uint32)
        out[f"ilv1_l{l}"] = ob.interleave(x1, [65], ld1[l], coarse, lib="ref")


NB_CASES = [  # (name, dtype, n, nb, tiny)
    ("nbf32", np.float32, 1000, 32, False),
    ("nbf32b", np.float32, 96, 16, False),
    ("nbf32t", np.float32, 333, 32, True),
    ("nbf64", np.float64, 700, 64, False),
    ("nbf64b", np.float64, 129, 40, False),
    ("nbf64t", np.float64, 200, 64, True),
]
GB_CASES = [  # (name, dtype, n, nb, tiny); n >= block size (the reference wraps below it)
    ("gbf32", np.float32, 1000, 32, False),
    ("gbf32b", np.float32, 64, 20, False),
    ("gbf32t", np.float32, 333, 32, True),
    ("gbf64", np.float64, 700, 60, False),
    ("gbf64c", np.float64, 300, 64, False),  # nb 64: the x86 saturation of the reference applies
    ("gbf64t", np.float64, 200, 60, True),
]


def splits(nb):
    a, b = nb // 5, nb // 2
    return [(0, nb), (0, a), (a, b - a), (b, nb - b)]


def gen_negabinary(out):
    for i, (name, dt, n, nb, tiny) in enumerate(NB_CASES):
        x = codec_input(n, dt, 100 + i, tiny)
        e = level_exp(x)
        planes = ob.nb_encode(x, e, nb, lib="ref")
        out[f"{name}_x"] = x
        out[f"{name}_exp"] = np.array(e)
        out[f"{name}_nb"] = np.array(nb)
        out[f"{name}_planes"] = planes
        for j, (s, k) in enumerate(splits(nb)):
            out[f"{name}_dec{j}"] = ob.nb_decode(planes[s:s + k], n, e, s, k, dt, lib="ref")


def gen_grouped(out):
    for i, (name, dt, n, nb, tiny) in enumerate(GB_CASES):
        x = codec_input(n, dt, 200 + i, tiny)
        e = level_exp(x)
        streams = ob.gb_encode(x, e, nb, lib="ref")
        out[f"{name}_x"] = x
        out[f"{name}_exp"] = np.array(e)
        out[f"{name}_nb"] = np.array(nb)
        out[f"{name}_sizes"] = np.array([len(s) for s in streams], np.uint32)
        out[f"{name}_bytes"] = np.frombuffer(b"".join(streams), np.uint8)
        dec = ob.GBDecoder(dt, lib="ref")
        for j, (s, k) in enumerate(splits(nb)[1:]):  # progressive: one level, consecutive windows
            out[f"{name}_dec{j}"] = dec.decode(0, streams[s:s + k], n, e, s, k)
        one = ob.GBDecoder(dt, lib="ref")
        out[f"{name}_full"] = one.decode(0, streams, n, e, 0, nb)


def gen_perbit(out):
    """PerBitBPEncoder streams and progressive decodes (PerBitBPEncoder.hpp:72-391)"""
    cases = (("f32", np.float32, 32, 1000), ("f32", np.float32, 16, 777), ("f64", np.float64, 64, 1000),
             ("f64", np.float64, 40, 513), ("f32", np.float32, 32, 1))
    for i, (t, dt, nb, n) in enumerate(cases):
        x = codec_input(n, dt, 500 + i)
        e = level_exp(x)
        streams = ob.pb_encode(x, e, nb, lib="ref")
        name = f"pb{i}"
        out[f"{name}_x"] = x
        out[f"{name}_exp"] = np.array(e)
        out[f"{name}_nb"] = np.array(nb)
        out[f"{name}_sizes"] = np.array([len(s) for s in streams], np.uint32)
        out[f"{name}_bytes"] = np.frombuffer(b"".join(streams), np.uint8)
        dec = ob.PBDecoder(dt, lib="ref")
        for j, (s, k) in enumerate(splits(nb)[1:]):
            out[f"{name}_dec{j}"] = dec.decode(0, streams[s:s + k], n, e, s, k)
    out["pb_count"] = np.array(len(cases))


def gen_collectors(out):
    bounds = [1.0, 3.7, 0.3, 1e-30]
    for t, dt, nb in (("f32", np.float32, 32), ("f64", np.float64, 64)):
        out[f"max_{t}_bounds"] = np.array(bounds, dt)
        out[f"max_{t}"] = np.stack([ob.collect_max(dt(b), nb, dt, lib="ref") for b in bounds])
    for i, (n, nb) in enumerate(((3000, 32), (500, 20), (4096, 28))):
        x = codec_input(n, np.float32, 300 + i)
        m = np.float32(np.max(np.abs(x)))
        out[f"sq{i}_x"] = x
        out[f"sq{i}_nb"] = np.array(nb)
        out[f"sq{i}"] = ob.collect_squared(x, nb, m, mode="ref")
    out["sq_count"] = np.array(3)
    # the encoders' level_errors overloads (fp32): NegaBinary and GroupedBP
    fle = ((1000, 32, False), (777, 16, False), (64, 30, False), (4096, 32, True))
    for i, (n, nb, tiny) in enumerate(fle):
        x = codec_input(n, np.float32, 400 + i)
        if tiny:  # a wide dynamic range: elements far below 2^-nb of the maximum
            x[: n // 4] *= np.float32(2.0 ** -40)
        e = level_exp(x)
        out[f"fle{i}_x"] = x
        out[f"fle{i}_exp"] = np.array(e)
        out[f"fle{i}_nb"] = np.array(nb)
        out[f"fle{i}_negabinary"] = ob.fused_level_errors(x, e, nb, "nb", mode="ref")
        out[f"fle{i}_grouped"] = ob.fused_level_errors(x, e, nb, "gb", mode="ref")
    out["fle_count"] = np.array(len(fle))
    out["maxest_c"] = np.array([ob.ref_max_estimator_c(nd) for nd in (1, 2, 3)])


def interp_tables(rng, L, nb, kind):
    counts = [nb] * L
    sizes, errors = [], []
    for l in range(L):
        s = rng.integers(50, 5000, nb).astype(np.uint32)
        if kind == 2:  # negabinary: dense planes of equal size
            s[:] = s[0]
        sizes.append(s)
        e0 = float(rng.uniform(0.5, 4.0)) * 2.0 ** (-l)
        decay = rng.uniform(0.3, 0.7, nb)
        err = [e0]
        for k in range(nb):
            err.append(err[-1] * float(decay[k]))
        errors.append(np.array(err))
    return counts, sizes, errors


def gen_interpreters(out):
    rng = np.random.default_rng(400)
    i = 0
    for est in (0, 1, 2):
        for kind in (0, 1, 2, 3, 4):
            L, nb = 4, 16
            counts, sizes, errors = interp_tables(rng, L, nb, kind)
            for tol_rel in (1e-1, 1e-3, 1e-5):
                tol = tol_rel * float(errors[0][0])
                ret, idx = ob.ref_interpret(kind, est, 3, L, counts, sizes, errors, tol, np.zeros(L, np.uint8),
                                            s_param=0.0)
                out[f"it{i}_meta"] = np.array([kind, est, 3, L, nb])
                out[f"it{i}_tol"] = np.array(tol)
                out[f"it{i}_sizes"] = np.stack(sizes)
                out[f"it{i}_errors"] = np.stack(errors)
                out[f"it{i}_ret"] = ret
                out[f"it{i}_idx"] = idx
                i += 1
    out["it_count"] = np.array(i)


def main():
    if not ob.ref_available():
        sys.exit("oracle/_ref/libref.so is not built: run `make -C oracle ref`")
    out = {}
    gen_geometry(out)
    gen_interleave(out)
    gen_negabinary(out)
    gen_grouped(out)
    gen_perbit(out)
    gen_collectors(out)
    gen_interpreters(out)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
