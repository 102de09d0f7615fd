"""Host-side checks of libprodm_mdr.so that need no GPU: the library loads,
exports every entry point include/prodm_mdr.h declares, and its host-only
functions (geometry, max-error tables, retrieval planners) match the oracle
and the golden fixtures written from the compiled ProDM reference.
"""
import os
import re

import numpy as np
import pytest

import oracle_bind as ob
from prodm_amd import capi, interpret

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "prodm_mdr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(prodm_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert sorted(capi.exported_symbols()) == syms


def test_version_and_error_channel():
    lib = capi.lib()
    assert lib.prodm_version() >= 10000
    import ctypes as C
    dims = np.array([10, 10], np.uint32)
    out = np.zeros(3 * 2, np.uint32)
    el = np.zeros(3, np.uint64)
    ok = lib.prodm_level_geometry(dims.ctypes.data_as(C.c_void_p), 2, 2, out.ctypes.data_as(C.c_void_p),
                                  el.ctypes.data_as(C.c_void_p))
    assert ok == 0  # even dims are not a supported decomposition geometry
    rc = lib.prodm_interpret_retrieve_size(0, 9, 0, 3, 0, 0.0, 1, None, None, None, 1.0, None, None)
    assert rc == -1
    assert b"bad arguments" in lib.prodm_last_error()


def test_planes_of_2gib_refused_before_any_device_call():
    """The level sweeps address a plane of the last two dims with 31-bit byte
    offsets, so a plan (and a standalone decompose or recompose) refuses a
    geometry whose plane spans 2 GiB or more, with a message, before any
    device call: this runs without a GPU. Planes just under the limit are
    accepted by the geometry check."""
    import ctypes as C
    lib = capi.lib()
    plan = C.c_void_p()
    for dims, dt in (([3, 32769, 16385], 0), ([3, 16385, 16385], 1), ([32769, 16385], 0)):
        d = np.array(dims, np.uint32)
        for enc in (0, 1):
            rc = lib.prodm_mdr_plan_create(C.byref(plan), d.ctypes.data_as(C.c_void_p), len(dims), 1, 32, dt, enc)
            assert rc == -6 and not plan.value, (dims, dt, rc)
            assert b"2 GiB" in lib.prodm_last_error()
        dec = lib.prodm_decompose_f64 if dt else lib.prodm_decompose_f32
        rec = lib.prodm_recompose_f64 if dt else lib.prodm_recompose_f32
        for fn in (dec, rec):
            rc = fn(None, None, d.ctypes.data_as(C.c_void_p), len(dims), 1, None, None, 0, None)
            assert rc == -6 and b"2 GiB" in lib.prodm_last_error()
    # one element short of the limit passes the geometry check and fails on
    # the null pointers instead
    d = np.array([3, 16383, 32767], np.uint32)  # 16383 * 32767 * 4 < 2^31 - 1
    rc = lib.prodm_decompose_f32(None, None, d.ctypes.data_as(C.c_void_p), 3, 1, None, None, 0, None)
    assert rc == -1 and b"distinct device input" in lib.prodm_last_error()


@pytest.mark.parametrize("dims,L", [([513, 513, 513], 6), ([1025, 1025, 257], 7), ([129, 129, 129], 4),
                                    ([65], 5), ([33, 17], 3), ([1, 65, 33], 4)])
def test_level_geometry_matches_oracle(dims, L):
    import ctypes as C
    lib = capi.lib()
    nd = len(dims)
    out = np.zeros((L + 1) * nd, np.uint32)
    el = np.zeros(L + 1, np.uint64)
    d = np.array(dims, np.uint32)
    ok = lib.prodm_level_geometry(d.ctypes.data_as(C.c_void_p), nd, L, out.ctypes.data_as(C.c_void_p),
                                  el.ctypes.data_as(C.c_void_p))
    assert ok == 1
    np.testing.assert_array_equal(out.reshape(L + 1, nd), ob.level_dims(dims, L))
    np.testing.assert_array_equal(el, ob.level_elements(dims, L))
    assert int(el.sum()) == int(np.prod(dims))


def test_collect_max_matches_fixtures():
    import ctypes as C
    g = np.load(GOLD, allow_pickle=False)
    lib = capi.lib()
    for t, dt, nb in (("f32", np.float32, 32), ("f64", np.float64, 64)):
        fn = getattr(lib, "prodm_collect_max_" + t)
        for b, want in zip(g[f"max_{t}_bounds"], g[f"max_{t}"]):
            got = np.zeros(nb + 1)
            fn(dt(b), nb, got.ctypes.data_as(C.c_void_p))
            np.testing.assert_array_equal(got, want)


def test_max_estimator_constant():
    g = np.load(GOLD, allow_pickle=False)
    got = [interpret.max_estimator_ob_c(nd) for nd in (1, 2, 3)]
    np.testing.assert_array_equal(got, g["maxest_c"])


def test_planners_match_reference_fixtures():
    g = np.load(GOLD, allow_pickle=False)
    for i in range(int(g[f"it_count"])):
        kind, est, nd, L, nb = (int(v) for v in g[f"it{i}_meta"])
        sizes = list(g[f"it{i}_sizes"])
        errors = list(g[f"it{i}_errors"])
        idx = np.zeros(L, np.uint8)
        ret, idx = interpret.retrieve_sizes(kind, est, sizes, errors, float(g[f"it{i}_tol"]), idx,
                                            dtype=np.float32, ndims=nd, target_level=L - 1)
        np.testing.assert_array_equal(ret, g[f"it{i}_ret"], err_msg=f"case {i} kind {kind} est {est}")
        np.testing.assert_array_equal(idx, g[f"it{i}_idx"], err_msg=f"case {i} kind {kind} est {est}")


@pytest.mark.skipif(not ob.ref_available(), reason="compiled reference checker not built")
@pytest.mark.parametrize("kind", range(5))
@pytest.mark.parametrize("est", range(3))
def test_planners_match_reference_progressive(kind, est):
    """Several plans in a row, each starting from the previous index, against
    the compiled reference planner."""
    rng = np.random.default_rng(kind * 7 + est)
    L, nb = 5, 32
    sizes = [rng.integers(10, 4000, nb).astype(np.uint32) for _ in range(L)]
    if kind == 2:
        for s in sizes:
            s[:] = s[0]
    errors = []
    for l in range(L):
        e = [float(rng.uniform(0.5, 4)) * 2.0 ** -l]
        for _ in range(nb):
            e.append(e[-1] * float(rng.uniform(0.2, 0.8)))
        errors.append(np.array(e))
    mine = np.zeros(L, np.uint8)
    theirs = np.zeros(L, np.uint8)
    for tol in (1e-1, 1e-2, 1e-4, 1e-6):
        t = tol * errors[0][0]
        r1, mine = interpret.retrieve_sizes(kind, est, sizes, errors, t, mine, ndims=3, target_level=L - 1)
        r2, theirs = ob.ref_interpret(kind, est, 3, L, [nb] * L, sizes, errors, t, theirs)
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(mine, theirs)
