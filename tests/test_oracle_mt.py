"""The OpenMP build of the oracle (oracle/liboracle_mt.so, the all-core CPU
baseline) computes exactly what the single-thread checker computes: its
pragmas only split loops whose iterations write disjoint outputs."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from fields import bumps_field


@pytest.mark.skipif(not os.path.exists(ob.ORACLE_MT_SO), reason="liboracle_mt.so not built (make -C oracle)")
@pytest.mark.parametrize("dtype,dims,L,nb", [(np.float32, [65, 33, 65], 4, 32), (np.float64, [33, 65], 4, 52)])
def test_openmp_build_is_identical(dtype, dims, L, nb):
    u = bumps_field(dims, seed=3, dtype=dtype)

    def pipeline():
        dec = ob.decompose(u, dims, L)
        ld = ob.level_dims(dims, L)
        planes, vals = [], []
        for l in range(L + 1):
            coarse = ld[l - 1] if l else np.zeros(len(dims), np.uint32)
            c = ob.interleave(dec, dims, ld[l], coarse)
            e = ob.frexp_exp(ob.maxabs(c), dtype)
            pl = ob.nb_encode(c, e, nb)
            planes.append(pl)
            vals.append(ob.nb_decode(pl, c.size, e, 0, nb // 2, dtype))
        rec = ob.recompose(dec, dims, L)
        return dec, planes, vals, rec

    want = pipeline()
    prev = ob.use_library(ob.ORACLE_MT_SO)
    try:
        got = pipeline()
    finally:
        ob.restore_library(prev)
    np.testing.assert_array_equal(got[0], want[0])
    for a, b in zip(got[1], want[1]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(got[2], want[2]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[3], want[3])
