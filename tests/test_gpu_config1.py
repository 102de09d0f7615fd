"""The 513^3 float32 L2 configuration at full size: 6 levels, 32 bitplanes,
NegaBinary encoder, squared-error tables, retrieval at relative L2 1e-2 ..
1e-6 with the S-norm estimator (s = 0) and the negabinary greedy planner
(SNormErrorEstimator, SquaredErrorEstimator.hpp:42-68;
NegaBinaryGreedyBasedSizeInterpreter, GreedyBasedSizeInterpreter.hpp:158-250;
ComposedReconstructor.hpp:42-44).

* SquaredErrorCollector tables (SquaredErrorCollector.hpp:23-65) computed on
  the GPU equal the oracle's exactly rounded tables on the GPU's own
  coefficients, every level.
* The negabinary encoder's own tables (NegaBinaryBPEncoder.hpp:65-125) equal
  the oracle's, every level, and the finest level's 32 planes are
  bit-exact against the oracle's encoder.
* Retrievals planned on the encoder's tables meet every relative L2 target,
  ||u - u~||_2 / ||u||_2 <= rel, checked against the original field, through
  one progressive sequence of plans. The same planner on
  SquaredErrorCollector's tables (sign-magnitude truncation at the level
  maximum's exponent, two exponents below the negabinary alignment) is
  reported beside them: it reads fewer planes and can miss the target.
"""
import os
import sys

import numpy as np
import pytest
import torch

import oracle_bind as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N, L, NB = 513, 6, 32
RELS = (1e-2, 1e-3, 1e-4, 1e-5, 1e-6)


@pytest.fixture(scope="module")
def refactored(gpu):
    sys.path.insert(0, ROOT)
    from bench import make_field_gpu
    from prodm_amd import capi, mdr
    print(f"config 1: building the {N}^3 field and refactoring (L={L}, nb={NB}, negabinary)", flush=True)
    u = make_field_gpu(N, seed=0, nbumps=16, noise=1e-4)
    plan = mdr.Plan([N] * 3, L, NB, torch.float32, capi.NEGABINARY)
    plan.refactor(u)
    torch.cuda.synchronize()
    yield u, plan
    plan.close()


def _level(plan, l):
    m, e, n = plan.level_info(l)
    return np.float32(m), e, n


def test_config1_coefficients_match_oracle(refactored):
    """Every level's coefficients against the oracle's decompose + interleave
    of the same field (fp32 tolerance 1e-6 of the largest coefficient, as for
    the smaller plan tests); the finest level is the sweep of 7 x 65-node
    tiles that only fields of this size take."""
    u, plan = refactored
    host = u.cpu().numpy()
    prev = ob.use_library(ob.ORACLE_MT_SO)
    try:
        dec = ob.decompose(host, [N] * 3, L)
        ld = ob.level_dims([N] * 3, L)
        ref = [ob.interleave(dec, [N] * 3, ld[l], ld[l - 1] if l else np.zeros(3, np.uint32)) for l in range(L + 1)]
    finally:
        ob.restore_library(prev)
    scale = max(float(np.max(np.abs(r))) for r in ref)
    for l in range(L + 1):
        c = plan.coefficients(l).cpu().numpy()
        assert c.size == ref[l].size, l
        err = float(np.max(np.abs(c.astype(np.float64) - ref[l])))
        print(f"level {l}: {c.size} coefficients, max |gpu - oracle| = {err:.3e} (scale {scale:.3e})", flush=True)
        assert err <= 1e-6 * scale, l


def test_config1_squared_collector_tables(refactored):
    from prodm_amd import mdr
    _, plan = refactored
    for l in range(L + 1):
        m, _, n = _level(plan, l)
        c = plan.coefficients(l)
        got = mdr.collect_squared(c, NB, m)
        want = ob.collect_squared(c.cpu().numpy(), NB, m, mode="exact")
        print(f"level {l}: {n} coefficients, SquaredErrorCollector table bit-exact", flush=True)
        np.testing.assert_array_equal(got, want)


def test_config1_negabinary_tables_and_planes(refactored):
    _, plan = refactored
    for l in range(L + 1):
        _, e, n = _level(plan, l)
        c = plan.coefficients(l).cpu().numpy()
        np.testing.assert_array_equal(plan.level_errors(l), ob.fused_level_errors(c, e, NB, "nb"))
        print(f"level {l}: negabinary level-error table bit-exact", flush=True)
        if l == L:
            want = ob.nb_encode(c, e, NB)
            for p in range(NB):
                assert plan.plane_bytes(l, p) == want[p].tobytes(), p
            print(f"level {l}: {NB} planes bit-exact", flush=True)


def _plan_l2(tables, sizes, tol, index):
    from prodm_amd import interpret as ip
    _, idx = ip.retrieve_sizes(ip.NEGABINARY_GREEDY, ip.SNORM, sizes, tables, tol, index, ndims=3, target_level=L)
    index[:] = [int(k) for k in idx]
    return sum(2.0 ** (3 * (L - l)) * tables[l][index[l]] for l in range(L + 1))


def test_config1_l2_retrievals(refactored):
    from prodm_amd import mdr
    u, plan = refactored
    ud = u.double()
    norm2 = float(torch.sum(ud * ud))
    sizes, enc_tables, col_tables = [], [], []
    for l in range(L + 1):
        m, _, n = _level(plan, l)
        sizes.append([((n + 31) // 32) * 4] * NB)
        enc_tables.append(plan.level_errors(l))
        col_tables.append(mdr.collect_squared(plan.coefficients(l), NB, m))
    out = torch.empty_like(u)
    held_enc, held_col = [0] * (L + 1), [0] * (L + 1)
    for rel in RELS:
        tol = rel * rel * norm2  # S-norm estimate: squared error on the fine grid
        est = _plan_l2(enc_tables, sizes, tol, held_enc)
        plan.retrieve(held_enc, out)
        got = float(torch.sqrt(torch.sum((out.double() - ud) ** 2) / norm2))
        _plan_l2(col_tables, sizes, tol, held_col)
        plan.retrieve(held_col, out)
        got_col = float(torch.sqrt(torch.sum((out.double() - ud) ** 2) / norm2))
        print(f"rel L2 {rel:.0e}: encoder tables planes {held_enc} estimate {np.sqrt(est / norm2):.3e} "
              f"actual {got:.3e}; SquaredErrorCollector tables planes {held_col} actual {got_col:.3e}", flush=True)
        assert got <= rel, (rel, held_enc, got)
