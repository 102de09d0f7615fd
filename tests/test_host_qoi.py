"""Host checks of the certified amplification and the QoI planner
(include/prodm/qoi.hpp through libprodm_mdr.so), no GPU:

* prodm_recompose_amplification equals, level by level, the largest row
  1-norm of the level's block of the recompose operator assembled densely
  from the oracle's recompose of unit coefficient vectors (every node, not a
  sample), for 1-, 2- and 3-D grids; hierarchical: all 1.
* MaxErrorEstimatorOB's constant per dimension count bounds it (ProDM's
  L-infinity estimator is a bound without its extra factor 4).
* The QoI interpreter's plan brings the sampled |V| bound (ProDM's
  composition, include/qoi_utils.hpp:18-41) below the tolerance, charges the
  bytes of the planes it takes, takes nothing more once the planes held meet
  the tolerance, and makes the greedy's choice on a hand-checkable case.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from prodm_amd import interpret as ip


def dense_level_row_norms(dims, L, hierarchical=False):
    """max over rows x of sum_(j in level l) |R(x, j)|, R assembled column by
    column: R e_j = oracle recompose of the unit coefficient vector e_j."""
    N = int(np.prod(dims))
    R = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        R[:, j] = ob.recompose(e, dims, L, hierarchical=hierarchical)
    ld = ob.level_dims(dims, L)
    level_of = np.full(dims, -1)
    for l in range(L, -1, -1):
        level_of[tuple(slice(0, int(n)) for n in ld[l])] = l
    level_of = level_of.ravel()
    return np.array([np.max(np.abs(R[:, level_of == l]).sum(axis=1)) for l in range(L + 1)])


@pytest.mark.parametrize("dims,L", [([9, 17, 5], 2), ([5, 9, 9], 2), ([17, 9], 3), ([9, 33], 2), ([33], 4),
                                    ([65], 5), ([1, 17, 9], 2), ([9, 5, 1], 1), ([3, 5, 9], 1)])
def test_amplification_is_dense_row_norm(dims, L):
    want = dense_level_row_norms(dims, L)
    got = ip.recompose_amplification(dims, L)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=0)
    assert np.all(got >= want)  # rounded up: a bound


def test_amplification_hierarchical():
    dims, L = [9, 9, 5], 2
    np.testing.assert_allclose(dense_level_row_norms(dims, L, hierarchical=True), 1.0, rtol=1e-12)
    np.testing.assert_array_equal(ip.recompose_amplification(dims, L, hierarchical=True), np.ones(L + 1))


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_ob_constant_bounds_amplification(nd):
    dims = [513] * nd if nd < 3 else [257] * 3
    a = ip.recompose_amplification(dims, 5)
    c = ip.max_estimator_ob_c(nd, np.float64) / 4  # ProDM's constant before its factor 4
    assert np.max(a) <= c, (a, c)


def vtot_bound(v, E):
    """|V| bound of ProDM's composition at reconstructed values v (n, 3)."""
    v2 = np.sum(v * v, axis=1)
    e2 = np.sum(2 * np.abs(v) * E + E * E, axis=1)
    root = np.sqrt(v2)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(v2 > e2, e2 / (np.sqrt(np.maximum(v2 - e2, 0)) + root), e2 / root)
    return np.where(v2 == 0, np.sqrt(e2), out)


def comp_bounds(amp, tables, idx, L, rounding):
    return np.array([rounding[i] + sum(amp[l] * tables[i * (L + 1) + l][idx[i * (L + 1) + l]] for l in range(L + 1))
                     for i in range(3)])


def qoi_case(seed, L=3, nb=24):
    rng = np.random.default_rng(seed)
    amp = np.concatenate([[1.0], rng.uniform(2.0, 5.5, L)])
    tables, sizes = [], []
    for i in range(3):
        for l in range(L + 1):
            e = int(rng.integers(-3, 2))
            tables.append(np.array([4 * 2.0 ** (e - 1)] + [2.0 ** (e + 2 - k) for k in range(1, nb + 1)]))
            sizes.append([int(rng.integers(100, 4000)) * (l + 1)] * nb)
    sample = rng.standard_normal((500, 3))
    sample[0] = 0.0  # a zero-velocity point: the bound is sqrt(e2) there
    return amp, tables, sizes, sample


@pytest.mark.parametrize("seed", range(4))
def test_qoi_plan_meets_sampled_bound(seed):
    L = 3
    amp, tables, sizes, sample = qoi_case(seed, L)
    rounding = [1e-7, 2e-7, 0.0]
    for tol in (1e-1, 1e-2, 1e-3):
        idx = np.zeros(3 * (L + 1), np.uint8)
        out, idx = ip.qoi_retrieve_sizes(sizes, tables, amp, sample, tol, idx, rounding=rounding)
        E = comp_bounds(amp, tables, idx, L, rounding)
        assert np.max(vtot_bound(sample, E)) < tol
        # bytes returned are the sizes of the planes taken
        for k in range(3 * (L + 1)):
            assert out[k] == sum(sizes[k][:idx[k]])
        assert idx.sum() > 0
        # planes held already are kept and not charged again
        idx2 = idx.copy()
        out2, idx2 = ip.qoi_retrieve_sizes(sizes, tables, amp, sample, tol, idx2, rounding=rounding)
        assert np.all(out2 == 0) and np.all(idx2 == idx)


def test_qoi_plan_greedy_choice_by_hand():
    """One point v = (1, 0, 0), one level (A = 1), tables 2^-k: the bound is
    e2 / (sqrt(1 - e2) + 1), e2 = 2 E_0 + sum E_i^2. Component 0 weighs 2|v_0|
    E_0 and dominates, so its planes are taken first; a plane of component 1
    lowers e2 by E_1^2 - E_1'^2 only. With equal plane sizes the greedy takes
    component 0's planes until its term is below components 1 and 2's
    squares."""
    nb = 12
    amp = np.array([1.0])
    t = np.array([2.0 ** -k for k in range(nb + 1)])
    tables = [t, t, t]
    sizes = [[100] * nb] * 3
    sample = np.array([[1.0, 0.0, 0.0]])
    idx = np.zeros(3, np.uint8)
    _, idx = ip.qoi_retrieve_sizes(sizes, tables, amp, sample, 1e-3, idx)
    E = t[idx]
    e2 = 2 * E[0] + np.sum(E * E)
    assert e2 / (math.sqrt(1 - e2) + 1) < 1e-3
    assert idx[0] >= 10 and idx[1] <= 6 and idx[2] <= 6, idx
