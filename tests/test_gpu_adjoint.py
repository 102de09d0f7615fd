"""GPU adjoint (transpose) of recompose: against the oracle's adjoint (which
the CPU suite pins to the dense transpose), the dot-product identity, a dense
transpose assembled from GPU recomposes, 64 random probes, and central
finite differences (h = 1e-3, fp64) through the autograd wrapper."""
import numpy as np
import pytest
import torch

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,L", [([33, 17, 65], 3), ([65, 33], 4), ([129], 6), ([1, 17, 33], 3),
                                    # lines past the register-resident solves (the sequential solve)
                                    ([3, 2049], 1), ([2049, 5], 2), ([1153, 3, 5], 1)])
def test_adjoint_matches_oracle_f64(gpu, dims, L):
    from prodm_amd import mdr
    rng = np.random.default_rng(1)
    g = rng.standard_normal(int(np.prod(dims)))
    got = mdr.recompose_adjoint(torch.from_numpy(g).to(gpu), dims, L).cpu().numpy()
    want = ob.recompose_adjoint(g, dims, L)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(g).sum() / np.sqrt(g.size))


def test_adjoint_f32_matches_oracle(gpu):
    from prodm_amd import mdr
    dims, L = [33, 33, 33], 3
    rng = np.random.default_rng(2)
    g = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    got = mdr.recompose_adjoint(torch.from_numpy(g).to(gpu), dims, L).cpu().numpy().astype(np.float64)
    want = ob.recompose_adjoint(g.astype(np.float64), dims, L)
    assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))


@pytest.mark.parametrize("dims,L", [([65, 65, 65], 4), ([5, 2049], 1)])
def test_dot_product_identity(gpu, dims, L):
    """<R x, y> = <x, R^T y> for the GPU recompose R."""
    from prodm_amd import mdr
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).to(gpu)
    y = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).to(gpu)
    lhs = float(torch.dot(mdr.recompose(x, dims, L), y))
    rhs = float(torch.dot(x, mdr.recompose_adjoint(y, dims, L)))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + 1.0)


def test_dense_transpose_small(gpu):
    """Assemble R column by column from GPU recomposes on a 9x9x9 grid and
    compare R^T with the adjoint applied to every unit vector."""
    from prodm_amd import mdr
    dims, L = [9, 9, 9], 2
    N = int(np.prod(dims))
    eye = torch.eye(N, dtype=torch.float64, device=gpu)
    R = torch.stack([mdr.recompose(eye[j].contiguous(), dims, L) for j in range(N)], dim=1)
    RT = torch.stack([mdr.recompose_adjoint(eye[j].contiguous(), dims, L) for j in range(N)], dim=1)
    assert float(torch.max(torch.abs(R.T - RT))) <= 1e-12


def test_random_probes_against_oracle(gpu):
    """64 random probe directions on a 33^3 grid: the GPU adjoint against the
    CPU transpose (oracle adjoint)."""
    from prodm_amd import mdr
    dims, L = [33, 33, 33], 4
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(64):
        d = rng.standard_normal(int(np.prod(dims)))
        got = mdr.recompose_adjoint(torch.from_numpy(d).to(gpu), dims, L).cpu().numpy()
        want = ob.recompose_adjoint(d, dims, L)
        worst = max(worst, float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    assert worst <= 1e-12, worst


def test_autograd_finite_differences(gpu):
    """J(c) = sum(sin(recompose(c))): the autograd gradient against central
    differences with h = 1e-3 in fp64 along 64 random directions."""
    from prodm_amd import autograd, mdr
    dims, L = [17, 17, 17], 3
    rng = np.random.default_rng(5)
    u = torch.from_numpy(rng.standard_normal(int(np.prod(dims)))).to(gpu)
    c0 = mdr.decompose(u, dims, L).requires_grad_(True)
    J = torch.sin(autograd.recompose(c0, dims, L)).sum()
    J.backward()
    grad = c0.grad.detach()
    h = 1e-3
    with torch.no_grad():
        for _ in range(64):
            d = torch.from_numpy(rng.standard_normal(c0.numel())).to(gpu)
            jp = torch.sin(mdr.recompose(c0 + h * d, dims, L)).sum()
            jm = torch.sin(mdr.recompose(c0 - h * d, dims, L)).sum()
            fd = float((jp - jm) / (2 * h))
            an = float(torch.dot(grad, d))
            # central-difference truncation: the third derivative of sin is
            # -cos, at most 1 per point, so the error is at most
            # h^2 / 6 * sum |R d|^3 (plus rounding)
            rd = mdr.recompose(d, dims, L)
            bound = h * h / 6 * float(torch.sum(torch.abs(rd) ** 3)) + 1e-9 * (abs(an) + 1.0)
            assert abs(fd - an) <= bound, (fd, an, bound)
