"""QoI-guided retrieval of |V| from three refactored velocity components:
the plain halving loop (ProDM's app/GE/sequential/halving_Vtot.cpp) and the
adjoint-guided planner both meet the QoI tolerance against the original
field; the bytes each reads are printed."""
import json

import pytest
import torch

pytestmark = pytest.mark.gpu


def test_qoi_retrieval_meets_tolerance(gpu):
    from prodm_amd import qoi
    res = qoi.run(n=65, L=4, nb=32, rel=1e-3)
    print(json.dumps(res))
    for k in ("plain", "adjoint_guided"):
        assert res[k]["max_error"] <= res[k]["tau"], (k, res[k])
    assert res["adjoint_guided"]["bytes"] > 0 and res["plain"]["bytes"] > 0


def test_velocity_field_is_divergence_free(gpu):
    """Every mode's amplitude is perpendicular to its wavevector (so the
    divergence vanishes identically), and the field equals the mode sum
    evaluated independently at random nodes."""
    import numpy as np
    from prodm_amd import qoi
    ks, amps, ph = qoi.velocity_modes()
    assert np.all(np.linalg.norm(ks, axis=1) <= 32) and np.all(np.abs(np.sum(ks * amps, axis=1)) < 1e-12)
    n = 33
    V = [v.cpu().numpy() for v in qoi.velocity_field(n, device=gpu)]
    rng = np.random.default_rng(0)
    for _ in range(20):
        i, j, k = rng.integers(0, n, 3)
        x = np.array([i, j, k]) / (n - 1)
        want = (amps * np.sin(2 * np.pi * ks @ x + ph)[:, None]).sum(axis=0)
        got = [V[c][(i * n + j) * n + k] for c in range(3)]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
