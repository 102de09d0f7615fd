"""Graph capture / replay of the refactor and retrieve sequences: every
replay, with eager work on the same plan and other kernels between replays,
reproduces the eager results bit for bit (plane streams and fields)."""
import numpy as np
import pytest
import torch

from fields import bumps_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("enc", [0, 1])
def test_graph_replays_match_eager(gpu, enc):
    from prodm_amd import mdr
    dims, L, nb = [65, 65, 65], 4, 32
    u = torch.from_numpy(bumps_field(dims, seed=11)).to(gpu)
    v = torch.from_numpy(bumps_field(dims, seed=12)).to(gpu)
    ks = [32, 30, 24, 18, 12]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan = mdr.Plan(dims, L, nb, torch.float32, enc)
        plan.refactor(u)
        want_planes = [plan.plane_stream(l, p) for l in range(L + 1) for p in (0, 7, nb - 1)]
        want = plan.retrieve(ks).clone()
        out = torch.empty_like(u)
        for rep in range(4):
            # eager work on the same plan (another field) between replays
            plan.refactor(v)
            plan.retrieve([nb] * (L + 1))
            torch.cuda.synchronize()
            plan.refactor_graph(u)
            got_planes = [plan.plane_stream(l, p) for l in range(L + 1) for p in (0, 7, nb - 1)]
            assert got_planes == want_planes, rep
            plan.retrieve_graph(ks, out)
            torch.cuda.synchronize()
            assert torch.equal(out, want), rep
            out.mul_(0.5)  # unrelated eager kernel between replays
        plan.close()


def test_graph_needs_a_stream(gpu):
    from prodm_amd import capi, mdr
    dims, L, nb = [33, 33, 33], 3, 32
    u = torch.zeros(33 ** 3, device=gpu)
    plan = mdr.Plan(dims, L, nb, torch.float32, 0)
    with pytest.raises(capi.ProdmError):
        plan.refactor_graph(u)  # default stream: capture is refused
    plan.close()
