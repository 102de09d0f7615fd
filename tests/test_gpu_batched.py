"""GPU parity of the batched refactor (prodm_mdr_batch): every field's level
coefficients, level maxima and bitplanes are bit-identical to a single-field
plan refactor of the same input, for eager grouped launches, graph replays,
autotuned shapes and cached shapes; the 96-field workload of 2-D / 3-D
fields of uneven shapes is checked in full."""
import os

import numpy as np
import pytest
import torch

import bench
from fields import bumps_field

pytestmark = pytest.mark.gpu

TD = {np.float32: torch.float32, np.float64: torch.float64}

MIXED = [  # (dims, L): 2-D and 3-D fields of uneven shapes
    ([65, 65, 65], 5), ([33, 129], 4), ([129, 33, 65], 4), ([257, 257], 6), ([33, 33, 33], 4),
    ([65, 257], 5), ([129, 129, 33], 4), ([513, 65], 5), ([17, 17, 17], 3), ([65, 33, 129], 4),
    # contiguous lines of 257 and 513 coarse nodes (5 and 9 per lane)
    ([5, 513], 2), ([9, 1025], 3),
]


def level_block(plan, level, nb, elem):
    """the level's plane storage (nb planes of `stride` words) as a tensor,
    plus the words per plane that hold data"""
    from prodm_amd import mdr
    p0, nbytes = plan.plane(level, 0)
    stride = (plan.plane(level, 1)[0] - p0) if nb > 1 else nbytes
    buf = torch.empty(nb * stride, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    mdr._d2d(buf, p0, nb * stride)
    return buf.view(nb, stride)[:, :nbytes], nbytes


def assert_same_results(bplan, pplan, L, nb, elem, encoder, what):
    for l in range(L + 1):
        assert bplan.level_info(l) == pplan.level_info(l), (what, l)
        cb, cp = bplan.coefficients(l), pplan.coefficients(l)
        assert torch.equal(cb.view(torch.uint8), cp.view(torch.uint8)), (what, "coefficients", l)
        if encoder == 0:
            a, na = level_block(bplan, l, nb, elem)
            b, nbb = level_block(pplan, l, nb, elem)
            assert na == nbb and torch.equal(a, b), (what, "planes", l)
        else:
            for p in range(nb):
                assert bplan.plane_bytes(l, p) == pplan.plane_bytes(l, p), (what, l, p)
            assert np.array_equal(bplan.level_rb(l), pplan.level_rb(l)), (what, "rb", l)


def make_inputs(gpu, shapes, dt, seed0):
    return [torch.from_numpy(bumps_field(d, seed=seed0 + i, dtype=dt)).to(gpu) for i, (d, L) in enumerate(shapes)]


def plain_plans(xs, shapes, nb, dt, enc, decomposer=0):
    from prodm_amd import mdr
    plans = []
    for x, (d, L) in zip(xs, shapes):
        p = mdr.Plan(d, L, nb, TD[dt], enc, decomposer)
        p.refactor(x)
        plans.append(p)
    torch.cuda.synchronize()
    return plans


@pytest.mark.parametrize("dt,enc,nb,decomp", [
    (np.float32, 0, 32, 0),
    (np.float32, 1, 32, 0),
    (np.float64, 0, 64, 0),
    (np.float64, 1, 60, 0),
    (np.float32, 0, 32, 1),
])
def test_batched_equals_plain(gpu, dt, enc, nb, decomp):
    from prodm_amd.batched import Batch
    xs = make_inputs(gpu, MIXED, dt, 100)
    plans = plain_plans(xs, MIXED, nb, dt, enc, decomp)
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, MIXED)], nb, TD[dt], enc, decomp)
    b.refactor()
    torch.cuda.synchronize()
    desc = b.describe().strip().splitlines()
    # grouped: far fewer launches than the fields' single-field sequences
    assert len(desc) <= 4 * max(L for _, L in MIXED) + 3, desc
    elem = 4 if dt == np.float32 else 8
    for f, (d, L) in enumerate(MIXED):
        assert_same_results(b.plan(f), plans[f], L, nb, elem, enc, f"field {f} {d}")


@pytest.mark.parametrize("dt,enc,nb", [(np.float32, 0, 32), (np.float64, 1, 60)])
def test_batched_fields_match_oracle(gpu, dt, enc, nb):
    """Every batched field against the oracle directly (not through the plain
    GPU path): level coefficients against the oracle's decompose + interleave
    (1e-6 relative for fp32, 1e-14 for fp64), level maxima and exponents, and
    the bitplane streams bit-exact against the oracle encoders fed the
    batched coefficients. The shapes include lines of 9 and 17 coarse nodes
    (16- and 32-lane groups of the contiguous solve) and strided lines packed
    several groups to a workgroup."""
    import oracle_bind as ob
    from prodm_amd.batched import Batch
    tol = 1e-6 if dt == np.float32 else 1e-14
    us = [bumps_field(d, seed=500 + i, dtype=dt) for i, (d, L) in enumerate(MIXED)]
    b = Batch([(torch.from_numpy(u).to(gpu), d, L) for u, (d, L) in zip(us, MIXED)], nb, TD[dt], enc)
    b.refactor()
    torch.cuda.synchronize()
    for f, ((d, L), u) in enumerate(zip(MIXED, us)):
        plan = b.plan(f)
        dec = ob.decompose(u, d, L)
        ld = ob.level_dims(d, L)
        scale = float(np.max(np.abs(dec)))
        for l in range(L + 1):
            coarse = ld[l - 1] if l else np.zeros(len(d), np.uint32)
            want = ob.interleave(dec, d, ld[l], coarse)
            c = plan.coefficients(l).cpu().numpy()
            assert c.size == want.size
            assert np.max(np.abs(c.astype(np.float64) - want)) <= tol * scale, (f, d, l)
            m, e, n = plan.level_info(l)
            assert m == ob.maxabs(c) and e == ob.frexp_exp(m, dt), (f, l)
            if enc == 0:
                op = ob.nb_encode(c, e, nb)
                for p in range(nb):
                    assert plan.plane_bytes(l, p) == op[p].tobytes(), (f, l, p)
            else:
                os_ = ob.gb_encode(c, e, nb)
                rb = plan.level_rb(l)
                assert os_[0][4:4 + rb.size] == rb.tobytes(), (f, l)
                for p in range(nb):
                    want_p = os_[p][4 + rb.size:] if p == 0 else os_[p]
                    assert plan.plane_bytes(l, p) == want_p, (f, l, p)
    b.close()


def test_batched_graph_autotune_and_cache(gpu, tmp_path):
    from prodm_amd.batched import Batch
    shapes = MIXED[:7]
    xs = make_inputs(gpu, shapes, np.float32, 200)
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], 32)
    s = torch.cuda.Stream()
    cache = str(tmp_path / "shapes.txt")
    with torch.cuda.stream(s):
        b.autotune(cache, reps=2)
        tuned = b.describe()
        # graph replays, with the inputs changed in place between them
        for rep in range(3):
            for i, x in enumerate(xs):
                x.copy_(torch.from_numpy(bumps_field(shapes[i][0], seed=300 + 10 * rep + i, dtype=np.float32)))
            b.refactor_graph()
            s.synchronize()
            plans = plain_plans(xs, shapes, 32, np.float32, 0)
            for f, (d, L) in enumerate(shapes):
                assert_same_results(b.plan(f), plans[f], L, 32, 4, 0, f"replay {rep} field {f}")
    assert os.path.exists(cache)
    text = open(cache).read()
    assert text.startswith("prodm-batch-shapes 1")
    # a new batch over the same shapes reads the cache and launches the same shapes
    b2 = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], 32)
    b2.autotune(cache)
    assert b2.describe() == tuned
    b2.refactor()
    torch.cuda.synchronize()
    for f, (d, L) in enumerate(shapes):
        assert_same_results(b2.plan(f), b.plan(f), L, 32, 4, 0, f"cached field {f}")


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float64, 64)])
def test_batched_capped_grids(gpu, tmp_path, dt, nb):
    """Every grouped launch capped to 40 workgroups through a hand-edited
    shape cache: each workgroup walks many tiles / lines in its persistent
    loop (the load sweep's tile body then runs in a loop, with the per-slot
    set-up redone per tile), and every field stays bit-identical to its
    plain plan."""
    from prodm_amd.batched import Batch
    shapes = MIXED[:7]
    xs = make_inputs(gpu, shapes, dt, 500)
    fields = [(x, d, L) for x, (d, L) in zip(xs, shapes)]
    cache = str(tmp_path / "shapes.txt")
    Batch(fields, nb, TD[dt]).autotune(cache, reps=1)
    lines = []
    for ln in open(cache).read().splitlines():
        if ln.startswith("launch "):
            _, i, a, _cap = ln.split()
            ln = f"launch {i} {a} 40"
        lines.append(ln)
    open(cache, "w").write("\n".join(lines) + "\n")
    b = Batch(fields, nb, TD[dt])
    b.autotune(cache)
    desc = b.describe().strip().splitlines()
    assert desc and all("cap=40" in d for d in desc), desc
    assert any("load" in d and "grid=40" in d for d in desc), desc
    b.refactor()
    torch.cuda.synchronize()
    plans = plain_plans(xs, shapes, nb, dt, 0)
    elem = 4 if dt == np.float32 else 8
    for f, (d, L) in enumerate(shapes):
        assert_same_results(b.plan(f), plans[f], L, nb, elem, 0, f"capped field {f} {d}")


def test_batched_retrieve_through_field_plan(gpu):
    from prodm_amd.batched import Batch
    shapes = MIXED[:4]
    xs = make_inputs(gpu, shapes, np.float32, 400)
    plans = plain_plans(xs, shapes, 32, np.float32, 0)
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], 32)
    b.refactor()
    for f, (d, L) in enumerate(shapes):
        k = [min(32, 6 + 3 * l) for l in range(L + 1)]
        ob = b.plan(f).retrieve(k)
        op = plans[f].retrieve(k)
        torch.cuda.synchronize()
        assert torch.equal(ob, op), f


def test_batched_rejects_bad_fields(gpu):
    from prodm_amd import capi
    from prodm_amd.batched import Batch
    x = torch.zeros(33 * 33, device=gpu)
    with pytest.raises(capi.ProdmError):
        Batch([(x, [33, 33], 0)], 32)          # target level 0
    with pytest.raises(capi.ProdmError):
        Batch([(x, [33, 33], 9)], 32)          # too many levels for the shape
    with pytest.raises(ValueError):
        Batch([(x, [33, 34], 4)], 32)          # element count mismatch
    # the grouped solves hold a line in registers: a 2049-point axis (1025
    # coarse nodes) is refused with a message (the single-field plan takes it)
    y = torch.zeros(3 * 2049, device=gpu)
    b = Batch([(y, [3, 2049], 1)], 32)
    with pytest.raises(capi.ProdmError, match="1024 coarse nodes"):
        b.refactor()  # the launch tables are built (and checked) before anything runs
    b.close()


def test_batched_equals_plain_at_513(gpu):
    """One 513^3 float32 field (6 levels): the plan's finest decompose sweep
    runs 7 x 65-node tiles of 512 threads, the grouped launches the 7 x 33
    tiles of 256; the coefficients, maxima and planes are bit-identical."""
    from prodm_amd import mdr
    from prodm_amd.batched import Batch
    d, L = [513] * 3, 6
    print(f"batched vs plain at {d}, L={L}", flush=True)
    x = bench.make_field_gpu(513)
    b = Batch([(x, d, L)], 32)
    b.refactor()
    p = mdr.Plan(d, L, 32, torch.float32, 0)
    p.refactor(x)
    torch.cuda.synchronize()
    assert_same_results(b.plan(0), p, L, 32, 4, 0, "513^3")
    p.close()
    b.close()


def test_batched_full_workload(gpu):
    """96 fields (47 2-D, 49 3-D), 203 M points: bit-identical to plain runs"""
    from prodm_amd import mdr
    from prodm_amd.batched import Batch
    shapes = bench.batch_shapes()
    assert len(shapes) == 96
    xs = [bench.make_batch_field_gpu(d, 1000 + i) for i, (d, L) in enumerate(shapes)]
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], 32)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.refactor_graph()
    s.synchronize()
    for f, (d, L) in enumerate(shapes):
        p = mdr.Plan(d, L, 32, torch.float32, 0)
        p.refactor(xs[f])
        torch.cuda.synchronize()
        assert_same_results(b.plan(f), p, L, 32, 4, 0, f"field {f} {d}")
        p.close()
