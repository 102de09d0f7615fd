#!/usr/bin/env python3
"""Refactor + retrieve throughput of the MDR hot path on one MI355X.

Workload (one "step"): refactor a 513^3 float32 field (6 levels, 32 bitplanes,
NegaBinary encoder) and retrieve it at relative L-inf 1e-4 (planes per level
chosen by MaxErrorEstimatorOB(3) + greedy interpreter from the level error
tables), inputs resident in HBM. GPU stages only: lossless compression and
file I/O stay on the host and are outside the timed region.

value = 2 * input bytes / (t_refactor + t_retrieve)  [GB/s]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_field_gpu(n, seed=0, nbumps=16, noise=1e-4):
    """sin(2pi 3x) cos(2pi 5y) sin(2pi 2z) + Gaussian bumps + N(0, noise) on
    the unit cube (x along the slowest dim); numpy PCG64 draws A, w, c, noise
    in that order, evaluated in fp64 on the GPU, returned as float32."""
    import torch
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.5, 2.0, nbumps)
    w = rng.uniform(0.02, 0.1, nbumps)
    c = rng.uniform(0.0, 1.0, (nbumps, 3))
    noise_h = rng.normal(0.0, noise, n * n * n)
    dev = torch.device("cuda")
    t = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=dev)
    x, y, z = t.view(n, 1, 1), t.view(1, n, 1), t.view(1, 1, n)
    u = torch.sin(2 * math.pi * 3 * x) * torch.cos(2 * math.pi * 5 * y) * torch.sin(2 * math.pi * 2 * z)
    for b in range(nbumps):
        gx = torch.exp(-((x - c[b, 0]) ** 2) / (2 * w[b] ** 2))
        gy = torch.exp(-((y - c[b, 1]) ** 2) / (2 * w[b] ** 2))
        gz = torch.exp(-((z - c[b, 2]) ** 2) / (2 * w[b] ** 2))
        u = u + A[b] * (gx * gy * gz)
    u = u + torch.from_numpy(noise_h).to(dev).view(n, n, n)
    return u.to(torch.float32).contiguous().view(-1)


def batch_shapes(nfields=96, seed=3):
    """Shapes of the batched workload: per field ndim in {2, 3}, each dim
    2^k + 1 with k ~ U{5..9}, redrawn while the field has more than 2^25
    points; L = min(6, k_min - 1), at least 4. numpy PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nfields):
        nd = int(rng.integers(2, 4))
        while True:
            ks = [int(k) for k in rng.integers(5, 10, nd)]
            dims = [2 ** k + 1 for k in ks]
            if int(np.prod(dims)) <= 2 ** 25:
                break
        out.append((dims, max(4, min(6, min(ks) - 1))))
    return out


def make_batch_field_gpu(dims, seed, nbumps=8, noise=1e-3):
    """The 129^3 workload's generator on any 2-D / 3-D box: sin(2pi 3x)
    cos(2pi 5y) [sin(2pi 2z)] + 8 Gaussian bumps (A, w, c drawn by numpy
    PCG64 in that order) + N(0, noise) from torch's generator seeded with
    `seed`; fp64 on the GPU, returned as dense float32."""
    import torch
    rng = np.random.default_rng(seed)
    nd = len(dims)
    A = rng.uniform(0.5, 2.0, nbumps)
    w = rng.uniform(0.02, 0.1, nbumps)
    c = rng.uniform(0.0, 1.0, (nbumps, nd))
    dev = torch.device("cuda")
    ax = []
    for d, n in enumerate(dims):
        shape = [1] * nd
        shape[d] = n
        ax.append(torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=dev).view(shape))
    u = torch.sin(2 * math.pi * 3 * ax[0]) * torch.cos(2 * math.pi * 5 * ax[1])
    if nd > 2:
        u = u * torch.sin(2 * math.pi * 2 * ax[2])
    for b in range(nbumps):
        g = None
        for d in range(nd):
            t = torch.exp(-((ax[d] - c[b, d]) ** 2) / (2 * w[b] ** 2))
            g = t if g is None else g * t
        u = u + A[b] * g
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    u = u + noise * torch.randn(tuple(dims), dtype=torch.float64, device=dev, generator=gen)
    return u.to(torch.float32).contiguous().view(-1)


def batched_report(steps, warmup, nb=32):
    """The 96-field batched workload: per-field plan refactors (eager and
    per-plan graphs) against the batched refactor (default shapes, autotuned
    shapes, graph replay). Wall clock over `steps` runs on one non-default
    stream, inputs resident; finest-level planes and all level maxima of
    every field are checked bit-identical between the plain and batched runs."""
    import torch
    from prodm_amd import capi, mdr
    from prodm_amd.batched import Batch
    shapes = batch_shapes()
    xs = [make_batch_field_gpu(d, 1000 + i) for i, (d, L) in enumerate(shapes)]
    pts = sum(int(np.prod(d)) for d, L in shapes)
    in_bytes = 4 * pts
    plans = [mdr.Plan(d, L, nb, torch.float32, capi.NEGABINARY) for d, L in shapes]
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], nb)
    s = torch.cuda.Stream()

    def timed(fn):
        for _ in range(max(1, warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    def plain():
        for p, x in zip(plans, xs):
            p.refactor(x)

    def plain_graph():
        for p, x in zip(plans, xs):
            p.refactor_graph(x)

    with torch.cuda.stream(s):
        t_plain = timed(plain)
        t_plain_graph = timed(plain_graph)
        t_default = timed(b.refactor)
        t0 = time.perf_counter()
        b.autotune(None, 3)
        t_tune = time.perf_counter() - t0
        t_tuned = timed(b.refactor)
        t_graph = timed(b.refactor_graph)
        plain()
        b.refactor_graph()
    torch.cuda.synchronize()
    same = True
    for f, (d, L) in enumerate(shapes):
        bp = b.plan(f)
        for l in range(L + 1):
            same &= bp.level_info(l) == plans[f].level_info(l)
        a, na = bp.plane(L, 0)
        c, nc = plans[f].plane(L, 0)
        stride = bp.plane(L, 1)[0] - a
        ta = torch.empty(nb * stride, dtype=torch.uint8, device="cuda")
        tc = torch.empty_like(ta)
        mdr._d2d(ta, a, nb * stride)
        mdr._d2d(tc, c, nb * stride)
        same &= na == nc and torch.equal(ta.view(nb, stride)[:, :na], tc.view(nb, stride)[:, :nc])
    launches = b.describe().strip().splitlines()
    best = min(t_tuned, t_graph)
    res = {
        "workload": f"{len(shapes)} float32 fields ({sum(len(d) == 2 for d, _ in shapes)} 2-D, "
                    f"{sum(len(d) == 3 for d, _ in shapes)} 3-D), dims 2^k+1 with k ~ U{{5..9}} (<= 2^25 points "
                    f"each), L = min(6, k_min - 1) >= 4, {nb} bitplanes, NegaBinary; config-0 generator, "
                    f"seeds 1000+f; refactor only",
        "fields": len(shapes), "points": pts, "input_bytes": in_bytes,
        "plain_ms": round(t_plain, 4), "plain_graph_ms": round(t_plain_graph, 4),
        "batched_default_ms": round(t_default, 4), "batched_tuned_ms": round(t_tuned, 4),
        "batched_graph_ms": round(t_graph, 4), "autotune_s": round(t_tune, 2),
        "plain_GBps": round(in_bytes / (t_plain * 1e6), 2), "batched_GBps": round(in_bytes / (best * 1e6), 2),
        "speedup_vs_plain": round(t_plain / best, 3), "speedup_vs_plain_graph": round(t_plain_graph / best, 3),
        "grouped_launches": len(launches), "bit_identical": bool(same),
    }
    log(f"[bench] batched: plain {t_plain:.3f} ms (graphs {t_plain_graph:.3f}), batched default {t_default:.3f}, "
        f"tuned {t_tuned:.3f}, graph {t_graph:.3f} ms; {len(launches)} grouped launches; identical={same}")
    for p in plans:
        p.close()
    b.close()
    return res


def perbit_report(steps, warmup, n=513, L=6, nb=32):
    """PerBit encode (PerBitBPEncoder.hpp:81-143) of the finest level of the
    n^3 workload under the hierarchical decomposer (the pairing of ProDM's GE
    applications): the level's coefficients are resident in HBM; a step is one
    encode of all nb streams (count, scan, write and boundary passes), timed
    with HIP events on the launch stream. Bytes = coefficients read + stream
    bytes written."""
    import torch
    from prodm_amd import capi, mdr
    u = make_field_gpu(n, seed=0)
    plan = mdr.Plan([n, n, n], L, nb, torch.float32, capi.NEGABINARY, capi.HIERARCHICAL)
    plan.refactor(u)
    x = plan.coefficients(L).contiguous()
    _, e, _ = plan.level_info(L)  # the level's exponent (frexp of its max)
    plan.close()
    del u
    m = x.numel()
    stride = int(capi.lib().prodm_perbit_stream_words(m))
    streams = torch.empty((nb, stride), dtype=torch.int64, device="cuda")
    sizes = torch.zeros(nb, dtype=torch.int32, device="cuda")
    ws = torch.empty(int(capi.lib().prodm_perbit_workspace_bytes(m, nb, capi.F32)), dtype=torch.uint8, device="cuda")
    s = mdr._stream()

    def enc():
        capi.check(capi.lib().prodm_perbit_encode_f32(mdr._ptr(x), m, e, nb, 0, mdr._ptr(streams), stride,
                                                      mdr._ptr(sizes), mdr._ptr(ws), ws.numel(), s), "perbit encode")
    for _ in range(max(warmup, 1)):
        enc()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        enc()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    out_bytes = int(sizes.to(torch.int64).sum().item())
    total = 4 * m + out_bytes
    gbs = total / (ms * 1e6)
    log(f"[bench] perbit: {m} coefficients, {ms:.3f} ms per encode, {gbs:.1f} GB/s (in + out)")
    return {"workload": f"PerBit encode of the finest level ({m} float32 coefficients) of the {n}^3 field, "
                        f"hierarchical decomposer, {L} levels, {nb} planes",
            "elements": m, "ms": round(ms, 4), "in_bytes": 4 * m, "stream_bytes": out_bytes,
            "GBps": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}


def end_to_end_report(n=513, L=6, nb=32, reps=3, tols=(1e-2, 1e-4, 1e-6)):
    """Host buffers and files end to end, with the C++ applications over the
    library: prodm_refactor (upload, GPU stages, plane streams copied out
    overlapped with host ZSTD, ProDM's files) and prodm_reconstruct (plan,
    read, decompress, upload, decode + recompose, download, error check) on
    the n^3 workload written as a raw file in a temporary directory. Reports
    the best of `reps` refactors (each app run is its own process)."""
    import re
    import subprocess
    import tempfile
    here = os.path.dirname(os.path.abspath(__file__))
    binr = os.path.join(here, "prodm_amd", "bin")
    u = make_field_gpu(n, seed=0).cpu().numpy()
    out = {"workload": f"{n}^3 float32 raw file -> ProDM files ({L} levels, {nb} planes, NegaBinary, "
                       f"AdaptiveLevelCompressor(64)) -> progressive retrieval", "refactor": [], "reconstruct": []}
    with tempfile.TemporaryDirectory() as tmp:
        raw, rd = os.path.join(tmp, "field.raw"), os.path.join(tmp, "rd")
        u.tofile(raw)
        for _ in range(reps):
            r = subprocess.run([os.path.join(binr, "prodm_refactor"), raw, str(L), str(nb), "3", str(n), str(n), str(n),
                                "--out", rd], capture_output=True, text=True, timeout=300)
            if r.returncode:
                raise RuntimeError(r.stdout + r.stderr)
            st = [float(v) for v in re.findall(r"([0-9.]+)s", r.stdout)]
            out["refactor"].append({"device_s": st[0], "copyout_lossless_s": st[1], "files_s": st[2], "total_s": st[3]})
        files = sum(os.path.getsize(os.path.join(rd, f)) for f in os.listdir(rd))
        r = subprocess.run([os.path.join(binr, "prodm_reconstruct"), raw, str(len(tols)), *[f"{t:g}" for t in tols],
                            "--in", rd], capture_output=True, text=True, timeout=300)
        if r.returncode:
            raise RuntimeError(r.stdout + r.stderr)
        for line in r.stdout.splitlines():
            m = re.search(r"tolerance ([0-9.e+-]+): reconstruct ([0-9.]+)s, retrieved ([0-9]+) bytes.*rel ([0-9.e+-]+)\)",
                          line)
            if m:
                out["reconstruct"].append({"tolerance": float(m.group(1)), "s": float(m.group(2)),
                                           "file_bytes": int(m.group(3)), "rel_linf": float(m.group(4))})
    best = min(out["refactor"], key=lambda d: d["total_s"])
    out["file_bytes"] = files
    out["refactor_GBps"] = round(u.nbytes / best["total_s"] / 1e9, 3)
    log(f"[bench] end to end: refactor {best['total_s']:.3f} s ({out['refactor_GBps']} GB/s of input), "
        f"reconstruct {[d['s'] for d in out['reconstruct']]} s")
    return out


def qoi_report(n=257, L=5, nb=32, rel=1e-3):
    """The adjoint-guided QoI workload: |V| of a divergence-free n^3 float32
    velocity (24 Fourier modes, |k| <= 32, seeds 10, 11, 12), L levels, nb
    planes, NegaBinary, |V| L-inf tolerance rel * range(|V|). Plane bytes
    read by ProDM's halving loop (halving_Vtot.cpp) and by the adjoint-guided
    planner, each checked against the original field (bound over every node
    and actual error at most tau)."""
    from prodm_amd import qoi
    log(f"[bench] QoI: |V| of a {n}^3 velocity, L={L}, nb={nb}, rel {rel:g} ...")
    t0 = time.perf_counter()
    res = qoi.run(n, L, nb, rel)
    res["seconds"] = round(time.perf_counter() - t0, 2)
    for k in ("plain", "adjoint_guided"):
        r = res[k]
        if not (r["max_error"] <= r["tau"] and r["max_bound"] <= r["tau"]):
            log(f"[bench] FAIL: QoI {k} misses tau: {r}")
            sys.exit(1)
    log(f"[bench] QoI bytes: plain {res['plain']['bytes']}, adjoint-guided {res['adjoint_guided']['bytes']} "
        f"({100 * res['bytes_reduction']:.1f}% fewer)")
    return res


def plan_planes(level_bounds, level_plane_bytes, nb, tol):
    """Planes per level for an L-infinity tolerance: the host C++ greedy
    planner (GreedyBasedSizeInterpreter.hpp:23-70) over the MaxErrorCollector
    tables (MaxErrorCollector.hpp:14-25) with MaxErrorEstimatorOB(3)
    (MaxErrorEstimator.hpp:13-40), as ComposedReconstructor plans a first
    retrieval (ComposedReconstructor.hpp:33-53)."""
    import ctypes as C
    from prodm_amd import capi, interpret
    tables = []
    for b in level_bounds:
        t = np.zeros(nb + 1)
        capi.lib().prodm_collect_max_f32(C.c_float(b), nb, t.ctypes.data_as(C.c_void_p))
        tables.append(t)
    sizes = [[pb] * nb for pb in level_plane_bytes]
    idx = np.zeros(len(level_bounds), np.uint8)
    _, idx = interpret.retrieve_sizes(interpret.GREEDY, interpret.MAX_OB, sizes, tables, tol, idx, ndims=3)
    c = interpret.max_estimator_ob_c(3)
    est = sum(float(np.float32(c) * np.float32(tables[l][idx[l]])) for l in range(len(level_bounds)))
    return [int(k) for k in idx], est


def measured_traffic(stage, n, L, nb):
    """HBM bytes per launch of a stage's kernel as measured by rocprofv3
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy-kernel calibration),
    recorded in profiles/hbm_traffic_513_f32_pmc.md for the 513^3 workload;
    None for other workloads or when absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_513_f32_pmc.md")
    if (n, L, nb) != (513, 6, 32) or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes" and f[1] == stage:
            return int(f[2])
    return None


def calibrate_copy(mib=512, reps=3):
    """A plain copy kernel of known traffic (torch elementwise y = x + 0 over
    `mib` MiB of fp32: reads and writes `mib` MiB per launch), for scaling
    the FETCH_SIZE / WRITE_SIZE counters of the MDR kernels."""
    import torch
    x = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": mib * 2 ** 20,
                      "bytes_written_per_launch": mib * 2 ** 20, "launches": reps}))


def cpu_baseline(u_host, dims, L, nb, planes, all_cores=False):
    """The same refactor + retrieve on the host cores. Single core: ProDM's
    own code, compiled from its headers (oracle/_ref/libref.so), for the
    stages it holds (interleave, max, NegaBinary encode and decode,
    reposition), and the single-threaded restatement (oracle/liboracle.so)
    for the MGARD decompose / recompose, whose arithmetic ProDM takes from the
    absent MGARDx. all_cores: the restatement built with OpenMP over
    independent loops (oracle/liboracle_mt.so, identical results) for every
    stage. Returns (refactor s, retrieve s, field, per-stage code)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    prev = ob.use_library(ob.ORACLE_MT_SO) if all_cores else None
    lib = "ref" if (not all_cores and ob.ref_available()) else "oracle"
    code = {"decompose/recompose": "restatement (OpenMP)" if all_cores else "restatement",
            "interleave/max/encode/decode/reposition": "ProDM (compiled reference)" if lib == "ref" else
            ("restatement (OpenMP)" if all_cores else "restatement")}
    try:
        t0 = time.perf_counter()
        dec = ob.decompose(u_host, dims, L)
        ld = ob.level_dims(dims, L)
        streams, exps, bufs = [], [], []
        for l in range(L + 1):
            coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
            buf = ob.interleave(dec, dims, ld[l], coarse, lib=lib)
            e = ob.frexp_exp(ob.maxabs(buf, lib=lib), np.float32)
            streams.append(ob.nb_encode(buf, e, nb, lib=lib))
            exps.append(e)
            bufs.append(buf.size)
        t1 = time.perf_counter()
        rec = np.zeros_like(u_host)
        for l in range(L + 1):
            coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
            vals = ob.nb_decode(streams[l], bufs[l], exps[l], 0, planes[l], np.float32, lib=lib)
            ob.reposition(vals, rec, dims, ld[l], coarse, lib=lib)
        rec = ob.recompose(rec, dims, L)
        t2 = time.perf_counter()
    finally:
        if prev is not None:
            ob.restore_library(prev)
    return t1 - t0, t2 - t1, rec, code


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=513)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--bitplanes", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true", help="time HIP graph replays (the default)")
    ap.add_argument("--eager", action="store_true", help="time eager launches instead of HIP graph replays")
    ap.add_argument("--no-batched", action="store_true", help="skip the 96-field batched refactor report")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched report (one JSON line)")
    ap.add_argument("--no-perbit", action="store_true", help="skip the PerBit encode report")
    ap.add_argument("--perbit-only", action="store_true", help="run only the PerBit encode report (one JSON line)")
    ap.add_argument("--no-qoi", action="store_true", help="skip the adjoint-guided QoI retrieval report")
    ap.add_argument("--qoi-only", action="store_true", help="run only the QoI retrieval report (one JSON line)")
    ap.add_argument("--end-to-end", action="store_true",
                    help="run only the host-buffer + files report of the C++ applications (one JSON line)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a plain copy kernel of known traffic (for counter scaling) and exit")
    args = ap.parse_args()
    args.graph = not args.eager  # graph replays unless --eager
    if args.calibrate:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the refactor/retrieve path runs on a "
              "single GPU and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import torch
    from prodm_amd import capi, mdr

    if args.batched_only:
        print(json.dumps(batched_report(args.steps, args.warmup)), flush=True)
        return
    if args.perbit_only:
        print(json.dumps(perbit_report(args.steps, args.warmup)), flush=True)
        return
    if args.end_to_end:
        print(json.dumps(end_to_end_report()), flush=True)
        return
    if args.qoi_only:
        print(json.dumps(qoi_report()), flush=True)
        return

    n, L, nb = args.size, args.levels, args.bitplanes
    dims = [n, n, n]
    N = n ** 3
    in_bytes = 4 * N
    log(f"[bench] device {torch.cuda.get_device_name(0)}; field {n}^3 f32, L={L}, nb={nb}")
    u = make_field_gpu(n)
    vmin, vmax = torch.aminmax(u)
    vrange = float(vmax - vmin)
    tau = args.tol * vrange
    plan = mdr.Plan(dims, L, nb, torch.float32, capi.NEGABINARY)
    out = torch.empty_like(u)
    log(f"[bench] plan holds {plan.device_bytes / 2**30:.2f} GiB on the device")

    # retrieval plan from the level error bounds (4 * max for negabinary)
    plan.refactor(u)
    torch.cuda.synchronize()
    bounds, plane_bytes = [], []
    for l in range(L + 1):
        m, e, el = plan.level_info(l)
        bounds.append(float(np.float32(m) * np.float32(4.0)))
        plane_bytes.append(((el + 31) // 32) * 4)
    planes, est = plan_planes(bounds, plane_bytes, nb, tau)
    log(f"[bench] tau={tau:.4e} planes per level {planes} (estimated error {est:.4e})")
    plan.retrieve(planes, out)
    torch.cuda.synchronize()
    err = float(torch.max(torch.abs(out.double() - u.double())))
    log(f"[bench] retrieval max error {err:.4e} (bound {tau:.4e})")
    if not err <= tau:
        log("[bench] FAIL: requested tolerance not met")
        sys.exit(1)
    retrieved_bytes = sum(planes[l] * plane_bytes[l] for l in range(L + 1))

    # warmup
    for _ in range(args.warmup):
        plan.refactor(u)
        plan.retrieve(planes, out)
    torch.cuda.synchronize()

    # timed region: exactly K steps of refactor + retrieve, nothing else on
    # the stream, bracketed by synchronizations. Eager launches by default;
    # --graph replays the same kernels from captured HIP graphs.
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        # warm the stream itself (first launches on a new stream pay set-up)
        for _ in range(max(1, args.warmup)):
            if args.graph:
                plan.refactor_graph(u)
                plan.retrieve_graph(planes, out)
            else:
                plan.refactor(u)
                plan.retrieve(planes, out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if (not args.graph):
            for _ in range(args.steps):
                plan.refactor(u)
                plan.retrieve(planes, out)
        else:
            for _ in range(args.steps):
                plan.refactor_graph(u)
                plan.retrieve_graph(planes, out)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = 2 * in_bytes / (ms_step * 1e6)
    err_graph = float(torch.max(torch.abs(out.double() - u.double())))
    if not err_graph <= tau:
        log("[bench] FAIL: timed retrieval misses the tolerance")
        sys.exit(1)
    s = torch.cuda.current_stream()
    # the other mode, for the record (warmed first: the first graph call
    # captures and instantiates the graphs)
    with torch.cuda.stream(gs):
        for _ in range(max(1, args.warmup)):
            if (not args.graph):
                plan.refactor_graph(u)
                plan.retrieve_graph(planes, out)
            else:
                plan.refactor(u)
                plan.retrieve(planes, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(gs):
        for _ in range(args.steps):
            if (not args.graph):
                plan.refactor_graph(u)
                plan.retrieve_graph(planes, out)
            else:
                plan.refactor(u)
                plan.retrieve(planes, out)
        torch.cuda.synchronize()
    other_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # refactor / retrieve split: K more steps with one event between the two
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for k in range(args.steps):
        ev[k][0].record(s)
        plan.refactor(u)
        ev[k][1].record(s)
        plan.retrieve(planes, out)
        ev[k][2].record(s)
    torch.cuda.synchronize()
    t_ref = sum(a.elapsed_time(b) for a, b, _ in ev) / args.steps
    t_ret = sum(b.elapsed_time(c) for _, b, c in ev) / args.steps

    # per-stage kernel times: K more steps with HIP events recorded around
    # each kernel group on the launch stream (the events' own cost falls
    # between kernels, outside the measured spans)
    plan.profile(True)
    for _ in range(args.steps):
        plan.refactor(u)
        plan.retrieve(planes, out)
    stages = plan.stage_times()
    plan.profile(False)
    gbs_ref = in_bytes / (t_ref * 1e6)
    gbs_ret = in_bytes / (t_ret * 1e6)
    log(f"[bench] {ms_step:.3f} ms/step -> {value:.1f} GB/s; refactor {t_ref:.3f} ms ({gbs_ref:.1f} GB/s), "
        f"retrieve {t_ret:.3f} ms ({gbs_ret:.1f} GB/s)")

    # algorithmic bytes (refactor: decompose 2 s n_l per level + encoder read s N
    # + planes written; retrieve: planes read + s N written + recompose 2 s n_l)
    ld, el, _ = mdr.level_geometry(dims, L)
    pts = [int(np.prod(ld[l].astype(np.int64))) for l in range(L + 1)]
    P = sum(nb * plane_bytes[l] for l in range(L + 1))
    b_ref = sum(2 * 4 * pts[l] for l in range(1, L + 1)) + 4 * N + P
    b_ret = retrieved_bytes + 4 * N + sum(2 * 4 * pts[l] for l in range(1, L + 1))
    # per-kernel roofline: algorithmic bytes of one launch at the finest level
    eltop = int(el[L])
    kern = {
        "decompose_load_top": ("k_load<float,decompose>", 2 * 4 * pts[L]),
        "encode_top": ("k_nb_encode<float,32>", 4 * eltop + nb * plane_bytes[L]),
        "decode_top": ("k_nb_decode<float>", planes[L] * plane_bytes[L] + 4 * eltop),
        "recompose_load_top": ("k_load<float,recompose>", 4 * eltop),
        "recompose_interp_top": ("k_interp<float>", 4 * eltop + 4 * pts[L - 1] + 4 * pts[L]),
    }
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in stages.items()}
    dom = max(kern, key=lambda k: stage_ms.get(k, 0.0))
    dname, dbytes = kern[dom]
    dms = stage_ms[dom]
    achieved = dbytes / (dms * 1e6)
    traffic = measured_traffic(dom, n, L, nb)
    roof = {"bound": "hbm", "kernel": dname, "stage": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": dbytes, "avg_launch_ms": round(dms, 5)}
    per_kernel = {k: {"avg_ms": round(stage_ms[k], 5), "algorithmic_GBps": round(v[1] / (stage_ms[k] * 1e6), 1)}
                  for k, v in kern.items() if stage_ms.get(k, 0) > 0}

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline: one core of the host, same field ...")
        u_host = u.cpu().numpy()
        tr, tt, rec, code = cpu_baseline(u_host, dims, L, nb, planes)
        cerr = float(np.max(np.abs(rec.astype(np.float64) - u_host)))
        cpu_gbs = 2 * in_bytes / ((tr + tt) * 1e9)
        cpu = {"value": round(cpu_gbs, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"the full {n}^3 field, one refactor + one retrieve with the same planes per level "
                         f"(refactor {tr:.2f} s, retrieve {tt:.2f} s, max error {cerr:.3e})",
               "code_per_stage": code}
        log(f"[bench] CPU 1 core: refactor {tr:.2f} s, retrieve {tt:.2f} s -> {cpu_gbs:.3f} GB/s")
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        tr, tt, rec, code = cpu_baseline(u_host, dims, L, nb, planes, all_cores=True)
        cerr = float(np.max(np.abs(rec.astype(np.float64) - u_host)))
        cpu_gbs = 2 * in_bytes / ((tr + tt) * 1e9)
        cpu_all = {"value": round(cpu_gbs, 4), "unit": "GB/s", "cores": threads, "kind": "port",
                   "sample": f"the full {n}^3 field, one refactor + one retrieve (refactor {tr:.2f} s, retrieve "
                             f"{tt:.2f} s, max error {cerr:.3e}); os.cpu_count() = {os.cpu_count()}",
                   "code_per_stage": code}
        log(f"[bench] CPU {threads} threads: refactor {tr:.2f} s, retrieve {tt:.2f} s -> {cpu_gbs:.3f} GB/s")

    res = {
        "metric": "refactor+retrieve GB/s at 513^3 fp32, rel Linf 1e-4",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "launch_mode": "eager" if (not args.graph) else "hip_graph",
        ("graph_ms_per_step" if (not args.graph) else "eager_ms_per_step"): round(other_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"MDR refactor + retrieve of a {n}^3 float32 field (sin*cos*sin + 16 Gaussian bumps + "
                        f"N(0,1e-4), seed 0), {L} levels, {nb} bitplanes, NegaBinary encoder, retrieval at "
                        f"relative L-inf {args.tol:g} (MaxErrorEstimatorOB + greedy interpreter)",
            "dims": dims, "levels": L, "bitplanes": nb, "encoder": "negabinary", "tolerance_rel_linf": args.tol,
        },
        "refactor_ms": round(t_ref, 4),
        "retrieve_ms": round(t_ret, 4),
        "refactor_GBps": round(gbs_ref, 2),
        "retrieve_GBps": round(gbs_ret, 2),
        "refactor_hbm_frac": round(b_ref / (t_ref * 1e6) / PEAK_HBM_GBS, 4),
        "retrieve_hbm_frac": round(b_ret / (t_ret * 1e6) / PEAK_HBM_GBS, 4),
        "algorithmic_bytes": {"refactor": b_ref, "retrieve": b_ret},
        "planes_per_level": planes,
        "retrieved_plane_bytes": retrieved_bytes,
        "max_abs_error": err,
        "tolerance_abs": tau,
        "stages_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in stages.items() if v[1]},
        "kernels": per_kernel,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    plan.close()
    if not args.no_batched:
        res["batched"] = batched_report(args.steps, args.warmup)
    if not args.no_perbit:
        res["perbit"] = perbit_report(args.steps, args.warmup)
    if not args.no_qoi:
        res["qoi"] = qoi_report()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
