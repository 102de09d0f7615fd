"""QoI-guided retrieval of a velocity magnitude |V| = sqrt(Vx^2 + Vy^2 + Vz^2)
from three refactored components, against the halving loop of ProDM's GE
application (app/GE/sequential/halving_Vtot.cpp:31-84,149-189).

Error model. For one component, a retrieval that holds k_l planes of level l
leaves coefficient errors |dc| <= t_l = 2^(e_l + 2 - k_l) + 2^(e_l + 2 - nb)
(negabinary truncation; t_l = max |c| for k_l = 0), so the field error at a
point x is bounded by
    |dV(x)| <= sum_l a_l(x) t_l,   a_l(x) = || row x of R, level-l part ||_1
with R the recompose operator. a_l(x) is read off the adjoint: R^T e_x is the
row of x. The QoI error follows ProDM's composition of bounds
(include/qoi_utils.hpp:18-41): e(|V|^2) <= sum_i 2 |V_i| E_i + E_i^2 and
e(|V|) <= e2 / (sqrt(|V|^2 - e2) + |V|) (or e2 / |V|, sqrt(e2) at 0).

* plain: ProDM's loop. Each component is retrieved to an L-inf tolerance
  eps_i (greedy planner, MaxErrorEstimatorOB); the QoI bound is evaluated
  with E_i = eps_i at every point; while its maximum exceeds tau, all eps_i
  shrink by 1.5 (halving_Vtot.cpp:56-70) and the components are retrieved
  again.
* adjoint-guided: the per-level amplification A_l = max a_l(x) over probe
  points (interior parity classes and the boundary), read off adjoint rows,
  replaces the estimator constant, and one greedy over all (component,
  level) planes adds the plane that lowers the largest QoI bound (composed
  at the reconstructed field, so each component counts by its sensitivity
  |V_i| / |V|) the most per byte, until the bound holds at every point.

Both report the plane bytes read and the actual QoI error against the
original field.
"""
import math

import numpy as np
import torch

from . import capi, interpret, mdr


# ---- synthetic field -------------------------------------------------------------
def velocity_modes(nmodes=24, kmax=32, seeds=(10, 11, 12)):
    """Integer wavevectors k_m with 0 < |k_m| <= kmax (seed seeds[0]),
    amplitude vectors a_m perpendicular to k_m with |a_m| = |k_m|^(-5/6)
    (seed seeds[1]) and phases (seed seeds[2])."""
    rk, ra, rp = (np.random.default_rng(s) for s in seeds)
    ks = []
    while len(ks) < nmodes:
        k = rk.integers(-kmax, kmax + 1, 3)
        if 0 < np.linalg.norm(k) <= kmax:
            ks.append(k.astype(np.float64))
    ks = np.array(ks)
    amps = []
    for k in ks:
        v = ra.normal(size=3)
        v -= v.dot(k) / k.dot(k) * k  # perpendicular to k
        amps.append(v / np.linalg.norm(v) * np.linalg.norm(k) ** (-5.0 / 6.0))
    return ks, np.array(amps), rp.uniform(0, 2 * np.pi, nmodes)


def velocity_field(n, nmodes=24, kmax=32, seeds=(10, 11, 12), device="cuda"):
    """Divergence-free random Fourier velocity on [0,1]^3 (n^3 per component,
    float32): V = sum_m a_m sin(2 pi k_m . x + phi_m); each mode has
    a_m . k_m = 0, so div V = 0."""
    ks, amps, phases = velocity_modes(nmodes, kmax, seeds)
    ax = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=device)
    V = torch.zeros(3, n, n, n, dtype=torch.float64, device=device)
    for m in range(nmodes):
        ph = (2 * np.pi * ks[m, 0] * ax)[:, None, None] + (2 * np.pi * ks[m, 1] * ax)[None, :, None] \
            + (2 * np.pi * ks[m, 2] * ax)[None, None, :] + phases[m]
        s = torch.sin(ph)
        for c in range(3):
            V[c] += amps[m, c] * s
    return [V[c].reshape(-1).to(torch.float32).contiguous() for c in range(3)]


# ---- QoI bound (include/qoi_utils.hpp:18-41, restated) ---------------------------
def vtot_bound(V, E):
    """Per-point bound on | |V| - |V~| | from per-component bounds E_i (tensors
    or scalars) at the reconstructed values V (list of 3 tensors)."""
    v2 = sum(v.double() ** 2 for v in V)
    e2 = sum(2 * v.double().abs() * e + e * e for v, e in zip(V, E))
    root = torch.sqrt(v2)
    out = torch.where(v2 > e2, e2 / (torch.sqrt(torch.clamp(v2 - e2, min=0)) + root), e2 / torch.clamp(root, min=1e-300))
    return torch.where(v2 == 0, torch.sqrt(e2), out)


# ---- per-component refactor and retrieval state -----------------------------------------
class Component:
    def __init__(self, x, dims, L, nb):
        self.dims, self.L, self.nb = dims, L, nb
        self.plan = mdr.Plan(dims, L, nb, torch.float32, capi.NEGABINARY)
        self.plan.refactor(x)
        torch.cuda.synchronize()
        self.info = [self.plan.level_info(l) for l in range(L + 1)]  # (max, exp, n)
        self.plane_bytes = [((n + 31) // 32) * 4 for _, _, n in self.info]
        self.k = [0] * (L + 1)

    def t(self, l, k):
        """Bound on the coefficient error of level l holding k planes."""
        m, e, _ = self.info[l]
        if k == 0:
            return float(m)
        return math.ldexp(1.0, e + 2 - k) + math.ldexp(1.0, e + 2 - self.nb)

    def bytes(self):
        return sum(k * b for k, b in zip(self.k, self.plane_bytes))

    def retrieve(self):
        return self.plan.retrieve(self.k)

    def plan_linf(self, tol):
        """Planes for an L-inf tolerance: the greedy planner over
        MaxErrorCollector tables with MaxErrorEstimatorOB(3), continuing from
        the planes held (ComposedReconstructor::progressive_reconstruct)."""
        import ctypes as C
        tables = []
        for m, _, _ in self.info:
            t = np.zeros(self.nb + 1)
            capi.lib().prodm_collect_max_f32(C.c_float(np.float32(m) * np.float32(4.0)), self.nb,
                                             t.ctypes.data_as(C.c_void_p))
            tables.append(t)
        sizes = [[b] * self.nb for b in self.plane_bytes]
        idx = np.array(self.k, np.uint8)
        _, idx = interpret.retrieve_sizes(interpret.GREEDY, interpret.MAX_OB, sizes, tables, tol, idx, ndims=3)
        self.k = [int(v) for v in idx]


# ---- adjoint rows ------------------------------------------------------------------
def level_boxes(dims, L):
    ld, _, _ = mdr.level_geometry(dims, L)
    return [tuple(int(v) for v in ld[l]) for l in range(L + 1)]


def row_level_norms(dims, L, points, device):
    """a_l(x) = L1 norm of the level-l part of row x of the recompose
    operator, from one adjoint per point: returns (len(points), L+1)."""
    boxes = level_boxes(dims, L)
    N = int(np.prod(dims))
    out = torch.zeros(len(points), L + 1, dtype=torch.float64, device=device)
    e = torch.zeros(N, dtype=torch.float64, device=device)
    for p, x in enumerate(points):
        e.zero_()
        e[x] = 1.0
        r = mdr.recompose_adjoint(e, dims, L).abs().reshape(dims)
        prev = 0.0
        for l in range(L + 1):
            b = boxes[l]
            s = float(r[:b[0], :b[1], :b[2]].sum())
            out[p, l] = s - prev
            prev = s
    return out


def probe_points(dims):
    """Points covering the interior parity classes and the boundary: the 27
    parity/position combinations around the centre, and faces, edges and
    corners."""
    pts = set()
    c = [n // 2 for n in dims]
    for d0 in range(3):
        for d1 in range(3):
            for d2 in range(3):
                pts.add((c[0] + d0, c[1] + d1, c[2] + d2))
    for a in (0, 1, 2, dims[0] // 2, dims[0] - 2, dims[0] - 1):
        for b in (0, 1, dims[1] // 2, dims[1] - 1):
            for q in (0, 1, dims[2] // 2, dims[2] - 1):
                pts.add((a, b, q))
    return [(i * dims[1] + j) * dims[2] + k for i, j, k in sorted(pts)]


# ---- the two retrieval strategies ------------------------------------------------------------
def plain_halving(comps, V_true, rel, max_iter=8):
    """ProDM's loop (halving_Vtot.cpp:99-189): eps_i = rel * range(V_i),
    retrieve, evaluate the QoI bound with E_i = eps_i, shrink by 1.5."""
    Q = torch.sqrt(sum(v.double() ** 2 for v in V_true))
    tau = rel * float(Q.max() - Q.min())
    eps = [rel * float(v.max() - v.min()) for v in V_true]
    for it in range(max_iter):
        for c, e in zip(comps, eps):
            c.plan_linf(e)
        Vr = [c.retrieve() for c in comps]
        est = vtot_bound(Vr, eps)
        i = int(torch.argmax(est))
        if float(est[i]) <= tau:
            break
        v2 = float(sum(v[i].double() ** 2 for v in Vr))
        e_est = float(est[i])
        while e_est > tau:
            eps = [e / 1.5 for e in eps]
            e2 = sum(2 * abs(float(v[i])) * e + e * e for v, e in zip(Vr, eps))
            root = math.sqrt(v2)
            e_est = e2 / (math.sqrt(max(v2 - e2, 0)) + root) if v2 > e2 else (math.sqrt(e2) if v2 == 0 else e2 / root)
    Qr = torch.sqrt(sum(v.double() ** 2 for v in Vr))
    err = float(torch.max(torch.abs(Qr - Q)))
    return {"bytes": sum(c.bytes() for c in comps), "planes": [list(c.k) for c in comps], "tau": tau,
            "max_error": err, "max_bound": float(est.max()), "iterations": it + 1}


def adjoint_guided(comps, V_true, rel, topk=1 << 16, max_steps=2000):
    """Greedy over (component, level) planes on the QoI bound itself: the
    per-level amplifications A_l come from adjoint rows (probe points), the
    component weighting from the QoI bound composed at the reconstructed
    field (its sensitivity |V_i| / |V| to first order). Each step adds the
    plane that lowers the largest QoI bound the most per byte, the maximum
    taken over the points where the bound is largest."""
    dims, L = comps[0].dims, comps[0].L
    dev = V_true[0].device
    Q = torch.sqrt(sum(v.double() ** 2 for v in V_true))
    tau = rel * float(Q.max() - Q.min())
    A = row_level_norms(dims, L, probe_points(dims), dev).max(dim=0).values.cpu().numpy()

    def E_of(ks):
        return [sum(float(A[l]) * c.t(l, ks[i][l]) for l in range(L + 1)) for i, c in enumerate(comps)]

    # start from ProDM's first pass (each component to rel * range(V_i)) so
    # that the bound is evaluated on a meaningful field
    for c, v in zip(comps, V_true):
        c.k = [0] * (L + 1)
        c.plan_linf(rel * float(v.max() - v.min()))
    ks = [list(c.k) for c in comps]
    rounds = 0
    for rounds in range(1, 9):
        Vr = [c.retrieve() for c in comps]
        B = vtot_bound(Vr, E_of(ks))
        if float(B.max()) <= tau:
            break
        steps = 0
        while steps < max_steps:
            idx = torch.topk(B, min(topk, B.numel())).indices
            Vs = [v[idx] for v in Vr]
            cur = float(vtot_bound(Vs, E_of(ks)).max())
            if cur <= tau:
                break
            best, best_gain = None, -1.0
            for i, c in enumerate(comps):
                for l in range(L + 1):
                    if ks[i][l] >= c.nb:
                        continue
                    ks[i][l] += 1
                    val = float(vtot_bound(Vs, E_of(ks)).max())
                    ks[i][l] -= 1
                    gain = (cur - val) / c.plane_bytes[l]
                    if gain > best_gain:
                        best, best_gain = (i, l), gain
            if best is None:
                break
            ks[best[0]][best[1]] += 1
            steps += 1
            B = vtot_bound(Vr, E_of(ks))
        for i, c in enumerate(comps):
            c.k = list(ks[i])
    Vr = [c.retrieve() for c in comps]
    est = vtot_bound(Vr, E_of(ks))
    Qr = torch.sqrt(sum(v.double() ** 2 for v in Vr))
    err = float(torch.max(torch.abs(Qr - Q)))
    return {"bytes": sum(c.bytes() for c in comps), "planes": [list(c.k) for c in comps], "tau": tau,
            "max_error": err, "max_bound": float(est.max()), "rounds": rounds,
            "amplification": [round(float(a), 4) for a in A]}


def run(n=257, L=5, nb=32, rel=1e-3, device="cuda"):
    dims = [n, n, n]
    V = velocity_field(n, device=device)
    comps = [Component(v, dims, L, nb) for v in V]
    plain = plain_halving(comps, V, rel)
    for c in comps:
        c.k = [0] * (L + 1)
    guided = adjoint_guided(comps, V, rel)
    for c in comps:
        c.plan.close()
    return {"workload": f"|V| of a divergence-free {n}^3 float32 velocity, {L} levels, {nb} planes, rel L-inf {rel}",
            "plain": plain, "adjoint_guided": guided,
            "bytes_reduction": round(1.0 - guided["bytes"] / plain["bytes"], 4)}


if __name__ == "__main__":
    import json
    import sys
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 257
    print(json.dumps(run(n)))
