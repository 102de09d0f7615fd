"""QoI-guided retrieval of a velocity magnitude |V| = sqrt(Vx^2 + Vy^2 + Vz^2)
from three refactored components, against the halving loop of ProDM's GE
application (app/GE/sequential/halving_Vtot.cpp:31-84,149-189).

Error model. A retrieval holding k planes of a level leaves every
coefficient's error below t_l[k], the MaxErrorCollector table of the level's
stored bound (MaxErrorCollector.hpp:14-25; for negabinary streams the bound
is 4 max|c|, so t_l[k] = 2^(e_l + 2 - k) for k >= 1, which bounds the
negabinary truncation, and t_l[0] >= max|c|). The field error at every node
is then at most E = sum_l A_l t_l[k_l] + r, A_l the exact largest row
1-norm of recompose's level-l block (prodm_recompose_amplification,
include/prodm/qoi.hpp) and r an allowance for rounding: 2^-21 of each level
maximum for the decoder's float conversion (amplified by A_l) and 2^-19 of
the component's magnitude per level for the float32 recompose. The QoI error
follows ProDM's composition of bounds (include/qoi_utils.hpp:18-41):
e(|V|^2) <= sum_i 2 |V~_i| E_i + E_i^2 and e(|V|) <= e2 / (sqrt(|V~|^2 - e2)
+ |V~|) (or e2 / |V~|, sqrt(e2) at 0), at the reconstructed values V~.

* plain: ProDM's loop. Each component is retrieved to an L-inf tolerance
  eps_i (greedy planner, MaxErrorEstimatorOB); the QoI bound is evaluated
  with E_i = eps_i at every point; while its maximum exceeds tau, all eps_i
  shrink by 1.5 (halving_Vtot.cpp:56-70) and the components are retrieved
  again. MaxErrorEstimatorOB's constant bounds A_l, so eps_i bounds the
  component error.
* adjoint-guided: each round retrieves, evaluates the certified bound over
  every node, and hands the nodes where it is largest to the QoI greedy
  planner (QoIGreedyBasedSizeInterpreter, one plane of one component's level
  at a time, the largest drop of the bound per byte first), until the bound
  holds at every node.

Both report the plane bytes read, the largest certified bound over the field
and the actual QoI error against the original field.
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

    def tables(self):
        """Per level, the MaxErrorCollector table of the stored bound (4 max
        for negabinary): entry k bounds every coefficient's error after k
        planes."""
        import ctypes as C
        out = []
        for m, _, _ in self.info:
            t = np.zeros(self.nb + 1)
            capi.lib().prodm_collect_max_f32(C.c_float(np.float32(m) * np.float32(4.0)), self.nb,
                                             t.ctypes.data_as(C.c_void_p))
            out.append(t)
        return out

    def rounding(self, A, vmax):
        """Allowance for the decoder's float conversion (2^-21 of each level
        maximum, amplified) and the float32 recompose (2^-19 of the
        component's magnitude per level)."""
        dec = sum(A[l] * math.ldexp(float(m), -21) for l, (m, _, _) in enumerate(self.info))
        return dec + (self.L + 1) * math.ldexp(float(vmax), -19)

    def bytes(self):
        return sum(k * b for k, b in zip(self.k, self.plane_bytes))

    def retrieve(self):
        return self.plan.retrieve(self.k)

    def plan_linf(self, tol):
        """Planes for an L-inf tolerance: the greedy planner over
        MaxErrorCollector tables with MaxErrorEstimatorOB(3), continuing from
        the planes held (ComposedReconstructor::progressive_reconstruct)."""
        tables = self.tables()
        sizes = [[b] * self.nb for b in self.plane_bytes]
        idx = np.array(self.k, np.uint8)
        _, idx = interpret.retrieve_sizes(interpret.GREEDY, interpret.MAX_OB, sizes, tables, tol, idx, ndims=3)
        self.k = [int(v) for v in idx]


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


def adjoint_guided(comps, V_true, rel, sample_points=4096, max_rounds=12):
    """Certified QoI planning (see the module docstring): a warm start at
    4 rel range(V_i) per component with ProDM's L-inf planner (whose extra
    factor 4 makes it about the certified estimate at rel range(V_i)), then
    rounds of retrieve, bound over every node, and the QoI greedy planner on
    the sample_points nodes with the largest bound."""
    dims, L, nb = comps[0].dims, comps[0].L, comps[0].nb
    Q = torch.sqrt(sum(v.double() ** 2 for v in V_true))
    tau = rel * float(Q.max() - Q.min())
    A = interpret.recompose_amplification(dims, L)
    tables = [t for c in comps for t in c.tables()]
    sizes = [[b] * nb for c in comps for b in c.plane_bytes]
    for c, v in zip(comps, V_true):
        c.k = [0] * (L + 1)
        c.plan_linf(4 * rel * float(v.max() - v.min()))
    rounds = 0
    for rounds in range(1, max_rounds + 1):
        Vr = [c.retrieve() for c in comps]
        rnd = [c.rounding(A, float(v.abs().max())) for c, v in zip(comps, Vr)]
        E = [rnd[i] + sum(A[l] * tables[i * (L + 1) + l][c.k[l]] for l in range(L + 1)) for i, c in enumerate(comps)]
        B = vtot_bound(Vr, E)
        if float(B.max()) <= tau:
            break
        top = torch.topk(B, min(sample_points, B.numel())).indices
        sample = torch.stack([v[top].double() for v in Vr], dim=1).cpu().numpy()
        idx = np.array([k for c in comps for k in c.k], np.uint8)
        _, idx = interpret.qoi_retrieve_sizes(sizes, tables, A, sample, tau, idx, rounding=rnd)
        for i, c in enumerate(comps):
            c.k = [int(v) for v in idx[i * (L + 1):(i + 1) * (L + 1)]]
    Qr = torch.sqrt(sum(v.double() ** 2 for v in Vr))
    err = float(torch.max(torch.abs(Qr - Q)))
    return {"bytes": sum(c.bytes() for c in comps), "planes": [list(c.k) for c in comps], "tau": tau,
            "max_error": err, "max_bound": float(B.max()), "rounds": rounds,
            "amplification": [round(float(a), 6) for a in A]}


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
