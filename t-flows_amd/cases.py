"""The measured workloads: sizes, seeds and solver settings.

Master seed 20261202; workload k draws from seed 20261202 + k.  Tolerances
follow the reference convention (absolute on ||r||_2, Cg.f90:174): a relative
target eps is passed as ``tol = eps * ||b||_2`` with ``x0 = 0``.

1. Poisson 32^3, uniform, Neumann walls, cell 1 pinned, zero-mean RHS;
   Cg with prec 2 to 1e-8 (CPU reference and GPU).
2. Momentum 128^3: tanh clustering beta = 2 on every axis, +-10% node jitter,
   random cell numbering, upwind convection-diffusion (nu = 1e-3) from a
   divergence-free 8-mode field, inertia V/dt with dt = h_min (CFL 1);
   Bicg with prec 2 (the reference's only incomplete factorisation) to 1e-10.
3. Poisson 256^3 with tanh clustering giving a largest cell aspect ratio of
   100; FP32-inner PCG inside FP64 refinement to 1e-10, plain FP64 beside it.
4. 64 Poisson systems, per-axis cells log-uniform in [16, 96], beta uniform in
   [1, 3]; FP32+FP64 refinement to 1e-10.
5. Serving: Poisson arrivals, 70% pressure Cg / 30% momentum Bicg, cells
   log-uniform in [1e4, 4e6], 10 revisits per case with a perturbed RHS.
"""
import math

import numpy as np

from . import mesh

MASTER_SEED = 20261202


def seed(k):
    return MASTER_SEED + k


def beta_for_aspect(n, aspect):
    """tanh clustering strength whose 1-D grid of n cells has largest/smallest
    spacing ratio `aspect` (the largest cell aspect ratio of a box stretched
    the same way on every axis)."""
    def ratio(beta):
        i = np.arange(n + 1)
        x = 0.5 * (1 + np.tanh(beta * (2 * i / n - 1)) / np.tanh(beta))
        h = np.diff(x)
        return h.max() / h.min()
    lo, hi = 1e-3, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if ratio(mid) < aspect:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def workload_kwargs(k, scale=1.0):
    """make_case keyword arguments of workload k (1..3); `scale` shrinks the
    mesh per axis for parity runs at oracle-friendly sizes."""
    if k == 1:
        n = max(2, int(round(32 * scale)))
        return dict(kind=mesh.POISSON, nx=n, beta=0.0, pin_first=True, rhs_zero_mean=True,
                    seed=seed(1))
    if k == 2:
        n = max(2, int(round(128 * scale)))
        return dict(kind=mesh.MOMENTUM, nx=n, beta=2.0, jitter=0.1, permute=True,
                    pin_first=False, rhs_zero_mean=False, nu=1e-3, cfl=1.0, n_modes=8,
                    seed=seed(2))
    if k == 3:
        n = max(2, int(round(256 * scale)))
        return dict(kind=mesh.POISSON, nx=n, beta=beta_for_aspect(n, 100.0), pin_first=True,
                    rhs_zero_mean=True, seed=seed(3))
    raise ValueError(k)


SOLVE = {
    1: dict(method="cg", prec=2, rtol=1e-8, cap=10000),
    2: dict(method="bicg", prec=2, rtol=1e-10, cap=10000),
    3: dict(method="cg", prec=2, rtol=1e-10, cap=100000),
}


SERVING_SOLVE = {
    "pressure": dict(method="cg", prec=2, rtol=1e-10, cap=20000),
    "momentum": dict(method="bicg", prec=2, rtol=1e-10, cap=20000),
}


def serving_cases(count=24, lo=1e4, hi=4e6, p_pressure=0.7, k=5):
    """Workload 5: per-case (kind name, make_case kwargs).  Cells log-uniform
    in [lo, hi] on a cube-shaped hex mesh, tanh stretching uniform in [1, 3],
    +-10% node jitter, random numbering; 70% pressure (Poisson, Cg) and 30%
    momentum (Bicg)."""
    s = seed(k)
    out = []
    for i in range(count):
        cells = math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * mesh.rng_u01(s, 300 + i, 0))
        nx = max(2, int(round(cells ** (1.0 / 3.0))))
        beta = 1.0 + 2.0 * mesh.rng_u01(s, 500 + i, 0)
        if mesh.rng_u01(s, 400 + i, 0) < p_pressure:
            out.append(("pressure", dict(kind=mesh.POISSON, nx=nx, beta=beta, jitter=0.1,
                                         permute=True, pin_first=True, rhs_zero_mean=True,
                                         seed=s * 1000 + i)))
        else:
            out.append(("momentum", dict(kind=mesh.MOMENTUM, nx=nx, beta=beta, jitter=0.1,
                                         permute=True, pin_first=False, rhs_zero_mean=False,
                                         nu=1e-3, cfl=1.0, n_modes=8, seed=s * 1000 + i)))
    return out


def serving_schedule(n_cases, visits=10, rate=200.0, k=5):
    """Workload 5 request stream: (arrival time [s], case index, visit) in
    arrival order.  Exponential inter-arrival times at `rate` requests/s
    (a Poisson process); each case gets `visits` requests, the order of cases
    drawn from the seed."""
    rng = np.random.default_rng(seed(k))
    order = np.repeat(np.arange(n_cases), visits)
    rng.shuffle(order)
    t = np.cumsum(rng.exponential(1.0 / rate, size=order.size))
    seen = np.zeros(n_cases, int)
    out = []
    for ti, ci in zip(t, order):
        out.append((float(ti), int(ci), int(seen[ci])))
        seen[ci] += 1
    return out


def perturbed_rhs(b, case_index, visit, eps=0.01, k=5):
    """RHS of one time step: b scaled entrywise by 1 + eps*u, u uniform in
    [-1, 1) from the seed (case, visit); visit 0 is b itself."""
    if visit == 0:
        return b.copy()
    rng = np.random.default_rng([seed(k), case_index, visit])
    return b * (1.0 + eps * (2.0 * rng.random(b.size) - 1.0))


def batch_systems(count=64, lo=16, hi=96, k=4):
    """Workload 4: per-system make_case kwargs (log-uniform cells per axis)."""
    s = seed(k)
    out = []
    for i in range(count):
        dims = [int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo))
                                   * mesh.rng_u01(s, 100 + i, a)))) for a in range(3)]
        beta = 1.0 + 2.0 * mesh.rng_u01(s, 200 + i, 0)
        out.append(dict(kind=mesh.POISSON, nx=dims[0], ny=dims[1], nz=dims[2], beta=beta,
                        pin_first=True, rhs_zero_mean=True, seed=s * 1000 + i))
    return out
