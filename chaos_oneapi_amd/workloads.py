"""Named sweep workloads (descriptor dictionaries for lyap.make_desc).

Parameter orders: lorenz (sigma, rho, beta), rossler (a, c, b),
lorenz96_N (F, d), shimizu (alpha, lambda).  Grids follow the reference rule
p = start + i*|end-start|/(n-1), first parameter outer (tasklyapunov.cpp:45-63).
"""
from __future__ import annotations

from . import lyap as L


def grid_fields(idx1, start1, end1, n1, idx2, start2, end2, n2) -> dict:
    return dict(sweep_idx1=idx1, start1=float(start1), delta1=abs(end1 - start1) / (n1 - 1) if n1 > 1 else float("inf"),
                n1=n1, sweep_idx2=idx2, start2=float(start2),
                delta2=abs(end2 - start2) / (n2 - 1) if n2 > 1 else float("inf"), n2=n2)


def lorenz_fixed(n: int = 64, precision: int = L.PREC_FP64, t_transient: float = 100.0,
                 t_average: float = 1000.0) -> dict:
    """Lorenz (sigma=10), rho in [0,200] x beta in [0.5,4], Dopri5 fixed h=0.01,
    renormalise every 1.0, initial (1,1,1) + U(+-1e-3)."""
    d = dict(model=L.MODEL_LORENZ, dim=3, n_exps=3, n_params=3, method=L.METHOD_VAR,
             integrator=L.INT_DOPRI5, adaptive=0, precision=precision, h=0.01,
             t_transient=t_transient, t_slave=0.0, t_average=t_average, t_renorm=1.0,
             escape_radius=1e6, base_params=[10.0, 0.0, 0.0], init_state=[1.0, 1.0, 1.0],
             init_kind=L.INIT_UNIFORM, init_param_offset=-1, init_scale=1e-3, seed=20240601)
    d.update(grid_fields(1, 0.0, 200.0, n, 2, 0.5, 4.0, n))
    return d


def rossler_dverk(n: int = 512, precision: int = L.PREC_FP64, t_transient: float = 200.0,
                  t_average: float = 2000.0, tol: float = 1e-10) -> dict:
    """Rossler (b=0.2), a in [0.05,0.4] x c in [2,20], adaptive DVERK 6(5)
    rtol=atol=1e-10, renormalise every 0.5, initial state U[-1,1]^3.  Orbits
    leaving |x_i| < 1e4 (the attractors are O(c)) are marked escaped; a point
    may take 2e6 steps (~15x the grid's mean)."""
    d = dict(model=L.MODEL_ROSSLER, dim=3, n_exps=3, n_params=3, method=L.METHOD_VAR,
             integrator=L.INT_DVERK, adaptive=1, precision=precision, h=0.01, rtol=tol, atol=tol,
             escape_radius=1e4, max_steps=2_000_000,
             t_transient=t_transient, t_slave=0.0, t_average=t_average, t_renorm=0.5,
             base_params=[0.0, 0.0, 0.2], init_state=[0.0, 0.0, 0.0],
             init_kind=L.INIT_UNIFORM, init_param_offset=-1, init_scale=1.0, seed=20240602)
    d.update(grid_fields(0, 0.05, 0.4, n, 1, 2.0, 20.0, n))
    return d


def lorenz96_dopri(n: int = 256, dim: int = 16, precision: int = L.PREC_FP64,
                   t_transient: float = 100.0, t_average: float = 500.0, tol: float = 1e-9) -> dict:
    """Lorenz-96 (N=dim, periodic, damping d), F in [2,16] x d in [0.5,1.5],
    adaptive Dopri5 rtol=atol=1e-9, renormalise every 0.2, initial F + N(0, 0.1^2)."""
    d = dict(model=L.MODEL_L96, dim=dim, n_exps=dim, n_params=2, method=L.METHOD_VAR,
             integrator=L.INT_DOPRI5, adaptive=1, precision=precision, h=0.01, rtol=tol, atol=tol,
             escape_radius=1e4, max_steps=5_000_000,
             t_transient=t_transient, t_slave=0.0, t_average=t_average, t_renorm=0.2,
             base_params=[0.0, 0.0], init_state=[0.0] * dim, init_kind=L.INIT_NORMAL,
             init_param_offset=0, init_scale=0.1, seed=20240603)
    d.update(grid_fields(0, 2.0, 16.0, n, 1, 0.5, 1.5, n))
    return d


# class mix of the batched uneven sweep: (name, share)
BATCH_MIX = (("lorenz", 0.40), ("rossler", 0.30), ("lorenz96_8", 0.20), ("lorenz96_16", 0.08),
             ("lorenz96_32", 0.02))


def batch_mix(n_points: int = 1 << 20, seed: int = 20240604, precision: int = L.PREC_MIXED,
              time_scale: float = 1.0, tol: float = 1e-9) -> list:
    """Batched uneven sweep: n_points drawn with the seed as 40% Lorenz, 30%
    Rossler, 20% Lorenz-96 N=8, 8% N=16 and 2% N=32, each with parameters
    uniform over its sweep ranges (Lorenz rho in [0,200], beta in [0.5,4];
    Rossler a in [0.05,0.4], c in [2,20]; Lorenz-96 F in [2,16], d in
    [0.5,1.5]), all adaptive Dopri5 at rtol=atol=tol.  Each class keeps the
    times, renormalisation interval and initial-state rule of its single-model
    workload; time_scale shrinks the transient and averaging times (tests).

    Returns a list of dicts {name, spec, count, params} with params an
    [count, n_params] float64 numpy array, in the mix order."""
    import numpy as np

    rng = np.random.default_rng(seed)
    share = np.array([s for _, s in BATCH_MIX])
    cls = rng.choice(len(BATCH_MIX), size=n_points, p=share / share.sum())
    counts = np.bincount(cls, minlength=len(BATCH_MIX))
    out = []
    for j, (name, _) in enumerate(BATCH_MIX):
        n = int(counts[j])
        if name == "lorenz":
            spec = lorenz_fixed(2, precision=precision, t_transient=100.0 * time_scale,
                                t_average=1000.0 * time_scale)
            p = np.stack([np.full(n, 10.0), rng.uniform(0.0, 200.0, n), rng.uniform(0.5, 4.0, n)], 1)
        elif name == "rossler":
            spec = rossler_dverk(2, precision=precision, t_transient=200.0 * time_scale,
                                 t_average=2000.0 * time_scale)
            p = np.stack([rng.uniform(0.05, 0.4, n), rng.uniform(2.0, 20.0, n), np.full(n, 0.2)], 1)
        else:
            dim = int(name.split("_")[1])
            spec = lorenz96_dopri(2, dim=dim, precision=precision, t_transient=100.0 * time_scale,
                                  t_average=500.0 * time_scale)
            p = np.stack([rng.uniform(2.0, 16.0, n), rng.uniform(0.5, 1.5, n)], 1)
        spec.update(integrator=L.INT_DOPRI5, adaptive=1, rtol=tol, atol=tol)
        out.append(dict(name=name, spec=spec, count=n, params=np.ascontiguousarray(p)))
    return out


def shimizu_reference(n1: int = 200, n2: int = 200, method: int = L.METHOD_FD_REF) -> dict:
    """The reference's shipped configuration (ChaosMigration/config.json:2-36):
    shimizu (alpha, lambda) over [0.2, 2] x [0, 1.2], integration step 1e-2,
    skip times 1e3 (attractor) and 1e3 (slave trajectories), total time 1e4,
    normalisation every 1e-2, epsilon 1e-7, initial state (1e-4, 0, 0); the
    reference's FD mode (using_qr = using_var = false) unless another method
    is given."""
    d = dict(model=L.MODEL_SHIMIZU, dim=3, n_exps=3, n_params=2, method=method, integrator=L.INT_DVERK,
             adaptive=0, precision=L.PREC_FP64, h=1e-2, eps=1e-7, t_transient=1e3, t_slave=1e3,
             t_average=1e4, t_renorm=1e-2, base_params=[0.0, 0.0], init_state=[1e-4, 0.0, 0.0],
             init_kind=L.INIT_NONE, init_param_offset=-1)
    d.update(grid_fields(0, 0.2, 2.0, n1, 1, 0.0, 1.2, n2))
    return d


def henon(n: int = 256, precision: int = L.PREC_FP64, t_transient: float = 1000.0,
          t_average: float = 100000.0, t_renorm: float = 10.0) -> dict:
    """Henon map x' = 1 - a x^2 + y, y' = b x over a in [1.0, 1.42] x b in
    [0.05, 0.35]; times count iterations (h = 1), renormalise every t_renorm
    iterations; initial state (0.1, 0.1) + U(+-0.05).  Orbits leaving
    |x_i| < 1e3 are marked escaped."""
    d = dict(model=L.MODEL_HENON, dim=2, n_exps=2, n_params=2, method=L.METHOD_VAR,
             integrator=L.INT_DVERK, adaptive=0, precision=precision, h=1.0,
             t_transient=t_transient, t_slave=0.0, t_average=t_average, t_renorm=t_renorm,
             escape_radius=1e3, base_params=[0.0, 0.0], init_state=[0.1, 0.1],
             init_kind=L.INIT_UNIFORM, init_param_offset=-1, init_scale=0.05, seed=20240605)
    d.update(grid_fields(0, 1.0, 1.42, n, 1, 0.05, 0.35, n))
    return d


def logistic(n: int = 4096, precision: int = L.PREC_FP64, t_transient: float = 1000.0,
             t_average: float = 100000.0, t_renorm: float = 10.0) -> dict:
    """Logistic map x' = r x (1 - x) over r in [2.8, 4.0] (n points, one
    parameter); times count iterations; initial state 0.3 + U(+-0.1)."""
    d = dict(model=L.MODEL_LOGISTIC, dim=1, n_exps=1, n_params=1, method=L.METHOD_VAR,
             integrator=L.INT_DVERK, adaptive=0, precision=precision, h=1.0,
             t_transient=t_transient, t_slave=0.0, t_average=t_average, t_renorm=t_renorm,
             escape_radius=1e3, base_params=[0.0], init_state=[0.3], init_kind=L.INIT_UNIFORM,
             init_param_offset=-1, init_scale=0.1, seed=20240606)
    d.update(grid_fields(0, 2.8, 4.0, n, 0, 2.8, 2.8, 1))
    return d
