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
