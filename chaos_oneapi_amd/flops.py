"""Algorithmic flop count of the sweep (the numerator of the roofline figure).

Convention: an FMA counts 2, every other arithmetic operation 1; negations,
comparisons, conversions and the step-size controller's log/exp are not
counted.  Counts follow the operations the kernels perform:

  stage sums     first term a product (1), later terms fma (2), then
                 fma(h, acc, y) (2)  ->  2*nnz + 1 per stage and component
  solution       2*nnz(b) + 1 per component
  error estimate 2*nnz(e) per component (the last multiply by h included),
                 norm 7 per component + 2
  model          f and J v costs of the written formulas (below)
  MGS            per vector j: j*(4N-1) + (2N-1) + N + 2

Per point: n_acc * (base + K * tangent) + n_rej * base + renorms * MGS.

Precision split of the mixed form: the base step, error norm, MGS and log
accumulation are binary64; a tangent vector's stage arguments, Jacobian
products and stage sums are binary32 except the last add of its increment to
the binary64 deviation vector (N flops per vector and step).  The Lorenz-96
mixed form folds h into its binary32 Jacobian data (two binary64 products per
component and stage, h*d once per step): its stage arguments are fma chains
started on v (2*nnz per stage and component) and its increment sum(b_j K'_j)
has no multiply by h (2*nnz(b) - 1).  The 3-D flows' mixed form with adaptive
DVERK steps integrates its binary32 tangent with the embedded fifth-order
weights b - e (5 nonzeros) on stages 1-6 (a-row nonzeros 0..5), one stage
fewer than the sixth-order solution's.

Peaks (MI355X, published): binary64 vector 78.6 TFLOP/s, binary32 vector
157.3 TFLOP/s (the latter counts packed two-wide FMAs, which this build does
not emit; unpacked binary32 FMAs issue at the binary64 rate, measured by the
valu_peak probe).  A launch's precision-weighted peak is the rate at which
its flop mix would run if every flop ran at its precision's peak:
total / (f64 / P64 + f32 / P32).
"""
from __future__ import annotations

from . import lyap as L

# (f, J v per vector) flop counts of the written formulas
MODEL_COST = {
    L.MODEL_LORENZ: lambda n: (8, 12),
    L.MODEL_ROSSLER: lambda n: (6, 7),
    L.MODEL_SHIMIZU: lambda n: (7, 10),
    # f_i = fma(x+ - x--, x-, fma(-d, x, F)): 5;  J v_i: 1 + 1 + 2 + 2 + 1 = 7
    L.MODEL_L96: lambda n: (5 * n, 7 * n),
}

# nonzero counts of the tableaux
DVERK_A_ROWS = [0, 1, 2, 3, 4, 5, 5, 6]  # stage 6 (index 5) feeds the error only
DVERK_NNZ_B, DVERK_NNZ_E = 6, 7
DVERK_LOW_ROWS, DVERK_NNZ_BH = [0, 1, 2, 3, 4, 5], 5  # the fifth-order weights' stages 1-6
DOPRI_A_ROWS = [0, 1, 2, 3, 4, 5, 5]  # stage 7 (index 6) is the FSAL stage
DOPRI_NNZ_B, DOPRI_NNZ_E = 5, 6


def _stages(integ: int, adaptive: bool, for_tangent: bool):
    rows = DVERK_A_ROWS if integ == L.INT_DVERK else DOPRI_A_ROWS
    skip = 5 if integ == L.INT_DVERK else 6
    out = []
    for s, nnz in enumerate(rows):
        if s == skip and (for_tangent or not adaptive):
            continue
        out.append(nnz)
    return out


# the step's integral of tr J (the trace closure): h * (c + sum_j b_j g(X_j))
# for Rossler (g = x), h * c for the flows with a constant trace; plus the
# add into the interval's sum
def _trace_cost(model: int, integ: int) -> int:
    nnz_b = DVERK_NNZ_B if integ == L.INT_DVERK else DOPRI_NNZ_B
    return (2 * nnz_b - 1) + 3 if model == L.MODEL_ROSSLER else 2


def _step_parts(model: int, n: int, k: int, integ: int, adaptive: bool, mixed: bool,
                trace: bool = False):
    """(base step, binary32 tangent, binary64 tangent) flops of an accepted
    step of one point; a rejected step costs the base step alone.  trace: the
    base step also integrates tr J (k is then the number of vectors, N - 1)."""
    f_cost, jv_cost = MODEL_COST[model](n)
    nnz_b = DVERK_NNZ_B if integ == L.INT_DVERK else DOPRI_NNZ_B
    nnz_e = DVERK_NNZ_E if integ == L.INT_DVERK else DOPRI_NNZ_E
    base_st = _stages(integ, adaptive, for_tangent=False)
    # Dopri5 with error control reuses stage 7 as the next stage 1 (FSAL)
    n_f = len(base_st) - (1 if (integ == L.INT_DOPRI5 and adaptive) else 0)
    base = n_f * f_cost + n * sum(2 * z + 1 for z in base_st if z > 0) + n * (2 * nnz_b + 1)
    if adaptive:
        base += n * 2 * nnz_e + 7 * n + 2
    if trace:
        base += _trace_cost(model, integ)
    tan_st = _stages(integ, adaptive, for_tangent=True)
    if mixed and model == L.MODEL_L96:
        vec32 = len(tan_st) * jv_cost + n * sum(2 * z for z in tan_st if z > 0) + n * (2 * nnz_b - 1)
        return base, k * vec32, k * n + 2 * n * len(tan_st) + 1
    if mixed and adaptive and integ == L.INT_DVERK:
        tan_st, nnz_b = DVERK_LOW_ROWS, DVERK_NNZ_BH
    tangent = len(tan_st) * jv_cost + n * sum(2 * z + 1 for z in tan_st if z > 0) + n * (2 * nnz_b + 1)
    if mixed:
        return base, k * (tangent - n), k * n
    return base, 0, k * tangent


def per_step(model: int, n: int, k: int, integ: int, adaptive: bool, mixed: bool = False,
             trace: bool = False):
    """(flops of an accepted step, flops of a rejected step) for one point."""
    base, t32, t64 = _step_parts(model, n, k, integ, adaptive, mixed, trace)
    return base + t32 + t64, base


def desc_step(desc):
    """(accepted-step flops, rejected-step flops, binary32 flops of an accepted
    step, deviation vectors) of a finalized descriptor's method."""
    trace = desc.method == L.METHOD_VAR_TRACE
    k = desc.dim - 1 if trace else desc.dim
    mixed = desc.precision == L.PREC_MIXED
    base, t32, t64 = _step_parts(desc.model, desc.dim, k, desc.integrator, bool(desc.adaptive), mixed, trace)
    return base + t32 + t64, base, t32, k


PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def per_step_split(model: int, n: int, k: int, integ: int, adaptive: bool, mixed: bool,
                   trace: bool = False):
    """Binary32 flops of an accepted step (the rest of per_step's accepted
    count is binary64; a rejected step is all binary64)."""
    return _step_parts(model, n, k, integ, adaptive, mixed, trace)[1]


def weighted_peak(f64: float, f32: float) -> float:
    """Precision-weighted published peak (TFLOP/s) of a flop mix."""
    tot = f64 + f32
    if tot <= 0:
        return PEAK_FP64_TFLOPS
    return tot / (f64 / PEAK_FP64_TFLOPS + f32 / PEAK_FP32_TFLOPS)


def sweep_flops_split(desc, n_acc, n_rej):
    """(binary64 flops, binary32 flops) of a sweep from its step counters."""
    import numpy as np

    total = sweep_flops(desc, n_acc, n_rej)
    f32_step = desc_step(desc)[2]
    f32 = float(f32_step * np.asarray(n_acc, dtype=np.float64).sum())
    return total - f32, f32


def mgs(n: int, k: int) -> int:
    return sum(j * (4 * n - 1) + (2 * n - 1) + n + 2 for j in range(k))


def sweep_flops(desc, n_acc, n_rej) -> float:
    """Total algorithmic flops of a sweep from its per-point step counters."""
    import numpy as np

    acc, rej, _, k = desc_step(desc)
    renorms = desc.n_renorm_transient + desc.n_renorm_average
    na = np.asarray(n_acc, dtype=np.float64)
    nr = np.asarray(n_rej, dtype=np.float64)
    return float(acc * na.sum() + rej * nr.sum() + mgs(desc.dim, k) * renorms * len(na))


def ref_fd_point_flops(desc) -> float:
    """Algorithmic flops of one point in the reference's FD mode
    (cuda/lyapunov.dp.cpp:5-68 restated): the main trajectory alone over the
    transient, then the main and k slave trajectories (fixed-step DVERK, its
    seven live stages) over the slave and averaging phases, with the slaves
    made relative to the main state, Gram-Schmidt-orthogonalised and
    normalised (every step in the slave phase, every add_skip_steps steps in
    the averaging phase) and moved back."""
    n, k = desc.dim, desc.n_exps
    base = per_step(desc.model, n, 0, L.INT_DVERK, False)[1]
    renorm = mgs(n, k) + 2 * k * n
    every = max(1.0, float(desc.add_skip_steps))
    return float(desc.ref_steps_transient * base
                 + desc.ref_steps_slave * ((k + 1) * base + renorm)
                 + desc.ref_steps_average * ((k + 1) * base + renorm / every))
