"""Pins of the oracle's map-type models (Henon, logistic) and of its three
map methods: this build's VAR method (MGS every t_renorm iterations, Kahan
sums) and the reference's VAR and FD map branches (cuda/lyapunov.dp.cpp:
182-222 and 70-122, StepMAP/StepMAPVAR in cuda/tools/mapstep.dp.cpp:5-18).

The reference ships no map model and no test, so the pins are mathematical:
- J v against central differences of the map;
- Henon: det J = -b at every point, so lambda1 + lambda2 = ln b exactly (up to
  rounding), and at a stable fixed point the exponents are ln |mu_i| of the
  Jacobian's eigenvalues there;
- logistic: at a stable fixed point lambda = ln |2 - r|, on the stable 2-cycle
  lambda = 1/2 ln |4 + 2r - r^2| (closed forms), and at r = 4 lambda = ln 2;
- Henon a = 1.4, b = 0.3: lambda1 = 0.4192 (published), finite-time noise allowed.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def map_desc(model, params, method=O.METHOD_VAR, **kw):
    dim = 2 if model == O.MODEL_HENON else 1
    base = dict(model=model, dim=dim, n_exps=dim, n_params=len(params), method=method,
                integrator=O.INT_DVERK, adaptive=0, precision=O.PREC_FP64, h=1.0, eps=1e-8,
                t_transient=1000, t_slave=0, t_average=100000, t_renorm=10, max_steps=10**9,
                escape_radius=1e3, sweep_idx1=0, sweep_idx2=1 if dim == 2 else 0, n1=1, n2=1,
                start1=params[0], delta1=0.0, start2=params[-1], delta2=0.0,
                base_params=list(params), init_state=[0.1, 0.1][:dim], init_kind=O.INIT_NONE,
                init_param_offset=-1, init_scale=0.0, seed=1)
    base.update(kw)
    return O.make_desc(base)


def henon_fixed_point_exponents(a, b):
    x = (-(1 - b) + math.sqrt((1 - b) ** 2 + 4 * a)) / (2 * a)
    J = np.array([[-2 * a * x, 1.0], [b, 0.0]])
    return np.sort(np.log(np.abs(np.linalg.eigvals(J))))[::-1], x


@pytest.mark.parametrize("model,p,x", [(O.MODEL_HENON, [1.4, 0.3], [0.3, -0.2]),
                                       (O.MODEL_LOGISTIC, [3.7], [0.41])])
def test_map_and_jacobian(model, p, x):
    f = O.rhs(model, np.array(x), p)
    if model == O.MODEL_HENON:
        np.testing.assert_allclose(f, [1 - p[0] * x[0] ** 2 + x[1], p[1] * x[0]], rtol=1e-15)
    else:
        np.testing.assert_allclose(f, [p[0] * x[0] * (1 - x[0])], rtol=1e-15)
    n = len(x)
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        h = 1e-6
        fd = (O.rhs(model, np.array(x) + h * e, p) - O.rhs(model, np.array(x) - h * e, p)) / (2 * h)
        np.testing.assert_allclose(O.jvp(model, np.array(x), p, e), fd, atol=1e-8)


@pytest.mark.parametrize("method", [O.METHOD_VAR, O.METHOD_VAR_REF, O.METHOD_FD_REF])
def test_henon_sum_rule_and_published_lambda1(method):
    d = map_desc(O.MODEL_HENON, [1.4, 0.3], method, t_average=200000)
    lam, st, na, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    # lambda1 + lambda2 = <ln |det J|> = ln b
    tol = 1e-9 if method != O.METHOD_FD_REF else 1e-6
    assert abs(lam[0].sum() - math.log(0.3)) < tol
    assert abs(lam[0][0] - 0.4192) < 5e-3
    assert na[0] == (201000 if method != O.METHOD_VAR else 201000)


@pytest.mark.parametrize("method", [O.METHOD_VAR, O.METHOD_VAR_REF, O.METHOD_FD_REF])
def test_henon_stable_fixed_point(method):
    a, b = 0.2, 0.3
    ref, xs = henon_fixed_point_exponents(a, b)
    d = map_desc(O.MODEL_HENON, [a, b], method, t_transient=500, t_average=20000,
                 init_state=[xs + 0.01, b * xs - 0.01], init_kind=O.INIT_NONE)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    np.testing.assert_allclose(lam[0], ref, atol=2e-3)


@pytest.mark.parametrize("method", [O.METHOD_VAR, O.METHOD_VAR_REF, O.METHOD_FD_REF])
@pytest.mark.parametrize("r", [2.5, 3.2, 4.0])
def test_logistic_closed_forms(method, r):
    if r < 3:
        ref = math.log(abs(2 - r))
    elif r < 1 + math.sqrt(6):
        ref = 0.5 * math.log(abs(4 + 2 * r - r * r))
    else:
        ref = math.log(2.0)
    d = map_desc(O.MODEL_LOGISTIC, [r], method, init_state=[0.3], t_average=400000)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    tol = 1e-3 if r < 4 else 6e-3
    assert abs(lam[0][0] - ref) < tol


def test_map_var_mixed_needs_short_renormalisation():
    """The mixed form rounds J v to binary32 every iteration; between two
    renormalisations that error grows by exp((lambda1 - lambda2) * t_renorm)
    relative to the second vector (Henon: exp(2.04 per iteration)).  With
    renormalisation every iteration it stays at the float epsilon; over 10
    iterations (e^20 * 6e-8 >> 1) lambda2 is lost while lambda1 is not."""
    d64 = map_desc(O.MODEL_HENON, [1.4, 0.3], t_renorm=1)
    dmx = map_desc(O.MODEL_HENON, [1.4, 0.3], t_renorm=1, precision=O.PREC_MIXED)
    l64, s64, _, _ = O.sweep(d64, count=1)
    lmx, smx, _, _ = O.sweep(dmx, count=1)
    assert s64[0] == smx[0] == O.OK
    assert np.max(np.abs(l64[0] - lmx[0])) < 5e-3  # the same base orbit, FP32 tangent products
    assert abs(lmx[0].sum() - math.log(0.3)) < 1e-5
    lbad, _, _, _ = O.sweep(map_desc(O.MODEL_HENON, [1.4, 0.3], precision=O.PREC_MIXED), count=1)
    assert abs(lbad[0][0] - l64[0][0]) < 5e-3 and abs(lbad[0][1] - l64[0][1]) > 0.05


def test_map_escape():
    # a = 1.6 leaves the attractor basin: the orbit diverges
    desc = map_desc(O.MODEL_HENON, [1.6, 0.3])
    lam, st, _, _ = O.sweep(desc, count=1)
    assert st[0] == O.ESCAPED and np.all(np.isnan(lam[0]))
