"""Pins of the CPU oracle (oracle/lyap_oracle.c) against what the mathematics
and independent libraries fix.  The reference ships no tests, golden vectors
or fixtures for this path and cannot be built here (it needs a SYCL compiler
and the dpct headers), so every pin below is a closed form, an invariant, a
library routine (numpy/scipy) or a published value, chosen so that a dropped
term, a wrong sign/index or a transposed operand fails at least one of them.
"""
import math

import numpy as np
import pytest
import scipy.integrate
import scipy.linalg

from oracle import oracle as O

rng = np.random.default_rng(7)


# ---------------------------------------------------------------- elementary
def test_log_exp_against_libm():
    L = O.lib()
    xs = np.concatenate([np.exp(rng.uniform(-740, 709, 4000)), [1.0, 2.0, 0.5, 1e-300, 5e-324,
                                                                 1.4142135623730951, 0.7071067811865476]])
    for x in xs:
        ref = math.log(x)
        assert abs(L.orc_log(float(x)) - ref) <= 4.5e-16 * max(1.0, abs(ref))
    for y in rng.uniform(-700, 700, 4000):
        ref = math.exp(y)
        assert abs(L.orc_exp(float(y)) - ref) <= 4.5e-16 * ref
    assert L.orc_log(0.0) == -math.inf and math.isnan(L.orc_log(-1.0)) and L.orc_log(math.inf) == math.inf
    assert L.orc_exp(-800.0) == 0.0 and L.orc_exp(800.0) == math.inf and L.orc_pow(0.0, 0.17) == 0.0
    for x, y in [(0.3, 0.17), (1e-4, 0.04), (7.5, 1.0 / 6.0 - 0.03)]:
        assert abs(L.orc_pow(x, y) - x ** y) <= 2e-15 * x ** y


def test_step_control_log_exp():
    """The controller's shorter-series log/exp: ~1e-10 relative against libm;
    exp clamps its argument to [-1.6, 2.5] (the range the factor limits keep)."""
    L = O.lib()
    rng = np.random.default_rng(3)
    for x in np.concatenate([10.0 ** rng.uniform(-300, 300, 2000), [1e-4, 1.0, 0.5, 2.0, 5e-324]]):
        assert abs(L.orc_ctrl_log(x) - math.log(x)) <= 2e-10 * max(1.0, abs(math.log(x)))
    for x in rng.uniform(-1.6, 2.5, 2000):
        # degree-8 Taylor remainder on |r| <= ln2/2: e^(ln2/2) (ln2/2)^9 / 9! = 3.0e-10
        assert abs(L.orc_ctrl_exp(x) / math.exp(x) - 1.0) < 3.0e-10
    assert L.orc_ctrl_exp(-50.0) == L.orc_ctrl_exp(-1.6)
    assert L.orc_ctrl_exp(1e300) == L.orc_ctrl_exp(2.5)
    assert L.orc_ctrl_log(0.0) == -math.inf


def test_windowed_log_norm_accumulation():
    """The sweep's method multiplies the norms of up to LOG_WINDOW
    renormalisations into mant * 2^ex and takes one log per window.  Pins:
    split_exponent is exact (mant in [1, 2), mant * 2^ex == the product, by
    ldexp) and leaves zero/subnormal/inf/NaN alone; log_scaled is log(mant) +
    ex ln2 to ~1 ulp (math.log, math.ldexp); a window's log equals the sum
    of the logs of its norms (math.fsum) to ~1e-15 relative."""
    rng = np.random.default_rng(5)
    for m in np.concatenate([10.0 ** rng.uniform(-307, 307, 3000), [1.0, 2.0, 1.5, 0.75, 2.2250738585072014e-308]]):
        mant, ex = O.split_exponent(m, 7)
        assert 1.0 <= mant < 2.0
        assert math.ldexp(mant, ex - 7) == m
    for v in (0.0, 5e-324, 1e-310, math.inf):
        assert O.split_exponent(v, 3) == (v, 3)
    mant, ex = O.split_exponent(math.nan, 3)
    assert math.isnan(mant) and ex == 3
    for _ in range(2000):
        mant, ex = rng.uniform(1.0, 2.0), int(rng.integers(-3000, 3000))
        ref = math.log(mant) + ex * math.log(2.0)
        assert abs(O.log_scaled(mant, ex) - ref) <= 2.5e-16 * max(1.0, abs(ref))
    # a window of norms, as the kernels accumulate them
    for _ in range(200):
        norms = np.exp(rng.uniform(-8.0, 8.0, O.LOG_WINDOW))
        mant, ex = 1.0, 0
        for nv in norms:
            mant, ex = O.split_exponent(mant * nv, ex)
        ref = math.fsum(math.log(v) for v in norms)
        assert abs(O.log_scaled(mant, ex) - ref) <= 1e-15 * max(1.0, abs(ref)) * O.LOG_WINDOW


def test_squared_error_norm_of_the_3d_flows():
    """orc_err2_seq (one division) against numpy's mean of (ev/sk)^2 with
    sk = rtol*max(|x|, |xn|) + atol, for n = 1..4 and ranges the controller
    sees; and orc_err_norm (per-component divisions) squared agrees."""
    L = O.lib()
    rng = np.random.default_rng(9)
    for _ in range(3000):
        n = int(rng.integers(1, 5))
        x = rng.uniform(-50, 50, n) * 10.0 ** rng.uniform(-6, 2, n)
        xn = x + rng.normal(0, 1e-3, n)
        ev = rng.normal(0, 1, n) * 10.0 ** rng.uniform(-16, -6, n)
        rtol, atol = 10.0 ** rng.uniform(-12, -6), 10.0 ** rng.uniform(-12, -6)
        sk = rtol * np.maximum(np.abs(x), np.abs(xn)) + atol
        ref = float(np.mean((ev / sk) ** 2))
        got = L.orc_err2_seq(n, O._ptr(O.f64(ev)), O._ptr(O.f64(x)), O._ptr(O.f64(xn)), rtol, atol)
        assert abs(got - ref) <= 8e-16 * n * ref
        e1 = L.orc_err_norm(n, 0, O._ptr(O.f64(ev)), O._ptr(O.f64(x)), O._ptr(O.f64(xn)), rtol, atol)
        assert abs(e1 * e1 - got) <= 1e-14 * got
    z = np.zeros(3)
    assert L.orc_err2_seq(3, O._ptr(z), O._ptr(O.f64([1.0, 2.0, 3.0])), O._ptr(O.f64([1.0, 2.0, 3.0])),
                          1e-9, 1e-9) == 0.0


def test_loop_count_matches_float_loop():
    # cuda/lyapunov.dp.cpp iterates `for (double t = t0; t < T; t += h)`
    def brute(t0, t1, h):
        n, t = 0, t0
        while t < t1:
            n += 1
            t += h
        return n
    assert O.lib().orc_loop_count(0.0, 1000.0, 0.01) == 100001
    assert O.lib().orc_loop_count(0.0, 100.0, 0.01) == 10000
    for t0, t1, h in [(0, 7.3, 0.01), (1000, 1000 + 37.1, 0.01), (0, 5, 1e-3), (0, 1, 0.1)]:
        assert O.lib().orc_loop_count(float(t0), float(t1), h) == brute(float(t0), float(t1), h)


def test_uniform_generator_statistics_and_determinism():
    L = O.lib()
    u = np.array([L.orc_uniform(99, i, j) for i in range(3000) for j in range(3)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    assert L.orc_uniform(99, 5, 1) == L.orc_uniform(99, 5, 1) != L.orc_uniform(98, 5, 1)
    # adjacent streams and points are uncorrelated
    a = np.array([L.orc_uniform(3, i, 0) for i in range(5000)])
    b = np.array([L.orc_uniform(3, i, 1) for i in range(5000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05


def lorenz_desc(**kw):
    base = dict(model=O.MODEL_LORENZ, dim=3, n_exps=3, n_params=3, method=O.METHOD_VAR,
                integrator=O.INT_DOPRI5, adaptive=0, precision=O.PREC_FP64, h=0.01, rtol=1e-9,
                atol=1e-9, h_min=1e-12, h_max=1.0, eps=1e-7, t_transient=50, t_slave=0,
                t_average=200, t_renorm=1.0, max_steps=10**9, escape_radius=1e8, sweep_idx1=1,
                sweep_idx2=2, n1=1, n2=1, start1=28.0, delta1=0.0, start2=8 / 3, delta2=0.0,
                base_params=[10, 28, 8 / 3], init_state=[1, 1, 1], init_kind=O.INIT_UNIFORM,
                init_param_offset=-1, init_scale=1e-3, seed=1234)
    base.update(kw)
    return O.make_desc(base)


def test_grid_rule_p1_outer():
    # tasklyapunov.cpp:45-63: p = start + i*|end-start|/(n-1), first parameter outer
    n1, n2 = 5, 7
    d = lorenz_desc(n1=n1, n2=n2, start1=0.0, delta1=200.0 / (n1 - 1), start2=0.5,
                    delta2=3.5 / (n2 - 1))
    for idx in range(n1 * n2):
        p = O.point_params(d, idx)
        i1, i2 = divmod(idx, n2)
        assert p[0] == 10.0
        assert abs(p[1] - np.linspace(0, 200, n1)[i1]) <= 1e-12
        assert abs(p[2] - np.linspace(0.5, 4, n2)[i2]) <= 1e-12


def test_initial_state_rules():
    d = lorenz_desc(init_kind=O.INIT_UNIFORM, init_scale=1e-3)
    xs = np.array([O.point_init(d, i) for i in range(2000)])
    assert np.all(np.abs(xs - 1.0) <= 1e-3) and abs(xs.mean() - 1.0) < 5e-5
    d = lorenz_desc(init_kind=O.INIT_NORMAL, init_scale=0.1, init_state=[0, 0, 0],
                    init_param_offset=0)
    xs = np.array([O.point_init(d, i) for i in range(4000)])
    assert abs(xs.mean() - 10.0) < 0.01 and abs(xs.std() - 0.1) < 0.005  # F + N(0, 0.1^2)
    d = lorenz_desc(init_kind=O.INIT_NONE)
    assert np.array_equal(O.point_init(d, 17), [1.0, 1.0, 1.0])  # tasklyapunov.cpp:67-71


# ---------------------------------------------------------------- models
MODELS = [
    (O.MODEL_LORENZ, [10.0, 28.0, 8 / 3], 3,
     lambda x, p: [p[0] * (x[1] - x[0]), x[0] * (p[1] - x[2]) - x[1], x[0] * x[1] - p[2] * x[2]]),
    (O.MODEL_ROSSLER, [0.2, 5.7, 0.2], 3,
     lambda x, p: [-x[1] - x[2], x[0] + p[0] * x[1], p[2] + x[2] * (x[0] - p[1])]),
    # cuda/models/flows/shimizu.dp.cpp:15-19
    (O.MODEL_SHIMIZU, [0.75, 0.45], 3,
     lambda x, p: [x[1], -p[0] * x[1] - x[0] * x[2] + x[0], -p[1] * x[2] + x[0] ** 2]),
    (O.MODEL_L96, [8.0, 1.0], 8,
     lambda x, p: [(x[(i + 1) % 8] - x[i - 2]) * x[i - 1] - p[1] * x[i] + p[0] for i in range(8)]),
]


@pytest.mark.parametrize("model,p,n,f", MODELS)
def test_rhs_and_jacobian(model, p, n, f):
    for _ in range(5):
        x = rng.normal(size=n) * 3
        np.testing.assert_allclose(O.rhs(model, x, p), f(x, p), rtol=1e-14, atol=1e-13)
        v = rng.normal(size=n)
        d = 1e-6
        fd = (O.rhs(model, x + d * v, p) - O.rhs(model, x - d * v, p)) / (2 * d)
        np.testing.assert_allclose(O.jvp(model, x, p, v), fd, rtol=1e-7, atol=1e-7)


# ---------------------------------------------------------------- RK steps
def exact_flow(model, p, x0, h):
    sol = scipy.integrate.solve_ivp(lambda t, y: O.rhs(model, y, p), (0, h), x0, method="DOP853",
                                    rtol=1e-14, atol=1e-16)
    return sol.y[:, -1]


@pytest.mark.parametrize("integ,order", [(O.INT_DVERK, 6), (O.INT_DOPRI5, 5)])
@pytest.mark.parametrize("model,p,n,f", MODELS[:3])
def test_local_error_order(integ, order, model, p, n, f):
    x0 = np.array([1.0, 2.0, 3.0]) if model != O.MODEL_ROSSLER else np.array([1.0, -2.0, 0.5])
    # step pairs in each model's asymptotic range (above roundoff, below stiffness)
    hs = {O.MODEL_LORENZ: [0.016, 0.008], O.MODEL_ROSSLER: [0.064, 0.032],
          O.MODEL_SHIMIZU: [0.128, 0.064]}[model]
    errs, ests = [], []
    for h in hs:
        xn, ev, _ = O.rk_step(model, integ, x0, p, h)
        errs.append(np.linalg.norm(xn - exact_flow(model, p, x0, h)))
        ests.append(np.linalg.norm(ev))
    # local error of the propagated solution ~ h^(order+1), estimate ~ h^order
    assert math.log2(errs[0] / errs[1]) > order + 1 - 0.7
    assert math.log2(ests[0] / ests[1]) > order - 0.7


def test_dopri5_step_matches_scipy_tableau():
    from scipy.integrate._ivp import rk as srk
    for model, p, n, f in MODELS:
        x0 = rng.normal(size=n)
        h = 0.013
        fun = lambda t, y: O.rhs(model, y, p)
        K = np.empty((srk.RK45.n_stages + 1, n))
        y_new, f_new = srk.rk_step(fun, 0.0, x0, fun(0, x0), h, srk.RK45.A, srk.RK45.B, srk.RK45.C, K)
        xn, ev, _ = O.rk_step(model, O.INT_DOPRI5, x0, p, h)
        np.testing.assert_allclose(xn, y_new, rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(ev, -h * (K.T @ srk.RK45.E), rtol=1e-10, atol=1e-17)


@pytest.mark.parametrize("integ", [O.INT_DVERK, O.INT_DOPRI5])
@pytest.mark.parametrize("model,p,n,f", MODELS)
def test_coupled_tangent_is_derivative_of_step(integ, model, p, n, f):
    # J at every stage state: the tangent step is D(step)(x) v exactly
    x0 = rng.normal(size=n) + 1.0
    v = rng.normal(size=n)
    h = 0.02
    _, _, sx = O.rk_step(model, integ, x0, p, h)
    tv = O.tangent_step(model, integ, sx, p, h, v)
    d = 1e-6
    fd = (O.rk_step(model, integ, x0 + d * v, p, h)[0] - O.rk_step(model, integ, x0 - d * v, p, h)[0]) / (2 * d)
    np.testing.assert_allclose(tv, fd, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("model,p,x0,h", [(O.MODEL_LORENZ, [10.0, 28.0, 8 / 3], [5.0, 8.0, 20.0], 0.1),
                                          (O.MODEL_ROSSLER, [0.2, 5.7, 0.2], [1.0, 2.0, 3.0], 0.3),
                                          (O.MODEL_SHIMIZU, [0.75, 0.45], [1.0, 2.0, 3.0], 0.5)])
def test_low_order_tangent_is_derivative_of_the_fifth_order_step(model, p, x0, h):
    # the mixed form's tangent with adaptive DVERK steps takes the embedded
    # fifth-order weights: it is the derivative of x5 = x_new - err (b - e),
    # not of the sixth-order x_new; at a step large enough for the two
    # solutions' derivatives to differ far beyond binary32 rounding, only the
    # right one agrees with central differences
    x0 = np.array(x0)
    v = np.array([0.3, -0.7, 0.5])
    _, _, sx = O.rk_step(model, O.INT_DVERK, x0, p, h)

    def fd(g):
        d = 1e-6
        return (g(x0 + d * v) - g(x0 - d * v)) / (2 * d)

    def x6(x):
        return O.rk_step(model, O.INT_DVERK, x, p, h)[0]

    def x5(x):
        xn, ev, _ = O.rk_step(model, O.INT_DVERK, x, p, h)
        return xn - ev

    d5, d6 = fd(x5), fd(x6)
    scale = np.abs(d6).max()
    tol = 1e-6 * scale
    assert np.abs(d5 - d6).max() > 100 * tol  # the test can tell the two apart
    low = O.tangent_step_f32(model, O.INT_DVERK, sx, p, h, v, low=True)
    full = O.tangent_step_f32(model, O.INT_DVERK, sx, p, h, v, low=False)
    assert np.abs(low - d5).max() < tol
    assert np.abs(full - d6).max() < tol
    # the FP64 tangent keeps the sixth-order weights
    np.testing.assert_allclose(O.tangent_step(model, O.INT_DVERK, sx, p, h, v), d6, rtol=1e-7,
                               atol=1e-8 * scale)


def test_fifth_order_weights_are_b_minus_e_and_fifth_order():
    from fractions import Fraction as Fr

    A = [[0], [Fr(1, 6)], [Fr(4, 75), Fr(16, 75)], [Fr(5, 6), Fr(-8, 3), Fr(5, 2)],
         [Fr(-165, 64), Fr(55, 6), Fr(-425, 64), Fr(85, 96)],
         [Fr(12, 5), Fr(-8), Fr(4015, 612), Fr(-11, 36), Fr(88, 255)]]
    b = [Fr(3, 40), 0, Fr(875, 2244), Fr(23, 72), Fr(264, 1955), 0, Fr(125, 11592), Fr(43, 616)]
    e = [Fr(-1, 160), 0, Fr(-125, 17952), Fr(1, 144), Fr(-12, 1955), Fr(-3, 44), Fr(125, 11592),
         Fr(43, 616)]
    bh = [bj - ej for bj, ej in zip(b, e)]
    assert bh[6] == 0 and bh[7] == 0  # stages 7 and 8 drop out
    assert bh[:6] == [Fr(13, 160), 0, Fr(2375, 5984), Fr(5, 16), Fr(12, 85), Fr(3, 44)]
    c = [sum(row, Fr(0)) for row in A]
    for q in range(1, 6):
        assert sum(bh[i] * c[i] ** (q - 1) for i in range(6)) == Fr(1, q)
    assert sum(bh[i] * c[i] ** 5 for i in range(6)) != Fr(1, 6)  # fifth order, not sixth


@pytest.mark.parametrize("model,p,n,f", MODELS)
def test_frozen_tangent_is_dverk_on_linear_system(model, p, n, f):
    # cuda/tools/dverk.dp.cpp:46-80 applies DVERK to v' = J(x*) v: exp(hJ) v to O(h^7)
    x0 = rng.normal(size=n)
    J = np.column_stack([O.jvp(model, x0, p, e) for e in np.eye(n)])
    v = rng.normal(size=n)
    for h in (0.01, 0.005):
        got = O.frozen_tangent_step(model, x0, p, h, v)
        ref = scipy.linalg.expm(h * J) @ v
        tol = (h * np.linalg.norm(J, 2)) ** 7 / 5040 * np.linalg.norm(v) * 4 + 1e-14
        assert np.linalg.norm(got - ref) <= tol


# ---------------------------------------------------------------- orthogonalisation
def test_inverse_square_root_of_the_renormalisation():
    """orc_inv_sqrt (seed + four Newton steps) against libm's correctly rounded
    sqrt: within 3 * 2^-52 relative over squared norms from 1e-60 to 1e60 and
    exact on powers of four; three steps would leave ~3e-11, a wrong seed shift
    far more."""
    f = O.lib().orc_inv_sqrt
    s = np.exp(rng.uniform(-138.0, 138.0, 20000))
    s = np.concatenate([s, [1.0, 4.0, 0.25, 2.0, 3.0, 9.0, 1e-300 * 1e10, 7.0 ** 2]])
    y = np.array([f(float(v)) for v in s])
    ref = 1.0 / np.sqrt(s)
    rel = np.abs(y - ref) / ref
    assert rel.max() <= 3 * 2.0 ** -52
    # the norm the renormalisation records, s * y, against sqrt(s)
    nrm = s * y
    assert (np.abs(nrm - np.sqrt(s)) / np.sqrt(s)).max() <= 3 * 2.0 ** -52
    assert f(4.0) == 0.5 and f(0.25) == 2.0


def test_mgs_matches_numpy_qr():
    for n, k in [(3, 3), (8, 8), (16, 16), (32, 32), (5, 3)]:
        A = rng.normal(size=(n, k))
        Q, R = np.linalg.qr(A)
        for tree in (False, True):
            if tree and n & (n - 1):
                continue
            V, ln = O.mgs(A.T, tree=tree)
            np.testing.assert_allclose(ln, np.log(np.abs(np.diag(R))), rtol=0, atol=1e-12)
            np.testing.assert_allclose(np.abs(V), np.abs(Q.T), atol=1e-11)
            np.testing.assert_allclose(V @ V.T, np.eye(k), atol=1e-12)


def test_reference_cgs_normalize_and_qr():
    A = rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(A)
    # cuda/tools/linalg.dp.cpp:36-51: row i -> R_ii q_i (unnormalised orthogonal rows)
    V = O.cgs_ref(A.T)
    np.testing.assert_allclose(np.abs(V), np.abs(np.diag(R))[:, None] * np.abs(Q.T), atol=1e-12)
    # :20-28: rows scaled to eps
    W = O.normalize_ref(V, 1e-7)
    np.testing.assert_allclose(np.linalg.norm(W, axis=1), 1e-7, rtol=1e-14)
    # :78-128: columns are the vectors; diag(R) up to sign, Q orthonormal
    Qo, rd = O.qr_ref(A)
    np.testing.assert_allclose(np.abs(rd), np.abs(np.diag(R)), rtol=1e-12)
    np.testing.assert_allclose(Qo.T @ Qo, np.eye(3), atol=1e-12)


# ---------------------------------------------------------------- spectra
def lorenz_fixed_point_spectrum(sigma, rho, beta):
    if rho < 1:
        J = np.array([[-sigma, sigma, 0], [rho, -1, 0], [0, 0, -beta]])
    else:
        c = math.sqrt(beta * (rho - 1))
        J = np.array([[-sigma, sigma, 0], [1, -1, -c], [c, c, -beta]])
    return np.sort(np.linalg.eigvals(J).real)[::-1]


@pytest.mark.parametrize("rho", [0.5, 10.0])
@pytest.mark.parametrize("method", [O.METHOD_VAR, O.METHOD_VAR_REF, O.METHOD_QR_REF, O.METHOD_FD_REF])
def test_stable_fixed_point_spectrum(rho, method):
    # the reference QR mode has no tangent transient (cuda/lyapunov.dp.cpp:239-254): at a
    # decoupled fixed point its second column needs ~ln(1e12)/7.9 time units to turn onto
    # the z direction, an O(1/T) bias of ~28/T, so it is run ten times longer
    t_avg = 4000 if method == O.METHOD_QR_REF else 400
    d = lorenz_desc(method=method, integrator=O.INT_DVERK if method != O.METHOD_VAR else O.INT_DOPRI5,
                    start1=rho, base_params=[10, rho, 8 / 3], t_transient=60, t_slave=20,
                    t_average=t_avg, t_renorm=0.1 if method != O.METHOD_VAR else 1.0,
                    init_state=[1, 1, 1] if rho < 1 else [7.9, 7.9, 8.8], eps=1e-8)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    np.testing.assert_allclose(lam[0], lorenz_fixed_point_spectrum(10, rho, 8 / 3), atol=0.03)


@pytest.mark.parametrize("integ,adaptive", [(O.INT_DOPRI5, 0), (O.INT_DVERK, 0), (O.INT_DOPRI5, 1),
                                             (O.INT_DVERK, 1)])
def test_lorenz_chaotic_spectrum(integ, adaptive):
    d = lorenz_desc(integrator=integ, adaptive=adaptive, t_transient=100, t_average=1000)
    lam, st, na, nr = O.sweep(d, count=1)
    assert st[0] == O.OK
    # trace invariant: sum(lambda) = <tr J> = -(sigma + 1 + beta)
    assert abs(lam[0].sum() + (11 + 8 / 3)) < 1e-6
    assert abs(lam[0][1]) < 2e-3  # the flow direction
    assert abs(lam[0][0] - 0.9056) < 0.03  # Sprott's published value, finite-time noise allowed
    if adaptive:
        assert na[0] > 0


def test_rossler_and_l96_spectra():
    d = lorenz_desc(model=O.MODEL_ROSSLER, base_params=[0.2, 5.7, 0.2], sweep_idx1=0, sweep_idx2=1,
                    start1=0.2, start2=5.7, integrator=O.INT_DVERK, adaptive=1, rtol=1e-10, atol=1e-10,
                    t_transient=200, t_average=2000, t_renorm=0.5, init_state=[0, 0, 0], init_scale=1.0)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    assert abs(lam[0][0] - 0.0714) < 0.01 and abs(lam[0][1]) < 2e-3 and abs(lam[0][2] + 5.3943) < 0.05
    n = 8
    d = lorenz_desc(model=O.MODEL_L96, dim=n, n_exps=n, n_params=2, base_params=[8.0, 1.0],
                    sweep_idx1=0, sweep_idx2=1, start1=8.0, start2=1.0, integrator=O.INT_DOPRI5,
                    adaptive=1, t_transient=20, t_average=100, t_renorm=0.2, init_state=[0] * n,
                    init_kind=O.INIT_NORMAL, init_param_offset=0, init_scale=0.1)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] == O.OK
    assert abs(lam[0].sum() + n * 1.0) < 1e-6  # tr J = -N d
    assert lam[0][0] > 0.5 and np.all(np.diff(lam[0]) <= 1e-2)


def test_mixed_form_close_to_fp64():
    d = lorenz_desc(t_transient=50, t_average=300)
    lam64, _, _, _ = O.sweep(d, count=1)
    d.precision = O.PREC_MIXED
    lam32, _, _, _ = O.sweep(d, count=1)
    assert abs(lam32[0][0] - lam64[0][0]) < 1e-5


def test_adaptive_controller_rejects_and_meets_tolerance():
    # a deliberately huge first step must be rejected; tightening rtol moves
    # lambda_1,2 toward the small fixed-step value; at this stable focus the
    # step is held by the h*||J||_inf <= 1 cap, which bounds the lambda_3 error
    fp = dict(start1=10.0, base_params=[10, 10, 8 / 3], init_state=[7.9, 7.9, 8.8],
              t_transient=10, t_average=50)
    _, st, na, nr = O.sweep(lorenz_desc(adaptive=1, h=0.5, **fp), count=1)
    assert st[0] == O.OK and nr[0] >= 1
    ref = O.sweep(lorenz_desc(h=0.001, **fp), count=1)[0][0]
    errs = []
    for tol in (1e-4, 1e-9):
        lam = O.sweep(lorenz_desc(adaptive=1, rtol=tol, atol=tol, **fp), count=1)[0][0]
        errs.append(np.max(np.abs(lam[:2] - ref[:2])))
        assert abs(lam[2] - ref[2]) < 2e-3
    assert errs[1] <= errs[0] and errs[1] < 2e-5


def test_step_cap_bounds_steps_by_jacobian_norm():
    # without the cap the base error estimate vanishes at the fixed point and
    # steps grow to h_max; with it h*||J|| <= 1 (mean step ~ 1/||J||)
    d = lorenz_desc(adaptive=1, start1=10.0, base_params=[10, 10, 8 / 3],
                    init_state=[7.9, 7.9, 8.8], t_transient=10, t_average=50)
    _, st, na, _ = O.sweep(d, count=1)
    c = math.sqrt(8 / 3 * 9)
    jn = max(20.0, 1 + 1 + c, c + c + 8 / 3)
    assert st[0] == O.OK and na[0] >= 60 * jn * 0.9


# ------------------------------------------------------- step controller
def reference_controller(integ, err, fac_old, is_reject, step, max_step, beta=0.04, safe=0.9,
                         fac1=0.2, fac2=10.0, q=None):
    """dopriStep's step-size rule as the reference writes it, with pow
    (cuda/tools/dopri.dp.cpp:9-12,27-29,109-139): expo1 = 1/q - 0.75*beta,
    fac = max(1/fac2, min(1/fac1, err^expo1 / facOld^beta / safe)), newStep =
    step / fac; accepted (err <= 1): facOld = max(err, 1e-4), |newStep| capped
    at maxStep, and no growth right after a rejection; rejected: newStep =
    step / min(1/fac1, err^expo1 / safe).  q = 5 for Dopri5 as in the
    reference; 6 for DVERK's 6(5) pair.  Returns (accepted, new step, facOld,
    rejected flag)."""
    if q is None:
        q = 6.0 if integ == O.INT_DVERK else 5.0
    expo1 = 1.0 / q - beta * 0.75
    facc1, facc2 = 1.0 / fac1, 1.0 / fac2
    fac11 = err ** expo1
    fac = fac11 / fac_old ** beta
    fac = max(facc2, min(facc1, fac / safe))
    new = step / fac
    if err <= 1.0:
        fac_old = max(err, 1e-4)
        if abs(new) > max_step:
            new = max_step
        if is_reject:
            new = min(abs(new), abs(step))
        return True, new, fac_old, False
    return False, step / min(facc1, fac11 / safe), fac_old, True


def controller_cases():
    """Single decisions (random and boundary states) and sequences that carry
    the state (facOld, rejected flag) through accepts and rejects."""
    rng = np.random.default_rng(11)
    single = []
    for _ in range(3000):
        single.append((10.0 ** rng.uniform(-14, 4), 10.0 ** rng.uniform(-4, 0), bool(rng.integers(2)),
                       10.0 ** rng.uniform(-5, 0), 0.5))
    # err = 1 (accept edge), just above 1, the 1e-4 floor and either side of it,
    # growth limit (tiny err), shrink limit on rejection (huge err), the h_max
    # cap, reject after reject
    for e in (1.0, np.nextafter(1.0, 2.0), 1e-4, 0.99e-4, 1.01e-4, 1e-30, 1e-12, 1e4, 1e8):
        for fo in (1e-4, 0.3, 1.0):
            for rej in (False, True):
                single.append((e, fo, rej, 0.01, 0.5))
                single.append((e, fo, rej, 0.45, 0.5))
    seqs = [10.0 ** rng.uniform(-10, 3, 200) for _ in range(20)]
    seqs.append(np.array([5.0, 7.0, 2.0, 0.5, 3.0, 3.0, 1e-9, 1e-9, 0.9, 1e5]))
    return single, seqs


def controller_mismatch(integ, rule):
    """Largest relative difference between the oracle's controller and `rule`
    over controller_cases() (next step, and facOld on accepts)."""
    single, seqs = controller_cases()
    worst = 0.0
    for err, fo, rej, h, hmax in single:
        acc, hn, lfo, rj = O.ctrl_update(integ, err, O.lib().orc_ctrl_log(fo), rej, h, hmax)
        racc, rhn, rfo, rrj = rule(integ, err, fo, rej, h, hmax)
        if acc != racc or rj != rrj:
            return math.inf
        worst = max(worst, abs(hn - rhn) / abs(rhn), abs(lfo - math.log(rfo)) / max(1.0, abs(math.log(rfo))))
    for seq in seqs:
        lfo, rej, fo, rrej, h, rh = O.lib().orc_ctrl_log(1e-4), False, 1e-4, False, 0.01, 0.01
        for err in seq:
            acc, h, lfo, rej = O.ctrl_update(integ, err, lfo, rej, rh, 0.5)
            racc, rh, fo, rrej = rule(integ, err, fo, rrej, rh, 0.5)
            if acc != racc or rej != rrej:
                return math.inf
            worst = max(worst, abs(h - rh) / abs(rh))
            rh = h  # both continue from the oracle's step (no drift)
    return worst


@pytest.mark.parametrize("integ", [O.INT_DVERK, O.INT_DOPRI5])
def test_controller_matches_reference_pow_rule(integ):
    """The oracle's one-exponential controller (the form the kernels share)
    against the reference's pow form: same accept/reject decisions, next steps
    within 2e-9 relative (the controller's own log/exp series are 2e-10 /
    3e-10 accurate, pinned in test_step_control_log_exp, and enter through
    expo * log(err) with |log err| up to ~70)."""
    assert controller_mismatch(integ, reference_controller) <= 2e-9


@pytest.mark.parametrize("integ", [O.INT_DVERK, O.INT_DOPRI5])
def test_controller_comparison_catches_planted_mistakes(integ):
    """The comparison above is sharp: the reference rule with beta's sign
    flipped, with the other integrator's exponent, with the factor limits
    swapped, with no floor on facOld, or without the no-growth-after-reject
    rule differs from the oracle by far more than its 2e-9 bound."""
    mutants = [
        lambda i, e, f, r, h, m: reference_controller(i, e, f, r, h, m, beta=-0.04),
        lambda i, e, f, r, h, m: reference_controller(i, e, f, r, h, m,
                                                      q=5.0 if i == O.INT_DVERK else 6.0),
        lambda i, e, f, r, h, m: reference_controller(i, e, f, r, h, m, fac1=0.1, fac2=5.0),
        lambda i, e, f, r, h, m: (lambda a: (a[0], a[1], max(e, 1e-12) if a[0] else a[2], a[3]))(
            reference_controller(i, e, f, r, h, m)),
        lambda i, e, f, r, h, m: reference_controller(i, e, f, False, h, m)[:3] + (not (e <= 1.0),),
    ]
    for mut in mutants:
        assert controller_mismatch(integ, mut) > 1e-3


@pytest.mark.parametrize("model,n", [(O.MODEL_SHIMIZU, 3), (O.MODEL_LORENZ, 3), (O.MODEL_ROSSLER, 3),
                                     (O.MODEL_L96, 8), (O.MODEL_L96, 16), (O.MODEL_L96, 32)])
def test_step_cap_norm_is_the_jacobian_inf_norm(model, n):
    """The adaptive step cap h * ||J(x)||_inf <= 1 uses orc_jac_norm: the
    largest row sum of |J|, with J assembled column by column from the
    Jacobian-vector product (itself pinned against central differences of the
    right-hand side) and normed by numpy.linalg.norm(J, inf)."""
    rng = np.random.default_rng(100 + model * 7 + n)
    npar = {O.MODEL_SHIMIZU: 2, O.MODEL_LORENZ: 3, O.MODEL_ROSSLER: 3, O.MODEL_L96: 2}[model]
    for _ in range(50):
        x = rng.normal(0, 5, n)
        p = rng.uniform(-3, 30, npar)
        J = np.stack([O.jvp(model, x, p, np.eye(n)[j]) for j in range(n)], axis=1)
        ref = np.linalg.norm(J, np.inf)
        assert abs(O.jac_norm(model, x, p) - ref) <= 4e-16 * ref


# ---------------------------------------------------------------- trace closure
@pytest.mark.parametrize("model,n", [(O.MODEL_SHIMIZU, 3), (O.MODEL_LORENZ, 3), (O.MODEL_ROSSLER, 3)])
def test_trace_is_the_jacobian_trace(model, n):
    """orc_trace = the trace of J assembled from the Jacobian-vector product
    (pinned against central differences of the right-hand side)."""
    rng = np.random.default_rng(200 + model)
    npar = {O.MODEL_SHIMIZU: 2, O.MODEL_LORENZ: 3, O.MODEL_ROSSLER: 3}[model]
    for _ in range(50):
        x = rng.normal(0, 5, n)
        p = rng.uniform(-3, 30, npar)
        J = np.stack([O.jvp(model, x, p, np.eye(n)[j]) for j in range(n)], axis=1)
        assert abs(O.trace(model, x, p) - np.trace(J)) <= 1e-14 * (1 + np.abs(J).sum())


@pytest.mark.parametrize("integ", [O.INT_DVERK, O.INT_DOPRI5])
@pytest.mark.parametrize("model,p,x0", [(O.MODEL_ROSSLER, [0.2, 5.7, 0.2], [1.0, -2.0, 0.5]),
                                         (O.MODEL_LORENZ, [10.0, 28.0, 8 / 3], [1.0, 1.0, 20.0])])
def test_trace_step_integrates_the_divergence(integ, model, p, x0):
    """The sum of orc_trace_step over 300 steps of h = 0.01 equals the
    integral of tr J along the trajectory, computed by scipy's DOP853 on the
    system augmented with y' = tr J(x) (a wrong weight, stage or factor h
    would miss it by far more than the steps' own 1e-9)."""
    x = np.array(x0, dtype=np.float64)
    h, steps, acc = 0.01, 300, 0.0
    for _ in range(steps):
        xn, _, sx = O.rk_step(model, integ, x, p, h, want_err=True)
        acc += O.trace_step(model, integ, sx, p, h)
        x = xn

    def aug(t, y):
        return list(O.rhs(model, y[:3], p)) + [O.trace(model, y[:3], p)]

    sol = scipy.integrate.solve_ivp(aug, (0, h * steps), list(x0) + [0.0], method="DOP853",
                                    rtol=1e-13, atol=1e-14)
    assert abs(acc - sol.y[3, -1]) < 1e-8 * max(1.0, abs(sol.y[3, -1]))


def test_trace_closure_at_a_stable_fixed_point():
    """Lorenz rho = 0.5 (stable origin, J constant along the converged
    trajectory): the closed exponent is the third eigenvalue of J(0) (a
    closed form), the first two are the vector method's bit for bit."""
    ref = lorenz_fixed_point_spectrum(10, 0.5, 8 / 3)
    kw = dict(start1=0.5, base_params=[10, 0.5, 8 / 3], t_transient=60, t_average=400,
              init_state=[1, 1, 1])
    lam_v, st_v, _, _ = O.sweep(lorenz_desc(method=O.METHOD_VAR, **kw), count=1)
    lam_t, st_t, _, _ = O.sweep(lorenz_desc(method=O.METHOD_VAR_TRACE, **kw), count=1)
    assert st_v[0] == O.OK and st_t[0] == O.OK
    assert np.array_equal(lam_v[0][:2], lam_t[0][:2])
    assert abs(lam_t[0][2] - ref[2]) < 1e-8
    np.testing.assert_allclose(lam_t[0], ref, atol=1e-6)


@pytest.mark.parametrize("prec", [O.PREC_FP64, O.PREC_MIXED])
def test_trace_closure_on_chaotic_orbits(prec):
    """Chaotic Lorenz and Rossler orbits: the closed spectrum sums to the
    trace's time average (Lorenz: the constant -(sigma + 1 + beta)), its
    first two exponents are the vector method's (same precision) bit for bit,
    and at a small fixed step its last exponent agrees with the one the third
    deviation vector measures in FP64.  In binary32 the third vector itself
    loses the most contracting direction (off by ~3e-2 on this Lorenz orbit);
    the closure is off only by the first two exponents' binary32 error."""
    def pair(**kw):
        v64 = O.sweep(lorenz_desc(method=O.METHOD_VAR, **dict(kw, precision=O.PREC_FP64)), count=1)
        v = O.sweep(lorenz_desc(method=O.METHOD_VAR, **dict(kw, precision=prec)), count=1)
        t = O.sweep(lorenz_desc(method=O.METHOD_VAR_TRACE, **dict(kw, precision=prec)), count=1)
        assert v64[1][0] == O.OK and v[1][0] == O.OK and t[1][0] == O.OK
        assert np.array_equal(v[0][0][:2], t[0][0][:2])
        return v64[0][0], v[0][0], t[0][0]

    tol = 1e-6 if prec == O.PREC_FP64 else 2e-5
    lam64, lam_v, lam_t = pair(t_transient=50, t_average=400, h=0.002)
    assert abs(lam_t.sum() + (11 + 8 / 3)) < 1e-9
    assert abs(lam_t[2] - lam64[2]) < tol
    if prec == O.PREC_MIXED:
        assert abs(lam_v[2] - lam64[2]) > 1e-2  # the binary32 third vector
    lam64, lam_v, lam_t = pair(model=O.MODEL_ROSSLER, base_params=[0.2, 5.7, 0.2], sweep_idx1=0,
                               sweep_idx2=1, start1=0.2, start2=5.7, integrator=O.INT_DVERK, h=0.005,
                               t_transient=100, t_average=400, t_renorm=0.5, init_state=[0, 0, 0],
                               init_scale=1.0)
    assert abs(lam_t[2] + 5.3943) < 0.05
    assert abs(lam_t[2] - lam64[2]) < tol


def test_status_words():
    # escaping orbit (Rossler with a large a), step underflow, step budget
    d = lorenz_desc(model=O.MODEL_ROSSLER, base_params=[0.45, 4.0, 0.2], sweep_idx1=0, sweep_idx2=1,
                    start1=0.45, start2=4.0, integrator=O.INT_DVERK, adaptive=1, rtol=1e-10,
                    atol=1e-10, t_renorm=0.5, init_state=[5, 5, 5], init_scale=0.0, escape_radius=1e3)
    lam, st, _, _ = O.sweep(d, count=1)
    assert st[0] in (O.ESCAPED, O.NONFINITE) and np.all(np.isnan(lam[0]))
    d = lorenz_desc(adaptive=1, rtol=1e-30, atol=1e-30, h_min=1e-6)
    assert O.sweep(d, count=1)[1][0] == O.STEP_UNDERFLOW
    d = lorenz_desc(adaptive=1, max_steps=100)
    assert O.sweep(d, count=1)[1][0] == O.MAXSTEPS
