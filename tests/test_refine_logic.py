"""Host logic of the surrogate refinement: which points the surrogate may
answer and the calibration of its std multiplier (no GPU)."""
import numpy as np

from chaos_oneapi_amd import refine as R

THR = dict(std_thr=0.01, lam_thr=0.02, bad_lo=0.02, bad_hi=0.98)


def pred_rows(mu1, std1, pbad):
    n = len(mu1)
    p = np.zeros((n, 5), dtype=np.float32)
    p[:, 0], p[:, 3], p[:, 4] = mu1, std1, pbad
    return p


def test_answerable_rules():
    mu1 = np.array([0.1, 0.1, 0.01, 0.1, 0.1, 0.1])
    std = np.array([0.005, 0.02, 0.001, 0.005, 0.005, 0.009])
    pbad = np.array([0.0, 0.0, 0.0, 0.5, 0.99, 0.0])
    a1 = R.answerable(mu1, std, pbad, 1.0, **THR)
    assert a1.tolist() == [True, False, False, False, True, True]
    a2 = R.answerable(mu1, std, pbad, 2.0, **THR)
    assert a2.tolist() == [True, False, False, False, True, False]  # 2*0.009 > 0.01


def test_calibration_raises_kappa_until_answers_agree():
    rng = np.random.default_rng(0)
    n = 20000
    lam1 = rng.uniform(0.03, 0.2, n)           # all chaotic
    mu1 = lam1.copy()
    std = rng.uniform(0.0, 0.01, n)
    # the surrogate is wrong (regular) on points it is least sure of
    wrong = std > 0.008
    mu1[wrong] = -0.05
    lam = np.stack([lam1, np.zeros(n), np.full(n, -1.0)], 1)
    st = np.zeros(n, dtype=np.uint32)
    kappa, escapes, agree, answered, table = R.calibrate_kappa(
        pred_rows(mu1, std, np.zeros(n)), lam, st, target=0.9995, **THR)
    # kappa 1 answers the wrong points (std in (0.008, 0.01]) and 1.1 those in
    # (0.008, 0.0091], far over the budget of 10 wrong answers;
    # 1.25*std <= 0.01 <=> std <= 0.008 has none
    assert kappa == 1.25 and agree == 1.0
    assert answered == int(np.sum(1.25 * std <= 0.01))
    assert len(table) == 2 * len(R.KAPPAS)


def test_calibration_gives_up_when_nothing_agrees():
    n = 1000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    p = pred_rows(np.full(n, -0.1), np.zeros(n), np.zeros(n))  # confidently wrong everywhere
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == float("inf") and answered == 0


def test_calibration_drops_wrong_escape_answers():
    """Escape answers are not scaled by kappa: when confidently-escaping
    predictions are wrong, the rule answers without them."""
    rng = np.random.default_rng(1)
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    pbad = np.zeros(n)
    pbad[:50] = 0.99  # predicted escaping, but they are chaotic and bounded
    p = pred_rows(np.full(n, 0.1), rng.uniform(0, 0.005, n), pbad)
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == 1.0 and not escapes and agree == 1.0 and answered == n - 50


def test_calibration_spends_the_wrong_answer_budget():
    """A few wrong classes are acceptable when the map's agreement stays at the
    target: 4 wrong of 10000 calibration points is within 0.05 %.  The wrong
    ones here are points the method found escaping while the surrogate
    answered a value (judged by class only: an escaped point has no lambda)."""
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    st[:4] = 4
    lam[:4] = np.nan
    p = pred_rows(np.full(n, 0.1), np.full(n, 0.001), np.zeros(n))
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == 1.0 and answered == n and abs(agree - 0.9996) < 1e-12


def test_calibration_keeps_answered_values_within_the_accuracy_gate():
    """Every lambda1 the rule answers on the calibration subsample must be
    within dlam_cal (5e-4) of the method's value.  With errors of 0.3 * std the
    rule must stop at std <= 5e-4 / 0.3 = 1.67e-3, i.e. kappa >= 0.01 / 1.67e-3
    = 6: the smallest multiplier of the grid that keeps the gate is 6, even
    though every answer's class is right at kappa 1."""
    rng = np.random.default_rng(2)
    n = 20000
    std = rng.uniform(0.0, 0.01, n)
    lam1 = rng.uniform(0.05, 0.2, n)
    mu1 = lam1 + 0.3 * std * rng.choice([-1.0, 1.0], n)
    lam = np.stack([lam1, np.zeros(n), np.full(n, -1.0)], 1)
    st = np.zeros(n, dtype=np.uint32)
    kappa, escapes, agree, answered, table = R.calibrate_kappa(
        pred_rows(mu1, std, np.zeros(n)), lam, st, target=0.9995, **THR)
    assert kappa == 6.0 and agree == 1.0
    assert answered == int(np.sum(6.0 * std <= 0.01))
    row = [r for r in table if r[0] == 6.0 and r[1]][0]
    assert row[4] <= 5e-4
    # at kappa 5 (std <= 2e-3) the largest error exceeds 5e-4: rejected
    assert [r for r in table if r[0] == 5.0 and r[1]][0][4] > 5e-4
    # a looser calibration bound admits smaller multipliers
    k2 = R.calibrate_kappa(pred_rows(mu1, std, np.zeros(n)), lam, st, target=0.9995,
                           dlam_cal=1e-3, **THR)[0]
    assert k2 == 3.0


def test_calibration_falls_back_to_escape_answers_only():
    """When no lambda answer meets the accuracy bound, the rule that answers
    only the confident escapes (kappa inf) is taken."""
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    st[:500] = 4
    lam[:500] = np.nan
    pbad = np.zeros(n)
    pbad[:500] = 0.995
    mu1 = np.full(n, 0.1 + 2e-3)  # every value 2e-3 off: over the gate
    p = pred_rows(mu1, np.full(n, 1e-5), pbad)
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == float("inf") and escapes and answered == 500 and agree == 1.0


def test_calibration_needs_evidence_for_valued_answers():
    """A rule whose lambda answers on the calibration subsample are all within
    the bound but number fewer than min_valued is not taken: 20 well-predicted
    values are no evidence for the rest of the plane.  The escapes-only rule
    is taken instead; with min_valued 10 the 20 values are admitted."""
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    st[:500] = 4
    lam[:500] = np.nan
    pbad = np.zeros(n)
    pbad[:500] = 0.995
    std = np.full(n, 0.05)  # unsure everywhere ...
    std[500:520] = 1e-5     # ... but at 20 points, whose values are exact
    p = pred_rows(np.full(n, 0.1), std, pbad)
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == float("inf") and escapes and answered == 500
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995,
                                                          min_valued=10, **THR)
    assert kappa == 1.0 and escapes and answered == 520


def test_calibration_rejects_values_the_calibration_never_saw():
    """A rule with no valued answer on the calibration subsample, but which
    would answer values elsewhere on the plane, has no evidence for them: it
    is not taken (plane_valued counts the plane's valued answers per rule)."""
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    st[:500] = 4
    lam[:500] = np.nan
    pbad = np.zeros(n)
    pbad[:500] = 0.995
    p = pred_rows(np.full(n, 0.1), np.full(n, 0.05), pbad)  # no value answered here
    seen = []

    def plane_valued(kappa, escapes):
        seen.append((kappa, escapes))
        return 0 if kappa >= 4.0 else 50  # the plane has confident values below kappa 4

    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995,
                                                          plane_valued=plane_valued, **THR)
    assert kappa == 4.0 and escapes and answered == 500
    assert (1.0, True) in seen


def test_verify_gates_use_the_surrogate_where_it_answers():
    """The chosen form's answer at a verification point is the surrogate's
    where the rule answers it, the method's elsewhere; |dlambda1| is taken
    where that answer is a value and the FP64 judge's point is OK."""
    flagged = np.array([True, False, False, False, True])
    pred = pred_rows(np.array([0.5, 0.1008, 0.3, -0.2, 0.0]), np.zeros(5),
                     np.array([0.0, 0.0, 0.99, 0.0, 0.0]))
    lam_m = np.array([[0.2, 0, -1], [0.1, 0, -1], [0.3, 0, -1], [-0.2, 0, -1], [0.0, 0, -1]], float)
    lam_m[0, 0] = 0.2 + 1e-6
    st_m = np.zeros(5, dtype=np.uint32)
    lam_j = lam_m.copy()
    lam_j[0, 0] = 0.2
    st_j = np.array([0, 0, 4, 0, 0], dtype=np.uint32)
    lam_j[2] = np.nan
    g = R.verify_gates(flagged, pred, lam_m, st_m, lam_j, st_j)
    # point 1: surrogate 0.1008 vs judge 0.1 -> 8e-4; point 0: method 1e-6 off;
    # point 2: surrogate says escaped, judge escaped -> class only
    assert abs(g["max_dlambda1_chosen_vs_fp64"] - 8e-4) < 1e-8  # pred is binary32
    assert abs(g["max_dlambda1_surrogate_answers"] - 8e-4) < 1e-8
    assert abs(g["max_dlambda1_computed_vs_fp64"] - 1e-6) < 1e-12
    assert g["surrogate_answered_verify"] == 3 and g["surrogate_valued_answers_verify"] == 2
    assert g["class_agreement"] == 1.0 and g["gate_met"]
    # a surrogate value 2e-3 off breaks the gate
    pred[3, 0] = -0.198
    g = R.verify_gates(flagged, pred, lam_m, st_m, lam_j, st_j)
    assert g["chosen_points_over_gate"] == 1 and not g["gate_met"]


def test_classify():
    lam1 = np.array([0.02, 0.005, -0.3, 0.5])
    st = np.array([0, 0, 0, 4])
    assert R.classify(lam1, st).tolist() == [1, 0, 0, 2]


def test_spread_order_is_a_permutation_with_the_kernels_stride():
    """refine.spread_order is the kernels' default item order (position
    i * s mod n, s the first integer >= n/phi coprime to n): a permutation of
    range(n), for sizes with and without a common factor at n/phi."""
    import math

    for n in (1, 2, 3, 10, 64, 1000, 4096, 65536, 3_980_001):
        p = R.spread_order(n, "cpu").numpy()
        assert sorted(p.tolist()) == list(range(n)) if n <= 65536 else len(np.unique(p)) == n
        if n > 2:
            s = int(n * 0.6180339887498949)
            while math.gcd(s, n) != 1:
                s += 1
            assert p[1] == s % n and p[2] == (2 * s) % n
