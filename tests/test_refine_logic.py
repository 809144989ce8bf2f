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
    """A few wrong answers are acceptable when the map's agreement stays at the
    target: 4 wrong of 10000 calibration points is within 0.05 %."""
    n = 10000
    lam = np.tile([0.1, 0.0, -1.0], (n, 1))
    st = np.zeros(n, dtype=np.uint32)
    mu1 = np.full(n, 0.1)
    mu1[:4] = -0.1  # confidently wrong
    p = pred_rows(mu1, np.full(n, 0.001), np.zeros(n))
    kappa, escapes, agree, answered, _ = R.calibrate_kappa(p, lam, st, target=0.9995, **THR)
    assert kappa == 1.0 and answered == n and abs(agree - 0.9996) < 1e-12


def test_classify():
    lam1 = np.array([0.02, 0.005, -0.3, 0.5])
    st = np.array([0, 0, 0, 4])
    assert R.classify(lam1, st).tolist() == [1, 0, 0, 2]
