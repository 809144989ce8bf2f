"""Host logic of bench.py and the flop model behind its roofline: the
precision split and weighted peak, the roofline record, the per-exponent
error-map statistics (no GPU)."""
import numpy as np

import bench
from chaos_oneapi_amd import flops as F
from chaos_oneapi_amd import lyap as L
from chaos_oneapi_amd import workloads as W


def test_precision_split_of_a_rossler_dverk_step():
    acc, rej = F.per_step(L.MODEL_ROSSLER, 3, 3, L.INT_DVERK, True, True)
    f32 = F.per_step_split(L.MODEL_ROSSLER, 3, 3, L.INT_DVERK, True, True)
    # the tangent part is binary32 but the last add of each of the 3 vectors'
    # increments (3 components each) to the binary64 deviation vector
    assert f32 == (acc - rej) - 9
    assert F.per_step_split(L.MODEL_ROSSLER, 3, 3, L.INT_DVERK, True, False) == 0
    desc = L.finalize(L.make_desc(W.rossler_dverk(4, precision=L.PREC_MIXED)))
    na, nr = np.full(16, 1000), np.full(16, 10)
    f64, f32s = F.sweep_flops_split(desc, na, nr)
    assert abs((f64 + f32s) - F.sweep_flops(desc, na, nr)) < 1e-6
    assert f32s == f32 * 16000


def test_l96_mixed_step_counts_the_folded_tangent():
    # Dopri5 tangent: 6 stages, a-row nonzeros 1..5 (15), 5 nonzero b.  Folding h
    # into the Jacobian data saves, per vector and component, the fma(h, acc, v)
    # of the 5 stages with a stage sum and the multiply by h of the increment
    # (6 flops), and costs 2 binary64 products per component and stage plus h*d.
    n = k = 16
    acc64, rej64 = F.per_step(L.MODEL_L96, n, k, L.INT_DOPRI5, True)
    acc32, rej32 = F.per_step(L.MODEL_L96, n, k, L.INT_DOPRI5, True, True)
    assert rej32 == rej64
    assert acc64 - acc32 == 6 * k * n - (2 * n * 6 + 1)
    f32 = F.per_step_split(L.MODEL_L96, n, k, L.INT_DOPRI5, True, True)
    # binary32: 6 Jacobian products (7n) + stage chains (2*15n) + b sum (9n) per vector
    assert f32 == k * (6 * 7 * n + 30 * n + 9 * n)
    # the 3-D flows' mixed form counts as the FP64 form (only the precision
    # moves) where its tangent keeps the solution's weights
    for integ, adaptive in ((L.INT_DOPRI5, True), (L.INT_DVERK, False)):
        assert F.per_step(L.MODEL_ROSSLER, 3, 3, integ, adaptive, True) == \
            F.per_step(L.MODEL_ROSSLER, 3, 3, integ, adaptive)


def test_low_order_tangent_of_adaptive_dverk_counts_six_stages():
    # sixth-order tangent: stages 1-5, 7, 8 (a-row nonzeros 0+1+2+3+4+5+6 = 21),
    # 6 nonzero b; fifth-order: stages 1-6 (0+1+2+3+4+5 = 15), 5 nonzero b - e.
    # Per vector: Rossler J v costs 7 a stage; stage sums 2*nnz + 1 per component
    # with a nonzero row; the increment 2*nnz + 1.
    n = 3
    full = 7 * 7 + n * sum(2 * z + 1 for z in (1, 2, 3, 4, 5, 6)) + n * (2 * 6 + 1)
    low = 6 * 7 + n * sum(2 * z + 1 for z in (1, 2, 3, 4, 5)) + n * (2 * 5 + 1)
    acc64, rej64 = F.per_step(L.MODEL_ROSSLER, n, 2, L.INT_DVERK, True)
    acc32, rej32 = F.per_step(L.MODEL_ROSSLER, n, 2, L.INT_DVERK, True, True)
    assert rej32 == rej64
    assert acc64 - rej64 == 2 * full and acc32 - rej32 == 2 * low


def test_trace_closure_step_counts_two_vectors_and_the_trace():
    desc = L.finalize(L.make_desc(dict(W.rossler_dverk(4, precision=L.PREC_MIXED),
                                       method=L.METHOD_VAR_TRACE)))
    acc, rej, f32, k = F.desc_step(desc)
    acc3, rej3 = F.per_step(L.MODEL_ROSSLER, 3, 3, L.INT_DVERK, True, True)
    assert k == 2
    # DVERK: b has 6 nonzeros -> the quadrature 2*6 - 1, plus c, h and the sum
    assert rej == rej3 + 14
    tangent = (acc3 - rej3) // 3
    assert acc == rej + 2 * tangent
    assert f32 == 2 * (tangent - 3)
    na, nr = np.full(4, 100), np.full(4, 2)
    renorms = desc.n_renorm_transient + desc.n_renorm_average
    assert F.sweep_flops(desc, na, nr) == acc * 400 + rej * 8 + F.mgs(3, 2) * renorms * 4


def test_weighted_peak_is_the_harmonic_mix():
    assert F.weighted_peak(1.0, 0.0) == F.PEAK_FP64_TFLOPS
    assert abs(F.weighted_peak(0.0, 1.0) - F.PEAK_FP32_TFLOPS) < 1e-12
    # equal flops: 2 / (1/78.6 + 1/157.3)
    assert abs(F.weighted_peak(1.0, 1.0) - 2.0 / (1 / 78.6 + 1 / 157.3)) < 1e-9


def test_roofline_record():
    probe = {"fp64_fma_tflops": 60.0, "fp32_unpacked_fma_tflops": 60.0}
    r = bench.roofline(3e13, 6e13, 1000.0, "k", probe, traffic=123)
    assert abs(r["achieved"] - 90.0) < 1e-9 and r["unit"] == "TFLOP/s" and r["traffic"] == 123
    assert abs(r["frac"] - 90.0 / F.weighted_peak(3e13, 6e13)) < 1e-12
    assert abs(r["measured_issue_peak"] - 60.0) < 1e-9 and abs(r["frac_of_measured_issue_peak"] - 1.5) < 1e-12


def test_error_map_statistics():
    n1, n2 = 8, 8
    a = np.zeros((64, 2))
    b = np.zeros((64, 2))
    b[5, 0] = 2e-3   # p1 index 0, p2 index 5: block (0, 2) of a 4 x 4 block map
    b[63, 1] = 5e-4  # block (3, 3)
    ok = np.ones(64, bool)
    ok[10] = False
    e = bench.error_map_stats(a, b, ok, n1, n2, blocks=4)
    assert e["max"] == [2e-3, 5e-4] and e["over_1e-3"] == [1, 0]
    assert e["block_max"][0][0][2] == 2e-3 and e["block_max"][1][3][3] == 5e-4
    assert e["block_max"][0][1][1] == 0.0
