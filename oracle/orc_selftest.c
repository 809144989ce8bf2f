/*
 * Self-test driver of the oracle for sanitizer builds (ASan + UBSan) —
 * TEST INFRASTRUCTURE ONLY, built by `make -C oracle selftest-san`.
 *
 * Runs a few points of every model through every method the oracle has
 * (VAR fixed/adaptive DVERK/Dopri5 FP64/mixed, the reference VAR/QR/FD modes,
 * the map branches) and checks each status word is one of the defined ones.
 * Any out-of-bounds access or undefined behaviour aborts the run.
 */
#include <stdio.h>
#include <string.h>

#include "lyap_oracle.h"

static int run(orc_desc* d, int64_t count) {
    double lam[8 * ORC_MAX_DIM];
    uint32_t st[8], na[8], nr[8];
    orc_sweep(d, 0, count, NULL, NULL, NULL, lam, st, na, nr, 2);
    for (int64_t i = 0; i < count; ++i)
        if (st[i] > ORC_ESCAPED) return 1;
    return 0;
}

static void base(orc_desc* d, int model, int dim, int np) {
    memset(d, 0, sizeof(*d));
    d->model = model;
    d->dim = dim;
    d->n_exps = dim;
    d->n_params = np;
    d->h = 0.01;
    d->rtol = d->atol = 1e-8;
    d->h_min = 1e-12;
    d->h_max = 0.5;
    d->eps = 1e-8;
    d->t_transient = 1.0;
    d->t_average = 4.0;
    d->t_renorm = 0.5;
    d->max_steps = 100000;
    d->escape_radius = 1e6;
    d->sweep_idx1 = 0;
    d->sweep_idx2 = np > 1 ? 1 : 0;
    d->n1 = 2;
    d->n2 = 2;
    d->init_kind = ORC_INIT_UNIFORM;
    d->init_param_offset = -1;
    d->init_scale = 0.1;
    d->seed = 7;
    for (int i = 0; i < dim; ++i) d->init_state[i] = 0.5;
}

int main(void) {
    int bad = 0;
    orc_desc d;
    const int flows[4][3] = {{ORC_MODEL_SHIMIZU, 3, 2}, {ORC_MODEL_LORENZ, 3, 3},
                             {ORC_MODEL_ROSSLER, 3, 3}, {ORC_MODEL_L96, 8, 2}};
    const double p0[4][3] = {{0.75, 0.45, 0}, {10, 28, 8.0 / 3.0}, {0.2, 5.7, 0.2}, {8, 1, 0}};
    for (int f = 0; f < 4; ++f) {
        for (int integ = 0; integ < 2; ++integ)
            for (int adaptive = 0; adaptive < 2; ++adaptive)
                for (int prec = 0; prec < 2; ++prec) {
                    base(&d, flows[f][0], flows[f][1], flows[f][2]);
                    for (int j = 0; j < 3; ++j) d.base_params[j] = p0[f][j];
                    d.start1 = p0[f][0];
                    d.delta1 = 0.01;
                    d.start2 = p0[f][1];
                    d.delta2 = 0.01;
                    d.method = ORC_METHOD_VAR;
                    d.integrator = integ;
                    d.adaptive = adaptive;
                    d.precision = prec;
                    d.n_renorm_transient = 2;
                    d.n_renorm_average = 8;
                    d.steps_per_renorm = 50;
                    bad |= run(&d, 4);
                }
        if (flows[f][0] == ORC_MODEL_L96) continue;
        for (int method = ORC_METHOD_VAR_REF; method <= ORC_METHOD_FD_REF; ++method) {
            base(&d, flows[f][0], flows[f][1], flows[f][2]);
            for (int j = 0; j < 3; ++j) d.base_params[j] = p0[f][j];
            d.start1 = p0[f][0];
            d.delta1 = 0.01;
            d.start2 = p0[f][1];
            d.delta2 = 0.01;
            d.method = method;
            bad |= run(&d, 4);
        }
    }
    const int maps[2][3] = {{ORC_MODEL_HENON, 2, 2}, {ORC_MODEL_LOGISTIC, 1, 1}};
    const double mp[2][2] = {{1.4, 0.3}, {3.9, 0}};
    for (int mi = 0; mi < 2; ++mi)
        for (int method = ORC_METHOD_VAR; method <= ORC_METHOD_FD_REF; ++method) {
            if (method == ORC_METHOD_QR_REF) continue;
            for (int prec = 0; prec < (method == ORC_METHOD_VAR ? 2 : 1); ++prec) {
                base(&d, maps[mi][0], maps[mi][1], maps[mi][2]);
                d.h = 1.0;
                d.t_transient = 10;
                d.t_average = 100;
                d.t_renorm = 2;
                d.base_params[0] = mp[mi][0];
                d.base_params[1] = mp[mi][1];
                d.start1 = mp[mi][0];
                d.delta1 = 0.001;
                d.start2 = mp[mi][1];
                d.delta2 = 0.001;
                d.init_scale = 0.01;
                d.init_state[0] = 0.1;
                d.method = method;
                d.precision = prec;
                bad |= run(&d, 4);
            }
        }
    printf(bad ? "FAILED\n" : "OK\n");
    return bad;
}
