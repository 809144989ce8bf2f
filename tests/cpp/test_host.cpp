// CPU unit tests of the C++ host layer (no GPU): JSON reader/writer,
// LyapunovParameters parsing (reference tool_params.cpp:77-122 rules plus the
// optional keys), TaskLyapunov::init's descriptor, saveCalculatedData's format.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "filehandler.h"
#include "singleton.h"
#include "tasklyapunov.h"

static int g_fail = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            ++g_fail;                                                     \
        }                                                                 \
    } while (0)

static const char* kConfig = R"({
  "hardware": {"num_GPUs": 1},
  "tool": {"lyapunov": {"samples_num": 1, "using_qr": false, "using_var": true}},
  "systems": [
    {"model_name": "lorenz", "lyapunov_exponent_num": 3, "epsilon": 1e-7,
     "integration_step": 1e-2, "skip_time_attractor": 10, "skip_time_slave_trajectory": 5,
     "skip_time_normalization": 0.5, "total_time": 100, "initial_state": [1, 1, 1],
     "initial_model_params": [10, 0, 0], "first_sweep_param": [1, 0, 200, 3],
     "second_sweep_param": [2, 0.5, 4, 4],
     "method": "var", "integrator": "dopri5", "adaptive": true, "rtol": 1e-9, "atol": 1e-9,
     "precision": "mixed", "seed": 42, "order_probe": 10,
     "init_perturbation": {"kind": "uniform", "scale": 0.001}},
    {"model_name": "shimizu", "lyapunov_exponent_num": 3, "epsilon": 1e-7,
     "integration_step": 1e-2, "skip_time_attractor": 1e+3, "skip_time_slave_trajectory": 1e+3,
     "skip_time_normalization": 1e-2, "total_time": 1e+4, "skip_time_normalization": 0.02,
     "initial_state": [1e-4, 0, 0], "initial_model_params": [0, 0],
     "first_sweep_param": [0, 0.2, 2, 200], "second_sweep_param": [1, 0, 1.2, 200]},
    {"model_name": "rossler", "lyapunov_exponent_num": 3, "epsilon": 1e-7,
     "integration_step": 1e-2, "skip_time_attractor": 1, "skip_time_slave_trajectory": 1,
     "skip_time_normalization": 0.01, "total_time": 1, "initial_state": [0, 0, 0],
     "initial_model_params": [0, 0, 0.3], "first_sweep_param": [0, 0.4, 0.6],
     "second_sweep_param": [1, 3, 8, 200]}
  ]
})";

static void test_json() {
    json j = json::parse(R"({"b": [1, 2.5, -3e-7, true, null, "x\n\"y"], "a": {"k": 0.01}})");
    CHECK(j.at("a").at("k").as_double() == 0.01);
    CHECK(j.at("b").size() == 6);
    CHECK(j.at("b").at(0).as_int() == 1);
    CHECK(j.at("b").at(5).as_string() == "x\n\"y");
    bool threw = false;
    try {
        j.at("missing");
    } catch (const std::out_of_range&) {
        threw = true;
    }
    CHECK(threw);
    // nlohmann-style output: sorted keys, shortest round-trip doubles, ".0" on integral doubles
    json o = json::object();
    o["z"] = 1e-07;
    o["a"] = 1000.0;
    o["m"] = 2.666666666667;
    o["i"] = 3;
    o["v"] = json(std::vector<double>{0.01, 0.1, 1.0 / 3.0});
    const std::string s = o.dump(4);
    CHECK(s == "{\n    \"a\": 1000.0,\n    \"i\": 3,\n    \"m\": 2.666666666667,\n    \"v\": [\n"
               "        0.01,\n        0.1,\n        0.3333333333333333\n    ],\n    \"z\": 1e-07\n}");
    CHECK(json::parse(s).at("v").at(2).as_double() == 1.0 / 3.0);
    CHECK(chaos::format_double(100.0) == "100.0");
    CHECK(chaos::format_double(0.30000000000000004) == "0.30000000000000004");
    CHECK(chaos::format_double(1.5e20) == "1.5e+20" && chaos::format_double(-0.0001) == "-0.0001");
    CHECK(chaos::format_double(1e-05) == "1e-05" && chaos::format_double(123456.5) == "123456.5");
    threw = false;
    try {
        json::parse("{\"a\": [1, 2}");
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
}

static void test_params() {
    json cfg = json::parse(kConfig);
    LyapunovParameters p(cfg, "lorenz");
    CHECK(p.num_GPUs == 1 && p.samples_num == 1 && !p.is_qr && p.is_var);
    CHECK(p.lyapunov_exponent_num == 3 && p.total_time == 100.0);
    CHECK(p.method == LYAP_METHOD_VAR && p.integrator == LYAP_INT_DOPRI5 && p.adaptive);
    CHECK(p.precision == LYAP_PREC_MIXED && p.seed == 42 && p.init_kind == LYAP_INIT_UNIFORM);
    CHECK(p.order_probe == 10);
    // reference defaults: using_var -> VAR_REF; a repeated key keeps the last value
    LyapunovParameters s(cfg, "shimizu");
    CHECK(s.method == LYAP_METHOD_VAR_REF && s.skip_time_normalization == 0.02);
    CHECK(s.order_probe == 0);
    // tool_params.cpp:108-121: missing model and a 3-element sweep array throw
    int thrown = 0;
    try {
        LyapunovParameters x(cfg, "lorenz4d");
    } catch (const std::runtime_error& e) {
        thrown += std::string(e.what()) == "config error";
    }
    try {
        LyapunovParameters x(cfg, "rossler");
    } catch (const std::runtime_error& e) {
        thrown += std::string(e.what()) == "config error";
    }
    CHECK(thrown == 2);
    // paramsToJson keeps the reference's keys
    json pj = s.paramsToJson();
    CHECK(pj.at("using_var").as_bool() && !pj.at("using_qr").as_bool());
    CHECK(pj.at("first_sweep_param").size() == 4);
    CHECK(ConfigFactory::paramsFromJson(ToolID::KNEADING, "shimizu", cfg) == nullptr);
}

static void test_task_init_and_output() {
    json cfg = json::parse(kConfig);
    auto params = ConfigFactory::paramsFromJson(ToolID::LYAPUNOV, "lorenz", cfg);
    TaskLyapunov task;
    CHECK(task.init(params, "lorenz") == 0);
    const lyap_sweep_desc& d = task.descriptor();
    CHECK(d.model == LYAP_MODEL_LORENZ && d.n1 == 3 && d.n2 == 4 && d.sweep_idx1 == 1 && d.sweep_idx2 == 2);
    CHECK(d.delta1 == 100.0 && d.start2 == 0.5 && std::fabs(d.delta2 - 3.5 / 3.0) == 0.0);
    CHECK(d.base_params[0] == 10.0 && d.init_kind == LYAP_INIT_UNIFORM && d.init_scale == 0.001);
    CHECK(d.n_renorm_transient == 30 && d.n_renorm_average == 200);

    // saveCalculatedData: rows printed with std::to_string, Model block, .npy sidecar
    struct FakeTask : public TaskLyapunov {
        void fill() {
            mResults = {{0.0, 0.5, 0.123456789, -1.0, std::nan("")}, {100.0, 4.0, 1e-9, 2.0, 3.0}};
            mReady = true;
        }
    } ft;
    ft.fill();
    CHECK(saveCalculatedData(ToolID::LYAPUNOV, "lorenz", params.get(), &ft) == 0);
    std::ifstream f("Data/Lyapunov lorenz.json");
    std::stringstream ss;
    ss << f.rdbuf();
    json out = json::parse(ss.str());
    CHECK(out.at("ID_Tool").as_string() == "Lyapunov");
    CHECK(out.at("_Data").as_string() ==
          "0.000000 0.500000 0.123457 -1.000000 nan\n100.000000 4.000000 0.000000 2.000000 3.000000\n");
    CHECK(out.at("Model").at("Name").as_string() == "lorenz");
    CHECK(out.at("Model").at("Dimension").as_int() == 3);
    CHECK(out.at("Model").at("systemType").as_int() == 0);
    CHECK(out.at("Parameters").at("total_time").as_double() == 100.0);
    std::ifstream npy("Data/Lyapunov lorenz.npy", std::ios::binary);
    std::string raw((std::istreambuf_iterator<char>(npy)), std::istreambuf_iterator<char>());
    CHECK(raw.size() == 128 + 10 * 8);
    CHECK(raw.compare(0, 6, "\x93NUMPY") == 0);
    double v;
    std::memcpy(&v, raw.data() + 128 + 2 * 8, 8);
    CHECK(v == 0.123456789);
}

// A registrar in the reference's shape (cuda/models/flows/shimizu.dp.cpp:70-211):
// five device-function-pointer fetchers and the model strings.  "shimizu" binds
// to the compiled model of that name; "lorenz4d" (named in the reference's
// config.json, with no code there or here) registers without a kernel.
static int g_fetches = 0;
static struct ReferenceStyleRegistry : public ModelFactoryRegistry {
    ReferenceStyleRegistry() {
        auto func = [](diffSysFunc* p) { ++g_fetches; *p = nullptr; };
        auto func_var = [](diffSysFuncVar* p) { ++g_fetches; *p = nullptr; };
        ModelType type = ModelType::flow;
        RegisterModel("shimizu", 3, 2, type, func, func_var, func_var, func_var, func_var, "alpha, lamda",
                      "x, y, z", "x' = y;y' = - alpha * y - x * z + x;z' = - lamda * z + x ** 2.0");
        RegisterModel("lorenz4d", 4, 8, type, func, func_var, func_var, func_var, func_var, "a, b", "x, y, z, w",
                      "(not compiled)");
    }
} g_reference_style;

static void test_reference_style_registration() {
    auto& st = Singleton<ModelFactoryRegistry::Storage>::Instance();
    const auto& sh = st._data.at("shimizu");
    CHECK(sh._model_id == LYAP_MODEL_SHIMIZU && sh._dimension == 3 && sh._param_number == 2);
    CHECK(sh._model_param == "alpha, lamda" && sh._type == ModelType::flow);
    CHECK(sh._model_func && sh._func_Q);  // kept, never called
    CHECK(st._data.at("lorenz4d")._model_id == -1);
    CHECK(g_fetches == 0);
    // a config of the reference's shimizu entry runs through the bound model;
    // the uncompiled one is refused by init
    json cfg = json::parse(kConfig);
    auto params = ConfigFactory::paramsFromJson(ToolID::LYAPUNOV, "shimizu", cfg);
    TaskLyapunov ok;
    CHECK(ok.init(params, "shimizu") == 0 && ok.descriptor().model == LYAP_MODEL_SHIMIZU);
    TaskLyapunov bad;
    CHECK(bad.init(params, "lorenz4d") == -1);
    CHECK(bad.init(params, "no_such_model") == -1);
}

int main() {
    test_json();
    test_params();
    test_task_init_and_output();
    test_reference_style_registration();
    // models registered at start-up, sorted as the reference's std::map
    auto& st = Singleton<ModelFactoryRegistry::Storage>::Instance();
    CHECK(st._data.size() == 9 && st._data.begin()->first == "henon");
    CHECK(st._data.at("shimizu")._param_number == 2);
    std::printf(g_fail ? "FAILED %d\n" : "OK\n", g_fail);
    return g_fail ? 1 : 0;
}
