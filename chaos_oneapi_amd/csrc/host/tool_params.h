// Sweep configuration read from config.json.  Public names, fields and error
// behaviour follow the reference's tool_params.h:10-71 / tool_params.cpp:
// a missing key or a sweep array whose length is not 4 throws
// std::runtime_error("config error") after printing the reason; a missing or
// unparsable file makes ConfigFactory return nullptr.
//
// Optional keys this build adds (in the system entry, else under
// tool.lyapunov; when absent the reference behaviour is kept):
//   "method": "var" | "var_ref" | "qr_ref" | "fd_ref"   (default from using_qr/using_var)
//   "integrator": "dverk" | "dopri5", "adaptive": bool, "rtol", "atol",
//   "h_min", "h_max", "max_steps", "escape_radius",
//   "precision": "fp64" | "mixed",
//   "seed", "init_perturbation": {"kind": "uniform"|"normal", "scale": s, "add_param": i},
//   "order_probe": thousandths of the integration times of the heavy-first
//   probe of adaptive sweeps (0: off; the lyap_ctx_set_option key of that name)
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "chaos_lyap.h"
#include "json.h"

using json = chaos::Json;

enum ToolID {
    UNDEFINED = -1,
    LYAPUNOV = 0,
    ANGLE = 2,
    KNEADING = 1,
};

class Parameters {
public:
    virtual ~Parameters() = default;
    virtual void Print() const {}
    virtual json paramsToJson() const = 0;
};

class LyapunovParameters : public Parameters {
public:
    LyapunovParameters(const json& jsonConfig, const std::string& modelName);
    void Print() const override;
    json paramsToJson() const override;

    bool is_qr = true;
    bool is_var = true;
    int32_t lyapunov_exponent_num = 0;
    int32_t num_GPUs = 1;
    int32_t samples_num = 1;
    double epsilon = 0.0;
    double integration_step = 0.0;
    double skip_time_attractor = 0.0;
    double skip_time_slave_trajectory = 0.0;
    double skip_time_normalization = 0.0;
    double total_time = 0.0;
    std::vector<double> initial_state;
    std::vector<double> initial_model_params;
    std::vector<double> first_sweep_param;
    std::vector<double> second_sweep_param;

    // this build's optional keys
    int32_t method = LYAP_METHOD_FD_REF;
    int32_t integrator = LYAP_INT_DVERK;
    bool adaptive = false;
    int32_t precision = LYAP_PREC_FP64;
    double rtol = 0.0, atol = 0.0, h_min = 0.0, h_max = 0.0;
    int64_t max_steps = 0;
    double escape_radius = 0.0;
    uint64_t seed = 0;
    int32_t init_kind = LYAP_INIT_NONE;
    double init_scale = 0.0;
    int32_t init_add_param = -1;
    int32_t order_probe = 0;
};

struct ConfigFactory {
    static std::shared_ptr<Parameters> paramsFromConfigFile(ToolID toolId, const std::string& modelName);
    static std::shared_ptr<Parameters> paramsFromJson(ToolID toolId, const std::string& modelName,
                                                      const json& jsonConfig);
};
