#include "tool_params.h"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace {

const json* find_key(const json& system, const json* tool, const char* key) {
    if (system.contains(key)) return &system.at(key);
    if (tool && tool->contains(key)) return &tool->at(key);
    return nullptr;
}

int32_t method_from_name(const std::string& s) {
    if (s == "var") return LYAP_METHOD_VAR;
    if (s == "var_ref") return LYAP_METHOD_VAR_REF;
    if (s == "qr_ref") return LYAP_METHOD_QR_REF;
    if (s == "fd_ref") return LYAP_METHOD_FD_REF;
    if (s == "var_trace") return LYAP_METHOD_VAR_TRACE;
    throw std::invalid_argument("unknown method '" + s + "'");
}

const char* method_name(int32_t m) {
    switch (m) {
        case LYAP_METHOD_VAR: return "var";
        case LYAP_METHOD_VAR_REF: return "var_ref";
        case LYAP_METHOD_QR_REF: return "qr_ref";
        case LYAP_METHOD_VAR_TRACE: return "var_trace";
        default: return "fd_ref";
    }
}

void print_vec(const char* title, const std::vector<double>& v) {
    std::cout << title;
    int count = 0;
    for (double e : v) std::cout << " [" << count++ << "] = " << e << " ";
    std::cout << std::endl;
}

}  // namespace

// tool_params.cpp:5-51 of the reference
void LyapunovParameters::Print() const {
    std::cout << "************************************************************************" << std::endl;
    std::cout << "Lyapunov's parameters: " << std::endl;
    std::cout << "Number of lyapunov's exponent: " << lyapunov_exponent_num << std::endl;
    std::cout << "Epsilon: " << epsilon << std::endl;
    std::cout << "Step of integration: " << integration_step << std::endl;
    std::cout << "Time skip to reach attractor: " << skip_time_attractor << std::endl;
    std::cout << "Time skip for complementary trajectories: " << skip_time_slave_trajectory << std::endl;
    std::cout << "Time skip for Lyapunov's normalization: " << skip_time_normalization << std::endl;
    std::cout << "Total time: " << total_time << std::endl;
    std::cout << "Samples num: " << samples_num << std::endl;
    print_vec("Initial state: ", initial_state);
    print_vec("Initial parameters of model: ", initial_model_params);
    print_vec("First sweep parameter: ", first_sweep_param);
    print_vec("Second sweep parameter: ", second_sweep_param);
    std::cout << "Number of GPUs: " << num_GPUs << std::endl;
    std::cout << "Using QR: " << std::boolalpha << is_qr << std::endl;
    std::cout << "Using VAR: " << std::boolalpha << is_var << std::endl;
    std::cout << "Method: " << method_name(method)
              << ((method == LYAP_METHOD_VAR || method == LYAP_METHOD_VAR_TRACE)
                      ? (std::string(" (") + (integrator == LYAP_INT_DVERK ? "dverk" : "dopri5") +
                         (adaptive ? ", adaptive" : ", fixed step") +
                         (precision == LYAP_PREC_MIXED ? ", mixed)" : ", fp64)"))
                      : std::string())
              << std::endl;
    std::cout << "************************************************************************" << std::endl
              << std::endl;
}

// tool_params.cpp:53-75 of the reference
json LyapunovParameters::paramsToJson() const {
    json params = json::object();
    params["lyapunov_exponent_num"] = lyapunov_exponent_num;
    params["epsilon"] = epsilon;
    params["integration_step"] = integration_step;
    params["skip_time_attractor"] = skip_time_attractor;
    params["skip_time_slave_trajectory"] = skip_time_slave_trajectory;
    params["skip_time_normalization"] = skip_time_normalization;
    params["total_time"] = total_time;
    params["samples_num"] = samples_num;
    params["initial_state"] = json(initial_state);
    params["initial_model_params"] = json(initial_model_params);
    params["first_sweep_param"] = json(first_sweep_param);
    params["second_sweep_param"] = json(second_sweep_param);
    params["num_GPUs"] = num_GPUs;
    params["using_qr"] = is_qr;
    params["using_var"] = is_var;
    return params;
}

// tool_params.cpp:77-122 of the reference, plus the optional keys
LyapunovParameters::LyapunovParameters(const json& jsonConfig, const std::string& modelName) {
    bool isParamsInFile = false;
    try {
        jsonConfig.at("hardware").at("num_GPUs").get_to(num_GPUs);
        const json& tool = jsonConfig.at("tool").at("lyapunov");
        tool.at("samples_num").get_to(samples_num);
        tool.at("using_qr").get_to(is_qr);
        tool.at("using_var").get_to(is_var);
        const json& systems = jsonConfig.at("systems");
        for (size_t i = 0; i < systems.size(); ++i) {
            const json& c = systems.at(i);
            if (c.at("model_name") != modelName) continue;
            isParamsInFile = true;
            c.at("lyapunov_exponent_num").get_to(lyapunov_exponent_num);
            c.at("epsilon").get_to(epsilon);
            c.at("integration_step").get_to(integration_step);
            c.at("skip_time_attractor").get_to(skip_time_attractor);
            c.at("skip_time_slave_trajectory").get_to(skip_time_slave_trajectory);
            c.at("skip_time_normalization").get_to(skip_time_normalization);
            c.at("total_time").get_to(total_time);
            c.at("initial_state").get_to(initial_state);
            c.at("initial_model_params").get_to(initial_model_params);
            c.at("first_sweep_param").get_to(first_sweep_param);
            c.at("second_sweep_param").get_to(second_sweep_param);

            // reference mode selection (cuda/lyapunov.dp.cpp:274-317)
            method = is_qr ? LYAP_METHOD_QR_REF : (is_var ? LYAP_METHOD_VAR_REF : LYAP_METHOD_FD_REF);
            if (const json* v = find_key(c, &tool, "method")) method = method_from_name(v->as_string());
            if (const json* v = find_key(c, &tool, "integrator")) {
                const std::string& s = v->as_string();
                if (s == "dverk") integrator = LYAP_INT_DVERK;
                else if (s == "dopri5") integrator = LYAP_INT_DOPRI5;
                else throw std::invalid_argument("unknown integrator '" + s + "'");
            }
            if (const json* v = find_key(c, &tool, "adaptive")) adaptive = v->as_bool();
            if (const json* v = find_key(c, &tool, "precision")) {
                const std::string& s = v->as_string();
                if (s == "fp64") precision = LYAP_PREC_FP64;
                else if (s == "mixed") precision = LYAP_PREC_MIXED;
                else throw std::invalid_argument("unknown precision '" + s + "'");
            }
            if (const json* v = find_key(c, &tool, "rtol")) rtol = v->as_double();
            if (const json* v = find_key(c, &tool, "atol")) atol = v->as_double();
            if (const json* v = find_key(c, &tool, "h_min")) h_min = v->as_double();
            if (const json* v = find_key(c, &tool, "h_max")) h_max = v->as_double();
            if (const json* v = find_key(c, &tool, "max_steps")) max_steps = v->as_int();
            if (const json* v = find_key(c, &tool, "escape_radius")) escape_radius = v->as_double();
            if (const json* v = find_key(c, &tool, "seed")) seed = (uint64_t)v->as_int();
            if (const json* v = find_key(c, &tool, "order_probe")) order_probe = (int32_t)v->as_int();
            if (const json* v = find_key(c, &tool, "init_perturbation")) {
                const std::string& kind = v->at("kind").as_string();
                if (kind == "uniform") init_kind = LYAP_INIT_UNIFORM;
                else if (kind == "normal") init_kind = LYAP_INIT_NORMAL;
                else if (kind == "none") init_kind = LYAP_INIT_NONE;
                else throw std::invalid_argument("unknown init_perturbation kind '" + kind + "'");
                init_scale = v->contains("scale") ? v->at("scale").as_double() : 0.0;
                init_add_param = v->contains("add_param") ? (int32_t)v->at("add_param").as_int() : -1;
            }
            break;
        }
    } catch (std::exception& e) {
        printf("error: model: %s config exception: %s\n", modelName.c_str(), e.what());
        throw std::runtime_error("config error");
    }
    if (!isParamsInFile) {
        printf("error: model: %s config not found\n", modelName.c_str());
        throw std::runtime_error("config error");
    }
    if (first_sweep_param.size() != 4 || second_sweep_param.size() != 4) {
        printf("error: model: %s config sweep parameters size error\n", modelName.c_str());
        throw std::runtime_error("config error");
    }
}

// tool_params.cpp:238-286 of the reference (the kneading tool is not built:
// its ids stay in ToolID and yield nullptr)
std::shared_ptr<Parameters> ConfigFactory::paramsFromConfigFile(ToolID toolId,
                                                                const std::string& modelName) {
    std::string configPath;
    switch (toolId) {
        case ToolID::LYAPUNOV:
        case ToolID::ANGLE:
        case ToolID::KNEADING:
            configPath = "config.json";
            break;
        default:
            return nullptr;
    }
    std::ifstream config(configPath);
    if (!config.good()) {
        printf("error: config file: %s not found\n", configPath.c_str());
        return nullptr;
    }
    json jsonConfig;
    try {
        std::stringstream ss;
        ss << config.rdbuf();
        jsonConfig = json::parse(ss.str());
    } catch (...) {
        printf("error: config file: %s parse failure\n", configPath.c_str());
        return nullptr;
    }
    return paramsFromJson(toolId, modelName, jsonConfig);
}

std::shared_ptr<Parameters> ConfigFactory::paramsFromJson(ToolID toolId, const std::string& modelName,
                                                          const json& jsonConfig) {
    switch (toolId) {
        case ToolID::LYAPUNOV:
        case ToolID::ANGLE:
            return std::make_shared<LyapunovParameters>(jsonConfig, modelName);
        default:
            return nullptr;
    }
}
