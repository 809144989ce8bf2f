#include "model_factory.h"

#include "chaos_lyap.h"
#include "singleton.h"

void ModelFactoryRegistry::RegisterModel(const std::string& name, int32_t dim, int32_t param_num,
                                         ModelType type, int32_t model_id, const std::string& model_param,
                                         const std::string& model_phase, const std::string& model_equat) {
    Storage& storage = Singleton<Storage>::Instance();
    ModelData data(name, dim, param_num, type, model_id, model_param, model_phase, model_equat);
    // the first registration of a name wins (a reference-shaped registrar that
    // ran earlier in static initialisation keeps its entry)
    storage._data.insert(std::make_pair(name, std::move(data)));
}

void ModelFactoryRegistry::RegisterModel(const std::string& name, int32_t dim, int32_t param_num,
                                         ModelType type, ModelFunction func, ModelFunctionVar func_var,
                                         ModelFunctionVar func_var_rev, ModelFunctionVar func_var_trans_rev,
                                         ModelFunctionQ func_Q, const std::string& model_param,
                                         const std::string& model_phase, const std::string& model_equat) {
    ModelData data(name, dim, param_num, type, func, func_var, func_var_rev, func_var_trans_rev, func_Q,
                   model_param, model_phase, model_equat);
    lyap_model_info info;
    if (lyap_model_lookup(name.c_str(), &info) == LYAP_SUCCESS && info.dim == dim &&
        info.n_params == param_num &&
        (info.type == LYAP_TYPE_MAP) == (type == ModelType::map))
        data._model_id = info.model;
    Storage& storage = Singleton<Storage>::Instance();
    // a registrar of the reference's shape replaces the table's entry of the
    // same name (its strings are the caller's); the kernel stays the table's
    storage._data.erase(name);
    storage._data.insert(std::make_pair(name, std::move(data)));
}

namespace {

// registers every model of the library's table before main(), like the
// reference's per-model static registrars (cuda/models/flows/shimizu.dp.cpp:70-211)
struct CompiledModelsRegistry : public ModelFactoryRegistry {
    CompiledModelsRegistry() {
        const int n = lyap_model_count();
        for (int i = 0; i < n; ++i) {
            lyap_model_info info;
            if (lyap_model_at(i, &info) != LYAP_SUCCESS) continue;
            RegisterModel(info.name, info.dim, info.n_params,
                          info.type == LYAP_TYPE_MAP ? ModelType::map : ModelType::flow, info.model,
                          info.param_names, info.phase_names, info.equations);
        }
    }
};

const CompiledModelsRegistry g_registry;

}  // namespace
