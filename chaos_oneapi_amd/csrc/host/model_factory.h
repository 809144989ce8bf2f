// Model registry with the reference's public names (model_factory.h:22-69).
// The reference stored per-model fetchers of device function pointers; here a
// model is a compiled-in kernel specialisation named by its LYAP_MODEL_* id,
// so ModelData carries that id instead.  Models register themselves at static
// initialisation from the library's model table (lyap_model_at).
#pragma once

#include <cstdint>
#include <map>
#include <string>

enum class ModelType { flow, map };

class ModelFactoryRegistry {
protected:
    static void RegisterModel(const std::string& name, int32_t dim, int32_t param_num, ModelType type,
                              int32_t model_id, const std::string& model_param,
                              const std::string& model_phase, const std::string& model_equat);

public:
    struct ModelData {
        ModelData(const std::string& name, int32_t dim, int32_t param_num, ModelType type, int32_t model_id,
                  const std::string& model_param, const std::string& model_phase,
                  const std::string& model_equat)
            : _dimension(dim), _param_number(param_num), _model_name(name), _type(type),
              _model_id(model_id), _model_param(model_param), _model_phase(model_phase),
              _model_equat(model_equat) {}

        int32_t _dimension;
        int32_t _param_number;
        std::string _model_name;
        ModelType _type;
        int32_t _model_id;  // LYAP_MODEL_* of include/chaos_lyap.h
        std::string _model_param;
        std::string _model_phase;
        std::string _model_equat;
    };

    struct Storage {
        std::map<std::string, ModelData> _data;
    };
};
