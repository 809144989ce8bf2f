// Model registry with the reference's public names (model_factory.h:22-69).
// The reference stored per-model fetchers of device function pointers; here a
// model is a compiled-in kernel specialisation named by its LYAP_MODEL_* id,
// so ModelData carries that id.  Models of the library's table register
// themselves at static initialisation (lyap_model_at).
//
// A registrar written against the reference's signature (five fetchers,
// model_factory.h:36-42, e.g. cuda/models/flows/shimizu.dp.cpp:70-211) still
// compiles and registers: the fetchers are kept in ModelData but never called
// (device function pointers have no counterpart here), and the name is bound
// to the compiled model of the same name and dimension.  A name with no
// compiled model registers with _model_id = -1, and TaskLyapunov::init refuses
// it with a message saying that the model's kernel must be compiled in.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <string>

enum class ModelType { flow, map };

// the reference's right-hand-side ABIs (types.h:17-24): f(x, dxdt, p) and
// fvar(v, dvdt, p, x_base); only their types are needed here
typedef void (*diffSysFunc)(double*, double*, double*);
typedef void (*diffSysFuncVar)(double*, double*, double*, double*);

class ModelFactoryRegistry {
protected:
    using ModelFunction = std::function<void(diffSysFunc*)>;
    using ModelFunctionVar = std::function<void(diffSysFuncVar*)>;
    using ModelFunctionQ = std::function<void(diffSysFuncVar*)>;

    // this build's registration: a compiled model id
    static void RegisterModel(const std::string& name, int32_t dim, int32_t param_num, ModelType type,
                              int32_t model_id, const std::string& model_param,
                              const std::string& model_phase, const std::string& model_equat);

    // the reference's registration (model_factory.h:36-42): bound by name to
    // the compiled model of that name and dimension (else _model_id = -1)
    static void RegisterModel(const std::string& name, int32_t dim, int32_t param_num, ModelType type,
                              ModelFunction func, ModelFunctionVar func_var, ModelFunctionVar func_var_rev,
                              ModelFunctionVar func_var_trans_rev, ModelFunctionQ func_Q,
                              const std::string& model_param, const std::string& model_phase,
                              const std::string& model_equat);

public:
    struct ModelData {
        ModelData(const std::string& name, int32_t dim, int32_t param_num, ModelType type, int32_t model_id,
                  const std::string& model_param, const std::string& model_phase,
                  const std::string& model_equat)
            : _dimension(dim), _param_number(param_num), _model_name(name), _type(type),
              _model_id(model_id), _model_param(model_param), _model_phase(model_phase),
              _model_equat(model_equat) {}

        // the reference's constructor (model_factory.h:47-54)
        ModelData(const std::string& name, int32_t dim, int32_t param_num, ModelType type, ModelFunction func,
                  ModelFunctionVar func_var, ModelFunctionVar func_var_rev, ModelFunctionVar func_var_trans_rev,
                  ModelFunctionQ func_Q, const std::string& model_param, const std::string& model_phase,
                  const std::string& model_equat)
            : _dimension(dim), _param_number(param_num), _model_name(name), _type(type), _model_id(-1),
              _model_func(func), _func_var(func_var), _func_var_rev(func_var_rev),
              _func_var_trans_rev(func_var_trans_rev), _func_Q(func_Q), _model_param(model_param),
              _model_phase(model_phase), _model_equat(model_equat) {}

        int32_t _dimension;
        int32_t _param_number;
        std::string _model_name;
        ModelType _type;
        int32_t _model_id;  // LYAP_MODEL_* of include/chaos_lyap.h; -1: no compiled kernel
        // the reference's fetchers, kept for source compatibility (never called)
        ModelFunction _model_func;
        ModelFunctionVar _func_var;
        ModelFunctionVar _func_var_rev;
        ModelFunctionVar _func_var_trans_rev;
        ModelFunctionQ _func_Q;
        std::string _model_param;
        std::string _model_phase;
        std::string _model_equat;
    };

    struct Storage {
        std::map<std::string, ModelData> _data;
    };
};
