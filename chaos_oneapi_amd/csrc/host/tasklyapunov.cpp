#include "tasklyapunov.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "raiitimer.h"
#include "singleton.h"

namespace {

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    printf("error: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

TaskLyapunov::~TaskLyapunov() { clear(); }

std::vector<std::vector<double>>& TaskLyapunov::accessResult() { return mResults; }

// tasklyapunov.cpp:13-35 (point count, per-device batch, launch report) and
// :45-46,65 (grid deltas, normalisation interval); builds the descriptor
int TaskLyapunov::init(std::shared_ptr<Parameters> paramsBase, const std::string& modelName) {
    mParameters = std::dynamic_pointer_cast<LyapunovParameters>(paramsBase);
    if (!mParameters) {
        printf("error: TaskLyapunov needs LyapunovParameters\n");
        return -1;
    }
    ModelFactoryRegistry::Storage& storage = Singleton<ModelFactoryRegistry::Storage>::Instance();
    auto it = storage._data.find(modelName);
    if (it == storage._data.end()) {
        printf("error: model %s is not registered\n", modelName.c_str());
        return -1;
    }
    if (it->second._model_id < 0) {
        printf("error: model %s is registered but has no compiled kernel in liblyap.so: its right-hand "
               "side and Jacobian must be added as a device model (the reference's device function "
               "pointers are not used)\n", modelName.c_str());
        return -1;
    }
    mModelPtr = &it->second;
    const auto& P = *mParameters;
    mKNodsNum = (size_t)(P.first_sweep_param[3] * P.second_sweep_param[3]);
    const int gpus = P.num_GPUs > 0 ? P.num_GPUs : 1;
    mKNodsGPUBatch = (mKNodsNum + gpus - 1) / gpus;
    while (mBlockNum * mBlockSize > 1.5 * mKNodsGPUBatch && (mBlockNum > 1 || mBlockSize > 1)) {
        if (mBlockNum >= 2) {
            mBlockNum /= 2;
            continue;
        }
        if (mBlockSize >= 2) mBlockSize /= 2;
    }
    printf("block num %zu, block size %zu, total threads %llu, nodes num %zu.\n", mBlockNum, mBlockSize,
           (unsigned long long)(mBlockNum * mBlockSize), mKNodsNum);

    if ((int)P.initial_state.size() < mModelPtr->_dimension ||
        (int)P.initial_model_params.size() < mModelPtr->_param_number) {
        printf("error: model %s needs %d initial_state values and %d initial_model_params\n",
               modelName.c_str(), mModelPtr->_dimension, mModelPtr->_param_number);
        return -1;
    }
    param1_delta = std::fabs(P.first_sweep_param[2] - P.first_sweep_param[1]) / (P.first_sweep_param[3] - 1);
    param2_delta = std::fabs(P.second_sweep_param[2] - P.second_sweep_param[1]) / (P.second_sweep_param[3] - 1);

    lyap_sweep_desc& d = mDesc;
    std::memset(&d, 0, sizeof(d));
    d.model = mModelPtr->_model_id;
    d.dim = mModelPtr->_dimension;
    d.n_exps = P.lyapunov_exponent_num;
    d.n_params = mModelPtr->_param_number;
    d.method = P.method;
    const bool var = P.method == LYAP_METHOD_VAR || P.method == LYAP_METHOD_VAR_TRACE;
    d.integrator = var ? P.integrator : LYAP_INT_DVERK;
    d.adaptive = var ? (P.adaptive ? 1 : 0) : 0;
    d.precision = var ? P.precision : LYAP_PREC_FP64;
    d.h = P.integration_step;
    d.rtol = P.rtol;
    d.atol = P.atol;
    d.h_min = P.h_min;
    d.h_max = P.h_max;
    d.eps = P.epsilon;
    d.t_transient = P.skip_time_attractor;
    d.t_slave = P.skip_time_slave_trajectory;
    d.t_average = P.total_time;
    d.t_renorm = P.skip_time_normalization;
    d.max_steps = P.max_steps;
    d.escape_radius = P.escape_radius;
    d.sweep_idx1 = (int32_t)P.first_sweep_param[0];
    d.sweep_idx2 = (int32_t)P.second_sweep_param[0];
    d.n1 = (int32_t)P.first_sweep_param[3];
    d.n2 = (int32_t)P.second_sweep_param[3];
    d.start1 = P.first_sweep_param[1];
    d.delta1 = param1_delta;
    d.start2 = P.second_sweep_param[1];
    d.delta2 = param2_delta;
    for (int j = 0; j < d.n_params && j < LYAP_MAX_PARAMS; ++j) d.base_params[j] = P.initial_model_params[j];
    for (int i = 0; i < d.dim && i < LYAP_MAX_DIM; ++i) d.init_state[i] = P.initial_state[i];
    d.init_kind = P.init_kind;
    d.init_param_offset = P.init_add_param;
    d.init_scale = P.init_scale;
    d.seed = P.seed;
    if (lyap_desc_finalize(&d) != LYAP_SUCCESS) {
        printf("error: %s\n", lyap_last_error());
        return -1;
    }
    return 0;
}

// tasklyapunov.cpp:37-181: one context and output buffers per device shard;
// inputs are generated on the device, so nothing is copied in
int TaskLyapunov::allocate() {
    auto t = RaiiTimer(&mClocks.allocate);
    if (mModelPtr == nullptr || !mParameters) return -1;
    clear();
    const int gpus = mParameters->num_GPUs > 0 ? mParameters->num_GPUs : 1;
    const int64_t n = (int64_t)mKNodsNum;
    int64_t first = 0;
    for (int g = 0; g < gpus; ++g) {
        Shard s;
        s.device = g;
        s.first = first;
        s.count = n / gpus + (g < n % gpus ? 1 : 0);
        first += s.count;
        if (lyap_ctx_create(g, &s.ctx) != LYAP_SUCCESS) {
            printf("error: %s\n", lyap_last_error());
            return -1;
        }
        if (mParameters->order_probe > 0 &&
            lyap_ctx_set_option(s.ctx, "order_probe", mParameters->order_probe) != LYAP_SUCCESS) {
            printf("error: %s\n", lyap_last_error());
            mShards.push_back(s);
            return -1;
        }
        const size_t cnt = (size_t)(s.count > 0 ? s.count : 1);
        if (!hip_ok(hipMalloc(&s.d_lambda, cnt * mDesc.n_exps * sizeof(double)), "hipMalloc") ||
            !hip_ok(hipMalloc(&s.d_status, cnt * sizeof(uint32_t)), "hipMalloc") ||
            !hip_ok(hipMalloc(&s.d_acc, cnt * sizeof(uint32_t)), "hipMalloc") ||
            !hip_ok(hipMalloc(&s.d_rej, cnt * sizeof(uint32_t)), "hipMalloc")) {
            mShards.push_back(s);
            return -1;
        }
        mShards.push_back(s);
    }
    mResults.reserve(mKNodsNum);
    return 0;
}

// tasklyapunov.cpp:183-250: launch on every device, wait, copy back, rows
int TaskLyapunov::execute() {
    auto t = RaiiTimer(&mClocks.execute, &mExecuteMs);
    mResults.clear();
    mReady = false;
    for (auto& s : mShards) {
        if (s.count == 0) continue;
        if (lyap_sweep_run(s.ctx, &mDesc, s.first, s.count, nullptr, nullptr, nullptr, s.d_lambda, s.d_status,
                           s.d_acc, s.d_rej, nullptr) != LYAP_SUCCESS) {
            printf("error: %s\n", lyap_last_error());
            return -1;
        }
    }
    const int k = mDesc.n_exps;
    lyapExps.assign(mKNodsNum * k, 0.0);
    mStatus.assign(mKNodsNum, 0);
    mAcc.assign(mKNodsNum, 0);
    mRej.assign(mKNodsNum, 0);
    for (auto& s : mShards) {
        if (s.count == 0) continue;
        if (!hip_ok(hipSetDevice(s.device), "hipSetDevice") ||
            !hip_ok(hipDeviceSynchronize(), "sweep kernel") ||
            !hip_ok(hipMemcpy(&lyapExps[s.first * k], s.d_lambda, s.count * k * sizeof(double),
                              hipMemcpyDeviceToHost), "hipMemcpy") ||
            !hip_ok(hipMemcpy(&mStatus[s.first], s.d_status, s.count * sizeof(uint32_t), hipMemcpyDeviceToHost),
                    "hipMemcpy") ||
            !hip_ok(hipMemcpy(&mAcc[s.first], s.d_acc, s.count * sizeof(uint32_t), hipMemcpyDeviceToHost),
                    "hipMemcpy") ||
            !hip_ok(hipMemcpy(&mRej[s.first], s.d_rej, s.count * sizeof(uint32_t), hipMemcpyDeviceToHost),
                    "hipMemcpy"))
            return -1;
    }
    const auto& P = *mParameters;
    size_t paramCount = 0;
    for (int p1 = 0; p1 < (int)P.first_sweep_param[3]; p1++) {
        for (int p2 = 0; p2 < (int)P.second_sweep_param[3]; p2++) {
            const double firstParam = P.first_sweep_param[1] + p1 * param1_delta;
            const double secondParam = P.second_sweep_param[1] + p2 * param2_delta;
            std::vector<double> res;
            res.reserve(2 + k);
            res.push_back(firstParam);
            res.push_back(secondParam);
            for (int i = 0; i < k; i++) res.push_back(lyapExps[paramCount * k + i]);
            mResults.push_back(std::move(res));
            paramCount++;
        }
    }
    mReady = true;
    return (int)mKNodsNum;
}

// tasklyapunov.cpp:252-327
int TaskLyapunov::clear() {
    if (mShards.empty()) return 0;
    auto t = RaiiTimer(&mClocks.clear);
    for (auto& s : mShards) {
        hipSetDevice(s.device);
        if (s.d_lambda) hipFree(s.d_lambda);
        if (s.d_status) hipFree(s.d_status);
        if (s.d_acc) hipFree(s.d_acc);
        if (s.d_rej) hipFree(s.d_rej);
        if (s.ctx) lyap_ctx_destroy(s.ctx);
    }
    mShards.clear();
    return 0;
}
