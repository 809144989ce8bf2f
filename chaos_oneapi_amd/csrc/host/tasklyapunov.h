// The Lyapunov sweep task (reference tasklyapunov.h:8-59 / tasklyapunov.cpp).
// It builds the sweep descriptor from LyapunovParameters, splits the grid into
// contiguous per-device shards (remainder included, unlike the reference's
// truncating split at tasklyapunov.cpp:19), runs lyap_sweep_run on each
// device, and assembles rows [p1, p2, lambda_1..lambda_k], p1 outer.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "chaos_lyap.h"
#include "task.h"

class TaskLyapunov : public Task {
public:
    ~TaskLyapunov() override;

    std::vector<std::vector<double>>& accessResult() override;
    int init(std::shared_ptr<Parameters> paramsBase, const std::string& modelName) override;
    int allocate() override;
    int execute() override;
    int clear() override;

    // the descriptor the sweep runs with (valid after init)
    const lyap_sweep_desc& descriptor() const { return mDesc; }
    // per-point status words and accepted/rejected step counts (after execute)
    const std::vector<uint32_t>& pointStatus() const { return mStatus; }
    const std::vector<uint32_t>& acceptedSteps() const { return mAcc; }
    const std::vector<uint32_t>& rejectedSteps() const { return mRej; }

private:
    struct Shard {
        int device = 0;
        int64_t first = 0, count = 0;
        lyap_ctx* ctx = nullptr;
        double* d_lambda = nullptr;
        uint32_t* d_status = nullptr;
        uint32_t* d_acc = nullptr;
        uint32_t* d_rej = nullptr;
    };

    size_t mKNodsNum = 0, mKNodsGPUBatch = 0;
    double param1_delta = 0.0, param2_delta = 0.0;
    std::shared_ptr<LyapunovParameters> mParameters;
    ModelFactoryRegistry::ModelData* mModelPtr = nullptr;
    lyap_sweep_desc mDesc{};
    std::vector<Shard> mShards;
    std::vector<double> lyapExps;
    std::vector<uint32_t> mStatus, mAcc, mRej;
};
