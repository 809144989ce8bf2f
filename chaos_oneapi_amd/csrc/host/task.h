// Task lifecycle interface with the reference's public surface (task.h:11-73):
// init -> allocate -> execute -> (saveCalculatedData) -> clear -> printTimeReport.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "consts.h"
#include "model_factory.h"
#include "tool_params.h"

class Task {
    struct Clocks {
        uint32_t allocate;
        uint32_t execute;
        uint32_t clear;
    };

public:
    Task() : mBlockSize(DEFAULT_BLOCK_SIZE), mBlockNum(DEFAULT_BLOCK_NUM), mClocks({0, 0, 0}), mReady(false) {}
    virtual ~Task() = default;

    virtual int init(std::shared_ptr<Parameters> paramsBase, const std::string& modelName) = 0;
    virtual int allocate() = 0;
    virtual int execute() = 0;
    virtual int clear() = 0;

    virtual bool isReady() final { return mReady; }
    virtual uint32_t getAllocationTime() { return mClocks.allocate; }
    virtual uint32_t getExecutionTime() { return mClocks.execute; }
    virtual std::vector<std::vector<double>>& accessResult() = 0;

    // milliseconds of the last execute() (the reference only kept whole seconds)
    double getExecutionTimeMs() const { return mExecuteMs; }

    virtual void printTimeReport() final {
        printf("\n========TIME REPORT==========\n");
        printf(" allocation: %u\n", mClocks.allocate);
        printf(" execution:  %u\n", mClocks.execute);
        printf(" clear:      %u\n", mClocks.clear);
        printf(" Total:      %u\n", mClocks.allocate + mClocks.execute + mClocks.clear);
        printf("==============================\n\n");
    }

protected:
    size_t mBlockSize;
    size_t mBlockNum;
    Clocks mClocks;
    double mExecuteMs = 0.0;
    std::vector<std::string> mResultsStr;
    std::vector<std::vector<double>> mResults;
    bool mReady;

    friend int32_t saveCalculatedData(ToolID tool_id, const std::string& modelName, const Parameters* params,
                                      const Task*);
};
