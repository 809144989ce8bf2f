// chaos_lyap: command-line driver with the reference's flow (main.cpp:11-80):
// list the registered models, read config.json for the chosen model, run the
// Task lifecycle, save Data/Lyapunov <model>.json, print the time report.
//
//   chaos_lyap [model_name_or_index] [--config path]
// Without arguments it picks model 0 and ./config.json, as the reference did.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "filehandler.h"
#include "singleton.h"
#include "tasklyapunov.h"

int main(int argc, char** argv) {
    std::string choice, config_path;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--config" && i + 1 < argc) config_path = argv[++i];
        else choice = a;
    }
    ModelFactoryRegistry::Storage& storage = Singleton<ModelFactoryRegistry::Storage>::Instance();
    std::cout << "Available models: " << std::endl;
    int count = 0;
    std::vector<std::string> model_names;
    for (const auto& element : storage._data) {
        model_names.push_back(element.first);
        std::cout << "Model[" << count << "]: \"" << element.first << "\"";
        std::cout << "  Dimension: " << element.second._dimension;
        std::cout << " Number of parameters: " << element.second._param_number << std::endl;
        count++;
    }
    std::string modelName;
    if (choice.empty()) {
        modelName = model_names.at(0);
    } else if (storage._data.count(choice)) {
        modelName = choice;
    } else {
        char* end = nullptr;
        long id = std::strtol(choice.c_str(), &end, 10);
        if (*end != 0 || id < 0 || id >= (long)model_names.size()) {
            std::cout << "Number of model is incorrect: " << choice << " Please try again" << std::endl;
            return 1;
        }
        modelName = model_names.at((size_t)id);
    }
    const auto toolId = ToolID::LYAPUNOV;

    std::shared_ptr<Parameters> params;
    try {
        if (config_path.empty()) {
            params = ConfigFactory::paramsFromConfigFile(toolId, modelName);
        } else {
            std::ifstream f(config_path);
            if (!f.good()) {
                printf("error: config file: %s not found\n", config_path.c_str());
                return 1;
            }
            std::stringstream ss;
            ss << f.rdbuf();
            params = ConfigFactory::paramsFromJson(toolId, modelName, json::parse(ss.str()));
        }
    } catch (std::exception& e) {
        printf("error: %s\n", e.what());
        return 1;
    }
    if (!params) {
        printf("error: config error\n");
        return 1;
    }
    params->Print();

    std::unique_ptr<Task> task = std::make_unique<TaskLyapunov>();
    if (task->init(params, modelName) != 0 || task->allocate() != 0) {
        printf("error: task set-up failed\n");
        return 1;
    }
    const int done = task->execute();
    int rc = 0;
    if (task->isReady()) {
        if (saveCalculatedData(toolId, modelName, params.get(), task.get()) != 0) rc = 1;
    } else {
        printf("error: sweep failed (%d)\n", done);
        rc = 1;
    }
    task->clear();
    task->printTimeReport();
    printf(" execution (ms): %.3f\n", task->getExecutionTimeMs());
    return rc;
}
