#include "filehandler.h"

#include <sys/stat.h>

#include <cassert>
#include <cstdio>
#include <fstream>
#include <iostream>

#include "singleton.h"

// filehandler.cpp:5-33 of the reference
std::string GetToolName(int tool_id) {
    switch (tool_id) {
        case 0: return "Lyapunov";
        case 1: return "Kneadings";
        case 2: return "Angle";
        case 3: return "minDistance";
        default: return "Unknown";
    }
}

int32_t saveNpy(const std::string& path, const std::vector<std::vector<double>>& rows) {
    const size_t nr = rows.size(), nc = nr ? rows[0].size() : 0;
    std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(nr) + ", " +
                         std::to_string(nc) + "), }";
    while ((10 + header.size() + 1) % 64 != 0) header += ' ';
    header += '\n';
    std::ofstream f(path, std::ios::binary);
    if (!f.good()) return -1;
    const char magic[] = "\x93NUMPY\x01\x00";
    f.write(magic, 8);
    const uint16_t hl = (uint16_t)header.size();
    f.write(reinterpret_cast<const char*>(&hl), 2);
    f.write(header.data(), (std::streamsize)header.size());
    for (const auto& r : rows) {
        if (r.size() != nc) return -1;
        f.write(reinterpret_cast<const char*>(r.data()), (std::streamsize)(nc * sizeof(double)));
    }
    return f.good() ? 0 : -1;
}

// filehandler.cpp:35-89 of the reference
int32_t saveCalculatedData(ToolID tool_id, const std::string& modelName, const Parameters* params,
                           const Task* task) {
    if (!task->mResultsStr.empty()) assert(task->mResultsStr.size() == task->mResults.size());
    assert(params != nullptr);
    json outputData = json::object();
    outputData["Parameters"] = params->paramsToJson();
    const std::string tool_name = GetToolName(tool_id);
    outputData["ID_Tool"] = tool_name;

    std::string data_str;
    if (!task->mResults.empty()) data_str.reserve(task->mResults.size() * task->mResults.at(0).size() * 9);
    size_t row_cnt = 0;
    for (const auto& row : task->mResults) {
        for (double val : row) {
            data_str.append(std::to_string(val));
            data_str.append(" ");
        }
        if (!task->mResultsStr.empty()) {
            data_str.append(task->mResultsStr.at(row_cnt));
            data_str.append(" ");
        }
        data_str.at(data_str.size() - 1) = '\n';
        row_cnt += 1;
    }
    outputData["_Data"] = data_str;

    ModelFactoryRegistry::Storage& storage = Singleton<ModelFactoryRegistry::Storage>::Instance();
    const auto& m = storage._data.at(modelName);
    json modelData = json::object();
    modelData["Name"] = m._model_name;
    modelData["Dimension"] = m._dimension;
    modelData["Number of parameters"] = m._param_number;
    modelData["Name of parameters"] = m._model_param;
    modelData["Name of phase"] = m._model_phase;
    modelData["Equations"] = m._model_equat;
    modelData["systemType"] = (int)m._type;
    outputData["Model"] = modelData;

    mkdir("Data", 0755);  // the reference required Data/ to exist already
    const std::string stem = "Data/" + tool_name + " " + m._model_name;
    std::ofstream jsonOStream(stem + ".json");
    if (!jsonOStream.good()) {
        std::cout << "Unable to write the file: " << stem << ".json" << std::endl;
        return -1;
    }
    jsonOStream << outputData.dump(4) << std::endl;
    if (saveNpy(stem + ".npy", task->mResults) != 0) {
        std::cout << "Unable to write the file: " << stem << ".npy" << std::endl;
        return -1;
    }
    return 0;
}
