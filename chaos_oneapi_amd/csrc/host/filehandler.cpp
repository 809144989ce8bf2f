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

namespace {

// One text row of the _Data block: every value as std::to_string prints it
// (fixed notation, six decimals), then the row's optional label, separated by
// single spaces and closed by a newline.
void append_row(std::string& text, const std::vector<double>& values, const std::string* label) {
    bool first = true;
    for (double v : values) {
        if (!first) text += ' ';
        text += std::to_string(v);
        first = false;
    }
    if (label) {
        if (!first) text += ' ';
        text += *label;
    }
    text += '\n';
}

json model_block(const ModelFactoryRegistry::ModelData& m) {
    json block = json::object();
    block["Name"] = m._model_name;
    block["Dimension"] = m._dimension;
    block["Number of parameters"] = m._param_number;
    block["Name of parameters"] = m._model_param;
    block["Name of phase"] = m._model_phase;
    block["Equations"] = m._model_equat;
    block["systemType"] = (int)m._type;
    return block;
}

}  // namespace

// The output of the reference's saveCalculatedData (filehandler.cpp:35-89):
// one JSON object {Parameters, ID_Tool, _Data, Model} pretty-printed with a
// four-space indent into "Data/<tool> <model>.json".  Here the folder is
// created when missing and the rows also go, losslessly, to a .npy sidecar.
int32_t saveCalculatedData(ToolID tool_id, const std::string& modelName, const Parameters* params,
                           const Task* task) {
    assert(params != nullptr);
    const auto& rows = task->mResults;
    const auto& labels = task->mResultsStr;
    assert(labels.empty() || labels.size() == rows.size());
    const std::string tool = GetToolName(tool_id);

    std::string table;
    table.reserve(rows.empty() ? 0 : rows.size() * rows.front().size() * 12);
    for (size_t r = 0; r < rows.size(); ++r) append_row(table, rows[r], labels.empty() ? nullptr : &labels[r]);

    const auto& models = Singleton<ModelFactoryRegistry::Storage>::Instance()._data;
    const auto& model = models.at(modelName);
    json doc = json::object();
    doc["Parameters"] = params->paramsToJson();
    doc["ID_Tool"] = tool;
    doc["_Data"] = table;
    doc["Model"] = model_block(model);

    mkdir("Data", 0755);
    const std::string base = "Data/" + tool + " " + model._model_name;
    std::ofstream out(base + ".json");
    if (!out.good()) {
        std::cout << "Unable to write the file: " << base << ".json" << std::endl;
        return -1;
    }
    out << doc.dump(4) << std::endl;
    if (saveNpy(base + ".npy", rows) != 0) {
        std::cout << "Unable to write the file: " << base << ".npy" << std::endl;
        return -1;
    }
    return 0;
}
