// Result output (reference filehandler.h:23-24, filehandler.cpp:35-89):
// Data/<Tool> <model>.json with Parameters, ID_Tool, _Data (rows printed with
// std::to_string, space separated) and Model.  This build also creates Data/
// when it is missing and writes a lossless FP64 sidecar, Data/<Tool> <model>.npy.
#pragma once

#include <cstdint>
#include <string>

#include "task.h"
#include "tool_params.h"

std::string GetToolName(int tool_id);

int32_t saveCalculatedData(ToolID tool_id, const std::string& modelName, const Parameters* params,
                           const Task* task);

// writes rows (all of equal length) as a 2-D little-endian float64 .npy file
int32_t saveNpy(const std::string& path, const std::vector<std::vector<double>>& rows);
