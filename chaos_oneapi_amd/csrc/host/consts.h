// Launch-shape defaults of the reference (consts.h:6-11).  The kernels size
// their own persistent grids from the occupancy of the device; the defaults
// are kept for Task's public mBlockSize/mBlockNum and its launch report.
#pragma once

#define MAX_DIMENSION 32
#define DEFAULT_BLOCK_SIZE 32
#define DEFAULT_BLOCK_NUM 512
