// Internal layout of a dtcf_view (shared by the kernels' host launchers).
#pragma once
#include <stdint.h>

struct dtcf_view_s {
    int64_t rows, cols, rpb, cpb, ncb, nblk;
    int64_t *rofs = nullptr, *cofs = nullptr, *blk = nullptr;  // device
    int64_t span = 0;
    bool colpairs = false, rowpairs = false;
    bool colunit = false;  // no block table and cofs[c] = cofs[0] + c: rows are contiguous
    bool owner = true;  // sub-views share their parent's tables
};
