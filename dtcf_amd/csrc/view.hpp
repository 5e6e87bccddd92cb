// Internal layout of a dtcf_view (shared by the kernels' host launchers).
#pragma once
#include <stdint.h>

struct dtcf_view_s {
    int64_t rows, cols, rpb, cpb, ncb, nblk;
    int64_t *rofs = nullptr, *cofs = nullptr, *blk = nullptr;  // device
    int64_t span = 0;
    bool colpairs = false, rowpairs = false;
    bool rowrun16 = false;  // no block table and rofs[r] = rofs[r & ~15] + (r & 15)
    bool owner = true;  // sub-views share their parent's tables
};
