// Dependency-level schedules for in-place sequential sweeps run in parallel.
//
// A reference sweep visits rows one at a time in a fixed order and updates a
// vector in place; row i reads the entries of the columns colidx[jb[i] ..
// je[i]) - values already produced for rows visited earlier, original values
// for rows visited later.  Rows are grouped into levels so that every row
// runs after the rows whose new values it reads and before the rows whose old
// values it reads; rows of one level never read each other.  Running the
// levels in order, each as one kernel launch, reproduces the sequential sweep
// exactly.  On a colour-ordered numbering each colour is one level.
#pragma once

#include <vector>

#include "ucfd_internal.h"

struct UcfdLevel {
    int base, cnt;
    bool ordered;   // rows are order[base .. base+cnt) instead of base .. base+cnt-1
    int maxw = -1;  // wave-staged kernels: most blocks one wave's rows use (-1: unknown)
};

struct UcfdSchedule {
    std::vector<UcfdLevel> lv;
    int *order = nullptr;   // device array of the rows of ordered levels
};

// levels[k] = rows of level k, ascending.  *reads_unvisited: some row reads
// a row not yet produced (or itself) in the sweep.
UCFD_HIDDEN void ucfd_build_levels(int nrows, const std::vector<int> &seq, const std::vector<int> &jb,
                                   const std::vector<int> &je, const std::vector<int> &colidx,
                                   std::vector<std::vector<int>> &levels, bool *reads_unvisited);
// Turns levels into launches: contiguous levels become ranges, the others
// index a device order array (allocated here).
UCFD_HIDDEN int ucfd_finish_schedule(const std::vector<std::vector<int>> &levels, UcfdSchedule &S,
                                     const char *fn);
UCFD_HIDDEN void ucfd_free_schedule(UcfdSchedule &S);
// Copy n ints of a host or device array to host memory.
UCFD_HIDDEN int ucfd_fetch_ints(const int *p, size_t n, std::vector<int> &out, const char *fn);
