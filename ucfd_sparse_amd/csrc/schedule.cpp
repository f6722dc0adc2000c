// Dependency-level schedules (see schedule.h).
#include "schedule.h"

#include <algorithm>
#include <string.h>

void ucfd_build_levels(int nrows, const std::vector<int> &seq, const std::vector<int> &jb,
                       const std::vector<int> &je, const std::vector<int> &ci,
                       std::vector<std::vector<int>> &levels, bool *reads_unvisited)
{
    std::vector<int> pos(nrows, -1), lev(nrows, 0), need(nrows, 0);
    for (size_t k = 0; k < seq.size(); ++k) pos[seq[k]] = (int)k;
    *reads_unvisited = false;
    int maxlev = 0;
    for (size_t k = 0; k < seq.size(); ++k) {
        const int i = seq[k];
        int l = need[i];
        for (int j = jb[i]; j < je[i]; ++j) {
            const int c = ci[j];
            if (c == i) {
                *reads_unvisited = true;
                continue;
            }
            if (pos[c] >= 0 && pos[c] < (int)k) l = std::max(l, lev[c]);
            else *reads_unvisited = true;
        }
        lev[i] = l + 1;
        maxlev = std::max(maxlev, lev[i]);
        for (int j = jb[i]; j < je[i]; ++j) {
            const int c = ci[j];
            if (c != i && pos[c] > (int)k) need[c] = std::max(need[c], lev[i]);
        }
    }
    levels.assign(maxlev, {});
    for (int i : seq) levels[lev[i] - 1].push_back(i);
    for (auto &v : levels) std::sort(v.begin(), v.end());
}

int ucfd_finish_schedule(const std::vector<std::vector<int>> &levels, UcfdSchedule &S,
                         const char *fn)
{
    std::vector<int> ord;
    S.lv.clear();
    for (const auto &v : levels) {
        if (v.empty()) continue;
        const bool contig = v.back() - v.front() + 1 == (int)v.size();
        if (contig) {
            S.lv.push_back(UcfdLevel{v.front(), (int)v.size(), false});
        } else {
            S.lv.push_back(UcfdLevel{(int)ord.size(), (int)v.size(), true});
            ord.insert(ord.end(), v.begin(), v.end());
        }
    }
    if (S.order) (void)hipFree(S.order);
    S.order = nullptr;
    if (ord.empty()) return UCFD_OK;
    int rc = ucfd_check_hip(hipMalloc((void **)&S.order, ord.size()*sizeof(int)), fn,
                            "hipMalloc(schedule)");
    if (rc) return rc;
    return ucfd_check_hip(hipMemcpy(S.order, ord.data(), ord.size()*sizeof(int),
                                    hipMemcpyHostToDevice), fn, "hipMemcpy(schedule)");
}

void ucfd_free_schedule(UcfdSchedule &S)
{
    if (S.order) (void)hipFree(S.order);
    S.order = nullptr;
    S.lv.clear();
}

int ucfd_fetch_ints(const int *p, size_t n, std::vector<int> &out, const char *fn)
{
    out.resize(n);
    if (n == 0) return UCFD_OK;
    if (ucfd_is_device_ptr(p))
        return ucfd_check_hip(hipMemcpy(out.data(), p, n*sizeof(int), hipMemcpyDeviceToHost), fn,
                              "hipMemcpy(structure)");
    memcpy(out.data(), p, n*sizeof(int));
    return UCFD_OK;
}
