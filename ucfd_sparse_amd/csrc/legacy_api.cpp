// Legacy LU-SGS array API (the one the reference examples and pyBaram call):
// {serial,parallel}_pre_lusgs, the C in-place sweeps {serial,parallel}_ns_*_sweep,
// the Python two-array sweeps ns_{serial,parallel}_*_sweep and the update
// spellings.  Signatures are recovered from the reference call sites (the
// implementation files are absent from the reference):
//   examples/omp3d.c:180-189, examples/mpi3d.c:166-173,
//   examples/omp3d.py:22-35,76-94, examples/mpi3d.py:21-34,71-81.
//
// Colored calls launch one kernel over the cells icolor[n0:ne); a neighbour is
// "lower" iff lcolor[nb] < lcolor[i] (upper: >).  Serial calls keep the
// sequential semantics of src/lusgs/lusgs.c:175,227 (lower iff nb < i) and run
// them level by level: a cell's level is one more than the highest level of
// its lower neighbours, so each level is independent and the result equals
// the sequential sweep exactly.
#include "ucfd_internal.h"

#include <algorithm>
#include <mutex>
#include <vector>

#include "coloredlusgs.h"
#include "lusgs.h"
#include "ucfd_ext.h"

namespace {

bool bad_dims(int neles, int nface) { return neles < 0 || nface < 1 || nface > 32; }

// Colour-contiguous layouts declared by the caller (ucfd_declare_colored_layout):
// icolor is the identity and lcolor is non-decreasing, so colour c occupies
// the index range [off[c], off[c+1]).  A coloured call on such a layout whose
// [n0, ne) is one whole colour takes the neighbour test "nb < n0" / "nb >= ne"
// instead of reading lcolor[nb]; that is the same test, without the gather.
// Layouts are found by their pointers; since an address can be freed and
// reused for other arrays, every range-mode launch samples one cell per
// workgroup (icolor[i] == i, lcolor[i] == the colour's declared value) and a
// mismatch fails the call with UCFD_ERR_STATE instead of sweeping silently
// with the wrong neighbour test.
struct DeclaredLayout {
    const int *icolor;
    const int *lcolor;
    int neles;
    std::vector<int> off;      // colour c = [off[c], off[c+1])
    std::vector<int> value;    // lcolor value of colour c
};
std::mutex g_layout_mu;
std::vector<DeclaredLayout> g_layouts;

// true if [n0, ne) is one whole colour of a layout declared with these
// pointers; *lc_value = that colour's lcolor value
bool whole_colour_of_declared(const int *icolor, const int *lcolor, int neles, int n0, int ne,
                              int *lc_value)
{
    std::lock_guard<std::mutex> g(g_layout_mu);
    for (const DeclaredLayout &l : g_layouts) {
        if (l.icolor != icolor || l.lcolor != lcolor || l.neles != neles) continue;
        auto it = std::lower_bound(l.off.begin(), l.off.end(), n0);
        if (it != l.off.end() && *it == n0 && it + 1 != l.off.end() && *(it + 1) == ne) {
            *lc_value = l.value[it - l.off.begin()];
            return true;
        }
        return false;
    }
    return false;
}

int do_pre(const char *fn, int neles, int nface, double factor, const double *fnorm_vol,
           const double *dt, double *diag, const double *fspr)
{
    if (bad_dims(neles, nface)) return ucfd_report(UCFD_ERR_ARG, fn, "bad neles/nface");
    if (neles == 0) return UCFD_OK;
    const size_t N = (size_t)neles, F = (size_t)nface;
    UcfdStage sg(fn);
    const int a_fnv = sg.add(fnorm_vol, F*N*8, UCFD_IN);
    const int a_dt = sg.add(dt, N*8, UCFD_IN);
    const int a_dg = sg.add(diag, N*8, UCFD_OUT);
    const int a_sp = sg.add(fspr, F*N*8, UCFD_IN);
    int rc = sg.enter();
    if (rc) return rc;
    rc = ucfd_launch_pre(ucfd_stream(), neles, nface, factor, (const double *)sg.dev(a_fnv),
                         (const double *)sg.dev(a_dt), (double *)sg.dev(a_dg),
                         (const double *)sg.dev(a_sp));
    if (rc) ucfd_report(rc, fn, "kernel launch failed");
    return sg.leave(rc);
}

// Turbulence rows of the RANS sweeps: NVARS - NFVARS, 1 (Spalart-Allmaras)
// or 2 (k-omega); 0 when the configuration has none.
int rans_rows(const char *fn, int nface)
{
    const UcfdConfig &c = ucfd_config();
    const int nt = c.nvars - c.nfvars;
    if (nt < 1 || nt > 2) {
        ucfd_report(UCFD_ERR_ARG, fn, "RANS sweeps need NVARS - NFVARS turbulence rows in "
                                      "{1, 2} (ucfd_set_config)");
        return 0;
    }
    if (nface != 3 && nface != 4 && nface != 6) {
        ucfd_report(UCFD_ERR_ARG, fn, "RANS sweeps support nface 3, 4 or 6");
        return 0;
    }
    return nt;
}

// One coloured sweep over icolor[n0:ne).  lower: out = (in - 0.5 df)/diag;
// upper: out = in - 0.5 df/diag.  in == out is the in-place C form.  rans:
// the turbulence rows (diag / fspr are then the turbulence diagonal and
// spectral radius), else the NFVARS flow rows.
int do_colored(const char *fn, int dir, int n0, int ne, int neles, int nface, const int *nei,
               const int *icolor, const int *lcolor, const double *fnorm_vol,
               const double *vec_fnorm, const double *upts, const double *in, double *out,
               const double *diag, const double *fspr, bool rans = false)
{
    if (bad_dims(neles, nface) || n0 < 0 || ne < n0 || ne > neles)
        return ucfd_report(UCFD_ERR_ARG, fn, "bad n0/ne/neles/nface");
    const int nt = rans ? rans_rows(fn, nface) : 0;
    if (rans && !nt) return UCFD_ERR_ARG;
    if (ne == n0) return UCFD_OK;
    const UcfdConfig &c = ucfd_config();
    const size_t N = (size_t)neles, F = (size_t)nface, D = (size_t)c.ndims,
                 V = rans ? (size_t)c.nvars : (size_t)c.nfvars;
    UcfdStage sg(fn);
    const int a_nei = sg.add(nei, F*N*4, UCFD_IN);
    const int a_ic = sg.add(icolor, N*4, UCFD_IN);
    const int a_lc = sg.add(lcolor, N*4, UCFD_IN);
    const int a_fnv = sg.add(fnorm_vol, F*N*8, UCFD_IN);
    const int a_vfn = sg.add(vec_fnorm, F*D*N*8, UCFD_IN);
    const int a_u = sg.add(upts, V*N*8, UCFD_IN);
    const int a_in = sg.add(in, V*N*8, UCFD_IN);
    const int a_out = sg.add(out, V*N*8, UCFD_INOUT);
    const int a_dg = sg.add(diag, N*8, UCFD_IN);
    const int a_sp = sg.add(fspr, F*N*8, UCFD_IN);
    int lc_value = 0;
    const int mode = whole_colour_of_declared(icolor, lcolor, neles, n0, ne, &lc_value)
                         ? UCFD_SWEEP_RANGE : UCFD_SWEEP_COLORED;
    int rc = sg.enter();
    if (rc) return rc;
    if (rans)
        rc = ucfd_launch_rans_sweep(ucfd_stream(), dir, mode, n0, ne, neles, nface, c.ndims, nt,
                                    (const int *)sg.dev(a_nei), (const int *)sg.dev(a_ic),
                                    (const int *)sg.dev(a_lc), (const double *)sg.dev(a_fnv),
                                    (const double *)sg.dev(a_vfn), (const double *)sg.dev(a_u),
                                    (const double *)sg.dev(a_in), (double *)sg.dev(a_out),
                                    (const double *)sg.dev(a_dg), (const double *)sg.dev(a_sp),
                                    ucfd_device_flag(), lc_value);
    else
        rc = ucfd_launch_sweep(ucfd_stream(), dir, mode, n0, ne, neles, nface, c.ndims,
                               c.nfvars, (const int *)sg.dev(a_nei), (const int *)sg.dev(a_ic),
                               (const int *)sg.dev(a_lc), (const double *)sg.dev(a_fnv),
                               (const double *)sg.dev(a_vfn), (const double *)sg.dev(a_u),
                               (const double *)sg.dev(a_in), (double *)sg.dev(a_out),
                               (const double *)sg.dev(a_dg), (const double *)sg.dev(a_sp),
                               ucfd_device_flag(), lc_value);
    if (rc) ucfd_report(rc, fn, "kernel launch failed");
    return sg.leave(rc);
}

// Level schedules of the serial sweeps, one per direction, kept for the
// last neighbour table seen: the key is the array's address, the sizes and a
// content hash of nei_ele (about 0.1 ms on the device), so a schedule is
// reused only for the same table and rebuilt whenever an entry changes.
struct SerialSchedule {
    const void *nei = nullptr;
    int neles = -1, nface = -1;
    bool dev = false;
    unsigned long long hash = 0;
    std::vector<int> off;
    int *order_d = nullptr;     // device copy of the level order
    size_t cap = 0;
};
SerialSchedule g_serial[2];
std::mutex g_serial_mu;
unsigned long long *g_hash_d = nullptr;

// Dependency levels of the sequential sweep: lower (nb < i) or upper (nb > i).
// order[] lists cells level by level, off[l]..off[l+1] is level l.
void build_levels(int dir, int N, int F, const int *nei, std::vector<int> &order,
                  std::vector<int> &off)
{
    std::vector<int> lev(N, 0);
    int maxl = 0;
    if (dir == 0) {
        for (int i = 0; i < N; ++i) {
            int l = 0;
            for (int f = 0; f < F; ++f) {
                const int nb = nei[(size_t)f*N + i];
                if (nb >= 0 && nb < i) l = std::max(l, lev[nb] + 1);
            }
            lev[i] = l;
            maxl = std::max(maxl, l);
        }
    } else {
        for (int i = N - 1; i >= 0; --i) {
            int l = 0;
            for (int f = 0; f < F; ++f) {
                const int nb = nei[(size_t)f*N + i];
                if (nb > i && nb < N) l = std::max(l, lev[nb] + 1);
            }
            lev[i] = l;
            maxl = std::max(maxl, l);
        }
    }
    off.assign(maxl + 2, 0);
    for (int i = 0; i < N; ++i) off[lev[i] + 1]++;
    for (int l = 0; l <= maxl; ++l) off[l + 1] += off[l];
    order.resize(N);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int i = 0; i < N; ++i) order[pos[lev[i]]++] = i;
}

int do_serial(const char *fn, int dir, int neles, int nface, const int *nei,
              const double *fnorm_vol, const double *vec_fnorm, const double *upts,
              const double *in, double *out, const double *diag, const double *fspr,
              bool rans = false)
{
    if (bad_dims(neles, nface)) return ucfd_report(UCFD_ERR_ARG, fn, "bad neles/nface");
    const int nt = rans ? rans_rows(fn, nface) : 0;
    if (rans && !nt) return UCFD_ERR_ARG;
    if (neles == 0) return UCFD_OK;
    const UcfdConfig &c = ucfd_config();
    const size_t N = (size_t)neles, F = (size_t)nface, D = (size_t)c.ndims,
                 V = rans ? (size_t)c.nvars : (size_t)c.nfvars;
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;

    // the level schedule is built on the host from nei_ele, or reused; one
    // lock covers the hash buffer, the hash and the schedule cache, so
    // concurrent serial calls neither allocate nor read the hash twice
    if (!nei) return ucfd_report(UCFD_ERR_ARG, fn, "NULL nei_ele");
    std::lock_guard<std::mutex> lock(g_serial_mu);
    const bool dev = ucfd_is_device_ptr(nei);
    unsigned long long h = 0;
    if (dev) {
        if (!g_hash_d && (rc = ucfd_check_hip(hipMalloc((void **)&g_hash_d, 64), fn,
                                              "hipMalloc(hash)")))
            return rc;
        rc = ucfd_launch_hash(ucfd_stream(), nei, (int64_t)(F*N), g_hash_d);
        if (!rc) rc = ucfd_check_hip(hipMemcpyAsync(&h, g_hash_d, sizeof h, hipMemcpyDeviceToHost,
                                                    ucfd_stream()), fn, "hipMemcpyAsync(hash)");
        if (!rc) rc = ucfd_check_hip(hipStreamSynchronize(ucfd_stream()), fn, "sync");
        if (rc) return rc;
    } else {
        h = ucfd_hash_host(nei, (int64_t)(F*N));
    }
    SerialSchedule &C = g_serial[dir];
    if (!(C.nei == nei && C.neles == neles && C.nface == nface && C.dev == dev && C.hash == h)) {
        std::vector<int> hnei;
        const int *nei_h = nei;
        if (dev) {
            hnei.resize(F*N);
            rc = ucfd_check_hip(hipMemcpy(hnei.data(), nei, F*N*4, hipMemcpyDeviceToHost), fn,
                                "hipMemcpy(nei_ele)");
            if (rc) return rc;
            nei_h = hnei.data();
        }
        std::vector<int> order;
        C.nei = nullptr;
        build_levels(dir, neles, nface, nei_h, order, C.off);
        if (C.cap < N) {
            if (C.order_d) (void)hipFree(C.order_d);
            C.order_d = nullptr;
            C.cap = 0;
            if ((rc = ucfd_check_hip(hipMalloc((void **)&C.order_d, N*4), fn, "hipMalloc(order)")))
                return rc;
            C.cap = N;
        }
        if ((rc = ucfd_check_hip(hipMemcpy(C.order_d, order.data(), N*4, hipMemcpyHostToDevice),
                                 fn, "hipMemcpy(order)")))
            return rc;
        C.nei = nei;
        C.neles = neles;
        C.nface = nface;
        C.dev = dev;
        C.hash = h;
    }
    const std::vector<int> &off = C.off;

    UcfdStage sg(fn);
    const int a_nei = sg.add(nei, F*N*4, UCFD_IN);
    const int a_ord = sg.add(C.order_d, N*4, UCFD_IN);
    const int a_fnv = sg.add(fnorm_vol, F*N*8, UCFD_IN);
    const int a_vfn = sg.add(vec_fnorm, F*D*N*8, UCFD_IN);
    const int a_u = sg.add(upts, V*N*8, UCFD_IN);
    const int a_in = sg.add(in, V*N*8, UCFD_IN);
    const int a_out = sg.add(out, V*N*8, UCFD_INOUT);
    const int a_dg = sg.add(diag, N*8, UCFD_IN);
    const int a_sp = sg.add(fspr, F*N*8, UCFD_IN);
    rc = sg.enter();
    if (rc) return rc;
    const int nlev = (int)off.size() - 1;
    for (int l = 0; l < nlev && !rc; ++l) {
        if (rans)
            rc = ucfd_launch_rans_sweep(ucfd_stream(), dir, UCFD_SWEEP_ORDERED, off[l], off[l + 1],
                                        neles, nface, c.ndims, nt, (const int *)sg.dev(a_nei),
                                        (const int *)sg.dev(a_ord), nullptr,
                                        (const double *)sg.dev(a_fnv),
                                        (const double *)sg.dev(a_vfn),
                                        (const double *)sg.dev(a_u), (const double *)sg.dev(a_in),
                                        (double *)sg.dev(a_out), (const double *)sg.dev(a_dg),
                                        (const double *)sg.dev(a_sp), ucfd_device_flag(), 0);
        else
            rc = ucfd_launch_sweep(ucfd_stream(), dir, UCFD_SWEEP_ORDERED, off[l], off[l + 1],
                                   neles, nface, c.ndims, c.nfvars, (const int *)sg.dev(a_nei),
                                   (const int *)sg.dev(a_ord), nullptr,
                                   (const double *)sg.dev(a_fnv), (const double *)sg.dev(a_vfn),
                                   (const double *)sg.dev(a_u), (const double *)sg.dev(a_in),
                                   (double *)sg.dev(a_out), (const double *)sg.dev(a_dg),
                                   (const double *)sg.dev(a_sp), ucfd_device_flag(), 0);
    }
    if (rc) ucfd_report(rc, fn, "kernel launch failed");
    return sg.leave(rc);
}

int do_update(const char *fn, int neles, double *upts, const double *x)
{
    if (neles < 0) return ucfd_report(UCFD_ERR_ARG, fn, "bad neles");
    if (neles == 0) return UCFD_OK;
    const size_t n = (size_t)neles*(size_t)ucfd_config().nvars;
    UcfdStage sg(fn);
    const int a_u = sg.add(upts, n*8, UCFD_INOUT);
    const int a_x = sg.add(x, n*8, UCFD_IN);
    int rc = sg.enter();
    if (rc) return rc;
    rc = ucfd_launch_update(ucfd_stream(), neles, ucfd_config().nvars, (double *)sg.dev(a_u),
                            (const double *)sg.dev(a_x));
    if (rc) ucfd_report(rc, fn, "kernel launch failed");
    return sg.leave(rc);
}

}  // namespace

// ------------------------------------------------------------------ coloured

UCFD_API int parallel_pre_lusgs(int neles, int nface, double factor, double *fnorm_vol,
                                double *dt, double *diag, double *fspr)
{
    return do_pre("parallel_pre_lusgs", neles, nface, factor, fnorm_vol, dt, diag, fspr);
}

UCFD_API int parallel_ns_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                     int *icolor, int *lcolor, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *dub, double *diag,
                                     double *fspr)
{
    return do_colored("parallel_ns_lower_sweep", 0, n0, ne, neles, nface, nei_ele, icolor, lcolor,
                      fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr);
}

UCFD_API int parallel_ns_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                     int *icolor, int *lcolor, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *dub, double *diag,
                                     double *fspr)
{
    return do_colored("parallel_ns_upper_sweep", 1, n0, ne, neles, nface, nei_ele, icolor, lcolor,
                      fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr);
}

UCFD_API int ns_parallel_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                     int *icolor, int *lcolor, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                     double *diag, double *fspr)
{
    return do_colored("ns_parallel_lower_sweep", 0, n0, ne, neles, nface, nei_ele, icolor, lcolor,
                      fnorm_vol, vec_fnorm, upts, rhsb, dub, diag, fspr);
}

UCFD_API int ns_parallel_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                     int *icolor, int *lcolor, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                     double *diag, double *fspr)
{
    return do_colored("ns_parallel_upper_sweep", 1, n0, ne, neles, nface, nei_ele, icolor, lcolor,
                      fnorm_vol, vec_fnorm, upts, dub, rhsb, diag, fspr);
}

UCFD_API int lusgs_parallel_ns_update(int neles, double *upts, double *x)
{
    return do_update("lusgs_parallel_ns_update", neles, upts, x);
}

UCFD_API int parallel_ns_lusgs_update(int neles, double *upts, double *x)
{
    return do_update("parallel_ns_lusgs_update", neles, upts, x);
}

// -------------------------------------------------------------------- serial

UCFD_API int serial_pre_lusgs(int neles, int nface, double factor, double *fnorm_vol, double *dt,
                              double *diag, double *fspr)
{
    return do_pre("serial_pre_lusgs", neles, nface, factor, fnorm_vol, dt, diag, fspr);
}

UCFD_API int serial_ns_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                   double *vec_fnorm, double *upts, double *dub, double *diag,
                                   double *fspr)
{
    return do_serial("serial_ns_lower_sweep", 0, neles, nface, nei_ele, fnorm_vol, vec_fnorm, upts,
                     dub, dub, diag, fspr);
}

UCFD_API int serial_ns_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                   double *vec_fnorm, double *upts, double *dub, double *diag,
                                   double *fspr)
{
    return do_serial("serial_ns_upper_sweep", 1, neles, nface, nei_ele, fnorm_vol, vec_fnorm, upts,
                     dub, dub, diag, fspr);
}

UCFD_API int ns_serial_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                   double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                   double *diag, double *fspr)
{
    return do_serial("ns_serial_lower_sweep", 0, neles, nface, nei_ele, fnorm_vol, vec_fnorm, upts,
                     rhsb, dub, diag, fspr);
}

UCFD_API int ns_serial_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                   double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                   double *diag, double *fspr)
{
    return do_serial("ns_serial_upper_sweep", 1, neles, nface, nei_ele, fnorm_vol, vec_fnorm, upts,
                     dub, rhsb, diag, fspr);
}

UCFD_API int serial_ns_lusgs_update(int neles, double *upts, double *dub)
{
    return do_update("serial_ns_lusgs_update", neles, upts, dub);
}

UCFD_API int lusgs_serial_ns_update(int neles, double *upts, double *rhsb)
{
    return do_update("lusgs_serial_ns_update", neles, upts, rhsb);
}

// ---------------------------------------------------------------- RANS rows
// The {rans}_{serial,parallel}_{lower,upper}_sweep family named by
// doc/1_introduction.md:19-22,40-41 (no call site in the reference): the NS
// argument lists with diag / fspr taken as the turbulence diagonal and
// spectral radius (the diagonal from {serial,parallel}_pre_lusgs with the
// turbulence spectral radius and factor = turb_factor, as lusgs.c:23-36,
// 127-137 form the turbulence rows of diag), acting on the turbulence rows
// NFVARS .. NVARS-1 of dub / rhsb with the neighbour state upts; arithmetic
// of lusgs.c:274-321 with rans_flux_container (flux.c:77-90).

UCFD_API int parallel_rans_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                       int *icolor, int *lcolor, double *fnorm_vol,
                                       double *vec_fnorm, double *upts, double *dub, double *tdiag,
                                       double *tfspr)
{
    return do_colored("parallel_rans_lower_sweep", 0, n0, ne, neles, nface, nei_ele, icolor,
                      lcolor, fnorm_vol, vec_fnorm, upts, dub, dub, tdiag, tfspr, true);
}

UCFD_API int parallel_rans_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                       int *icolor, int *lcolor, double *fnorm_vol,
                                       double *vec_fnorm, double *upts, double *dub, double *tdiag,
                                       double *tfspr)
{
    return do_colored("parallel_rans_upper_sweep", 1, n0, ne, neles, nface, nei_ele, icolor,
                      lcolor, fnorm_vol, vec_fnorm, upts, dub, dub, tdiag, tfspr, true);
}

UCFD_API int rans_parallel_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                       int *icolor, int *lcolor, double *fnorm_vol,
                                       double *vec_fnorm, double *upts, double *rhsb,
                                       double *dub, double *tdiag, double *tfspr)
{
    return do_colored("rans_parallel_lower_sweep", 0, n0, ne, neles, nface, nei_ele, icolor,
                      lcolor, fnorm_vol, vec_fnorm, upts, rhsb, dub, tdiag, tfspr, true);
}

UCFD_API int rans_parallel_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                                       int *icolor, int *lcolor, double *fnorm_vol,
                                       double *vec_fnorm, double *upts, double *rhsb,
                                       double *dub, double *tdiag, double *tfspr)
{
    return do_colored("rans_parallel_upper_sweep", 1, n0, ne, neles, nface, nei_ele, icolor,
                      lcolor, fnorm_vol, vec_fnorm, upts, dub, rhsb, tdiag, tfspr, true);
}

UCFD_API int serial_rans_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *dub, double *tdiag,
                                     double *tfspr)
{
    return do_serial("serial_rans_lower_sweep", 0, neles, nface, nei_ele, fnorm_vol, vec_fnorm,
                     upts, dub, dub, tdiag, tfspr, true);
}

UCFD_API int serial_rans_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *dub, double *tdiag,
                                     double *tfspr)
{
    return do_serial("serial_rans_upper_sweep", 1, neles, nface, nei_ele, fnorm_vol, vec_fnorm,
                     upts, dub, dub, tdiag, tfspr, true);
}

UCFD_API int rans_serial_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                     double *tdiag, double *tfspr)
{
    return do_serial("rans_serial_lower_sweep", 0, neles, nface, nei_ele, fnorm_vol, vec_fnorm,
                     upts, rhsb, dub, tdiag, tfspr, true);
}

UCFD_API int rans_serial_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                                     double *vec_fnorm, double *upts, double *rhsb, double *dub,
                                     double *tdiag, double *tfspr)
{
    return do_serial("rans_serial_upper_sweep", 1, neles, nface, nei_ele, fnorm_vol, vec_fnorm,
                     upts, dub, rhsb, tdiag, tfspr, true);
}

// ------------------------------------------------------- declared layouts

UCFD_API int ucfd_declare_colored_layout(int neles, const int *icolor, const int *lcolor,
                                         int *ncolors_out)
{
    const char *fn = "ucfd_declare_colored_layout";
    if (neles <= 0 || !icolor || !lcolor) return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments");
    std::vector<int> ic(neles), lc(neles);
    const size_t bytes = (size_t)neles*4;
    const bool dev_i = ucfd_is_device_ptr(icolor), dev_l = ucfd_is_device_ptr(lcolor);
    if (dev_i || dev_l) {
        int rc = ucfd_ensure_device(fn);
        if (rc) return rc;
        if (dev_i && (rc = ucfd_check_hip(hipMemcpy(ic.data(), icolor, bytes, hipMemcpyDeviceToHost),
                                          fn, "hipMemcpy(icolor)")))
            return rc;
        if (dev_l && (rc = ucfd_check_hip(hipMemcpy(lc.data(), lcolor, bytes, hipMemcpyDeviceToHost),
                                          fn, "hipMemcpy(lcolor)")))
            return rc;
    }
    if (!dev_i) std::copy(icolor, icolor + neles, ic.begin());
    if (!dev_l) std::copy(lcolor, lcolor + neles, lc.begin());
    DeclaredLayout l{icolor, lcolor, neles, {0}, {lc[0]}};
    for (int p = 0; p < neles; ++p) {
        if (ic[p] != p)
            return ucfd_report(UCFD_ERR_ARG, fn, "icolor is not the identity permutation");
        if (p > 0 && lc[p] < lc[p - 1])
            return ucfd_report(UCFD_ERR_ARG, fn, "lcolor is not non-decreasing");
        if (p > 0 && lc[p] != lc[p - 1]) {
            l.off.push_back(p);
            l.value.push_back(lc[p]);
        }
    }
    l.off.push_back(neles);
    std::lock_guard<std::mutex> g(g_layout_mu);
    for (DeclaredLayout &o : g_layouts)
        if (o.icolor == icolor && o.lcolor == lcolor) { o = l; goto done; }
    g_layouts.push_back(l);
done:
    if (ncolors_out) *ncolors_out = (int)l.off.size() - 1;
    return UCFD_OK;
}

UCFD_API int ucfd_release_colored_layout(const int *icolor, const int *lcolor)
{
    std::lock_guard<std::mutex> g(g_layout_mu);
    for (size_t k = 0; k < g_layouts.size(); ++k)
        if (g_layouts[k].icolor == icolor && g_layouts[k].lcolor == lcolor) {
            g_layouts.erase(g_layouts.begin() + k);
            return UCFD_OK;
        }
    return UCFD_OK;
}
