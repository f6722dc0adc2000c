// FlowSys API with the block LU-SGS method on the GPU
// (src/lusgs/include/ucfdlusgs.h, src/lusgs/flowsys.c, src/lusgs/blusgs.c).
//
// The system keeps the caller's structure, geometry and element arrays (as
// the reference does, flowsys.c:18-54, 120-140); the working arrays rhs, du,
// dup, diag and tdiag are library-owned device memory (the reference mallocs
// them, flowsys.c:91-117).  At the first Prepare the face-slot structure is
// validated, copied to the device, and the sweep schedule is built: the
// reference visits cells one at a time and updates du in place, reading every
// face neighbour (new values for cells visited before, old values for the
// others); cells are grouped into dependency levels (schedule.h) so that each
// level is one kernel launch and the result equals the sequential sweep bit
// for bit.  On a colour-ordered numbering each colour is one level.  Prepare
// also packs every position's neighbour block and fv next to each other in
// position order: the sweeps then stream them instead of gathering 25
// separate face-major arrays through the slot index.  A changed jmat
// therefore needs Prepare again (the reference's diag already does).
#include "flowsys_internal.h"
#include "schedule.h"
#include "ucfdlusgs.h"

#include <new>
#include <string.h>
#include <vector>

namespace {

struct FsElem {
    int neles = 0, nface = 0;
    UCFDInt *cell_ids = nullptr;
    UCFDReal *uptsb = nullptr, *rhs = nullptr, *dt = nullptr, *dsrc = nullptr;
    UCFDReal *vol = nullptr, *resid_out = nullptr;
};

struct Packed {               // per-position blocks of one variable group
    double *jp = nullptr, *fv = nullptr;
    float *jp32 = nullptr, *fv32 = nullptr;
    bool ready = false;
};

enum { FS_NONE = 0, FS_LUSGS = 1, FS_BLUSGS = 2 };

}  // namespace

struct _FlowSys {
    int nlocal, nvars, nfvars, nturbvars, ndims, nfaces, nelem;
    UCFDInt *rowptr, *colidx, *slots;
    UCFDInt8 *sides;
    UCFDReal *face_area, *face_normal, *rcp_vol;
    std::vector<FsElem> eles;
    int method = FS_NONE;
    UCFDReal *jmat = nullptr, *tjmat = nullptr;
    // device state
    double *rhs = nullptr, *du = nullptr, *dup = nullptr, *diag = nullptr, *tdiag = nullptr;
    float *du32 = nullptr;
    int *rowptr_d = nullptr, *colidx_d = nullptr;
    // (position, its cell) pairs in ascending face slot (stable); slot s's
    // are pairs slot_ptr_d[s] .. slot_ptr_d[s+1]-1
    int *pos_sorted_d = nullptr;
    int *slot_ptr_d = nullptr;
    int *pos_partner_d = nullptr;   // the other position of each position's face slot, or -1
    std::vector<int> rowptr_h;      // host copy of the structure's row pointer
    // per block size nv: the most positions 64/nv consecutive rows of each
    // level hold (lower, upper schedules), computed at the first sweep
    std::vector<int> wave_pos[2][8];
    // scalar LU-SGS: rank arrays [nvars, nlocal] and the caller's per-slot
    // spectral radii (src/lusgs/flowsys.h:20-24)
    double *lu_u = nullptr, *lu_du = nullptr, *lu_diag = nullptr;
    UCFDReal *fspr = nullptr, *tfspr = nullptr;
    // per-position face records written by the last NS / RANS Prepare
    // (geometry, spectral radius), read by the sweeps that follow
    double *lu_geo = nullptr, *lu_lam = nullptr, *lu_tlam = nullptr;
    int *lu_inv = nullptr;          // Pack's inverse of cell_ids [nlocal] (both methods)
    double *lu_rec = nullptr;       // Pack / Update records [nlocal][fs_rec_words(nvars)] (both)
    bool lu_geo_ok = false, lu_lam_ok = false, lu_tlam_ok = false;
    int64_t npos = 0;
    bool fp32 = false;
    bool structured = false;
    Packed ns, rans;
    UcfdSchedule lower, upper;
};

namespace {

ucfd_status_t fail(const char *fn, const char *msg, int code = UCFD_ERR_ARG)
{
    ucfd_report(code, fn, msg);
    return UCFD_FAILED;
}

ucfd_status_t status(int rc) { return rc == UCFD_OK ? UCFD_SUCCESS : UCFD_FAILED; }

int hip(hipError_t e, const char *fn, const char *what) { return ucfd_check_hip(e, fn, what); }

template <typename T> int dalloc(T **p, size_t n, const char *fn)
{
    if (*p) return UCFD_OK;
    return hip(hipMalloc((void **)p, (n ? n : 1)*sizeof(T)), fn, "hipMalloc");
}

template <typename T> void dfree(T **p)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}

void free_packed(Packed &P)
{
    dfree(&P.jp);
    dfree(&P.fv);
    dfree(&P.jp32);
    dfree(&P.fv32);
    P.ready = false;
}

void free_device(FlowSys s)
{
    (void)hipStreamSynchronize(ucfd_stream());
    dfree(&s->rhs);
    dfree(&s->du);
    dfree(&s->dup);
    dfree(&s->diag);
    dfree(&s->tdiag);
    dfree(&s->du32);
    dfree(&s->rowptr_d);
    dfree(&s->colidx_d);
    dfree(&s->pos_sorted_d);
    dfree(&s->slot_ptr_d);
    dfree(&s->pos_partner_d);
    dfree(&s->lu_u);
    dfree(&s->lu_du);
    dfree(&s->lu_diag);
    dfree(&s->lu_geo);
    dfree(&s->lu_inv);
    dfree(&s->lu_rec);
    dfree(&s->lu_lam);
    dfree(&s->lu_tlam);
    s->lu_geo_ok = s->lu_lam_ok = s->lu_tlam_ok = false;
    free_packed(s->ns);
    free_packed(s->rans);
    ucfd_free_schedule(s->lower);
    ucfd_free_schedule(s->upper);
    s->structured = false;
}

bool is_blusgs(FlowSys s) { return s && s->method == FS_BLUSGS; }
bool is_lusgs(FlowSys s) { return s && s->method == FS_LUSGS; }

// Validate the face-slot structure, copy it to the device and build the
// forward (cells ascending) and backward (descending) schedules.
int build_structure(FlowSys s, const char *fn)
{
    if (s->structured) return UCFD_OK;
    const int n = s->nlocal;
    std::vector<int> rp, ci, sl;
    int rc;
    if (!s->rowptr || !s->colidx || !s->slots || !s->sides || !s->face_area || !s->rcp_vol)
        return ucfd_report(UCFD_ERR_ARG, fn, "NULL structure or geometry array");
    if ((rc = ucfd_fetch_ints(s->rowptr, (size_t)n + 1, rp, fn))) return rc;
    if (rp[0] != 0) return ucfd_report(UCFD_ERR_INDEX, fn, "rowptr[0] must be 0");
    for (int i = 0; i < n; ++i)
        if (rp[i + 1] < rp[i]) return ucfd_report(UCFD_ERR_INDEX, fn, "rowptr must be nondecreasing");
    s->npos = rp[n];
    s->rowptr_h = rp;
    if ((rc = ucfd_fetch_ints(s->colidx, (size_t)s->npos, ci, fn))) return rc;
    if ((rc = ucfd_fetch_ints(s->slots, (size_t)s->npos, sl, fn))) return rc;
    for (int64_t p = 0; p < s->npos; ++p) {
        if (ci[p] < 0 || ci[p] >= n)
            return ucfd_report(UCFD_ERR_INDEX, fn, "a face neighbour lies outside [0, nlocal)");
        if (sl[p] < 0 || sl[p] >= s->nfaces)
            return ucfd_report(UCFD_ERR_INDEX, fn, "a face slot lies outside [0, nfaces)");
    }
    if ((rc = dalloc(&s->rowptr_d, (size_t)n + 1, fn)) || (rc = dalloc(&s->colidx_d, (size_t)s->npos, fn)))
        return rc;
    if ((rc = hip(hipMemcpy(s->rowptr_d, rp.data(), rp.size()*sizeof(int), hipMemcpyHostToDevice),
                  fn, "hipMemcpy(rowptr)")) ||
        (rc = hip(hipMemcpy(s->colidx_d, ci.data(), ci.size()*sizeof(int), hipMemcpyHostToDevice),
                  fn, "hipMemcpy(colidx)")))
        return rc;
    // positions in ascending face slot (stable) with each slot's range, for
    // Prepare's coalesced pass over the face-major Jacobians, and each
    // position's partner: the other position of a slot held by exactly two
    {
        // sorted[2k], sorted[2k+1]: the k-th position in slot order and its
        // cell, side by side so the counting sort scatters one line per position
        std::vector<int> cnt((size_t)s->nfaces + 1, 0), sorted(2*(size_t)s->npos),
            partner((size_t)s->npos, -1);
        for (int64_t p = 0; p < s->npos; ++p) cnt[(size_t)sl[p] + 1]++;
        for (int f = 0; f < s->nfaces; ++f) cnt[(size_t)f + 1] += cnt[f];
        // rows cover positions 0 .. npos-1 in order (rowptr[0] = 0, nondecreasing)
        for (int i = 0; i < n; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int k = cnt[sl[p]]++;
                sorted[2*(size_t)k] = p;
                sorted[2*(size_t)k + 1] = i;
            }
        // cnt[f] is now the end of slot f's range
        for (int f = 0; f < s->nfaces; ++f) {
            const int e = cnt[f], b = f ? cnt[f - 1] : 0;
            if (e - b == 2) {
                const int pa = sorted[2*(size_t)b], pb = sorted[2*(size_t)b + 2];
                partner[(size_t)pa] = pb;
                partner[(size_t)pb] = pa;
            }
        }
        // slot ranges: cnt[f] ends slot f's range, so it begins slot f + 1's
        for (int64_t f = s->nfaces; f > 0; --f) cnt[(size_t)f] = cnt[(size_t)f - 1];
        cnt[0] = 0;
        if ((rc = dalloc(&s->pos_sorted_d, 2*(size_t)s->npos, fn)) ||
            (rc = dalloc(&s->slot_ptr_d, (size_t)s->nfaces + 1, fn)) ||
            (rc = dalloc(&s->pos_partner_d, (size_t)s->npos, fn)))
            return rc;
        if (s->npos &&
            ((rc = hip(hipMemcpy(s->pos_sorted_d, sorted.data(), sorted.size()*sizeof(int),
                                 hipMemcpyHostToDevice), fn, "hipMemcpy(pos_sorted)")) ||
             (rc = hip(hipMemcpy(s->slot_ptr_d, cnt.data(), cnt.size()*sizeof(int),
                                 hipMemcpyHostToDevice), fn, "hipMemcpy(slot_ptr)")) ||
             (rc = hip(hipMemcpy(s->pos_partner_d, partner.data(), partner.size()*sizeof(int),
                                 hipMemcpyHostToDevice), fn, "hipMemcpy(pos_partner)"))))
            return rc;
    }
    std::vector<int> seq(n), jb(n), je(n);
    for (int i = 0; i < n; ++i) {
        jb[i] = rp[i];
        je[i] = rp[i + 1];
    }
    std::vector<std::vector<int>> levels;
    bool unvisited;
    for (int i = 0; i < n; ++i) seq[i] = i;
    ucfd_build_levels(n, seq, jb, je, ci, levels, &unvisited);
    if ((rc = ucfd_finish_schedule(levels, s->lower, fn))) return rc;
    for (int i = 0; i < n; ++i) seq[i] = n - 1 - i;
    ucfd_build_levels(n, seq, jb, je, ci, levels, &unvisited);
    if ((rc = ucfd_finish_schedule(levels, s->upper, fn))) return rc;
    s->structured = true;
    return UCFD_OK;
}

FsElem *elem(FlowSys s, UCFDInt eidx, const char *fn)
{
    if (!s || eidx < 0 || eidx >= (int)s->eles.size()) {
        ucfd_report(UCFD_ERR_ARG, fn, "element index out of range");
        return nullptr;
    }
    FsElem *e = &s->eles[eidx];
    if (!e->cell_ids) {
        ucfd_report(UCFD_ERR_STATE, fn, "element arrays not set (UCFDFlowSysSetElement)");
        return nullptr;
    }
    return e;
}

// NS (turb = false) or turbulence (turb = true) diagonal assembly + packing
ucfd_status_t prepare(FlowSys s, bool turb, const char *fn)
{
    if (!is_blusgs(s)) return fail(fn, "block LU-SGS is not set (UCFDFlowSysSetBLUSGS)", UCFD_ERR_STATE);
    const int nv = turb ? s->nturbvars : s->nfvars;
    UCFDReal *jm = turb ? s->tjmat : s->jmat;
    double *dg = turb ? s->tdiag : s->diag;
    Packed &P = turb ? s->rans : s->ns;
    if (nv < 1 || nv > 7) return fail(fn, "block size must lie in 1 .. 7");
    if (!jm || !dg) return fail(fn, "Jacobian matrix is not set", UCFD_ERR_STATE);
    int rc = build_structure(s, fn);
    if (rc) return UCFD_FAILED;
    const size_t d2 = (size_t)nv*nv;
    if ((rc = dalloc(&P.jp, (size_t)s->npos*d2, fn)) || (rc = dalloc(&P.fv, (size_t)s->npos, fn)))
        return UCFD_FAILED;
    if (s->fp32 && ((rc = dalloc(&P.jp32, (size_t)s->npos*d2, fn)) ||
                    (rc = dalloc(&P.fv32, (size_t)s->npos, fn))))
        return UCFD_FAILED;
    UcfdStage sg(fn);
    const int isl = sg.add(s->slots, sizeof(int)*(size_t)s->npos, UCFD_IN);
    const int isd = sg.add(s->sides, (size_t)s->npos, UCFD_IN);
    const int ifa = sg.add(s->face_area, sizeof(double)*(size_t)s->nfaces, UCFD_IN);
    const int irv = sg.add(s->rcp_vol, sizeof(double)*(size_t)s->nlocal, UCFD_IN);
    const int ijm = sg.add(jm, sizeof(double)*2*d2*(size_t)s->nfaces, UCFD_IN);
    if ((rc = sg.enter())) return UCFD_FAILED;
    rc = fs_launch_prepare(ucfd_stream(), nv, s->nlocal, s->nfaces, s->npos, s->rowptr_d,
                           s->slot_ptr_d, (const int2 *)s->pos_sorted_d, s->pos_partner_d,
                           (const int *)sg.dev(isl), (const signed char *)sg.dev(isd),
                           (const double *)sg.dev(ifa), (const double *)sg.dev(irv), dg,
                           (const double *)sg.dev(ijm), P.jp, P.fv, s->fp32 ? P.jp32 : nullptr,
                           s->fp32 ? P.fv32 : nullptr, ucfd_device_flag());
    // the factorisation is complete and checked for zero pivots on return
    if (!rc) rc = ucfd_sync_and_check(fn);
    rc = sg.leave(rc);
    P.ready = rc == UCFD_OK;
    return status(rc);
}

ucfd_status_t sweep(FlowSys s, bool turb, bool upper, const char *fn)
{
    if (!is_blusgs(s)) return fail(fn, "block LU-SGS is not set (UCFDFlowSysSetBLUSGS)", UCFD_ERR_STATE);
    const Packed &P = turb ? s->rans : s->ns;
    if (!P.ready) return fail(fn, "Prepare must be called before the sweeps", UCFD_ERR_STATE);
    const int nv = turb ? s->nturbvars : s->nfvars, var0 = turb ? s->nfvars : 0;
    const double *dg = turb ? s->tdiag : s->diag;
    const bool f32 = s->fp32 && P.jp32;
    UcfdStage sg(fn);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    const UcfdSchedule &S = upper ? s->upper : s->lower;
    std::vector<int> &wp = s->wave_pos[upper ? 1 : 0][nv];
    if (wp.size() != S.lv.size()) {
        // staging sizes of the wave kernels: 64/nv consecutive rows per wave
        const int G = 64/nv;
        const std::vector<int> &rp = s->rowptr_h;
        wp.assign(S.lv.size(), -1);
        for (size_t k = 0; k < S.lv.size(); ++k) {
            const UcfdLevel &l = S.lv[k];
            if (l.ordered) continue;
            int mx = 0;
            for (int w0 = 0; w0 < l.cnt; w0 += G) {
                const int a = l.base + w0, b = l.base + std::min(w0 + G, l.cnt);
                mx = std::max(mx, rp[b] - rp[a]);
            }
            wp[k] = mx;
        }
    }
    for (size_t k = 0; k < S.lv.size(); ++k) {
        const UcfdLevel &l = S.lv[k];
        if ((rc = fs_launch_sweep(ucfd_stream(), nv, f32, l.base, l.cnt, l.ordered ? S.order : nullptr,
                                  var0, s->nvars, s->rowptr_d, s->colidx_d, P.jp, P.fv, P.jp32,
                                  P.fv32, s->rhs, dg, s->du, f32 ? s->du32 : nullptr, wp[k])))
            break;
    }
    return status(sg.leave(rc));
}

}  // namespace

// ================================================================ common

UCFD_API ucfd_status_t UCFDFlowSysCreate(FlowSys *sys, UCFDInt nlocal, UCFDInt nvars,
                                         UCFDInt nfvars, UCFDInt ndims, UCFDInt nfaces,
                                         UCFDInt nelem, UCFDInt *rowptr, UCFDInt *colidx,
                                         UCFDInt *slots, UCFDInt8 *sides, UCFDReal *face_area,
                                         UCFDReal *face_normal, UCFDReal *rcp_vol)
{
    const char *fn = "UCFDFlowSysCreate";
    if (!sys) return fail(fn, "NULL system handle");
    if (nlocal < 0 || nvars < 1 || nvars > 16 || nfvars < 1 || nfvars > nvars || nfaces < 0 ||
        nelem < 0)
        return fail(fn, "need nlocal >= 0, 1 <= nfvars <= nvars <= 16, nfaces >= 0, nelem >= 0");
    FlowSys s = new (std::nothrow) _FlowSys();
    if (!s) return fail(fn, "Flow system allocation failed", UCFD_ERR_ALLOC);
    s->nlocal = nlocal;
    s->nvars = nvars;
    s->nfvars = nfvars;
    s->nturbvars = nvars - nfvars;
    s->ndims = ndims;
    s->nfaces = nfaces;
    s->nelem = nelem;
    s->rowptr = rowptr;
    s->colidx = colidx;
    s->slots = slots;
    s->sides = sides;
    s->face_area = face_area;
    s->face_normal = face_normal;
    s->rcp_vol = rcp_vol;
    s->eles.resize(nelem);
    *sys = s;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDFlowSysSetElement(FlowSys *sys, UCFDInt eidx, UCFDInt neles,
                                             UCFDInt nface, UCFDInt *cell_ids, UCFDReal *uptsb,
                                             UCFDReal *rhs, UCFDReal *dt, UCFDReal *dsrc,
                                             UCFDReal *vol, UCFDReal *resid_out)
{
    const char *fn = "UCFDFlowSysSetElement";
    if (!sys || !*sys) return fail(fn, "NULL system handle");
    FlowSys s = *sys;
    if (eidx < 0 || eidx >= (int)s->eles.size()) return fail(fn, "element index out of range");
    if (neles < 0) return fail(fn, "need neles >= 0");
    FsElem &e = s->eles[eidx];
    e.neles = neles;
    e.nface = nface;
    e.cell_ids = cell_ids;
    e.uptsb = uptsb;
    e.rhs = rhs;
    e.dt = dt;
    e.dsrc = dsrc;
    e.vol = vol;
    e.resid_out = resid_out;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDFlowSysDestroy(FlowSys *sys)
{
    if (!sys || !*sys) return UCFD_SUCCESS;
    free_device(*sys);
    delete *sys;
    *sys = nullptr;
    return UCFD_SUCCESS;
}

// ============================================================ block LU-SGS

UCFD_API ucfd_status_t UCFDFlowSysSetBLUSGS(FlowSys *sys, UCFDReal *jmat, UCFDReal *tjmat)
{
    const char *fn = "UCFDFlowSysSetBLUSGS";
    if (!sys || !*sys) return fail(fn, "NULL system handle");
    FlowSys s = *sys;
    int rc = ucfd_ensure_device(fn);
    if (rc) return UCFD_FAILED;
    free_device(s);
    s->method = FS_BLUSGS;
    s->jmat = jmat;
    s->tjmat = tjmat;
    const size_t nv = (size_t)s->nlocal*s->nvars;
    if ((rc = dalloc(&s->rhs, nv, fn)) || (rc = dalloc(&s->du, nv, fn)) ||
        (rc = dalloc(&s->dup, nv, fn)) ||
        (rc = dalloc(&s->diag, (size_t)s->nlocal*s->nfvars*s->nfvars, fn)) ||
        (s->nturbvars && (rc = dalloc(&s->tdiag, (size_t)s->nlocal*s->nturbvars*s->nturbvars, fn))))
        return UCFD_FAILED;
    // the reference's buffers come from malloc; start from zeros here
    if ((rc = fs_launch_zero(ucfd_stream(), (int64_t)nv, s->du, s->dup, nullptr)) ||
        (rc = fs_launch_zero(ucfd_stream(), (int64_t)nv, s->rhs, s->rhs, nullptr)) ||
        (rc = ucfd_sync_and_check(fn)))
        return UCFD_FAILED;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDBLUSGS_SetPrecision(FlowSys sys, UCFDInt fp32)
{
    const char *fn = "UCFDBLUSGS_SetPrecision";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    const bool want = fp32 != 0;
    if (want == sys->fp32) return UCFD_SUCCESS;
    sys->fp32 = want;
    sys->ns.ready = sys->rans.ready = false;     // Prepare again to pack FP32 blocks
    if (!want) {
        dfree(&sys->ns.jp32);
        dfree(&sys->ns.fv32);
        dfree(&sys->rans.jp32);
        dfree(&sys->rans.fv32);
        dfree(&sys->du32);
        return UCFD_SUCCESS;
    }
    const size_t nv = (size_t)sys->nlocal*sys->nvars;
    int rc = dalloc(&sys->du32, nv, fn);
    if (rc) return UCFD_FAILED;
    // the FP32 shadow of du starts from the current du
    std::vector<double> h(nv);
    if (nv && ((rc = hip(hipMemcpy(h.data(), sys->du, nv*sizeof(double), hipMemcpyDeviceToHost), fn,
                         "hipMemcpy(du)"))))
        return UCFD_FAILED;
    std::vector<float> f(h.begin(), h.end());
    if (nv && (rc = hip(hipMemcpy(sys->du32, f.data(), nv*sizeof(float), hipMemcpyHostToDevice), fn,
                        "hipMemcpy(du32)")))
        return UCFD_FAILED;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDBLUSGS_Pack(FlowSys sys, UCFDInt eidx, UCFDReal a0)
{
    const char *fn = "UCFDBLUSGS_Pack";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    UcfdStage sg(fn);
    const size_t ne = (size_t)e->neles;
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int ir = sg.add(e->rhs, sizeof(double)*ne*sys->nvars, UCFD_IN);
    const int idt = sg.add(e->dt, sizeof(double)*ne, UCFD_IN);
    if (dalloc(&sys->lu_inv, (size_t)sys->nlocal, fn) ||
        dalloc(&sys->lu_rec, (size_t)sys->nlocal*fs_rec_words(sys->nvars), fn))
        return UCFD_FAILED;
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_pack(ucfd_stream(), e->neles, sys->nlocal, sys->nvars, sys->nfvars, a0,
                        (const int *)sg.dev(ic), (const double *)sg.dev(ir),
                        (const double *)sg.dev(idt), sys->rhs, sys->diag, ucfd_device_flag(),
                        sys->lu_inv, sys->lu_rec);
    return status(sg.leave(rc));
}

static ucfd_status_t tpack(FlowSys sys, UCFDInt eidx, UCFDReal factor, UCFDReal a0, int kind,
                           const char *fn)
{
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    if (sys->nturbvars < (kind == 2 ? 2 : 1))
        return fail(fn, "the system has too few turbulence variables", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    UcfdStage sg(fn);
    const size_t ne = (size_t)e->neles;
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int iu = sg.add(e->uptsb, sizeof(double)*ne*sys->nvars, UCFD_IN);
    const int isrc = sg.add(e->dsrc, sizeof(double)*ne*sys->nvars, UCFD_IN);
    const int idt = sg.add(e->dt, sizeof(double)*ne, UCFD_IN);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_tpack(ucfd_stream(), e->neles, sys->nlocal, sys->nvars, sys->nfvars, a0, factor,
                         kind, (const int *)sg.dev(ic), (const double *)sg.dev(iu),
                         (const double *)sg.dev(isrc), (const double *)sg.dev(idt), sys->tdiag,
                         ucfd_device_flag());
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDBLUSGS_SA_Pack(FlowSys sys, UCFDInt eidx, UCFDReal turb_factor,
                                          UCFDReal a0)
{
    return tpack(sys, eidx, turb_factor, a0, 1, "UCFDBLUSGS_SA_Pack");
}

UCFD_API ucfd_status_t UCFDBLUSGS_KWSST_Pack(FlowSys sys, UCFDInt eidx, UCFDReal turb_factor,
                                             UCFDReal a0)
{
    return tpack(sys, eidx, turb_factor, a0, 2, "UCFDBLUSGS_KWSST_Pack");
}

UCFD_API ucfd_status_t UCFDBLUSGS_Update(FlowSys sys, UCFDInt eidx)
{
    const char *fn = "UCFDBLUSGS_Update";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    UcfdStage sg(fn);
    const size_t ne = (size_t)e->neles;
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int iu = sg.add(e->uptsb, sizeof(double)*ne*sys->nvars, UCFD_INOUT);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_update(ucfd_stream(), e->neles, sys->nlocal, sys->nvars, (const int *)sg.dev(ic),
                          sys->du, (double *)sg.dev(iu), ucfd_device_flag());
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDBLUSGS_SubResidual(FlowSys sys, UCFDInt eidx)
{
    const char *fn = "UCFDBLUSGS_SubResidual";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    UcfdStage sg(fn);
    const size_t ne = (size_t)e->neles;
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int iv = sg.add(e->vol, sizeof(double)*ne, UCFD_IN);
    const int io = sg.add(e->resid_out, sizeof(double)*ne*sys->nvars, UCFD_OUT);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_subres(ucfd_stream(), e->neles, sys->nlocal, sys->nvars, (const int *)sg.dev(ic),
                          (const double *)sg.dev(iv), sys->du, sys->dup, (double *)sg.dev(io),
                          ucfd_device_flag());
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDBLUSGS_NSPrepare(FlowSys sys)
{
    return prepare(sys, false, "UCFDBLUSGS_NSPrepare");
}

UCFD_API ucfd_status_t UCFDBLUSGS_RANSPrepare(FlowSys sys)
{
    return prepare(sys, true, "UCFDBLUSGS_RANSPrepare");
}

UCFD_API ucfd_status_t UCFDBLUSGS_NSLowerSweep(FlowSys sys)
{
    return sweep(sys, false, false, "UCFDBLUSGS_NSLowerSweep");
}

UCFD_API ucfd_status_t UCFDBLUSGS_RANSLowerSweep(FlowSys sys)
{
    return sweep(sys, true, false, "UCFDBLUSGS_RANSLowerSweep");
}

UCFD_API ucfd_status_t UCFDBLUSGS_NSUpperSweep(FlowSys sys)
{
    return sweep(sys, false, true, "UCFDBLUSGS_NSUpperSweep");
}

UCFD_API ucfd_status_t UCFDBLUSGS_RANSUpperSweep(FlowSys sys)
{
    return sweep(sys, true, true, "UCFDBLUSGS_RANSUpperSweep");
}

UCFD_API ucfd_status_t UCFDBLUSGS_Reset(FlowSys sys)
{
    const char *fn = "UCFDBLUSGS_Reset";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    UcfdStage sg(fn);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_zero(ucfd_stream(), (int64_t)sys->nlocal*sys->nvars, sys->du, sys->dup,
                        sys->fp32 ? sys->du32 : nullptr);
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDBLUSGS_Export(FlowSys sys, UCFDInt which, UCFDReal *out)
{
    const char *fn = "UCFDBLUSGS_Export";
    if (!is_blusgs(sys)) return fail(fn, "block LU-SGS is not set", UCFD_ERR_STATE);
    const size_t nv = (size_t)sys->nlocal*sys->nvars;
    const double *src = nullptr;
    size_t n = 0;
    switch (which) {
    case 0: src = sys->rhs; n = nv; break;
    case 1: src = sys->du; n = nv; break;
    case 2: src = sys->dup; n = nv; break;
    case 3: src = sys->diag; n = (size_t)sys->nlocal*sys->nfvars*sys->nfvars; break;
    case 4: src = sys->tdiag; n = (size_t)sys->nlocal*sys->nturbvars*sys->nturbvars; break;
    default: return fail(fn, "which must be 0 (rhs), 1 (du), 2 (dup), 3 (diag) or 4 (tdiag)");
    }
    if (!src) return fail(fn, "the array is not allocated", UCFD_ERR_STATE);
    UcfdStage sg(fn);
    const int io = sg.add(out, sizeof(double)*n, UCFD_OUT);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = hip(hipMemcpyAsync(sg.dev(io), src, sizeof(double)*n, hipMemcpyDeviceToDevice,
                            ucfd_stream()), fn, "hipMemcpyAsync(export)");
    return status(sg.leave(rc));
}

// ======================================= scalar LU-SGS methods of the FlowSys
//
// src/lusgs/lusgs.c on the GPU: Pack / Update per element, Prepare and the
// sweeps on the rank arrays u / du / diag [nvars, nlocal].  The sweeps are
// the reference's sequential in-place Gauss-Seidel visits (rows ascending /
// descending, lower / upper neighbours only), run level by level on the
// schedules of the face-slot structure (one launch per colour on a
// colour-ordered numbering): results equal the reference bit for bit.
// fspr / tfspr (per face slot) are the caller's arrays.  Each Prepare also
// writes per-position face records (side*normal, face_area*rcp_vol and the
// spectral radius of its rows) that the sweeps read in position order: the
// sweeps use the spectral radius their rows' last Prepare saw and the
// geometry of the last NS Prepare (the reference reads them again in the
// sweeps; a caller that changes them calls Prepare again, as it must for the
// diagonal anyway).  Sweeps with no Prepare before them read the caller's
// arrays directly.

UCFD_API ucfd_status_t UCFDFlowSysSetLUSGS(FlowSys *sys, UCFDReal *fspr, UCFDReal *tfspr)
{
    const char *fn = "UCFDFlowSysSetLUSGS";
    if (!sys || !*sys) return fail(fn, "NULL system handle");
    FlowSys s = *sys;
    int rc = ucfd_ensure_device(fn);
    if (rc) return UCFD_FAILED;
    free_device(s);
    s->method = FS_LUSGS;
    s->fspr = fspr;
    s->tfspr = tfspr;
    const size_t nv = (size_t)s->nlocal*s->nvars;
    if ((rc = dalloc(&s->lu_u, nv, fn)) || (rc = dalloc(&s->lu_du, nv, fn)) ||
        (rc = dalloc(&s->lu_diag, nv, fn)))
        return UCFD_FAILED;
    // the reference's buffers come from malloc; start from zeros here
    if ((rc = fs_launch_zero(ucfd_stream(), (int64_t)nv, s->lu_u, s->lu_du, nullptr)) ||
        (rc = fs_launch_zero(ucfd_stream(), (int64_t)nv, s->lu_diag, s->lu_diag, nullptr)) ||
        (rc = ucfd_sync_and_check(fn)))
        return UCFD_FAILED;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDLUSGS_Pack(FlowSys sys, UCFDInt eidx, UCFDReal turb_factor, UCFDReal a0)
{
    const char *fn = "UCFDLUSGS_Pack";
    if (!is_lusgs(sys)) return fail(fn, "LU-SGS is not set (UCFDFlowSysSetLUSGS)", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    const size_t ne = (size_t)e->neles, nv = (size_t)sys->nvars;
    UcfdStage sg(fn);
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int iu = sg.add(e->uptsb, sizeof(double)*ne*nv, UCFD_IN);
    const int ir = sg.add(e->rhs, sizeof(double)*ne*nv, UCFD_IN);
    const int it = sg.add(e->dt, sizeof(double)*ne, UCFD_IN);
    const int isr = sg.add(e->dsrc, sizeof(double)*ne*nv, UCFD_IN);
    if (dalloc(&sys->lu_inv, (size_t)sys->nlocal, fn) ||
        dalloc(&sys->lu_rec, (size_t)sys->nlocal*fs_rec_words(sys->nvars), fn))
        return UCFD_FAILED;
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_lpack(ucfd_stream(), e->neles, sys->nlocal, sys->nvars, sys->nfvars, turb_factor,
                         a0, (const int *)sg.dev(ic), (const double *)sg.dev(iu),
                         (const double *)sg.dev(ir), (const double *)sg.dev(it),
                         (const double *)sg.dev(isr), sys->lu_u, sys->lu_du, sys->lu_diag,
                         ucfd_device_flag(), sys->lu_inv, sys->lu_rec);
    if (!rc) rc = ucfd_sync_and_check(fn);
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDLUSGS_Update(FlowSys sys, UCFDInt eidx)
{
    const char *fn = "UCFDLUSGS_Update";
    if (!is_lusgs(sys)) return fail(fn, "LU-SGS is not set (UCFDFlowSysSetLUSGS)", UCFD_ERR_STATE);
    FsElem *e = elem(sys, eidx, fn);
    if (!e) return UCFD_FAILED;
    const size_t ne = (size_t)e->neles, nv = (size_t)sys->nvars;
    UcfdStage sg(fn);
    const int ic = sg.add(e->cell_ids, sizeof(int)*ne, UCFD_IN);
    const int iu = sg.add(e->uptsb, sizeof(double)*ne*nv, UCFD_INOUT);
    if (dalloc(&sys->lu_rec, (size_t)sys->nlocal*fs_rec_words(sys->nvars), fn))
        return UCFD_FAILED;
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = fs_launch_lupdate(ucfd_stream(), e->neles, sys->nlocal, sys->nvars,
                           (const int *)sg.dev(ic), sys->lu_du, (double *)sg.dev(iu),
                           ucfd_device_flag(), sys->lu_rec);
    if (!rc) rc = ucfd_sync_and_check(fn);
    return status(sg.leave(rc));
}

namespace {

// NS rows (turb = false) or turbulence rows: the arguments both Prepare and
// the sweeps check
int lusgs_rows(FlowSys s, bool turb, const char *fn, const UCFDReal **lam)
{
    if (!is_lusgs(s)) return ucfd_report(UCFD_ERR_STATE, fn, "LU-SGS is not set (UCFDFlowSysSetLUSGS)");
    if (turb) {
        if (s->nturbvars < 1 || s->nturbvars > 4)
            return ucfd_report(UCFD_ERR_ARG, fn, "the turbulence rows must number 1 .. 4");
        if (!s->tfspr) return ucfd_report(UCFD_ERR_STATE, fn, "Turbulent spectral radius is not set");
        *lam = s->tfspr;
    } else {
        if (!s->fspr) return ucfd_report(UCFD_ERR_STATE, fn, "spectral radius is not set");
        *lam = s->fspr;
    }
    if ((s->ndims != 2 && s->ndims != 3) || s->nfvars != s->ndims + 2)
        return ucfd_report(UCFD_ERR_ARG, fn, "the flux needs ndims 2 or 3 and nfvars = ndims + 2");
    return build_structure(s, fn);
}

ucfd_status_t lusgs_prepare(FlowSys s, bool turb, double kappa, const char *fn)
{
    const UCFDReal *lam = nullptr;
    if (lusgs_rows(s, turb, fn, &lam)) return UCFD_FAILED;
    UcfdStage sg(fn);
    const int isl = sg.add(s->slots, sizeof(int)*(size_t)s->npos, UCFD_IN);
    const int ifa = sg.add(s->face_area, sizeof(double)*(size_t)s->nfaces, UCFD_IN);
    const int irv = sg.add(s->rcp_vol, sizeof(double)*(size_t)s->nlocal, UCFD_IN);
    const int ilm = sg.add(lam, sizeof(double)*(size_t)s->nfaces, UCFD_IN);
    const int isd = sg.add(s->sides, (size_t)s->npos, UCFD_IN);
    const int ifn = sg.add(s->face_normal, sizeof(double)*(size_t)s->nfaces*s->ndims, UCFD_IN);
    double *&lamp = turb ? s->lu_tlam : s->lu_lam;
    if (dalloc(&s->lu_geo, 4*(size_t)s->npos, fn) || dalloc(&lamp, (size_t)s->npos, fn))
        return UCFD_FAILED;
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    // the geometry records are written by every NS Prepare; a RANS Prepare
    // after one (the reference's order) writes only its spectral radius
    const bool geo = !(turb && s->lu_geo_ok && s->lu_lam_ok);
    rc = fs_launch_lprep_pos(ucfd_stream(), s->ndims, s->nlocal, s->nfaces,
                             turb ? s->nfvars : 0, turb ? s->nvars : s->nfvars, kappa, s->rowptr_d,
                             (const int *)sg.dev(isl), (const signed char *)sg.dev(isd),
                             (const double *)sg.dev(ifa), (const double *)sg.dev(ifn),
                             (const double *)sg.dev(irv), (const double *)sg.dev(ilm), s->lu_diag,
                             geo ? s->lu_geo : nullptr, lamp);
    rc = sg.leave(rc);
    s->lu_geo_ok = rc == UCFD_OK;
    (turb ? s->lu_tlam_ok : s->lu_lam_ok) = rc == UCFD_OK;
    return status(rc);
}

// The first level of a schedule holds the cells with no neighbour visited
// before them (ucfd_build_levels gives a cell the level after its earlier
// neighbours'), so the scalar sweep takes no face there: a kernel that reads
// only du and diag.
int lusgs_first_level(FlowSys s, bool turb, bool upper, const UcfdLevel &l, const UcfdSchedule &S)
{
    return fs_launch_lsweep_free(ucfd_stream(), upper, l.base, l.cnt, l.ordered ? S.order : nullptr,
                                 turb ? s->nfvars : 0, turb ? s->nvars : s->nfvars, s->nlocal,
                                 s->lu_diag, s->lu_du);
}

ucfd_status_t lusgs_sweep(FlowSys s, bool turb, bool upper, double kappa, const char *fn)
{
    const UCFDReal *lam = nullptr;
    if (lusgs_rows(s, turb, fn, &lam)) return UCFD_FAILED;
    const UcfdSchedule &S = upper ? s->upper : s->lower;
    if (s->lu_geo_ok && (turb ? s->lu_tlam_ok : s->lu_lam_ok)) {
        // the face records of the last Prepare (no staging: all device data)
        int rc = UCFD_OK;
        for (const UcfdLevel &l : S.lv)
            if ((rc = &l == &S.lv.front() ? lusgs_first_level(s, turb, upper, l, S) :
                      fs_launch_lsweep_packed(ucfd_stream(), upper, s->ndims,
                                              turb ? s->nturbvars : 0, l.base, l.cnt,
                                              l.ordered ? S.order : nullptr, s->nfvars, kappa,
                                              s->nlocal, s->rowptr_d, s->colidx_d, s->lu_geo,
                                              turb ? s->lu_tlam : s->lu_lam, s->lu_diag, s->lu_u,
                                              s->lu_du)))
                break;
        if (!rc) {
            if (!ucfd_stream_is_user()) rc = ucfd_sync_and_check(fn);
            else rc = ucfd_check_hip(hipGetLastError(), fn, "kernel launch");
        }
        return status(rc);
    }
    UcfdStage sg(fn);
    const int isd = sg.add(s->sides, (size_t)s->npos, UCFD_IN);
    const int isl = sg.add(s->slots, sizeof(int)*(size_t)s->npos, UCFD_IN);
    const int ifa = sg.add(s->face_area, sizeof(double)*(size_t)s->nfaces, UCFD_IN);
    const int ifn = sg.add(s->face_normal, sizeof(double)*(size_t)s->nfaces*s->ndims, UCFD_IN);
    const int irv = sg.add(s->rcp_vol, sizeof(double)*(size_t)s->nlocal, UCFD_IN);
    const int ilm = sg.add(lam, sizeof(double)*(size_t)s->nfaces, UCFD_IN);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    for (const UcfdLevel &l : S.lv)
        if ((rc = &l == &S.lv.front() ? lusgs_first_level(s, turb, upper, l, S) :
                  fs_launch_lsweep(ucfd_stream(), upper, s->ndims, turb ? s->nturbvars : 0, l.base,
                                   l.cnt, l.ordered ? S.order : nullptr, s->nvars, s->nfvars,
                                   kappa, (const double *)sg.dev(ilm), s->nlocal, s->nfaces,
                                   s->rowptr_d, s->colidx_d, (const signed char *)sg.dev(isd),
                                   (const int *)sg.dev(isl), (const double *)sg.dev(ifa),
                                   (const double *)sg.dev(ifn), (const double *)sg.dev(irv),
                                   s->lu_diag, s->lu_u, s->lu_du)))
            break;
    return status(sg.leave(rc));
}

}  // namespace

UCFD_API ucfd_status_t UCFDLUSGS_NSPrepare(FlowSys sys, UCFDReal kappa)
{
    return lusgs_prepare(sys, false, kappa, "UCFDLUSGS_NSPrepare");
}

UCFD_API ucfd_status_t UCFDLUSGS_RANSPrepare(FlowSys sys, UCFDReal kappa)
{
    return lusgs_prepare(sys, true, kappa, "UCFDLUSGS_RANSPrepare");
}

UCFD_API ucfd_status_t UCFDLUSGS_NSLowerSweep(FlowSys sys, UCFDReal kappa)
{
    return lusgs_sweep(sys, false, false, kappa, "UCFDLUSGS_NSLowerSweep");
}

UCFD_API ucfd_status_t UCFDLUSGS_NSUpperSweep(FlowSys sys, UCFDReal kappa)
{
    return lusgs_sweep(sys, false, true, kappa, "UCFDLUSGS_NSUpperSweep");
}

UCFD_API ucfd_status_t UCFDLUSGS_RANSLowerSweep(FlowSys sys, UCFDReal kappa)
{
    return lusgs_sweep(sys, true, false, kappa, "UCFDLUSGS_RANSLowerSweep");
}

UCFD_API ucfd_status_t UCFDLUSGS_RANSUpperSweep(FlowSys sys, UCFDReal kappa)
{
    return lusgs_sweep(sys, true, true, kappa, "UCFDLUSGS_RANSUpperSweep");
}

UCFD_API ucfd_status_t UCFDLUSGS_Export(FlowSys sys, UCFDInt which, UCFDReal *out)
{
    const char *fn = "UCFDLUSGS_Export";
    if (!is_lusgs(sys)) return fail(fn, "LU-SGS is not set", UCFD_ERR_STATE);
    const double *src = which == 0 ? sys->lu_u : which == 1 ? sys->lu_du : which == 2 ? sys->lu_diag
                                                                                    : nullptr;
    if (!src) return fail(fn, "which must be 0 (u), 1 (du) or 2 (diag)");
    const size_t n = (size_t)sys->nlocal*sys->nvars;
    UcfdStage sg(fn);
    const int io = sg.add(out, sizeof(double)*n, UCFD_OUT);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = hip(hipMemcpyAsync(sg.dev(io), src, sizeof(double)*n, hipMemcpyDeviceToDevice,
                            ucfd_stream()), fn, "hipMemcpyAsync(export)");
    return status(sg.leave(rc));
}
