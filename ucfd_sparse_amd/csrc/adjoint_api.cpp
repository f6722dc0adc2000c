// Differentiation form of the coloured LU-SGS pass: the C-ABI of the adjoint
// (include/ucfd_ext.h).  The reference has no adjoint; the forward it
// differentiates is the legacy pre + lower + upper pass on a
// colour-contiguous numbering (the layout ucfd_declare_colored_layout takes).
//
// ucfd_adjoint_create validates the colour ranges (no two cells of one
// colour are face neighbours), copies the colour offsets and builds the
// face-pair map on the device.  ucfd_lusgs_adjoint then takes w = dJ/d dQ
// and the forward's dQ* (lower-sweep output) and dQ, and returns the
// gradients of J with respect to rhsb (rows < NFVARS), dt and fspr.
#include "adjoint_internal.h"
#include "schedule.h"
#include "ucfd_ext.h"

#include <new>
#include <vector>

namespace {

struct Adjoint {
    int neles = 0, nface = 0, ndims = 3;
    std::vector<int> off;
    const int *nei = nullptr;
    const double *fnv = nullptr, *vfn = nullptr, *upts = nullptr;
    int *pair = nullptr;           // device [nface, neles]
    double *diag = nullptr;        // device [neles]
    double *work = nullptr;        // device [nfvars, neles]
};

}  // namespace

UCFD_API int ucfd_adjoint_create(void **handle, int neles, int nface, int ncolors,
                                 const int *color_off, const int *nei_ele, const double *fnorm_vol,
                                 const double *vec_fnorm, const double *upts)
{
    const char *fn = "ucfd_adjoint_create";
    if (!handle || neles <= 0 || nface < 1 || nface > 8 || ncolors <= 0 || !color_off)
        return ucfd_report(UCFD_ERR_ARG, fn, "bad sizes or NULL colour offsets");
    if (color_off[0] != 0 || color_off[ncolors] != neles)
        return ucfd_report(UCFD_ERR_ARG, fn, "color_off must run from 0 to neles");
    for (int c = 0; c < ncolors; ++c)
        if (color_off[c + 1] < color_off[c])
            return ucfd_report(UCFD_ERR_ARG, fn, "color_off must be non-decreasing");
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;
    const UcfdConfig &cfg = ucfd_config();
    std::vector<int> nh;
    if ((rc = ucfd_fetch_ints(nei_ele, (size_t)nface*neles, nh, fn))) return rc;
    for (int c = 0; c < ncolors; ++c)
        for (int i = color_off[c]; i < color_off[c + 1]; ++i)
            for (int f = 0; f < nface; ++f) {
                const int nb = nh[(size_t)f*neles + i];
                if (nb < 0 || nb >= neles)
                    return ucfd_report(UCFD_ERR_INDEX, fn, "nei_ele holds an entry outside [0, neles)");
                if (nb != i && nb >= color_off[c] && nb < color_off[c + 1])
                    return ucfd_report(UCFD_ERR_ARG, fn, "two cells of one colour are face "
                                                         "neighbours: not a colour-contiguous layout");
            }
    Adjoint *A = new (std::nothrow) Adjoint();
    if (!A) return ucfd_report(UCFD_ERR_ALLOC, fn, "handle allocation failed");
    A->neles = neles;
    A->nface = nface;
    A->ndims = cfg.ndims;
    A->off.assign(color_off, color_off + ncolors + 1);
    A->nei = nei_ele;
    A->fnv = fnorm_vol;
    A->vfn = vec_fnorm;
    A->upts = upts;
    const size_t nf = (size_t)cfg.nfvars*neles;
    if ((rc = ucfd_check_hip(hipMalloc((void **)&A->pair, sizeof(int)*(size_t)nface*neles), fn,
                             "hipMalloc(pair)")) ||
        (rc = ucfd_check_hip(hipMalloc((void **)&A->diag, sizeof(double)*neles), fn,
                             "hipMalloc(diag)")) ||
        (rc = ucfd_check_hip(hipMalloc((void **)&A->work, sizeof(double)*nf), fn,
                             "hipMalloc(work)"))) {
        ucfd_adjoint_destroy(A);
        return rc;
    }
    UcfdStage sg(fn);
    const int in = sg.add(nei_ele, sizeof(int)*(size_t)nface*neles, UCFD_IN);
    if ((rc = sg.enter())) {
        ucfd_adjoint_destroy(A);
        return rc;
    }
    rc = adj_launch_pairs(ucfd_stream(), neles, nface, (const int *)sg.dev(in), A->pair,
                          ucfd_device_flag());
    if (!rc) rc = ucfd_sync_and_check(fn);
    rc = sg.leave(rc);
    if (rc) {
        ucfd_adjoint_destroy(A);
        return rc;
    }
    *handle = A;
    return UCFD_OK;
}

UCFD_API int ucfd_adjoint_destroy(void *handle)
{
    Adjoint *A = (Adjoint *)handle;
    if (!A) return UCFD_OK;
    (void)hipStreamSynchronize(ucfd_stream());
    if (A->pair) (void)hipFree(A->pair);
    if (A->diag) (void)hipFree(A->diag);
    if (A->work) (void)hipFree(A->work);
    delete A;
    return UCFD_OK;
}

UCFD_API int ucfd_lusgs_adjoint(void *handle, double factor, const double *dt, const double *fspr,
                                const double *dstar, const double *dq, const double *w,
                                double *rhsbar, double *dtbar, double *fsprbar)
{
    const char *fn = "ucfd_lusgs_adjoint";
    Adjoint *A = (Adjoint *)handle;
    if (!A) return ucfd_report(UCFD_ERR_ARG, fn, "NULL handle");
    if (!(factor > 0.0)) return ucfd_report(UCFD_ERR_ARG, fn, "factor must be positive");
    const UcfdConfig &cfg = ucfd_config();
    if (cfg.ndims != A->ndims || cfg.nfvars != A->ndims + 2)
        return ucfd_report(UCFD_ERR_STATE, fn, "configuration changed since create");
    const size_t N = (size_t)A->neles, F = (size_t)A->nface, NV = (size_t)cfg.nfvars;
    UcfdStage sg(fn);
    const int inei = sg.add(A->nei, sizeof(int)*F*N, UCFD_IN);
    const int ifnv = sg.add(A->fnv, sizeof(double)*F*N, UCFD_IN);
    const int ivfn = sg.add(A->vfn, sizeof(double)*F*A->ndims*N, UCFD_IN);
    const int iup = sg.add(A->upts, sizeof(double)*NV*N, UCFD_IN);
    const int idt = sg.add(dt, sizeof(double)*N, UCFD_IN);
    const int ifs = sg.add(fspr, sizeof(double)*F*N, UCFD_IN);
    const int ids = sg.add(dstar, sizeof(double)*NV*N, UCFD_IN);
    const int idq = sg.add(dq, sizeof(double)*NV*N, UCFD_IN);
    const int iw = sg.add(w, sizeof(double)*NV*N, UCFD_IN);
    const int irb = sg.add(rhsbar, sizeof(double)*NV*N, UCFD_OUT);
    const int idb = sg.add(dtbar, sizeof(double)*N, UCFD_OUT);
    const int ifb = sg.add(fsprbar, sizeof(double)*F*N, UCFD_OUT);
    int rc = sg.enter();
    if (rc) return rc;
    hipStream_t s = ucfd_stream();
    const int *nei = (const int *)sg.dev(inei);
    const double *fnv = (const double *)sg.dev(ifnv), *vfn = (const double *)sg.dev(ivfn);
    const double *fs = (const double *)sg.dev(ifs), *dtd = (const double *)sg.dev(idt);
    rc = ucfd_launch_pre(s, A->neles, A->nface, factor, fnv, dtd, A->diag, fs);
    const int nc = (int)A->off.size() - 1;
    for (int c = 0; c < nc && !rc; ++c)
        rc = adj_launch_colour(s, 0, A->off[c], A->off[c + 1], A->neles, A->nface, A->ndims, nei,
                               A->pair, fnv, vfn, fs, dtd, factor, A->diag,
                               (const double *)sg.dev(iup), (const double *)sg.dev(ids),
                               (const double *)sg.dev(idq), (const double *)sg.dev(iw), A->work,
                               nullptr, nullptr, nullptr);
    for (int c = nc - 1; c >= 0 && !rc; --c)
        rc = adj_launch_colour(s, 1, A->off[c], A->off[c + 1], A->neles, A->nface, A->ndims, nei,
                               A->pair, fnv, vfn, fs, dtd, factor, A->diag,
                               (const double *)sg.dev(iup), (const double *)sg.dev(ids),
                               (const double *)sg.dev(idq), (const double *)sg.dev(iw), A->work,
                               (double *)sg.dev(irb), (double *)sg.dev(idb), (double *)sg.dev(ifb));
    return sg.leave(rc);
}
