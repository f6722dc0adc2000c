// Krylov API: sparse matrices, block LU-SGS and ILU(0)-family
// preconditioners and restarted GMRES / BiCGSTAB, executed on the GPU behind
// the reference's C interface (src/krylov/include/ucfdmat.h, ucfdprecon.h,
// ucfdsolver.h).
//
// Handles keep the caller's array pointers, as the reference does
// (sparsemat.c:32-50, precon.c:13-27): values written between calls are seen
// by the next call.  Each entry point resolves the arrays it reads through
// the runtime's staging (device pointers in place, host arrays copied for the
// duration of the call).  Library-owned device memory holds the factored
// diagonal blocks, the FP32 copy of the blocks (mixed form), the sweep
// schedules and the Krylov work vectors.
//
// Algorithms follow the reference step for step (gmres.c:4-129,
// bicgstab.c:5-131, blusgs.c:7-100, pblusgs.c:7-104, ilu.c:4-86,
// bilu.c:5-130, pbilu.c:5-180); the dot products are
// two-stage device reductions in a fixed order, so results are reproducible
// but not bitwise equal to a sequential CPU sum.
#include "krylov_internal.h"
#include "schedule.h"
#include "ucfdmat.h"
#include "ucfdprecon.h"
#include "ucfdsolver.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

struct _SpMat {
    SpMatType type_name;
    int n;          // scalar rows
    int bn, blk;    // block rows and block size (CSR: n and 1)
    bool csr;
    UCFDInt *rowptr, *colidx;
    UCFDReal *values;
    int64_t nnz;    // stored blocks; -1 until read from rowptr
};

struct _Precon {
    PreconType type_name;
    UCFDInt *rowptr, *colidx, *diagslots;
    UCFDReal *values;
    int bn, blk;
    int ncolors;             // PBLUSGS / PBILU: number of colour ranges
    const UCFDInt *icolors;  // PBLUSGS / PBILU: ncolors+1 offsets (caller's array)
    bool fp32;               // mixed apply: FP32 blocks and work vector, FP64 diagonal LU
    bool prepared;
    bool scheduled;          // structure validated and schedules built (first Prepare)
    bool self_free;          // no lower slot of a row refers to the row itself
    int64_t nnz;
    double *diagvalues;      // device, bn*blk*blk: LU-SGS the factored diagonal
                             // blocks, ILU the factored U_ii (scalar: the pivot)
    float *vals32;           // device, nnz*blk*blk (fp32 only)
    float *x32;              // device, bn*blk (fp32 only)
    bool init32;             // the lower pass reads rows it has not yet produced
    UcfdSchedule lower, upper;
};

namespace {

constexpr int HCAP = 64;      // Krylov dimension limit of the device reductions
constexpr int DB_H = 0, DB_H2 = 64, DB_NSQ = 128, DB_Y = 136, DB_N = 208;

}  // namespace

struct _Solver {
    SolverType type_name;
    UCFDReal tol, haptol;
    UCFDInt maxiter, itnum;
    UCFDReal residual;
    ucfd_solver_t stat;
    UCFDReal *hist_residual;
    bool right;              // extension: right preconditioning / relative tolerance
    int n, m;
    double *vec;             // device work vectors (one allocation)
    double *V, *w, *r, *z;   // GMRES
    double *Z;               // GMRES, right mode: the preconditioned basis M^-1 v_j
    double *rt, *p, *pt, *v, *s, *shat, *t;   // BiCGSTAB (r shared)
    double *part;            // reduction partials
    double *dbuf;            // device scalars
    double *hbuf;            // pinned host scalars
};

namespace {

ucfd_status_t fail(const char *fn, const char *msg, int code = UCFD_ERR_ARG)
{
    ucfd_report(code, fn, msg);
    return UCFD_FAILED;
}

ucfd_status_t status(int rc) { return rc == UCFD_OK ? UCFD_SUCCESS : UCFD_FAILED; }

int hip(hipError_t e, const char *fn, const char *what) { return ucfd_check_hip(e, fn, what); }

// Validate a row pointer (0 at the start, nondecreasing) and return its last entry.
int read_nnz(const int *rowptr, int rows, int64_t *nnz, const char *fn)
{
    if (!rowptr) return ucfd_report(UCFD_ERR_ARG, fn, "NULL rowptr");
    std::vector<int> rp;
    int rc = ucfd_fetch_ints(rowptr, (size_t)rows + 1, rp, fn);
    if (rc) return rc;
    if (rp[0] != 0) return ucfd_report(UCFD_ERR_INDEX, fn, "rowptr[0] must be 0");
    for (int i = 0; i < rows; ++i)
        if (rp[i + 1] < rp[i]) return ucfd_report(UCFD_ERR_INDEX, fn, "rowptr must be nondecreasing");
    *nnz = rp[rows];
    return UCFD_OK;
}

int mat_ready(SpMat A, const char *fn)
{
    if (!A || !A->type_name) return ucfd_report(UCFD_ERR_STATE, fn, "Matrix must be constructed");
    if (A->nnz >= 0) return UCFD_OK;
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;
    return read_nnz(A->rowptr, A->bn, &A->nnz, fn);
}

struct MatDev {
    const int *rowptr, *colidx;
    const double *values;
};

int stage_mat(UcfdStage &sg, SpMat A, int idx[3])
{
    const size_t b2 = (size_t)A->blk*A->blk;
    idx[0] = sg.add(A->rowptr, sizeof(int)*((size_t)A->bn + 1), UCFD_IN);
    idx[1] = sg.add(A->colidx, sizeof(int)*(size_t)A->nnz, UCFD_IN);
    idx[2] = sg.add(A->values, sizeof(double)*(size_t)A->nnz*b2, UCFD_IN);
    return UCFD_OK;
}

MatDev mat_dev(const UcfdStage &sg, const int idx[3])
{
    return MatDev{(const int *)sg.dev(idx[0]), (const int *)sg.dev(idx[1]),
                  (const double *)sg.dev(idx[2])};
}

int spmv(hipStream_t s, SpMat A, const MatDev &d, double alpha, const double *x, double beta,
         double *y)
{
    if (A->csr)
        return kry_spmv_csr(s, A->n, d.rowptr, d.colidx, d.values, alpha, x, beta, y,
                            ucfd_device_flag());
    return kry_spmv_bsr(s, A->bn, A->blk, d.rowptr, d.colidx, d.values, alpha, x, beta, y,
                        ucfd_device_flag());
}

struct PcDev {
    const int *rowptr, *colidx, *diagslots;
    const double *values;
};

// Sweep kind of the preconditioner's apply (KRY_SW_*), -1 when it has none.
int pc_kind(Precon pc)
{
    const char *t = pc->type_name;
    if (!t) return -1;
    if (strcmp(t, BLUSGS) == 0 || strcmp(t, PBLUSGS) == 0) return KRY_SW_LUSGS;
    if (strcmp(t, BILU) == 0 || strcmp(t, PBILU) == 0) return KRY_SW_BILU;
    if (strcmp(t, ILU) == 0) return KRY_SW_ILU;
    return -1;
}

bool pc_active(Precon pc) { return pc_kind(pc) >= 0; }

bool pc_colored(Precon pc)
{
    return strcmp(pc->type_name, PBLUSGS) == 0 || strcmp(pc->type_name, PBILU) == 0;
}

int stage_pc(UcfdStage &sg, Precon pc, int idx[4])
{
    const size_t b2 = (size_t)pc->blk*pc->blk;
    idx[0] = sg.add(pc->rowptr, sizeof(int)*((size_t)pc->bn + 1), UCFD_IN);
    idx[1] = sg.add(pc->colidx, sizeof(int)*(size_t)pc->nnz, UCFD_IN);
    idx[2] = sg.add(pc->diagslots, sizeof(int)*(size_t)pc->bn, UCFD_IN);
    idx[3] = sg.add(pc->values, sizeof(double)*(size_t)pc->nnz*b2, UCFD_IN);
    return UCFD_OK;
}

PcDev pc_dev(const UcfdStage &sg, const int idx[4])
{
    return PcDev{(const int *)sg.dev(idx[0]), (const int *)sg.dev(idx[1]),
                 (const int *)sg.dev(idx[2]), (const double *)sg.dev(idx[3])};
}

// Apply the preconditioner in place on the device vector b (n = bn*blk).
int pc_apply(hipStream_t s, Precon pc, const PcDev &d, double *b)
{
    const int kind = pc_kind(pc);
    if (kind < 0) return UCFD_OK;
    int rc = UCFD_OK;
    if (!pc->fp32) {
        for (const UcfdLevel &l : pc->lower.lv)
            if ((rc = kry_bsgs_sweep64(s, kind, 0, pc->blk, l.base, l.cnt,
                                       l.ordered ? pc->lower.order : nullptr, d.rowptr, d.colidx,
                                       d.diagslots, d.values, pc->diagvalues, b, l.maxw)))
                return rc;
        for (const UcfdLevel &l : pc->upper.lv)
            if ((rc = kry_bsgs_sweep64(s, kind, 1, pc->blk, l.base, l.cnt,
                                       l.ordered ? pc->upper.order : nullptr, d.rowptr, d.colidx,
                                       d.diagslots, d.values, pc->diagvalues, b, l.maxw)))
                return rc;
        return UCFD_OK;
    }
    if (pc->init32 && (rc = kry_cast32(s, (int64_t)pc->bn*pc->blk, b, pc->x32))) return rc;
    for (const UcfdLevel &l : pc->lower.lv)
        if ((rc = kry_bsgs_sweep32(s, 0, pc->blk, l.base, l.cnt,
                                   l.ordered ? pc->lower.order : nullptr, d.rowptr, d.colidx,
                                   d.diagslots, pc->vals32, pc->diagvalues, pc->x32, b, l.maxw)))
            return rc;
    for (const UcfdLevel &l : pc->upper.lv)
        if ((rc = kry_bsgs_sweep32(s, 1, pc->blk, l.base, l.cnt,
                                   l.ordered ? pc->upper.order : nullptr, d.rowptr, d.colidx,
                                   d.diagslots, pc->vals32, pc->diagvalues, pc->x32, b, l.maxw)))
            return rc;
    return UCFD_OK;
}

// ------------------------------------------------------------ schedules
//
// Visit orders of the reference (see schedule.h for the levels):
// blusgs.c:43-99 and bilu.c:67-130 rows ascending, then descending;
// pblusgs.c:90-104 and pbilu.c:80-180 at one thread: colour ranges ascending
// with rows ascending, then colour ranges descending with rows ascending.  A
// colour-ordered matrix whose ranges hold no internal couplings (the
// intended use of the coloured forms) gets one launch per colour range
// without building levels.
//
// The ILU(0) factorisations (ilu.c:4-48, bilu.c:5-65, pbilu.c:5-76) visit
// rows in the lower sweep's order, and row i reads exactly the rows of its
// lower slots (whole rows, where the sweep reads their vector entries) and
// writes only itself: the lower schedule is also the factorisation's.

// Blocks that the rows of one wave of the wave-staged sweep kernels use, at
// most, over the waves of each contiguous level (rows base .. base+cnt-1 in
// groups of 64/blk): sizes the LDS of each launch, so that levels whose rows
// have no block in the sweep's direction take no block staging space.
void set_wave_blocks(UcfdSchedule &S, bool upper, int blk, const std::vector<int> &rp,
                     const std::vector<int> &ds)
{
    const int G = blk > 1 ? 64/blk : 64;
    for (UcfdLevel &l : S.lv) {
        if (l.ordered) continue;
        int mx = 0;
        for (int w0 = 0; w0 < l.cnt; w0 += G) {
            int sum = 0;
            for (int t = w0; t < l.cnt && t < w0 + G; ++t) {
                const int i = l.base + t;
                sum += upper ? rp[i + 1] - ds[i] - 1 : ds[i] - rp[i];
            }
            mx = sum > mx ? sum : mx;
        }
        l.maxw = mx;
    }
}

int make_schedules(Precon pc, const char *fn)
{
    const int bn = pc->bn;
    std::vector<int> rp, ci, ds;
    int rc;
    if ((rc = ucfd_fetch_ints(pc->rowptr, (size_t)bn + 1, rp, fn))) return rc;
    if ((rc = ucfd_fetch_ints(pc->colidx, (size_t)pc->nnz, ci, fn))) return rc;
    if ((rc = ucfd_fetch_ints(pc->diagslots, (size_t)bn, ds, fn))) return rc;
    for (int64_t j = 0; j < pc->nnz; ++j)
        if (ci[j] < 0 || ci[j] >= bn)
            return ucfd_report(UCFD_ERR_INDEX, fn, "colidx holds an entry outside [0, bn)");
    for (int i = 0; i < bn; ++i)
        if (ds[i] < rp[i] || ds[i] >= rp[i + 1])
            return ucfd_report(UCFD_ERR_INDEX, fn, "diagslots[i] must lie in rowptr[i] .. rowptr[i+1]-1");
    pc->self_free = true;
    for (int i = 0; i < bn && pc->self_free; ++i)
        for (int j = rp[i]; j < ds[i]; ++j)
            if (ci[j] == i) {
                pc->self_free = false;
                break;
            }

    std::vector<int> lseq, useq;
    std::vector<int> off;
    const bool colored = pc_colored(pc);
    if (colored) {
        if (pc->ncolors < 0 || !pc->icolors) return ucfd_report(UCFD_ERR_ARG, fn, "bad colour offsets");
        if ((rc = ucfd_fetch_ints(pc->icolors, (size_t)pc->ncolors + 1, off, fn))) return rc;
        for (int c = 0; c < pc->ncolors; ++c)
            if (off[c] < 0 || off[c] > off[c + 1] || off[c + 1] > bn)
                return ucfd_report(UCFD_ERR_INDEX, fn, "icolors must be nondecreasing offsets in [0, bn]");
        for (int c = 0; c < pc->ncolors; ++c)
            for (int i = off[c]; i < off[c + 1]; ++i) lseq.push_back(i);
        for (int c = pc->ncolors - 1; c >= 0; --c)
            for (int i = off[c]; i < off[c + 1]; ++i) useq.push_back(i);
    } else {
        for (int i = 0; i < bn; ++i) lseq.push_back(i);
        for (int i = bn - 1; i >= 0; --i) useq.push_back(i);
    }

    bool unvis_l = false, unvis_u = false;
    bool fast = colored;
    if (colored) {
        // one level per colour range when no range couples to itself
        for (int c = 0; c < pc->ncolors && fast; ++c)
            for (int i = off[c]; i < off[c + 1] && fast; ++i)
                for (int j = rp[i]; j < rp[i + 1]; ++j) {
                    if (j == ds[i]) continue;
                    const int col = ci[j];
                    if (col >= off[c] && col < off[c + 1]) { fast = false; break; }
                    if (j < ds[i] && col >= off[c]) unvis_l = true;
                }
    }
    if (fast) {
        ucfd_free_schedule(pc->lower);
        ucfd_free_schedule(pc->upper);
        for (int c = 0; c < pc->ncolors; ++c)
            if (off[c + 1] > off[c]) pc->lower.lv.push_back(UcfdLevel{off[c], off[c + 1] - off[c], false});
        for (int c = pc->ncolors - 1; c >= 0; --c)
            if (off[c + 1] > off[c]) pc->upper.lv.push_back(UcfdLevel{off[c], off[c + 1] - off[c], false});
        set_wave_blocks(pc->lower, false, pc->blk, rp, ds);
        set_wave_blocks(pc->upper, true, pc->blk, rp, ds);
        pc->init32 = unvis_l;
        return UCFD_OK;
    }
    std::vector<std::vector<int>> levels;
    std::vector<int> jb(bn), je(bn);
    for (int i = 0; i < bn; ++i) {
        jb[i] = rp[i];
        je[i] = ds[i];
    }
    ucfd_build_levels(bn, lseq, jb, je, ci, levels, &unvis_l);
    if ((rc = ucfd_finish_schedule(levels, pc->lower, fn))) return rc;
    for (int i = 0; i < bn; ++i) {
        jb[i] = ds[i] + 1;
        je[i] = rp[i + 1];
    }
    ucfd_build_levels(bn, useq, jb, je, ci, levels, &unvis_u);
    if ((rc = ucfd_finish_schedule(levels, pc->upper, fn))) return rc;
    set_wave_blocks(pc->lower, false, pc->blk, rp, ds);
    set_wave_blocks(pc->upper, true, pc->blk, rp, ds);
    pc->init32 = unvis_l;
    return UCFD_OK;
}

int pc_alloc(Precon pc, const char *fn)
{
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;
    if (!pc->diagvalues) {
        const size_t bytes = sizeof(double)*(size_t)pc->bn*pc->blk*pc->blk;
        if ((rc = hip(hipMalloc((void **)&pc->diagvalues, bytes ? bytes : 8), fn,
                      "hipMalloc(diagvalues)")))
            return rc;
    }
    return UCFD_OK;
}

void pc_free_device(Precon pc)
{
    if (pc->diagvalues) (void)hipFree(pc->diagvalues);
    if (pc->vals32) (void)hipFree(pc->vals32);
    if (pc->x32) (void)hipFree(pc->x32);
    pc->diagvalues = nullptr;
    pc->vals32 = nullptr;
    pc->x32 = nullptr;
    ucfd_free_schedule(pc->lower);
    ucfd_free_schedule(pc->upper);
    pc->prepared = false;
    pc->scheduled = false;
}

ucfd_status_t set_block(Precon *precon, UCFDInt bn, UCFDInt block, PreconType type, int ncolors,
                        const UCFDInt *icolors, const char *fn)
{
    if (!precon || !*precon) return fail(fn, "Preconditioner must be initialized");
    if (bn < 0 || block < 1 || block > 7) return fail(fn, "need bn >= 0 and 1 <= block <= 7");
    Precon pc = *precon;
    pc_free_device(pc);
    pc->type_name = type;
    if (pc_kind(pc) != KRY_SW_LUSGS) pc->fp32 = false;
    pc->bn = bn;
    pc->blk = block;
    pc->ncolors = ncolors;
    pc->icolors = icolors;
    return status(pc_alloc(pc, fn));
}

// --------------------------------------------------------------- solvers

int d2h(Solver S, int off, int count, hipStream_t s, const char *fn)
{
    int rc = hip(hipMemcpyAsync(S->hbuf + off, S->dbuf + off, sizeof(double)*count,
                                hipMemcpyDeviceToHost, s), fn, "hipMemcpyAsync(scalars)");
    if (rc) return rc;
    return hip(hipStreamSynchronize(s), fn, "hipStreamSynchronize");
}

void record(Solver S, int iter, double res)
{
    if (S->hist_residual && iter >= 0 && iter < S->maxiter) S->hist_residual[iter] = res;
}

// gmres.c:42-129; right mode: flexible right preconditioning (the
// preconditioned basis z_j = M^-1 v_j is kept and x += Z y, so the Arnoldi
// relation A Z = V H holds in FP64 even when M^-1 is applied in FP32 and is
// not exactly linear), tolerance tol*||b||, exit inside a cycle when the
// Givens estimate reaches it.
int gmres(Solver S, Precon pc, SpMat A, const MatDev &ma, const PcDev &mp, double *x,
          const double *b, hipStream_t s, const char *fn)
{
    const int n = S->n, m = S->m, ld = m + 1;
    const double tol = S->tol, haptol = S->haptol;
    const int64_t N = n;
    std::vector<double> H((size_t)ld*m, 0.0), y(ld, 0.0), cs(ld, 0.0), sn(ld, 0.0);
    double *V = S->V, *w = S->w, *r = S->r, *part = S->part, *db = S->dbuf;
    double *hb = S->hbuf;
    int rc, iter = 0, k = 0;
    double abeta = 0.0, tolv = tol;
    if (S->right) {
        if ((rc = kry_dot(s, N, b, b, part, db + DB_NSQ)) || (rc = d2h(S, DB_NSQ, 1, s, fn)))
            return rc;
        tolv = tol*sqrt(hb[DB_NSQ]);
    }
    while (iter < S->maxiter) {
        if ((rc = kry_copy(s, N, b, r)) || (rc = spmv(s, A, ma, -1.0, x, 1.0, r))) return rc;
        if ((rc = kry_dot(s, N, r, r, part, db + DB_NSQ)) || (rc = d2h(S, DB_NSQ, 1, s, fn)))
            return rc;
        abeta = sqrt(hb[DB_NSQ]);
        record(S, iter, abeta);
        if (abeta <= tolv) {
            S->stat = CONVERGED;
            break;
        }
        double beta = abeta;
        int nsq_slot = DB_NSQ;
        if (!S->right) {
            if ((rc = pc_apply(s, pc, mp, r))) return rc;
            nsq_slot = DB_NSQ + 1;
            if ((rc = kry_dot(s, N, r, r, part, db + nsq_slot)) || (rc = d2h(S, nsq_slot, 1, s, fn)))
                return rc;
            beta = sqrt(hb[nsq_slot]);
        }
        y[0] = beta;
        // right mode: z_0 starts as a copy of v_0 (M^-1 is applied in place)
        if ((rc = kry_scale_by_inv_sqrt(s, N, r, db + nsq_slot, V, S->right ? S->Z : nullptr)))
            return rc;
        k = 0;
        bool happy = false;
        for (int j = 0; j < m; ++j) {
            const int kk = j + 1;
            double *vj = V + (size_t)j*n;
            if (S->right) {
                // flexible form: z_j = M^-1 v_j is kept, x is updated with Z y;
                // z_j holds a copy of v_j written with it
                double *zj = S->Z + (size_t)j*n;
                if ((rc = pc_apply(s, pc, mp, zj)) ||
                    (rc = spmv(s, A, ma, 1.0, zj, 0.0, w)))
                    return rc;
            } else {
                if ((rc = spmv(s, A, ma, 1.0, vj, 0.0, w)) || (rc = pc_apply(s, pc, mp, w)))
                    return rc;
            }
            // CGS2 (gmres.c:20-29); the first update and the second projection
            // share one pass over V
            if ((rc = kry_gemvt(s, N, kk, V, N, w, part, db + DB_H)) ||
                (rc = kry_gemv_gemvt(s, N, kk, V, N, -1.0, db + DB_H, w, part, db + DB_H2)) ||
                (rc = kry_gemv(s, N, kk, V, N, -1.0, db + DB_H2, w, part, db + DB_NSQ)) ||
                (rc = kry_scale_by_inv_sqrt(s, N, w, db + DB_NSQ, V + (size_t)(j + 1)*n,
                                            S->right && j + 1 < m ? S->Z + (size_t)(j + 1)*n
                                                                  : nullptr)) ||
                (rc = d2h(S, 0, DB_NSQ + 1, s, fn)))
                return rc;
            double *Hcol = H.data() + (size_t)j*ld;
            for (int i = 0; i < kk; ++i) Hcol[i] = hb[DB_H + i] + 1.0*hb[DB_H2 + i];
            const double hsub = sqrt(hb[DB_NSQ]);
            Hcol[j + 1] = hsub;
            // Givens rotations (gmres.c:86-101)
            for (int i = 0; i < j; ++i) {
                const double c = cs[i], sv = sn[i];
                const double t = c*Hcol[i] + sv*Hcol[i + 1];
                Hcol[i + 1] = -sv*Hcol[i] + c*Hcol[i + 1];
                Hcol[i] = t;
            }
            const double h1 = Hcol[j], h2 = Hcol[j + 1];
            const double rr = hypot(h1, h2);
            cs[j] = h1/rr;
            sn[j] = h2/rr;
            Hcol[j] = cs[j]*Hcol[j] + sn[j]*Hcol[j + 1];
            Hcol[j + 1] = 0.0;
            y[j + 1] = -sn[j]*y[j];
            y[j] = cs[j]*y[j];
            k = j + 1;
            if (hsub < haptol*beta) {
                happy = true;
                break;
            }
            if (S->right && fabs(y[j + 1]) <= tolv) break;
        }
        // back substitution (gmres.c:112-118)
        for (int i = k - 1; i >= 0; --i) {
            double sum = y[i];
            for (int jj = i + 1; jj < k; ++jj) sum -= H[i + (size_t)jj*ld]*y[jj];
            y[i] = sum/H[i + (size_t)i*ld];
        }
        if (k > 0) {
            memcpy(hb + DB_Y, y.data(), sizeof(double)*k);
            if ((rc = hip(hipMemcpyAsync(db + DB_Y, hb + DB_Y, sizeof(double)*k,
                                         hipMemcpyHostToDevice, s), fn, "hipMemcpyAsync(y)")))
                return rc;
            if (S->right) {
                if ((rc = kry_gemv(s, N, k, S->Z, N, 1.0, db + DB_Y, x, part, nullptr)))
                    return rc;
            } else if ((rc = kry_gemv(s, N, k, V, N, 1.0, db + DB_Y, x, part, nullptr))) {
                return rc;
            }
            // the host buffer of y is rewritten only after the next synchronisation
        }
        if (happy) S->stat = HAPPYBREAKDOWN;
        iter++;
    }
    if (iter == S->maxiter) S->stat = REACH_ITERMAX;
    S->residual = abeta;
    S->itnum = iter;
    return UCFD_OK;
}

// bicgstab.c:24-131 (right preconditioned); right mode: tolerance tol*||b||.
int bicgstab(Solver S, Precon pc, SpMat A, const MatDev &ma, const PcDev &mp, double *x,
             const double *b, hipStream_t s, const char *fn)
{
    const int64_t N = S->n;
    double *r = S->r, *rt = S->rt, *p = S->p, *pt = S->pt, *v = S->v, *sv = S->s;
    double *shat = S->shat, *t = S->t, *part = S->part, *db = S->dbuf, *hb = S->hbuf;
    int rc, iter = 0;
    double rho, rhoprev, alpha, beta, omega, resnorm = 0.0, tolv = S->tol;
    if (S->right) {
        if ((rc = kry_dot(s, N, b, b, part, db + DB_NSQ)) || (rc = d2h(S, DB_NSQ, 1, s, fn)))
            return rc;
        tolv = S->tol*sqrt(hb[DB_NSQ]);
    }
    if ((rc = kry_copy(s, N, b, r)) || (rc = spmv(s, A, ma, -1.0, x, 1.0, r)) ||
        (rc = kry_copy(s, N, r, rt)) || (rc = kry_dot(s, N, rt, r, part, db + DB_NSQ)) ||
        (rc = d2h(S, DB_NSQ, 1, s, fn)))
        return rc;
    rho = hb[DB_NSQ];
    rhoprev = 1.0;
    alpha = 1.0;
    omega = 1.0;
    if (fabs(rho) < 1e-30) {
        S->stat = RHOBREAKDOWN;
        goto done;
    }
    while (iter < S->maxiter) {
        if (iter == 0) {
            if ((rc = kry_copy(s, N, r, p)) || (rc = kry_copy(s, N, r, pt))) return rc;
        } else {
            beta = (rho/rhoprev)*(alpha/omega);
            if ((rc = kry_bicg_p(s, N, beta, omega, r, v, p, pt))) return rc;
        }
        // pi = rt.v stays on the device: s = r - (rho/pi) v is formed there, and
        // pi comes back with the five dots in one copy (one host wait per
        // iteration).  A pi breakdown then stops before x and r change.
        if ((rc = pc_apply(s, pc, mp, pt)) || (rc = spmv(s, A, ma, 1.0, pt, 0.0, v)) ||
            (rc = kry_dot(s, N, rt, v, part, db + DB_NSQ)) ||
            (rc = kry_bicg_s(s, N, rho, db + DB_NSQ, r, v, sv, shat)) ||
            (rc = pc_apply(s, pc, mp, shat)) || (rc = spmv(s, A, ma, 1.0, shat, 0.0, t)) ||
            (rc = kry_dot5(s, N, rt, sv, t, part, db + DB_H)) ||
            (rc = d2h(S, DB_H, DB_NSQ + 1 - DB_H, s, fn)))
            return rc;
        const double pi = hb[DB_NSQ];
        if (fabs(pi) < 1e-30) {
            S->stat = PIBREAKDOWN;
            break;
        }
        alpha = rho/pi;
        const double phi = hb[DB_H], rtt = hb[DB_H + 1], ts = hb[DB_H + 2], tt = hb[DB_H + 3],
                     ss = hb[DB_H + 4];
        omega = ts/tt;
        if ((rc = kry_bicg_xr(s, N, alpha, pt, omega, shat, sv, t, x, r))) return rc;
        const double rho_new = phi - omega*rtt;
        resnorm = sqrt(fabs(ss - 2.0*omega*ts + omega*omega*tt));
        record(S, iter, resnorm);
        if (resnorm < tolv) {
            S->stat = CONVERGED;
            goto done;
        }
        rhoprev = rho;
        rho = rho_new;
        iter++;
    }
done:
    if (iter == S->maxiter) S->stat = REACH_ITERMAX;
    S->itnum = iter;
    S->residual = resnorm;
    return UCFD_OK;
}

ucfd_status_t solver_create(Solver *solver, SolverType type, UCFDInt n, UCFDInt m, UCFDInt maxiter,
                            UCFDReal tol, const char *fn)
{
    if (!solver) return fail(fn, "NULL solver handle");
    if (n < 0) return fail(fn, "need n >= 0");
    if (strcmp(type, GMRES) == 0 && (m < 1 || m > HCAP - 1))
        return fail(fn, "the restart length must lie in 1 .. 63");
    int rc = ucfd_ensure_device(fn);
    if (rc) return UCFD_FAILED;
    if (UCFDSolverInit(solver) != UCFD_SUCCESS) return UCFD_FAILED;
    Solver S = *solver;
    S->type_name = type;
    S->n = n;
    S->m = m;
    S->tol = tol;
    S->maxiter = maxiter;
    const size_t nv = strcmp(type, GMRES) == 0 ? (size_t)(m + 1) + 3 : 8;
    const size_t bytes = sizeof(double)*((size_t)n*nv + (size_t)HCAP*1024 + DB_N);
    if ((rc = hip(hipMalloc((void **)&S->vec, bytes), fn, "hipMalloc(work vectors)")) ||
        (rc = hip(hipMemset(S->vec, 0, bytes), fn, "hipMemset(work vectors)")) ||
        (rc = hip(hipHostMalloc((void **)&S->hbuf, sizeof(double)*DB_N, hipHostMallocDefault), fn,
                  "hipHostMalloc(scalars)"))) {
        UCFDSolverDestroy(solver);
        return UCFD_FAILED;
    }
    double *q = S->vec;
    auto take = [&](size_t cnt) { double *o = q; q += cnt; return o; };
    if (strcmp(type, GMRES) == 0) {
        S->V = take((size_t)n*(m + 1));
        S->w = take(n);
        S->r = take(n);
        S->z = take(n);
    } else {
        S->r = take(n);
        S->rt = take(n);
        S->p = take(n);
        S->pt = take(n);
        S->v = take(n);
        S->s = take(n);
        S->shat = take(n);
        S->t = take(n);
    }
    S->part = take((size_t)HCAP*1024);
    S->dbuf = take(DB_N);
    return UCFD_SUCCESS;
}

}  // namespace

// ================================================================ matrices

UCFD_API ucfd_status_t UCFDMatInit(SpMat *mat)
{
    if (!mat) return fail("UCFDMatInit", "NULL matrix handle");
    SpMat m = (SpMat)calloc(1, sizeof(*m));
    if (!m) return fail("UCFDMatInit", "Matrix allocation failed", UCFD_ERR_ALLOC);
    m->nnz = -1;
    *mat = m;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDMatDestroy(SpMat *mat)
{
    if (!mat || !*mat) return UCFD_SUCCESS;
    free(*mat);
    *mat = nullptr;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDMatCreateCSR(SpMat *mat, UCFDInt n, UCFDInt *rowptr, UCFDInt *colidx,
                                        UCFDReal *values)
{
    if (n < 0) return fail("UCFDMatCreateCSR", "need n >= 0");
    if (UCFDMatInit(mat) != UCFD_SUCCESS) return UCFD_FAILED;
    SpMat m = *mat;
    m->type_name = CSR;
    m->csr = true;
    m->n = n;
    m->bn = n;
    m->blk = 1;
    m->rowptr = rowptr;
    m->colidx = colidx;
    m->values = values;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDMatCreateBSR(SpMat *mat, UCFDInt bn, UCFDInt blk, UCFDInt *rowptr,
                                        UCFDInt *colidx, UCFDReal *values)
{
    if (bn < 0 || blk < 1 || blk > 7) return fail("UCFDMatCreateBSR", "need bn >= 0, 1 <= blk <= 7");
    if (UCFDMatInit(mat) != UCFD_SUCCESS) return UCFD_FAILED;
    SpMat m = *mat;
    m->type_name = BSR;
    m->csr = false;
    m->n = bn*blk;
    m->bn = bn;
    m->blk = blk;
    m->rowptr = rowptr;
    m->colidx = colidx;
    m->values = values;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDMatCreateCUDACSR(SpMat *mat, UCFDInt n, UCFDInt *rowptr,
                                            UCFDInt *colidx, UCFDReal *values)
{
    return UCFDMatCreateCSR(mat, n, rowptr, colidx, values);
}

UCFD_API ucfd_status_t UCFDMatCreateCUDABSR(SpMat *mat, UCFDInt bn, UCFDInt blk, UCFDInt *rowptr,
                                            UCFDInt *colidx, UCFDReal *values)
{
    return UCFDMatCreateBSR(mat, bn, blk, rowptr, colidx, values);
}

UCFD_API ucfd_status_t UCFDMatMult(UCFDReal alpha, SpMat mat, UCFDReal *x, UCFDReal beta,
                                   UCFDReal *y)
{
    const char *fn = "UCFDMatMult";
    int rc = mat_ready(mat, fn);
    if (rc) return UCFD_FAILED;
    UcfdStage sg(fn);
    int mi[3];
    stage_mat(sg, mat, mi);
    const int ix = sg.add(x, sizeof(double)*(size_t)mat->n, UCFD_IN);
    const int iy = sg.add(y, sizeof(double)*(size_t)mat->n, UCFD_INOUT);
    if ((rc = sg.enter())) return UCFD_FAILED;
    rc = spmv(ucfd_stream(), mat, mat_dev(sg, mi), alpha, (const double *)sg.dev(ix), beta,
              (double *)sg.dev(iy));
    return status(sg.leave(rc));
}

// ========================================================== preconditioners

UCFD_API ucfd_status_t UCFDPreconCreatefromArrays(Precon *precon, UCFDInt *rowptr, UCFDInt *colidx,
                                                  UCFDInt *diagslots, UCFDReal *values)
{
    if (!precon) return fail("UCFDPreconCreatefromArrays", "NULL preconditioner handle");
    Precon pc = new (std::nothrow) _Precon();
    if (!pc) return fail("UCFDPreconCreatefromArrays", "Precon allocation failed", UCFD_ERR_ALLOC);
    pc->rowptr = rowptr;
    pc->colidx = colidx;
    pc->diagslots = diagslots;
    pc->values = values;
    pc->nnz = -1;
    *precon = pc;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDPreconCreateNone(Precon *precon)
{
    if (UCFDPreconCreatefromArrays(precon, nullptr, nullptr, nullptr, nullptr) != UCFD_SUCCESS)
        return UCFD_FAILED;
    (*precon)->type_name = NONE;
    (*precon)->prepared = true;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDPreconSetBLUSGS(Precon *precon, UCFDInt bn, UCFDInt block)
{
    return set_block(precon, bn, block, BLUSGS, 0, nullptr, "UCFDPreconSetBLUSGS");
}

UCFD_API ucfd_status_t UCFDPreconSetPBLUSGS(Precon *precon, UCFDInt bn, UCFDInt block,
                                            UCFDInt ncolors, UCFDInt *icolors)
{
    return set_block(precon, bn, block, PBLUSGS, ncolors, icolors, "UCFDPreconSetPBLUSGS");
}

UCFD_API ucfd_status_t UCFDPreconSetCUDABLUSGS(Precon *precon, UCFDInt bn, UCFDInt block,
                                               UCFDInt ncolors, UCFDInt *icolors)
{
    return set_block(precon, bn, block, PBLUSGS, ncolors, icolors, "UCFDPreconSetCUDABLUSGS");
}

UCFD_API ucfd_status_t UCFDPreconSetILU(Precon *precon, UCFDInt n)
{
    return set_block(precon, n, 1, ILU, 0, nullptr, "UCFDPreconSetILU");
}

UCFD_API ucfd_status_t UCFDPreconSetBILU(Precon *precon, UCFDInt bn, UCFDInt block)
{
    return set_block(precon, bn, block, BILU, 0, nullptr, "UCFDPreconSetBILU");
}

UCFD_API ucfd_status_t UCFDPreconSetPBILU(Precon *precon, UCFDInt bn, UCFDInt block,
                                          UCFDInt ncolors, UCFDInt *icolors)
{
    return set_block(precon, bn, block, PBILU, ncolors, icolors, "UCFDPreconSetPBILU");
}

UCFD_API ucfd_status_t UCFDPreconSetCUDABILU(Precon *precon, UCFDInt bn, UCFDInt block,
                                             UCFDInt ncolors, UCFDInt *icolors)
{
    return set_block(precon, bn, block, PBILU, ncolors, icolors, "UCFDPreconSetCUDABILU");
}

UCFD_API ucfd_status_t UCFDPreconSetPrecision(Precon precon, UCFDInt fp32)
{
    if (!precon || pc_kind(precon) != KRY_SW_LUSGS)
        return fail("UCFDPreconSetPrecision", "a block LU-SGS preconditioner is required");
    if (precon->fp32 != (fp32 != 0)) precon->prepared = false;
    precon->fp32 = fp32 != 0;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDPreconPrepare(Precon precon)
{
    const char *fn = "UCFDPreconPrepare";
    if (!precon || !precon->type_name) return fail(fn, "Preconditioner type must be set", UCFD_ERR_STATE);
    if (!pc_active(precon)) return UCFD_SUCCESS;
    Precon pc = precon;
    int rc = pc_alloc(pc, fn);
    if (rc) return UCFD_FAILED;
    if (!pc->scheduled) {
        // the structure is validated and scheduled once per preconditioner
        // type; later Prepares read only the values
        if ((rc = read_nnz(pc->rowptr, pc->bn, &pc->nnz, fn))) return UCFD_FAILED;
        if (!pc->colidx || !pc->diagslots || !pc->values) return fail(fn, "NULL preconditioner array");
        if ((rc = make_schedules(pc, fn))) return UCFD_FAILED;
        pc->scheduled = true;
    }
    const size_t b2 = (size_t)pc->blk*pc->blk;
    const int kind = pc_kind(pc);
    if (kind != KRY_SW_LUSGS) {
        // ILU(0): factored in place on the caller's values, as the reference does
        UcfdStage sg(fn);
        int pi[4];
        pi[0] = sg.add(pc->rowptr, sizeof(int)*((size_t)pc->bn + 1), UCFD_IN);
        pi[1] = sg.add(pc->colidx, sizeof(int)*(size_t)pc->nnz, UCFD_IN);
        pi[2] = sg.add(pc->diagslots, sizeof(int)*(size_t)pc->bn, UCFD_IN);
        pi[3] = sg.add(pc->values, sizeof(double)*(size_t)pc->nnz*b2, UCFD_INOUT);
        if ((rc = sg.enter())) return UCFD_FAILED;
        const PcDev d = pc_dev(sg, pi);
        for (const UcfdLevel &l : pc->lower.lv)
            if ((rc = kry_ilu_factor(ucfd_stream(), kind == KRY_SW_ILU, pc->self_free, pc->blk,
                                     l.base, l.cnt,
                                     l.ordered ? pc->lower.order : nullptr, d.rowptr, d.colidx,
                                     d.diagslots, (double *)d.values, pc->diagvalues,
                                     ucfd_device_flag())))
                break;
        if (!rc) rc = ucfd_sync_and_check(fn);
        rc = sg.leave(rc);
        pc->prepared = rc == UCFD_OK;
        return status(rc);
    }
    if (pc->fp32 && !pc->vals32) {
        const size_t vb = sizeof(float)*(size_t)pc->nnz*b2, xb = sizeof(float)*(size_t)pc->bn*pc->blk;
        if ((rc = hip(hipMalloc((void **)&pc->vals32, vb ? vb : 4), fn, "hipMalloc(vals32)")) ||
            (rc = hip(hipMalloc((void **)&pc->x32, xb ? xb : 4), fn, "hipMalloc(x32)")))
            return UCFD_FAILED;
    }
    UcfdStage sg(fn);
    const int ids = sg.add(pc->diagslots, sizeof(int)*(size_t)pc->bn, UCFD_IN);
    const int iv = sg.add(pc->values, sizeof(double)*(size_t)pc->nnz*b2, UCFD_IN);
    if ((rc = sg.enter())) return UCFD_FAILED;
    rc = kry_bsgs_prepare(ucfd_stream(), pc->bn, pc->blk, pc->nnz, (const int *)sg.dev(ids),
                          (const double *)sg.dev(iv), pc->diagvalues, pc->fp32 ? pc->vals32 : nullptr,
                          ucfd_device_flag());
    // the factorisation is complete and checked for zero pivots on return
    if (!rc) rc = ucfd_sync_and_check(fn);
    rc = sg.leave(rc);
    pc->prepared = rc == UCFD_OK;
    return status(rc);
}

UCFD_API ucfd_status_t UCFDPreconApply(Precon precon, UCFDReal *b)
{
    const char *fn = "UCFDPreconApply";
    if (!precon || !precon->type_name) return fail(fn, "Preconditioner type must be set", UCFD_ERR_STATE);
    if (!pc_active(precon)) return UCFD_SUCCESS;
    if (!precon->prepared) return fail(fn, "UCFDPreconPrepare must be called first", UCFD_ERR_STATE);
    UcfdStage sg(fn);
    int pi[4];
    stage_pc(sg, precon, pi);
    const int ib = sg.add(b, sizeof(double)*(size_t)precon->bn*precon->blk, UCFD_INOUT);
    int rc = sg.enter();
    if (rc) return UCFD_FAILED;
    rc = pc_apply(ucfd_stream(), precon, pc_dev(sg, pi), (double *)sg.dev(ib));
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDPreconDestroy(Precon *precon)
{
    if (!precon || !*precon) return UCFD_SUCCESS;
    Precon pc = *precon;
    if (pc->diagvalues || pc->vals32 || pc->x32 || pc->lower.order || pc->upper.order)
        (void)hipStreamSynchronize(ucfd_stream());
    pc_free_device(pc);
    delete pc;
    *precon = nullptr;
    return UCFD_SUCCESS;
}

// ================================================================= solvers

UCFD_API ucfd_status_t UCFDSolverInit(Solver *solver)
{
    if (!solver) return fail("UCFDSolverInit", "NULL solver handle");
    Solver s = (Solver)calloc(1, sizeof(*s));
    if (!s) return fail("UCFDSolverInit", "Solver allocation failed", UCFD_ERR_ALLOC);
    s->type_name = nullptr;
    s->tol = 1e-5;
    s->haptol = 1e-12;
    s->maxiter = -1;
    s->itnum = 0;
    s->residual = 0.0;
    s->stat = INITIALIZED;
    *solver = s;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDSolverDestroy(Solver *solver)
{
    if (!solver || !*solver) return UCFD_SUCCESS;
    Solver s = *solver;
    if (s->vec || s->Z) {
        (void)hipStreamSynchronize(ucfd_stream());
        if (s->vec) (void)hipFree(s->vec);
        if (s->Z) (void)hipFree(s->Z);
    }
    if (s->hbuf) (void)hipHostFree(s->hbuf);
    free(s->hist_residual);
    free(s);
    *solver = nullptr;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDSolverCreateGMRES(Solver *solver, UCFDInt n, UCFDInt m, UCFDInt maxiter,
                                             UCFDReal tol)
{
    return solver_create(solver, GMRES, n, m, maxiter, tol, "UCFDSolverCreateGMRES");
}

UCFD_API ucfd_status_t UCFDSolverCreateCUDAGMRES(Solver *solver, UCFDInt n, UCFDInt m,
                                                 UCFDInt maxiter, UCFDReal tol)
{
    return solver_create(solver, GMRES, n, m, maxiter, tol, "UCFDSolverCreateCUDAGMRES");
}

UCFD_API ucfd_status_t UCFDSolverCreateBICGSTAB(Solver *solver, UCFDInt n, UCFDInt maxiter,
                                                UCFDReal tol)
{
    return solver_create(solver, BICGSTAB, n, 0, maxiter, tol, "UCFDSolverCreateBICGSTAB");
}

UCFD_API ucfd_status_t UCFDSolverSetRightPrecon(Solver solver, UCFDBool right)
{
    if (!solver || !solver->type_name)
        return fail("UCFDSolverSetRightPrecon", "Solver must be initialized", UCFD_ERR_STATE);
    solver->right = right;
    // flexible GMRES keeps the m preconditioned basis vectors
    if (right && !solver->Z && strcmp(solver->type_name, GMRES) == 0 &&
        hip(hipMalloc((void **)&solver->Z, sizeof(double)*((size_t)solver->n*solver->m + 1)),
            "UCFDSolverSetRightPrecon", "hipMalloc(preconditioned basis)"))
        return UCFD_FAILED;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDSolve(Solver solver, Precon pc, SpMat A, UCFDReal *x, UCFDReal *b)
{
    const char *fn = "UCFDSolve";
    if (!solver || !solver->type_name) return fail(fn, "Solver must be initialized", UCFD_ERR_STATE);
    if (!pc || !pc->type_name) return fail(fn, "Preconditioner must be initialized", UCFD_ERR_STATE);
    int rc = mat_ready(A, fn);
    if (rc) return UCFD_FAILED;
    if (A->n != solver->n) return fail(fn, "matrix and solver sizes differ");
    const bool blockpc = pc_active(pc);
    if (blockpc) {
        if (!pc->prepared) return fail(fn, "UCFDPreconPrepare must be called first", UCFD_ERR_STATE);
        if (pc->bn*pc->blk != solver->n) return fail(fn, "preconditioner and solver sizes differ");
    }
    UcfdStage sg(fn);
    int mi[3], pi[4];
    stage_mat(sg, A, mi);
    if (blockpc) stage_pc(sg, pc, pi);
    const int ix = sg.add(x, sizeof(double)*(size_t)solver->n, UCFD_INOUT);
    const int ib = sg.add(b, sizeof(double)*(size_t)solver->n, UCFD_IN);
    if ((rc = sg.enter())) return UCFD_FAILED;
    const MatDev md = mat_dev(sg, mi);
    const PcDev pd = blockpc ? pc_dev(sg, pi) : PcDev{nullptr, nullptr, nullptr, nullptr};
    double *xd = (double *)sg.dev(ix);
    const double *bd = (const double *)sg.dev(ib);
    if (strcmp(solver->type_name, GMRES) == 0)
        rc = gmres(solver, pc, A, md, pd, xd, bd, ucfd_stream(), fn);
    else
        rc = bicgstab(solver, pc, A, md, pd, xd, bd, ucfd_stream(), fn);
    // a solve is synchronous: the result and the device error word are final on return
    if (!rc) rc = ucfd_sync_and_check(fn);
    return status(sg.leave(rc));
}

UCFD_API ucfd_status_t UCFDSolverGetResult(Solver solver, UCFDInt *stat, UCFDInt *iter,
                                           UCFDReal *residual)
{
    if (!solver) return fail("UCFDSolverGetResult", "NULL solver");
    if (stat) *stat = solver->stat;
    if (iter) *iter = solver->itnum;
    if (residual) *residual = solver->residual;
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDSolverTraceResidualHistory(Solver solver)
{
    if (!solver) return fail("UCFDSolverTraceResidualHistory", "NULL solver");
    free(solver->hist_residual);
    const size_t n = solver->maxiter > 0 ? (size_t)solver->maxiter : 0;
    solver->hist_residual = (UCFDReal *)calloc(n ? n : 1, sizeof(UCFDReal));
    if (!solver->hist_residual)
        return fail("UCFDSolverTraceResidualHistory", "allocation failed", UCFD_ERR_ALLOC);
    return UCFD_SUCCESS;
}

UCFD_API ucfd_status_t UCFDSolverGetResidualHistory(Solver solver, UCFDReal *hist)
{
    if (!solver || !solver->hist_residual)
        return fail("UCFDSolverGetResidualHistory", "Tracing residual unset", UCFD_ERR_STATE);
    if (solver->maxiter > 0) memcpy(hist, solver->hist_residual, solver->maxiter*sizeof(UCFDReal));
    return UCFD_SUCCESS;
}
