/*
 * Extensions of the MI355X LU-SGS library that have no counterpart in the
 * reference's public headers.  They cover what a GPU caller needs around the
 * legacy array API (stream control, configuration of the compile-time
 * NVARS / NFVARS / NDIMS of src/lusgs, colouring and colour-contiguous
 * renumbering).  All functions return 0 on success unless stated otherwise.
 */
#ifndef UCFD_EXT_H
#define UCFD_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- execution control ---------------------------------------------------
 * By default every call is synchronous.  After ucfd_set_stream(s) (a
 * hipStream_t), calls whose arrays are all device resident only enqueue on
 * s; ucfd_synchronize() waits for s and reports kernel-detected errors.
 * ucfd_set_stream(NULL) restores the synchronous default. */
int ucfd_set_stream(void *hip_stream);
void *ucfd_get_stream(void);
int ucfd_synchronize(void);
/* number of visible HIP devices (0: every compute entry point will fail) */
int ucfd_device_available(void);
/* free the staging workspace used for host-pointer calls */
int ucfd_release_workspace(void);

/* ---- configuration --------------------------------------------------------
 * nvars: rows touched by update (NVARS of the reference build,
 * Makefile.inc:21); nfvars = ndims + 2: rows touched by the NS sweeps;
 * ndims in {2, 3}.  Defaults 5/5/3, or the environment variables
 * UCFD_NVARS / UCFD_NFVARS / UCFD_NDIMS. */
int ucfd_set_config(int nvars, int nfvars, int ndims);
int ucfd_get_config(int *nvars, int *nfvars, int *ndims);

/* ---- colouring (host) -----------------------------------------------------
 * First-fit greedy colouring in ascending cell id on the symmetrised face
 * graph (self faces ignored).  Writes lcolor[neles] (1-based), icolor[neles]
 * (cells colour by colour, ascending id within a colour; nullable) and
 * color_off[ncolors+1] (colour c+1 is icolor[color_off[c] .. color_off[c+1]);
 * nullable).  Returns ncolors, or < 0 on bad input.  On an unpermuted
 * structured hex grid the result is the red-black colouring of
 * examples/utils/pbutils.c:98-121. */
int ucfd_greedy_coloring(int neles, int nface, const int *nei_ele, int *icolor,
                         int *lcolor, int *color_off);
/* First-fit greedy colouring visiting the cells in order[0..neles) (a
 * permutation; no counterpart in the reference), e.g. along a space-filling
 * curve of the cell centres.  Same outputs as ucfd_greedy_coloring; returns
 * ncolors, or < 0 on bad input (order not a permutation). */
int ucfd_greedy_coloring_ordered(int neles, int nface, const int *nei_ele, const int *order,
                                 int *icolor, int *lcolor, int *color_off);
/* Breadth-first visiting order of the face graph (no counterpart in the
 * reference), components entered at their first cell in seeds (NULL:
 * ascending id).  Greedy colouring in this order 2-colours any bipartite
 * face graph.  Returns the number of components, or -1 on bad input. */
int ucfd_bfs_order(int neles, int nface, const int *nei_ele, const int *seeds, int *order);
/* Removes the highest colour where possible (no counterpart in the
 * reference): each of its cells takes a colour no neighbour uses, or one
 * freed by swapping a two-colour Kempe chain of at most max_chain cells.
 * lcolor (a valid colouring) is updated in place; icolor and color_off
 * (nullable) are rebuilt.  Returns the new number of colours. */
int ucfd_reduce_colors(int neles, int nface, const int *nei_ele, int max_chain, int *icolor,
                       int *lcolor, int *color_off);
/* Balanced colouring (no counterpart in the reference): cells in ascending
 * id take the least-used colour among 1..C not used by a coloured neighbour,
 * C = max(ncolors, max degree + 1).  Same outputs as ucfd_greedy_coloring
 * (color_off holds C+1 entries); returns C. */
int ucfd_balanced_coloring(int neles, int nface, const int *nei_ele, int ncolors, int *icolor,
                           int *lcolor, int *color_off);
/* face pairs with equal colours (0 = valid colouring) */
long long ucfd_coloring_conflicts(int neles, int nface, const int *nei_ele,
                                  const int *lcolor);

/* ---- colour-contiguous renumbering (host) ---------------------------------
 * perm[old] = new with cells sorted by (lcolor, key, id); key (nullable) is a
 * locality key such as a Morton code of the cell centre.  After permuting
 * every per-cell array with perm (and nei_ele with ucfd_permute_nei), icolor
 * becomes the identity and each colour is a contiguous range, so the sweeps
 * read every array coalesced.  Returns ncolors. */
int ucfd_color_reorder(int neles, const int *lcolor, const unsigned long long *key,
                       int *perm, int *color_off);
/* out[r*neles + perm[i]] = in[r*neles + i], r < nrows, esize 4 or 8 bytes */
int ucfd_permute_rows(int neles, int nrows, int esize, const int *perm,
                      const void *in, void *out);
/* out[f*neles + perm[i]] = perm[nei[f*neles + i]] */
int ucfd_permute_nei(int neles, int nface, const int *perm, const int *nei, int *out);

/* ---- declared colour-contiguous layouts ------------------------------------
 * Declares that icolor (the identity) and lcolor (non-decreasing) describe a
 * colour-contiguous mesh; the library checks both arrays once (host or device
 * memory) and remembers the pointer pair.  Coloured sweeps called with these
 * exact pointers and an [n0, ne) covering one whole colour then decide
 * "lower" / "upper" from the index range (nb < n0 / nb >= ne), which equals
 * the lcolor test on such a layout and saves the lcolor[nb] gathers.  The
 * arrays must not change while declared: release them first.  Each such
 * call samples one cell per workgroup against the declaration and fails with
 * status 5 (UCFD_ERR_STATE) if icolor / lcolor no longer match it (arrays
 * changed, or freed and the addresses reused).  A colour in which no face
 * can be taken (the first colour of a lower sweep, the last of an upper
 * sweep) reads no neighbour index on such a layout, so an out-of-range
 * nei_ele entry of that colour is reported by the sweep in the other
 * direction rather than by this call.  Writes the number of colours to
 * *ncolors (nullable). */
int ucfd_declare_colored_layout(int neles, const int *icolor, const int *lcolor,
                                int *ncolors);
int ucfd_release_colored_layout(const int *icolor, const int *lcolor);

/* ---- defect-correction solver (FP64 and mixed FP32/FP64 forms) -------------
 * Solves A(dq) = b, the LU-SGS system of the legacy API,
 *   A(dq)_i = D_i dq_i + 1/2 sum_{f, nb != i} fnorm_vol_f [F(Q_nb + dq_nb; n_f)
 *             - F(Q_nb; n_f) - fspr_f dq_nb],  D_i = 1/dt_i + 1/2 sum_f fspr_f fnorm_vol_f,
 * by defect correction with one coloured LU-SGS pass as the preconditioner.
 * The mesh must be colour-contiguous: colour c = cells color_off[c] ..
 * color_off[c+1]-1 (see ucfd_color_reorder).  Inputs may be host or device
 * arrays in the legacy layouts (upts: the first 5 rows are used); the solver
 * keeps device copies (FP64 and packed FP32 face data) until destroyed.
 * UCFD_SOLVE_MIXED runs the sweeps in FP32 inside an FP64 outer residual
 * loop (iterative refinement); dq is accumulated in FP64. */
typedef struct ucfd_lusgs_solver ucfd_lusgs_solver;
#define UCFD_SOLVE_FP64 0
#define UCFD_SOLVE_MIXED 1
#define UCFD_SOLVE_HIST 128
typedef struct {
    int converged;          /* 1 if ||b - A(dq)|| <= rtol ||b|| */
    int outer;              /* FP64 residual evaluations (outer iterations) */
    int passes;             /* LU-SGS passes (lower + upper sweep) */
    int nhist;              /* entries of history[] */
    double rnorm0;          /* ||b|| */
    double rel_residual;    /* ||b - A(dq)|| / ||b|| of the returned dq */
    double history[UCFD_SOLVE_HIST];   /* relative residuals as computed */
} ucfd_solve_stats;
int ucfd_lusgs_solver_create(ucfd_lusgs_solver **solver, int neles, int nface,
                             int ncolors, const int *color_off, const int *nei_ele,
                             const double *fnorm_vol, const double *vec_fnorm,
                             const double *fspr, const double *dt, const double *upts);
int ucfd_lusgs_solver_destroy(ucfd_lusgs_solver *solver);
/* r = b - A(dq) in FP64 (r nullable); *rnorm = ||r||_2 (nullable) */
int ucfd_lusgs_residual(ucfd_lusgs_solver *solver, const double *b, const double *dq,
                        double *r, double *rnorm);
/* No-allocation variant of ucfd_lusgs_residual.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same r = b - A(dq) (r nullable) and *rnorm = ||r||_2 (nullable), bit for bit,
 * but without a solver handle: the operator is formed from the legacy arrays
 * (nei_ele, fnorm_vol, vec_fnorm, fspr, dt, upts [>= 5 rows]) of a
 * colour-contiguous mesh (colour c = cells color_off[c] .. color_off[c+1]-1;
 * color_off is host memory).  Every array must be device (or managed)
 * memory.  The call allocates nothing: its scratch (block partial sums, norm,
 * error word) is a fixed library array of 1 MiB per GPU, held until the
 * process ends, so the colour ranges may hold at most 131072 workgroups of
 * 256 cells (sum over colours of ceil(cells/256); about 33.5M cells) and
 * larger meshes are refused with status 1.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (several host threads or streams)
 * race and must be serialised by the caller.  Runs on the library stream and
 * returns after it finished (the norm is read back).  Status 3 for an index
 * outside [0, neles), 1 for a face joining two cells of one colour. */
int ucfd_lusgs_residual_noalloc(int neles, int nface, int ncolors, const int *color_off,
                                const int *nei_ele, const double *fnorm_vol,
                                const double *vec_fnorm, const double *fspr, const double *dt,
                                const double *upts, const double *b, const double *dq, double *r,
                                double *rnorm);
/* dq (output, [5][neles]) with ||b - A(dq)|| <= rtol ||b||; maxiter bounds the
 * number of LU-SGS passes; stats nullable */
int ucfd_lusgs_solve(ucfd_lusgs_solver *solver, int mode, const double *b, double *dq,
                     double rtol, int maxiter, ucfd_solve_stats *stats);

/* Measurement extension (no counterpart in the reference): HIP-event times
 * of the solver's kernels over one full pass each, averaged over reps, on the
 * library stream, as in the solve: ms[0] lower sweep with residual of the
 * colours after the first (FP64; with more than one colour the last
 * colour's launch also does that colour's upper step and update), ms[1]
 * upper sweep with update of the other colours (FP64; the first colour's
 * launch also forms that colour's residual and lower step for the next
 * pass), ms[2] / ms[3] the same
 * kernels in FP32 (the inner kernels of UCFD_SOLVE_MIXED), ms[4] the FP64
 * residual and its norm.  b is a device array [5][neles]; the solver's work
 * arrays are overwritten. */
int ucfd_lusgs_solver_time_kernels(ucfd_lusgs_solver *solver, const double *b, int reps,
                                   double *ms);

/* ---- Differentiation form (no counterpart in the reference).
 * Adjoint of one FP64 coloured pass dQ = upper(lower(pre(dt, fspr), rhsb))
 * on a colour-contiguous numbering (colour c = cells color_off[c] ..
 * color_off[c+1]-1, the layout of ucfd_declare_colored_layout): the upper
 * sweep transposed (colours ascending), then the lower sweep transposed
 * (colours descending), each cell gathering its neighbours' adjoints
 * through the face-pair map in face order (deterministic).  The handle
 * keeps the mesh pointers (nei_ele [nface][neles], fnorm_vol, vec_fnorm,
 * upts [NVARS][neles]); create validates the colour ranges and builds the
 * face-pair map. */
int ucfd_adjoint_create(void **handle, int neles, int nface, int ncolors, const int *color_off,
                        const int *nei_ele, const double *fnorm_vol, const double *vec_fnorm,
                        const double *upts);
int ucfd_adjoint_destroy(void *handle);
/* Given the forward's dstar (lower-sweep output) and dq (upper-sweep output),
 * both [NFVARS][neles], and w = dJ/d dq: rhsbar = dJ/d rhsb (rows < NFVARS),
 * dtbar = dJ/d dt [neles], fsprbar = dJ/d fspr [nface][neles].  factor as in
 * the pre call (diag = 1/(dt factor) + ...).  rhsbar may be the dstar array
 * (it is then overwritten), which saves one [NFVARS][neles] buffer. */
int ucfd_lusgs_adjoint(void *handle, double factor, const double *dt, const double *fspr,
                       const double *dstar, const double *dq, const double *w, double *rhsbar,
                       double *dtbar, double *fsprbar);

#ifdef __cplusplus
}
#endif

#endif /* UCFD_EXT_H */
