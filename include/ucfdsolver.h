/*
 * Krylov solvers (replaces src/krylov/include/ucfdsolver.h).
 *
 * Restarted GMRES(m) with CGS2 orthogonalisation and Givens rotations, and
 * BiCGSTAB with the fused five-way reduction, driven from the host with all
 * vectors on the GPU in FP64.  Default semantics are the reference's:
 * GMRES is left preconditioned, stops on the absolute true residual at the
 * start of a restart, and maxiter counts restarts (gmres.c:42-129); BiCGSTAB
 * is right preconditioned with an absolute tolerance (bicgstab.c:24-131).
 * UCFDSolve is synchronous.
 */
#pragma once

#include "ucfdprecon.h"
#include "ucfdtypes.h"

typedef struct _Solver *Solver;
typedef const char *SolverType;

#define GMRES "gmres"
#define BICGSTAB "bicgstab"

#if defined(__cplusplus)
extern "C" {
#endif

/* solver.c:5-22 */
UCFD_EXTERN ucfd_status_t UCFDSolverInit(Solver *solver);
/* solver.c:24-35 */
UCFD_EXTERN ucfd_status_t UCFDSolverDestroy(Solver *solver);
/* solver.c:36-40: x (initial guess, overwritten) and b of length n */
UCFD_EXTERN ucfd_status_t UCFDSolve(Solver solver, Precon pc, SpMat A, UCFDReal *x, UCFDReal *b);
/* solver.c:42-48 */
UCFD_EXTERN ucfd_status_t UCFDSolverGetResult(Solver solver, UCFDInt *stat, UCFDInt *iter,
                                              UCFDReal *residual);
/* solver.c:56-62: record one residual per restart (GMRES) or iteration
 * (BiCGSTAB), maxiter entries */
UCFD_EXTERN ucfd_status_t UCFDSolverTraceResidualHistory(Solver solver);
/* solver.c:64-69 */
UCFD_EXTERN ucfd_status_t UCFDSolverGetResidualHistory(Solver solver, UCFDReal *hist);
/* gmres.c:153-178: 1 <= m <= 63 */
UCFD_EXTERN ucfd_status_t UCFDSolverCreateGMRES(Solver *solver, UCFDInt n, UCFDInt m,
                                                UCFDInt maxiter, UCFDReal tol);
/* bicgstab.c:156-179 */
UCFD_EXTERN ucfd_status_t UCFDSolverCreateBICGSTAB(Solver *solver, UCFDInt n, UCFDInt maxiter,
                                                   UCFDReal tol);
/* ucfdsolver.h:28 (CUDA builds): GMRES under the CUDA name */
UCFD_EXTERN ucfd_status_t UCFDSolverCreateCUDAGMRES(Solver *solver, UCFDInt n, UCFDInt m,
                                                    UCFDInt maxiter, UCFDReal tol);

/* Extension: right = true switches GMRES to flexible right preconditioning
 * (the preconditioned basis M^-1 v_j is kept, m more vectors of n doubles,
 * so a preconditioner applied in FP32 keeps the FP64 Arnoldi relation), and
 * both solvers to the relative tolerance tol*||b||_2; GMRES then also leaves
 * a restart cycle as soon as the Givens residual estimate reaches it. */
UCFD_EXTERN ucfd_status_t UCFDSolverSetRightPrecon(Solver solver, UCFDBool right);

#if defined(__cplusplus)
}
#endif
