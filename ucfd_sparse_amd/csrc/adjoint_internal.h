// Internal interface between the adjoint host code (adjoint_api.cpp) and its
// kernels (adjoint_kernels.hip).  Not installed.
#pragma once

#include <stdint.h>

#include "ucfd_internal.h"

// pair[f*N + i] = face of nb = nei[f*N + i] pointing back to i (-1 on self
// faces); a missing back face or an index outside [0, N) raises
// UCFD_KFLAG_INDEX.
UCFD_HIDDEN int adj_launch_pairs(hipStream_t s, int64_t N, int F, const int *nei, int *pair,
                                 int *flag);
// One colour range [n0, ne) of the adjoint: lower == 0 -> transposed upper
// sweep (work = abar from w), lower == 1 -> transposed lower sweep (work =
// bbar, and the cells' rhsbar, dtbar, fsprbar).
UCFD_HIDDEN int adj_launch_colour(hipStream_t s, int lower, int n0, int ne, int64_t N, int nface,
                                  int ndims, const int *nei, const int *pair, const double *fnv,
                                  const double *vfn, const double *fspr, const double *dt,
                                  double factor, const double *diag, const double *upts,
                                  const double *dstar, const double *dq, const double *w,
                                  double *work, double *rhsbar, double *dtbar, double *fsprbar);
