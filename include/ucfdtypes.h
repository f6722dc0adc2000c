/*
 * Types of the Krylov API (matrices, preconditioners, solvers).
 *
 * Replaces src/krylov/include/ucfdtypes.h:17-41 and config.h:15-28 of the
 * reference: the same status and solver-state enumerators with the same
 * values, UCFDInt = int32_t and UCFDReal = double (the reference default; its
 * UCFD_INT64 / UCFD_FLOAT32 switches are not offered, the library computes in
 * FP64 with an opt-in FP32 preconditioner apply).
 */
#pragma once

#include <stdbool.h>
#include <stdint.h>

#if defined(UCFD_INT64) || defined(UCFD_FLOAT32)
#error "this library is built for 32-bit UCFDInt and double UCFDReal only"
#endif

#if defined(__cplusplus)
#define UCFD_EXTERN extern "C"
#else
#define UCFD_EXTERN extern
#endif

typedef int32_t UCFDInt;
typedef double UCFDReal;
typedef int8_t UCFDInt8;
typedef bool UCFDBool;

#define UCFD_FALSE false
#define UCFD_TRUE true

/* ucfdtypes.h:17-21 */
typedef enum {
    UCFD_SUCCESS = 0,
    UCFD_FAILED = 1
} ucfd_status_t;

/* ucfdtypes.h:23-31 */
typedef enum {
    INITIALIZED = -1,
    CONVERGED = 0,
    HAPPYBREAKDOWN = 1,
    REACH_ITERMAX = 2,
    RHOBREAKDOWN = 3,
    PIBREAKDOWN = 4
} ucfd_solver_t;
