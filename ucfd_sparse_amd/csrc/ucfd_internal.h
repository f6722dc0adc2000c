// Internal declarations shared by the host runtime (.cpp, built with g++) and
// the kernel translation units (.hip, built with hipcc).  Not installed.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#define UCFD_API extern "C" __attribute__((visibility("default")))
#define UCFD_HIDDEN __attribute__((visibility("hidden")))

// Library status codes (0 = success), returned by every C-ABI entry point.
enum {
    UCFD_OK = 0,
    UCFD_ERR_ARG = 1,        // invalid scalar argument or NULL pointer
    UCFD_ERR_HIP = 2,        // HIP runtime error (no device, launch failure, ...)
    UCFD_ERR_INDEX = 3,      // an index array held an entry outside [0, neles)
    UCFD_ERR_ALLOC = 4,      // device or host allocation failed
    UCFD_ERR_STATE = 5,      // call sequence / handle state error
    UCFD_ERR_NUMERIC = 6     // non-finite value or zero pivot detected
};

// Bits raised by kernels in the device error word.
enum { UCFD_KFLAG_INDEX = 1, UCFD_KFLAG_PIVOT = 2, UCFD_KFLAG_LAYOUT = 4, UCFD_KFLAG_DUP = 8 };

// Flow-variable configuration of the legacy API (compile-time in the
// reference: NVARS / NFVARS / NDIMS, Makefile.inc:21-24).
struct UcfdConfig {
    int nvars;
    int nfvars;
    int ndims;
};

// ---- host runtime (runtime.cpp)
UCFD_HIDDEN int ucfd_report(int code, const char *fn, const char *msg);
UCFD_HIDDEN int ucfd_check_hip(hipError_t e, const char *fn, const char *what);
UCFD_HIDDEN hipStream_t ucfd_stream();
UCFD_HIDDEN bool ucfd_stream_is_user();
UCFD_HIDDEN const UcfdConfig &ucfd_config();
UCFD_HIDDEN int *ucfd_device_flag();          // device error word (per process)
UCFD_HIDDEN int ucfd_ensure_device(const char *fn);
UCFD_HIDDEN bool ucfd_is_device_ptr(const void *p);
// Wait for the library stream; a raised device error word becomes a status
// (and is cleared).  Call with the runtime lock held or outside any staging.
UCFD_HIDDEN int ucfd_sync_and_check(const char *fn);

// Staging of host arrays for one call.  Device (or managed) pointers are used
// in place; host pointers are copied to a library workspace on entry and
// copied back on exit when written.  Arguments that alias share one buffer.
enum { UCFD_IN = 1, UCFD_OUT = 2, UCFD_INOUT = 3 };

struct UcfdStage {
    struct Arg { const void *host; size_t bytes; int dir; void *dev; bool staged; int owner; };
    Arg args[24];
    int nargs = 0;
    bool any_staged = false;
    const char *fn;
    explicit UcfdStage(const char *name) : fn(name) {}
    // register an argument; returns its index
    int add(const void *p, size_t bytes, int dir);
    // resolve device addresses and copy inputs; returns status
    int enter();
    // copy outputs back and synchronise as the call contract requires
    int leave(int status);
    void *dev(int idx) const { return args[idx].dev; }
};

// ---- kernel launchers (lusgs_kernels.hip)
enum { UCFD_SWEEP_COLORED = 0, UCFD_SWEEP_ORDERED = 1, UCFD_SWEEP_RANGE = 2 };

UCFD_HIDDEN int ucfd_launch_pre(hipStream_t s, int neles, int nface, double factor,
                                const double *fnorm_vol, const double *dt, double *diag,
                                const double *fspr);

// dir 0 = lower, 1 = upper.  mode COLORED: cells order[n0:ne), neighbour
// taken iff lcolor[nb] </> lcolor[i].  mode ORDERED: cells order[n0:ne),
// neighbour taken iff nb </> i (sequential semantics, level scheduled).
// mode RANGE: cells n0..ne-1 of a declared colour-contiguous layout,
// neighbour taken iff nb < n0 / nb >= ne; order (= icolor) and lcolor are
// only sampled, one cell per workgroup, against the declared layout
// (icolor[i] == i, lcolor[i] == lc_expect; UCFD_KFLAG_LAYOUT otherwise).
UCFD_HIDDEN int ucfd_launch_sweep(hipStream_t s, int dir, int mode, int n0, int ne, int neles,
                                  int nface, int ndims, int nfvars, const int *nei,
                                  const int *order, const int *lcolor, const double *fnorm_vol,
                                  const double *vec_fnorm, const double *upts, const double *in,
                                  double *out, const double *diag, const double *fspr,
                                  int *flag, int lc_expect);

// The turbulence rows of the legacy RANS sweeps (rows ndims+2 ..
// ndims+2+nt-1, nt in {1, 2}; nface in {3, 4, 6}); modes and tests as
// ucfd_launch_sweep.  UCFD_ERR_ARG for an unsupported nface / nt / ndims.
UCFD_HIDDEN int ucfd_launch_rans_sweep(hipStream_t s, int dir, int mode, int n0, int ne,
                                       int neles, int nface, int ndims, int nt, const int *nei,
                                       const int *order, const int *lcolor,
                                       const double *fnorm_vol, const double *vec_fnorm,
                                       const double *upts, const double *in, double *out,
                                       const double *tdiag, const double *tfspr, int *flag,
                                       int lc_expect);

// y += x over nrows rows of neles cells ([nrows][neles])
UCFD_HIDDEN int ucfd_launch_update(hipStream_t s, int neles, int nrows, double *y, const double *x);
// *out = sum over k < n of a mix of (k, a[k]) mod 2^64 (device array a; out
// device memory): a content hash, order independent and deterministic.
UCFD_HIDDEN int ucfd_launch_hash(hipStream_t s, const int *a, int64_t n, unsigned long long *out);
// the same hash of a host array
UCFD_HIDDEN unsigned long long ucfd_hash_host(const int *a, int64_t n);
