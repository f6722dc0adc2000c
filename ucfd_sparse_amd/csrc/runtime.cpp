// Host runtime of the library: stream selection, pointer-kind dispatch,
// staging of host arrays, the device error word and error reporting.
//
// Call contract (all C-ABI entry points):
//  * Device or managed pointers are used in place.  Host pointers are copied
//    to a library workspace, the kernels run, written arrays are copied back.
//  * With no user stream set (default), every call is synchronous: it returns
//    after its kernels finished and reports index errors found by them.
//  * After ucfd_set_stream(s), calls whose arrays are all device-resident
//    only enqueue work on s and return; ucfd_synchronize() waits and reports.
//  * Errors return a nonzero status and print one line to stderr (the
//    reference prints and exits, macros.h:34-68); UCFD_ABORT_ON_ERROR=1 in the
//    environment restores exit-on-error.
#include "ucfd_internal.h"
#include "ucfd_ext.h"

#include <mutex>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace {

struct State {
    std::recursive_mutex mu;
    hipStream_t stream = nullptr;
    bool user_stream = false;
    bool inited = false;
    int device = -1;
    int *dflag = nullptr;        // device error word
    int *hflag = nullptr;        // pinned host copy
    void *ws = nullptr;          // staging workspace (grow only)
    size_t ws_bytes = 0;
    UcfdConfig cfg{0, 0, 0};
    bool cfg_set = false;
};

State &st()
{
    static State s;
    return s;
}

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    return atoi(v);
}

#ifndef UCFD_NVARS
#define UCFD_NVARS 5
#endif
#ifndef UCFD_NFVARS
#define UCFD_NFVARS 5
#endif
#ifndef UCFD_NDIMS
#define UCFD_NDIMS 3
#endif

}  // namespace

int ucfd_report(int code, const char *fn, const char *msg)
{
    if (code == UCFD_OK) return code;
    fprintf(stderr, "[ucfd] %s: %s (status %d)\n", fn ? fn : "?", msg ? msg : "", code);
    fflush(stderr);
    if (env_int("UCFD_ABORT_ON_ERROR", 0)) exit(EXIT_FAILURE);
    return code;
}

int ucfd_check_hip(hipError_t e, const char *fn, const char *what)
{
    if (e == hipSuccess) return UCFD_OK;
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ucfd_report(UCFD_ERR_HIP, fn, buf);
}

const UcfdConfig &ucfd_config()
{
    State &s = st();
    if (!s.cfg_set) {
        s.cfg.nvars = env_int("UCFD_NVARS", UCFD_NVARS);
        s.cfg.nfvars = env_int("UCFD_NFVARS", UCFD_NFVARS);
        s.cfg.ndims = env_int("UCFD_NDIMS", UCFD_NDIMS);
        s.cfg_set = true;
    }
    return s.cfg;
}

int ucfd_ensure_device(const char *fn)
{
    State &s = st();
    std::lock_guard<std::recursive_mutex> g(s.mu);
    if (s.inited) return UCFD_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return ucfd_report(UCFD_ERR_HIP, fn, "no HIP device is available (this library runs "
                                             "only on the GPU; there is no CPU path)");
    }
    int rc;
    if ((rc = ucfd_check_hip(hipGetDevice(&s.device), fn, "hipGetDevice"))) return rc;
    if ((rc = ucfd_check_hip(hipMalloc((void **)&s.dflag, 256), fn, "hipMalloc(flag)"))) return rc;
    if ((rc = ucfd_check_hip(hipMemset(s.dflag, 0, 256), fn, "hipMemset(flag)"))) return rc;
    if ((rc = ucfd_check_hip(hipHostMalloc((void **)&s.hflag, 256, hipHostMallocDefault), fn,
                             "hipHostMalloc(flag)")))
        return rc;
    *s.hflag = 0;
    s.inited = true;
    return UCFD_OK;
}

hipStream_t ucfd_stream() { return st().stream; }
bool ucfd_stream_is_user() { return st().user_stream; }
int *ucfd_device_flag() { return st().dflag; }

bool ucfd_is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Wait for the stream and turn a raised device error word into a status.
int ucfd_sync_and_check(const char *fn)
{
    State &s = st();
    int rc;
    if ((rc = ucfd_check_hip(hipMemcpyAsync(s.hflag, s.dflag, sizeof(int), hipMemcpyDeviceToHost,
                                            s.stream), fn, "hipMemcpyAsync(flag)")))
        return rc;
    if ((rc = ucfd_check_hip(hipStreamSynchronize(s.stream), fn, "hipStreamSynchronize")))
        return rc;
    const int f = *s.hflag;
    if (f) {
        (void)hipMemset(s.dflag, 0, sizeof(int));
        *s.hflag = 0;
        if (f & UCFD_KFLAG_INDEX)
            return ucfd_report(UCFD_ERR_INDEX, fn, "an index array holds an entry outside "
                                                   "[0, neles); the affected cells were skipped");
        if (f & UCFD_KFLAG_DUP)
            return ucfd_report(UCFD_ERR_INDEX, fn, "cell_ids maps two elements to the same cell; "
                                                   "the rank rows of that cell are undefined");
        if (f & UCFD_KFLAG_LAYOUT)
            return ucfd_report(UCFD_ERR_STATE, fn, "icolor/lcolor no longer hold the layout "
                                                   "declared with these pointers (changed, or "
                                                   "freed and reused): release and re-declare it");
        return ucfd_report(UCFD_ERR_NUMERIC, fn, "zero pivot in a diagonal block factorisation");
    }
    return UCFD_OK;
}

int UcfdStage::add(const void *p, size_t bytes, int dir)
{
    Arg &a = args[nargs];
    a.host = p;
    a.bytes = bytes;
    a.dir = dir;
    a.dev = nullptr;
    a.staged = false;
    return nargs++;
}

int UcfdStage::enter()
{
    State &s = st();
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;
    s.mu.lock();
    // pass 1: classify; host arrays that alias share the first one's buffer
    for (int k = 0; k < nargs; ++k) {
        Arg &a = args[k];
        a.owner = -1;
        if (!a.host) {
            s.mu.unlock();
            return ucfd_report(UCFD_ERR_ARG, fn, "NULL array argument");
        }
        if (ucfd_is_device_ptr(a.host)) {
            a.dev = const_cast<void *>(a.host);
            continue;
        }
        a.staged = true;
        a.owner = k;
        for (int j = 0; j < k; ++j)
            if (args[j].staged && args[j].host == a.host) { a.owner = args[j].owner; break; }
        if (a.owner != k) {
            Arg &o = args[a.owner];
            if (a.bytes > o.bytes) o.bytes = a.bytes;
            o.dir |= a.dir;
        }
    }
    // pass 2: place owners in the workspace
    size_t need = 0;
    for (int k = 0; k < nargs; ++k)
        if (args[k].staged && args[k].owner == k) need += (args[k].bytes + 255) & ~(size_t)255;
    if (need > s.ws_bytes) {
        if (s.ws) (void)hipFree(s.ws);
        s.ws = nullptr;
        s.ws_bytes = 0;
        const size_t sz = need + need/4;
        if ((rc = ucfd_check_hip(hipMalloc(&s.ws, sz), fn, "hipMalloc(workspace)"))) {
            s.mu.unlock();
            return rc;
        }
        s.ws_bytes = sz;
    }
    size_t off = 0;
    for (int k = 0; k < nargs; ++k) {
        Arg &a = args[k];
        if (!a.staged) continue;
        any_staged = true;
        if (a.owner == k) {
            a.dev = (char *)s.ws + off;
            off += (a.bytes + 255) & ~(size_t)255;
        } else {
            a.dev = args[a.owner].dev;
        }
    }
    // pass 3: copy inputs
    for (int k = 0; k < nargs; ++k) {
        Arg &a = args[k];
        if (!a.staged || a.owner != k || !(a.dir & UCFD_IN) || a.bytes == 0) continue;
        if ((rc = ucfd_check_hip(hipMemcpyAsync(a.dev, a.host, a.bytes, hipMemcpyHostToDevice,
                                                s.stream), fn, "hipMemcpyAsync(H2D)"))) {
            s.mu.unlock();
            return rc;
        }
    }
    return UCFD_OK;   // the lock is held until leave()
}

int UcfdStage::leave(int status)
{
    State &s = st();
    int rc = status;
    for (int k = 0; k < nargs; ++k) {
        Arg &a = args[k];
        if (!a.staged || a.owner != k || !(a.dir & UCFD_OUT) || a.bytes == 0) continue;
        int r = ucfd_check_hip(hipMemcpyAsync(const_cast<void *>(a.host), a.dev, a.bytes,
                                              hipMemcpyDeviceToHost, s.stream),
                               fn, "hipMemcpyAsync(D2H)");
        if (!rc) rc = r;
    }
    if (any_staged || !s.user_stream) {
        int r = ucfd_sync_and_check(fn);
        if (!rc) rc = r;
    } else {
        int r = ucfd_check_hip(hipGetLastError(), fn, "kernel launch");
        if (!rc) rc = r;
    }
    s.mu.unlock();
    return rc;
}

unsigned long long ucfd_hash_host(const int *a, int64_t n)
{
    unsigned long long s = 0ull;
    for (int64_t k = 0; k < n; ++k) {
        unsigned long long z = (unsigned long long)k*0x9E3779B97F4A7C15ull +
                               (unsigned long long)(unsigned)a[k];
        z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27))*0x94D049BB133111EBull;
        s += z ^ (z >> 31);
    }
    return s;
}

// ---------------------------------------------------------------- extensions

UCFD_API int ucfd_set_stream(void *stream)
{
    State &s = st();
    std::lock_guard<std::recursive_mutex> g(s.mu);
    s.stream = (hipStream_t)stream;
    s.user_stream = stream != nullptr;
    return UCFD_OK;
}

UCFD_API void *ucfd_get_stream(void) { return (void *)st().stream; }

UCFD_API int ucfd_synchronize(void)
{
    int rc = ucfd_ensure_device("ucfd_synchronize");
    if (rc) return rc;
    std::lock_guard<std::recursive_mutex> g(st().mu);
    return ucfd_sync_and_check("ucfd_synchronize");
}

UCFD_API int ucfd_set_config(int nvars, int nfvars, int ndims)
{
    if (ndims < 2 || ndims > 3 || nfvars != ndims + 2 || nvars < nfvars || nvars > 16)
        return ucfd_report(UCFD_ERR_ARG, "ucfd_set_config",
                           "need ndims in {2,3}, nfvars = ndims+2, nfvars <= nvars <= 16");
    State &s = st();
    std::lock_guard<std::recursive_mutex> g(s.mu);
    s.cfg = UcfdConfig{nvars, nfvars, ndims};
    s.cfg_set = true;
    return UCFD_OK;
}

UCFD_API int ucfd_get_config(int *nvars, int *nfvars, int *ndims)
{
    const UcfdConfig &c = ucfd_config();
    if (nvars) *nvars = c.nvars;
    if (nfvars) *nfvars = c.nfvars;
    if (ndims) *ndims = c.ndims;
    return UCFD_OK;
}

UCFD_API int ucfd_release_workspace(void)
{
    State &s = st();
    std::lock_guard<std::recursive_mutex> g(s.mu);
    if (s.ws) {
        int rc = ucfd_check_hip(hipStreamSynchronize(s.stream), "ucfd_release_workspace", "sync");
        if (rc) return rc;
        (void)hipFree(s.ws);
    }
    s.ws = nullptr;
    s.ws_bytes = 0;
    return UCFD_OK;
}

UCFD_API int ucfd_device_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}
