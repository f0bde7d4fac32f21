/*
 * og_runtime.c -- device context, error text and allocation helpers of
 * libopengraph_amd (host C over the HIP runtime C API).
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "og_internal.h"
#include "og_learn.h"
#include "opengraph_amd_ext.h"

static __thread char g_err[512];
static int g_ready = 0;
static int g_device = 0;
static int g_cus = 0;
static hipStream_t g_stream = NULL;
/* second stream and its fork / join events: the pull's long-row chains run
 * there beside the persistent kernel (og_pr.hip) */
static hipStream_t g_side = NULL;
static hipEvent_t g_fork = NULL, g_join = NULL;
static pthread_mutex_t g_ctx_lock = PTHREAD_MUTEX_INITIALIZER;

void og_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *og_get_error(void) { return g_err; }

int og_hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    og_set_error("%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return -1;
}

static __thread hipError_t g_launch_err = hipSuccess;

void og_launch_note(hipError_t e)
{
    if (e != hipSuccess && g_launch_err == hipSuccess) g_launch_err = e;
}

hipError_t og_launch_status(void)
{
    hipError_t e = g_launch_err;
    g_launch_err = hipSuccess;
    return e;
}

static void og_ctx_release(void)
{
    if (g_join) (void)hipEventDestroy(g_join);
    if (g_fork) (void)hipEventDestroy(g_fork);
    if (g_side) (void)hipStreamDestroy(g_side);
    if (g_stream) (void)hipStreamDestroy(g_stream);
    g_join = g_fork = NULL;
    g_side = g_stream = NULL;
}

/* The device, its streams (both non-blocking, default priority) and events,
 * created together once per process under a lock; on a failure nothing
 * made so far is kept and the next call starts again. */
static int og_ctx_create(void)
{
    int n = 0;
    if (og_hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return -1;
    if (n <= 0)
    {
        og_set_error("no HIP device is visible");
        return -1;
    }
    if (g_device >= n) g_device = 0;
    if (og_hip_check(hipSetDevice(g_device), "hipSetDevice")) return -1;
    hipDeviceProp_t prop;
    if (og_hip_check(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties")) return -1;
    g_cus = prop.multiProcessorCount;
    if (og_hip_check(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate") ||
        og_hip_check(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking), "hipStreamCreate (side)") ||
        og_hip_check(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming), "hipEventCreate (fork)") ||
        og_hip_check(hipEventCreateWithFlags(&g_join, hipEventDisableTiming), "hipEventCreate (join)") ||
        og_learn_noalloc_setup(g_device))
    {
        og_ctx_release();
        return -1;
    }
    return 0;
}

int og_device(void) { return g_device; }

int og_ctx_ensure(void)
{
    if (__atomic_load_n(&g_ready, __ATOMIC_ACQUIRE)) return 0;
    pthread_mutex_lock(&g_ctx_lock);
    int rc = 0;
    if (!g_ready)
    {
        rc = og_ctx_create();
        if (!rc) __atomic_store_n(&g_ready, 1, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&g_ctx_lock);
    return rc;
}

hipStream_t og_side_stream(void) { return g_side; }
hipEvent_t og_side_fork(void) { return g_fork; }
hipEvent_t og_side_join(void) { return g_join; }

hipStream_t og_stream(void) { return g_stream; }

hipError_t og_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, g_stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(g_stream);
}
int og_num_cus(void) { return g_cus; }

/* With OGPU_TRACE set (any value), each drop-in entry point names itself,
 * the device and the graph on stderr: a host program linked against the
 * library can show which definition it called.  Output only. */
void og_trace_entry(const char *fn, uint32_t n_vertices, uint32_t n_edges)
{
    if (!getenv("OGPU_TRACE")) return;
    fprintf(stderr, "opengraph_amd: %s on HIP device %d (%d CUs), %u vertices, %u edges\n", fn, g_device,
            g_cus, n_vertices, n_edges);
}

void *og_dalloc(size_t bytes)
{
    void *p = NULL;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
    {
        og_set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return NULL;
    }
    return p;
}

void og_dfree(void *p)
{
    if (p) (void)hipFree(p);
}

void *og_host_alloc(size_t bytes)
{
    void *p = NULL;
    if (bytes == 0) bytes = 1;
    if (posix_memalign(&p, 128, bytes) != 0)
    {
        fprintf(stderr, "Error, aligned_alloc of %zu bytes failed\n", bytes);
        exit(EXIT_FAILURE);
    }
    return p;
}

/* ------------------------- public helpers ------------------------- */

int ogpu_init(int device)
{
    if (g_ready && device != g_device)
    {
        og_set_error("ogpu_init: already initialised on device %d", g_device);
        return -1;
    }
    g_device = device < 0 ? 0 : device;
    return og_ctx_ensure();
}

const char *ogpu_last_error(void) { return g_err; }

int ogpu_device_info(char *name, size_t name_len, int *num_cus, size_t *total_mem)
{
    if (og_ctx_ensure()) return -1;
    hipDeviceProp_t prop;
    if (og_hip_check(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties")) return -1;
    if (name && name_len)
    {
        snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    return 0;
}

int ogpu_malloc(void **ptr, size_t bytes)
{
    if (og_ctx_ensure()) return -1;
    *ptr = og_dalloc(bytes);
    return *ptr ? 0 : -1;
}

int ogpu_free(void *ptr)
{
    if (!ptr) return 0;
    return og_hip_check(hipFree(ptr), "hipFree");
}

int ogpu_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (og_ctx_ensure()) return -1;
    if (og_hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream), "hipMemcpy H2D"))
        return -1;
    return og_hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
}

int ogpu_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (og_ctx_ensure()) return -1;
    if (og_hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream), "hipMemcpy D2H"))
        return -1;
    return og_hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
}

int og_launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s);

int ogpu_device_copy(void *d_dst, const void *d_src, size_t bytes)
{
    if (og_ctx_ensure()) return -1;
    if (bytes % 16)
    {
        og_set_error("ogpu_device_copy: size %zu is not a multiple of 16", bytes);
        return -1;
    }
    if (og_launch_copy(d_src, d_dst, bytes, g_stream)) return -1;
    return og_hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
}

int ogpu_synchronize(void)
{
    if (og_ctx_ensure()) return -1;
    if (og_hip_check(hipStreamSynchronize(g_stream), "hipStreamSynchronize")) return -1;
    return og_hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
