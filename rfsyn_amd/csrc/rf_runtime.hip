// HIP runtime wrappers behind the plain-C interface of rf_device.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "rf_device.h"
#include "rf_hip.h"

struct rf_graph { hipGraphExec_t exec; hipGraph_t graph; };

static thread_local char g_err[256];

static int fail(hipError_t e, const char *what) {
    std::snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

#define RF_CHECK(call)                                     \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(e_, #call);      \
    } while (0)

static hipStream_t raw(rf_stream_t *s) { return rf_raw_stream(s); }

extern "C" {

const char *rf_error_string(void) { return g_err; }

int rf_init(void) {
    static int done = 0;
    if (done) return 0;
    int n = 0;
    RF_CHECK(hipGetDeviceCount(&n));
    if (n < 1) {
        std::snprintf(g_err, sizeof g_err, "no HIP device visible");
        return -1;
    }
    RF_CHECK(hipSetDevice(0));
    done = 1;
    return 0;
}

int rf_malloc(void **p, size_t bytes) {
    *p = nullptr;
    if (rf_init() != 0) return -1;
    if (bytes == 0) bytes = 256;
    RF_CHECK(hipMalloc(p, bytes));
    // The library's streams are non-blocking and do not wait for the null
    // stream, so the fill must be complete before the buffer is handed out:
    // a fill still queued behind earlier large fills would otherwise land
    // after the caller's first upload into the buffer and erase it.
    hipError_t e = hipMemsetAsync(*p, 0, bytes, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        (void)hipFree(*p);
        *p = nullptr;
        return fail(e, "hipMemset");
    }
    return 0;
}

void rf_free(void *p) {
    if (p) (void)hipFree(p);
}

int rf_host_alloc(void **p, size_t bytes) {
    *p = nullptr;
    if (rf_init() != 0) return -1;
    if (bytes == 0) bytes = 64;
    RF_CHECK(hipHostMalloc(p, bytes, hipHostMallocDefault));
    std::memset(*p, 0, bytes);
    return 0;
}

void rf_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

int rf_h2d(void *dst, const void *src, size_t bytes, rf_stream_t *s) {
    if (!bytes) return 0;
    RF_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, raw(s)));
    return 0;
}

int rf_d2h(void *dst, const void *src, size_t bytes, rf_stream_t *s) {
    if (!bytes) return 0;
    RF_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, raw(s)));
    RF_CHECK(hipStreamSynchronize(raw(s)));
    return 0;
}

int rf_d2d(void *dst, const void *src, size_t bytes, rf_stream_t *s) {
    if (!bytes) return 0;
    RF_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, raw(s)));
    return 0;
}

int rf_memset(void *dst, int v, size_t bytes, rf_stream_t *s) {
    if (!bytes) return 0;
    RF_CHECK(hipMemsetAsync(dst, v, bytes, raw(s)));
    return 0;
}

rf_stream_t *rf_stream_create(void) {
    if (rf_init() != 0) return nullptr;
    rf_stream_t *s = new rf_stream;
    // Default priority only: differing stream priorities are not used.
    if (hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return nullptr;
    }
    return s;
}

void rf_stream_destroy(rf_stream_t *s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->s);
    (void)hipStreamDestroy(s->s);
    delete s;
}

int rf_sync(rf_stream_t *s) {
    RF_CHECK(hipStreamSynchronize(raw(s)));
    return 0;
}

int rf_graph_begin(rf_stream_t *s) {
    RF_CHECK(hipStreamBeginCapture(raw(s), hipStreamCaptureModeThreadLocal));
    return 0;
}

int rf_graph_end(rf_stream_t *s, rf_graph_t **out) {
    *out = nullptr;
    hipGraph_t g = nullptr;
    RF_CHECK(hipStreamEndCapture(raw(s), &g));
    rf_graph *r = new rf_graph;
    r->graph = g;
    hipError_t e = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete r;
        return fail(e, "hipGraphInstantiate");
    }
    *out = r;
    return 0;
}

int rf_graph_launch(rf_graph_t *g, rf_stream_t *s) {
    RF_CHECK(hipGraphLaunch(g->exec, raw(s)));
    return 0;
}

void rf_graph_destroy(rf_graph_t *g) {
    if (!g) return;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
}

}  // extern "C"

// ------------------------------------------------------------ events
struct rf_event { hipEvent_t e; };

extern "C" {

rf_event_t *rf_event_create(void) {
    if (rf_init() != 0) return nullptr;
    rf_event_t *e = new rf_event;
    if (hipEventCreate(&e->e) != hipSuccess) {
        delete e;
        return nullptr;
    }
    return e;
}

void rf_event_destroy(rf_event_t *e) {
    if (!e) return;
    (void)hipEventDestroy(e->e);
    delete e;
}

int rf_event_record(rf_event_t *e, rf_stream_t *s) {
    RF_CHECK(hipEventRecord(e->e, raw(s)));
    return 0;
}

rf_event_t *rf_event_create_sync(void) {
    if (rf_init() != 0) return nullptr;
    rf_event_t *e = new rf_event;
    if (hipEventCreateWithFlags(&e->e, hipEventDisableTiming) != hipSuccess) {
        delete e;
        return nullptr;
    }
    return e;
}

int rf_stream_wait(rf_stream_t *s, rf_event_t *e) {
    RF_CHECK(hipStreamWaitEvent(raw(s), e->e, 0));
    return 0;
}

int rf_event_elapsed_ms(rf_event_t *a, rf_event_t *b, float *ms) {
    RF_CHECK(hipEventElapsedTime(ms, a->e, b->e));
    return 0;
}

}  // extern "C"
