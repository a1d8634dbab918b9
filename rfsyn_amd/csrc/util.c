/* Error reporting and a small hash map for the host side of librfsyn_fdtd. */
#include "host_internal.h"
#include "fdtd_native.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_last_error[512];

void rf_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
}

const char *fdtd_last_error(void) { return g_last_error; }

static int g_gpu_errors;

int fdtd_gpu_errors(void) { return __atomic_load_n(&g_gpu_errors, __ATOMIC_RELAXED); }

void rf_fatal(const char *what) {
    const char *mode = getenv("RFSYN_GPU_ERRORS");
    if (mode && !strcmp(mode, "record")) {
        char gpu[256];
        snprintf(gpu, sizeof gpu, "%s", rf_error_string());
        rf_set_error("%s failed on the GPU: %s", what, gpu);
        __atomic_add_fetch(&g_gpu_errors, 1, __ATOMIC_RELAXED);
        return;
    }
    fprintf(stderr, "librfsyn_fdtd: %s failed on the GPU: %s (%s)\n", what, rf_error_string(),
            g_last_error);
    fflush(stderr);
    abort();
}

/* ------------------------------------------------------------ hash map */

static size_t mix(size_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

#define EMPTY_KEY ((size_t)-1)

static int rehash(rf_map_t *m, size_t cap) {
    size_t *keys = malloc(cap * sizeof *keys);
    uint32_t *vals = malloc(cap * sizeof *vals);
    if (!keys || !vals) { free(keys); free(vals); return -1; }
    for (size_t q = 0; q < cap; q++) keys[q] = EMPTY_KEY;
    for (size_t q = 0; q < m->cap; q++) {
        if (m->keys[q] == EMPTY_KEY) continue;
        size_t h = mix(m->keys[q]) & (cap - 1);
        while (keys[h] != EMPTY_KEY) h = (h + 1) & (cap - 1);
        keys[h] = m->keys[q];
        vals[h] = m->vals[q];
    }
    free(m->keys);
    free(m->vals);
    m->keys = keys;
    m->vals = vals;
    m->cap = cap;
    return 0;
}

int rf_map_put(rf_map_t *m, size_t key, uint32_t val) {
    if ((m->count + 1) * 2 > m->cap && rehash(m, m->cap ? m->cap * 2 : 64) != 0) return -1;
    size_t h = mix(key) & (m->cap - 1);
    while (m->keys[h] != EMPTY_KEY && m->keys[h] != key) h = (h + 1) & (m->cap - 1);
    if (m->keys[h] == EMPTY_KEY) m->count++;
    m->keys[h] = key;
    m->vals[h] = val;
    return 0;
}

int rf_map_get(const rf_map_t *m, size_t key, uint32_t *val) {
    if (!m->count) return 0;
    size_t h = mix(key) & (m->cap - 1);
    while (m->keys[h] != EMPTY_KEY) {
        if (m->keys[h] == key) {
            if (val) *val = m->vals[h];
            return 1;
        }
        h = (h + 1) & (m->cap - 1);
    }
    return 0;
}

int rf_map_del(rf_map_t *m, size_t key) {
    if (!m->count) return 0;
    size_t mask = m->cap - 1, h = mix(key) & mask;
    while (m->keys[h] != EMPTY_KEY && m->keys[h] != key) h = (h + 1) & mask;
    if (m->keys[h] == EMPTY_KEY) return 0;
    /* backward-shift deletion keeps probe chains intact */
    size_t hole = h, q = (h + 1) & mask;
    while (m->keys[q] != EMPTY_KEY) {
        size_t home = mix(m->keys[q]) & mask;
        int movable = (hole <= q) ? (home <= hole || home > q) : (home <= hole && home > q);
        if (movable) {
            m->keys[hole] = m->keys[q];
            m->vals[hole] = m->vals[q];
            hole = q;
        }
        q = (q + 1) & mask;
    }
    m->keys[hole] = EMPTY_KEY;
    m->count--;
    return 1;
}

void rf_map_free(rf_map_t *m) {
    free(m->keys);
    free(m->vals);
    m->keys = NULL;
    m->vals = NULL;
    m->cap = m->count = 0;
}
