/*
 * Grid lifetime and the vacuum H/E updates of the reference API.
 *
 * Replaces engine/fdtd/grid.c.  The six FP64 arrays stay in host memory
 * (the reference exposes them: grid.h:11-12) and each update call mirrors
 * them to the GPU, runs the HIP kernel and copies the updated family back.
 */
#include "host_internal.h"
#include "constants.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* engine/fdtd/grid.c:7-46 */
fdtd_grid_t *fdtd_grid_create(size_t nx, size_t ny, size_t nz, double dx,
                              double courant_safety_factor) {
    if (!nx || !ny || !nz) return NULL;
    if (!(isfinite(dx) && dx > 0.0)) return NULL;
    if (!(isfinite(courant_safety_factor) && courant_safety_factor > 0.0 &&
          courant_safety_factor <= 1.0)) return NULL;
    if (nx == SIZE_MAX || ny == SIZE_MAX || nz == SIZE_MAX) return NULL;
    size_t a = nx + 1, b = ny + 1, c = nz + 1;
    if (a > SIZE_MAX / b || a * b > SIZE_MAX / c) return NULL;
    size_t cells = a * b * c;
    /* dt = S * dx / (c0 sqrt 3), grid.c:27-28 */
    double dt = courant_safety_factor * (dx / (C0 * sqrt(3.0)));
    if (cells > SIZE_MAX / sizeof(double) || !isfinite(dt) || dt <= 0.0) return NULL;
    /* the kernels index with 32-bit extents */
    if (nx >= (1u << 30) || ny >= (1u << 30) || nz >= (1u << 30)) return NULL;

    fdtd_grid_t *g = calloc(1, sizeof *g);
    if (!g) return NULL;
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->dx = dx; g->dt = dt;
    double **f[6] = {&g->ex, &g->ey, &g->ez, &g->hx, &g->hy, &g->hz};
    for (int q = 0; q < 6; q++) {
        *f[q] = calloc(cells, sizeof(double));
        if (!*f[q]) { fdtd_grid_destroy(g); return NULL; }
    }
    return g;
}

void fdtd_grid_destroy(fdtd_grid_t *g) {
    if (!g) return;
    if (g->dev) {
        for (int q = 0; q < 6; q++) rf_free(g->dev->f[q]);
        rf_stream_destroy(g->dev->st);
        free(g->dev);
    }
    free(g->ex); free(g->ey); free(g->ez);
    free(g->hx); free(g->hy); free(g->hz);
    free(g);
}

/* engine/fdtd/grid.c:126-130 */
int fdtd_grid_set_periodic(fdtd_grid_t *g, unsigned axes) {
    if (!g || (axes & ~7u)) return -1;
    g->periodic_axes = axes;
    return 0;
}

rf_geom_t rf_geom_of(const fdtd_grid_t *g) {
    rf_geom_t o;
    o.n[0] = (int)g->nx; o.n[1] = (int)g->ny; o.n[2] = (int)g->nz;
    o.per = (int)g->periodic_axes;
    o.sy = (long long)g->nz + 1;
    o.sx = ((long long)g->ny + 1) * ((long long)g->nz + 1);
    return o;
}

size_t rf_cells(const fdtd_grid_t *g) { return (g->nx + 1) * (g->ny + 1) * (g->nz + 1); }

struct fdtd_grid_device *rf_grid_dev(fdtd_grid_t *g) {
    if (g->dev) return g->dev;
    struct fdtd_grid_device *d = calloc(1, sizeof *d);
    if (!d) return NULL;
    d->st = rf_stream_create();
    int ok = d->st != NULL;
    for (int q = 0; q < 6 && ok; q++) ok = rf_malloc(&d->f[q], rf_cells(g) * sizeof(double)) == 0;
    if (!ok) {
        for (int q = 0; q < 6; q++) rf_free(d->f[q]);
        rf_stream_destroy(d->st);
        free(d);
        return NULL;
    }
    g->dev = d;
    return d;
}

static double *host_field(fdtd_grid_t *g, int q) {
    double *f[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
    return f[q];
}

int rf_grid_upload(fdtd_grid_t *g, unsigned mask) {
    struct fdtd_grid_device *d = rf_grid_dev(g);
    if (!d) return -1;
    size_t bytes = rf_cells(g) * sizeof(double);
    for (int q = 0; q < 6; q++)
        if (((mask >> q) & 1u) && rf_h2d(d->f[q], host_field(g, q), bytes, d->st) != 0) return -1;
    return 0;
}

int rf_grid_download(fdtd_grid_t *g, unsigned mask) {
    struct fdtd_grid_device *d = g->dev;
    if (!d) return -1;
    size_t bytes = rf_cells(g) * sizeof(double);
    for (int q = 0; q < 6; q++)
        if (((mask >> q) & 1u) && rf_d2h(host_field(g, q), d->f[q], bytes, d->st) != 0) return -1;
    return 0;
}

/* engine/fdtd/grid.c:62-90 */
void fdtd_update_h(fdtd_grid_t *g) {
    if (!g) return;
    if (rf_grid_upload(g, 077u) != 0) RF_FAIL("fdtd_update_h upload");
    rf_geom_t geo = rf_geom_of(g);
    double cb = g->dt / (MU0 * g->dx);
    if (rf_launch_update_h(0, &geo, g->dev->f, cb, g->dx, NULL, g->dev->st) != 0)
        RF_FAIL("fdtd_update_h launch");
    if (rf_grid_download(g, 070u) != 0) RF_FAIL("fdtd_update_h download");
}

/* engine/fdtd/grid.c:92-118.  E += ca (curl) is the materials update with
 * Ca = 1, Cb = ca, which rounds identically (1*E is exact). */
void fdtd_update_e(fdtd_grid_t *g) {
    if (!g) return;
    if (rf_grid_upload(g, 077u) != 0) RF_FAIL("fdtd_update_e upload");
    rf_geom_t geo = rf_geom_of(g);
    rf_mat_t vac = {0};
    vac.vac_ca = g->dt / (EPS0 * g->dx);
    if (rf_launch_update_e(0, &geo, g->dev->f, g->dx, NULL, &vac, g->dev->st) != 0)
        RF_FAIL("fdtd_update_e launch");
    if (rf_grid_download(g, 007u) != 0) RF_FAIL("fdtd_update_e download");
}
