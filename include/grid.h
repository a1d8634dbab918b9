/*
 * Yee grid of the rfsyn FDTD core (drop-in for engine/fdtd/grid.h).
 *
 * Replaces: engine/fdtd/grid.h:6-33 (struct, create/destroy, set_periodic,
 * fdtd_index, grid_previous, fdtd_update_h/e).
 *
 * Storage: six FP64 arrays, one per component, each over the padded box
 * (nx+1)*(ny+1)*(nz+1), z fastest.  The pointers are host memory that the
 * caller may read and write between update calls; every update call of
 * this library copies what it reads to the GPU, runs the HIP kernel and
 * copies what it wrote back, so host edits made between calls are seen.
 * The device-resident fast path is the fdtd_sim handle (fdtd_native.h).
 */
#ifndef RFSYN_FDTD_GRID_H
#define RFSYN_FDTD_GRID_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct fdtd_grid_device;

typedef struct {
    size_t nx, ny, nz;
    double dx;
    double dt;
    unsigned periodic_axes;          /* bit 0 = x, bit 1 = y, bit 2 = z */
    double *ex, *ey, *ez;
    double *hx, *hy, *hz;
    /* Library-private device mirror; not part of the reference layout
     * and never touched by callers. */
    struct fdtd_grid_device *dev;
} fdtd_grid_t;

/* engine/fdtd/grid.h:15, grid.c:7-46.  dt = S*dx/(c0*sqrt(3)); NULL on
 * invalid sizes, dx or S outside (0,1]. */
fdtd_grid_t *fdtd_grid_create(size_t nx, size_t ny, size_t nz, double dx,
                              double courant_safety_factor);

/* engine/fdtd/grid.h:17, grid.c:126-130.  Must be set before CPML and
 * materials are created.  -1 on NULL grid or bits other than 0..7. */
int fdtd_grid_set_periodic(fdtd_grid_t *g, unsigned axes);

/* engine/fdtd/grid.h:19, grid.c:48-53.  NULL-safe. */
void fdtd_grid_destroy(fdtd_grid_t *g);

/* engine/fdtd/grid.h:21-23: backward neighbour with wrap at 0. */
static inline size_t grid_previous(size_t q, size_t extent) {
    return q ? q - 1 : extent - 1;
}

/* engine/fdtd/grid.h:25-30: padded linear index; periodic axes wrap. */
static inline size_t fdtd_index(const fdtd_grid_t *g, size_t i, size_t j,
                                size_t k) {
    if (g->periodic_axes & 1u) i %= g->nx;
    if (g->periodic_axes & 2u) j %= g->ny;
    if (g->periodic_axes & 4u) k %= g->nz;
    return (i * (g->ny + 1) + j) * (g->nz + 1) + k;
}

/* engine/fdtd/grid.c:62-90: vacuum H update, H -= dt/(mu0 dx) * curl E. */
void fdtd_update_h(fdtd_grid_t *g);
/* engine/fdtd/grid.c:92-118: vacuum E update with PEC on non-periodic
 * boundary planes. */
void fdtd_update_e(fdtd_grid_t *g);

#ifdef __cplusplus
}
#endif

#endif
