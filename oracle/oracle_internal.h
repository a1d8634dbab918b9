/*
 * Internal helpers shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see oracle_core.c).
 */
#ifndef RFSYN_ORACLE_INTERNAL_H
#define RFSYN_ORACLE_INTERNAL_H

#include "grid.h"
#include "cpml.h"

/*
 * Component-generic stencil bookkeeping.  For component c the two
 * "curl" axes are p = (c+1)%3 and q = (c+2)%3.  With F the opposite
 * field family:
 *   H_c update:  dA = F_q(+1 along p) - F_q,  dB = F_p(+1 along q) - F_p
 *   E_c update:  dA = F_q - F_q(-1 along p),  dB = F_p - F_p(-1 along q)
 * which is grid.c:67-89 / 95-117 written once for the three components.
 */
typedef struct {
    size_t n[3];
    size_t stride[3];
    unsigned per;
} oracle_geom_t;

static inline oracle_geom_t geom_of(const fdtd_grid_t *g) {
    oracle_geom_t o;
    o.n[0] = g->nx; o.n[1] = g->ny; o.n[2] = g->nz;
    o.stride[2] = 1;
    o.stride[1] = g->nz + 1;
    o.stride[0] = (g->ny + 1) * (g->nz + 1);
    o.per = g->periodic_axes;
    return o;
}

static inline size_t wrap_up(const oracle_geom_t *o, int ax, size_t q) {
    size_t r = q + 1;
    if ((o->per >> ax) & 1u) r %= o->n[ax];
    return r;
}

static inline size_t wrap_down(const oracle_geom_t *o, int ax, size_t q) {
    (void)o;
    return q == 0 ? o->n[ax] - 1 : q - 1;
}

static inline size_t lin(const oracle_geom_t *o, const size_t x[3]) {
    return x[0] * o->stride[0] + x[1] * o->stride[1] + x[2] * o->stride[2];
}

/* Update ranges: H_c covers [0, n_c + !periodic_c) on its own axis and
 * [0, n_a) on the others (grid.c:67, 76, 85); E_c covers [0, n_c) on its
 * own axis and [!periodic_a, n_a) on the others (grid.c:95-97 ...). */
static inline void h_range(const oracle_geom_t *o, int c, size_t lo[3], size_t hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = 0;
        hi[a] = o->n[a] + ((a == c && !((o->per >> a) & 1u)) ? 1u : 0u);
    }
}

static inline void e_range(const oracle_geom_t *o, int c, size_t lo[3], size_t hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = (a == c || ((o->per >> a) & 1u)) ? 0u : 1u;
        hi[a] = o->n[a];
    }
}

static inline double *e_of(fdtd_grid_t *g, int c) { return c == 0 ? g->ex : c == 1 ? g->ey : g->ez; }
static inline double *h_of(fdtd_grid_t *g, int c) { return c == 0 ? g->hx : c == 1 ? g->hy : g->hz; }

typedef struct {
    const double *kinv, *b, *a;   /* profile along one axis */
    double *psi;                  /* full-size psi for this (component, axis) */
} oracle_stretch_t;

void oracle_cpml_stretch_h(const cpml_t *p, oracle_stretch_t st[3][2]);
void oracle_cpml_stretch_e(const cpml_t *p, oracle_stretch_t st[3][2]);

#endif
