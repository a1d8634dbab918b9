/*
 * CPML profiles (host, FP64) and their device form.
 *
 * Replaces engine/fdtd/cpml.c.  The grading profiles are computed on the
 * host with the reference formulas (cpml.c:44-84, 103-105) so the libm
 * results are the reference bits.  On the device the psi memories exist
 * only in the slabs where a profile's a-coefficient is nonzero: outside
 * them the reference's full-size psi stays exactly zero (b*0 + 0*d), so
 * the stretched difference is d/kappa + dx*0 either way.
 */
#include "host_internal.h"
#include "constants.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* engine/fdtd/cpml.c:10-19 */
cpml_params_t cpml_default_params(size_t npml) {
    cpml_params_t p = {npml, 3.0, 1.0, 7.0, 0.05, 1e-6};
    return p;
}

/* One profile point at position x along an axis of length len with PML
 * depth d (cpml.c:46-61). */
static void grade(double x, double d, double len, const cpml_params_t *pp, double sigma_max,
                  double dx, double dt, double *kinv, double *b, double *a) {
    double rho = -1.0;
    if (x < d) rho = d - x;
    else if (x > len - d) rho = x - (len - d);
    if (rho < 0.0) { *kinv = 1.0; *b = 1.0; *a = 0.0; return; }
    double r = rho / d;
    if (r > 1.0) r = 1.0;
    double sigma = sigma_max * pow(r, pp->m);
    double kappa = 1.0 + (pp->kappa_max - 1.0) * pow(r, pp->m);
    double alpha = pp->alpha_max * pow(1.0 - r, pp->ma);
    double bb = exp(-(sigma / kappa + alpha) * dt / EPS0);
    *kinv = 1.0 / kappa;
    *b = bb;
    *a = sigma > 0.0 ? sigma * (bb - 1.0) / (dx * kappa * (sigma + kappa * alpha)) : 0.0;
}

static double **profile_slot(cpml_t *p, int half, int kind, int ax) {
    double **m[3][3] = {{&p->kappa_inv_m_x, &p->b_m_x, &p->a_m_x},
                        {&p->kappa_inv_m_y, &p->b_m_y, &p->a_m_y},
                        {&p->kappa_inv_m_z, &p->b_m_z, &p->a_m_z}};
    double **h[3][3] = {{&p->kappa_inv_h_x, &p->b_h_x, &p->a_h_x},
                        {&p->kappa_inv_h_y, &p->b_h_y, &p->a_h_y},
                        {&p->kappa_inv_h_z, &p->b_h_z, &p->a_h_z}};
    return half ? h[ax][kind] : m[ax][kind];
}

/* engine/fdtd/cpml.c:91-150 */
cpml_t *cpml_create(const fdtd_grid_t *g, cpml_params_t prm) {
    if (!g || prm.npml == 0) return NULL;
    if (!(isfinite(prm.r0) && prm.r0 > 0.0 && prm.r0 < 1.0)) return NULL;
    if (!(isfinite(prm.m) && prm.m > 0.0)) return NULL;
    if (!(isfinite(prm.ma) && prm.ma >= 0.0)) return NULL;
    if (!(isfinite(prm.kappa_max) && prm.kappa_max >= 1.0)) return NULL;
    if (!(isfinite(prm.alpha_max) && prm.alpha_max >= 0.0)) return NULL;
    size_t n[3] = {g->nx, g->ny, g->nz};
    for (int ax = 0; ax < 3; ax++)
        if (!((g->periodic_axes >> ax) & 1u) && prm.npml >= (n[ax] + 1) / 2) return NULL;

    cpml_t *p = calloc(1, sizeof *p);
    if (!p) return NULL;
    p->npml = prm.npml;
    p->dx = g->dx;
    p->nx = g->nx; p->ny = g->ny; p->nz = g->nz;
    p->periodic_axes = g->periodic_axes;
    p->dt = g->dt;

    double eta0 = sqrt(MU0 / EPS0);
    double sigma_max = -(prm.m + 1.0) * log(prm.r0) / (2.0 * eta0 * (double)prm.npml * g->dx);
    for (int ax = 0; ax < 3; ax++) {
        int periodic = (g->periodic_axes >> ax) & 1u;
        double d = (double)prm.npml * g->dx, len = (double)n[ax] * g->dx;
        for (int half = 0; half < 2; half++) {
            size_t cnt = half ? n[ax] : n[ax] + 1;
            double *arr[3];
            for (int kind = 0; kind < 3; kind++) {
                arr[kind] = malloc((cnt ? cnt : 1) * sizeof(double));
                *profile_slot(p, half, kind, ax) = arr[kind];
                if (!arr[kind]) { cpml_destroy(p); return NULL; }
            }
            for (size_t q = 0; q < cnt; q++) {
                if (periodic) { arr[0][q] = 1.0; arr[1][q] = 1.0; arr[2][q] = 0.0; continue; }
                double x = half ? ((double)q + 0.5) * g->dx : (double)q * g->dx;
                grade(x, d, len, &prm, sigma_max, g->dx, g->dt, &arr[0][q], &arr[1][q], &arr[2][q]);
            }
        }
    }
    return p;
}

void cpml_destroy(cpml_t *p) {
    if (!p) return;
    for (int half = 0; half < 2; half++)
        for (int kind = 0; kind < 3; kind++)
            for (int ax = 0; ax < 3; ax++) free(*profile_slot(p, half, kind, ax));
    if (p->dev) {
        rf_cpml_release((rf_cpml_dev_t *)p->dev);
        free(p->dev);
    }
    free(p);
}

int rf_cpml_matches(const cpml_t *p, const fdtd_grid_t *g) {
    return p && g && p->nx == g->nx && p->ny == g->ny && p->nz == g->nz &&
           p->periodic_axes == g->periodic_axes && p->dx == g->dx;
}

/* Slab bounds along one axis: the union of the integer- and half-node
 * profiles' nonzero a-coefficients, split at the middle of the axis. */
static void slab_bounds(const double *am, const double *ah, size_t n, int *lo, int *hi) {
    size_t mid = (n + 1) / 2;
    long last_low = -1;
    size_t first_high = n + 1;
    for (size_t q = 0; q <= n; q++) {
        int nz = am[q] != 0.0 || (q < n && ah[q] != 0.0);
        if (!nz) continue;
        if (q < mid) last_low = (long)q;
        else if (q < first_high) first_high = q;
    }
    *lo = (int)(last_low + 1);
    *hi = (int)first_high;
}

static int push_owned(rf_cpml_dev_t *c, void *p) {
    if (c->n_owned >= 64) return -1;
    c->owned[c->n_owned++] = p;
    return 0;
}

static int upload_profile(rf_cpml_dev_t *c, const double *src, size_t n, int prec, const void **out) {
    void *d = NULL;
    size_t es = prec ? sizeof(float) : sizeof(double);
    if (rf_malloc(&d, (n ? n : 1) * es) != 0 || push_owned(c, d) != 0) return -1;
    c->bytes += n * es;
    if (prec == 0) {
        if (rf_h2d(d, src, n * sizeof(double), NULL) != 0 || rf_sync(NULL) != 0) return -1;
    } else {
        float *tmp = malloc((n ? n : 1) * sizeof(float));
        if (!tmp) return -1;
        for (size_t q = 0; q < n; q++) tmp[q] = (float)src[q];
        int rc = rf_h2d(d, tmp, n * sizeof(float), NULL);
        rf_sync(NULL);
        free(tmp);
        if (rc != 0) return -1;
    }
    *out = d;
    return 0;
}

int rf_cpml_build(const cpml_t *p, int prec, int with_h2, rf_cpml_dev_t *c) {
    memset(c, 0, sizeof *c);
    size_t n[3] = {p->nx, p->ny, p->nz};
    cpml_t *pp = (cpml_t *)p;
    for (int ax = 0; ax < 3; ax++) {
        const double *km = *profile_slot(pp, 0, 0, ax), *bm = *profile_slot(pp, 0, 1, ax),
                     *am = *profile_slot(pp, 0, 2, ax);
        const double *kh = *profile_slot(pp, 1, 0, ax), *bh = *profile_slot(pp, 1, 1, ax),
                     *ah = *profile_slot(pp, 1, 2, ax);
        if (upload_profile(c, km, n[ax] + 1, prec, &c->d.kinv_m[ax]) ||
            upload_profile(c, bm, n[ax] + 1, prec, &c->d.b_m[ax]) ||
            upload_profile(c, am, n[ax] + 1, prec, &c->d.a_m[ax]) ||
            upload_profile(c, kh, n[ax], prec, &c->d.kinv_h[ax]) ||
            upload_profile(c, bh, n[ax], prec, &c->d.b_h[ax]) ||
            upload_profile(c, ah, n[ax], prec, &c->d.a_h[ax])) {
            rf_cpml_release(c);
            return -1;
        }
        slab_bounds(am, ah, n[ax], &c->d.lo[ax], &c->d.hi[ax]);
        c->d.len[ax] = c->d.lo[ax] + ((int)n[ax] + 1 - c->d.hi[ax]);
    }
    size_t es = prec ? sizeof(float) : sizeof(double);
    for (int comp = 0; comp < 3; comp++)
        for (int s = 0; s < 2; s++) {
            int ax = (comp + 1 + s) % 3;
            size_t ext[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
            ext[ax] = (size_t)c->d.len[ax];
            size_t cnt = ext[0] * ext[1] * ext[2];
            if (!cnt) continue;
            void *h = NULL, *h2 = NULL, *e = NULL;
            if (rf_malloc(&h, cnt * es) || push_owned(c, h) || rf_malloc(&e, cnt * es) || push_owned(c, e)) {
                rf_cpml_release(c);
                return -1;
            }
            c->bytes += 2 * cnt * es;
            if (with_h2) {   /* second H-psi set for the double-buffered fused step */
                if (rf_malloc(&h2, cnt * es) || push_owned(c, h2)) {
                    rf_cpml_release(c);
                    return -1;
                }
                c->bytes += cnt * es;
            }
            c->d.psi_h[comp][s] = h;
            c->psi_h2[comp][s] = h2;
            c->d.psi_e[comp][s] = e;
        }
    return 0;
}

void rf_cpml_release(rf_cpml_dev_t *c) {
    for (int q = 0; q < c->n_owned; q++) rf_free(c->owned[q]);
    c->n_owned = 0;
}

rf_cpml_dev_t *rf_cpml_dev(cpml_t *p) {
    if (p->dev) return (rf_cpml_dev_t *)p->dev;
    rf_cpml_dev_t *c = calloc(1, sizeof *c);
    if (!c) return NULL;
    if (rf_cpml_build(p, 0, 0, c) != 0) { free(c); return NULL; }
    p->dev = (struct cpml_device *)c;
    return c;
}

/* engine/fdtd/cpml.c:176-219 */
void fdtd_update_h_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml || !rf_cpml_matches(pml, g)) return;
    rf_cpml_dev_t *c = rf_cpml_dev(pml);
    if (!c || rf_grid_upload(g, 077u) != 0) rf_fatal("fdtd_update_h_cpml setup");
    rf_geom_t geo = rf_geom_of(g);
    double cb = g->dt / (MU0 * g->dx);
    if (rf_launch_update_h(0, &geo, g->dev->f, cb, g->dx, &c->d, g->dev->st) != 0)
        rf_fatal("fdtd_update_h_cpml launch");
    if (rf_grid_download(g, 070u) != 0) rf_fatal("fdtd_update_h_cpml download");
}

/* engine/fdtd/cpml.c:221-264 */
void fdtd_update_e_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml || !rf_cpml_matches(pml, g)) return;
    rf_cpml_dev_t *c = rf_cpml_dev(pml);
    if (!c || rf_grid_upload(g, 077u) != 0) rf_fatal("fdtd_update_e_cpml setup");
    rf_geom_t geo = rf_geom_of(g);
    rf_mat_t vac = {0};
    vac.vac_ca = g->dt / (EPS0 * g->dx);
    if (rf_launch_update_e(0, &geo, g->dev->f, g->dx, &c->d, &vac, g->dev->st) != 0)
        rf_fatal("fdtd_update_e_cpml launch");
    if (rf_grid_download(g, 007u) != 0) rf_fatal("fdtd_update_e_cpml download");
}
