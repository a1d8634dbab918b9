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

/* Largest [plo, phi) inside the non-slab range [lo, hi) where kinv of both
 * grids is exactly 1 in the device precision (empty if a kinv != 1 sits
 * strictly inside it). */
static int kinv_is_one(const double *k, size_t q, int prec) {
    return prec ? (float)k[q] == 1.0f : k[q] == 1.0;
}

static void plain_range(const double *km, const double *kh, size_t n, int prec, int lo, int hi, int *plo,
                        int *phi) {
    int a = lo < 0 ? 0 : lo, b = hi > (int)n + 1 ? (int)n + 1 : hi;
#define RF_PLAIN(q) (kinv_is_one(km, (size_t)(q), prec) && ((size_t)(q) >= n || kinv_is_one(kh, (size_t)(q), prec)))
    while (a < b && !RF_PLAIN(a)) a++;
    while (b > a && !RF_PLAIN(b - 1)) b--;
    for (int q = a; q < b; q++)
        if (!RF_PLAIN(q)) { a = b = 0; break; }
#undef RF_PLAIN
    *plo = a;
    *phi = b;
}

static int push_owned(rf_cpml_dev_t *c, void *p) {
    if (c->n_owned >= 64) return -1;
    c->owned[c->n_owned++] = p;
    return 0;
}

/* Offset of psi memory (comp, s) inside one psi section, in elements;
 * sec[a] is the size of one slab array for derivative axis a.  Order
 * (0,0) (0,1) (1,0) (1,1) (2,0) (2,1), derivative axes 1 2 2 0 0 1. */
static long long psi_offset(const long long sec[3], int comp, int s) {
    const long long o[6] = {0, sec[1], sec[1] + sec[2], sec[1] + 2 * sec[2], sec[0] + sec[1] + 2 * sec[2],
                            2 * sec[0] + sec[1] + 2 * sec[2]};
    return o[2 * comp + s];
}

int rf_cpml_build(const cpml_t *p, int prec, rf_cpml_dev_t *c) {
    memset(c, 0, sizeof *c);
    size_t n[3] = {p->nx, p->ny, p->nz};
    cpml_t *pp = (cpml_t *)p;
    size_t es = prec ? sizeof(float) : sizeof(double);

    /* one coefficient block: per axis the six profiles kinv_m b_m a_m
     * kinv_h b_h a_h, n+1 values each (the half-node ones padded) */
    size_t tot = 0, coff[3];
    for (int ax = 0; ax < 3; ax++) {
        coff[ax] = tot;
        tot += 6 * (n[ax] + 1);
    }
    double *h64 = calloc(tot, sizeof *h64);
    if (!h64) return -1;
    for (int ax = 0; ax < 3; ax++)
        for (int prof = 0; prof < 6; prof++) {
            const double *src = *profile_slot(pp, prof / 3, prof % 3, ax);
            size_t cnt = prof < 3 ? n[ax] + 1 : n[ax];
            memcpy(h64 + coff[ax] + prof * (n[ax] + 1), src, cnt * sizeof(double));
        }
    void *coef = NULL;
    int rc = rf_malloc(&coef, tot * es) || push_owned(c, coef);
    if (!rc) {
        if (prec == 0) {
            rc = rf_h2d(coef, h64, tot * sizeof(double), NULL) || rf_sync(NULL);
        } else {
            float *f32 = malloc(tot * sizeof *f32);
            if (!f32) rc = -1;
            else {
                for (size_t q = 0; q < tot; q++) f32[q] = (float)h64[q];
                rc = rf_h2d(coef, f32, tot * sizeof(float), NULL) || rf_sync(NULL);
                free(f32);
            }
        }
    }
    if (rc) {
        free(h64);
        rf_cpml_release(c);
        return -1;
    }
    c->bytes += tot * es;
    char *cb = coef;
    for (int ax = 0; ax < 3; ax++) {
        size_t st = (n[ax] + 1) * es;
        c->d.kinv_m[ax] = cb + coff[ax] * es;
        c->d.b_m[ax] = cb + coff[ax] * es + st;
        c->d.a_m[ax] = cb + coff[ax] * es + 2 * st;
        c->d.kinv_h[ax] = cb + coff[ax] * es + 3 * st;
        c->d.b_h[ax] = cb + coff[ax] * es + 4 * st;
        c->d.a_h[ax] = cb + coff[ax] * es + 5 * st;
        const double *km = h64 + coff[ax], *am = h64 + coff[ax] + 2 * (n[ax] + 1);
        const double *kh = h64 + coff[ax] + 3 * (n[ax] + 1), *ah = h64 + coff[ax] + 5 * (n[ax] + 1);
        slab_bounds(am, ah, n[ax], &c->d.lo[ax], &c->d.hi[ax]);
        c->d.len[ax] = c->d.lo[ax] + ((int)n[ax] + 1 - c->d.hi[ax]);
        plain_range(km, kh, n[ax], prec, c->d.lo[ax], c->d.hi[ax], &c->d.plo[ax], &c->d.phi[ax]);
    }
    free(h64);

    /* psi: one zero-filled section each for E and H */
    long long sec[3];
    for (int ax = 0; ax < 3; ax++) {
        long long ext[3] = {(long long)n[0] + 1, (long long)n[1] + 1, (long long)n[2] + 1};
        ext[ax] = c->d.len[ax];
        sec[ax] = ext[0] * ext[1] * ext[2];
    }
    size_t psec = (size_t)(2 * (sec[0] + sec[1] + sec[2]));
    void *pe = NULL, *ph = NULL;
    if (rf_malloc(&pe, psec * es) || push_owned(c, pe) || rf_malloc(&ph, psec * es) || push_owned(c, ph)) {
        rf_cpml_release(c);
        return -1;
    }
    c->bytes += 2 * psec * es;
    c->psi_sec[0] = pe;
    c->psi_sec[1] = ph;
    c->psi_sec_bytes = psec * es;
    for (int comp = 0; comp < 3; comp++)
        for (int s2 = 0; s2 < 2; s2++) {
            size_t off = (size_t)psi_offset(sec, comp, s2) * es;
            c->d.psi_e[comp][s2] = (char *)pe + off;
            c->d.psi_h[comp][s2] = (char *)ph + off;
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
    if (rf_cpml_build(p, 0, c) != 0) { free(c); return NULL; }
    p->dev = (struct cpml_device *)c;
    return c;
}

/* engine/fdtd/cpml.c:176-219 */
void fdtd_update_h_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml || !rf_cpml_matches(pml, g)) return;
    rf_cpml_dev_t *c = rf_cpml_dev(pml);
    if (!c || rf_grid_upload(g, 077u) != 0) RF_FAIL("fdtd_update_h_cpml setup");
    rf_geom_t geo = rf_geom_of(g);
    double cb = g->dt / (MU0 * g->dx);
    if (rf_launch_update_h(0, &geo, g->dev->f, cb, g->dx, &c->d, g->dev->st) != 0)
        RF_FAIL("fdtd_update_h_cpml launch");
    if (rf_grid_download(g, 070u) != 0) RF_FAIL("fdtd_update_h_cpml download");
}

/* engine/fdtd/cpml.c:221-264 */
void fdtd_update_e_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml || !rf_cpml_matches(pml, g)) return;
    rf_cpml_dev_t *c = rf_cpml_dev(pml);
    if (!c || rf_grid_upload(g, 077u) != 0) RF_FAIL("fdtd_update_e_cpml setup");
    rf_geom_t geo = rf_geom_of(g);
    rf_mat_t vac = {0};
    vac.vac_ca = g->dt / (EPS0 * g->dx);
    if (rf_launch_update_e(0, &geo, g->dev->f, g->dx, &c->d, &vac, g->dev->st) != 0)
        RF_FAIL("fdtd_update_e_cpml launch");
    if (rf_grid_download(g, 007u) != 0) RF_FAIL("fdtd_update_e_cpml download");
}
