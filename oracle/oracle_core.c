/*
 * CPU oracle for the rfsyn FDTD core: grid, vacuum updates, CPML, sources
 * and probes, in FP64 with the reference's exact operation order.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load the library built from this
 * directory, and only as the checker or the timed CPU baseline.  The
 * product (rfsyn_amd/) never links or calls it.
 *
 * Each function restates the reference algorithm cited beside it
 * (paths relative to the rfsyn repository).  Parity pins:
 *   - bitwise against the reference compiled from its own sources
 *     (oracle/Makefile target `ref`, tests/test_oracle_vs_ref.py) and
 *     against the committed golden vectors in tests/golden/;
 *   - dt and CPML profile closed forms (tests/test_oracle_closed_forms.py).
 *
 * Built with -O3 -ffp-contract=off so no multiply-add is fused, exactly
 * like the reference's x86-64 Release build.
 */
#include "grid.h"
#include "cpml.h"
#include "source.h"
#include "probe.h"
#include "constants.h"
#include "oracle_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ grid */

/* engine/fdtd/grid.c:7-46 */
fdtd_grid_t *fdtd_grid_create(size_t nx, size_t ny, size_t nz, double dx,
                              double courant_safety_factor) {
    if (nx == 0 || ny == 0 || nz == 0) return NULL;
    if (!(isfinite(dx) && dx > 0.0)) return NULL;
    if (!(isfinite(courant_safety_factor) && courant_safety_factor > 0.0 &&
          courant_safety_factor <= 1.0)) return NULL;
    if (nx == SIZE_MAX || ny == SIZE_MAX || nz == SIZE_MAX) return NULL;
    size_t a = nx + 1, b = ny + 1, c = nz + 1;
    if (a > SIZE_MAX / b || a * b > SIZE_MAX / c) return NULL;
    size_t cells = a * b * c;
    double dt = courant_safety_factor * (dx / (C0 * sqrt(3.0)));
    if (cells > SIZE_MAX / sizeof(double) || !isfinite(dt) || dt <= 0.0)
        return NULL;

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


/* Shared H sweep: vacuum (st == NULL) or CPML (st[c][0..1] for the p/q
 * derivatives).  Order of operations per cell follows grid.c:70 and
 * cpml.c:185-189. */
static void sweep_h(fdtd_grid_t *g, const oracle_stretch_t st[3][2]) {
    oracle_geom_t o = geom_of(g);
    const double cb = g->dt / (MU0 * g->dx);
    const double dx = g->dx;
    for (int c = 0; c < 3; c++) {
        int p = (c + 1) % 3, q = (c + 2) % 3;
        double *H = h_of(g, c);
        const double *Fq = e_of(g, q), *Fp = e_of(g, p);
        size_t lo[3], hi[3];
        h_range(&o, c, lo, hi);
        size_t x[3];
        for (x[0] = lo[0]; x[0] < hi[0]; x[0]++)
        for (x[1] = lo[1]; x[1] < hi[1]; x[1]++)
        for (x[2] = lo[2]; x[2] < hi[2]; x[2]++) {
            size_t idx = lin(&o, x);
            size_t xp[3] = {x[0], x[1], x[2]}, xq[3] = {x[0], x[1], x[2]};
            xp[p] = wrap_up(&o, p, x[p]);
            xq[q] = wrap_up(&o, q, x[q]);
            double dA = Fq[lin(&o, xp)] - Fq[idx];
            double dB = Fp[lin(&o, xq)] - Fp[idx];
            if (st) {
                const oracle_stretch_t *sa = &st[c][0], *sb = &st[c][1];
                size_t ia = x[p], ib = x[q];
                dA = cpml_term(dA, sa->kinv[ia], sa->b[ia], sa->a[ia], dx, &sa->psi[idx]);
                dB = cpml_term(dB, sb->kinv[ib], sb->b[ib], sb->a[ib], dx, &sb->psi[idx]);
            }
            H[idx] -= cb * (dA - dB);
        }
    }
}

/* Shared vacuum / CPML E sweep (grid.c:92-118, cpml.c:221-264). */
static void sweep_e(fdtd_grid_t *g, const oracle_stretch_t st[3][2]) {
    oracle_geom_t o = geom_of(g);
    const double ca = g->dt / (EPS0 * g->dx);
    const double dx = g->dx;
    for (int c = 0; c < 3; c++) {
        int p = (c + 1) % 3, q = (c + 2) % 3;
        double *E = e_of(g, c);
        const double *Fq = h_of(g, q), *Fp = h_of(g, p);
        size_t lo[3], hi[3];
        e_range(&o, c, lo, hi);
        size_t x[3];
        for (x[0] = lo[0]; x[0] < hi[0]; x[0]++)
        for (x[1] = lo[1]; x[1] < hi[1]; x[1]++)
        for (x[2] = lo[2]; x[2] < hi[2]; x[2]++) {
            size_t idx = lin(&o, x);
            size_t xp[3] = {x[0], x[1], x[2]}, xq[3] = {x[0], x[1], x[2]};
            xp[p] = wrap_down(&o, p, x[p]);
            xq[q] = wrap_down(&o, q, x[q]);
            double dA = Fq[idx] - Fq[lin(&o, xp)];
            double dB = Fp[idx] - Fp[lin(&o, xq)];
            if (st) {
                const oracle_stretch_t *sa = &st[c][0], *sb = &st[c][1];
                size_t ia = x[p], ib = x[q];
                dA = cpml_term(dA, sa->kinv[ia], sa->b[ia], sa->a[ia], dx, &sa->psi[idx]);
                dB = cpml_term(dB, sb->kinv[ib], sb->b[ib], sb->a[ib], dx, &sb->psi[idx]);
            }
            E[idx] += ca * (dA - dB);
        }
    }
}

void fdtd_update_h(fdtd_grid_t *g) { if (g) sweep_h(g, NULL); }
void fdtd_update_e(fdtd_grid_t *g) { if (g) sweep_e(g, NULL); }

/* ------------------------------------------------------------------ cpml */

/* engine/fdtd/cpml.c:10-19 */
cpml_params_t cpml_default_params(size_t npml) {
    cpml_params_t p = {npml, 3.0, 1.0, 7.0, 0.05, 1e-6};
    return p;
}

/* One graded profile point at physical depth position x (cpml.c:44-84).
 * The depth rho is measured from the inner PML interface toward the wall. */
static void profile_point(double x, double d, double len, const cpml_params_t *pp,
                          double sigma_max, double dx, double dt,
                          double *kinv, double *b, double *a) {
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

static double *new_arr(size_t n) { return malloc((n ? n : 1) * sizeof(double)); }

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

    double **km[3] = {&p->kappa_inv_m_x, &p->kappa_inv_m_y, &p->kappa_inv_m_z};
    double **bm[3] = {&p->b_m_x, &p->b_m_y, &p->b_m_z};
    double **am[3] = {&p->a_m_x, &p->a_m_y, &p->a_m_z};
    double **kh[3] = {&p->kappa_inv_h_x, &p->kappa_inv_h_y, &p->kappa_inv_h_z};
    double **bh[3] = {&p->b_h_x, &p->b_h_y, &p->b_h_z};
    double **ah[3] = {&p->a_h_x, &p->a_h_y, &p->a_h_z};
    for (int ax = 0; ax < 3; ax++) {
        size_t nn = n[ax];
        *km[ax] = new_arr(nn + 1); *bm[ax] = new_arr(nn + 1); *am[ax] = new_arr(nn + 1);
        *kh[ax] = new_arr(nn);     *bh[ax] = new_arr(nn);     *ah[ax] = new_arr(nn);
        if (!*km[ax] || !*bm[ax] || !*am[ax] || !*kh[ax] || !*bh[ax] || !*ah[ax]) {
            cpml_destroy(p);
            return NULL;
        }
        int periodic = (g->periodic_axes >> ax) & 1u;
        double d = (double)prm.npml * g->dx, len = (double)nn * g->dx;
        for (size_t q = 0; q <= nn; q++) {
            if (periodic) { (*km[ax])[q] = 1; (*bm[ax])[q] = 1; (*am[ax])[q] = 0; continue; }
            profile_point((double)q * g->dx, d, len, &prm, sigma_max, g->dx, g->dt,
                          &(*km[ax])[q], &(*bm[ax])[q], &(*am[ax])[q]);
        }
        for (size_t q = 0; q < nn; q++) {
            if (periodic) { (*kh[ax])[q] = 1; (*bh[ax])[q] = 1; (*ah[ax])[q] = 0; continue; }
            profile_point(((double)q + 0.5) * g->dx, d, len, &prm, sigma_max, g->dx, g->dt,
                          &(*kh[ax])[q], &(*bh[ax])[q], &(*ah[ax])[q]);
        }
    }

    size_t cells = (g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    double **psi[12] = {&p->psi_hx_y, &p->psi_hx_z, &p->psi_hy_z, &p->psi_hy_x,
                        &p->psi_hz_x, &p->psi_hz_y, &p->psi_ex_y, &p->psi_ex_z,
                        &p->psi_ey_z, &p->psi_ey_x, &p->psi_ez_x, &p->psi_ez_y};
    for (int q = 0; q < 12; q++) {
        *psi[q] = calloc(cells, sizeof(double));
        if (!*psi[q]) { cpml_destroy(p); return NULL; }
    }
    return p;
}

void cpml_destroy(cpml_t *p) {
    if (!p) return;
    double *all[] = {p->kappa_inv_m_x, p->b_m_x, p->a_m_x, p->kappa_inv_m_y, p->b_m_y, p->a_m_y,
                     p->kappa_inv_m_z, p->b_m_z, p->a_m_z, p->kappa_inv_h_x, p->b_h_x, p->a_h_x,
                     p->kappa_inv_h_y, p->b_h_y, p->a_h_y, p->kappa_inv_h_z, p->b_h_z, p->a_h_z,
                     p->psi_hx_y, p->psi_hx_z, p->psi_hy_z, p->psi_hy_x, p->psi_hz_x, p->psi_hz_y,
                     p->psi_ex_y, p->psi_ex_z, p->psi_ey_z, p->psi_ey_x, p->psi_ez_x, p->psi_ez_y};
    for (size_t q = 0; q < sizeof all / sizeof all[0]; q++) free(all[q]);
    free(p);
}

/* Profile/psi table per (component, p|q derivative) for the H and E sweeps. */
void oracle_cpml_stretch_h(const cpml_t *p, oracle_stretch_t st[3][2]) {
    const double *k[3] = {p->kappa_inv_h_x, p->kappa_inv_h_y, p->kappa_inv_h_z};
    const double *b[3] = {p->b_h_x, p->b_h_y, p->b_h_z};
    const double *a[3] = {p->a_h_x, p->a_h_y, p->a_h_z};
    /* Hx: (y, z) ; Hy: (z, x) ; Hz: (x, y) */
    double *psi[3][2] = {{p->psi_hx_y, p->psi_hx_z}, {p->psi_hy_z, p->psi_hy_x}, {p->psi_hz_x, p->psi_hz_y}};
    for (int c = 0; c < 3; c++)
        for (int s = 0; s < 2; s++) {
            int ax = (c + 1 + s) % 3;
            st[c][s].kinv = k[ax]; st[c][s].b = b[ax]; st[c][s].a = a[ax];
            st[c][s].psi = psi[c][s];
        }
}

void oracle_cpml_stretch_e(const cpml_t *p, oracle_stretch_t st[3][2]) {
    const double *k[3] = {p->kappa_inv_m_x, p->kappa_inv_m_y, p->kappa_inv_m_z};
    const double *b[3] = {p->b_m_x, p->b_m_y, p->b_m_z};
    const double *a[3] = {p->a_m_x, p->a_m_y, p->a_m_z};
    double *psi[3][2] = {{p->psi_ex_y, p->psi_ex_z}, {p->psi_ey_z, p->psi_ey_x}, {p->psi_ez_x, p->psi_ez_y}};
    for (int c = 0; c < 3; c++)
        for (int s = 0; s < 2; s++) {
            int ax = (c + 1 + s) % 3;
            st[c][s].kinv = k[ax]; st[c][s].b = b[ax]; st[c][s].a = a[ax];
            st[c][s].psi = psi[c][s];
        }
}

/* engine/fdtd/cpml.c:176-219 */
void fdtd_update_h_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml) return;
    oracle_stretch_t st[3][2];
    oracle_cpml_stretch_h(pml, st);
    sweep_h(g, (const oracle_stretch_t (*)[2])st);
}

/* engine/fdtd/cpml.c:221-264 */
void fdtd_update_e_cpml(fdtd_grid_t *g, cpml_t *pml) {
    if (!g || !pml) return;
    oracle_stretch_t st[3][2];
    oracle_cpml_stretch_e(pml, st);
    sweep_e(g, (const oracle_stretch_t (*)[2])st);
}

/* ---------------------------------------------------------------- source */

/* engine/fdtd/source.c:8-26 */
double source_waveform_gaussian_derivative(double t, const void *params) {
    const source_gaussian_deriv_params_t *p = params;
    double u = (t - p->t0) / p->tau;
    return -2.0 * u * exp(-u * u);
}

double source_waveform_ricker(double t, const void *params) {
    const source_ricker_params_t *p = params;
    double a = M_PI * p->f0_hz * (t - p->t0);
    double a2 = a * a;
    return (1.0 - 2.0 * a2) * exp(-a2);
}

double source_waveform_modulated_gaussian(double t, const void *params) {
    const source_modulated_gaussian_params_t *p = params;
    double u = (t - p->t0) / p->tau;
    double env = exp(-u * u);
    return env * sin(2.0 * M_PI * p->freq_hz * (t - p->t0));
}

source_t source_make_point(size_t i, size_t j, size_t k, field_component_t component,
                           source_mode_t mode, double amplitude,
                           source_waveform_fn waveform, const void *params) {
    source_t s = {i, j, k, component, mode, amplitude, waveform, params};
    return s;
}

/* engine/fdtd/source.c:40-57 */
void source_apply(const source_t *src, fdtd_grid_t *g, double t) {
    if (!src || !g || !src->waveform) return;
    if (src->i > g->nx || src->j > g->ny || src->k > g->nz) return;
    if (src->component > FIELD_EZ) return;
    double *f = e_of(g, (int)src->component);
    double v = src->amplitude * src->waveform(t, src->params);
    size_t idx = fdtd_index(g, src->i, src->j, src->k);
    if (src->mode == SRC_SOFT) f[idx] += v;
    else f[idx] = v;
}

/* ----------------------------------------------------------------- probe */

static const double *field_of(const fdtd_grid_t *g, field_component_t c) {
    switch (c) {
    case FIELD_EX: return g->ex;
    case FIELD_EY: return g->ey;
    case FIELD_EZ: return g->ez;
    case FIELD_HX: return g->hx;
    case FIELD_HY: return g->hy;
    case FIELD_HZ: return g->hz;
    }
    return NULL;
}

/* engine/fdtd/probe.c:21-49 */
int probe_td_create(probe_td_t *p, size_t i, size_t j, size_t k,
                    field_component_t component, size_t n_steps) {
    if (!p || n_steps == 0) return -1;
    double *buf = malloc(n_steps * sizeof(double));
    if (!buf) return -1;
    p->i = i; p->j = j; p->k = k;
    p->component = component;
    p->buffer = buf;
    p->n_steps = n_steps;
    p->next_index = 0;
    return 0;
}

void probe_td_destroy(probe_td_t *p) {
    if (!p) return;
    free(p->buffer);
    p->buffer = NULL;
}

void probe_td_record(probe_td_t *p, const fdtd_grid_t *g) {
    if (!p || !g || !p->buffer || p->next_index >= p->n_steps) return;
    const double *f = field_of(g, p->component);
    if (!f) return;
    p->buffer[p->next_index++] = f[fdtd_index(g, p->i, p->j, p->k)];
}

/* engine/fdtd/probe.c:51-105 */
int probe_dft_create(probe_dft_t *p, size_t i, size_t j, size_t k,
                     field_component_t component, const double *freqs_hz, size_t n_freqs) {
    if (!p || !freqs_hz || n_freqs == 0) return -1;
    double *re = calloc(n_freqs, sizeof(double));
    double *im = calloc(n_freqs, sizeof(double));
    if (!re || !im) { free(re); free(im); return -1; }
    p->i = i; p->j = j; p->k = k;
    p->component = component;
    p->freqs_hz = freqs_hz;
    p->n_freqs = n_freqs;
    p->acc_re = re;
    p->acc_im = im;
    return 0;
}

void probe_dft_destroy(probe_dft_t *p) {
    if (!p) return;
    free(p->acc_re); free(p->acc_im);
    p->acc_re = p->acc_im = NULL;
}

void probe_dft_record(probe_dft_t *p, const fdtd_grid_t *g, double t, double dt) {
    if (!p || !g) return;
    const double *f = field_of(g, p->component);
    if (!f) return;
    double v = f[fdtd_index(g, p->i, p->j, p->k)];
    for (size_t q = 0; q < p->n_freqs; q++) {
        double w = 2.0 * M_PI * p->freqs_hz[q];
        p->acc_re[q] += v * cos(w * t) * dt;
        p->acc_im[q] -= v * sin(w * t) * dt;
    }
}

void probe_dft_result(const probe_dft_t *p, size_t f, double *mag, double *phase) {
    if (!p || f >= p->n_freqs) {
        if (mag) *mag = 0.0;
        if (phase) *phase = 0.0;
        return;
    }
    double re = p->acc_re[f], im = p->acc_im[f];
    if (mag) *mag = sqrt(re * re + im * im);
    if (phase) *phase = atan2(im, re);
}
