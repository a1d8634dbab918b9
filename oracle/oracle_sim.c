/*
 * CPU oracle: fdtd_simulate_echo and the device-handle API of
 * fdtd_native.h restated as a host loop over the oracle's own low-level
 * calls, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_core.c for the usage rule).
 *
 * fdtd_simulate_echo_with_setup follows engine/fdtd/simulate.c:10-69.
 * fdtd_sim_* follow the step order of simulate.c:48-54 and
 * validate_probe_dft.c:66-82: H -> pre-E sources at t = n dt -> E ->
 * post-E plane sources at t = (n+1) dt -> probes and DFTs at t = (n+1) dt.
 * The DFT accumulation is probe.c:87-91 (plain) or the same terms summed
 * with Neumaier compensation.
 */
#include "simulate.h"
#include "cpml.h"
#include "materials.h"
#include "source.h"
#include "probe.h"
#include "fdtd_native.h"
#include "constants.h"
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* engine/fdtd/simulate.c:10-66 */
int fdtd_simulate_echo_with_setup(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out,
                                  fdtd_material_setup_fn setup, void *context) {
    if (!cfg || !out_td || cfg->n_steps == 0) return -1;
    fdtd_grid_t *g = fdtd_grid_create(cfg->nx, cfg->ny, cfg->nz, cfg->dx, cfg->courant_safety);
    if (!g) return -1;
    int rc = -1;
    cpml_t *pml = NULL;
    materials_t *mat = NULL;
    probe_td_t probe = {0};
    if (cfg->src_i > g->nx || cfg->src_j > g->ny || cfg->src_k > g->nz ||
        cfg->probe_i > g->nx || cfg->probe_j > g->ny || cfg->probe_k > g->nz) goto out;
    pml = cpml_create(g, cpml_default_params(cfg->npml));
    if (!pml) goto out;
    mat = materials_create(g);
    if (!mat) goto out;
    if (setup && setup(g, mat, context) != 0) goto out;
    source_gaussian_deriv_params_t sp = {cfg->t0_over_dt * g->dt, cfg->tau_over_dt * g->dt};
    source_t src = source_make_point(cfg->src_i, cfg->src_j, cfg->src_k, cfg->src_component, SRC_SOFT,
                                     1.0, source_waveform_gaussian_derivative, &sp);
    if (probe_td_create(&probe, cfg->probe_i, cfg->probe_j, cfg->probe_k, cfg->probe_component,
                        cfg->n_steps) != 0) goto out;
    for (size_t n = 0; n < cfg->n_steps; n++) {
        fdtd_update_h_cpml(g, pml);
        source_apply(&src, g, (double)n * g->dt);
        fdtd_update_e_cpml_materials(g, pml, mat);
        probe_td_record(&probe, g);
    }
    memcpy(out_td, probe.buffer, cfg->n_steps * sizeof(double));
    if (sample_rate_out) *sample_rate_out = 1.0 / g->dt;
    rc = 0;
out:
    probe_td_destroy(&probe);
    materials_destroy(mat);
    cpml_destroy(pml);
    fdtd_grid_destroy(g);
    return rc;
}

int fdtd_simulate_echo(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out) {
    return fdtd_simulate_echo_with_setup(cfg, out_td, sample_rate_out, NULL, NULL);
}

/* ---------------------------------------------------- fdtd_sim on CPU */

materials_t *oracle_materials_clone(const materials_t *m, const fdtd_grid_t *g);

typedef struct { double s, c; } kahan_t;

static void nadd(kahan_t *k, double x) {       /* Neumaier */
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
    else k->c += (x - t) + k->s;
    k->s = t;
}

typedef struct {
    int comp; size_t i, j, k;
    double *td; size_t td_cap, td_n;
    double *freqs; size_t nf;
    kahan_t *re, *im;
} oprobe_t;

typedef struct {
    int axis; size_t index; int comp[2]; size_t na, nb;
    double *freqs; size_t nf;
    kahan_t *re[2], *im[2];      /* [f][a][b] */
} omon_t;

typedef struct {
    int comp, axis; size_t index; source_mode_t mode; int after_e;
    double amp; source_waveform_fn wf; const void *params;
} oplane_t;

struct fdtd_sim {
    fdtd_grid_t *g;
    cpml_t *pml;
    materials_t *mat;
    fdtd_sim_options_t opt;
    size_t steps;
    source_t *pts; size_t n_pts;
    oplane_t *planes; size_t n_planes;
    oprobe_t *probes; size_t n_probes;
    omon_t *mons; size_t n_mons;
    /* state n kept beside state n+1 for the energy audit and flux planes:
     * E^n, H^{n-1/2} copied at the start of every step */
    int track;
    double *prev[6];
    int (*flux)[2]; size_t n_flux;
    double *flux_acc;
};

static cpml_t *clone_cpml(const cpml_t *p, const fdtd_grid_t *g) {
    cpml_t *q = calloc(1, sizeof *q);
    if (!q) return NULL;
    *q = *p;
    size_t n[3] = {g->nx, g->ny, g->nz};
    double **dst[18] = {&q->kappa_inv_m_x, &q->b_m_x, &q->a_m_x, &q->kappa_inv_m_y, &q->b_m_y, &q->a_m_y,
                        &q->kappa_inv_m_z, &q->b_m_z, &q->a_m_z, &q->kappa_inv_h_x, &q->b_h_x, &q->a_h_x,
                        &q->kappa_inv_h_y, &q->b_h_y, &q->a_h_y, &q->kappa_inv_h_z, &q->b_h_z, &q->a_h_z};
    for (int a = 0; a < 18; a++) {
        int ax = (a / 3) % 3, half = a >= 9;
        size_t len = half ? n[ax] : n[ax] + 1;
        double *src = *dst[a];
        *dst[a] = malloc((len ? len : 1) * sizeof(double));
        if (!*dst[a]) return NULL;
        memcpy(*dst[a], src, len * sizeof(double));
    }
    size_t cells = (g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    double **psi[12] = {&q->psi_hx_y, &q->psi_hx_z, &q->psi_hy_z, &q->psi_hy_x, &q->psi_hz_x, &q->psi_hz_y,
                        &q->psi_ex_y, &q->psi_ex_z, &q->psi_ey_z, &q->psi_ey_x, &q->psi_ez_x, &q->psi_ez_y};
    for (int a = 0; a < 12; a++) {
        double *src = *psi[a];
        *psi[a] = malloc(cells * sizeof(double));
        if (!*psi[a]) return NULL;
        memcpy(*psi[a], src, cells * sizeof(double));
    }
    return q;
}

static fdtd_grid_t *clone_grid(const fdtd_grid_t *g) {
    fdtd_grid_t *c = fdtd_grid_create(g->nx, g->ny, g->nz, g->dx, 1.0);
    if (!c) return NULL;
    c->dt = g->dt;
    c->periodic_axes = g->periodic_axes;
    size_t cells = (g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    memcpy(c->ex, g->ex, cells * sizeof(double)); memcpy(c->ey, g->ey, cells * sizeof(double));
    memcpy(c->ez, g->ez, cells * sizeof(double)); memcpy(c->hx, g->hx, cells * sizeof(double));
    memcpy(c->hy, g->hy, cells * sizeof(double)); memcpy(c->hz, g->hz, cells * sizeof(double));
    return c;
}

fdtd_sim_t *fdtd_sim_create(const fdtd_grid_t *g, const cpml_t *pml, const materials_t *mat,
                            const fdtd_sim_options_t *opt) {
    if (!g) return NULL;
    fdtd_sim_t *s = calloc(1, sizeof *s);
    if (!s) return NULL;
    if (opt) s->opt = *opt;
    s->g = clone_grid(g);
    s->pml = pml ? clone_cpml(pml, g) : NULL;
    s->mat = mat ? oracle_materials_clone(mat, s->g) : materials_create(s->g);
    if (!s->g || (pml && !s->pml) || !s->mat) { fdtd_sim_destroy(s); return NULL; }
    return s;
}

void fdtd_sim_destroy(fdtd_sim_t *s) {
    if (!s) return;
    for (size_t p = 0; p < s->n_probes; p++) {
        free(s->probes[p].td); free(s->probes[p].freqs); free(s->probes[p].re); free(s->probes[p].im);
    }
    for (size_t p = 0; p < s->n_mons; p++) {
        free(s->mons[p].freqs);
        for (int w = 0; w < 2; w++) { free(s->mons[p].re[w]); free(s->mons[p].im[w]); }
    }
    free(s->probes); free(s->mons); free(s->pts); free(s->planes);
    for (int q = 0; q < 6; q++) free(s->prev[q]);
    free(s->flux); free(s->flux_acc);
    materials_destroy(s->mat);
    cpml_destroy(s->pml);
    fdtd_grid_destroy(s->g);
    free(s);
}

int fdtd_sim_add_point_source(fdtd_sim_t *s, const source_t *src) {
    if (!s || !src || !src->waveform || src->component > FIELD_EZ) return -1;
    source_t *n = realloc(s->pts, (s->n_pts + 1) * sizeof *n);
    if (!n) return -1;
    s->pts = n;
    s->pts[s->n_pts++] = *src;
    return 0;
}

int fdtd_sim_add_plane_source(fdtd_sim_t *s, field_component_t component, int axis, size_t index,
                              source_mode_t mode, int after_e, double amplitude,
                              source_waveform_fn waveform, const void *params) {
    if (!s || component > FIELD_EZ || axis < 0 || axis > 2 || !waveform) return -1;
    size_t n[3] = {s->g->nx, s->g->ny, s->g->nz};
    if (index > n[axis]) return -1;
    oplane_t *p = realloc(s->planes, (s->n_planes + 1) * sizeof *p);
    if (!p) return -1;
    s->planes = p;
    oplane_t pl = {(int)component, axis, index, mode, after_e ? 1 : 0, amplitude, waveform, params};
    s->planes[s->n_planes++] = pl;
    return 0;
}

static int kahan_alloc(kahan_t **k, size_t n) {
    *k = calloc(n ? n : 1, sizeof(kahan_t));
    return *k ? 0 : -1;
}

int fdtd_sim_add_probe(fdtd_sim_t *s, field_component_t component, size_t i, size_t j, size_t k,
                       size_t td_capacity, const double *freqs_hz, size_t n_freqs) {
    if (!s || component > FIELD_HZ || i > s->g->nx || j > s->g->ny || k > s->g->nz) return -1;
    if (n_freqs && !freqs_hz) return -1;
    oprobe_t *np = realloc(s->probes, (s->n_probes + 1) * sizeof *np);
    if (!np) return -1;
    s->probes = np;
    oprobe_t *p = &s->probes[s->n_probes];
    memset(p, 0, sizeof *p);
    p->comp = (int)component; p->i = i; p->j = j; p->k = k;
    p->td_cap = td_capacity;
    if (td_capacity && !(p->td = malloc(td_capacity * sizeof(double)))) return -1;
    p->nf = n_freqs;
    if (n_freqs) {
        p->freqs = malloc(n_freqs * sizeof(double));
        if (!p->freqs || kahan_alloc(&p->re, n_freqs) || kahan_alloc(&p->im, n_freqs)) return -1;
        memcpy(p->freqs, freqs_hz, n_freqs * sizeof(double));
    }
    return (int)s->n_probes++;
}

int fdtd_sim_add_plane_monitor(fdtd_sim_t *s, int axis, size_t index, const double *freqs_hz, size_t n_freqs) {
    if (!s || axis < 0 || axis > 2 || !freqs_hz || !n_freqs) return -1;
    size_t n[3] = {s->g->nx, s->g->ny, s->g->nz};
    if (index > n[axis]) return -1;
    omon_t *nm = realloc(s->mons, (s->n_mons + 1) * sizeof *nm);
    if (!nm) return -1;
    s->mons = nm;
    omon_t *m = &s->mons[s->n_mons];
    memset(m, 0, sizeof *m);
    m->axis = axis; m->index = index;
    m->comp[0] = (axis + 1) % 3 < (axis + 2) % 3 ? (axis + 1) % 3 : (axis + 2) % 3;
    m->comp[1] = 3 - axis - m->comp[0];
    m->na = n[m->comp[0]]; m->nb = n[m->comp[1]];
    m->nf = n_freqs;
    m->freqs = malloc(n_freqs * sizeof(double));
    if (!m->freqs) return -1;
    memcpy(m->freqs, freqs_hz, n_freqs * sizeof(double));
    size_t tot = n_freqs * m->na * m->nb;
    for (int w = 0; w < 2; w++)
        if (kahan_alloc(&m->re[w], tot) || kahan_alloc(&m->im[w], tot)) return -1;
    return (int)s->n_mons++;
}

static void apply_plane(fdtd_sim_t *s, const oplane_t *p, double t) {
    fdtd_grid_t *g = s->g;
    double *F = e_of(g, p->comp);
    double v = p->amp * p->wf(t, p->params);
    int a = (p->axis + 1) % 3 < (p->axis + 2) % 3 ? (p->axis + 1) % 3 : (p->axis + 2) % 3;
    int b = 3 - p->axis - a;
    size_t n[3] = {g->nx, g->ny, g->nz};
    size_t x[3];
    x[p->axis] = p->index;
    for (x[a] = 0; x[a] < n[a]; x[a]++)
        for (x[b] = 0; x[b] < n[b]; x[b]++) {
            size_t idx = fdtd_index(g, x[0], x[1], x[2]);
            if (p->mode == SRC_SOFT) F[idx] += v; else F[idx] = v;
        }
}

static double field_value(const fdtd_grid_t *g, int comp, size_t idx) {
    const double *f = comp < 3 ? e_of((fdtd_grid_t *)g, comp) : h_of((fdtd_grid_t *)g, comp - 3);
    return f[idx];
}

static void dft_add(kahan_t *re, kahan_t *im, double v, double w, double t, double dt, int comp) {
    double tr = v * cos(w * t) * dt;
    double ti = v * sin(w * t) * dt;
    if (comp) { nadd(re, tr); nadd(im, -ti); }
    else { re->s += tr; im->s -= ti; }
}

/* Flux along +axis through the plane of E nodes at index p between state
 * n (prev) and n+1 (grid): sum over the plane's cells of
 * 0.5 (E_b^n + E_b^{n+1}) H_c^{n+1/2} - 0.5 (E_c^n + E_c^{n+1}) H_b^{n+1/2}
 * with (a, b, c) cyclic; unscaled by dx^2 dt. */
static double plane_flux(const fdtd_sim_t *s, int a, size_t p) {
    const fdtd_grid_t *g = s->g;
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const size_t n[3] = {g->nx, g->ny, g->nz};
    const double *en[3] = {s->prev[0], s->prev[1], s->prev[2]};
    const double *ep[3] = {g->ex, g->ey, g->ez};
    const double *hp[3] = {g->hx, g->hy, g->hz};
    double sum = 0.0;
    size_t x[3];
    x[a] = p;
    for (x[b] = 0; x[b] < n[b]; x[b]++)
        for (x[c] = 0; x[c] < n[c]; x[c]++) {
            const size_t idx = fdtd_index(g, x[0], x[1], x[2]);
            const double eb = 0.5 * (en[b][idx] + ep[b][idx]), ec = 0.5 * (en[c][idx] + ep[c][idx]);
            sum += eb * hp[c][idx] - ec * hp[b][idx];
        }
    return sum;
}

int fdtd_sim_add_flux_plane(fdtd_sim_t *s, int axis, size_t index) {
    if (!s || axis < 0 || axis > 2 || s->steps) return -1;
    const size_t n[3] = {s->g->nx, s->g->ny, s->g->nz};
    if (index >= n[axis]) return -1;
    int (*f)[2] = realloc(s->flux, (s->n_flux + 1) * sizeof *f);
    double *acc = realloc(s->flux_acc, (s->n_flux + 1) * sizeof *acc);
    if (f) s->flux = f;
    if (acc) s->flux_acc = acc;
    if (!f || !acc) return -1;
    s->flux[s->n_flux][0] = axis;
    s->flux[s->n_flux][1] = (int)index;
    s->flux_acc[s->n_flux] = 0.0;
    return (int)s->n_flux++;
}

int fdtd_sim_flux(fdtd_sim_t *s, int plane, double *energy) {
    if (!s || !energy || plane < 0 || (size_t)plane >= s->n_flux) return -1;
    *energy = s->flux_acc[plane] * s->g->dx * s->g->dx * s->g->dt;
    return 0;
}

/* The state is tracked when flux planes exist or the grid is small (the
 * audit of large grids is left to the device library). */
static int start_tracking(fdtd_sim_t *s) {
    const fdtd_grid_t *g = s->g;
    const size_t total = (g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    if (s->track || !(s->n_flux || g->nx * g->ny * g->nz <= ((size_t)1 << 22))) return 0;
    for (int q = 0; q < 6; q++)
        if (!(s->prev[q] = malloc(total * sizeof(double)))) return -1;
    s->track = 1;
    return 0;
}

int fdtd_sim_run(fdtd_sim_t *s, size_t n_steps) {
    if (!s) return -1;
    fdtd_grid_t *g = s->g;
    if (start_tracking(s)) return -1;
    const size_t bytes = (g->nx + 1) * (g->ny + 1) * (g->nz + 1) * sizeof(double);
    for (size_t it = 0; it < n_steps; it++) {
        size_t n = s->steps;
        if (s->track) {
            double *cur[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
            for (int q = 0; q < 6; q++) memcpy(s->prev[q], cur[q], bytes);
        }
        double t = (double)n * g->dt;
        double tp = (double)(n + 1) * g->dt;
        if (s->pml) fdtd_update_h_cpml(g, s->pml); else fdtd_update_h(g);
        for (size_t q = 0; q < s->n_pts; q++) source_apply(&s->pts[q], g, t);
        for (size_t q = 0; q < s->n_planes; q++) if (!s->planes[q].after_e) apply_plane(s, &s->planes[q], t);
        if (s->pml) fdtd_update_e_cpml_materials(g, s->pml, s->mat);
        else fdtd_update_e_materials(g, s->mat);
        for (size_t q = 0; q < s->n_planes; q++) if (s->planes[q].after_e) apply_plane(s, &s->planes[q], tp);
        for (size_t q = 0; q < s->n_probes; q++) {
            oprobe_t *p = &s->probes[q];
            double v = field_value(g, p->comp, fdtd_index(g, p->i, p->j, p->k));
            if (p->td_n < p->td_cap) p->td[p->td_n++] = v;
            for (size_t f = 0; f < p->nf; f++)
                dft_add(&p->re[f], &p->im[f], v, 2.0 * M_PI * p->freqs[f], tp, g->dt, s->opt.dft_compensated);
        }
        for (size_t q = 0; q < s->n_mons; q++) {
            omon_t *m = &s->mons[q];
            size_t x[3];
            x[m->axis] = m->index;
            for (int w = 0; w < 2; w++)
                for (size_t f = 0; f < m->nf; f++) {
                    /* the twiddle depends on (f, t) only: same terms as
                     * dft_add, evaluated once per frequency */
                    double om = 2.0 * M_PI * m->freqs[f];
                    double cs = cos(om * tp), sn = sin(om * tp);
                    for (x[m->comp[0]] = 0; x[m->comp[0]] < m->na; x[m->comp[0]]++)
                        for (x[m->comp[1]] = 0; x[m->comp[1]] < m->nb; x[m->comp[1]]++) {
                            double v = field_value(g, m->comp[w], fdtd_index(g, x[0], x[1], x[2]));
                            size_t o = (f * m->na + x[m->comp[0]]) * m->nb + x[m->comp[1]];
                            m->re[w][o].s += v * cs * g->dt;
                            m->im[w][o].s -= v * sn * g->dt;
                        }
                }
        }
        for (size_t q = 0; q < s->n_flux; q++) s->flux_acc[q] += plane_flux(s, s->flux[q][0], (size_t)s->flux[q][1]);
        s->steps++;
    }
    return 0;
}

size_t fdtd_sim_steps_done(const fdtd_sim_t *s) { return s ? s->steps : 0; }

long fdtd_sim_probe_td(fdtd_sim_t *s, int probe, double *out, size_t n) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || (n && !out)) return -1;
    oprobe_t *p = &s->probes[probe];
    size_t c = n < p->td_n ? n : p->td_n;
    if (c) memcpy(out, p->td, c * sizeof(double));
    return (long)c;
}

int fdtd_sim_probe_dft(fdtd_sim_t *s, int probe, double *re, double *im) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || !re || !im) return -1;
    oprobe_t *p = &s->probes[probe];
    for (size_t f = 0; f < p->nf; f++) {
        re[f] = p->re[f].s + p->re[f].c;
        im[f] = p->im[f].s + p->im[f].c;
    }
    return 0;
}

int fdtd_sim_plane_dft(fdtd_sim_t *s, int monitor, int which, double *re, double *im) {
    if (!s || monitor < 0 || (size_t)monitor >= s->n_mons || which < 0 || which > 1 || !re || !im) return -1;
    omon_t *m = &s->mons[monitor];
    size_t tot = m->nf * m->na * m->nb;
    for (size_t o = 0; o < tot; o++) {
        re[o] = m->re[which][o].s + m->re[which][o].c;
        im[o] = m->im[which][o].s + m->im[which][o].c;
    }
    return 0;
}

static int same_shape(const fdtd_grid_t *a, const fdtd_grid_t *b) {
    return a && b && a->nx == b->nx && a->ny == b->ny && a->nz == b->nz;
}

int fdtd_sim_download(fdtd_sim_t *s, fdtd_grid_t *g) {
    if (!s || !same_shape(s->g, g)) return -1;
    size_t cells = (g->nx + 1) * (g->ny + 1) * (g->nz + 1) * sizeof(double);
    memcpy(g->ex, s->g->ex, cells); memcpy(g->ey, s->g->ey, cells); memcpy(g->ez, s->g->ez, cells);
    memcpy(g->hx, s->g->hx, cells); memcpy(g->hy, s->g->hy, cells); memcpy(g->hz, s->g->hz, cells);
    return 0;
}

int fdtd_sim_upload(fdtd_sim_t *s, const fdtd_grid_t *g) {
    if (!s || !same_shape(s->g, g)) return -1;
    size_t cells = (g->nx + 1) * (g->ny + 1) * (g->nz + 1) * sizeof(double);
    memcpy(s->g->ex, g->ex, cells); memcpy(s->g->ey, g->ey, cells); memcpy(s->g->ez, g->ez, cells);
    memcpy(s->g->hx, g->hx, cells); memcpy(s->g->hy, g->hy, cells); memcpy(s->g->hz, g->hz, cells);
    return 0;
}

/* 0.5 sum eps E^2 dV and 0.5 sum mu0 H^2 dV over i<nx, j<ny, k<nz with
 * eps = dt/(dx Cb) of each edge's scalar coefficient
 * (audit_periodic_energy.c:24-61). */
int fdtd_sim_energy(fdtd_sim_t *s, double *electric, double *magnetic) {
    if (!s || !electric || !magnetic) return -1;
    fdtd_grid_t *g = s->g;
    double ue = 0.0, uh = 0.0;
    for (size_t i = 0; i < g->nx; i++)
        for (size_t j = 0; j < g->ny; j++)
            for (size_t k = 0; k < g->nz; k++) {
                size_t idx = fdtd_index(g, i, j, k);
                for (int c = 0; c < 3; c++) {
                    material_coeff_t co;
                    if (materials_get_coefficients(s->mat, g, (field_component_t)c, i, j, k, &co) != 0) return -1;
                    double eps = g->dt / (g->dx * co.Cb);
                    double e = e_of(g, c)[idx], h = h_of(g, c)[idx];
                    ue += 0.5 * eps * e * e;
                    uh += 0.5 * MU0 * h * h;
                }
            }
    double vol = g->dx * g->dx * g->dx;
    *electric = ue * vol;
    *magnetic = uh * vol;
    return 0;
}

/* Energy audit of step n = steps - 1 from the tracked state n and the
 * current H^{n+1/2} (audit_periodic_energy.c:43-71), same sums and cells
 * as fdtd_sim_energy. */
int fdtd_sim_energy_audit(fdtd_sim_t *s, double out[5], size_t *step) {
    if (!s || !out || !s->track || s->steps == 0) return -1;
    fdtd_grid_t *g = s->g;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t i = 0; i < g->nx; i++)
        for (size_t j = 0; j < g->ny; j++)
            for (size_t k = 0; k < g->nz; k++) {
                size_t idx = fdtd_index(g, i, j, k);
                for (int c = 0; c < 3; c++) {
                    material_coeff_t co;
                    if (materials_get_coefficients(s->mat, g, (field_component_t)c, i, j, k, &co) != 0) return -1;
                    const double eps = g->dt / (g->dx * co.Cb);
                    const double e = s->prev[c][idx], hm = s->prev[3 + c][idx], hp = h_of(g, c)[idx];
                    const double avg = 0.5 * (hm + hp);
                    acc[0] += 0.5 * eps * e * e;
                    acc[1] += 0.5 * MU0 * hm * hm;
                    acc[2] += 0.5 * MU0 * hp * hp;
                    acc[3] += 0.5 * MU0 * avg * avg;
                    acc[4] += 0.5 * MU0 * hm * hp;
                }
            }
    const double vol = g->dx * g->dx * g->dx;
    for (int q = 0; q < 5; q++) out[q] = acc[q] * vol;
    if (step) *step = s->steps - 1;
    return 0;
}

size_t fdtd_sim_device_bytes(const fdtd_sim_t *s) { (void)s; return 0; }

const char *fdtd_last_error(void) { return ""; }

/* ---- Dataset output stage (test infrastructure) ----
 * PCG32 (engine/core/rng.c:3-23) and the AWGN stage (engine/noise/
 * awgn.c:16-44) restated on a host I/Q buffer, with glibc's pow, sqrt,
 * sqrtf, logf, cosf and sinf as the reference calls them; the samples of
 * the FDTD stage are cli/fdtd_stage.c:20-24 (I = (float)v + 0, Q = 0). */
typedef struct { uint64_t state, inc; } opcg_t;

static uint32_t opcg_next(opcg_t *r) {
    const uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    const uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (x >> rot) | (x << ((-rot) & 31u));
}

static float opcg_uniform(opcg_t *r) { return (float)(opcg_next(r) >> 8) / (float)(1u << 24); }

/* AWGN on n interleaved I/Q samples in place (awgn.c:20-41). */
void oracle_awgn_iq(float *iq, size_t n, float snr_db, uint64_t seed, uint64_t stream) {
    if (!iq || n == 0) return;
    double signal_power = 0.0;
    for (size_t i = 0; i < n; i++) {
        const float mag = hypotf(iq[2 * i], iq[2 * i + 1]);   /* cabsf */
        signal_power += (double)mag * (double)mag;
    }
    signal_power /= (double)n;
    const double snr_linear = pow(10.0, (double)snr_db / 10.0);
    const double noise_power = (snr_linear > 0.0) ? signal_power / snr_linear : 0.0;
    const double per_axis_std = sqrt(noise_power / 2.0);
    opcg_t r = {0u, (stream << 1u) | 1u};   /* rng_seed, rng.c:3-9 */
    opcg_next(&r);
    r.state += seed;
    opcg_next(&r);
    for (size_t i = 0; i < n; i++) {
        float u1 = opcg_uniform(&r);
        const float u2 = opcg_uniform(&r);
        if (u1 <= 0.0f) u1 = 1e-9f;
        const float rad = sqrtf(-2.0f * logf(u1));
        const float nre = (float)per_axis_std * rad * cosf(2.0f * (float)3.14159265358979323846 * u2);
        const float nim = (float)per_axis_std * rad * sinf(2.0f * (float)3.14159265358979323846 * u2);
        /* samples[i] += nre + nim * I */
        const float cre = nre + nim * 0.0f, cim = nim * 1.0f;
        iq[2 * i] = iq[2 * i] + cre;
        iq[2 * i + 1] = iq[2 * i + 1] + cim;
    }
}

int fdtd_sim_probe_iq(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                      uint64_t stream, float *out_iq) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || (n && !out_iq)) return -1;
    oprobe_t *p = &s->probes[probe];
    if (n > p->td_n) return -1;
    for (size_t i = 0; i < n; i++) {
        out_iq[2 * i] = (float)p->td[i] + 0.0f;
        out_iq[2 * i + 1] = 0.0f;
    }
    if (add_noise) oracle_awgn_iq(out_iq, n, snr_db, seed, stream);
    return 0;
}
