/*
 * Waveforms, point sources and probes on the grid's host arrays.
 *
 * Replaces engine/fdtd/source.c and engine/fdtd/probe.c.  These calls act
 * between update calls on the host-visible fields, which is the contract of
 * the low-level API (validators add into g->ez directly,
 * validation_common.h:16-18).  The device-resident loop (sim.c) tabulates
 * the same waveform functions on the host and applies them in kernels.
 */
#include "source.h"
#include "probe.h"

#include <math.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* source.c:8-12 */
double source_waveform_gaussian_derivative(double t, const void *params) {
    const source_gaussian_deriv_params_t *p = params;
    double u = (t - p->t0) / p->tau;
    return -2.0 * u * exp(-u * u);
}

/* source.c:14-19 */
double source_waveform_ricker(double t, const void *params) {
    const source_ricker_params_t *p = params;
    double a = M_PI * p->f0_hz * (t - p->t0);
    double a2 = a * a;
    return (1.0 - 2.0 * a2) * exp(-a2);
}

/* source.c:21-26 */
double source_waveform_modulated_gaussian(double t, const void *params) {
    const source_modulated_gaussian_params_t *p = params;
    double u = (t - p->t0) / p->tau;
    double env = exp(-u * u);
    return env * sin(2.0 * M_PI * p->freq_hz * (t - p->t0));
}

source_t source_make_point(size_t i, size_t j, size_t k, field_component_t component,
                           source_mode_t mode, double amplitude, source_waveform_fn waveform,
                           const void *params) {
    source_t s;
    s.i = i; s.j = j; s.k = k;
    s.component = component;
    s.mode = mode;
    s.amplitude = amplitude;
    s.waveform = waveform;
    s.params = params;
    return s;
}

static double *e_field(fdtd_grid_t *g, field_component_t c) {
    switch (c) {
    case FIELD_EX: return g->ex;
    case FIELD_EY: return g->ey;
    case FIELD_EZ: return g->ez;
    default: return NULL;
    }
}

/* source.c:40-57 */
void source_apply(const source_t *src, fdtd_grid_t *g, double t) {
    if (!src || !g || !src->waveform) return;
    if (src->i > g->nx || src->j > g->ny || src->k > g->nz) return;
    double *f = e_field(g, src->component);
    if (!f) return;
    double v = src->amplitude * src->waveform(t, src->params);
    size_t idx = fdtd_index(g, src->i, src->j, src->k);
    if (src->mode == SRC_SOFT) f[idx] += v;
    else f[idx] = v;
}

static const double *any_field(const fdtd_grid_t *g, field_component_t c) {
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

/* probe.c:21-38 */
int probe_td_create(probe_td_t *p, size_t i, size_t j, size_t k, field_component_t component,
                    size_t n_steps) {
    if (!p || n_steps == 0) return -1;
    p->buffer = malloc(n_steps * sizeof(double));
    if (!p->buffer) return -1;
    p->i = i; p->j = j; p->k = k;
    p->component = component;
    p->n_steps = n_steps;
    p->next_index = 0;
    return 0;
}

void probe_td_destroy(probe_td_t *p) {
    if (!p) return;
    free(p->buffer);
    p->buffer = NULL;
}

/* probe.c:40-49 */
void probe_td_record(probe_td_t *p, const fdtd_grid_t *g) {
    if (!p || !g || !p->buffer || p->next_index >= p->n_steps) return;
    const double *f = any_field(g, p->component);
    if (!f) return;
    p->buffer[p->next_index++] = f[fdtd_index(g, p->i, p->j, p->k)];
}

/* probe.c:51-76 */
int probe_dft_create(probe_dft_t *p, size_t i, size_t j, size_t k, field_component_t component,
                     const double *freqs_hz, size_t n_freqs) {
    if (!p || !freqs_hz || n_freqs == 0) return -1;
    p->acc_re = calloc(n_freqs, sizeof(double));
    p->acc_im = calloc(n_freqs, sizeof(double));
    if (!p->acc_re || !p->acc_im) {
        free(p->acc_re);
        free(p->acc_im);
        return -1;
    }
    p->i = i; p->j = j; p->k = k;
    p->component = component;
    p->freqs_hz = freqs_hz;
    p->n_freqs = n_freqs;
    return 0;
}

void probe_dft_destroy(probe_dft_t *p) {
    if (!p) return;
    free(p->acc_re);
    free(p->acc_im);
    p->acc_re = NULL;
    p->acc_im = NULL;
}

/* probe.c:78-92 */
void probe_dft_record(probe_dft_t *p, const fdtd_grid_t *g, double t, double dt) {
    if (!p || !g) return;
    const double *f = any_field(g, p->component);
    if (!f) return;
    double v = f[fdtd_index(g, p->i, p->j, p->k)];
    for (size_t q = 0; q < p->n_freqs; q++) {
        double w = 2.0 * M_PI * p->freqs_hz[q];
        p->acc_re[q] += v * cos(w * t) * dt;
        p->acc_im[q] -= v * sin(w * t) * dt;
    }
}

/* probe.c:94-104 */
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
