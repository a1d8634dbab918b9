/*
 * Point sources and waveforms (drop-in for engine/fdtd/source.h).
 *
 * Replaces: engine/fdtd/source.h:8-35 and source.c:8-57.
 * source_apply adds (soft) or writes (hard) amplitude*waveform(t) into one
 * E cell of the host field arrays; the device-resident path tabulates the
 * same waveform values on the host and applies them inside the E kernel.
 */
#ifndef RFSYN_FDTD_SOURCE_H
#define RFSYN_FDTD_SOURCE_H

#include <stddef.h>
#include "grid.h"
#include "field_component.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SRC_SOFT,
    SRC_HARD
} source_mode_t;

typedef double (*source_waveform_fn)(double t, const void *params);

typedef struct {
    size_t i, j, k;
    field_component_t component;   /* E components only */
    source_mode_t mode;
    double amplitude;
    source_waveform_fn waveform;
    const void *params;            /* borrowed */
} source_t;

/* -2u exp(-u^2), u = (t - t0)/tau  (source.c:8-12) */
typedef struct { double t0, tau; } source_gaussian_deriv_params_t;
double source_waveform_gaussian_derivative(double t, const void *params);

/* (1 - 2a^2) exp(-a^2), a = pi f0 (t - t0)  (source.c:14-19) */
typedef struct { double f0_hz, t0; } source_ricker_params_t;
double source_waveform_ricker(double t, const void *params);

/* exp(-((t-t0)/tau)^2) sin(2 pi f (t - t0))  (source.c:21-26) */
typedef struct { double freq_hz, t0, tau; } source_modulated_gaussian_params_t;
double source_waveform_modulated_gaussian(double t, const void *params);

source_t source_make_point(size_t i, size_t j, size_t k, field_component_t component, source_mode_t mode, double amplitude, source_waveform_fn waveform, const void *params);

/* Ignores NULLs, positions beyond (nx,ny,nz) and H components. */
void source_apply(const source_t *src, fdtd_grid_t *g, double t);

#ifdef __cplusplus
}
#endif

#endif
