/*
 * One-call simulation entry point (drop-in for engine/fdtd/simulate.h).
 *
 * Replaces: engine/fdtd/simulate.h:9-29 and simulate.c:10-69; called by
 * cli/fdtd_stage.c:20.  The step loop runs device-resident on the GPU in
 * FP64: H (CPML) -> soft Gaussian-derivative source at t = n dt -> E
 * (CPML + materials) -> time-domain probe, exactly the reference order.
 */
#ifndef RFSYN_FDTD_SIMULATE_H
#define RFSYN_FDTD_SIMULATE_H

#include <stddef.h>
#include "field_component.h"
#include "grid.h"
#include "materials.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    size_t nx, ny, nz;
    double dx;
    double courant_safety;
    size_t n_steps;

    size_t npml;

    size_t src_i, src_j, src_k;
    field_component_t src_component;
    double t0_over_dt, tau_over_dt;

    size_t probe_i, probe_j, probe_k;
    field_component_t probe_component;
} fdtd_sim_config_t;

/* 0 on success; -1 on NULL cfg/out_td, zero steps, invalid grid, source or
 * probe beyond (nx,ny,nz), CPML/materials failure, setup failure or
 * allocation failure.  out_td receives n_steps samples; sample_rate_out
 * (nullable) receives 1/dt. */
int fdtd_simulate_echo(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out);

typedef int (*fdtd_material_setup_fn)(const fdtd_grid_t *grid, materials_t *materials, void *context);

int fdtd_simulate_echo_with_setup(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out, fdtd_material_setup_fn setup, void *context);

#ifdef __cplusplus
}
#endif

#endif
