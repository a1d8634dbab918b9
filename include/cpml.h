/*
 * Convolutional PML (drop-in for engine/fdtd/cpml.h).
 *
 * Replaces: engine/fdtd/cpml.h:8-51 and cpml.c:10-264.
 *
 * The 1D grading profiles are public host arrays exactly as in the
 * reference: *_m_* at integer nodes (n+1 entries), *_h_* at half nodes
 * (n entries).  The psi memories live on the GPU only, compacted to the
 * slabs where the profile's a-coefficient is nonzero; the psi_* pointers
 * below are kept for layout compatibility and are NULL in this library.
 */
#ifndef RFSYN_FDTD_CPML_H
#define RFSYN_FDTD_CPML_H

#include <stddef.h>
#include "grid.h"

#ifdef __cplusplus
extern "C" {
#endif

struct cpml_device;

struct cpml {
    size_t npml;
    double dx;

    double *kappa_inv_m_x, *b_m_x, *a_m_x;
    double *kappa_inv_m_y, *b_m_y, *a_m_y;
    double *kappa_inv_m_z, *b_m_z, *a_m_z;

    double *kappa_inv_h_x, *b_h_x, *a_h_x;
    double *kappa_inv_h_y, *b_h_y, *a_h_y;
    double *kappa_inv_h_z, *b_h_z, *a_h_z;

    double *psi_hx_y, *psi_hx_z;
    double *psi_hy_z, *psi_hy_x;
    double *psi_hz_x, *psi_hz_y;
    double *psi_ex_y, *psi_ex_z;
    double *psi_ey_z, *psi_ey_x;
    double *psi_ez_x, *psi_ez_y;

    /* Library-private: grid extents and periodic mask captured at
     * creation, and the device-side profiles and psi slabs. */
    size_t nx, ny, nz;
    unsigned periodic_axes;
    double dt;
    struct cpml_device *dev;
};

typedef struct cpml cpml_t;

typedef struct {
    size_t npml;
    double m;
    double ma;
    double kappa_max;
    double alpha_max;
    double r0;
} cpml_params_t;

/* cpml.c:10-19: m=3, ma=1, kappa_max=7, alpha_max=0.05, R0=1e-6. */
cpml_params_t cpml_default_params(size_t npml);

/* cpml.c:91-150.  NULL on invalid params or npml >= (n+1)/2 on a
 * non-periodic axis.  Periodic axes get identity profiles. */
cpml_t *cpml_create(const fdtd_grid_t *g, cpml_params_t params);
void cpml_destroy(cpml_t *pml);

/* cpml.c:176-219 / 221-264: H and vacuum-E updates with CPML stretching. */
void fdtd_update_h_cpml(fdtd_grid_t *g, cpml_t *pml);
void fdtd_update_e_cpml(fdtd_grid_t *g, cpml_t *pml);

/* cpml.h:46-51: psi <- b psi + a d;  returns d/kappa + dx psi. */
static inline double cpml_term(double diff, double kappa_inv, double b,
                               double a, double dx, double *psi) {
    double next = b * (*psi) + a * diff;
    *psi = next;
    return diff * kappa_inv + dx * next;
}

#ifdef __cplusplus
}
#endif

#endif
