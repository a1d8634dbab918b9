/*
 * Materials for the E update (drop-in for engine/fdtd/materials.h).
 *
 * Replaces: engine/fdtd/materials.h:10-36 and materials.c:92-1082.
 *
 * Semantics kept from the reference:
 *   - a 256-entry table of registered (eps_r, sigma) materials and one
 *     material id per padded node, shared by the Ex/Ey/Ez edges owned by
 *     that node (materials.c:36-38, 496-502);
 *   - per-component scalar overrides that win over the node id, cleared
 *     by materials_assign_box inside the box (materials.c:504-526);
 *   - per-component tensor entries that win over scalars (materials.c:658);
 *   - Ca = (1-L)/(1+L), Cb = dt/(eps dx)/(1+L), L = sigma dt/(2 eps)
 *     (materials.c:92-125).
 *
 * Extensions (not in the reference):
 *   - dispersive materials (Drude and Lorentz poles, trapezoidal ADE);
 *   - full symmetric 3x3 inverse-permittivity tensors on any E component.
 */
#ifndef RFSYN_FDTD_MATERIALS_H
#define RFSYN_FDTD_MATERIALS_H

#include <stddef.h>
#include <stdint.h>
#include "grid.h"
#include "cpml.h"
#include "field_component.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct materials materials_t;
typedef uint8_t material_id_t;
typedef struct { double Ca, Cb; } material_coeff_t;

materials_t *materials_create(const fdtd_grid_t *g);          /* materials.c:310-375 */
void materials_destroy(materials_t *mat);                      /* materials.c:377-386 */

/* materials.c:388-423: dedups by exact (eps_r, sigma); -1 when invalid
 * (eps_r < 1, sigma < 0, non-finite) or the table is full. */
int materials_register(materials_t *mat, double eps_r, double sigma, material_id_t *id_out);

/* materials.c:475-527: [i0,i1)x[j0,j1)x[k0,k1), padded ends accepted,
 * clipped on periodic axes; clears overrides and tensors inside. */
int materials_assign_box(materials_t *mat, const fdtd_grid_t *g, size_t i0, size_t i1, size_t j0, size_t j1, size_t k0, size_t k1, material_id_t id);
int materials_set_box(materials_t *mat, const fdtd_grid_t *g, size_t i0, size_t i1, size_t j0, size_t j1, size_t k0, size_t k1, double eps_r, double sigma);

/* materials.c:552-656: per-component scalar overrides. */
int materials_set_component(materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, double eps_r, double sigma);
int materials_clear_component(materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k);
int materials_get_coefficients(const materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, material_coeff_t *out);

/* materials.c:1062-1082: E updates; silently no-op on grid mismatch. */
void fdtd_update_e_materials(fdtd_grid_t *g, const materials_t *mat);
void fdtd_update_e_cpml_materials(fdtd_grid_t *g, cpml_t *pml, const materials_t *mat);

typedef struct {
    double m11, m12, m22;
    double scale; /* dt/(eps0*dx) */
} material_tensor_t;

/* materials.c:673-799: 2x2 x-y inverse-permittivity on Ex/Ey edges.
 * Component-local convention: Ex (xx, xy, yy), Ey (yy, xy, xx).
 * Requires SPD and an x/y interior margin; get returns 1 when absent. */
int materials_set_component_tensor(materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, double m11, double m12, double m22);
int materials_clear_component_tensor(materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k);
int materials_get_tensor(const materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, material_tensor_t *out);

/* ---- Extension: full 3x3 symmetric inverse-permittivity tensors ----
 *
 * m = {xx, xy, xz, yy, yz, zz}, relative (eps0 factored into scale).
 * The E_c update on a tensor edge is
 *   E_c += scale * ( M_cc curl_c
 *                    + sum over the 4 nearest d-edges (d != c) that also
 *                      carry a tensor:  w_cd * curl_d(neighbour) )
 *   w_cd = (0.25 * rho * sqrt(M_aa(a-edge))) * sqrt(M_bb(b-edge)),
 *   rho  = 0.5 * r(a-edge) + 0.5 * r(b-edge),
 *   r    = (M_cd / sqrt(M_cc)) / sqrt(M_dd) at that edge,
 * with (a, b) the pair in x<y<z order and curl the full discrete curl.
 * This is the reference's reciprocal pair rule (materials.c:836-903)
 * extended to the three component pairs; it is symmetric by construction
 * and equals the reference bit for bit for z-invariant x-y tensors.
 * Requires SPD; the tensor stencil must stay outside CPML slabs. */
int materials_set_component_tensor3(materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, const double m[6]);
/* Returns 0 and the 6 entries when present, 1 when absent, -1 on error. */
int materials_get_tensor3(const materials_t *mat, const fdtd_grid_t *g, field_component_t component, size_t i, size_t j, size_t k, double m_out[6]);

/* ---- Extension: dispersive materials (ADE) ----
 *
 * With time dependence exp(-i w t):
 * eps(w) = eps_inf + i sigma/(w eps0)
 *          - sum_d wp_d^2 / (w^2 + i gamma_d w)                   (Drude)
 *          + sum_l deps_l w0_l^2 / (w0_l^2 - w^2 - 2 i delta_l w)   (Lorentz)
 * Every pole is the state-space system dP/dt = J,
 * dJ/dt = -w0^2 P - g J + eps0 A E (Drude: w0 = 0, g = gamma, A = wp^2;
 * Lorentz: g = 2 delta, A = deps w0^2), integrated together with
 * Ampere's law by the trapezoidal rule (bilinear transform), which is
 * unconditionally stable for any wp*dt and gamma*dt.  Per step and edge:
 *   u_p   = kj_p J_p - kp_p P_p,    kj = 2/D, kp = dt w0^2/D,
 *           D = 1 + g dt/2 + w0^2 dt^2/4
 *   E'    = Ca E + Cb (curl terms) - Cu sum_p u_p
 *   S_p   = u_p + v_p (E' + E),     v = dt eps0 A/(2 D)
 *   J_p'  = S_p - J_p,   P_p' = P_p + (dt/2) S_p
 * with Ca = (a-b)/(a+b), Cb = 1/((a+b) dx), Cu = Cb dx/2,
 * a = eps0 eps_inf/dt, b = (sigma + sum_p v_p)/2. */
#define MATERIALS_MAX_POLES 4
typedef struct {
    double eps_inf;            /* >= 1 */
    double sigma;              /* S/m, >= 0 */
    size_t n_drude;            /* 0..MATERIALS_MAX_POLES */
    double drude_wp[MATERIALS_MAX_POLES];     /* rad/s */
    double drude_gamma[MATERIALS_MAX_POLES];  /* 1/s */
    size_t n_lorentz;          /* 0..MATERIALS_MAX_POLES */
    double lorentz_deps[MATERIALS_MAX_POLES];
    double lorentz_w0[MATERIALS_MAX_POLES];   /* rad/s */
    double lorentz_delta[MATERIALS_MAX_POLES];/* 1/s */
} material_dispersive_t;

/* Registers a dispersive material in the same id table (no dedup).
 * Per-component scalar overrides on its cells replace it by a plain
 * (non-dispersive) coefficient, as overrides do for plain ids. */
int materials_register_dispersive(materials_t *mat, const material_dispersive_t *d, material_id_t *id_out);

/* Introspection used by the device builder and the tests. */
size_t materials_table_count(const materials_t *mat);
int materials_get_dispersive(const materials_t *mat, material_id_t id, material_dispersive_t *out); /* 1 if id is plain */
size_t materials_override_count(const materials_t *mat);
size_t materials_tensor_count(const materials_t *mat);

#ifdef __cplusplus
}
#endif

#endif
