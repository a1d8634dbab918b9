/*
 * Scene painter with subpixel smoothing (extension; the reference has the
 * 2D rectangle helpers of engine/fdtd/subpixel_te.h and the tilted-slab
 * set-up of validate_materials_tilted.c:163-380, but no 3D geometry).
 *
 * Shapes are painted in order (a later shape wins where shapes overlap)
 * into a materials_t through the public materials API only:
 *   - every node takes the material at its own position (node ids);
 *   - every E edge whose dual cell (the unit cube centred on the edge) is
 *     uniform but differs from its node's material gets a scalar override;
 *   - every E edge whose dual cell straddles an interface gets the smoothed
 *     inverse permittivity  M = ln n n^T + lt (I - n n^T)  with
 *     ln = (1-f)/eps_out + f/eps_in,  lt = 1/((1-f) eps_out + f eps_in),
 *     f the fill fraction of the inside material over the dual cell
 *     (s x s x s sub-samples) and n the shape's surface normal at the edge:
 *     the 3D form of subpixel_te_mix_epsilon_tensor (subpixel_te.c:106-132);
 *   - edges within `keepout` cells of a non-periodic boundary (the CPML and
 *     the tensor stencil margin) get the scalar 1/M_cc instead of a tensor,
 *     with f and n_c^2 quantised to 1/64 steps so that band holds few
 *     distinct coefficients.
 * Materials are lossless here.  Coordinates are in cell units with node
 * (i,j,k) at (i,j,k).
 */
#ifndef RFSYN_FDTD_SCENE_H
#define RFSYN_FDTD_SCENE_H

#include <stddef.h>
#include "grid.h"
#include "materials.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FDTD_SHAPE_SPHERE = 0,   /* centre p, radius */
    FDTD_SHAPE_SLAB = 1,     /* |normal . (x - p)| <= half_thickness, clipped to [lo, hi] */
    FDTD_SHAPE_BOX = 2       /* axis-aligned [lo, hi] */
} fdtd_shape_kind_t;

typedef struct {
    int kind;
    double eps_r;
    double p[3];
    double radius;
    double normal[3];
    double half_thickness;
    double lo[3], hi[3];
} fdtd_shape_t;

typedef struct {
    int supersample;          /* sub-samples per axis in a dual cell (0: 4) */
    size_t keepout;           /* scalar-only band at non-periodic faces */
} fdtd_scene_options_t;

typedef struct {
    size_t nodes_painted;
    size_t override_edges;
    size_t tensor_edges;
    size_t scalar_mixed_edges;
} fdtd_scene_stats_t;

/* 0 on success, -1 on invalid input or materials failure. */
int fdtd_scene_paint(materials_t *mat, const fdtd_grid_t *g, const fdtd_shape_t *shapes, size_t n_shapes,
                     const fdtd_scene_options_t *opt, fdtd_scene_stats_t *stats);

#ifdef __cplusplus
}
#endif

#endif
