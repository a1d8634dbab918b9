/*
 * Scene painter with subpixel smoothing (include/fdtd_scene.h), painted on
 * the device.
 *
 * The device computes, for the whole grid at once (scene_kernels.hip):
 *   1. the node ids: the last shape containing each node;
 *   2. the E edges whose dual cell may contain a shape boundary, listed in
 *      (i, j, k, component) order;
 *   3. each listed edge's coefficient from ss^3 sub-samples of its dual
 *      cell: a scalar override, the quantised scalar of the face band, or
 *      the smoothed 3x3 inverse-permittivity tensor.
 * The host then applies the node ids as k-runs of materials_assign_box and
 * the edge records, in list order, through materials_set_component and
 * materials_set_component_tensor3: the materials end in the same state as
 * painting shape by shape and edge by edge on the host.  The public
 * materials API is the only way in, so any library exporting it can be
 * painted.
 */
#include "fdtd_scene.h"
#include "host_internal.h"
#include "rf_device.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int shape_ok(const fdtd_shape_t *s) {
    const double v[] = {s->eps_r, s->p[0], s->p[1], s->p[2], s->radius, s->normal[0], s->normal[1], s->normal[2],
                        s->half_thickness, s->lo[0], s->lo[1], s->lo[2], s->hi[0], s->hi[1], s->hi[2]};
    for (size_t q = 0; q < sizeof v / sizeof v[0]; q++)
        if (!isfinite(v[q])) return 0;
    return s->eps_r >= 1.0;
}

int fdtd_scene_paint(materials_t *mat, const fdtd_grid_t *g, const fdtd_shape_t *shapes, size_t n_shapes,
                     const fdtd_scene_options_t *opt, fdtd_scene_stats_t *stats) {
    if (!mat || !g || (n_shapes && !shapes) || n_shapes > 250) return -1;
    int ss = (opt && opt->supersample > 0) ? opt->supersample : 4;
    if (ss > 16) return -1;
    size_t keep = opt ? opt->keepout : 0;
    if (keep < 2) keep = 2;     /* the tensor stencil needs one cell of margin */
    fdtd_scene_stats_t st = {0, 0, 0, 0};
    const size_t n[3] = {g->nx, g->ny, g->nz};
    const unsigned per = g->periodic_axes;
    if (n[0] > 65535 || n[1] > 65535 || n[2] > 65535 || keep > (1u << 30)) return -1;

    rf_paint_shape_t *sh = calloc(n_shapes ? n_shapes : 1, sizeof *sh);
    material_id_t *mid = calloc(n_shapes ? n_shapes : 1, sizeof *mid);
    if (!sh || !mid) {
        free(sh);
        free(mid);
        return -1;
    }
    int rc = -1;
    rf_stream_t *s = NULL;
    rf_paint_shape_t *d_sh = NULL;
    uint8_t *d_ids = NULL, *d_flags = NULL, *ids = NULL;
    unsigned long long *d_painted = NULL;
    int *d_count = NULL, *count = NULL;
    long long *d_off = NULL, *off = NULL, *d_codes = NULL, *codes = NULL;
    rf_paint_rec_t *d_recs = NULL, *recs = NULL;

    for (size_t q = 0; q < n_shapes; q++) {
        const fdtd_shape_t *in = &shapes[q];
        rf_paint_shape_t *o = &sh[q];
        if (!shape_ok(in)) goto out;
        if (materials_register(mat, in->eps_r, 0.0, &mid[q]) != 0) goto out;
        o->kind = in->kind;
        o->eps_r = in->eps_r;
        o->radius = in->radius;
        o->half_thickness = in->half_thickness;
        for (int a = 0; a < 3; a++) {
            o->p[a] = in->p[a];
            o->lo[a] = in->lo[a];
            o->hi[a] = in->hi[a];
        }
        if (in->kind == FDTD_SHAPE_SPHERE) {
            if (!(in->radius > 0.0)) goto out;
            for (int a = 0; a < 3; a++) {
                o->bb_lo[a] = in->p[a] - in->radius;
                o->bb_hi[a] = in->p[a] + in->radius;
            }
        } else {
            for (int a = 0; a < 3; a++) {
                if (!(in->hi[a] >= in->lo[a])) goto out;
                o->bb_lo[a] = in->lo[a];
                o->bb_hi[a] = in->hi[a];
            }
            if (in->kind == FDTD_SHAPE_SLAB) {
                double l = sqrt(in->normal[0] * in->normal[0] + in->normal[1] * in->normal[1] +
                                in->normal[2] * in->normal[2]);
                if (!(l > 0.0) || !(in->half_thickness > 0.0)) goto out;
                for (int a = 0; a < 3; a++) o->nrm[a] = in->normal[a] / l;
            } else if (in->kind != FDTD_SHAPE_BOX) {
                goto out;
            }
        }
    }
    if (!n_shapes) {
        rc = 0;
        goto out;
    }
    if (rf_init() != 0) {
        rf_set_error("no GPU: %s", rf_error_string());
        goto out;
    }
    s = rf_stream_create();
    if (!s) goto out;

    /* node extents: the padded node exists on non-periodic axes */
    int ext[3], ni[3];
    for (int a = 0; a < 3; a++) {
        ext[a] = ((per >> a) & 1u) ? (int)n[a] : (int)n[a] + 1;
        ni[a] = (int)n[a];
    }
    const size_t nodes = (size_t)ext[0] * ext[1] * ext[2];
    const size_t cells = n[0] * n[1] * n[2];
    const size_t rows = n[0] * n[1];

    /* 1. node ids, applied as k-runs of one material */
    unsigned long long painted = 0;
    if (rf_malloc((void **)&d_sh, n_shapes * sizeof *d_sh) || rf_malloc((void **)&d_ids, nodes) ||
        rf_malloc((void **)&d_painted, sizeof *d_painted) || !(ids = malloc(nodes)))
        goto out;
    if (rf_h2d(d_sh, sh, n_shapes * sizeof *sh, s) ||
        rf_launch_paint_nodes(d_sh, (int)n_shapes, ext, d_ids, d_painted, s) || rf_d2h(ids, d_ids, nodes, s) ||
        rf_d2h(&painted, d_painted, sizeof painted, s) || rf_sync(s))
        goto out;
    rf_free(d_ids);
    d_ids = NULL;
    st.nodes_painted = (size_t)painted;
    for (size_t i = 0; i < (size_t)ext[0]; i++)
        for (size_t j = 0; j < (size_t)ext[1]; j++) {
            const uint8_t *row = ids + (i * ext[1] + j) * ext[2];
            size_t k = 0;
            while (k < (size_t)ext[2]) {
                if (row[k] == 255) {
                    k++;
                    continue;
                }
                const material_id_t m = mid[row[k]];
                const size_t k0 = k;
                while (k < (size_t)ext[2] && row[k] != 255 && mid[row[k]] == m) k++;
                if (materials_assign_box(mat, g, i, i + 1, j, j + 1, k0, k, m) != 0) goto out;
            }
        }
    free(ids);
    ids = NULL;

    /* 2. the edges that may straddle an interface, listed in order */
    if (rf_malloc((void **)&d_flags, cells) || rf_malloc((void **)&d_count, rows * sizeof *d_count) ||
        !(count = malloc(rows * sizeof *count)) || !(off = malloc(rows * sizeof *off)))
        goto out;
    if (rf_launch_paint_flags(d_sh, (int)n_shapes, ni, per, d_flags, s) ||
        rf_launch_paint_rows(d_flags, ni, d_count, s) || rf_d2h(count, d_count, rows * sizeof *count, s) ||
        rf_sync(s))
        goto out;
    long long n_edges = 0;
    for (size_t r = 0; r < rows; r++) {
        off[r] = n_edges;
        n_edges += count[r];
    }
    if (n_edges > 0) {
        /* 3. classify them and apply the records in list order */
        if (rf_malloc((void **)&d_off, rows * sizeof *d_off) ||
            rf_malloc((void **)&d_codes, (size_t)n_edges * sizeof *d_codes) ||
            rf_malloc((void **)&d_recs, (size_t)n_edges * sizeof *d_recs) ||
            !(codes = malloc((size_t)n_edges * sizeof *codes)) ||
            !(recs = malloc((size_t)n_edges * sizeof *recs)))
            goto out;
        if (rf_h2d(d_off, off, rows * sizeof *off, s) || rf_launch_paint_list(d_flags, ni, d_off, d_codes, s) ||
            rf_launch_paint_classify(d_sh, (int)n_shapes, d_codes, n_edges, ni, per, ss, (long long)keep, d_recs,
                                     s) ||
            rf_d2h(codes, d_codes, (size_t)n_edges * sizeof *codes, s) ||
            rf_d2h(recs, d_recs, (size_t)n_edges * sizeof *recs, s) || rf_sync(s))
            goto out;
        for (long long e = 0; e < n_edges; e++) {
            const rf_paint_rec_t *r = &recs[e];
            if (r->kind == 0) continue;
            const long long node = codes[e] / 3;
            const field_component_t comp = (field_component_t)(codes[e] % 3);
            const size_t k = (size_t)(node % ni[2]), j = (size_t)((node / ni[2]) % ni[1]),
                         i = (size_t)(node / ((long long)ni[1] * ni[2]));
            if (r->kind == 3) {
                if (materials_set_component_tensor3(mat, g, comp, i, j, k, r->v) != 0) goto out;
                st.tensor_edges++;
            } else {
                if (materials_set_component(mat, g, comp, i, j, k, r->v[0], 0.0) != 0) goto out;
                if (r->kind == 1)
                    st.override_edges++;
                else
                    st.scalar_mixed_edges++;
            }
        }
    }
    rc = 0;
out:
    if (s) rf_sync(s);
    rf_free(d_sh);
    rf_free(d_ids);
    rf_free(d_painted);
    rf_free(d_flags);
    rf_free(d_count);
    rf_free(d_off);
    rf_free(d_codes);
    rf_free(d_recs);
    if (s) rf_stream_destroy(s);
    free(ids);
    free(count);
    free(off);
    free(codes);
    free(recs);
    free(sh);
    free(mid);
    if (stats) *stats = st;
    return rc;
}
