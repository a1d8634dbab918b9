/*
 * Scene painter with subpixel smoothing (include/fdtd_scene.h): the
 * oracle's host restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_core.c).  The product paints on the
 * device (rfsyn_amd/csrc/scene_kernels.hip) and must leave the materials
 * in exactly the state this painter leaves them in.  Written against the
 * public materials API only.  O(shape volume) for node painting and
 * O(interface edges x samples) for smoothing.  The painter is an extension
 * (the reference has 2D rectangle helpers only, subpixel_te.c:27-132), so
 * it is pinned by closed forms of the mixing rule (tests/test_scene.py)
 * rather than by reference vectors.
 */
#include "fdtd_scene.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    fdtd_shape_t s;
    double nrm[3];
    double bb_lo[3], bb_hi[3];     /* bounding box in cell units */
    material_id_t id;
} shape_t;

static int inside(const shape_t *sh, const double x[3]) {
    const fdtd_shape_t *s = &sh->s;
    if (s->kind == FDTD_SHAPE_SPHERE) {
        double d0 = x[0] - s->p[0], d1 = x[1] - s->p[1], d2 = x[2] - s->p[2];
        return d0 * d0 + d1 * d1 + d2 * d2 <= s->radius * s->radius;
    }
    for (int a = 0; a < 3; a++)
        if (x[a] < s->lo[a] || x[a] > s->hi[a]) return 0;
    if (s->kind == FDTD_SHAPE_BOX) return 1;
    double d = sh->nrm[0] * (x[0] - s->p[0]) + sh->nrm[1] * (x[1] - s->p[1]) + sh->nrm[2] * (x[2] - s->p[2]);
    return fabs(d) <= s->half_thickness;
}

/* 1 when the unit cube centred at x may contain a boundary of the shape. */
static int near_boundary(const shape_t *sh, const double x[3]) {
    const fdtd_shape_t *s = &sh->s;
    for (int a = 0; a < 3; a++)
        if (x[a] < sh->bb_lo[a] - 1.0 || x[a] > sh->bb_hi[a] + 1.0) return 0;
    if (s->kind == FDTD_SHAPE_SPHERE) {
        double d0 = x[0] - s->p[0], d1 = x[1] - s->p[1], d2 = x[2] - s->p[2];
        return fabs(sqrt(d0 * d0 + d1 * d1 + d2 * d2) - s->radius) <= 1.0;
    }
    for (int a = 0; a < 3; a++)
        if (fabs(x[a] - s->lo[a]) <= 1.0 || fabs(x[a] - s->hi[a]) <= 1.0) return 1;
    if (s->kind == FDTD_SHAPE_BOX) return 0;
    double d = sh->nrm[0] * (x[0] - s->p[0]) + sh->nrm[1] * (x[1] - s->p[1]) + sh->nrm[2] * (x[2] - s->p[2]);
    return fabs(fabs(d) - s->half_thickness) <= 1.0;
}

static void normal_at(const shape_t *sh, const double x[3], double n[3]) {
    const fdtd_shape_t *s = &sh->s;
    if (s->kind == FDTD_SHAPE_SPHERE) {
        double d[3] = {x[0] - s->p[0], x[1] - s->p[1], x[2] - s->p[2]};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r == 0.0) { n[0] = 0; n[1] = 0; n[2] = 1; return; }
        for (int a = 0; a < 3; a++) n[a] = d[a] / r;
        return;
    }
    if (s->kind == FDTD_SHAPE_SLAB) {
        /* the slab faces, unless a clip face is nearer */
        double d = sh->nrm[0] * (x[0] - s->p[0]) + sh->nrm[1] * (x[1] - s->p[1]) + sh->nrm[2] * (x[2] - s->p[2]);
        double best = fabs(fabs(d) - s->half_thickness);
        memcpy(n, sh->nrm, sizeof sh->nrm);
        for (int a = 0; a < 3; a++) {
            double dl = fabs(x[a] - s->lo[a]), dh = fabs(x[a] - s->hi[a]);
            double m = dl < dh ? dl : dh;
            if (m < best) { best = m; n[0] = n[1] = n[2] = 0.0; n[a] = 1.0; }
        }
        return;
    }
    double best = INFINITY;
    for (int a = 0; a < 3; a++) {
        double dl = fabs(x[a] - s->lo[a]), dh = fabs(x[a] - s->hi[a]);
        double m = dl < dh ? dl : dh;
        if (m < best) { best = m; n[0] = n[1] = n[2] = 0.0; n[a] = 1.0; }
    }
}

/* Index of the topmost shape containing x among cand[0..nc), or -1. */
static int material_at(const shape_t *sh, const int *cand, int nc, const double x[3]) {
    for (int q = nc - 1; q >= 0; q--)
        if (inside(&sh[cand[q]], x)) return cand[q];
    return -1;
}

int fdtd_scene_paint(materials_t *mat, const fdtd_grid_t *g, const fdtd_shape_t *shapes, size_t n_shapes,
                     const fdtd_scene_options_t *opt, fdtd_scene_stats_t *stats) {
    if (!mat || !g || (n_shapes && !shapes) || n_shapes > 250) return -1;
    int ss = (opt && opt->supersample > 0) ? opt->supersample : 4;
    if (ss > 16) return -1;
    size_t keep = opt ? opt->keepout : 0;
    if (keep < 2) keep = 2;     /* the tensor stencil needs one cell of margin */
    fdtd_scene_stats_t st = {0, 0, 0, 0};
    size_t n[3] = {g->nx, g->ny, g->nz};
    unsigned per = g->periodic_axes;

    shape_t *sh = calloc(n_shapes ? n_shapes : 1, sizeof *sh);
    if (!sh) return -1;
    int rc = -1;
    uint8_t *flag = NULL;
    int *node_mat = NULL;
    for (size_t q = 0; q < n_shapes; q++) {
        sh[q].s = shapes[q];
        const fdtd_shape_t *s = &shapes[q];
        if (!(isfinite(s->eps_r) && s->eps_r >= 1.0)) goto out;
        if (materials_register(mat, s->eps_r, 0.0, &sh[q].id) != 0) goto out;
        if (s->kind == FDTD_SHAPE_SPHERE) {
            if (!(s->radius > 0.0)) goto out;
            for (int a = 0; a < 3; a++) {
                sh[q].bb_lo[a] = s->p[a] - s->radius;
                sh[q].bb_hi[a] = s->p[a] + s->radius;
            }
        } else {
            for (int a = 0; a < 3; a++) {
                if (!(s->hi[a] >= s->lo[a])) goto out;
                sh[q].bb_lo[a] = s->lo[a];
                sh[q].bb_hi[a] = s->hi[a];
            }
            if (s->kind == FDTD_SHAPE_SLAB) {
                double l = sqrt(s->normal[0] * s->normal[0] + s->normal[1] * s->normal[1] + s->normal[2] * s->normal[2]);
                if (!(l > 0.0) || !(s->half_thickness > 0.0)) goto out;
                for (int a = 0; a < 3; a++) sh[q].nrm[a] = s->normal[a] / l;
            } else if (s->kind != FDTD_SHAPE_BOX) {
                goto out;
            }
        }
    }

    /* node extents: the padded node exists on non-periodic axes */
    size_t ext[3];
    for (int a = 0; a < 3; a++) ext[a] = ((per >> a) & 1u) ? n[a] : n[a] + 1;
    size_t stride_j = g->nz + 1, stride_i = (g->ny + 1) * (g->nz + 1);
    size_t total = (g->nx + 1) * stride_i;
    node_mat = malloc(total * sizeof *node_mat);
    flag = calloc(total, 1);
    if (!node_mat || !flag) goto out;
    for (size_t q = 0; q < total; q++) node_mat[q] = -1;

    /* 1. node ids: material at the node position, later shapes win */
    for (size_t q = 0; q < n_shapes; q++) {
        size_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            double l = floor(sh[q].bb_lo[a]), h = ceil(sh[q].bb_hi[a]);
            lo[a] = l < 0 ? 0 : (size_t)l;
            hi[a] = h < 0 ? 0 : (size_t)h + 1;
            if (hi[a] > ext[a]) hi[a] = ext[a];
        }
        for (size_t i = lo[0]; i < hi[0]; i++)
            for (size_t j = lo[1]; j < hi[1]; j++) {
                size_t k = lo[2];
                while (k < hi[2]) {
                    double x[3] = {(double)i, (double)j, (double)k};
                    if (!inside(&sh[q], x)) { k++; continue; }
                    size_t k0 = k;
                    while (k < hi[2]) {
                        double y[3] = {(double)i, (double)j, (double)k};
                        if (!inside(&sh[q], y)) break;
                        node_mat[i * stride_i + j * stride_j + k] = (int)q;
                        k++;
                    }
                    if (materials_assign_box(mat, g, i, i + 1, j, j + 1, k0, k, sh[q].id) != 0) goto out;
                    st.nodes_painted += k - k0;
                }
            }
    }

    /* 2. edges whose dual cell may cross a boundary (bit c per node) */
    for (size_t q = 0; q < n_shapes; q++) {
        size_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            double l = floor(sh[q].bb_lo[a]) - 2, h = ceil(sh[q].bb_hi[a]) + 2;
            lo[a] = l < 0 ? 0 : (size_t)l;
            hi[a] = h < 0 ? 0 : (size_t)h + 1;
            if (hi[a] > n[a]) hi[a] = n[a];
        }
        for (size_t i = lo[0]; i < hi[0]; i++)
            for (size_t j = lo[1]; j < hi[1]; j++)
                for (size_t k = lo[2]; k < hi[2]; k++)
                    for (int c = 0; c < 3; c++) {
                        double x[3] = {(double)i, (double)j, (double)k};
                        x[c] += 0.5;
                        if (near_boundary(&sh[q], x)) flag[i * stride_i + j * stride_j + k] |= (uint8_t)(1u << c);
                    }
    }

    /* 3. classify the flagged edges by sub-sampling their dual cells */
    int cand[256];
    int counts[257];
    for (size_t i = 0; i < n[0]; i++)
        for (size_t j = 0; j < n[1]; j++)
            for (size_t k = 0; k < n[2]; k++) {
                size_t idx = i * stride_i + j * stride_j + k;
                if (!flag[idx]) continue;
                for (int c = 0; c < 3; c++) {
                    if (!((flag[idx] >> c) & 1u)) continue;
                    /* the E_c edge exists only inside its component range */
                    size_t x_int[3] = {i, j, k};
                    int valid = 1;
                    for (int a = 0; a < 3; a++)
                        if (a != c && !((per >> a) & 1u) && (x_int[a] == 0 || x_int[a] >= n[a])) valid = 0;
                    if (!valid) continue;
                    double xe[3] = {(double)i, (double)j, (double)k};
                    xe[c] += 0.5;
                    int nc = 0;
                    for (size_t q = 0; q < n_shapes; q++) {
                        int hit = 1;
                        for (int a = 0; a < 3; a++)
                            if (xe[a] < sh[q].bb_lo[a] - 1.0 || xe[a] > sh[q].bb_hi[a] + 1.0) hit = 0;
                        if (hit) cand[nc++] = (int)q;
                    }
                    memset(counts, 0, sizeof counts);
                    for (int a0 = 0; a0 < ss; a0++)
                        for (int a1 = 0; a1 < ss; a1++)
                            for (int a2 = 0; a2 < ss; a2++) {
                                double y[3] = {xe[0] + ((a0 + 0.5) / ss - 0.5), xe[1] + ((a1 + 0.5) / ss - 0.5),
                                               xe[2] + ((a2 + 0.5) / ss - 0.5)};
                                counts[material_at(sh, cand, nc, y) + 1]++;
                            }
                    /* the two most frequent materials (index -1 = background) */
                    int m1 = -2, m2 = -2;
                    for (int m = -1; m < (int)n_shapes; m++) {
                        if (!counts[m + 1]) continue;
                        if (m1 == -2 || counts[m + 1] > counts[m1 + 1]) { m2 = m1; m1 = m; }
                        else if (m2 == -2 || counts[m + 1] > counts[m2 + 1]) m2 = m;
                    }
                    int nm = node_mat[idx];
                    field_component_t comp = (field_component_t)c;
                    if (m2 == -2) {
                        if (m1 != nm) {
                            double eps = m1 < 0 ? 1.0 : sh[m1].s.eps_r;
                            if (materials_set_component(mat, g, comp, i, j, k, eps, 0.0) != 0) goto out;
                            st.override_edges++;
                        }
                        continue;
                    }
                    int in = m1 > m2 ? m1 : m2, outm = m1 > m2 ? m2 : m1;
                    double e_in = sh[in].s.eps_r, e_out = outm < 0 ? 1.0 : sh[outm].s.eps_r;
                    double f = (double)counts[in + 1] / (double)(counts[in + 1] + counts[outm + 1]);
                    double nv[3];
                    normal_at(&sh[in], xe, nv);
                    double ln = (1.0 - f) / e_out + f / e_in;
                    double lt = 1.0 / ((1.0 - f) * e_out + f * e_in);
                    double M[3][3];
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            M[a][b] = (ln - lt) * nv[a] * nv[b] + (a == b ? lt : 0.0);
                    int near_face = 0;
                    for (int a = 0; a < 3; a++)
                        if (!((per >> a) & 1u) && (xe[a] < (double)keep || xe[a] > (double)n[a] - (double)keep))
                            near_face = 1;
                    if (near_face) {
                        /* scalar 1/M_cc from the fill fraction and normal
                         * quantised to 1/64 steps, so the band holds a small
                         * set of distinct coefficients (compact device tags) */
                        double fq = floor(f * 64.0 + 0.5) / 64.0;
                        double nc = floor(nv[c] * nv[c] * 64.0 + 0.5) / 64.0;
                        double lnq = (1.0 - fq) / e_out + fq / e_in;
                        double ltq = 1.0 / ((1.0 - fq) * e_out + fq * e_in);
                        double mcc = (lnq - ltq) * nc + ltq;
                        double eps_c = 1.0 / mcc < 1.0 ? 1.0 : 1.0 / mcc;
                        if (materials_set_component(mat, g, comp, i, j, k, eps_c, 0.0) != 0) goto out;
                        st.scalar_mixed_edges++;
                        continue;
                    }
                    double m6[6] = {M[0][0], M[0][1], M[0][2], M[1][1], M[1][2], M[2][2]};
                    if (materials_set_component_tensor3(mat, g, comp, i, j, k, m6) != 0) goto out;
                    st.tensor_edges++;
                }
            }
    rc = 0;
out:
    free(sh);
    free(flag);
    free(node_mat);
    if (stats) *stats = st;
    return rc;
}
