/*
 * CPU oracle: materials table, overrides, tensors, dispersive (ADE) poles
 * and the materials E update, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_core.c for the usage rule).
 *
 * Reference parts (bitwise parity pinned against the compiled reference):
 *   make_coeff                  engine/fdtd/materials.c:92-125
 *   register / box / overrides  materials.c:388-656
 *   2x2 tensor set/get/clear    materials.c:673-799
 *   tensor pair rule            materials.c:836-903
 *   update_e_impl               materials.c:911-1082
 * Extensions (no reference; pinned by closed forms and invariants in
 * tests/test_oracle_extensions.py):
 *   3x3 tensors   reduction to the 2x2 rule, operator symmetry, diagonal
 *                 tensor == scalar update;
 *   ADE poles     single-cell z-transform closed form of the trapezoidal
 *                 recursion and zero-pole == plain-material update.
 */
#include "materials.h"
#include "constants.h"
#include "oracle_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TABLE_LIMIT 256
#define MAX_SLOTS (2 * MATERIALS_MAX_POLES)

typedef struct {
    double eps_r, sigma;
    material_coeff_t coeff;
    int dispersive;
    material_dispersive_t disp;
    /* derived trapezoidal-ADE coefficients */
    size_t n_poles;
    int has_p[MAX_SLOTS];
    double kj[MAX_SLOTS], kp[MAX_SLOTS], v[MAX_SLOTS];
    double Cu, hdt;
} entry_t;

typedef struct { size_t key; material_coeff_t coeff; } ovr_t;

typedef struct {
    size_t key;
    int kind;           /* 2: reference x-y tensor, 3: full 3x3 */
    double m[6];        /* xx xy xz yy yz zz */
    double scale;
} ten_t;

struct materials {
    size_t nx, ny, nz, total;
    double dx, dt;
    uint8_t *ids;
    entry_t table[TABLE_LIMIT];
    size_t count;
    ovr_t *ovr; size_t n_ovr, cap_ovr;
    ten_t *ten; size_t n_ten, cap_ten;
    /* ADE state: J and P per pole slot, indexed idx*3 + component */
    size_t n_slots;
    double *J[MAX_SLOTS], *P[MAX_SLOTS];
};

/* materials.c:50-59 */
static int matches(const materials_t *m, const fdtd_grid_t *g) {
    return m && g && m->nx == g->nx && m->ny == g->ny && m->nz == g->nz &&
           m->dx == g->dx && m->dt == g->dt;
}

static int comp_no(field_component_t c) {
    return (c == FIELD_EX || c == FIELD_EY || c == FIELD_EZ) ? (int)c : -1;
}

/* materials.c:70-90: component coordinate ranges (periodic axes omit the
 * duplicate endpoint). */
static int comp_ok(const fdtd_grid_t *g, int c, size_t i, size_t j, size_t k) {
    if (!g || c < 0 || c > 2) return 0;
    size_t lim[3] = {g->nx, g->ny, g->nz}, x[3] = {i, j, k};
    for (int a = 0; a < 3; a++) {
        if (a != c && !((g->periodic_axes >> a) & 1u)) lim[a]++;
        if (x[a] >= lim[a]) return 0;
    }
    return 1;
}

/* materials.c:92-125 */
static int coeff_of(const materials_t *m, double eps_r, double sigma, material_coeff_t *out) {
    if (!m || !out || !isfinite(eps_r) || eps_r < 1.0 || !isfinite(sigma) || sigma < 0.0) return -1;
    double eps = eps_r * EPS0;
    double L = (sigma * m->dt) / (2.0 * eps);
    double cb = m->dt / (eps * m->dx);
    if (!isfinite(eps) || eps <= 0.0 || !isfinite(L) || !isfinite(cb) || cb <= 0.0) return -1;
    material_coeff_t r;
    r.Ca = (1.0 - L) / (1.0 + L);
    r.Cb = cb / (1.0 + L);
    if (!isfinite(r.Ca) || !isfinite(r.Cb) || r.Cb <= 0.0) return -1;
    *out = r;
    return 0;
}

/* Binary search for the first key >= key in a sorted array of stride sz. */
static size_t find_key(const void *base, size_t n, size_t sz, size_t key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        size_t mk = *(const size_t *)((const char *)base + mid * sz);
        if (mk < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int grow(void **arr, size_t *cap, size_t n, size_t sz) {
    if (n < *cap) return 0;
    size_t nc = *cap ? *cap * 2 : 16;
    void *p = realloc(*arr, nc * sz);
    if (!p) return -1;
    *arr = p;
    *cap = nc;
    return 0;
}

static const ovr_t *ovr_find(const materials_t *m, size_t key) {
    if (!m->n_ovr) return NULL;
    size_t pos = find_key(m->ovr, m->n_ovr, sizeof(ovr_t), key);
    return (pos < m->n_ovr && m->ovr[pos].key == key) ? &m->ovr[pos] : NULL;
}

static const ten_t *ten_find(const materials_t *m, size_t key) {
    if (!m->n_ten) return NULL;
    size_t pos = find_key(m->ten, m->n_ten, sizeof(ten_t), key);
    return (pos < m->n_ten && m->ten[pos].key == key) ? &m->ten[pos] : NULL;
}

materials_t *materials_create(const fdtd_grid_t *g) {
    if (!g || !g->nx || !g->ny || !g->nz) return NULL;
    if (!(isfinite(g->dx) && g->dx > 0.0 && isfinite(g->dt) && g->dt > 0.0)) return NULL;
    size_t total = (g->nx + 1) * (g->ny + 1) * (g->nz + 1);
    materials_t *m = calloc(1, sizeof *m);
    if (!m) return NULL;
    m->nx = g->nx; m->ny = g->ny; m->nz = g->nz; m->total = total;
    m->dx = g->dx; m->dt = g->dt;
    m->ids = calloc(total, 1);
    material_id_t vac;
    if (!m->ids || materials_register(m, 1.0, 0.0, &vac) != 0) { materials_destroy(m); return NULL; }
    return m;
}

void materials_destroy(materials_t *m) {
    if (!m) return;
    free(m->ids); free(m->ovr); free(m->ten);
    for (size_t s = 0; s < MAX_SLOTS; s++) { free(m->J[s]); free(m->P[s]); }
    free(m);
}

/* materials.c:388-423 */
int materials_register(materials_t *m, double eps_r, double sigma, material_id_t *id_out) {
    material_coeff_t c;
    if (!m || !id_out || coeff_of(m, eps_r, sigma, &c) != 0) return -1;
    for (size_t n = 0; n < m->count; n++)
        if (!m->table[n].dispersive && m->table[n].eps_r == eps_r && m->table[n].sigma == sigma) {
            *id_out = (material_id_t)n;
            return 0;
        }
    if (m->count >= TABLE_LIMIT) return -1;
    entry_t *e = &m->table[m->count];
    memset(e, 0, sizeof *e);
    e->eps_r = eps_r; e->sigma = sigma; e->coeff = c;
    *id_out = (material_id_t)m->count++;
    return 0;
}

/* Trapezoidal ADE coefficients (materials.h describes the scheme). */
int materials_register_dispersive(materials_t *m, const material_dispersive_t *d, material_id_t *id_out) {
    if (!m || !d || !id_out) return -1;
    if (!(isfinite(d->eps_inf) && d->eps_inf >= 1.0 && isfinite(d->sigma) && d->sigma >= 0.0)) return -1;
    if (d->n_drude > MATERIALS_MAX_POLES || d->n_lorentz > MATERIALS_MAX_POLES) return -1;
    if (m->count >= TABLE_LIMIT) return -1;
    entry_t e;
    memset(&e, 0, sizeof e);
    e.dispersive = 1;
    e.disp = *d;
    e.eps_r = d->eps_inf;
    e.sigma = d->sigma;
    double dt = m->dt, dx = m->dx, sumv = 0.0;
    size_t s = 0;
    for (size_t p = 0; p < d->n_drude; p++, s++) {
        double wp = d->drude_wp[p], g = d->drude_gamma[p];
        if (!(isfinite(wp) && wp >= 0.0 && isfinite(g) && g >= 0.0)) return -1;
        double D = 1.0 + 0.5 * g * dt;
        e.has_p[s] = 0;
        e.kj[s] = 2.0 / D;
        e.kp[s] = 0.0;
        e.v[s] = 0.5 * dt * EPS0 * (wp * wp) / D;
        sumv += e.v[s];
    }
    for (size_t p = 0; p < d->n_lorentz; p++, s++) {
        double de = d->lorentz_deps[p], w0 = d->lorentz_w0[p], dl = d->lorentz_delta[p];
        if (!(isfinite(de) && de >= 0.0 && isfinite(w0) && w0 > 0.0 && isfinite(dl) && dl >= 0.0)) return -1;
        double w0sq = w0 * w0;
        double D = 1.0 + 0.5 * (2.0 * dl) * dt + 0.25 * w0sq * dt * dt;
        e.has_p[s] = 1;
        e.kj[s] = 2.0 / D;
        e.kp[s] = dt * w0sq / D;
        e.v[s] = 0.5 * dt * EPS0 * (de * w0sq) / D;
        sumv += e.v[s];
    }
    e.n_poles = s;
    double a = EPS0 * d->eps_inf / dt;
    double b = 0.5 * (d->sigma + sumv);
    e.coeff.Ca = (a - b) / (a + b);
    e.coeff.Cb = 1.0 / ((a + b) * dx);
    e.Cu = 0.5 * e.coeff.Cb * dx;
    e.hdt = 0.5 * dt;
    if (!isfinite(e.coeff.Ca) || !isfinite(e.coeff.Cb) || !(e.coeff.Cb > 0.0)) return -1;
    m->table[m->count] = e;
    *id_out = (material_id_t)m->count++;
    return 0;
}

size_t materials_table_count(const materials_t *m) { return m ? m->count : 0; }
size_t materials_override_count(const materials_t *m) { return m ? m->n_ovr : 0; }
size_t materials_tensor_count(const materials_t *m) { return m ? m->n_ten : 0; }

int materials_get_dispersive(const materials_t *m, material_id_t id, material_dispersive_t *out) {
    if (!m || !out || (size_t)id >= m->count) return -1;
    if (!m->table[id].dispersive) return 1;
    *out = m->table[id].disp;
    return 0;
}

/* materials.c:426-452 */
static int box_ok(const materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1,
                  size_t j0, size_t j1, size_t k0, size_t k1) {
    if (!matches(m, g)) return 0;
    if (i0 >= i1 || j0 >= j1 || k0 >= k1) return 0;
    if (i1 > g->nx + 1 || j1 > g->ny + 1 || k1 > g->nz + 1) return 0;
    if ((g->periodic_axes & 1u) && i0 >= g->nx) return 0;
    if ((g->periodic_axes & 2u) && j0 >= g->ny) return 0;
    if ((g->periodic_axes & 4u) && k0 >= g->nz) return 0;
    return 1;
}

static int key_in_box(const fdtd_grid_t *g, size_t key, const size_t lo[3], const size_t hi[3]) {
    size_t idx = key / 3;
    size_t k = idx % (g->nz + 1), r = idx / (g->nz + 1);
    size_t j = r % (g->ny + 1), i = r / (g->ny + 1);
    return i >= lo[0] && i < hi[0] && j >= lo[1] && j < hi[1] && k >= lo[2] && k < hi[2];
}

/* materials.c:475-527 (the box is clipped before override clearing, so
 * the clearing uses the clipped box as well). */
int materials_assign_box(materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1,
                         size_t j0, size_t j1, size_t k0, size_t k1, material_id_t id) {
    if (!box_ok(m, g, i0, i1, j0, j1, k0, k1) || (size_t)id >= m->count) return -1;
    if ((g->periodic_axes & 1u) && i1 > g->nx) i1 = g->nx;
    if ((g->periodic_axes & 2u) && j1 > g->ny) j1 = g->ny;
    if ((g->periodic_axes & 4u) && k1 > g->nz) k1 = g->nz;
    for (size_t i = i0; i < i1; i++)
        for (size_t j = j0; j < j1; j++)
            for (size_t k = k0; k < k1; k++)
                m->ids[fdtd_index(g, i, j, k)] = id;
    size_t lo[3] = {i0, j0, k0}, hi[3] = {i1, j1, k1};
    size_t w = 0;
    for (size_t n = 0; n < m->n_ovr; n++)
        if (!key_in_box(g, m->ovr[n].key, lo, hi)) m->ovr[w++] = m->ovr[n];
    m->n_ovr = w;
    w = 0;
    for (size_t n = 0; n < m->n_ten; n++)
        if (!key_in_box(g, m->ten[n].key, lo, hi)) m->ten[w++] = m->ten[n];
    m->n_ten = w;
    return 0;
}

int materials_set_box(materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1,
                      size_t j0, size_t j1, size_t k0, size_t k1, double eps_r, double sigma) {
    if (!box_ok(m, g, i0, i1, j0, j1, k0, k1)) return -1;
    material_id_t id;
    if (materials_register(m, eps_r, sigma, &id) != 0) return -1;
    return materials_assign_box(m, g, i0, i1, j0, j1, k0, k1, id);
}

/* materials.c:552-595 */
int materials_set_component(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                            size_t i, size_t j, size_t k, double eps_r, double sigma) {
    int c = comp_no(component);
    material_coeff_t co;
    if (!matches(m, g) || !comp_ok(g, c, i, j, k) || coeff_of(m, eps_r, sigma, &co) != 0) return -1;
    size_t key = fdtd_index(g, i, j, k) * 3 + (size_t)c;
    size_t pos = find_key(m->ovr, m->n_ovr, sizeof(ovr_t), key);
    if (pos < m->n_ovr && m->ovr[pos].key == key) { m->ovr[pos].coeff = co; return 0; }
    if (grow((void **)&m->ovr, &m->cap_ovr, m->n_ovr, sizeof(ovr_t)) != 0) return -1;
    memmove(m->ovr + pos + 1, m->ovr + pos, (m->n_ovr - pos) * sizeof(ovr_t));
    m->ovr[pos].key = key;
    m->ovr[pos].coeff = co;
    m->n_ovr++;
    return 0;
}

int materials_clear_component(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                              size_t i, size_t j, size_t k) {
    int c = comp_no(component);
    if (!matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    size_t key = fdtd_index(g, i, j, k) * 3 + (size_t)c;
    size_t pos = find_key(m->ovr, m->n_ovr, sizeof(ovr_t), key);
    if (pos < m->n_ovr && m->ovr[pos].key == key) {
        memmove(m->ovr + pos, m->ovr + pos + 1, (m->n_ovr - pos - 1) * sizeof(ovr_t));
        m->n_ovr--;
    }
    return 0;
}

/* Scalar coefficient of one edge, ignoring tensors (materials.c:165-180). */
static material_coeff_t scalar_at(const materials_t *m, size_t idx, int c, const entry_t **entry) {
    const ovr_t *o = ovr_find(m, idx * 3 + (size_t)c);
    if (o) { if (entry) *entry = NULL; return o->coeff; }
    const entry_t *e = &m->table[m->ids[idx]];
    if (entry) *entry = e;
    return e->coeff;
}

int materials_get_coefficients(const materials_t *m, const fdtd_grid_t *g, field_component_t component,
                               size_t i, size_t j, size_t k, material_coeff_t *out) {
    int c = comp_no(component);
    if (!out || !matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    *out = scalar_at(m, fdtd_index(g, i, j, k), c, NULL);
    return 0;
}

/* materials.c:277-308 */
static int margin_ok(const fdtd_grid_t *g, int axes_mask, size_t i, size_t j, size_t k) {
    size_t x[3] = {i, j, k}, n[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; a++) {
        if (!((axes_mask >> a) & 1) || ((g->periodic_axes >> a) & 1u)) continue;
        if (n[a] < 3 || x[a] == 0 || x[a] >= n[a] - 1) return 0;
    }
    return 1;
}

static int spd2(double a, double b, double c) {
    if (!isfinite(a) || !isfinite(b) || !isfinite(c) || a <= 0.0 || c <= 0.0) return 0;
    return fabs(b) / sqrt(a) < sqrt(c);
}

/* Sylvester's criterion on the symmetric 3x3. */
static int spd3(const double m[6]) {
    for (int q = 0; q < 6; q++) if (!isfinite(m[q])) return 0;
    double xx = m[0], xy = m[1], xz = m[2], yy = m[3], yz = m[4], zz = m[5];
    if (!(xx > 0.0)) return 0;
    double d2 = xx * yy - xy * xy;
    if (!(d2 > 0.0)) return 0;
    double d3 = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    return d3 > 0.0;
}

static int ten_insert(materials_t *m, size_t key, const ten_t *t) {
    size_t pos = find_key(m->ten, m->n_ten, sizeof(ten_t), key);
    if (pos < m->n_ten && m->ten[pos].key == key) { m->ten[pos] = *t; m->ten[pos].key = key; return 0; }
    if (grow((void **)&m->ten, &m->cap_ten, m->n_ten, sizeof(ten_t)) != 0) return -1;
    memmove(m->ten + pos + 1, m->ten + pos, (m->n_ten - pos) * sizeof(ten_t));
    m->ten[pos] = *t;
    m->ten[pos].key = key;
    m->n_ten++;
    return 0;
}

/* materials.c:673-738 */
int materials_set_component_tensor(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                   size_t i, size_t j, size_t k, double m11, double m12, double m22) {
    int c = comp_no(component);
    if (c != 0 && c != 1) return -1;
    if (!matches(m, g) || !comp_ok(g, c, i, j, k) || !margin_ok(g, 3, i, j, k) || !spd2(m11, m12, m22))
        return -1;
    double scale = m->dt / (EPS0 * m->dx);
    if (!isfinite(scale) || scale <= 0.0) return -1;
    ten_t t;
    memset(&t, 0, sizeof t);
    t.kind = 2;
    t.scale = scale;
    /* component-local (m11, m12, m22) -> global xx xy xz yy yz zz */
    if (c == 0) { t.m[0] = m11; t.m[1] = m12; t.m[3] = m22; }
    else        { t.m[0] = m22; t.m[1] = m12; t.m[3] = m11; }
    t.m[5] = 1.0;
    return ten_insert(m, fdtd_index(g, i, j, k) * 3 + (size_t)c, &t);
}

int materials_set_component_tensor3(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                    size_t i, size_t j, size_t k, const double mm[6]) {
    int c = comp_no(component);
    if (c < 0 || !mm || !matches(m, g) || !comp_ok(g, c, i, j, k) || !margin_ok(g, 7, i, j, k) || !spd3(mm))
        return -1;
    double scale = m->dt / (EPS0 * m->dx);
    if (!isfinite(scale) || scale <= 0.0) return -1;
    ten_t t;
    memset(&t, 0, sizeof t);
    t.kind = 3;
    t.scale = scale;
    memcpy(t.m, mm, sizeof t.m);
    return ten_insert(m, fdtd_index(g, i, j, k) * 3 + (size_t)c, &t);
}

int materials_clear_component_tensor(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                     size_t i, size_t j, size_t k) {
    int c = comp_no(component);
    if (!matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    size_t key = fdtd_index(g, i, j, k) * 3 + (size_t)c;
    size_t pos = find_key(m->ten, m->n_ten, sizeof(ten_t), key);
    if (pos < m->n_ten && m->ten[pos].key == key) {
        memmove(m->ten + pos, m->ten + pos + 1, (m->n_ten - pos - 1) * sizeof(ten_t));
        m->n_ten--;
    }
    return 0;
}

/* Diagonal index of component c in xx xy xz yy yz zz. */
static const int DIAG[3] = {0, 3, 5};
static int offdiag(int a, int b) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 1 : 2) : 4;
}

int materials_get_tensor(const materials_t *m, const fdtd_grid_t *g, field_component_t component,
                         size_t i, size_t j, size_t k, material_tensor_t *out) {
    int c = comp_no(component);
    if (!out || !matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    const ten_t *t = ten_find(m, fdtd_index(g, i, j, k) * 3 + (size_t)c);
    if (!t) return 1;
    int other = c == 1 ? 0 : 1;
    out->m11 = t->m[DIAG[c]];
    out->m12 = c == 2 ? t->m[offdiag(2, 0)] : t->m[1];
    out->m22 = t->m[DIAG[c == 2 ? 0 : other]];
    out->scale = t->scale;
    return 0;
}

int materials_get_tensor3(const materials_t *m, const fdtd_grid_t *g, field_component_t component,
                          size_t i, size_t j, size_t k, double out[6]) {
    int c = comp_no(component);
    if (!out || !matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    const ten_t *t = ten_find(m, fdtd_index(g, i, j, k) * 3 + (size_t)c);
    if (!t) return 1;
    memcpy(out, t->m, sizeof t->m);
    return 0;
}

/* ------------------------------------------------------- tensor update */

/* Normalised correlation r of the pair (c, d) at one edge, dividing by the
 * edge's own diagonal first (materials.c:838-839). */
static double corr(const ten_t *t, int c, int d) {
    return (t->m[offdiag(c, d)] / sqrt(t->m[DIAG[c]])) / sqrt(t->m[DIAG[d]]);
}

/* materials.c:836-845, generalised to the pair (a, b) with a < b. */
static double pair_weight(const ten_t *ta, int a, const ten_t *tb, int b) {
    double rho = 0.5 * corr(ta, a, b) + 0.5 * corr(tb, b, a);
    return (0.25 * rho * sqrt(ta->m[DIAG[a]])) * sqrt(tb->m[DIAG[b]]);
}

typedef struct {
    const fdtd_grid_t *g;
    oracle_geom_t o;
} tctx_t;

static size_t nxt(const tctx_t *t, int ax, size_t q) { return q + 1 == t->o.n[ax] ? 0 : q + 1; }
static size_t prv(const tctx_t *t, int ax, size_t q) { return q == 0 ? t->o.n[ax] - 1 : q - 1; }

/* Full discrete curl of H at the E_c edge x (no CPML). */
static double curl_full(const tctx_t *t, int c, const size_t x[3]) {
    int p = (c + 1) % 3, q = (c + 2) % 3;
    fdtd_grid_t *g = (fdtd_grid_t *)t->g;
    const double *Fq = h_of(g, q), *Fp = h_of(g, p);
    size_t xp[3] = {x[0], x[1], x[2]}, xq[3] = {x[0], x[1], x[2]};
    xp[p] = prv(t, p, x[p]);
    xq[q] = prv(t, q, x[q]);
    size_t idx = lin(&t->o, x);
    double dA = Fq[idx] - Fq[lin(&t->o, xp)];
    double dB = Fp[idx] - Fp[lin(&t->o, xq)];
    return dA - dB;
}


/* Neighbour d-edges of the c-edge at x: the d-edge positions
 * x + s_c e_c - s_d e_d for s_c, s_d in {0, 1}, with the offset along the
 * lower-numbered of the two axes varying fastest.  For the x-y pair this
 * is the reference order (materials.c:856-857 for Ex, 879-880 for Ey). */
static void neighbours(const tctx_t *t, int c, int d, const size_t x[3], size_t out[4][3]) {
    int lo = c < d ? c : d;
    for (int q = 0; q < 4; q++) {
        int s_lo = q & 1, s_hi = (q >> 1) & 1;
        int s_c = (c == lo) ? s_lo : s_hi;
        int s_d = (d == lo) ? s_lo : s_hi;
        out[q][0] = x[0]; out[q][1] = x[1]; out[q][2] = x[2];
        if (s_c) out[q][c] = nxt(t, c, x[c]);
        if (s_d) out[q][d] = prv(t, d, x[d]);
    }
}

/* Hz-only curls of the reference's z-invariant stencil
 * (materials.c:856-901): Ex uses dHz/dy, Ey uses -dHz/dx. */
static double curl_hz(const tctx_t *t, int c, const size_t x[3]) {
    const double *Hz = t->g->hz;
    size_t y[3] = {x[0], x[1], x[2]};
    if (c == 0) {
        y[1] = prv(t, 1, x[1]);
        return Hz[lin(&t->o, x)] - Hz[lin(&t->o, y)];
    }
    y[0] = prv(t, 0, x[0]);
    return -(Hz[lin(&t->o, x)] - Hz[lin(&t->o, y)]);
}

static void tensor_edge(const materials_t *m, const tctx_t *t, const ten_t *te, int c,
                        const size_t x[3], double *E) {
    size_t idx = lin(&t->o, x);
    double value;
    if (te->kind == 2) {
        const double *Hz = t->g->hz;
        size_t y[3] = {x[0], x[1], x[2]};
        if (c == 0) {
            y[1] = prv(t, 1, x[1]);
            value = te->m[DIAG[0]] * (Hz[idx] - Hz[lin(&t->o, y)]);
        } else {
            y[0] = prv(t, 0, x[0]);
            value = -te->m[DIAG[1]] * (Hz[idx] - Hz[lin(&t->o, y)]);
        }
    } else {
        value = te->m[DIAG[c]] * curl_full(t, c, x);
    }
    for (int d = 0; d < 3; d++) {
        if (d == c) continue;
        if (te->kind == 2 && d == 2) continue;
        size_t nb[4][3];
        neighbours(t, c, d, x, nb);
        for (int q = 0; q < 4; q++) {
            const ten_t *o = ten_find(m, lin(&t->o, nb[q]) * 3 + (size_t)d);
            if (!o) continue;
            if ((c == 2 || d == 2) && o->kind != 3) continue;
            double w = c < d ? pair_weight(te, c, o, d) : pair_weight(o, d, te, c);
            double cd = te->kind == 2 ? curl_hz(t, d, nb[q]) : curl_full(t, d, nb[q]);
            value += w * cd;
        }
    }
    E[idx] += te->scale * value;
}

static int ensure_ade(materials_t *m) {
    size_t need = 0;
    for (size_t n = 0; n < m->count; n++)
        if (m->table[n].dispersive && m->table[n].n_poles > need) need = m->table[n].n_poles;
    for (size_t s = 0; s < need; s++) {
        if (!m->J[s]) m->J[s] = calloc(3 * m->total, sizeof(double));
        if (!m->P[s]) m->P[s] = calloc(3 * m->total, sizeof(double));
        if (!m->J[s] || !m->P[s]) return -1;
    }
    if (need > m->n_slots) m->n_slots = need;
    return 0;
}

/* materials.c:911-1060, plus the tensor-3 and ADE extensions. */
static void update_e(fdtd_grid_t *g, cpml_t *pml, const materials_t *mc) {
    materials_t *m = (materials_t *)mc;   /* ADE state advances in place */
    if (ensure_ade(m) != 0) return;
    tctx_t t;
    t.g = g;
    t.o = geom_of(g);
    oracle_stretch_t st[3][2];
    if (pml) oracle_cpml_stretch_e(pml, st);
    const double dx = g->dx;
    for (int c = 0; c < 3; c++) {
        int p = (c + 1) % 3, q = (c + 2) % 3;
        double *E = e_of(g, c);
        const double *Fq = h_of(g, q), *Fp = h_of(g, p);
        size_t lo[3], hi[3];
        e_range(&t.o, c, lo, hi);
        size_t x[3];
        for (x[0] = lo[0]; x[0] < hi[0]; x[0]++)
        for (x[1] = lo[1]; x[1] < hi[1]; x[1]++)
        for (x[2] = lo[2]; x[2] < hi[2]; x[2]++) {
            size_t idx = lin(&t.o, x);
            const ten_t *te = ten_find(m, idx * 3 + (size_t)c);
            if (te) { tensor_edge(m, &t, te, c, x, E); continue; }
            const entry_t *ent;
            material_coeff_t co = scalar_at(m, idx, c, &ent);
            size_t xp[3] = {x[0], x[1], x[2]}, xq[3] = {x[0], x[1], x[2]};
            xp[p] = wrap_down(&t.o, p, x[p]);
            xq[q] = wrap_down(&t.o, q, x[q]);
            double dA = Fq[idx] - Fq[lin(&t.o, xp)];
            double dB = Fp[idx] - Fp[lin(&t.o, xq)];
            if (pml) {
                const oracle_stretch_t *sa = &st[c][0], *sb = &st[c][1];
                dA = cpml_term(dA, sa->kinv[x[p]], sa->b[x[p]], sa->a[x[p]], dx, &sa->psi[idx]);
                dB = cpml_term(dB, sb->kinv[x[q]], sb->b[x[q]], sb->a[x[q]], dx, &sb->psi[idx]);
            }
            if (!ent || !ent->dispersive) {
                E[idx] = co.Ca * E[idx] + co.Cb * (dA - dB);
                continue;
            }
            /* trapezoidal ADE edge (materials.h) */
            size_t sidx = idx * 3 + (size_t)c;
            double u[MAX_SLOTS], sum_u = 0.0;
            for (size_t s = 0; s < ent->n_poles; s++) {
                u[s] = ent->has_p[s] ? ent->kj[s] * m->J[s][sidx] - ent->kp[s] * m->P[s][sidx]
                                     : ent->kj[s] * m->J[s][sidx];
                sum_u += u[s];
            }
            double e_old = E[idx];
            double e_new = co.Ca * e_old + co.Cb * (dA - dB) - ent->Cu * sum_u;
            E[idx] = e_new;
            for (size_t s = 0; s < ent->n_poles; s++) {
                double S = u[s] + ent->v[s] * (e_new + e_old);
                m->J[s][sidx] = S - m->J[s][sidx];
                if (ent->has_p[s]) m->P[s][sidx] = m->P[s][sidx] + ent->hdt * S;
            }
        }
    }
}

/* materials.c:1062-1082 */
void fdtd_update_e_materials(fdtd_grid_t *g, const materials_t *m) {
    if (!matches(m, g)) return;
    update_e(g, NULL, m);
}

void fdtd_update_e_cpml_materials(fdtd_grid_t *g, cpml_t *pml, const materials_t *m) {
    if (!matches(m, g) || !pml || pml->dx != g->dx) return;
    update_e(g, pml, m);
}

/* Oracle-only introspection for tests: ADE state of one edge and slot. */
int oracle_materials_ade_state(const materials_t *m, size_t idx, int c, size_t slot, double *J, double *P) {
    if (!m || slot >= m->n_slots || c < 0 || c > 2 || idx >= m->total) return -1;
    *J = m->J[slot][idx * 3 + (size_t)c];
    *P = m->P[slot][idx * 3 + (size_t)c];
    return 0;
}

/* Deep copy for the oracle's fdtd_sim snapshot semantics. */
materials_t *oracle_materials_clone(const materials_t *m, const fdtd_grid_t *g) {
    if (!matches(m, g)) return NULL;
    materials_t *c = calloc(1, sizeof *c);
    if (!c) return NULL;
    *c = *m;
    c->ids = malloc(m->total);
    c->ovr = malloc((m->n_ovr ? m->n_ovr : 1) * sizeof(ovr_t));
    c->ten = malloc((m->n_ten ? m->n_ten : 1) * sizeof(ten_t));
    c->cap_ovr = m->n_ovr ? m->n_ovr : 1;
    c->cap_ten = m->n_ten ? m->n_ten : 1;
    for (size_t s = 0; s < MAX_SLOTS; s++) { c->J[s] = NULL; c->P[s] = NULL; }
    if (!c->ids || !c->ovr || !c->ten) { materials_destroy(c); return NULL; }
    memcpy(c->ids, m->ids, m->total);
    memcpy(c->ovr, m->ovr, m->n_ovr * sizeof(ovr_t));
    memcpy(c->ten, m->ten, m->n_ten * sizeof(ten_t));
    for (size_t s = 0; s < m->n_slots; s++) {
        c->J[s] = malloc(3 * m->total * sizeof(double));
        c->P[s] = malloc(3 * m->total * sizeof(double));
        if (!c->J[s] || !c->P[s]) { materials_destroy(c); return NULL; }
        memcpy(c->J[s], m->J[s], 3 * m->total * sizeof(double));
        memcpy(c->P[s], m->P[s], 3 * m->total * sizeof(double));
    }
    return c;
}
