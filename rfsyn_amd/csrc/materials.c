/*
 * Materials: host description with the reference semantics and its dense
 * device form for the E kernels.
 *
 * Replaces engine/fdtd/materials.c.  Host side: a 256-entry table and one
 * id per padded node (materials.c:36-38), per-component scalar overrides
 * and tensors in hash maps keyed idx*3+c (the reference keeps them in
 * sorted arrays with binary search, materials.c:127-199; a hash map makes
 * scene set-up linear instead of quadratic).  Device side: one tag per node
 * naming an entry that holds the three components' (Ca, Cb), a "tensor
 * edge" flag and a dispersive-material index, so the kernel does one tag
 * load and one small-table gather per node instead of binary searches.
 */
#include "host_internal.h"
#include "constants.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define TABLE_LIMIT 256
#define MAX_SLOTS RF_MAX_SLOTS

typedef struct {
    double eps_r, sigma;
    material_coeff_t coeff;
    int dispersive;
    material_dispersive_t disp;
    int n_poles;
    int has_p[MAX_SLOTS];
    double kj[MAX_SLOTS], kp[MAX_SLOTS], v[MAX_SLOTS];
    double Cu, hdt;
} mentry_t;

typedef struct {
    int kind;          /* 2: reference x-y tensor; 3: full 3x3 */
    double m[6];       /* xx xy xz yy yz zz */
    double scale;
} mtensor_t;

struct materials {
    size_t nx, ny, nz, total;
    double dx, dt;
    uint8_t *ids;
    mentry_t table[TABLE_LIMIT];
    size_t count;
    rf_map_t ovr;                     /* key -> index into ovr_vals */
    material_coeff_t *ovr_vals; size_t n_ovr_vals, cap_ovr_vals;
    rf_map_t ten;                     /* key -> index into ten_vals */
    mtensor_t *ten_vals; size_t n_ten_vals, cap_ten_vals;
    unsigned long version;
    rf_mat_dev_t *dev;                /* compat FP64 mirror */
    unsigned long dev_version;
    unsigned dev_per;
};

static const int DIAG[3] = {0, 3, 5};
static int offdiag(int a, int b) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 1 : 2) : 4;
}

int rf_materials_matches(const materials_t *m, const fdtd_grid_t *g) {
    return m && g && m->nx == g->nx && m->ny == g->ny && m->nz == g->nz && m->dx == g->dx &&
           m->dt == g->dt;
}

static int comp_no(field_component_t c) {
    return (c == FIELD_EX || c == FIELD_EY || c == FIELD_EZ) ? (int)c : -1;
}

/* materials.c:70-90 */
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

static size_t node_index(const fdtd_grid_t *g, size_t i, size_t j, size_t k) {
    return fdtd_index(g, i, j, k);
}

materials_t *materials_create(const fdtd_grid_t *g) {
    if (!g || !g->nx || !g->ny || !g->nz) return NULL;
    if (!(isfinite(g->dx) && g->dx > 0.0 && isfinite(g->dt) && g->dt > 0.0)) return NULL;
    materials_t *m = calloc(1, sizeof *m);
    if (!m) return NULL;
    m->nx = g->nx; m->ny = g->ny; m->nz = g->nz;
    m->total = rf_cells(g);
    m->dx = g->dx; m->dt = g->dt;
    m->ids = calloc(m->total, 1);
    material_id_t vac;
    if (!m->ids || materials_register(m, 1.0, 0.0, &vac) != 0) { materials_destroy(m); return NULL; }
    return m;
}

void materials_destroy(materials_t *m) {
    if (!m) return;
    if (m->dev) { rf_materials_release(m->dev); free(m->dev); }
    rf_map_free(&m->ovr);
    rf_map_free(&m->ten);
    free(m->ovr_vals);
    free(m->ten_vals);
    free(m->ids);
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
    mentry_t *e = &m->table[m->count];
    memset(e, 0, sizeof *e);
    e->eps_r = eps_r; e->sigma = sigma; e->coeff = c;
    *id_out = (material_id_t)m->count++;
    m->version++;
    return 0;
}

/* Trapezoidal ADE coefficients; the scheme is stated in materials.h. */
int materials_register_dispersive(materials_t *m, const material_dispersive_t *d, material_id_t *id_out) {
    if (!m || !d || !id_out) return -1;
    if (!(isfinite(d->eps_inf) && d->eps_inf >= 1.0 && isfinite(d->sigma) && d->sigma >= 0.0)) return -1;
    if (d->n_drude > MATERIALS_MAX_POLES || d->n_lorentz > MATERIALS_MAX_POLES) return -1;
    if (m->count >= TABLE_LIMIT) return -1;
    mentry_t e;
    memset(&e, 0, sizeof e);
    e.dispersive = 1;
    e.disp = *d;
    e.eps_r = d->eps_inf;
    e.sigma = d->sigma;
    double dt = m->dt, dx = m->dx, sumv = 0.0;
    int s = 0;
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
    m->version++;
    return 0;
}

size_t materials_table_count(const materials_t *m) { return m ? m->count : 0; }
size_t materials_override_count(const materials_t *m) { return m ? m->ovr.count : 0; }
size_t materials_tensor_count(const materials_t *m) { return m ? m->ten.count : 0; }

int materials_get_dispersive(const materials_t *m, material_id_t id, material_dispersive_t *out) {
    if (!m || !out || (size_t)id >= m->count) return -1;
    if (!m->table[id].dispersive) return 1;
    *out = m->table[id].disp;
    return 0;
}

/* materials.c:426-452 */
static int box_ok(const materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1, size_t j0,
                  size_t j1, size_t k0, size_t k1) {
    if (!rf_materials_matches(m, g)) return 0;
    if (i0 >= i1 || j0 >= j1 || k0 >= k1) return 0;
    if (i1 > g->nx + 1 || j1 > g->ny + 1 || k1 > g->nz + 1) return 0;
    if ((g->periodic_axes & 1u) && i0 >= g->nx) return 0;
    if ((g->periodic_axes & 2u) && j0 >= g->ny) return 0;
    if ((g->periodic_axes & 4u) && k0 >= g->nz) return 0;
    return 1;
}

static void decode(const materials_t *m, size_t idx, size_t x[3]) {
    x[2] = idx % (m->nz + 1);
    size_t r = idx / (m->nz + 1);
    x[1] = r % (m->ny + 1);
    x[0] = r / (m->ny + 1);
}

static void clear_in_box(materials_t *m, rf_map_t *map, const size_t lo[3], const size_t hi[3]) {
    size_t n_del = 0;
    size_t *del = malloc((map->count ? map->count : 1) * sizeof *del);
    if (!del) return;
    for (size_t q = 0; q < map->cap; q++) {
        if (map->keys[q] == (size_t)-1) continue;
        size_t x[3];
        decode(m, map->keys[q] / 3, x);
        if (x[0] >= lo[0] && x[0] < hi[0] && x[1] >= lo[1] && x[1] < hi[1] && x[2] >= lo[2] &&
            x[2] < hi[2])
            del[n_del++] = map->keys[q];
    }
    for (size_t q = 0; q < n_del; q++) rf_map_del(map, del[q]);
    free(del);
}

/* materials.c:475-527 */
int materials_assign_box(materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1, size_t j0,
                         size_t j1, size_t k0, size_t k1, material_id_t id) {
    if (!box_ok(m, g, i0, i1, j0, j1, k0, k1) || (size_t)id >= m->count) return -1;
    if ((g->periodic_axes & 1u) && i1 > g->nx) i1 = g->nx;
    if ((g->periodic_axes & 2u) && j1 > g->ny) j1 = g->ny;
    if ((g->periodic_axes & 4u) && k1 > g->nz) k1 = g->nz;
    for (size_t i = i0; i < i1; i++)
        for (size_t j = j0; j < j1; j++) {
            size_t base = node_index(g, i, j, k0);
            memset(m->ids + base, id, k1 - k0);
        }
    size_t lo[3] = {i0, j0, k0}, hi[3] = {i1, j1, k1};
    clear_in_box(m, &m->ovr, lo, hi);
    clear_in_box(m, &m->ten, lo, hi);
    m->version++;
    return 0;
}

int materials_set_box(materials_t *m, const fdtd_grid_t *g, size_t i0, size_t i1, size_t j0,
                      size_t j1, size_t k0, size_t k1, double eps_r, double sigma) {
    if (!box_ok(m, g, i0, i1, j0, j1, k0, k1)) return -1;
    material_id_t id;
    if (materials_register(m, eps_r, sigma, &id) != 0) return -1;
    return materials_assign_box(m, g, i0, i1, j0, j1, k0, k1, id);
}

static int pool_push(void **pool, size_t *n, size_t *cap, size_t sz, const void *v) {
    if (*n == *cap) {
        size_t nc = *cap ? *cap * 2 : 64;
        void *p = realloc(*pool, nc * sz);
        if (!p) return -1;
        *pool = p;
        *cap = nc;
    }
    memcpy((char *)*pool + *n * sz, v, sz);
    (*n)++;
    return 0;
}

/* materials.c:552-595 */
int materials_set_component(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                            size_t i, size_t j, size_t k, double eps_r, double sigma) {
    int c = comp_no(component);
    material_coeff_t co;
    if (!rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k) || coeff_of(m, eps_r, sigma, &co) != 0)
        return -1;
    size_t key = node_index(g, i, j, k) * 3 + (size_t)c;
    uint32_t slot;
    if (rf_map_get(&m->ovr, key, &slot)) {
        m->ovr_vals[slot] = co;
    } else {
        if (pool_push((void **)&m->ovr_vals, &m->n_ovr_vals, &m->cap_ovr_vals, sizeof co, &co) != 0)
            return -1;
        if (rf_map_put(&m->ovr, key, (uint32_t)(m->n_ovr_vals - 1)) != 0) return -1;
    }
    m->version++;
    return 0;
}

int materials_clear_component(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                              size_t i, size_t j, size_t k) {
    int c = comp_no(component);
    if (!rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    if (rf_map_del(&m->ovr, node_index(g, i, j, k) * 3 + (size_t)c)) m->version++;
    return 0;
}

/* Scalar coefficient of one edge, ignoring tensors (materials.c:165-180);
 * *ade_id is the dispersive table id or -1. */
static material_coeff_t scalar_at(const materials_t *m, size_t idx, int c, int *ade_id) {
    uint32_t slot;
    if (rf_map_get(&m->ovr, idx * 3 + (size_t)c, &slot)) {
        if (ade_id) *ade_id = -1;
        return m->ovr_vals[slot];
    }
    const mentry_t *e = &m->table[m->ids[idx]];
    if (ade_id) *ade_id = e->dispersive ? (int)m->ids[idx] : -1;
    return e->coeff;
}

int materials_get_coefficients(const materials_t *m, const fdtd_grid_t *g, field_component_t component,
                               size_t i, size_t j, size_t k, material_coeff_t *out) {
    int c = comp_no(component);
    if (!out || !rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    *out = scalar_at(m, node_index(g, i, j, k), c, NULL);
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

static int spd3(const double t[6]) {
    for (int q = 0; q < 6; q++) if (!isfinite(t[q])) return 0;
    double xx = t[0], xy = t[1], xz = t[2], yy = t[3], yz = t[4], zz = t[5];
    if (!(xx > 0.0)) return 0;
    if (!(xx * yy - xy * xy > 0.0)) return 0;
    double d3 = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    return d3 > 0.0;
}

static int tensor_put(materials_t *m, size_t key, const mtensor_t *t) {
    uint32_t slot;
    if (rf_map_get(&m->ten, key, &slot)) {
        m->ten_vals[slot] = *t;
    } else {
        if (pool_push((void **)&m->ten_vals, &m->n_ten_vals, &m->cap_ten_vals, sizeof *t, t) != 0)
            return -1;
        if (rf_map_put(&m->ten, key, (uint32_t)(m->n_ten_vals - 1)) != 0) return -1;
    }
    m->version++;
    return 0;
}

/* materials.c:673-738 */
int materials_set_component_tensor(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                   size_t i, size_t j, size_t k, double m11, double m12, double m22) {
    int c = comp_no(component);
    if (c != 0 && c != 1) return -1;
    if (!rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k) || !margin_ok(g, 3, i, j, k) ||
        !spd2(m11, m12, m22))
        return -1;
    double scale = m->dt / (EPS0 * m->dx);
    if (!isfinite(scale) || scale <= 0.0) return -1;
    mtensor_t t;
    memset(&t, 0, sizeof t);
    t.kind = 2;
    t.scale = scale;
    if (c == 0) { t.m[0] = m11; t.m[1] = m12; t.m[3] = m22; }
    else        { t.m[0] = m22; t.m[1] = m12; t.m[3] = m11; }
    t.m[5] = 1.0;
    return tensor_put(m, node_index(g, i, j, k) * 3 + (size_t)c, &t);
}

int materials_set_component_tensor3(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                    size_t i, size_t j, size_t k, const double mm[6]) {
    int c = comp_no(component);
    if (c < 0 || !mm || !rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k) ||
        !margin_ok(g, 7, i, j, k) || !spd3(mm))
        return -1;
    double scale = m->dt / (EPS0 * m->dx);
    if (!isfinite(scale) || scale <= 0.0) return -1;
    mtensor_t t;
    memset(&t, 0, sizeof t);
    t.kind = 3;
    t.scale = scale;
    memcpy(t.m, mm, sizeof t.m);
    return tensor_put(m, node_index(g, i, j, k) * 3 + (size_t)c, &t);
}

int materials_clear_component_tensor(materials_t *m, const fdtd_grid_t *g, field_component_t component,
                                     size_t i, size_t j, size_t k) {
    int c = comp_no(component);
    if (!rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    if (rf_map_del(&m->ten, node_index(g, i, j, k) * 3 + (size_t)c)) m->version++;
    return 0;
}

static const mtensor_t *tensor_at(const materials_t *m, size_t key) {
    uint32_t slot;
    return rf_map_get(&m->ten, key, &slot) ? &m->ten_vals[slot] : NULL;
}

int materials_get_tensor(const materials_t *m, const fdtd_grid_t *g, field_component_t component,
                         size_t i, size_t j, size_t k, material_tensor_t *out) {
    int c = comp_no(component);
    if (!out || !rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    const mtensor_t *t = tensor_at(m, node_index(g, i, j, k) * 3 + (size_t)c);
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
    if (!out || !rf_materials_matches(m, g) || !comp_ok(g, c, i, j, k)) return -1;
    const mtensor_t *t = tensor_at(m, node_index(g, i, j, k) * 3 + (size_t)c);
    if (!t) return 1;
    memcpy(out, t->m, sizeof t->m);
    return 0;
}

/* ------------------------------------------------------ device build */

/* Pair weight of the reference rule (materials.c:836-845) for the pair
 * (a-edge, b-edge), a < b. */
static double corr(const mtensor_t *t, int c, int d) {
    return (t->m[offdiag(c, d)] / sqrt(t->m[DIAG[c]])) / sqrt(t->m[DIAG[d]]);
}

static double pair_weight(const mtensor_t *ta, int a, const mtensor_t *tb, int b) {
    double rho = 0.5 * corr(ta, a, b) + 0.5 * corr(tb, b, a);
    return (0.25 * rho * sqrt(ta->m[DIAG[a]])) * sqrt(tb->m[DIAG[b]]);
}

typedef struct {
    rf_entry64_t *v;
    size_t n, cap;
    size_t *slots;          /* open-addressing table of entry indices */
    size_t nslots;
} combo_set_t;

static uint64_t fnv(const void *p, size_t n) {
    const unsigned char *b = p;
    uint64_t h = 1469598103934665603ULL;
    for (size_t q = 0; q < n; q++) { h ^= b[q]; h *= 1099511628211ULL; }
    return h;
}

static int combo_grow(combo_set_t *s) {
    size_t ns = s->nslots ? s->nslots * 2 : 1024;
    size_t *sl = malloc(ns * sizeof *sl);
    if (!sl) return -1;
    for (size_t q = 0; q < ns; q++) sl[q] = (size_t)-1;
    for (size_t e = 0; e < s->n; e++) {
        size_t h = fnv(&s->v[e], sizeof s->v[e]) & (ns - 1);
        while (sl[h] != (size_t)-1) h = (h + 1) & (ns - 1);
        sl[h] = e;
    }
    free(s->slots);
    s->slots = sl;
    s->nslots = ns;
    return 0;
}

/* Index of entry e in the set, inserting it when new; -1 on failure. */
static long combo_index(combo_set_t *s, const rf_entry64_t *e, int dedup) {
    if (dedup) {
        if ((s->n + 1) * 2 > s->nslots && combo_grow(s) != 0) return -1;
        size_t h = fnv(e, sizeof *e) & (s->nslots - 1);
        while (s->slots[h] != (size_t)-1) {
            if (!memcmp(&s->v[s->slots[h]], e, sizeof *e)) return (long)s->slots[h];
            h = (h + 1) & (s->nslots - 1);
        }
        s->slots[h] = s->n;
    }
    if (s->n == s->cap) {
        size_t nc = s->cap ? s->cap * 2 : 256;
        rf_entry64_t *v = realloc(s->v, nc * sizeof *v);
        if (!v) return -1;
        s->v = v;
        s->cap = nc;
    }
    s->v[s->n] = *e;
    return (long)s->n++;
}

static void entry_of_node(const materials_t *m, size_t idx, rf_entry64_t *e) {
    memset(e, 0, sizeof *e);
    for (int c = 0; c < 3; c++) {
        int ade_id;
        material_coeff_t co = scalar_at(m, idx, c, &ade_id);
        e->Ca[c] = co.Ca;
        e->Cb[c] = co.Cb;
        e->ade[c] = -1;
        if (ade_id >= 0) { e->flags[c] |= RF_EDGE_ADE; e->ade[c] = ade_id; }
        if (tensor_at(m, idx * 3 + (size_t)c)) {
            e->flags[c] = RF_EDGE_TENSOR;
            e->ade[c] = -1;
        }
    }
}

static int own(rf_mat_dev_t *d, void *p) {
    if (d->n_owned >= 16) return -1;
    d->owned[d->n_owned++] = p;
    return 0;
}

static int dev_upload(rf_mat_dev_t *d, const void *src, size_t bytes, void **out) {
    void *p = NULL;
    if (rf_malloc(&p, bytes) != 0 || own(d, p) != 0) return -1;
    d->bytes += bytes;
    if (src && (rf_h2d(p, src, bytes, NULL) != 0 || rf_sync(NULL) != 0)) return -1;
    *out = p;
    return 0;
}

static size_t wrap_next(size_t q, size_t n) { return q + 1 == n ? 0 : q + 1; }
static size_t wrap_prev(size_t q, size_t n) { return q == 0 ? n - 1 : q - 1; }

/* Tensor edge records in the order the kernel consumes neighbours. */
static int build_tensor_edges(const materials_t *m, const fdtd_grid_t *g, rf_tedge64_t **out, long long *n_out) {
    *out = NULL;
    *n_out = 0;
    size_t n = m->ten.count;
    if (!n) return 0;
    rf_tedge64_t *edges = calloc(n, sizeof *edges);
    size_t *keys = malloc(n * sizeof *keys);
    if (!edges || !keys) { free(edges); free(keys); return -1; }
    size_t w = 0;
    for (size_t q = 0; q < m->ten.cap; q++)
        if (m->ten.keys[q] != (size_t)-1) keys[w++] = m->ten.keys[q];
    /* deterministic order: keys ascending (shell sort) */
    for (size_t gap = n / 2; gap > 0; gap /= 2)
        for (size_t a = gap; a < n; a++) {
            size_t t = keys[a], b = a;
            while (b >= gap && keys[b - gap] > t) { keys[b] = keys[b - gap]; b -= gap; }
            keys[b] = t;
        }
    size_t ext[3] = {g->nx, g->ny, g->nz};
    for (size_t e = 0; e < n; e++) {
        size_t key = keys[e], idx = key / 3;
        int c = (int)(key % 3);
        const mtensor_t *te = tensor_at(m, key);
        size_t x[3];
        decode(m, idx, x);
        rf_tedge64_t *r = &edges[e];
        r->i = (int)x[0]; r->j = (int)x[1]; r->k = (int)x[2];
        r->comp = c;
        r->kind = te->kind;
        r->diag = te->m[DIAG[c]];
        r->scale = te->scale;
        int slot = 0;
        for (int d = 0; d < 3; d++) {
            if (d == c) continue;
            int lo = c < d ? c : d;
            for (int q = 0; q < 4; q++) {
                if (te->kind == 2 && d == 2) continue;
                int s_lo = q & 1, s_hi = (q >> 1) & 1;
                int s_c = (c == lo) ? s_lo : s_hi, s_d = (d == lo) ? s_lo : s_hi;
                size_t y[3] = {x[0], x[1], x[2]};
                if (s_c) y[c] = wrap_next(x[c], ext[c]);
                if (s_d) y[d] = wrap_prev(x[d], ext[d]);
                const mtensor_t *o = tensor_at(m, node_index(g, y[0], y[1], y[2]) * 3 + (size_t)d);
                if (!o) continue;
                if ((c == 2 || d == 2) && o->kind != 3) continue;
                r->mask |= 1u << (slot * 4 + q);
                r->w[slot * 4 + q] = c < d ? pair_weight(te, c, o, d) : pair_weight(o, d, te, c);
            }
            slot++;
        }
    }
    free(keys);
    *out = edges;
    *n_out = (long long)n;
    return 0;
}

/* Node box of the edges of one ADE set and its state region. */
typedef struct {
    int used;
    size_t lo[3], hi[3];
    size_t base, st;
} set_box_t;

static void set_layout(const set_box_t *b, long long *base, long long *st, int b0[3], int bn[3]) {
    const int used = b && b->used;
    *base = used ? (long long)b->base : 0;
    *st = used ? (long long)b->st : 0;
    for (int a = 0; a < 3; a++) {
        b0[a] = used ? (int)b->lo[a] : 0;
        bn[a] = used ? (int)(b->hi[a] - b->lo[a]) : 0;
    }
}

int rf_materials_build(const materials_t *m, const fdtd_grid_t *g, int prec, rf_mat_dev_t *d) {
    memset(d, 0, sizeof *d);
    int rc = -1;
    combo_set_t set = {0};
    set_box_t *sbox = NULL;
    uint32_t *tags = NULL;
    rf_tedge64_t *edges = NULL;
    size_t total = m->total;

    /* plain ids map to entries 0..count-1 (no dedup, so the index is the id) */
    for (size_t id = 0; id < m->count; id++) {
        rf_entry64_t e;
        memset(&e, 0, sizeof e);
        for (int c = 0; c < 3; c++) {
            e.Ca[c] = m->table[id].coeff.Ca;
            e.Cb[c] = m->table[id].coeff.Cb;
            e.ade[c] = -1;
            if (m->table[id].dispersive) { e.flags[c] = RF_EDGE_ADE; e.ade[c] = (int)id; }
        }
        if (combo_index(&set, &e, 0) < 0) goto out;
    }
    if (combo_grow(&set) != 0) goto out;
    tags = malloc(total * sizeof *tags);
    if (!tags) goto out;
    for (size_t q = 0; q < total; q++) tags[q] = m->ids[q];
    rf_map_t *maps[2] = {(rf_map_t *)&m->ovr, (rf_map_t *)&m->ten};
    for (int w = 0; w < 2; w++)
        for (size_t q = 0; q < maps[w]->cap; q++) {
            size_t key = maps[w]->keys[q];
            if (key == (size_t)-1) continue;
            size_t idx = key / 3;
            rf_entry64_t e;
            entry_of_node(m, idx, &e);
            long t = combo_index(&set, &e, 1);
            if (t < 0) goto out;
            tags[idx] = (uint32_t)t;
        }
    d->n_entries = set.n;
    d->d.n_entries = (int)set.n;

    /* ADE state: the union box of the nodes with a dispersive edge and,
     * per ADE set (= material id), the box of the nodes with an edge of
     * that set, in which the set's state is laid out */
    int any_ade = 0;
    size_t lo[3] = {(size_t)-1, (size_t)-1, (size_t)-1}, hi[3] = {0, 0, 0};
    for (size_t e = 0; e < set.n && !any_ade; e++)
        for (int c = 0; c < 3; c++) if (set.v[e].flags[c] & RF_EDGE_ADE) any_ade = 1;
    if (m->count) {
        sbox = calloc(m->count, sizeof *sbox);
        if (!sbox) goto out;
        for (size_t id = 0; id < m->count; id++)
            for (int a = 0; a < 3; a++) { sbox[id].lo[a] = (size_t)-1; sbox[id].hi[a] = 0; }
    }
    if (any_ade) {
        for (size_t q = 0; q < total; q++) {
            const rf_entry64_t *en = &set.v[tags[q]];
            if (!((en->flags[0] | en->flags[1] | en->flags[2]) & RF_EDGE_ADE)) continue;
            size_t x[3];
            decode(m, q, x);
            for (int a = 0; a < 3; a++) {
                if (x[a] < lo[a]) lo[a] = x[a];
                if (x[a] + 1 > hi[a]) hi[a] = x[a] + 1;
            }
            for (int c = 0; c < 3; c++) {
                if (!(en->flags[c] & RF_EDGE_ADE) || en->ade[c] < 0 || (size_t)en->ade[c] >= m->count) continue;
                set_box_t *b = &sbox[en->ade[c]];
                b->used = 1;
                for (int a = 0; a < 3; a++) {
                    if (x[a] < b->lo[a]) b->lo[a] = x[a];
                    if (x[a] + 1 > b->hi[a]) b->hi[a] = x[a] + 1;
                }
            }
        }
        int slots = 0;
        for (size_t id = 0; id < m->count; id++)
            if (m->table[id].dispersive && m->table[id].n_poles > slots) slots = m->table[id].n_poles;
        for (int a = 0; a < 3; a++) { d->d.b0[a] = (int)lo[a]; d->d.bn[a] = (int)(hi[a] - lo[a]); }
        d->d.n_slots = slots;
        size_t es = prec ? sizeof(float) : sizeof(double);
        /* each set's (pole, component) arrays: 64-element aligned plus an
         * odd number of 64-element blocks, so the state streams of a node
         * do not all start at one offset modulo a large power of two */
        size_t base = 0;
        for (size_t id = 0; id < m->count; id++) {
            set_box_t *b = &sbox[id];
            if (!b->used || m->table[id].n_poles < 1) continue;
            const size_t cells = (b->hi[0] - b->lo[0]) * (b->hi[1] - b->lo[1]) * (b->hi[2] - b->lo[2]);
            b->st = (cells + 63) / 64 * 64 + 64 * 37;
            b->base = base;
            base += (size_t)m->table[id].n_poles * 3 * b->st;
        }
        size_t sb = (base ? base : 64) * es;
        if (dev_upload(d, NULL, sb, &d->d.J) || dev_upload(d, NULL, sb, &d->d.P)) goto out;
        d->ade_bytes = sb;
    }

    /* tags */
    int tb = set.n <= 256 ? 1 : set.n <= 65536 ? 2 : 4;
    d->d.tag_bytes = tb;
    {
        void *packed = malloc(total * (size_t)tb);
        if (!packed) goto out;
        for (size_t q = 0; q < total; q++) {
            if (tb == 1) ((uint8_t *)packed)[q] = (uint8_t)tags[q];
            else if (tb == 2) ((uint16_t *)packed)[q] = (uint16_t)tags[q];
            else ((uint32_t *)packed)[q] = tags[q];
        }
        void *dp = NULL;
        int r = dev_upload(d, packed, total * (size_t)tb, &dp);
        free(packed);
        if (r) goto out;
        d->d.tag = dp;
    }

    /* entries, ADE coefficients and energy epsilons */
    {
        void *dp = NULL;
        double *eps = malloc(set.n * 3 * sizeof *eps);
        if (!eps) goto out;
        for (size_t e = 0; e < set.n; e++)
            for (int c = 0; c < 3; c++) eps[3 * e + c] = m->dt / (m->dx * set.v[e].Cb[c]);
        int r = dev_upload(d, eps, set.n * 3 * sizeof *eps, &dp);
        free(eps);
        if (r) goto out;
        d->eps_tab = dp;
        if (prec == 0) {
            if (dev_upload(d, set.v, set.n * sizeof(rf_entry64_t), &dp)) goto out;
        } else {
            rf_entry32_t *v32 = calloc(set.n, sizeof *v32);
            if (!v32) goto out;
            for (size_t e = 0; e < set.n; e++)
                for (int c = 0; c < 3; c++) {
                    v32[e].Ca[c] = (float)set.v[e].Ca[c];
                    v32[e].Cb[c] = (float)set.v[e].Cb[c];
                    v32[e].flags[c] = set.v[e].flags[c];
                    v32[e].ade[c] = set.v[e].ade[c];
                }
            r = dev_upload(d, v32, set.n * sizeof *v32, &dp);
            free(v32);
            if (r) goto out;
        }
        d->d.entries = dp;
        size_t na = m->count ? m->count : 1;
        d->d.n_ade = (int)na;
        if (prec == 0) {
            rf_ade64_t *a = calloc(na, sizeof *a);
            if (!a) goto out;
            for (size_t id = 0; id < m->count; id++) {
                const mentry_t *me = &m->table[id];
                a[id].n_poles = me->n_poles;
                for (int s = 0; s < me->n_poles; s++) {
                    a[id].has_p[s] = me->has_p[s];
                    a[id].kj[s] = me->kj[s]; a[id].kp[s] = me->kp[s]; a[id].v[s] = me->v[s];
                }
                a[id].Cu = me->Cu;
                a[id].hdt = me->hdt;
                set_layout(sbox ? &sbox[id] : NULL, &a[id].base, &a[id].st, a[id].b0, a[id].bn);
            }
            r = dev_upload(d, a, na * sizeof *a, &dp);
            free(a);
        } else {
            rf_ade32_t *a = calloc(na, sizeof *a);
            if (!a) goto out;
            for (size_t id = 0; id < m->count; id++) {
                const mentry_t *me = &m->table[id];
                a[id].n_poles = me->n_poles;
                for (int s = 0; s < me->n_poles; s++) {
                    a[id].has_p[s] = me->has_p[s];
                    a[id].kj[s] = (float)me->kj[s]; a[id].kp[s] = (float)me->kp[s];
                    a[id].v[s] = (float)me->v[s];
                }
                a[id].Cu = (float)me->Cu;
                a[id].hdt = (float)me->hdt;
                set_layout(sbox ? &sbox[id] : NULL, &a[id].base, &a[id].st, a[id].b0, a[id].bn);
            }
            r = dev_upload(d, a, na * sizeof *a, &dp);
            free(a);
        }
        if (r) goto out;
        d->d.ade = dp;
    }

    /* tensor edges */
    if (build_tensor_edges(m, g, &edges, &d->n_tensor) != 0) goto out;
    if (d->n_tensor) {
        void *dp = NULL;
        if (prec == 0) {
            if (dev_upload(d, edges, (size_t)d->n_tensor * sizeof *edges, &dp)) goto out;
        } else {
            rf_tedge32_t *e32 = calloc((size_t)d->n_tensor, sizeof *e32);
            if (!e32) goto out;
            for (long long e = 0; e < d->n_tensor; e++) {
                e32[e].i = edges[e].i; e32[e].j = edges[e].j; e32[e].k = edges[e].k;
                e32[e].comp = edges[e].comp; e32[e].kind = edges[e].kind; e32[e].mask = edges[e].mask;
                e32[e].diag = (float)edges[e].diag; e32[e].scale = (float)edges[e].scale;
                for (int q = 0; q < 8; q++) e32[e].w[q] = (float)edges[e].w[q];
            }
            int r = dev_upload(d, e32, (size_t)d->n_tensor * sizeof *e32, &dp);
            free(e32);
            if (r) goto out;
        }
        d->tensor_edges = dp;
    }
    d->d.vac_ca = m->dt / (EPS0 * m->dx);
    rc = 0;
out:
    free(edges);
    free(sbox);
    free(tags);
    free(set.v);
    free(set.slots);
    if (rc) rf_materials_release(d);
    return rc;
}

void rf_materials_release(rf_mat_dev_t *d) {
    for (int q = 0; q < d->n_owned; q++) rf_free(d->owned[q]);
    d->n_owned = 0;
}

rf_mat_dev_t *rf_materials_dev(materials_t *m, const fdtd_grid_t *g) {
    if (m->dev && m->dev_version == m->version && m->dev_per == g->periodic_axes) return m->dev;
    if (m->dev) { rf_materials_release(m->dev); free(m->dev); m->dev = NULL; }
    rf_mat_dev_t *d = calloc(1, sizeof *d);
    if (!d) return NULL;
    if (rf_materials_build(m, g, 0, d) != 0) { free(d); return NULL; }
    m->dev = d;
    m->dev_version = m->version;
    m->dev_per = g->periodic_axes;
    return d;
}

static void update_e_dev(fdtd_grid_t *g, cpml_t *pml, materials_t *m, const char *what) {
    rf_cpml_dev_t *c = NULL;
    if (pml) {
        c = rf_cpml_dev(pml);
        if (!c) RF_FAIL(what);
    }
    rf_mat_dev_t *d = rf_materials_dev(m, g);
    if (!d || rf_grid_upload(g, 077u) != 0) RF_FAIL(what);
    rf_geom_t geo = rf_geom_of(g);
    if (rf_launch_update_e(0, &geo, g->dev->f, g->dx, c ? &c->d : NULL, &d->d, g->dev->st) != 0)
        RF_FAIL(what);
    if (d->n_tensor &&
        rf_launch_tensor(0, &geo, g->dev->f, d->tensor_edges, d->n_tensor, g->dev->st) != 0)
        RF_FAIL(what);
    if (rf_grid_download(g, 007u) != 0) RF_FAIL(what);
}

/* materials.c:1062-1070 */
void fdtd_update_e_materials(fdtd_grid_t *g, const materials_t *m) {
    if (!rf_materials_matches(m, g)) return;
    update_e_dev(g, NULL, (materials_t *)m, "fdtd_update_e_materials");
}

/* materials.c:1072-1082 */
void fdtd_update_e_cpml_materials(fdtd_grid_t *g, cpml_t *pml, const materials_t *m) {
    if (!rf_materials_matches(m, g) || !pml || pml->dx != g->dx || !rf_cpml_matches(pml, g)) return;
    update_e_dev(g, pml, (materials_t *)m, "fdtd_update_e_cpml_materials");
}
