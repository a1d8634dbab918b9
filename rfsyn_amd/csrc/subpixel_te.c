/*
 * 2D subpixel smoothing helpers (include/subpixel_te.h), host code.
 *
 * The cell fraction clips the cell, in coordinates where it is the unit
 * square centred at the origin, by the four half-planes of the rectangle
 * (Sutherland-Hodgman), then takes the shoelace area of what is left.
 * Working in cell coordinates keeps every vertex of order one however
 * large or distant the rectangle is.  The mixing rules and the edge
 * fraction evaluate the reference's formulas (subpixel_te.c:27-132,
 * 410-455) term for term, so they return the same bits.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>

#include "subpixel_te.h"

typedef struct {
    double x, y;
} vtx_t;

/* a square cut by 4 half-planes has at most 8 vertices */
#define MAX_VTX 8

static int finite_rect(const subpixel_te_rectangle_t *r) {
    return isfinite(r->center_x) && isfinite(r->center_y) && isfinite(r->size_u) && isfinite(r->size_v) &&
           isfinite(r->angle_rad);
}

/* Keeps the part of polygon in[0..n) with a x + b y <= d; returns the
 * vertex count of out, or -1 on a non-finite intersection. */
static int cut(const vtx_t *in, int n, double a, double b, double d, vtx_t *out) {
    int m = 0;
    if (n == 0) return 0;
    /* normalise the plane so the distances stay of order one */
    const double scale = fmax(fabs(a), fmax(fabs(b), fabs(d)));
    if (scale == 0.0) {
        for (int q = 0; q < n; q++) out[q] = in[q];
        return n;
    }
    a /= scale;
    b /= scale;
    d /= scale;
    vtx_t p = in[n - 1];
    double fp = fma(a, p.x, b * p.y) - d;
    for (int q = 0; q < n; q++) {
        const vtx_t c = in[q];
        const double fc = fma(a, c.x, b * c.y) - d;
        if ((fc <= 0.0) != (fp <= 0.0)) {
            const double den = fp - fc;
            if (!isfinite(den) || den == 0.0) return -1;
            double t = fp / den;
            if (!isfinite(t)) return -1;
            t = fmin(1.0, fmax(0.0, t));
            if (m >= MAX_VTX) return -1;
            out[m].x = fma(t, c.x - p.x, p.x);
            out[m].y = fma(t, c.y - p.y, p.y);
            if (!isfinite(out[m].x) || !isfinite(out[m].y)) return -1;
            m++;
        }
        if (fc <= 0.0) {
            if (m >= MAX_VTX) return -1;
            out[m++] = c;
        }
        p = c;
        fp = fc;
    }
    return m;
}

int subpixel_te_rectangle_fraction(const subpixel_te_rectangle_t *r, double sx, double sy, double dx, double dy,
                                   double *fraction_out) {
    if (!r || !fraction_out || !finite_rect(r) || !(r->size_u > 0.0) || !(r->size_v > 0.0) || !isfinite(sx) ||
        !isfinite(sy) || !isfinite(dx) || !(dx > 0.0) || !isfinite(dy) || !(dy > 0.0))
        return SUBPIXEL_TE_INVALID_ARGUMENT;
    const double hu = 0.5 * r->size_u, hv = 0.5 * r->size_v;
    const double rx = sx - r->center_x, ry = sy - r->center_y;   /* cell centre in the rectangle frame */
    if (!(hu > 0.0) || !(hv > 0.0) || !isfinite(rx) || !isfinite(ry)) return SUBPIXEL_TE_NUMERIC_ERROR;
    const double c = cos(r->angle_rad), s = sin(r->angle_rad);
    /* rectangle coordinates of the cell point (X dx, Y dy) off its centre:
     * u = u0 + c dx X + s dy Y,  v = v0 - s dx X + c dy Y */
    const double u0 = fma(c, rx, s * ry), v0 = fma(-s, rx, c * ry);
    const double plane[4][3] = {
        {c * dx, s * dy, hu - u0},      /*  u <= hu */
        {-c * dx, -s * dy, hu + u0},    /* -u <= hu */
        {-s * dx, c * dy, hv - v0},     /*  v <= hv */
        {s * dx, -c * dy, hv + v0},     /* -v <= hv */
    };
    vtx_t buf[2][MAX_VTX] = {{{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}}};
    int n = 4, cur = 0;
    for (int p = 0; p < 4; p++) {
        n = cut(buf[cur], n, plane[p][0], plane[p][1], plane[p][2], buf[1 - cur]);
        if (n < 0) return SUBPIXEL_TE_NUMERIC_ERROR;
        cur = 1 - cur;
    }
    /* shoelace area as a fan from the first vertex */
    double twice = 0.0;
    const vtx_t *v = buf[cur];
    for (int q = 1; q + 1 < n; q++) {
        const double ax = v[q].x - v[0].x, ay = v[q].y - v[0].y;
        const double bx = v[q + 1].x - v[0].x, by = v[q + 1].y - v[0].y;
        twice += fma(ax, by, -ay * bx);
    }
    const double area = 0.5 * fabs(twice);
    if (!isfinite(area) || area > 1.0 + 128.0 * DBL_EPSILON) return SUBPIXEL_TE_NUMERIC_ERROR;
    *fraction_out = fmin(1.0, fmax(0.0, area));
    return SUBPIXEL_TE_OK;
}

/* [lo, hi] of t where |w0 + dw t| <= half; returns 0 if that set is empty
 * for every t (a direction too small to divide by). */
static int slab_interval(double w0, double dw, double half, double length, double *lo, double *hi) {
    if (fabs(dw) < 1e-300) {
        if (fabs(w0) > half * (1.0 + 1e-12)) return 0;
        *lo = 0.0;
        *hi = length;
        return 1;
    }
    const double t0 = (-half - w0) / dw, t1 = (half - w0) / dw;
    *lo = fmin(t0, t1);
    *hi = fmax(t0, t1);
    return 1;
}

int subpixel_te_rectangle_edge_fraction(const subpixel_te_rectangle_t *r, double sx, double sy, int edge_axis,
                                        double length, double *fraction_out) {
    if (!r || !fraction_out || !isfinite(sx) || !isfinite(sy) || (edge_axis != 1 && edge_axis != 2) ||
        !isfinite(length) || !(length > 0.0) || !finite_rect(r) || !(r->size_u >= 0.0) || !(r->size_v >= 0.0))
        return SUBPIXEL_TE_INVALID_ARGUMENT;
    const double rx = sx - r->center_x, ry = sy - r->center_y;
    if (!isfinite(rx) || !isfinite(ry)) return SUBPIXEL_TE_NUMERIC_ERROR;
    const double c = cos(r->angle_rad), s = sin(r->angle_rad);
    const double u0 = rx * c + ry * s, v0 = ry * c - rx * s;
    const double du = edge_axis == 1 ? c : -s, dv = edge_axis == 1 ? s : c;
    if (r->size_u == 0.0) return SUBPIXEL_TE_NUMERIC_ERROR;
    double alo, ahi, blo, bhi;
    if (!slab_interval(u0, du, 0.5 * r->size_u, length, &alo, &ahi) ||
        !slab_interval(v0, dv, 0.5 * r->size_v, length, &blo, &bhi)) {
        *fraction_out = 0.0;
        return SUBPIXEL_TE_OK;
    }
    const double lo = fmax(fmax(alo, blo), 0.0), hi = fmin(fmin(ahi, bhi), length);
    *fraction_out = hi > lo ? (hi - lo) / length : 0.0;
    return SUBPIXEL_TE_OK;
}

static int valid_mix(double f, double eb, double ei) {
    return isfinite(f) && f >= 0.0 && f <= 1.0 && isfinite(eb) && eb >= 1.0 && isfinite(ei) && ei >= 1.0;
}

int subpixel_te_mix_epsilon_harmonic(double f, double eb, double ei, double *epsilon_r_out) {
    if (!epsilon_r_out || !valid_mix(f, eb, ei)) return SUBPIXEL_TE_INVALID_ARGUMENT;
    const double inv = (1.0 - f) / eb + f / ei;
    if (!isfinite(inv) || !(inv > 0.0)) return SUBPIXEL_TE_NUMERIC_ERROR;
    *epsilon_r_out = 1.0 / inv;
    return SUBPIXEL_TE_OK;
}

int subpixel_te_mix_epsilon(double f, double eb, double ei, double *epsilon_r_out) {
    if (!epsilon_r_out || !valid_mix(f, eb, ei)) return SUBPIXEL_TE_INVALID_ARGUMENT;
    double e;
    if (f == 0.0)
        e = eb;
    else if (f == 1.0)
        e = ei;
    else if (ei >= eb)   /* step up from the smaller value: no overflow, monotone */
        e = fma(f, ei - eb, eb);
    else
        e = fma(1.0 - f, eb - ei, ei);
    if (!isfinite(e) || e < 1.0) return SUBPIXEL_TE_NUMERIC_ERROR;
    *epsilon_r_out = e;
    return SUBPIXEL_TE_OK;
}

int subpixel_te_rectangle_epsilon(const subpixel_te_rectangle_t *r, double sx, double sy, double dx, double dy,
                                  double eb, double ei, double *fraction_out, double *epsilon_r_out) {
    if (!fraction_out || !epsilon_r_out || fraction_out == epsilon_r_out) return SUBPIXEL_TE_INVALID_ARGUMENT;
    double f, e;
    int st = subpixel_te_rectangle_fraction(r, sx, sy, dx, dy, &f);
    if (st != SUBPIXEL_TE_OK) return st;
    st = subpixel_te_mix_epsilon(f, eb, ei, &e);
    if (st != SUBPIXEL_TE_OK) return st;
    *fraction_out = f;
    *epsilon_r_out = e;
    return SUBPIXEL_TE_OK;
}

int subpixel_te_mix_epsilon_tensor(double f, double eb, double ei, double normal_rad, double *minv_xx,
                                   double *minv_xy, double *minv_yy) {
    if (!minv_xx || !minv_xy || !minv_yy || !valid_mix(f, eb, ei) || !isfinite(normal_rad))
        return SUBPIXEL_TE_INVALID_ARGUMENT;
    const double lam_n = (1.0 - f) / eb + f / ei;    /* series (normal) */
    const double eps_t = (1.0 - f) * eb + f * ei;    /* parallel (tangential) */
    if (!isfinite(lam_n) || !(lam_n > 0.0) || !isfinite(eps_t) || !(eps_t > 0.0)) return SUBPIXEL_TE_NUMERIC_ERROR;
    const double lam_t = 1.0 / eps_t;
    const double c = cos(normal_rad), s = sin(normal_rad);
    *minv_xx = lam_n * c * c + lam_t * s * s;
    *minv_yy = lam_n * s * s + lam_t * c * c;
    *minv_xy = (lam_n - lam_t) * c * s;
    return SUBPIXEL_TE_OK;
}
