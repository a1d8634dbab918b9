// Device helpers shared by the Yee kernels (two-kernel and fused paths).
#ifndef RFSYN_YEE_DEVICE_H
#define RFSYN_YEE_DEVICE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rf_device.h"

namespace rfk {

struct Geo {
    int nx, ny, nz;
    bool px, py, pz;
    long long sx, sy;
};

// Block b of a grid of nb blocks runs on XCD b % RF_XCDS (blocks are dealt
// round-robin to the XCDs); renumbered so that each XCD takes one
// contiguous range of the grid's blocks in order (a bijection on [0, nb)).
// Node-per-thread sweeps whose threads read the rows j-1 and planes i-1
// that earlier blocks loaded then find them in the L2 of their own XCD.
constexpr unsigned RF_XCDS = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned x = b % RF_XCDS, r = b / RF_XCDS, per = nb / RF_XCDS, rem = nb % RF_XCDS;
    return x * per + (x < rem ? x : rem) + r;
}

__device__ __forceinline__ Geo geo_of(const rf_geom_t &g) {
    Geo o;
    o.nx = g.n[0]; o.ny = g.n[1]; o.nz = g.n[2];
    o.px = g.per & 1; o.py = g.per & 2; o.pz = g.per & 4;
    o.sx = g.sx; o.sy = g.sy;
    return o;
}

// Compressed slab coordinate of q along axis a, or -1 outside the slab.
__device__ __forceinline__ int slab_coord(const rf_cpml_t &c, int a, int q) {
    if (q < c.lo[a]) return q;
    if (q >= c.hi[a]) return c.lo[a] + q - c.hi[a];
    return -1;
}

__device__ __forceinline__ long long slab_index(const rf_cpml_t &c, const Geo &g, int a, int cq,
                                                int i, int j, int k) {
    long long ny1 = (a == 1) ? c.len[1] : g.ny + 1;
    long long nz1 = (a == 2) ? c.len[2] : g.nz + 1;
    long long ii = (a == 0) ? cq : i, jj = (a == 1) ? cq : j, kk = (a == 2) ? cq : k;
    return (ii * ny1 + jj) * nz1 + kk;
}

// cpml_term of cpml.h:46-51 with psi held only in slabs; outside the slab
// psi stays exactly zero, so the term is diff*kinv + dx*0 as in the
// reference's full-size arrays.
template <typename T>
__device__ __forceinline__ T stretch(T diff, const rf_cpml_t &c, const Geo &g, bool h, int axis,
                                     int q, T *psi_arr, int i, int j, int k, T dx) {
    // device memory, accessed as global memory (a pointer read from a
    // device table would otherwise be accessed with FLAT instructions)
    typedef const __attribute__((address_space(1))) T gc_t;
    typedef __attribute__((address_space(1))) T g_t;
    gc_t *kinv = (gc_t *)(h ? c.kinv_h[axis] : c.kinv_m[axis]);
    gc_t *bb = (gc_t *)(h ? c.b_h[axis] : c.b_m[axis]);
    gc_t *aa = (gc_t *)(h ? c.a_h[axis] : c.a_m[axis]);
    int cq = slab_coord(c, axis, q);
    T psi = T(0);
    if (cq >= 0) {
        long long s = slab_index(c, g, axis, cq, i, j, k);
        g_t *pa = (g_t *)psi_arr;
        psi = bb[q] * pa[s] + aa[q] * diff;
        pa[s] = psi;
    }
    return diff * kinv[q] + dx * psi;
}

// Global-memory views: loads and stores through a pointer the compiler
// cannot prove global (one read from a device table) compile to FLAT
// instructions, which also count against the LDS counter; device arrays
// reached through such pointers are accessed through these.
template <typename T> using gmem_c = const __attribute__((address_space(1))) T;
template <typename T> using gmem = __attribute__((address_space(1))) T;

// A CPML term (the arithmetic of stretch) in three steps, so that a node's
// loads can all be issued before any of its stores: term_load reads the
// profile values and psi, term_apply gives the stretched difference and
// leaves the new psi in the term, term_store writes it.  Without the split
// a psi store would keep the next component's loads behind it (the psi
// arrays may alias the fields as far as the compiler knows).
template <typename T>
struct Term {
    T kinv, b, a, psi;
    long long rd, wr;   // psi read and write index, -1 outside the slab
};

template <typename T>
__device__ __forceinline__ Term<T> term_load(const rf_cpml_t &c, const Geo &g, bool h, int axis, int q,
                                             const void *psi_arr, int i, int j, int k) {
    Term<T> o;
    o.kinv = ((gmem_c<T> *)(h ? c.kinv_h[axis] : c.kinv_m[axis]))[q];
    o.b = o.a = o.psi = T(0);
    o.rd = o.wr = -1;
    const int cq = slab_coord(c, axis, q);
    if (cq >= 0) {
        o.rd = o.wr = slab_index(c, g, axis, cq, i, j, k);
        o.b = ((gmem_c<T> *)(h ? c.b_h[axis] : c.b_m[axis]))[q];
        o.a = ((gmem_c<T> *)(h ? c.a_h[axis] : c.a_m[axis]))[q];
        o.psi = ((gmem_c<T> *)psi_arr)[o.rd];
    }
    return o;
}

template <typename T>
__device__ __forceinline__ T term_apply(Term<T> &t, T diff, T dx) {
    T psi = T(0);
    if (t.rd >= 0) {
        psi = t.b * t.psi + t.a * diff;
        t.psi = psi;
    }
    return diff * t.kinv + dx * psi;
}

template <typename T>
__device__ __forceinline__ void term_store(const Term<T> &t, void *psi_arr) {
    if (t.wr >= 0) ((gmem<T> *)psi_arr)[t.wr] = t.psi;
}

template <typename T> struct EntryOf;
template <> struct EntryOf<double> { using type = rf_entry64_t; using ade = rf_ade64_t; };
template <> struct EntryOf<float> { using type = rf_entry32_t; using ade = rf_ade32_t; };

struct NoTag { };

template <typename TagT>
__device__ __forceinline__ unsigned load_tag(const void *p, long long idx) {
    return static_cast<unsigned>(((const __attribute__((address_space(1))) TagT *)p)[idx]);
}

// Element t of the plane captures cs (capture q takes the elements
// [start, start + na nb), cell (ia, ib) with ib fastest): its field value
// into slot step % ring_len of the capture's ring.
template <typename T>
__device__ __forceinline__ void plane_capture_elem(const rf_geom_t &gg, const rf_caps_t &cs, long long t,
                                                   int ring_len, long long step) {
    int q = 0;
    while (q + 1 < cs.n && t >= cs.c[q + 1].start) q++;
    if (q >= cs.n) return;
    const rf_cap_t &c = cs.c[q];
    const long long local = t - c.start;
    if (local < 0 || local >= (long long)c.na * c.nb) return;
    const Geo g = geo_of(gg);
    int index = c.index;
    if ((gg.per >> c.axis) & 1) index %= (c.axis == 0 ? g.nx : c.axis == 1 ? g.ny : g.nz);
    const int ia = static_cast<int>(local / c.nb), ib = static_cast<int>(local % c.nb);
    // the other two axes in increasing order: (y, z), (x, z) or (x, y)
    const int i = c.axis == 0 ? index : ia;
    const int j = c.axis == 0 ? ia : (c.axis == 1 ? index : ib);
    const int k = c.axis == 2 ? index : ib;
    const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;
    static_cast<T *>(c.buf)[(step % ring_len) * (long long)c.na * c.nb + local] = static_cast<const T *>(c.f)[idx];
}

// Element of the state of pole 0, component c, node (i,j,k) in the box of
// ADE set `a`; pole s is s * 3 * a.st further on.
template <typename Ade>
__device__ __forceinline__ long long ade_off(const Ade &a, int c, int i, int j, int k) {
    return a.base + (long long)c * a.st +
           ((long long)(i - a.b0[0]) * a.bn[1] + (j - a.b0[1])) * a.bn[2] + (k - a.b0[2]);
}

// Trapezoidal ADE edge (materials.h): returns E^{n+1} and advances J, P.
template <typename T>
__device__ __forceinline__ T ade_edge(const typename EntryOf<T>::ade &a, T ca, T cb, T e_old,
                                      T curl_a, T curl_b, T *J, T *P, long long slot_stride,
                                      long long off) {
    T u[RF_MAX_SLOTS];
    T sum_u = T(0);
    for (int s = 0; s < a.n_poles; s++) {
        const long long o = s * slot_stride + off;
        u[s] = a.has_p[s] ? a.kj[s] * J[o] - a.kp[s] * P[o] : a.kj[s] * J[o];
        sum_u += u[s];
    }
    const T e_new = ca * e_old + cb * (curl_a - curl_b) - a.Cu * sum_u;
    for (int s = 0; s < a.n_poles; s++) {
        const long long o = s * slot_stride + off;
        const T S = u[s] + a.v[s] * (e_new + e_old);
        J[o] = S - J[o];
        if (a.has_p[s]) P[o] = P[o] + a.hdt * S;
    }
    return e_new;
}

}  // namespace rfk

#endif
