// Yee H/E update kernels for gfx950, FP64 (bitwise the reference arithmetic)
// and FP32 (fields, CPML and ADE in float).
//
// Layout: the reference padded layout idx = i*sx + j*sy + k, z contiguous
// (engine/fdtd/grid.h:25-30).  One thread owns one node (i,j,k) and updates
// the three components that live there, so every field value a wave loads
// along z is a coalesced 64-lane row.  The step is memory-bound: each node
// reads 6 field values plus its neighbours (served by L1/L2 from adjacent
// rows and planes) and writes 3.
//
// Compiled with -ffp-contract=off: every multiply and add rounds separately,
// in the order of engine/fdtd/grid.c:70-89, cpml.h:49-50, cpml.c:185-260 and
// materials.c:975-1055, so the FP64 kernels reproduce the reference bits.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rf_device.h"
#include "rf_hip.h"
#include "yee_device.h"

namespace {

using namespace rfk;

// Node (k, j, i) of the thread: blocks of 64 x 4 nodes over (k, j), one
// plane i per grid z, with the blocks renumbered per XCD (xcd_block) so that
// the rows j-1 and plane i-1 a thread reads were loaded on its own XCD.
__device__ __forceinline__ void grid_node(int &k, int &j, int &i) {
    const unsigned nbx = gridDim.x, nby = gridDim.y;
    const unsigned b = xcd_block(blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z), nbx * nby * gridDim.z);
    const unsigned q = b / nbx;
    k = static_cast<int>((b - q * nbx) * blockDim.x + threadIdx.x);
    j = static_cast<int>((q % nby) * blockDim.y + threadIdx.y);
    i = static_cast<int>(q / nby);
}

// ----------------------------------------------------------------- H

template <typename T, bool CPML>
__global__ void __launch_bounds__(256)
k_update_h(rf_geom_t gg, const T *__restrict__ ex, const T *__restrict__ ey,
           const T *__restrict__ ez, T *__restrict__ hx, T *__restrict__ hy, T *__restrict__ hz,
           T cb, T dx, rf_cpml_t cp) {
    const Geo g = geo_of(gg);
    int k, j, i;
    grid_node(k, j, i);
    if (k > g.nz || j > g.ny) return;
    const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;

    const bool do_x = (i < g.nx + !g.px) && j < g.ny && k < g.nz;
    const bool do_y = i < g.nx && (j < g.ny + !g.py) && k < g.nz;
    const bool do_z = i < g.nx && j < g.ny && (k < g.nz + !g.pz);
    if (!(do_x | do_y | do_z)) return;

    // +1 neighbours with periodic wrap (grid.h:26-28).
    const long long di = (g.px && i + 1 == g.nx) ? -(long long)i * g.sx : g.sx;
    const long long dj = (g.py && j + 1 == g.ny) ? -(long long)j * g.sy : g.sy;
    const long long dk = (g.pz && k + 1 == g.nz) ? -(long long)k : 1;

    // every load of the node first, then the updates, then the stores
    // (term_load, yee_device.h)
    const T exc = ex[idx], eyc = ey[idx], ezc = ez[idx];
    T ez_j = T(0), ey_k = T(0), ex_k = T(0), ez_i = T(0), ey_i = T(0), ex_j = T(0);
    T hxo = T(0), hyo = T(0), hzo = T(0);
    if (do_x) {
        ez_j = ez[idx + dj];
        ey_k = ey[idx + dk];
        hxo = hx[idx];
    }
    if (do_y) {
        ex_k = ex[idx + dk];
        ez_i = ez[idx + di];
        hyo = hy[idx];
    }
    if (do_z) {
        ey_i = ey[idx + di];
        ex_j = ex[idx + dj];
        hzo = hz[idx];
    }
    Term<T> ta[3] = {}, tb[3] = {};
    if (CPML) {
        if (do_x) {
            ta[0] = term_load<T>(cp, g, true, 1, j, cp.psi_h[0][0], i, j, k);
            tb[0] = term_load<T>(cp, g, true, 2, k, cp.psi_h[0][1], i, j, k);
        }
        if (do_y) {
            ta[1] = term_load<T>(cp, g, true, 2, k, cp.psi_h[1][0], i, j, k);
            tb[1] = term_load<T>(cp, g, true, 0, i, cp.psi_h[1][1], i, j, k);
        }
        if (do_z) {
            ta[2] = term_load<T>(cp, g, true, 0, i, cp.psi_h[2][0], i, j, k);
            tb[2] = term_load<T>(cp, g, true, 1, j, cp.psi_h[2][1], i, j, k);
        }
    }
    if (do_x) {   // grid.c:70, cpml.c:185-189
        T dA = ez_j - ezc;
        T dB = ey_k - eyc;
        if (CPML) {
            dA = term_apply<T>(ta[0], dA, dx);
            dB = term_apply<T>(tb[0], dB, dx);
            term_store<T>(ta[0], cp.psi_h[0][0]);
            term_store<T>(tb[0], cp.psi_h[0][1]);
        }
        hx[idx] = hxo - cb * (dA - dB);
    }
    if (do_y) {   // grid.c:78, cpml.c:198-202
        T dA = ex_k - exc;
        T dB = ez_i - ezc;
        if (CPML) {
            dA = term_apply<T>(ta[1], dA, dx);
            dB = term_apply<T>(tb[1], dB, dx);
            term_store<T>(ta[1], cp.psi_h[1][0]);
            term_store<T>(tb[1], cp.psi_h[1][1]);
        }
        hy[idx] = hyo - cb * (dA - dB);
    }
    if (do_z) {   // grid.c:86, cpml.c:211-215
        T dA = ey_i - eyc;
        T dB = ex_j - exc;
        if (CPML) {
            dA = term_apply<T>(ta[2], dA, dx);
            dB = term_apply<T>(tb[2], dB, dx);
            term_store<T>(ta[2], cp.psi_h[2][0]);
            term_store<T>(tb[2], cp.psi_h[2][1]);
        }
        hz[idx] = hzo - cb * (dA - dB);
    }
}

// ----------------------------------------------------------------- E

// E update of component c with old value e (materials.c:975-1058 for one
// component): false for a tensor edge, which this kernel leaves alone.
template <typename T, typename TagT>
__device__ __forceinline__ bool e_value(int c, T &e, T dA, T dB, const rf_mat_t &m, uint32_t fl, T ca, T cb,
                                        int aid, int i, int j, int k) {
    using Ade = typename EntryOf<T>::ade;
    if constexpr (std::is_same<TagT, NoTag>::value) {
        e = T(1) * e + static_cast<T>(m.vac_ca) * (dA - dB);
    } else {
        if (fl & RF_EDGE_TENSOR) return false;
        if (fl & RF_EDGE_ADE) {
            const Ade &a = static_cast<const Ade *>(m.ade)[aid];
            e = ade_edge<T>(a, ca, cb, e, dA, dB, static_cast<T *>(m.J), static_cast<T *>(m.P), 3 * a.st,
                            ade_off(a, c, i, j, k));
            return true;
        }
        e = ca * e + cb * (dA - dB);
    }
    return true;
}

template <typename T, typename TagT, bool CPML>
__global__ void __launch_bounds__(256)
k_update_e(rf_geom_t gg, T *__restrict__ ex, T *__restrict__ ey, T *__restrict__ ez,
           const T *__restrict__ hx, const T *__restrict__ hy, const T *__restrict__ hz,
           T dx, rf_cpml_t cp, rf_mat_t m, const rf_point_src_t *__restrict__ pts, int n_pts,
           const long long *__restrict__ d_step) {
    const Geo g = geo_of(gg);
    int k, j, i;
    grid_node(k, j, i);
    if (k >= g.nz || j >= g.ny) return;
    const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;

    // E ranges (grid.c:95-117): own axis from 0, the other axes from 1
    // unless periodic (PEC planes at 0 and n are never updated).
    const bool i_ok = g.px || i >= 1, j_ok = g.py || j >= 1, k_ok = g.pz || k >= 1;
    const bool do_x = j_ok && k_ok;
    const bool do_y = i_ok && k_ok;
    const bool do_z = i_ok && j_ok;
    if (!(do_x | do_y | do_z)) return;

    // -1 neighbours with wrap at 0 (materials.c:808-810).
    const long long di = (i == 0) ? (long long)(g.nx - 1) * g.sx : -g.sx;
    const long long dj = (j == 0) ? (long long)(g.ny - 1) * g.sy : -g.sy;
    const long long dk = (k == 0) ? (long long)(g.nz - 1) : -1;

    unsigned tag = 0;
    if constexpr (!std::is_same<TagT, NoTag>::value) tag = load_tag<TagT>(m.tag, idx);

    // every load of the node first (fields, psi and the material entry),
    // then the updates, then the stores (term_load, yee_device.h)
    const T hxc = hx[idx], hyc = hy[idx], hzc = hz[idx];
    T hz_j = T(0), hy_k = T(0), hx_k = T(0), hz_i = T(0), hy_i = T(0), hx_j = T(0);
    T e[3] = {T(0), T(0), T(0)};
    if (do_x) {
        hz_j = hz[idx + dj];
        hy_k = hy[idx + dk];
        e[0] = ex[idx];
    }
    if (do_y) {
        hx_k = hx[idx + dk];
        hz_i = hz[idx + di];
        e[1] = ey[idx];
    }
    if (do_z) {
        hy_i = hy[idx + di];
        hx_j = hx[idx + dj];
        e[2] = ez[idx];
    }
    // the point sources on E^n first, in insertion order (k_point_sources;
    // the host passes them only when each sits on a component this thread
    // updates)
    if (n_pts > 0) {
        const long long step = *d_step;
        for (int q = 0; q < n_pts; q++) {
            const rf_point_src_t &p = pts[q];
            if (p.idx == idx) {
                const double v = ((gmem_c<double> *)p.wave)[step];
#pragma unroll
                for (int c = 0; c < 3; c++)
                    if (p.comp == c) e[c] = p.hard ? static_cast<T>(v) : static_cast<T>(static_cast<double>(e[c]) + v);
            }
        }
    }
    Term<T> ta[3] = {}, tb[3] = {};
    if (CPML) {
        if (do_x) {
            ta[0] = term_load<T>(cp, g, false, 1, j, cp.psi_e[0][0], i, j, k);
            tb[0] = term_load<T>(cp, g, false, 2, k, cp.psi_e[0][1], i, j, k);
        }
        if (do_y) {
            ta[1] = term_load<T>(cp, g, false, 2, k, cp.psi_e[1][0], i, j, k);
            tb[1] = term_load<T>(cp, g, false, 0, i, cp.psi_e[1][1], i, j, k);
        }
        if (do_z) {
            ta[2] = term_load<T>(cp, g, false, 0, i, cp.psi_e[2][0], i, j, k);
            tb[2] = term_load<T>(cp, g, false, 1, j, cp.psi_e[2][1], i, j, k);
        }
    }
    uint32_t fl[3] = {0u, 0u, 0u};
    T ca[3] = {T(0), T(0), T(0)}, cbv[3] = {T(0), T(0), T(0)};
    int aid[3] = {0, 0, 0};
    if constexpr (!std::is_same<TagT, NoTag>::value) {
        typedef const __attribute__((address_space(1))) typename EntryOf<T>::type g_entry;
        const g_entry &en = ((g_entry *)m.entries)[tag];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            fl[c] = en.flags[c];
            ca[c] = en.Ca[c];
            cbv[c] = en.Cb[c];
            aid[c] = en.ade[c];
        }
    }
    bool put[3] = {false, false, false};
    if (do_x) {   // materials.c:975-1000
        T dA = hzc - hz_j;
        T dB = hyc - hy_k;
        if (CPML) {
            dA = term_apply<T>(ta[0], dA, dx);
            dB = term_apply<T>(tb[0], dB, dx);
        }
        put[0] = e_value<T, TagT>(0, e[0], dA, dB, m, fl[0], ca[0], cbv[0], aid[0], i, j, k);
    }
    if (do_y) {   // materials.c:1015-1040
        T dA = hxc - hx_k;
        T dB = hzc - hz_i;
        if (CPML) {
            dA = term_apply<T>(ta[1], dA, dx);
            dB = term_apply<T>(tb[1], dB, dx);
        }
        put[1] = e_value<T, TagT>(1, e[1], dA, dB, m, fl[1], ca[1], cbv[1], aid[1], i, j, k);
    }
    if (do_z) {   // materials.c:1050-1058
        T dA = hyc - hy_i;
        T dB = hxc - hx_j;
        if (CPML) {
            dA = term_apply<T>(ta[2], dA, dx);
            dB = term_apply<T>(tb[2], dB, dx);
        }
        put[2] = e_value<T, TagT>(2, e[2], dA, dB, m, fl[2], ca[2], cbv[2], aid[2], i, j, k);
    }
    if (CPML) {
        if (do_x) {
            term_store<T>(ta[0], cp.psi_e[0][0]);
            term_store<T>(tb[0], cp.psi_e[0][1]);
        }
        if (do_y) {
            term_store<T>(ta[1], cp.psi_e[1][0]);
            term_store<T>(tb[1], cp.psi_e[1][1]);
        }
        if (do_z) {
            term_store<T>(ta[2], cp.psi_e[2][0]);
            term_store<T>(tb[2], cp.psi_e[2][1]);
        }
    }
    if (put[0]) ex[idx] = e[0];
    if (put[1]) ey[idx] = e[1];
    if (put[2]) ez[idx] = e[2];
}

// ----------------------------------------------------------- tensors

template <typename T> struct TEdgeOf;
template <> struct TEdgeOf<double> { using type = rf_tedge64_t; };
template <> struct TEdgeOf<float> { using type = rf_tedge32_t; };

struct Pos { int x[3]; };

__device__ __forceinline__ int nxt(const Geo &g, int ax, int q) {
    const int n = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
    return q + 1 == n ? 0 : q + 1;
}
__device__ __forceinline__ int prv(const Geo &g, int ax, int q) {
    const int n = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
    return q == 0 ? n - 1 : q - 1;
}
__device__ __forceinline__ long long lin(const Geo &g, const Pos &p) {
    return (long long)p.x[0] * g.sx + (long long)p.x[1] * g.sy + p.x[2];
}

template <typename T>
__device__ __forceinline__ T curl_full(const Geo &g, const T *const H[3], int c, const Pos &x) {
    const int p = (c + 1) % 3, q = (c + 2) % 3;
    Pos xp = x, xq = x;
    xp.x[p] = prv(g, p, x.x[p]);
    xq.x[q] = prv(g, q, x.x[q]);
    const long long idx = lin(g, x);
    const T dA = H[q][idx] - H[q][lin(g, xp)];
    const T dB = H[p][idx] - H[p][lin(g, xq)];
    return dA - dB;
}

// Hz-only curls of the reference's z-invariant stencil (materials.c:856-901).
template <typename T>
__device__ __forceinline__ T curl_hz(const Geo &g, const T *Hz, int c, const Pos &x) {
    Pos y = x;
    if (c == 0) {
        y.x[1] = prv(g, 1, x.x[1]);
        return Hz[lin(g, x)] - Hz[lin(g, y)];
    }
    y.x[0] = prv(g, 0, x.x[0]);
    return -(Hz[lin(g, x)] - Hz[lin(g, y)]);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_tensor(rf_geom_t gg, T *ex, T *ey, T *ez, const T *hx, const T *hy, const T *hz,
         const typename TEdgeOf<T>::type *__restrict__ edges, long long n_edges) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_edges) return;
    const Geo g = geo_of(gg);
    const typename TEdgeOf<T>::type e = edges[t];
    const int c = e.comp;
    const bool hz_only = e.kind == 2;   // materials.c:858, 881
    const int x0 = e.i, x1 = e.j, x2 = e.k;
    // The nine curls of the edge (n = 0 its own, n = 1 + 4 slot + q the
    // neighbour curls, every neighbour position a wrapped in-grid node) are
    // gathered first, with the positions, arrays and indices chosen by
    // selects so that all their loads issue together; the masked neighbour
    // curls are then added in slot order, the same operations as adding each
    // present curl in turn.  A curl of component cc at y reads
    // a = H[(cc+2)%3] at y and at y less one along (cc+1)%3, and
    // b = H[(cc+1)%3] at y and at y less one along (cc+2)%3; the Hz-only
    // curls read a = Hz at y and at y less one along axis 1 (cc = 0) or
    // axis 0, and negate for cc != 0 (their b loads are made and unused).
    T a0[9], a1[9], b0[9], b1[9];
    int cc_of[9];
#pragma unroll
    for (int n = 0; n < 9; n++) {
        int y0 = x0, y1 = x1, y2 = x2, cc = c;
        if (n > 0) {
            const int slot = (n - 1) >> 2, q = (n - 1) & 3;
            // the other two components in increasing order
            const int d = slot == 0 ? (c == 0 ? 1 : 0) : (c == 2 ? 1 : 2);
            const int lo = c < d ? c : d;
            const int s_lo = q & 1, s_hi = (q >> 1) & 1;
            const bool s_c = ((c == lo) ? s_lo : s_hi) != 0;
            const bool s_d = ((d == lo) ? s_lo : s_hi) != 0;
            y0 = (s_c && c == 0) ? nxt(g, 0, x0) : (s_d && d == 0) ? prv(g, 0, x0) : x0;
            y1 = (s_c && c == 1) ? nxt(g, 1, x1) : (s_d && d == 1) ? prv(g, 1, x1) : x1;
            y2 = (s_c && c == 2) ? nxt(g, 2, x2) : (s_d && d == 2) ? prv(g, 2, x2) : x2;
            cc = d;
        }
        cc_of[n] = cc;
        const int pa = (cc + 1) % 3, qa = (cc + 2) % 3;
        // the axis of a's second node and of b's
        const int ax_a = hz_only ? (cc == 0 ? 1 : 0) : pa;
        const int ax_b = qa;
        const long long idx_y = (long long)y0 * g.sx + (long long)y1 * g.sy + y2;
        const long long ia1 = (long long)(ax_a == 0 ? prv(g, 0, y0) : y0) * g.sx +
                              (long long)(ax_a == 1 ? prv(g, 1, y1) : y1) * g.sy + (ax_a == 2 ? prv(g, 2, y2) : y2);
        const long long ib1 = (long long)(ax_b == 0 ? prv(g, 0, y0) : y0) * g.sx +
                              (long long)(ax_b == 1 ? prv(g, 1, y1) : y1) * g.sy + (ax_b == 2 ? prv(g, 2, y2) : y2);
        const T *ha = hz_only ? hz : (qa == 0 ? hx : qa == 1 ? hy : hz);
        const T *hb = pa == 0 ? hx : pa == 1 ? hy : hz;
        a0[n] = ha[idx_y];
        a1[n] = ha[ia1];
        b0[n] = hb[idx_y];
        b1[n] = hb[ib1];
    }
    T value;
    if (hz_only)
        value = c == 0 ? e.diag * (a0[0] - a1[0]) : -e.diag * (a0[0] - a1[0]);
    else
        value = e.diag * ((a0[0] - a1[0]) - (b0[0] - b1[0]));
    T cd[8];
#pragma unroll
    for (int t8 = 0; t8 < 8; t8++) {
        const int n = t8 + 1;
        const T da = a0[n] - a1[n];
        cd[t8] = hz_only ? (cc_of[n] == 0 ? da : -da) : da - (b0[n] - b1[n]);
    }
    for (int t8 = 0; t8 < 8; t8++)
        if ((e.mask >> t8) & 1u) value += e.w[t8] * cd[t8];
    T *const ec = c == 0 ? ex : c == 1 ? ey : ez;
    const long long idx = (long long)x0 * g.sx + (long long)x1 * g.sy + x2;
    ec[idx] += e.scale * value;
}

// ----------------------------------------------------------- sources

template <typename T>
__global__ void k_point_sources(T *ex, T *ey, T *ez, const rf_point_src_t *srcs, int n,
                                const long long *d_step, int step_offset) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long step = *d_step + step_offset;
    T *const E[3] = {ex, ey, ez};
    // Sequential in insertion order, like repeated source_apply calls.
    for (int s = 0; s < n; s++) {
        const rf_point_src_t p = srcs[s];
        const double v = p.wave[step];
        T *f = E[p.comp];
        if (p.hard) f[p.idx] = static_cast<T>(v);
        else f[p.idx] = static_cast<T>(static_cast<double>(f[p.idx]) + v);
    }
}

template <typename T>
__global__ void k_plane_source(rf_geom_t gg, T *ex, T *ey, T *ez, rf_plane_src_t p,
                               const long long *d_step, int step_offset) {
    const Geo g = geo_of(gg);
    const int a = (p.axis + 1) % 3 < (p.axis + 2) % 3 ? (p.axis + 1) % 3 : (p.axis + 2) % 3;
    const int b = 3 - p.axis - a;
    const int na = a == 0 ? g.nx : a == 1 ? g.ny : g.nz;
    const int nb = b == 0 ? g.nx : b == 1 ? g.ny : g.nz;
    const int ib = blockIdx.x * blockDim.x + threadIdx.x;
    const int ia = blockIdx.y;
    if (ib >= nb || ia >= na) return;
    Pos x;
    x.x[p.axis] = p.index;
    x.x[a] = ia;
    x.x[b] = ib;
    // fdtd_index wraps the plane coordinate on a periodic axis.
    const int nplane = p.axis == 0 ? g.nx : p.axis == 1 ? g.ny : g.nz;
    const bool wrap = (gg.per >> p.axis) & 1;
    if (wrap) x.x[p.axis] %= nplane;
    const long long idx = lin(g, x);
    const double v = p.wave[*d_step + step_offset];
    T *f = p.comp == 0 ? ex : p.comp == 1 ? ey : ez;
    if (p.hard) f[idx] = static_cast<T>(v);
    else f[idx] = static_cast<T>(static_cast<double>(f[idx]) + v);
}

// ------------------------------------------------------------ probes

template <typename T>
__global__ void __launch_bounds__(256)
k_probe_record(const T *ex, const T *ey, const T *ez, const T *hx, const T *hy, const T *hz,
               const long long *idx, const int *comp, int n, double *ring, int ring_len,
               double *td, long long td_cap, long long *d_step) {
    const long long step = *d_step;
    const T *const F[6] = {ex, ey, ez, hx, hy, hz};
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const double v = static_cast<double>(F[comp[p]][idx[p]]);
        ring[(step % ring_len) * n + p] = v;
        if (step < td_cap) td[(long long)p * td_cap + step] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) *d_step = step + 1;
}

// Probe gathers of many scenes: block b records the probes of job b and
// advances its step counter (k_probe_record per job).
template <typename T>
__global__ void __launch_bounds__(256)
k_probe_record_packed(const rf_probe_job_t *__restrict__ jobs) {
    const rf_probe_job_t &jb = jobs[blockIdx.x];
    const long long step = *jb.d_step;
    const T *const *F = reinterpret_cast<const T *const *>(jb.f);
    for (int p = threadIdx.x; p < jb.n; p += blockDim.x) {
        const double v = static_cast<double>(F[jb.comp[p]][jb.idx[p]]);
        jb.ring[(step % jb.ring_len) * jb.n + p] = v;
        if (step < jb.td_cap) jb.td[(long long)p * jb.td_cap + step] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) *jb.d_step = step + 1;
}

__device__ __forceinline__ void neumaier(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// One thread per (probe, frequency); sums over the buffered steps in time
// order, as probe_dft_record does step by step (probe.c:87-91).
__global__ void __launch_bounds__(256)
k_probe_dft(const double *ring, int ring_len, int n_probes, const int *freq_probe, int n_freq,
            const double *freq_w, double dt, const long long *d_step, int count, double *acc_re,
            double *acc_im, double *cmp_re, double *cmp_im, int comp_sum) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_freq) return;
    const int p = freq_probe[t];
    const double w = freq_w[t];
    const long long end = *d_step;
    double re = acc_re[t], im = acc_im[t];
    double cr = comp_sum ? cmp_re[t] : 0.0, ci = comp_sum ? cmp_im[t] : 0.0;
    for (long long n = end - count; n < end; n++) {
        const double v = ring[(n % ring_len) * n_probes + p];
        const double tt = static_cast<double>(n + 1) * dt;
        double sn, cs;
        sincos(w * tt, &sn, &cs);
        const double tr = v * cs * dt;
        const double ti = v * sn * dt;
        if (comp_sum) {
            neumaier(re, cr, tr);
            neumaier(im, ci, -ti);
        } else {
            re += tr;
            im -= ti;
        }
    }
    acc_re[t] = re;
    acc_im[t] = im;
    if (comp_sum) {
        cmp_re[t] = cr;
        cmp_im[t] = ci;
    }
}

// Point-probe DFTs of many scenes: blocks (x, y) take the pairs
// 256 x .. 256 x + 255 of job y, each summed as k_probe_dft sums it.
__global__ void __launch_bounds__(256)
k_probe_dft_packed(const rf_pdft_job_t *__restrict__ jobs, int count) {
    const rf_pdft_job_t &jb = jobs[blockIdx.y];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= jb.n_freq) return;
    const int p = jb.freq_probe[t];
    const double w = jb.freq_w[t];
    const double dt = jb.dt;
    const long long end = *jb.d_step;
    double re = jb.acc_re[t], im = jb.acc_im[t];
    double cr = jb.comp_sum ? jb.cmp_re[t] : 0.0, ci = jb.comp_sum ? jb.cmp_im[t] : 0.0;
    for (long long n = end - count; n < end; n++) {
        const double v = jb.ring[(n % jb.ring_len) * jb.n_probes + p];
        const double tt = static_cast<double>(n + 1) * dt;
        double sn, cs;
        sincos(w * tt, &sn, &cs);
        const double tr = v * cs * dt;
        const double ti = v * sn * dt;
        if (jb.comp_sum) {
            neumaier(re, cr, tr);
            neumaier(im, ci, -ti);
        } else {
            re += tr;
            im -= ti;
        }
    }
    jb.acc_re[t] = re;
    jb.acc_im[t] = im;
    if (jb.comp_sum) {
        jb.cmp_re[t] = cr;
        jb.cmp_im[t] = ci;
    }
}

// ----------------------------------------------------- plane monitors

// Every monitor component of a step in one launch: capture q takes the
// threads [start[q], start[q] + na nb) and copies its plane, cell (ia, ib)
// with ib fastest, into slot step % ring_len of its ring.  One launch of
// all planes keeps more of the plane gathers (one line per value for a z
// plane) in flight than a launch per plane.
template <typename T>
__global__ void __launch_bounds__(256)
k_plane_captures(rf_geom_t gg, rf_caps_t cs, int ring_len, const long long *d_step) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    plane_capture_elem<T>(gg, cs, t, ring_len, *d_step);
}

// Twiddles for the buffered block, scaled by dt, frequency-major:
// tw[f][s] = dt (cos, sin)(w_f (n_s+1) dt), rows of RF_PDFT_TWS steps.
__global__ void k_twiddles(int nf, const double *freq_w, double dt, const long long *d_step,
                           int count, double2 *tw) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * count) return;
    const int s = t % count, f = t / count;
    const long long n = *d_step - count + s;
    const double tt = static_cast<double>(n + 1) * dt;
    double sn, cs;
    sincos(freq_w[f] * tt, &sn, &cs);
    tw[(long long)f * RF_PDFT_TWS + s] = make_double2(cs * dt, sn * dt);
}

// Time-blocked plane DFT as a register-tiled product: a block takes 64
// cells of the plane and, frequency tile by frequency tile (32 at a time),
// adds the buffered samples off .. off + cnt - 1 of the block of `count`
// steps to the FP64 accumulators [f][cell], which are read and written
// once per pass.  The samples of the block's cells for the pass and the
// tile's twiddles tw[f][s] = dt (cos, sin)(w_f t_s) are staged in LDS;
// thread (cg, fg) holds the accumulators of cells cg + 16 i (i < 4) and
// frequencies fg + 16 j (j < 2), and adds the terms of each output in
// time order with one fused multiply-add each, so every accumulator sees
// the same sequence of operations whatever the block length.
constexpr int PD_CELLS = 64, PD_FT = 32;
template <typename B, int KP>
__global__ void __launch_bounds__(256)
k_plane_dft(const B *__restrict__ buf, long long cells, int ring_len, int nf, const double2 *__restrict__ tw,
            const long long *__restrict__ d_step, int count, int off, int cnt, double *__restrict__ acc_re,
            double *__restrict__ acc_im) {
    __shared__ B s_x[PD_CELLS][KP + 1];          // samples, padded rows
    __shared__ double2 s_w[KP][PD_FT];           // twiddles of the tile
    const int t = threadIdx.x;
    const long long c0 = (long long)blockIdx.x * PD_CELLS;
    const long long first = *d_step - count + off;
    for (int e = t; e < PD_CELLS * KP; e += 256) {
        const int c = e % PD_CELLS, q = e / PD_CELLS;
        const long long cell = c0 + c;
        s_x[c][q] = q < cnt && cell < cells ? buf[((first + q) % ring_len) * cells + cell] : B(0);
    }
    const int cg = t % 16, fg = t / 16;
    for (int ft = 0; ft < nf; ft += PD_FT) {
        // the tile's accumulators, in flight while its twiddles are staged
        double re[4][2], im[4][2];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const long long cell = c0 + cg + 16 * i;
                const int f = ft + fg + 16 * j;
                const bool ok = cell < cells && f < nf;
                re[i][j] = ok ? acc_re[(long long)f * cells + cell] : 0.0;
                im[i][j] = ok ? acc_im[(long long)f * cells + cell] : 0.0;
            }
        __syncthreads();   // the previous tile's twiddles are read
        for (int e = t; e < KP * PD_FT; e += 256) {
            const int q = e % KP, f = e / KP;
            s_w[q][f] = q < cnt && ft + f < nf ? tw[(long long)(ft + f) * RF_PDFT_TWS + off + q] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int q = 0; q < cnt; q++) {
            double x[4];
#pragma unroll
            for (int i = 0; i < 4; i++) x[i] = static_cast<double>(s_x[cg + 16 * i][q]);
            const double2 w0 = s_w[q][fg], w1 = s_w[q][fg + 16];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                re[i][0] = __builtin_fma(x[i], w0.x, re[i][0]);
                im[i][0] = __builtin_fma(-x[i], w0.y, im[i][0]);
                re[i][1] = __builtin_fma(x[i], w1.x, re[i][1]);
                im[i][1] = __builtin_fma(-x[i], w1.y, im[i][1]);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const long long cell = c0 + cg + 16 * i;
                const int f = ft + fg + 16 * j;
                if (cell < cells && f < nf) {
                    acc_re[(long long)f * cells + cell] = re[i][j];
                    acc_im[(long long)f * cells + cell] = im[i][j];
                }
            }
    }
}

// ------------------------------------------------------------ energy

template <typename T, typename TagT>
__global__ void __launch_bounds__(256)
k_energy(rf_geom_t gg, const T *ex, const T *ey, const T *ez, const T *hx, const T *hy,
         const T *hz, const void *tag, const double *eps_tab, double vac_eps, double mu0,
         double *partials) {
    const Geo g = geo_of(gg);
    const long long cells = (long long)g.nx * g.ny * g.nz;
    double ue = 0.0, uh = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long long r = t / g.nz;
        const int j = (int)(r % g.ny), i = (int)(r / g.ny);
        const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;
        double eps[3] = {vac_eps, vac_eps, vac_eps};
        if constexpr (!std::is_same<TagT, NoTag>::value) {
            const unsigned tg = load_tag<TagT>(tag, idx);
            eps[0] = eps_tab[3 * tg]; eps[1] = eps_tab[3 * tg + 1]; eps[2] = eps_tab[3 * tg + 2];
        }
        const double e0 = ex[idx], e1 = ey[idx], e2 = ez[idx];
        const double h0 = hx[idx], h1 = hy[idx], h2 = hz[idx];
        ue += 0.5 * eps[0] * e0 * e0 + 0.5 * eps[1] * e1 * e1 + 0.5 * eps[2] * e2 * e2;
        uh += 0.5 * mu0 * h0 * h0 + 0.5 * mu0 * h1 * h1 + 0.5 * mu0 * h2 * h2;
    }
    __shared__ double se[256], sh[256];
    se[threadIdx.x] = ue;
    sh[threadIdx.x] = uh;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sh[threadIdx.x] += sh[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = se[0];
        partials[2 * blockIdx.x + 1] = sh[0];
    }
}

__global__ void k_energy_final(const double *partials, int n, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double e = 0.0, h = 0.0;
    for (int b = 0; b < n; b++) {
        e += partials[2 * b];
        h += partials[2 * b + 1];
    }
    out[0] = e;
    out[1] = h;
}

// Energy audit of step n (audit_periodic_energy.c:43-71): from E^n and
// H^{n-1/2} (state n) and H^{n+1/2} (state n+1), per block the sums over
// i<nx, j<ny, k<nz of 0.5 eps E^2, 0.5 mu0 (H-)^2, 0.5 mu0 (H+)^2,
// 0.5 mu0 ((H- + H+)/2)^2 and 0.5 mu0 H- H+ (the last one makes the
// modified energy E-part + it, invariant in a closed lossless grid).
template <typename T, typename TagT>
__global__ void __launch_bounds__(256)
k_audit(rf_geom_t gg, const T *ex, const T *ey, const T *ez, const T *hxm, const T *hym, const T *hzm,
        const T *hxp, const T *hyp, const T *hzp, const void *tag, const double *eps_tab, double vac_eps,
        double mu0, double *partials) {
    const Geo g = geo_of(gg);
    const long long cells = (long long)g.nx * g.ny * g.nz;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % g.nz);
        const long long r = t / g.nz;
        const int j = (int)(r % g.ny), i = (int)(r / g.ny);
        const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;
        double eps[3] = {vac_eps, vac_eps, vac_eps};
        if constexpr (!std::is_same<TagT, NoTag>::value) {
            const unsigned tg = load_tag<TagT>(tag, idx);
            eps[0] = eps_tab[3 * tg]; eps[1] = eps_tab[3 * tg + 1]; eps[2] = eps_tab[3 * tg + 2];
        }
        const double e[3] = {(double)ex[idx], (double)ey[idx], (double)ez[idx]};
        const double hm[3] = {(double)hxm[idx], (double)hym[idx], (double)hzm[idx]};
        const double hp[3] = {(double)hxp[idx], (double)hyp[idx], (double)hzp[idx]};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double avg = 0.5 * (hm[c] + hp[c]);
            acc[0] += 0.5 * eps[c] * e[c] * e[c];
            acc[1] += 0.5 * mu0 * hm[c] * hm[c];
            acc[2] += 0.5 * mu0 * hp[c] * hp[c];
            acc[3] += 0.5 * mu0 * avg * avg;
            acc[4] += 0.5 * mu0 * hm[c] * hp[c];
        }
    }
    __shared__ double sm[5][256];
#pragma unroll
    for (int q = 0; q < 5; q++) sm[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 5; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) partials[5 * blockIdx.x + q] = sm[q][0];
}

// Sum of n_partials blocks of `width` partial sums, in block order, into
// out[0..width) (add = 0) or added to it (add = 1).
__global__ void k_sum_partials(const double *partials, int n, int width, double *out, int add) {
    if (blockIdx.x != 0 || threadIdx.x >= width) return;
    const int q = threadIdx.x;
    double s = 0.0;
    for (int b = 0; b < n; b++) s += partials[width * b + q];
    out[q] = add ? out[q] + s : s;
}

// Poynting flux through the plane of E nodes at index p along axis a,
// at time n+1/2: the sum over the plane's cells of
// (0.5 (E^n + E^{n+1})) x H^{n+1/2} along +a, with E tangential at node p
// and H tangential at array index p (half a cell above the plane).  With
// this pairing the change of the modified energy between two such planes
// is exactly dt times the difference of their fluxes (the discrete
// Poynting theorem of the Yee scheme).
template <typename T>
__global__ void __launch_bounds__(256)
k_flux(rf_geom_t gg, int a, int p, const T *e0n, const T *e1n, const T *e0p, const T *e1p, const T *h0,
       const T *h1, double *partials) {
    const Geo g = geo_of(gg);
    // the two tangential axes b, c (in cyclic order after a)
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const int n[3] = {g.nx, g.ny, g.nz};
    const long long st[3] = {g.sx, g.sy, 1};
    const long long cells = (long long)n[b] * n[c];
    double s = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cells;
         t += (long long)gridDim.x * blockDim.x) {
        const long long ib = t / n[c], ic = t % n[c];
        const long long idx = (long long)p * st[a] + ib * st[b] + ic * st[c];
        // E_b, E_c time-averaged; H_b, H_c at n+1/2: S_a = E_b H_c - E_c H_b
        const double eb = 0.5 * ((double)e0n[idx] + (double)e0p[idx]);
        const double ec = 0.5 * ((double)e1n[idx] + (double)e1p[idx]);
        s += eb * (double)h1[idx] - ec * (double)h0[idx];
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sm[0];
}

template <typename D, typename S>
__global__ void k_convert(D *dst, const S *src, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        dst[t] = static_cast<D>(src[t]);
}

int launched() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -1;
}

template <typename T>
int launch_h(const rf_geom_t *g, void *const f[6], double cb, double dx, const rf_cpml_t *cp,
             hipStream_t st) {
    dim3 blk(64, 4, 1);
    dim3 grd((g->n[2] + 1 + 63) / 64, (g->n[1] + 1 + 3) / 4, g->n[0] + 1);
    rf_cpml_t none{};
    T *F[6];
    for (int q = 0; q < 6; q++) F[q] = static_cast<T *>(f[q]);
    if (cp)
        hipLaunchKernelGGL((k_update_h<T, true>), grd, blk, 0, st, g[0], F[0], F[1], F[2], F[3], F[4],
                           F[5], static_cast<T>(cb), static_cast<T>(dx), *cp);
    else
        hipLaunchKernelGGL((k_update_h<T, false>), grd, blk, 0, st, g[0], F[0], F[1], F[2], F[3],
                           F[4], F[5], static_cast<T>(cb), static_cast<T>(dx), none);
    return launched();
}

template <typename T, typename TagT>
int launch_e_tag(const rf_geom_t *g, T *const F[6], double dx, const rf_cpml_t *cp,
                 const rf_mat_t *m, const rf_point_src_t *pts, int n_pts, const long long *d_step,
                 hipStream_t st) {
    dim3 blk(64, 4, 1);
    dim3 grd((g->n[2] + 63) / 64, (g->n[1] + 3) / 4, g->n[0]);
    rf_cpml_t none{};
    if (cp)
        hipLaunchKernelGGL((k_update_e<T, TagT, true>), grd, blk, 0, st, g[0], F[0], F[1], F[2], F[3],
                           F[4], F[5], static_cast<T>(dx), *cp, *m, pts, n_pts, d_step);
    else
        hipLaunchKernelGGL((k_update_e<T, TagT, false>), grd, blk, 0, st, g[0], F[0], F[1], F[2],
                           F[3], F[4], F[5], static_cast<T>(dx), none, *m, pts, n_pts, d_step);
    return launched();
}

template <typename T>
int launch_e(const rf_geom_t *g, void *const f[6], double dx, const rf_cpml_t *cp,
             const rf_mat_t *m, const rf_point_src_t *pts, int n_pts, const long long *d_step, hipStream_t st) {
    T *F[6];
    for (int q = 0; q < 6; q++) F[q] = static_cast<T *>(f[q]);
    if (!m->tag) return launch_e_tag<T, NoTag>(g, F, dx, cp, m, pts, n_pts, d_step, st);
    switch (m->tag_bytes) {
    case 1: return launch_e_tag<T, uint8_t>(g, F, dx, cp, m, pts, n_pts, d_step, st);
    case 2: return launch_e_tag<T, uint16_t>(g, F, dx, cp, m, pts, n_pts, d_step, st);
    case 4: return launch_e_tag<T, uint32_t>(g, F, dx, cp, m, pts, n_pts, d_step, st);
    }
    return -1;
}

template <typename T, typename TagT>
int launch_energy_t(const rf_geom_t *g, T *const F[6], const void *tag, const double *eps_tab,
                    double vac_eps, double mu0, double *partials, int n_partials, double *out,
                    hipStream_t st) {
    hipLaunchKernelGGL((k_energy<T, TagT>), dim3(n_partials), dim3(256), 0, st, g[0], F[0], F[1],
                       F[2], F[3], F[4], F[5], tag, eps_tab, vac_eps, mu0, partials);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, st, partials, n_partials, out);
    return launched();
}

}  // namespace

static hipStream_t hs(rf_stream_t *s) { return rf_raw_stream(s); }

template <typename T, typename TagT>
int launch_audit_t(const rf_geom_t *g, void *const fn[6], void *const fp[6], const void *tag, const double *eps_tab,
                   double vac_eps, double mu0, double *partials, int n_partials, double *out, hipStream_t st) {
    auto P = [](void *v) { return static_cast<const T *>(v); };
    hipLaunchKernelGGL((k_audit<T, TagT>), dim3(n_partials), dim3(256), 0, st, g[0], P(fn[0]), P(fn[1]), P(fn[2]),
                       P(fn[3]), P(fn[4]), P(fn[5]), P(fp[3]), P(fp[4]), P(fp[5]), tag, eps_tab, vac_eps, mu0,
                       partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partials, n_partials, 5, out, 0);
    return launched();
}

template <typename T>
int launch_audit_p(const rf_geom_t *g, void *const fn[6], void *const fp[6], const void *tag, int tag_bytes,
                   const double *eps_tab, double vac_eps, double mu0, double *partials, int n_partials,
                   double *out, hipStream_t st) {
    if (!tag) return launch_audit_t<T, NoTag>(g, fn, fp, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, st);
    switch (tag_bytes) {
    case 1: return launch_audit_t<T, uint8_t>(g, fn, fp, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, st);
    case 2: return launch_audit_t<T, uint16_t>(g, fn, fp, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, st);
    case 4: return launch_audit_t<T, uint32_t>(g, fn, fp, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, st);
    }
    return -1;
}

template <typename T>
int launch_flux_t(const rf_geom_t *g, int a, int p, void *const fn[6], void *const fp[6], double *partials,
                  int n_partials, double *acc, hipStream_t st) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    auto P = [](void *v) { return static_cast<const T *>(v); };
    hipLaunchKernelGGL((k_flux<T>), dim3(n_partials), dim3(256), 0, st, g[0], a, p, P(fn[b]), P(fn[c]), P(fp[b]),
                       P(fp[c]), P(fp[3 + b]), P(fp[3 + c]), partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partials, n_partials, 1, acc, 1);
    return launched();
}

// Divergence check: *flag |= 1 when any of the n values is a NaN or an
// infinity (one atomic per wavefront that saw one).
template <typename T>
__global__ void __launch_bounds__(256) k_nonfinite(const T *__restrict__ v, long long n, unsigned *flag) {
    bool bad = false;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        bad = bad || !__builtin_isfinite(v[t]);
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

extern "C" {

int rf_launch_update_h(int prec, const rf_geom_t *g, void *const f[6], double cb, double dx,
                       const rf_cpml_t *cpml, rf_stream_t *s) {
    return prec == 0 ? launch_h<double>(g, f, cb, dx, cpml, hs(s))
                     : launch_h<float>(g, f, cb, dx, cpml, hs(s));
}

int rf_launch_update_e(int prec, const rf_geom_t *g, void *const f[6], double dx,
                       const rf_cpml_t *cpml, const rf_mat_t *mat, rf_stream_t *s) {
    return rf_launch_update_e_src(prec, g, f, dx, cpml, mat, NULL, 0, NULL, s);
}

int rf_launch_update_e_src(int prec, const rf_geom_t *g, void *const f[6], double dx, const rf_cpml_t *cpml,
                           const rf_mat_t *mat, const rf_point_src_t *pts, int n_pts, const long long *d_step,
                           rf_stream_t *s) {
    if (n_pts > 0 && (!pts || !d_step)) return -1;
    return prec == 0 ? launch_e<double>(g, f, dx, cpml, mat, pts, n_pts, d_step, hs(s))
                     : launch_e<float>(g, f, dx, cpml, mat, pts, n_pts, d_step, hs(s));
}

int rf_launch_tensor(int prec, const rf_geom_t *g, void *const f[6], const void *edges,
                     long long n_edges, rf_stream_t *s) {
    if (n_edges <= 0) return 0;
    dim3 grd((unsigned)((n_edges + 255) / 256));
    if (prec == 0) {
        double *F[6];
        for (int q = 0; q < 6; q++) F[q] = static_cast<double *>(f[q]);
        hipLaunchKernelGGL(k_tensor<double>, grd, dim3(256), 0, hs(s), g[0], F[0], F[1], F[2], F[3],
                           F[4], F[5], static_cast<const rf_tedge64_t *>(edges), n_edges);
    } else {
        float *F[6];
        for (int q = 0; q < 6; q++) F[q] = static_cast<float *>(f[q]);
        hipLaunchKernelGGL(k_tensor<float>, grd, dim3(256), 0, hs(s), g[0], F[0], F[1], F[2], F[3],
                           F[4], F[5], static_cast<const rf_tedge32_t *>(edges), n_edges);
    }
    return launched();
}

int rf_launch_point_sources(int prec, void *const f[6], const rf_point_src_t *srcs_dev, int n,
                            const long long *d_step, int step_offset, rf_stream_t *s) {
    if (n <= 0) return 0;
    if (prec == 0)
        hipLaunchKernelGGL(k_point_sources<double>, dim3(1), dim3(64), 0, hs(s),
                           static_cast<double *>(f[0]), static_cast<double *>(f[1]),
                           static_cast<double *>(f[2]), srcs_dev, n, d_step, step_offset);
    else
        hipLaunchKernelGGL(k_point_sources<float>, dim3(1), dim3(64), 0, hs(s),
                           static_cast<float *>(f[0]), static_cast<float *>(f[1]),
                           static_cast<float *>(f[2]), srcs_dev, n, d_step, step_offset);
    return launched();
}

int rf_launch_plane_source(int prec, const rf_geom_t *g, void *const f[6], const rf_plane_src_t *p,
                           const long long *d_step, int step_offset, rf_stream_t *s) {
    const int a = (p->axis + 1) % 3 < (p->axis + 2) % 3 ? (p->axis + 1) % 3 : (p->axis + 2) % 3;
    const int b = 3 - p->axis - a;
    dim3 grd((g->n[b] + 255) / 256, g->n[a]);
    if (prec == 0)
        hipLaunchKernelGGL(k_plane_source<double>, grd, dim3(256), 0, hs(s), g[0],
                           static_cast<double *>(f[0]), static_cast<double *>(f[1]),
                           static_cast<double *>(f[2]), *p, d_step, step_offset);
    else
        hipLaunchKernelGGL(k_plane_source<float>, grd, dim3(256), 0, hs(s), g[0],
                           static_cast<float *>(f[0]), static_cast<float *>(f[1]),
                           static_cast<float *>(f[2]), *p, d_step, step_offset);
    return launched();
}

int rf_launch_probe_record(int prec, void *const f[6], const long long *idx_dev, const int *comp_dev,
                           int n, double *ring, int ring_len, double *td, long long td_cap,
                           long long *d_step, rf_stream_t *s) {
    if (prec == 0) {
        double *F[6];
        for (int q = 0; q < 6; q++) F[q] = static_cast<double *>(f[q]);
        hipLaunchKernelGGL(k_probe_record<double>, dim3(1), dim3(256), 0, hs(s), F[0], F[1], F[2],
                           F[3], F[4], F[5], idx_dev, comp_dev, n, ring, ring_len, td, td_cap, d_step);
    } else {
        float *F[6];
        for (int q = 0; q < 6; q++) F[q] = static_cast<float *>(f[q]);
        hipLaunchKernelGGL(k_probe_record<float>, dim3(1), dim3(256), 0, hs(s), F[0], F[1], F[2],
                           F[3], F[4], F[5], idx_dev, comp_dev, n, ring, ring_len, td, td_cap, d_step);
    }
    return launched();
}

int rf_launch_probe_dft(const double *ring, int ring_len, int n_probes, const int *freq_probe,
                        int n_freq, const double *freq_w, double dt, const long long *d_step,
                        int count, double *acc_re, double *acc_im, double *cmp_re, double *cmp_im,
                        int comp_sum, rf_stream_t *s) {
    if (n_freq <= 0 || count <= 0) return 0;
    hipLaunchKernelGGL(k_probe_dft, dim3((n_freq + 255) / 256), dim3(256), 0, hs(s), ring, ring_len,
                       n_probes, freq_probe, n_freq, freq_w, dt, d_step, count, acc_re, acc_im,
                       cmp_re, cmp_im, comp_sum);
    return launched();
}

int rf_launch_probe_record_packed(int prec, const rf_probe_job_t *jobs, int n_jobs, rf_stream_t *s) {
    if (n_jobs <= 0) return 0;
    if (!jobs) return -1;
    if (prec == 0)
        hipLaunchKernelGGL(k_probe_record_packed<double>, dim3(n_jobs), dim3(256), 0, hs(s), jobs);
    else
        hipLaunchKernelGGL(k_probe_record_packed<float>, dim3(n_jobs), dim3(256), 0, hs(s), jobs);
    return launched();
}

int rf_launch_probe_dft_packed(const rf_pdft_job_t *jobs, int n_jobs, int max_freq, int count, rf_stream_t *s) {
    if (n_jobs <= 0 || max_freq <= 0 || count <= 0) return 0;
    if (!jobs || n_jobs > 65535) return -1;
    hipLaunchKernelGGL(k_probe_dft_packed, dim3((max_freq + 255) / 256, n_jobs), dim3(256), 0, hs(s), jobs, count);
    return launched();
}

int rf_launch_plane_captures(int prec, const rf_geom_t *g, const rf_caps_t *cs, int ring_len,
                             const long long *d_step, rf_stream_t *s) {
    if (cs->n <= 0) return 0;
    if (cs->n > RF_MAX_CAPS) return -1;
    rf_caps_t c = *cs;
    long long total = 0;
    for (int q = 0; q < c.n; q++) {
        if (!c.c[q].f || !c.c[q].buf || c.c[q].na <= 0 || c.c[q].nb <= 0 || c.c[q].axis < 0 || c.c[q].axis > 2)
            return -1;
        c.c[q].start = total;
        total += (long long)c.c[q].na * c.c[q].nb;
    }
    if (total >= (1LL << 31)) return -1;
    const dim3 grd(static_cast<unsigned>((total + 255) / 256));
    if (prec == 0)
        hipLaunchKernelGGL((k_plane_captures<double>), grd, dim3(256), 0, hs(s), g[0], c, ring_len, d_step);
    else
        hipLaunchKernelGGL((k_plane_captures<float>), grd, dim3(256), 0, hs(s), g[0], c, ring_len, d_step);
    return launched();
}

int rf_launch_plane_dft(const void *buf, int buf_prec, long long cells, int ring_len, int nf,
                        const double *freq_w, double dt, const long long *d_step, int count,
                        double *tw_scratch, double *acc_re, double *acc_im, double *cmp_re,
                        double *cmp_im, int comp_sum, rf_stream_t *s) {
    if (count <= 0) return 0;
    if (count > RF_PDFT_TWS || nf <= 0) return -1;
    double2 *tw = reinterpret_cast<double2 *>(tw_scratch);
    hipLaunchKernelGGL(k_twiddles, dim3((nf * count + 255) / 256), dim3(256), 0, hs(s), nf, freq_w,
                       dt, d_step, count, tw);
    (void)cmp_re;
    (void)cmp_im;
    if (comp_sum) return -1;   /* plane monitors use plain double sums */
    dim3 grd((unsigned)((cells + PD_CELLS - 1) / PD_CELLS));
    // samples per pass: 64 (float) or 32 (double) staged in LDS
    const int pass = buf_prec == 0 ? 32 : 64;
    for (int off = 0; off < count; off += pass) {
        const int cnt = count - off < pass ? count - off : pass;
        if (buf_prec == 0)
            hipLaunchKernelGGL((k_plane_dft<double, 32>), grd, dim3(256), 0, hs(s), static_cast<const double *>(buf),
                               cells, ring_len, nf, tw, d_step, count, off, cnt, acc_re, acc_im);
        else
            hipLaunchKernelGGL((k_plane_dft<float, 64>), grd, dim3(256), 0, hs(s), static_cast<const float *>(buf),
                               cells, ring_len, nf, tw, d_step, count, off, cnt, acc_re, acc_im);
        if (launched()) return -1;
    }
    return 0;
}

int rf_launch_energy(int prec, const rf_geom_t *g, void *const f[6], const void *tag, int tag_bytes,
                     const double *eps_tab, double vac_eps, double mu0, double *partials,
                     int n_partials, double *out, rf_stream_t *s) {
    if (prec == 0) {
        double *F[6];
        for (int q = 0; q < 6; q++) F[q] = static_cast<double *>(f[q]);
        if (!tag) return launch_energy_t<double, NoTag>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
        if (tag_bytes == 1) return launch_energy_t<double, uint8_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
        if (tag_bytes == 2) return launch_energy_t<double, uint16_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
        return launch_energy_t<double, uint32_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
    }
    float *F[6];
    for (int q = 0; q < 6; q++) F[q] = static_cast<float *>(f[q]);
    if (!tag) return launch_energy_t<float, NoTag>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
    if (tag_bytes == 1) return launch_energy_t<float, uint8_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
    if (tag_bytes == 2) return launch_energy_t<float, uint16_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
    return launch_energy_t<float, uint32_t>(g, F, tag, eps_tab, vac_eps, mu0, partials, n_partials, out, hs(s));
}

int rf_launch_audit(int prec, const rf_geom_t *g, void *const fn[6], void *const fp[6], const void *tag,
                    int tag_bytes, const double *eps_tab, double vac_eps, double mu0, double *partials,
                    int n_partials, double *out, rf_stream_t *s) {
    if (n_partials < 1) return -1;
    return prec == 0 ? launch_audit_p<double>(g, fn, fp, tag, tag_bytes, eps_tab, vac_eps, mu0, partials, n_partials,
                                              out, hs(s))
                     : launch_audit_p<float>(g, fn, fp, tag, tag_bytes, eps_tab, vac_eps, mu0, partials, n_partials,
                                             out, hs(s));
}

int rf_launch_flux(int prec, const rf_geom_t *g, int axis, int index, void *const fn[6], void *const fp[6],
                   double *partials, int n_partials, double *acc, rf_stream_t *s) {
    if (axis < 0 || axis > 2 || index < 0 || index >= g->n[axis] || n_partials < 1) return -1;
    return prec == 0 ? launch_flux_t<double>(g, axis, index, fn, fp, partials, n_partials, acc, hs(s))
                     : launch_flux_t<float>(g, axis, index, fn, fp, partials, n_partials, acc, hs(s));
}

int rf_launch_nonfinite(int prec, const void *v, long long n, unsigned *flag, rf_stream_t *s) {
    if (n <= 0) return 0;
    if (!v || !flag) return -1;
    const unsigned blocks = (unsigned)((n + 255) / 256 > 8192 ? 8192 : (n + 255) / 256);
    if (prec == 0)
        hipLaunchKernelGGL((k_nonfinite<double>), dim3(blocks), dim3(256), 0, hs(s), static_cast<const double *>(v),
                           n, flag);
    else
        hipLaunchKernelGGL((k_nonfinite<float>), dim3(blocks), dim3(256), 0, hs(s), static_cast<const float *>(v), n,
                           flag);
    return launched();
}

int rf_launch_convert(void *dst, int dst_prec, const void *src, int src_prec, long long n,
                      rf_stream_t *s) {
    if (n <= 0) return 0;
    const unsigned blocks = (unsigned)((n + 255) / 256 > 65536 ? 65536 : (n + 255) / 256);
    if (dst_prec == 1 && src_prec == 0)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks), dim3(256), 0, hs(s),
                           static_cast<float *>(dst), static_cast<const double *>(src), n);
    else if (dst_prec == 0 && src_prec == 1)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks), dim3(256), 0, hs(s),
                           static_cast<double *>(dst), static_cast<const float *>(src), n);
    else
        return -1;
    return launched();
}

}  // extern "C"
