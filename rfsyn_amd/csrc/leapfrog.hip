// Fused leapfrog Yee step for gfx950: H^{n+1/2} then E^{n+1} in one pass.
//
// A block owns a tile of TJ rows (y) x 64 lanes (z, the contiguous axis) of
// node columns and marches along x.  At each plane i it
//   1. stages E^n(i) for the tile plus a one-cell frame in LDS (the tile's
//      own values come from the E^n(i+1) registers of the previous plane,
//      so each E value crosses HBM once),
//   2. computes H^{n+1/2}(i) on the tile and on the one-cell halo row
//      (j0-1) and column (k0-1) that the E update needs, keeping the result
//      in LDS and writing only the tile's own nodes to the B buffers,
//   3. computes E^{n+1}(i) on the tile from the LDS H plane and the H of
//      plane i-1 held in registers, and writes it to B.
// Reading state n from the A buffers and writing state n+1 to B makes the
// halo recomputation race-free; per node per step the kernel moves 6 reads
// and 6 writes instead of the 9 + 9 of separate H and E sweeps.
//
// The arithmetic of every update is the same expression, in the same order,
// as the two-kernel path (yee_kernels.hip), which follows
// engine/fdtd/grid.c:70-89, cpml.h:49-50, cpml.c:185-260 and
// materials.c:975-1055; compiled with -ffp-contract=off the FP64 variant
// reproduces the reference bits.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rf_device.h"
#include "rf_hip.h"
#include "yee_device.h"

namespace {

using namespace rfk;

constexpr int TK = 64;   // lanes along z

// Wrap a coordinate one step outside [0, n) on a periodic axis; -1 when the
// coordinate is outside the stored box of a non-periodic axis ([0, n]).
__device__ __forceinline__ int wrap_or_invalid(int q, int n, bool periodic) {
    if (periodic) {
        if (q < 0) return q + n;
        if (q >= n) return q - n;
        return q;
    }
    return (q < 0 || q > n) ? -1 : q;
}

// cpml_term with separate psi read/write slabs (write == nullptr: halo
// recomputation, nothing stored).
template <typename T>
__device__ __forceinline__ T stretch_rw(T diff, const rf_cpml_t &c, const Geo &g, bool h, int axis, int q,
                                        const T *psi_rd, T *psi_wr, int i, int j, int k, T dx) {
    const T *kinv = static_cast<const T *>(h ? c.kinv_h[axis] : c.kinv_m[axis]);
    const T *bb = static_cast<const T *>(h ? c.b_h[axis] : c.b_m[axis]);
    const T *aa = static_cast<const T *>(h ? c.a_h[axis] : c.a_m[axis]);
    const int cq = slab_coord(c, axis, q);
    T psi = T(0);
    if (cq >= 0) {
        const long long s = slab_index(c, g, axis, cq, i, j, k);
        psi = bb[q] * psi_rd[s] + aa[q] * diff;
        if (psi_wr) psi_wr[s] = psi;
    }
    return diff * kinv[q] + dx * psi;
}

// Applies the pre-E point and plane sources of this (node, component), in
// insertion order, to the old value (source_apply before the E update).
template <typename T>
__device__ __forceinline__ T pre_sources(const rf_leap_t &a, T v, long long idx, int c, int i, int j, int k,
                                         const Geo &g, long long step) {
    for (int s = 0; s < a.n_pts; s++) {
        const rf_point_src_t &p = a.pts[s];
        if (p.idx == idx && p.comp == c) {
            const double w = p.wave[step];
            v = p.hard ? static_cast<T>(w) : static_cast<T>(static_cast<double>(v) + w);
        }
    }
    for (int s = 0; s < a.n_planes; s++) {
        if (a.plane_after[s]) continue;
        const rf_plane_src_t &p = a.planes[s];
        if (p.comp != c) continue;
        const int x[3] = {i, j, k};
        const int n[3] = {g.nx, g.ny, g.nz};
        const int ax = p.axis, b0 = (ax + 1) % 3, b1 = (ax + 2) % 3;
        int pi = p.index;
        if ((a.g.per >> ax) & 1) pi %= n[ax];
        if (x[ax] == pi && x[b0] < n[b0] && x[b1] < n[b1]) {
            const double w = p.wave[step];
            v = p.hard ? static_cast<T>(w) : static_cast<T>(static_cast<double>(v) + w);
        }
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T post_sources(const rf_leap_t &a, T v, int c, int i, int j, int k, const Geo &g,
                                          long long step) {
    for (int s = 0; s < a.n_planes; s++) {
        if (!a.plane_after[s]) continue;
        const rf_plane_src_t &p = a.planes[s];
        if (p.comp != c) continue;
        const int x[3] = {i, j, k};
        const int n[3] = {g.nx, g.ny, g.nz};
        const int ax = p.axis, b0 = (ax + 1) % 3, b1 = (ax + 2) % 3;
        int pi = p.index;
        if ((a.g.per >> ax) & 1) pi %= n[ax];
        if (x[ax] == pi && x[b0] < n[b0] && x[b1] < n[b1]) {
            const double w = p.wave[step];
            v = p.hard ? static_cast<T>(w) : static_cast<T>(static_cast<double>(v) + w);
        }
    }
    return v;
}

// Material update of one E component (same arithmetic as e_component of the
// two-kernel path); returns the new value, or the old one for tensor edges.
template <typename T, typename TagT>
__device__ __forceinline__ T e_new_value(int c, T e_old, T dA, T dB, const rf_mat_t &m, unsigned tag, int i,
                                         int j, int k) {
    using Entry = typename EntryOf<T>::type;
    using Ade = typename EntryOf<T>::ade;
    if constexpr (std::is_same<TagT, NoTag>::value) {
        return T(1) * e_old + static_cast<T>(m.vac_ca) * (dA - dB);
    } else {
        const Entry &en = static_cast<const Entry *>(m.entries)[tag];
        const uint32_t fl = en.flags[c];
        if (fl & RF_EDGE_TENSOR) return e_old;
        if (fl & RF_EDGE_ADE) {
            const Ade &ad = static_cast<const Ade *>(m.ade)[en.ade[c]];
            const long long cells = (long long)m.bn[0] * m.bn[1] * m.bn[2];
            const long long cell =
                ((long long)(i - m.b0[0]) * m.bn[1] + (j - m.b0[1])) * m.bn[2] + (k - m.b0[2]);
            return ade_edge<T>(ad, en.Ca[c], en.Cb[c], e_old, dA, dB, static_cast<T *>(m.J),
                               static_cast<T *>(m.P), 3 * cells, (long long)c * cells + cell);
        }
        return en.Ca[c] * e_old + en.Cb[c] * (dA - dB);
    }
}

template <typename T, typename TagT, bool CPML, int TJ>
__global__ void __launch_bounds__(TK *(TJ + 2))
k_leapfrog(rf_leap_t a) {
    __shared__ T sE[3][TJ + 2][TK + 2];
    __shared__ T sH[3][TJ + 1][TK + 1];

    const Geo g = geo_of(a.g);
    const int lane = threadIdx.x;
    const int w = threadIdx.y;
    const int j0 = blockIdx.y * TJ;
    const int k0 = blockIdx.x * TK;
    const int jmax = g.py ? g.ny - 1 : g.ny;
    const int kmax = g.pz ? g.nz - 1 : g.nz;
    const int imax = g.px ? g.nx - 1 : g.nx;

    const T *EA[3] = {static_cast<const T *>(a.ea[0]), static_cast<const T *>(a.ea[1]),
                      static_cast<const T *>(a.ea[2])};
    const T *HA[3] = {static_cast<const T *>(a.ha[0]), static_cast<const T *>(a.ha[1]),
                      static_cast<const T *>(a.ha[2])};
    T *EB[3] = {static_cast<T *>(a.eb[0]), static_cast<T *>(a.eb[1]), static_cast<T *>(a.eb[2])};
    T *HB[3] = {static_cast<T *>(a.hb[0]), static_cast<T *>(a.hb[1]), static_cast<T *>(a.hb[2])};
    const T cb = static_cast<T>(a.cb), dx = static_cast<T>(a.dx);

    // ---- role of this thread -------------------------------------------
    // position: a node whose H (and, when owned, E) this thread computes;
    // (r, c): its cell in the LDS planes (row 0 = j0-1, col 0 = k0-1).
    int pj = -1, pk = -1, r = 0, c = 0;
    bool pos = false, owned = false;
    // a second LDS cell this thread refreshes from E^n(i) every plane
    int xr = -1, xc = -1, xj = -1, xk = -1;
    if (w < TJ) {
        pj = j0 + w;
        pk = k0 + lane;
        r = w + 1;
        c = lane + 1;
        owned = pos = pj <= jmax && pk <= kmax;
        if (!pos) {   // a periodic wrap cell just past the tile end still feeds neighbours
            const int wj = wrap_or_invalid(pj, g.ny, g.py), wk = wrap_or_invalid(pk, g.nz, g.pz);
            if (wj >= 0 && wk >= 0 && (pj <= jmax + 1) && (pk <= kmax + 1)) {
                xr = r; xc = c; xj = wj; xk = wk;
            }
        }
    } else if (w == TJ) {
        pj = j0 - 1;
        if (pj < 0) pj = g.py ? g.ny - 1 : -1;
        pk = k0 + lane;
        r = 0;
        c = lane + 1;
        pos = pj >= 0 && pk <= kmax;
        // the frame row below the tile: (j0+TJ, k0+lane)
        const int wj = wrap_or_invalid(j0 + TJ, g.ny, g.py), wk = wrap_or_invalid(k0 + lane, g.nz, g.pz);
        if (wj >= 0 && wk >= 0 && j0 + TJ <= jmax + 1 && k0 + lane <= kmax + 1) {
            xr = TJ + 1; xc = lane + 1; xj = wj; xk = wk;
        }
    } else {
        if (lane >= 1 && lane <= TJ) {         // halo column positions (j0+lane-1, k0-1)
            pj = j0 + lane - 1;
            pk = k0 - 1;
            if (pk < 0) pk = g.pz ? g.nz - 1 : -1;
            r = lane;
            c = 0;
            pos = pj <= jmax && pk >= 0;
        } else if (lane == 0 || lane == TJ + 1) {   // frame corners of column k0-1
            const int wj = wrap_or_invalid(j0 - 1 + lane, g.ny, g.py), wk = wrap_or_invalid(k0 - 1, g.nz, g.pz);
            if (wj >= 0 && wk >= 0) { xr = lane; xc = 0; xj = wj; xk = wk; }
        } else if (lane >= 32 && lane <= 33 + TJ) {   // frame column k0+64
            const int rr = lane - 32;
            const int wj = wrap_or_invalid(j0 - 1 + rr, g.ny, g.py), wk = wrap_or_invalid(k0 + TK, g.nz, g.pz);
            if (wj >= 0 && wk >= 0 && k0 + TK <= kmax + 1) { xr = rr; xc = TK + 1; xj = wj; xk = wk; }
        }
    }

    // ---- x range of this block ------------------------------------------
    const int chunk = a.chunk > 0 ? a.chunk : imax + 1;
    const int i_begin = blockIdx.z * chunk;
    if (i_begin > imax) return;
    const int i_end = min(i_begin + chunk, imax + 1);
    // E^{n+1}(i_begin) needs H^{n+1/2}(i_begin - 1): compute it first
    // without storing anything (plane nx-1 when x is periodic).
    const bool warm = i_begin > 0 || g.px;
    int i = warm ? (i_begin == 0 ? g.nx - 1 : i_begin - 1) : i_begin;
    const int iters = (i_end - i_begin) + (warm ? 1 : 0);

    const long long step = *a.d_step;
    unsigned tag = 0;
    T ecur[3] = {T(0), T(0), T(0)}, enx[3];
    T hp_y = T(0), hp_z = T(0);
    if (pos) {
        const long long idx = (long long)i * g.sx + (long long)pj * g.sy + pk;
        for (int q = 0; q < 3; q++) ecur[q] = EA[q][idx];
    }

    for (int it = 0; it < iters; it++) {
        const bool first = warm && it == 0;
        int ip = i + 1;
        bool has_ip = true;
        if (ip > g.nx || (g.px && ip == g.nx)) {
            if (g.px) ip = 0; else has_ip = false;
        }
        // 1. stage E^n(i) and fetch E^n(i+1) for the next plane
        if (pos) {
            for (int q = 0; q < 3; q++) sE[q][r][c] = ecur[q];
            if (has_ip) {
                const long long idxn = (long long)ip * g.sx + (long long)pj * g.sy + pk;
                for (int q = 0; q < 3; q++) enx[q] = EA[q][idxn];
            } else {
                for (int q = 0; q < 3; q++) enx[q] = T(0);
            }
        }
        if (xr >= 0) {
            const long long idx = (long long)i * g.sx + (long long)xj * g.sy + xk;
            for (int q = 0; q < 3; q++) sE[q][xr][xc] = EA[q][idx];
        }
        __syncthreads();

        // 2. H^{n+1/2}(i) at this position
        if (pos) {
            const long long idx = (long long)i * g.sx + (long long)pj * g.sy + pk;
            const bool do_x = (i < g.nx + !g.px) && pj < g.ny && pk < g.nz;
            const bool do_y = i < g.nx && (pj < g.ny + !g.py) && pk < g.nz;
            const bool do_z = i < g.nx && pj < g.ny && (pk < g.nz + !g.pz);
            const bool store = owned && !first;
            // halo nodes compute only what the tile's E update reads: Hx
            // on both, Hz on the halo row (Ex), Hy on the halo column (Ex)
            const bool need_y = owned || (w == TJ + 1);
            const bool need_z = owned || (w == TJ);
            T hn[3] = {T(0), T(0), T(0)};
            {
                T h = HA[0][idx];
                if (do_x) {
                    T dA = sE[2][r + 1][c] - sE[2][r][c];
                    T dB = sE[1][r][c + 1] - sE[1][r][c];
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, true, 1, pj, static_cast<const T *>(a.cp.psi_h[0][0]),
                                           store ? static_cast<T *>(a.psi_hb[0][0]) : nullptr, i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, true, 2, pk, static_cast<const T *>(a.cp.psi_h[0][1]),
                                           store ? static_cast<T *>(a.psi_hb[0][1]) : nullptr, i, pj, pk, dx);
                    }
                    h -= cb * (dA - dB);
                }
                hn[0] = h;
            }
            if (need_y) {
                T h = HA[1][idx];
                if (do_y) {
                    T dA = sE[0][r][c + 1] - sE[0][r][c];
                    T dB = enx[2] - sE[2][r][c];
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, true, 2, pk, static_cast<const T *>(a.cp.psi_h[1][0]),
                                           store ? static_cast<T *>(a.psi_hb[1][0]) : nullptr, i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, true, 0, i, static_cast<const T *>(a.cp.psi_h[1][1]),
                                           store ? static_cast<T *>(a.psi_hb[1][1]) : nullptr, i, pj, pk, dx);
                    }
                    h -= cb * (dA - dB);
                }
                hn[1] = h;
            }
            if (need_z) {
                T h = HA[2][idx];
                if (do_z) {
                    T dA = enx[1] - sE[1][r][c];
                    T dB = sE[0][r + 1][c] - sE[0][r][c];
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, true, 0, i, static_cast<const T *>(a.cp.psi_h[2][0]),
                                           store ? static_cast<T *>(a.psi_hb[2][0]) : nullptr, i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, true, 1, pj, static_cast<const T *>(a.cp.psi_h[2][1]),
                                           store ? static_cast<T *>(a.psi_hb[2][1]) : nullptr, i, pj, pk, dx);
                    }
                    h -= cb * (dA - dB);
                }
                hn[2] = h;
            }
            if (store)
                for (int q = 0; q < 3; q++) HB[q][idx] = hn[q];
            const int hr = r, hc = c;   // H plane uses the same (row, col) frame origin
            sH[0][hr][hc] = hn[0];
            if (need_y) sH[1][hr][hc] = hn[1];
            if (need_z) sH[2][hr][hc] = hn[2];
        }
        __syncthreads();

        // 3. E^{n+1}(i) on the tile
        if (owned) {
            const T hx_ = sH[0][r][c], hy_ = sH[1][r][c], hz_ = sH[2][r][c];
            if (!first) {
                const long long idx = (long long)i * g.sx + (long long)pj * g.sy + pk;
                if constexpr (!std::is_same<TagT, NoTag>::value) tag = load_tag<TagT>(a.mat.tag, idx);
                const bool in_x = i < g.nx, in_y = pj < g.ny, in_z = pk < g.nz;
                const bool i_ok = g.px || i >= 1, j_ok = g.py || pj >= 1, k_ok = g.pz || pk >= 1;
                const bool do_x = in_x && in_y && in_z && j_ok && k_ok;
                const bool do_y = in_x && in_y && in_z && i_ok && k_ok;
                const bool do_z = in_x && in_y && in_z && i_ok && j_ok;
                T en[3];
                for (int q = 0; q < 3; q++) en[q] = pre_sources<T>(a, ecur[q], idx, q, i, pj, pk, g, step);
                if (do_x) {   // materials.c:975-1000
                    T dA = hz_ - sH[2][r - 1][c];
                    T dB = hy_ - sH[1][r][c - 1];
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, false, 1, pj, static_cast<const T *>(a.cp.psi_e[0][0]),
                                           static_cast<T *>(a.cp.psi_e[0][0]), i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, false, 2, pk, static_cast<const T *>(a.cp.psi_e[0][1]),
                                           static_cast<T *>(a.cp.psi_e[0][1]), i, pj, pk, dx);
                    }
                    en[0] = e_new_value<T, TagT>(0, en[0], dA, dB, a.mat, tag, i, pj, pk);
                }
                if (do_y) {   // materials.c:1015-1040
                    T dA = hx_ - sH[0][r][c - 1];
                    T dB = hz_ - hp_z;
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, false, 2, pk, static_cast<const T *>(a.cp.psi_e[1][0]),
                                           static_cast<T *>(a.cp.psi_e[1][0]), i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, false, 0, i, static_cast<const T *>(a.cp.psi_e[1][1]),
                                           static_cast<T *>(a.cp.psi_e[1][1]), i, pj, pk, dx);
                    }
                    en[1] = e_new_value<T, TagT>(1, en[1], dA, dB, a.mat, tag, i, pj, pk);
                }
                if (do_z) {   // materials.c:1050-1058
                    T dA = hy_ - hp_y;
                    T dB = hx_ - sH[0][r - 1][c];
                    if (CPML) {
                        dA = stretch_rw<T>(dA, a.cp, g, false, 0, i, static_cast<const T *>(a.cp.psi_e[2][0]),
                                           static_cast<T *>(a.cp.psi_e[2][0]), i, pj, pk, dx);
                        dB = stretch_rw<T>(dB, a.cp, g, false, 1, pj, static_cast<const T *>(a.cp.psi_e[2][1]),
                                           static_cast<T *>(a.cp.psi_e[2][1]), i, pj, pk, dx);
                    }
                    en[2] = e_new_value<T, TagT>(2, en[2], dA, dB, a.mat, tag, i, pj, pk);
                }
                for (int q = 0; q < 3; q++) EB[q][idx] = post_sources<T>(a, en[q], q, i, pj, pk, g, step);
            }
            hp_y = hy_;
            hp_z = hz_;
        }
        for (int q = 0; q < 3; q++) ecur[q] = enx[q];
        i = first ? i_begin : i + 1;
        __syncthreads();
    }
}

template <typename T, typename TagT, int TJ>
int launch_tag(int has_cpml, const rf_leap_t *a, hipStream_t st) {
    const int imax = (a->g.per & 1) ? a->g.n[0] - 1 : a->g.n[0];
    const int chunk = a->chunk > 0 ? a->chunk : imax + 1;
    const int jn = (a->g.per & 2) ? a->g.n[1] : a->g.n[1] + 1;
    const int kn = (a->g.per & 4) ? a->g.n[2] : a->g.n[2] + 1;
    dim3 grd((kn + TK - 1) / TK, (jn + TJ - 1) / TJ, (imax + 1 + chunk - 1) / chunk);
    dim3 blk(TK, TJ + 2, 1);
    if (has_cpml)
        hipLaunchKernelGGL((k_leapfrog<T, TagT, true, TJ>), grd, blk, 0, st, *a);
    else
        hipLaunchKernelGGL((k_leapfrog<T, TagT, false, TJ>), grd, blk, 0, st, *a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, int TJ>
int launch_prec(int has_cpml, const rf_leap_t *a, hipStream_t st) {
    if (!a->mat.tag) return launch_tag<T, NoTag, TJ>(has_cpml, a, st);
    switch (a->mat.tag_bytes) {
    case 1: return launch_tag<T, uint8_t, TJ>(has_cpml, a, st);
    case 2: return launch_tag<T, uint16_t, TJ>(has_cpml, a, st);
    case 4: return launch_tag<T, uint32_t, TJ>(has_cpml, a, st);
    }
    return -1;
}

}  // namespace

extern "C" int rf_launch_leapfrog(int prec, int has_cpml, const rf_leap_t *a, rf_stream_t *s) {
    if (a->n_pts > RF_LEAP_MAX_PTS || a->n_planes > RF_LEAP_MAX_PLANES) return -1;
    hipStream_t st = rf_raw_stream(s);
    return prec == 0 ? launch_prec<double, 8>(has_cpml, a, st) : launch_prec<float, 8>(has_cpml, a, st);
}
