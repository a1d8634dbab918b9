// Fused leapfrog Yee step for gfx950: H^{n+1/2} then E^{n+1} in one pass.
//
// One wavefront owns a tile of R rows (y) by 62 columns (z, the contiguous
// axis) of nodes and marches along x through a chunk of planes.  Lane l
// holds column k0-1+l: lanes 1..62 own their column, lanes 0 and 63 are
// halo columns whose values reach the k-1 / k+1 neighbours through DPP
// wavefront shifts.  Registers hold E^n on rows j0-1 .. j0+R of the current
// and the next plane and the (hy, hz) of the previous plane, so a field value
// crosses HBM once per step in each direction: 6 reads and 6 writes per node
// instead of the 9 + 9 of separate H and E sweeps.  H^{n+1/2} is computed on
// rows j0-1 .. j0+R-1 (row j0-1 and lane 0 are the halo that the E update of
// the tile reads) and E^{n+1} on rows j0 .. j0+R-1.
//
// State n is read from the A buffers and state n+1 written to the B buffers,
// so recomputed halos never see a neighbour's new values and no wavefront
// waits for another: the only barrier is the one after the block stages the
// material table in LDS.  psi_h is double-buffered for the same reason;
// psi_e and the ADE state are touched only by the owner of the node and are
// updated in place.  Blocks are numbered XCD-major so that the tiles one XCD
// runs at a time are neighbours and share halo rows in its L2.
//
// The kernel owns the box of nodes that lies inside the CPML plain range on
// every axis (kinv = 1 and psi stays 0 there, so the stretched difference
// equals the plain one up to the sign of a zero) with a one-node margin, so
// it evaluates no CPML term; the shell kernels at the end of this file
// update the nodes around that box.  The common path is branch-free:
// sources are applied only on planes whose tile their bounding box
// touches, dispersive and tensor edges only where the material flags say so.
//
// Every update is the same expression, in the same order, as the two-kernel
// path (yee_kernels.hip), which follows engine/fdtd/grid.c:70-89,
// cpml.h:49-50, cpml.c:185-260 and materials.c:975-1055; compiled with
// -ffp-contract=off the FP64 variant reproduces the reference values.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rf_device.h"
#include "rf_hip.h"
#include "yee_device.h"

namespace {

using namespace rfk;

constexpr int WL = 64;       // lanes per wavefront
constexpr int KW = WL - 2;   // columns owned per wavefront
static_assert(KW == RF_LEAP_COLS, "the host plans z tiles of RF_LEAP_COLS columns");
constexpr int WPB = 4;       // wavefronts per block (independent mode)
constexpr int WPB_MAX = 4;   // wavefronts per block (cooperative mode)
constexpr int N_XCD = RF_XCDS;   // blocks are dealt round-robin to the XCDs

// DPP wavefront shifts: lane l receives lane l+1 (from_upper) or lane l-1
// (from_lower); the edge lane receives 0 and is a halo lane whose result is
// never stored.
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }

__device__ __forceinline__ float from_upper(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
__device__ __forceinline__ float from_lower(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ double from_upper(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(dpp_shl1(static_cast<int>(b)));
    const unsigned hi = static_cast<unsigned>(dpp_shl1(static_cast<int>(b >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ double from_lower(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(dpp_shr1(static_cast<int>(b)));
    const unsigned hi = static_cast<unsigned>(dpp_shr1(static_cast<int>(b >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Buffer loads and stores: one resource per state block (the six field
// arrays of a state, contiguous) with 32-bit byte offsets, so a lane keeps
// its column offset in one register and the row, plane and component
// offsets are scalar; no 64-bit address arithmetic per access.
using rsrc_t = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, static_cast<int>(bytes), 0x00020000);
}
template <typename T> __device__ __forceinline__ T bld(rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ float bld<float>(rsrc_t r, unsigned vo, unsigned so) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(vo), static_cast<int>(so), 0));
}
template <> __device__ __forceinline__ double bld<double>(rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(vo),
                                                                           static_cast<int>(so), 0));
}
__device__ __forceinline__ void bst(float v, rsrc_t r, unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, static_cast<int>(vo), static_cast<int>(so), 0);
}
__device__ __forceinline__ void bst(double v, rsrc_t r, unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, static_cast<int>(vo),
                                          static_cast<int>(so), 0);
}
template <typename TagT> __device__ __forceinline__ unsigned btag(rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ unsigned btag<uint8_t>(rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, static_cast<int>(vo), static_cast<int>(so), 0);
}
template <> __device__ __forceinline__ unsigned btag<uint16_t>(rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, static_cast<int>(vo), static_cast<int>(so), 0);
}
template <> __device__ __forceinline__ unsigned btag<uint32_t>(rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(vo), static_cast<int>(so), 0);
}

// The cold data of the step, read through a pointer the compiler cannot
// see through, so its loads stay inside the slab / flagged-edge branches
// that use them instead of being hoisted into registers for the whole loop.
__device__ __forceinline__ const rf_leap_cold_t *cold_view(const rf_leap_cold_t *p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Sources of the device table that hit node (i,j,k), component c, phase
// `after`, in table order (point sources, then plane sources, as the
// separate source kernels apply them).
template <typename T>
__device__ __forceinline__ T apply_sources(const rf_leap_src_t *tab_p, int n, T v, int c, int i, int j, int k,
                                        int gnx, int gny, int gnz, long long step, int after) {
    // the table and the waveforms are device memory: global accesses, not
    // FLAT ones (the table pointer may come from a device job table)
    typedef const __attribute__((address_space(1))) rf_leap_src_t gsrc_t;
    gsrc_t *tab = (gsrc_t *)tab_p;
    for (int s = 0; s < n; s++) {
        gsrc_t &p = tab[s];
        if (p.after != after || p.comp != c) continue;
        bool hit;
        if (p.plane) {
            const int ax = p.axis;
            const int x = ax == 0 ? i : (ax == 1 ? j : k);
            const bool in_x = ax == 0 || i < gnx, in_y = ax == 1 || j < gny, in_z = ax == 2 || k < gnz;
            hit = x == p.index && in_x && in_y && in_z;
        } else {
            hit = p.i == i && p.j == j && p.k == k;
        }
        if (hit) {
            const double w = ((const __attribute__((address_space(1))) double *)p.wave)[step];
            v = p.hard ? static_cast<T>(w) : static_cast<T>(static_cast<double>(v) + w);
        }
    }
    return v;
}

// Whether any source of the table sits on node (i, j, k), whatever its
// component and phase (apply_sources leaves every value of a node it
// returns false for unchanged).
__device__ __forceinline__ bool sources_touch(const rf_leap_src_t *tab_p, int n, int i, int j, int k, int gnx,
                                              int gny, int gnz) {
    typedef const __attribute__((address_space(1))) rf_leap_src_t gsrc_t;
    gsrc_t *tab = (gsrc_t *)tab_p;
    bool hit = false;
    for (int s = 0; s < n; s++) {
        // both tests without branches, so a source's fields are read in one
        // batch of loads
        gsrc_t &p = tab[s];
        const int plane = p.plane, ax = p.axis, index = p.index, pi = p.i, pj = p.j, pk = p.k;
        const int x = ax == 0 ? i : (ax == 1 ? j : k);
        const bool in_x = ax == 0 || i < gnx, in_y = ax == 1 || j < gny, in_z = ax == 2 || k < gnz;
        const bool on_plane = x == index && in_x && in_y && in_z;
        const bool on_point = pi == i && pj == j && pk == k;
        hit = hit || (plane ? on_plane : on_point);
    }
    return hit;
}

// A source of the table with its value at the current step, staged in
// LDS at the start of the kernel.
struct LSrc {
    double w;
    int plane, comp, hard, after, axis, index, i, j, k;
};
constexpr int MAX_LSRC = 16;

template <typename T>
__device__ __forceinline__ T apply_lsrc(const LSrc *tab, int n, T v, int c, int i, int j, int k, int gnx, int gny,
                                        int gnz, int after) {
    for (int s = 0; s < n; s++) {
        const LSrc &p = tab[s];
        if (p.after != after || p.comp != c) continue;
        bool hit;
        if (p.plane) {
            const int ax = p.axis;
            const int x = ax == 0 ? i : (ax == 1 ? j : k);
            const bool in_x = ax == 0 || i < gnx, in_y = ax == 1 || j < gny, in_z = ax == 2 || k < gnz;
            hit = x == p.index && in_x && in_y && in_z;
        } else {
            hit = p.i == i && p.j == j && p.k == k;
        }
        if (hit) v = p.hard ? static_cast<T>(p.w) : static_cast<T>(static_cast<double>(v) + p.w);
    }
    return v;
}

// Coefficients of one material entry: Ca, Cb per component and the flags
// of the three components packed one byte each.  Entries below `ncache`
// come from the block's LDS copy of the table, the rest from HBM.
template <typename T>
struct Coefs {
    T ca[3], cb[3];
    uint32_t fl;
};

// When every active lane's entry is cached (a wavefront-uniform test) the
// reads are LDS reads only: one merged path for both sources compiles to
// FLAT loads, which count against the vector memory counter too, so the
// wait for the coefficients would also wait for every field load the
// wavefront has in flight.
template <typename T>
__device__ __forceinline__ Coefs<T> entry_coefs(const void *entries, const T *s_ca, const T *s_cb,
                                                const uint32_t *s_fl, int ncache, unsigned tag) {
    Coefs<T> o;
    if (__ballot(static_cast<int>(tag) >= ncache) == 0) {
        typedef __attribute__((address_space(3))) const T lds_t;
        typedef __attribute__((address_space(3))) const uint32_t lds_u;
        o.fl = ((lds_u *)s_fl)[tag];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o.ca[c] = ((lds_t *)s_ca)[3 * tag + c];
            o.cb[c] = ((lds_t *)s_cb)[3 * tag + c];
        }
        return o;
    }
    if (static_cast<int>(tag) < ncache) {
        o.fl = s_fl[tag];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o.ca[c] = s_ca[3 * tag + c];
            o.cb[c] = s_cb[3 * tag + c];
        }
    } else {
        const typename EntryOf<T>::type &en = static_cast<const typename EntryOf<T>::type *>(entries)[tag];
        o.fl = en.flags[0] | (en.flags[1] << 8) | (en.flags[2] << 16);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o.ca[c] = en.Ca[c];
            o.cb[c] = en.Cb[c];
        }
    }
    return o;
}

// Trapezoidal ADE update of one edge with at most two poles whose state
// (j[s], p[s]) is already in registers: the arithmetic, in order, of
// ade_edge (yee_device.h).  Returns E^{n+1} and the new state.
template <typename T, typename Ade>
__device__ __forceinline__ T ade2(const Ade &ad, T ca, T cb, T e_old, T dA, T dB, const T j[2], const T p[2],
                                  T nj[2], T np[2]) {
    T u[2];
    T sum_u = T(0);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s < ad.n_poles) {
            u[s] = ad.has_p[s] ? ad.kj[s] * j[s] - ad.kp[s] * p[s] : ad.kj[s] * j[s];
            sum_u += u[s];
        }
    }
    const T e_new = ca * e_old + cb * (dA - dB) - ad.Cu * sum_u;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s < ad.n_poles) {
            const T S = u[s] + ad.v[s] * (e_new + e_old);
            nj[s] = S - j[s];
            np[s] = ad.has_p[s] ? p[s] + ad.hdt * S : p[s];
        }
    }
    return e_new;
}

// The three components of a node whose entry has flags: tensor edges keep
// the old value (the tensor kernel updates them), dispersive edges take
// the trapezoidal ADE step with their state loaded for all components
// before any is updated, the rest the plain update.  ADE coefficient sets
// below n_adec come from LDS.
template <typename T>
__device__ __forceinline__ void e_flagged3(T e[3], const T dA[3], const T dB[3], const bool on[3],
                                           const T ca[3], const T cb[3], uint32_t fl, uint32_t ai,
                                           const typename EntryOf<T>::ade *s_ade, int n_adec, int n_slots,
                                           const rf_leap_cold_t *cold, int i, int j, int k) {
    using Ade = typename EntryOf<T>::ade;
    bool ade_c[3];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const uint32_t f = (fl >> (8 * c)) & 0xffu;
        ade_c[c] = on[c] && (f & RF_EDGE_ADE) && !(f & RF_EDGE_TENSOR);
        any = any || ade_c[c];
        if (on[c] && !(f & (RF_EDGE_ADE | RF_EDGE_TENSOR))) e[c] = ca[c] * e[c] + cb[c] * (dA[c] - dB[c]);
    }
    if (!any) return;
    const rf_leap_cold_t &m = *cold_view(cold);
    T *J = static_cast<T *>(m.J), *P = static_cast<T *>(m.P);
    if (n_slots <= 2) {
        // the state of every dispersive edge of the node, loaded before any
        // edge is updated (each edge in the box of its own ADE set)
        T jv[3][2], pv[3][2];
        long long off[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            off[c] = 0;
#pragma unroll
            for (int s = 0; s < 2; s++) jv[c][s] = pv[c][s] = T(0);
            if (!ade_c[c]) continue;
            const int id = (ai >> (8 * c)) & 0xff;
            const Ade &ad = id < n_adec ? s_ade[id] : static_cast<const Ade *>(m.ade)[id];
            off[c] = ade_off(ad, c, i, j, k);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                if (s < ad.n_poles) {
                    const long long o = s * 3 * ad.st + off[c];
                    jv[c][s] = J[o];
                    if (ad.has_p[s]) pv[c][s] = P[o];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (!ade_c[c]) continue;
            const int id = (ai >> (8 * c)) & 0xff;
            const Ade &ad = id < n_adec ? s_ade[id] : static_cast<const Ade *>(m.ade)[id];
            T nj[2], np[2];
            e[c] = ade2<T>(ad, ca[c], cb[c], e[c], dA[c], dB[c], jv[c], pv[c], nj, np);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                if (s < ad.n_poles) {
                    const long long o = s * 3 * ad.st + off[c];
                    J[o] = nj[s];
                    if (ad.has_p[s]) P[o] = np[s];
                }
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (!ade_c[c]) continue;
            const int id = (ai >> (8 * c)) & 0xff;
            const Ade &ad = id < n_adec ? s_ade[id] : static_cast<const Ade *>(m.ade)[id];
            e[c] = ade_edge<T>(ad, ca[c], cb[c], e[c], dA[c], dB[c], J, P, 3 * ad.st, ade_off(ad, c, i, j, k));
        }
    }
}

// Packed ADE set ids of an entry read from the table in HBM.
template <typename T>
__device__ __forceinline__ uint32_t entry_ade_ids(const void *entries, unsigned tag) {
    const typename EntryOf<T>::type &en = static_cast<const typename EntryOf<T>::type *>(entries)[tag];
    uint32_t ai = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) ai |= (static_cast<uint32_t>(en.ade[c]) & 0xffu) << (8 * c);
    return ai;
}

// LDS bytes per cached material entry: Ca, Cb per component, packed flags
// and packed ADE set ids.
template <typename T>
constexpr int coef_bytes() { return 6 * static_cast<int>(sizeof(T)) + 8; }

// Per-lane CPML profile values along z (the lane axis): loaded once per
// wavefront, since a lane keeps its column for the whole march.
template <typename T>
struct ZProf {
    T kh, bh, ah, km, bm, am;
    int cq;        // compressed slab coordinate, -1 outside the psi slab
    bool on;       // outside the plain range: the stretched term applies
};

// One z CPML term (cpml.h:46-51) for a lane outside the plain range, with
// the psi value already loaded (0 outside the psi slab).
template <typename T>
__device__ __forceinline__ T zterm(T diff, T kinv, T b, T aa, bool slab, T old, T &psi_new, T dx) {
    const T psi = slab ? b * old + aa * diff : T(0);
    psi_new = psi;
    return diff * kinv + dx * psi;
}

// LDS view of the tables the fused step stages: the sources with their
// value at the step, the ADE coefficient sets and the first ncache entries
// of the material table (Ca, Cb per component, packed flags, packed ADE
// set ids).
template <typename T>
struct LeapLds {
    LSrc *src;
    typename EntryOf<T>::ade *ade;
    T *ca, *cb;
    uint32_t *fl, *ai;
    int nls, ncache, n_adec;
};

template <typename T>
__device__ __forceinline__ LeapLds<T> leap_lds_view(unsigned char *base, int n_srcs, int n_adec, int ncache) {
    using Ade = typename EntryOf<T>::ade;
    LeapLds<T> v;
    v.nls = n_srcs <= MAX_LSRC ? n_srcs : 0;   // more sources: read from HBM
    v.ncache = ncache;
    v.n_adec = n_adec;
    v.src = reinterpret_cast<LSrc *>(base);
    v.ade = reinterpret_cast<Ade *>(base + (size_t)v.nls * sizeof(LSrc));
    v.ca = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(v.ade) + (size_t)n_adec * sizeof(Ade));
    v.cb = v.ca + 3 * ncache;
    v.fl = reinterpret_cast<uint32_t *>(v.cb + 3 * ncache);
    v.ai = v.fl + ncache;
    return v;
}

// Stages the tables of `a` at `step` into v: the caller's threads t0,
// t0 + nt, ... share the work; the caller orders the stores before the
// reads (a block barrier, or a wavefront fence within one wavefront).
template <typename T, bool TAGS>
__device__ __forceinline__ void leap_stage(const rf_leap_t &a, const LeapLds<T> &v, long long step, int t0, int nt) {
    using Ade = typename EntryOf<T>::ade;
    for (int e = t0; e < v.nls; e += nt) {
        const rf_leap_src_t &p = a.srcs[e];
        LSrc q;
        q.w = p.wave[step];
        q.plane = p.plane;
        q.comp = p.comp;
        q.hard = p.hard;
        q.after = p.after;
        q.axis = p.axis;
        q.index = p.index;
        q.i = p.i;
        q.j = p.j;
        q.k = p.k;
        v.src[e] = q;
    }
    if constexpr (TAGS) {
        using Entry = typename EntryOf<T>::type;
        const Entry *ents = static_cast<const Entry *>(a.entries);
        for (int e = t0; e < v.ncache; e += nt) {
            const Entry &en = ents[e];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                v.ca[3 * e + c] = en.Ca[c];
                v.cb[3 * e + c] = en.Cb[c];
            }
            v.fl[e] = en.flags[0] | (en.flags[1] << 8) | (en.flags[2] << 16);
            uint32_t ai = 0;
#pragma unroll
            for (int c = 0; c < 3; c++) ai |= (static_cast<uint32_t>(en.ade[c]) & 0xffu) << (8 * c);
            v.ai[e] = ai;
        }
        const Ade *ades = static_cast<const Ade *>(a.ade);
        for (int e = t0; e < v.n_adec; e += nt) v.ade[e] = ades[e];
    }
}

// One work item of the fused step: z tile tl (of the launch's part), row
// tile tj and x chunk tc, marched plane by plane through the chunk.  In
// cooperative mode the wavefront keeps pace with the others of its block
// (lane_wave, wid and s_prog as in k_leapfrog).
template <typename T, typename TagT, bool ZC, int R, bool COOP>
__device__ __forceinline__ void leap_item(const rf_leap_t &a, const LeapLds<T> &lv, long long step, int tl, int tj,
                                          int tc, int lane, bool lane_wave, int wid, int *s_prog) {
    constexpr bool TAGS = !std::is_same<TagT, NoTag>::value;
    using Ade = typename EntryOf<T>::ade;
    const Geo g = geo_of(a.g);
    const int nls = lv.nls, ncache = lv.ncache, n_adec = lv.n_adec;
    const LSrc *s_src = lv.src;
    const Ade *s_ade = lv.ade;
    const T *s_ca = lv.ca, *s_cb = lv.cb;
    const uint32_t *s_fl = lv.fl, *s_ai = lv.ai;
    (void)lane_wave;
    (void)wid;
    (void)s_prog;
    const int tk = tl < a.tkr[1] ? a.tkr[0] + tl : a.tkr[2] + (tl - a.tkr[1]);

    // the box of nodes this launch owns, [ib0, ib1) per axis; along z the
    // tiles restart at kbrk (the start of the upper CPML slab)
    const int nt0 = a.ntk0;
    const int k0 = tk < nt0 ? a.ib0[2] + tk * KW : a.kbrk + (tk - nt0) * KW;
    const int kend = tk < nt0 ? min(k0 + KW, a.kbrk) : a.ib1[2];
    const int j0 = a.ib0[1] + tj * R;
    const int jend = a.ib1[1];
    const int i_begin = a.ib0[0] + tc * a.chunk;
    const int i_end = min(i_begin + a.chunk, a.ib1[0]);
    const int sy = static_cast<int>(g.sy);

    // this lane's column: true coordinate, stored coordinate (wrapped on a
    // periodic axis, clamped into the stored box otherwise), ownership
    const int kt = k0 - 1 + lane;
    const int kc = g.pz ? (kt + g.nz) % g.nz : min(max(kt, 0), g.nz);
    const bool col_own = lane >= 1 && lane <= KW && kt < kend;
    // rows j0-1 .. j0+R, stored coordinates
    int jr[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
        const int jt = j0 - 1 + r;
        jr[r] = g.py ? (jt + g.ny) % g.ny : min(max(jt, 0), g.ny);
    }
    // z CPML profile of this column (grid.c stretched differences along z
    // for hx, hy, ex, ey); any such lane in the wavefront
    ZProf<T> zp;
    bool zs = false;
    if constexpr (ZC) {
        zp.on = kc < a.zplo || kc >= a.zphi;
        zp.cq = kc < a.zlo ? kc : (kc >= a.zhi ? a.zlo + kc - a.zhi : -1);
        const int qh = min(kc, g.nz - 1);   // half-grid profiles have nz entries
        zp.kh = zp.on ? static_cast<const T *>(a.zk[0])[qh] : T(1);
        zp.bh = zp.on ? static_cast<const T *>(a.zk[1])[qh] : T(0);
        zp.ah = zp.on ? static_cast<const T *>(a.zk[2])[qh] : T(0);
        zp.km = zp.on ? static_cast<const T *>(a.zk[3])[kc] : T(1);
        zp.bm = zp.on ? static_cast<const T *>(a.zk[4])[kc] : T(0);
        zp.am = zp.on ? static_cast<const T *>(a.zk[5])[kc] : T(0);
        zs = __ballot(zp.on) != 0;
    }
    // z-psi store: per (i, j) node row six slot rows of zlen values,
    // slot-major, so the slab lanes of a wavefront access one contiguous run
    const int zr = 2 * a.zpar, zw = 2 - 2 * a.zpar;   // H psi slots read / written
    const rsrc_t RZ = make_rsrc(a.zpsi, ZC ? static_cast<unsigned>(a.zbytes) : 0u);
    const unsigned zsb = static_cast<unsigned>(a.zlen) * sizeof(T);   // bytes per slot row
    auto zrow = [&](int ii, int jj) {
        return (static_cast<unsigned>(ii) * static_cast<unsigned>(g.ny + 1) + static_cast<unsigned>(jj)) * 6 * zsb;
    };
    // sources: does their bounding box meet this tile's rows and columns;
    // the components they touch before and after the E update
    const bool src_tile = a.n_srcs > 0 && a.src_box[2] < j0 + R && a.src_box[3] >= j0 &&
                          a.src_box[4] < k0 + KW && a.src_box[5] >= k0;
    const bool src_lane = kc >= a.src_box[4] && kc <= a.src_box[5];

    // state blocks: component q at byte q * fsb (ex ey ez hx hy hz)
    constexpr unsigned ES = sizeof(T);
    const unsigned fsb = static_cast<unsigned>(a.fstride) * ES;
    const rsrc_t RA = make_rsrc(a.blk_a, 6 * fsb), RB = make_rsrc(a.blk_b, 6 * fsb);
    // tags (vacuum: an empty resource that is never read)
    const rsrc_t RT = make_rsrc(a.tag, TAGS ? static_cast<unsigned>(a.fstride) * sizeof(TagT) : 0u);
    const unsigned vk = static_cast<unsigned>(kc) * ES;          // this lane's column
    const unsigned sxb = static_cast<unsigned>(g.sx) * ES, syb = static_cast<unsigned>(g.sy) * ES;
    unsigned rowb[R + 2];                                          // row offsets
#pragma unroll
    for (int r = 0; r < R + 2; r++) rowb[r] = static_cast<unsigned>(jr[r]) * syb;
    const T cb = static_cast<T>(a.cb), dx = static_cast<T>(a.dx);
    const T vac = static_cast<T>(a.vac_ca);

    // E^{n+1}(i_begin) needs H^{n+1/2}(i_begin - 1): a warm-up plane
    // computes it without storing (plane nx-1 when x is periodic).
    const bool warm = i_begin > 0 || g.px;
    int i = warm ? (i_begin == 0 ? g.nx - 1 : i_begin - 1) : i_begin;
    const int iters = (i_end - i_begin) + (warm ? 1 : 0);

    T cx[R + 2], cy[R + 2], cz[R + 2];   // E^n, plane i
    T nx[R + 2], ny[R + 2], nz[R + 2];   // E^n, plane i+1
    T py[R + 1], pz[R + 1];              // H^{n+1/2} hy, hz of plane i-1
    {
        const unsigned pb = static_cast<unsigned>(i) * sxb;
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            cx[r] = bld<T>(RA, vk, pb + rowb[r]);
            cy[r] = r <= R ? bld<T>(RA, vk, fsb + pb + rowb[r]) : T(0);
            cz[r] = bld<T>(RA, vk, 2 * fsb + pb + rowb[r]);
        }
    }
#pragma unroll
    for (int r = 0; r <= R; r++) py[r] = pz[r] = T(0);

    for (int it = 0; it < iters; it++) {
        if constexpr (COOP) {
            // plane slack: start plane `it` once every wavefront of the
            // group has completed plane it - gk.  Only a hint for the
            // caches (no wavefront reads what another writes in this
            // launch), so the wait is bounded and then proceeds anyway.
            if (a.gk > 0 && it > a.gk) {
                const int need = it - a.gk;
                for (int spin = 0; spin < (1 << 14); spin++) {
                    const int p = lane_wave ? *reinterpret_cast<volatile int *>(&s_prog[lane]) : need;
                    if (__ballot(p < need) == 0) break;
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        const bool first = warm && it == 0;
        int ip = i + 1;
        if (g.px) {
            if (ip == g.nx) ip = 0;
        } else if (ip > g.nx) {
            ip = g.nx;   // plane nx updates only hx, which needs no i+1
        }
        const unsigned pb = static_cast<unsigned>(i) * sxb, npb = static_cast<unsigned>(ip) * sxb;
        // row j0+R feeds only hx, hz of row j0+R-1 (ez, ex at j+1)
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            nx[r] = bld<T>(RA, vk, npb + rowb[r]);
            ny[r] = r <= R ? bld<T>(RA, vk, fsb + npb + rowb[r]) : T(0);
            nz[r] = bld<T>(RA, vk, 2 * fsb + npb + rowb[r]);
        }
        T hx[R + 1], hy[R + 1], hz[R + 1];
#pragma unroll
        for (int r = 0; r <= R; r++) {
            hx[r] = bld<T>(RA, vk, 3 * fsb + pb + rowb[r]);
            hy[r] = r > 0 ? bld<T>(RA, vk, 4 * fsb + pb + rowb[r]) : T(0);   // the halo row needs hx, hz only
            hz[r] = bld<T>(RA, vk, 5 * fsb + pb + rowb[r]);
        }
        // z-psi of the slab lanes, issued with the field loads: H of hx and
        // hy (read slots) on rows 0..R, E of ex and ey on rows 1..R
        // (wavefront-uniform test, loads from every lane at a clamped
        // node, zero outside the slab: no divergent branch around loads)
        T zh0[R + 1], zh1[R + 1], ze0[R + 1], ze1[R + 1];
#pragma unroll
        for (int r = 0; r <= R; r++) zh0[r] = zh1[r] = ze0[r] = ze1[r] = T(0);
        if constexpr (ZC) {
            if (zs) {
                const bool slab = zp.cq >= 0;
                const unsigned zl = static_cast<unsigned>(slab ? zp.cq : 0) * ES;
#pragma unroll
                for (int r = 0; r <= R; r++) {
                    const unsigned zb = zrow(i, jr[r]);
                    const T h0 = bld<T>(RZ, zl, zb + zr * zsb), h1 = bld<T>(RZ, zl, zb + (zr + 1) * zsb);
                    zh0[r] = slab ? h0 : T(0);
                    zh1[r] = slab ? h1 : T(0);
                    if (r >= 1) {
                        const T e0 = bld<T>(RZ, zl, zb + 4 * zsb), e1 = bld<T>(RZ, zl, zb + 5 * zsb);
                        ze0[r] = slab ? e0 : T(0);
                        ze1[r] = slab ? e1 : T(0);
                    }
                }
            }
        }
        // material tags of the owned rows, issued with the field loads
        unsigned tg[R + 1];
#pragma unroll
        for (int r = 0; r <= R; r++) {
            tg[r] = 0;
            if constexpr (TAGS)
                if (r >= 1)
                    tg[r] = btag<TagT>(RT, static_cast<unsigned>(kc) * sizeof(TagT),
                                       (static_cast<unsigned>(i) * static_cast<unsigned>(g.sx) +
                                        static_cast<unsigned>(jr[r]) * static_cast<unsigned>(sy)) *
                                           sizeof(TagT));
        }

        // ---- H^{n+1/2}(i), rows j0-1 .. j0+R-1 (grid.c:70-89, cpml.c:185-215)
        const bool hx_i = i < g.nx + !g.px, hyz_i = i < g.nx;
#pragma unroll
        for (int r = 0; r <= R; r++) {
            const int jc = jr[r];
            const T eyk = from_upper(cy[r]);   // ey(i, j, k+1)
            const T exk = from_upper(cx[r]);   // ex(i, j, k+1)
            const bool own = !first && r >= 1 && j0 - 1 + r < jend && col_own;
            const bool do_x = hx_i && jc < g.ny && kc < g.nz;
            const bool do_y = r >= 1 && hyz_i && jc < g.ny + !g.py && kc < g.nz;
            const bool do_z = hyz_i && jc < g.ny && kc < g.nz + !g.pz;
            const T dAx = cz[r + 1] - cz[r];
            T dBx = eyk - cy[r];
            T dAy = exk - cx[r];
            const T dBy = nz[r] - cz[r];
            const T dAz = ny[r] - cy[r], dBz = cx[r + 1] - cx[r];
            if constexpr (ZC) {
                if (zs && zp.on && (do_x || do_y)) {   // dBx, dAy are z differences
                    const bool slab = zp.cq >= 0;
                    T px_, py_;
                    if (do_x) dBx = zterm<T>(dBx, zp.kh, zp.bh, zp.ah, slab, zh0[r], px_, dx);
                    if (do_y) dAy = zterm<T>(dAy, zp.kh, zp.bh, zp.ah, slab, zh1[r], py_, dx);
                    if (slab && own) {
                        const unsigned zl = static_cast<unsigned>(zp.cq) * ES, zb = zrow(i, jc);
                        if (do_x) bst(px_, RZ, zl, zb + zw * zsb);
                        if (do_y) bst(py_, RZ, zl, zb + (zw + 1) * zsb);
                    }
                }
            }
            const T nhx = hx[r] - cb * (dAx - dBx);
            const T nhy = hy[r] - cb * (dAy - dBy);
            const T nhz = hz[r] - cb * (dAz - dBz);
            hx[r] = do_x ? nhx : hx[r];
            hy[r] = do_y ? nhy : hy[r];
            hz[r] = do_z ? nhz : hz[r];
            if (own) {
                bst(hx[r], RB, vk, 3 * fsb + pb + rowb[r]);
                bst(hy[r], RB, vk, 4 * fsb + pb + rowb[r]);
                bst(hz[r], RB, vk, 5 * fsb + pb + rowb[r]);
            }
        }

        // ---- sources and E^{n+1}(i), rows j0 .. j0+R-1 (materials.c:975-1058)
        if (!first) {
            const bool i_ok = g.px || i >= 1;
            const bool src = src_tile && a.src_box[0] <= i && a.src_box[1] >= i;
#pragma unroll
            for (int r = 1; r <= R; r++) {
                const T hxk = from_lower(hx[r]);   // hx(i, j, k-1)
                const T hyk = from_lower(hy[r]);   // hy(i, j, k-1)
                const int jc = jr[r];
                if (j0 - 1 + r < jend && col_own) {
                    const bool in = i < g.nx && jc < g.ny && kc < g.nz;
                    const bool j_ok = g.py || jc >= 1, k_ok = g.pz || kc >= 1;
                    const bool on[3] = {in && j_ok && k_ok, in && i_ok && k_ok, in && i_ok && j_ok};
                    T e[3] = {cx[r], cy[r], cz[r]};
                    if (src && src_lane) {
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            if ((a.src_pre >> c) & 1)
                                e[c] = nls == a.n_srcs
                                           ? apply_lsrc<T>(s_src, nls, e[c], c, i, jc, kc, g.nx, g.ny, g.nz, 0)
                                           : apply_sources<T>(a.srcs, a.n_srcs, e[c], c, i, jc, kc, g.nx, g.ny,
                                                              g.nz, step, 0);
                    }
                    T dA[3] = {hz[r] - hz[r - 1], hx[r] - hxk, hy[r] - py[r]};
                    T dB[3] = {hy[r] - hyk, hz[r] - pz[r], hx[r] - hx[r - 1]};
                    if constexpr (ZC) {
                        if (zs && zp.on && (on[0] || on[1])) {   // dB[0], dA[1] are z differences
                            const bool slab = zp.cq >= 0;
                            T px_, py_;
                            if (on[0]) dB[0] = zterm<T>(dB[0], zp.km, zp.bm, zp.am, slab, ze0[r], px_, dx);
                            if (on[1]) dA[1] = zterm<T>(dA[1], zp.km, zp.bm, zp.am, slab, ze1[r], py_, dx);
                            if (slab) {
                                const unsigned zl = static_cast<unsigned>(zp.cq) * ES, zb = zrow(i, jc);
                                if (on[0]) bst(px_, RZ, zl, zb + 4 * zsb);
                                if (on[1]) bst(py_, RZ, zl, zb + 5 * zsb);
                            }
                        }
                    }
                    if constexpr (!TAGS) {
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            const T n = T(1) * e[c] + vac * (dA[c] - dB[c]);
                            e[c] = on[c] ? n : e[c];
                        }
                    } else {
                        const unsigned tag = tg[r];
                        const Coefs<T> kf = entry_coefs<T>(a.entries, s_ca, s_cb, s_fl, ncache, tag);
                        // split ADE step: dispersive edges take the plain
                        // part here and the ADE pass completes them
                        const uint32_t fl = a.ade_split ? kf.fl & ~(RF_EDGE_ADE * 0x010101u) : kf.fl;
                        if (fl == 0) {
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                const T n = kf.ca[c] * e[c] + kf.cb[c] * (dA[c] - dB[c]);
                                e[c] = on[c] ? n : e[c];
                            }
                        } else {
                            const uint32_t ai = static_cast<int>(tag) < ncache ? s_ai[tag]
                                                                                : entry_ade_ids<T>(a.entries, tag);
                            e_flagged3<T>(e, dA, dB, on, kf.ca, kf.cb, fl, ai, s_ade, n_adec, a.n_slots, a.cold,
                                          i, jc, kc);
                        }
                    }
                    if (src && src_lane) {
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            if ((a.src_post >> c) & 1)
                                e[c] = nls == a.n_srcs
                                           ? apply_lsrc<T>(s_src, nls, e[c], c, i, jc, kc, g.nx, g.ny, g.nz, 1)
                                           : apply_sources<T>(a.srcs, a.n_srcs, e[c], c, i, jc, kc, g.nx, g.ny,
                                                              g.nz, step, 1);
                    }
                    bst(e[0], RB, vk, pb + rowb[r]);
                    bst(e[1], RB, vk, fsb + pb + rowb[r]);
                    bst(e[2], RB, vk, 2 * fsb + pb + rowb[r]);
                }
            }
        }

#pragma unroll
        for (int r = 1; r <= R; r++) {
            py[r] = hy[r];
            pz[r] = hz[r];
        }
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            cx[r] = nx[r];
            cy[r] = ny[r];
            cz[r] = nz[r];
        }
        i = first ? i_begin : i + 1;
        // the block's tiles stay on the same plane: a barrier without a
        // memory fence (no wavefront reads what another writes in this
        // launch), so no wavefront waits for its stores to complete
        if constexpr (COOP) {
            if (a.gk == 0)
                __builtin_amdgcn_s_barrier();
            else if (lane == 0)
                *reinterpret_cast<volatile int *>(&s_prog[wid]) = it + 1;
        }
    }
}

template <typename T, typename TagT, bool ZC, int R, bool COOP>
__global__ void __launch_bounds__(WL *WPB)
k_leapfrog(rf_leap_t a) {
    constexpr bool TAGS = !std::is_same<TagT, NoTag>::value;
    // the sources, the ADE coefficient sets and the first a.ncache entries
    // of the material table, staged in LDS before any wavefront of the
    // block may leave
    extern __shared__ __align__(16) unsigned char s_raw[];
    const LeapLds<T> lv = leap_lds_view<T>(s_raw, a.n_srcs, a.n_adec, a.ncache);
    const long long step = *a.d_step;
    leap_stage<T, TAGS>(a, lv, step, threadIdx.x, blockDim.x);
    if (TAGS || lv.nls > 0) __syncthreads();
    const int lane = threadIdx.x & (WL - 1);
    // Work queue: the items are split in N_XCD contiguous ranges, one per
    // XCD (blocks are dealt round-robin to the XCDs), so the tiles an XCD
    // runs at a time are neighbours that share halo rows in its L2; each
    // taker gets the next item of its XCD's counter until the range is
    // exhausted, which balances tiles of unequal cost (dispersive, source
    // and CPML columns).  The counters of this launch are a.ctr[a.qpar][*]
    // (a.ctr points at the pair of sets of its part); the other set is
    // zeroed here for the next step.
    //
    // Independent mode (a.gz == 0): an item is one (tile, x chunk) and each
    // wavefront takes its own.  Cooperative mode: an item is a group of
    // gz x gy tiles (gz neighbouring z tiles by gy neighbouring row tiles)
    // of one x chunk, taken by a block of gz * gy wavefronts that march the
    // planes together (a block barrier per plane, or at most gk planes
    // apart), so the lines that neighbouring tiles share (the halo rows)
    // are fetched once and hit in L1/L2 for the other wavefront instead of
    // being fetched again planes later.  The plan uses gz = 1: tiles of
    // one z tile cost the same per plane (z CPML columns), so no wavefront
    // of a block idles waiting for a costlier neighbour.
    const int xcd = blockIdx.x % N_XCD;
    if (blockIdx.x == 0 && threadIdx.x < N_XCD) a.ctr[(1 - a.qpar) * N_XCD + threadIdx.x] = 0u;
    const int item_lo = (int)(((long long)a.n_items * xcd) / N_XCD);
    const int item_hi = (int)(((long long)a.n_items * (xcd + 1)) / N_XCD);
    unsigned *my_ctr = a.ctr + a.qpar * N_XCD + xcd;
    const Geo g = geo_of(a.g);
    const int wid = static_cast<int>(threadIdx.x) / WL;
    // cooperative mode: the block's ticket and, with a plane slack
    // (a.gk > 0), the planes each wavefront of the group has completed
    __shared__ unsigned s_ticket;
    __shared__ int s_prog[WPB_MAX];
    for (;;) {
    const int ntl = a.tkr[1] + a.tkr[3];   // z tiles of this part
    int tl, tj, tc;
    bool valid = true;
    bool lane_wave = false;   // lane < group size and that wavefront has a tile
    if constexpr (COOP) {
        __syncthreads();   // every wavefront has read the previous ticket
        if (threadIdx.x == 0) s_ticket = atomicAdd(my_ctr, 1u);
        if (lane == 0) s_prog[wid] = 0;
        __syncthreads();
        const int grp = item_lo + static_cast<int>(s_ticket);
        if (grp >= item_hi) break;   // block-uniform
        const int nzg = (ntl + a.gz - 1) / a.gz, nyg = (a.ntj + a.gy - 1) / a.gy;
        const int tl0 = (grp % nzg) * a.gz, tj0 = ((grp / nzg) % nyg) * a.gy;
        tl = tl0 + wid % a.gz;
        tj = tj0 + wid / a.gz;
        tc = grp / (nzg * nyg);
        valid = tl < ntl && tj < a.ntj;
        lane_wave = lane < a.gz * a.gy && tl0 + lane % a.gz < ntl && tj0 + lane / a.gz < a.ntj;
    } else {
        unsigned ticket = 0;
        if (lane == 0) ticket = atomicAdd(my_ctr, 1u);
        const int item = item_lo + static_cast<int>(__builtin_amdgcn_readfirstlane(ticket));
        if (item >= item_hi) break;
        tl = item % ntl;
        tj = (item / ntl) % a.ntj;
        tc = item / (ntl * a.ntj);
    }
    if (!valid) {
        // a wavefront of the block without a tile: only the barriers of
        // the group's planes (the count depends on the chunk alone)
        if (a.gk == 0) {
            const int ib = a.ib0[0] + tc * a.chunk, ie = min(ib + a.chunk, a.ib1[0]);
            const int n_it = (ie - ib) + ((ib > 0 || g.px) ? 1 : 0);
            for (int it = 0; it < n_it; it++) __builtin_amdgcn_s_barrier();
        }
        continue;
    }
    leap_item<T, TagT, ZC, R, COOP>(a, lv, step, tl, tj, tc, lane, lane_wave, wid, s_prog);
    }
}

// ------------------------------------------------------- packed scenes
// The fused steps of many scenes in one launch.  The work items of all
// scenes form one queue (scene sc owns the items item_base[sc] ..
// item_base[sc] + tab[sc].n_items, item_scene names the scene of each),
// split in N_XCD contiguous ranges as in k_leapfrog, so an XCD works
// through whole scenes and keeps their halos in its L2.  The four
// wavefronts of a block take items on their own, from any scene: each
// stages the tables of the scene it works on in its own LDS region when
// it moves to an item of another scene.  LDS operations of one wavefront
// complete in order, so its staging stores precede its later reads and
// no block barrier is needed.
template <typename T, typename TagT, bool ZC, int R>
__global__ void __launch_bounds__(WL *WPB)
k_leapfrog_packed(const rf_leap_t *__restrict__ tab, const int *__restrict__ item_base,
                  const unsigned short *__restrict__ item_scene, int n_items, unsigned *ctr, int qpar,
                  int lds_wave) {
    constexpr bool TAGS = !std::is_same<TagT, NoTag>::value;
    extern __shared__ __align__(16) unsigned char s_raw[];
    const int lane = threadIdx.x & (WL - 1);
    const int wid = static_cast<int>(threadIdx.x) / WL;
    unsigned char *mine = s_raw + (size_t)wid * lds_wave;
    const int xcd = blockIdx.x % N_XCD;
    if (blockIdx.x == 0 && threadIdx.x < N_XCD) ctr[(1 - qpar) * N_XCD + threadIdx.x] = 0u;
    const int item_lo = (int)(((long long)n_items * xcd) / N_XCD);
    const int item_hi = (int)(((long long)n_items * (xcd + 1)) / N_XCD);
    unsigned *my_ctr = ctr + qpar * N_XCD + xcd;
    int cur = -1;
    LeapLds<T> lv = leap_lds_view<T>(mine, 0, 0, 0);
    long long step = 0;
    for (;;) {
        unsigned ticket = 0;
        if (lane == 0) ticket = atomicAdd(my_ctr, 1u);
        const int item = item_lo + static_cast<int>(__builtin_amdgcn_readfirstlane(ticket));
        if (item >= item_hi) break;
        const int sc = __builtin_amdgcn_readfirstlane(static_cast<int>(item_scene[item]));
        const rf_leap_t &a = tab[sc];
        if (sc != cur) {
            lv = leap_lds_view<T>(mine, a.n_srcs, a.n_adec, a.ncache);
            step = *a.d_step;
            leap_stage<T, TAGS>(a, lv, step, lane, WL);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            cur = sc;
        }
        const int local = item - item_base[sc];
        const int ntl = a.tkr[1] + a.tkr[3];
        leap_item<T, TagT, ZC, R, false>(a, lv, step, local % ntl, (local / ntl) % a.ntj, local / (ntl * a.ntj),
                                         lane, false, wid, nullptr);
    }
}

// ------------------------------------------------------------ row mode
// FP32, z CPML inside the box and the box spanning every z column [0, nz]:
// one wavefront owns R rows (y) of all the columns and marches along x
// through a chunk of planes.  Lane l of chunk c holds column 64 c + l, so
// a row is loaded and stored as whole 128-byte lines that no other
// wavefront touches, and every wavefront does the same work per plane (the
// z CPML slabs lie in its first and last chunk): the wavefronts of an XCD
// stay at nearby planes and their halo rows meet in its L2.  A chunk's
// edge lane takes its z neighbour from the next or previous chunk by
// readlane.  When nz = 64 NCH, column nz lies past the chunks; there
// only hz changes (the boundary's tangential E and the H components with a
// z difference keep their values, equal in both state buffers), so its E
// values are read as wavefront-uniform values and its hz update is done by
// every lane and stored by one.  The expressions and their order are those
// of leap_item.
constexpr int ROW_NCH_MAX = 4;   // NCH chunks of 64 columns: 64 (NCH - 1) < nz <= 64 NCH

__device__ __forceinline__ float lane_of(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ float uniform_of(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <typename TagT, int R, int NCH>
__device__ __forceinline__ void leap_row(const rf_leap_t &a, const LeapLds<float> &lv, const float *s_zk,
                                         long long step, int tj, int tc, int lane) {
    using T = float;
    constexpr bool TAGS = !std::is_same<TagT, NoTag>::value;
    const Geo g = geo_of(a.g);
    const int nls = lv.nls, ncache = lv.ncache;
    const LSrc *s_src = lv.src;
    const T *s_ca = lv.ca, *s_cb = lv.cb;
    const uint32_t *s_fl = lv.fl;

    const int j0 = a.ib0[1] + tj * R;
    const int jend = a.ib1[1];
    const int i_begin = a.ib0[0] + tc * a.chunk;
    const int i_end = min(i_begin + a.chunk, a.ib1[0]);
    const int sy = static_cast<int>(g.sy);
    const bool bnd = g.nz == 64 * NCH;   // column nz past the chunks

    // this lane's column in each chunk (clamped to nz past it), ownership
    // (nz > 64 (NCH - 1): only the last chunk reaches past nz)
    int kc[NCH];
    bool col_own[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const int k = 64 * c + lane;
        kc[c] = c < NCH - 1 ? k : min(k, g.nz);
        col_own[c] = c < NCH - 1 || k <= g.nz;
    }
    int jr[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
        const int jt = j0 - 1 + r;
        jr[r] = g.py ? (jt + g.ny) % g.ny : min(max(jt, 0), g.ny);
    }
    // z CPML of the first (e = 0) and last (e = 1) chunk's column: outside
    // the plain range, slab coordinate; the profile values are read from
    // the block's LDS copy (zprof) where they are used
    bool zon[2], zs[2];
    int zcq[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int q = e == 0 ? kc[0] : kc[NCH - 1];
        zon[e] = q < a.zplo || q >= a.zphi;
        zcq[e] = q < a.zlo ? q : (q >= a.zhi ? a.zlo + q - a.zhi : -1);
        zs[e] = __ballot(zon[e]) != 0;
    }
    auto zprof = [&](int e, int p) { return s_zk[(e * 6 + p) * WL + lane]; };
    const int zr = 2 * a.zpar, zw = 2 - 2 * a.zpar;
    const rsrc_t RZ = make_rsrc(a.zpsi, static_cast<unsigned>(a.zbytes));
    const unsigned zsb = static_cast<unsigned>(a.zlen) * sizeof(T);
    auto zrow = [&](int ii, int jj) {
        return (static_cast<unsigned>(ii) * static_cast<unsigned>(g.ny + 1) + static_cast<unsigned>(jj)) * 6 * zsb;
    };
    const bool src_tile = a.n_srcs > 0 && a.src_box[2] < j0 + R && a.src_box[3] >= j0;
    bool src_lane[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) src_lane[c] = kc[c] >= a.src_box[4] && kc[c] <= a.src_box[5];

    constexpr unsigned ES = sizeof(T);
    const unsigned fsb = static_cast<unsigned>(a.fstride) * ES;
    const rsrc_t RA = make_rsrc(a.blk_a, 6 * fsb), RB = make_rsrc(a.blk_b, 6 * fsb);
    const rsrc_t RT = make_rsrc(a.tag, TAGS ? static_cast<unsigned>(a.fstride) * sizeof(TagT) : 0u);
    unsigned vk[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) vk[c] = static_cast<unsigned>(kc[c]) * ES;
    const unsigned vb = static_cast<unsigned>(g.nz) * ES;   // column nz
    const unsigned sxb = static_cast<unsigned>(g.sx) * ES, syb = static_cast<unsigned>(g.sy) * ES;
    unsigned rowb[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++) rowb[r] = static_cast<unsigned>(jr[r]) * syb;
    const T cb = static_cast<T>(a.cb), dx = static_cast<T>(a.dx);
    const T vac = static_cast<T>(a.vac_ca);

    const bool warm = i_begin > 0 || g.px;
    int i = warm ? (i_begin == 0 ? g.nx - 1 : i_begin - 1) : i_begin;
    const int iters = (i_end - i_begin) + (warm ? 1 : 0);

    T cx[NCH][R + 2], cy[NCH][R + 2], cz[NCH][R + 2];   // E^n, plane i
    T nx[NCH][R + 2], ny[NCH][R + 2], nz[NCH][R + 2];   // E^n, plane i+1
    T py[NCH][R + 1], pz[NCH][R + 1];                   // hy, hz of plane i-1
    T bx[R + 2], by[R + 1];                             // ex, ey of column nz, plane i
    {
        const unsigned pb = static_cast<unsigned>(i) * sxb;
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int r = 0; r < R + 2; r++) {
                cx[c][r] = bld<T>(RA, vk[c], pb + rowb[r]);
                cy[c][r] = r <= R ? bld<T>(RA, vk[c], fsb + pb + rowb[r]) : T(0);
                cz[c][r] = bld<T>(RA, vk[c], 2 * fsb + pb + rowb[r]);
            }
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            bx[r] = bnd ? uniform_of(bld<T>(RA, vb, pb + rowb[r])) : T(0);
            if (r <= R) by[r] = bnd ? uniform_of(bld<T>(RA, vb, fsb + pb + rowb[r])) : T(0);
        }
    }
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int r = 0; r <= R; r++) py[c][r] = pz[c][r] = T(0);

    for (int it = 0; it < iters; it++) {
        const bool first = warm && it == 0;
        int ip = i + 1;
        if (g.px) {
            if (ip == g.nx) ip = 0;
        } else if (ip > g.nx) {
            ip = g.nx;
        }
        const unsigned pb = static_cast<unsigned>(i) * sxb, npb = static_cast<unsigned>(ip) * sxb;
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int r = 0; r < R + 2; r++) {
                nx[c][r] = bld<T>(RA, vk[c], npb + rowb[r]);
                ny[c][r] = r <= R ? bld<T>(RA, vk[c], fsb + npb + rowb[r]) : T(0);
                nz[c][r] = bld<T>(RA, vk[c], 2 * fsb + npb + rowb[r]);
            }
        T hx[NCH][R + 1], hy[NCH][R + 1], hz[NCH][R + 1];
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int r = 0; r <= R; r++) {
                hx[c][r] = bld<T>(RA, vk[c], 3 * fsb + pb + rowb[r]);
                hy[c][r] = r > 0 ? bld<T>(RA, vk[c], 4 * fsb + pb + rowb[r]) : T(0);
                hz[c][r] = bld<T>(RA, vk[c], 5 * fsb + pb + rowb[r]);
            }
        T nbx[R + 2], nby[R + 1], bhz[R + 1];
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            nbx[r] = bnd ? uniform_of(bld<T>(RA, vb, npb + rowb[r])) : T(0);
            if (r <= R) {
                nby[r] = bnd ? uniform_of(bld<T>(RA, vb, fsb + npb + rowb[r])) : T(0);
                bhz[r] = bnd && r >= 1 ? uniform_of(bld<T>(RA, vb, 5 * fsb + pb + rowb[r])) : T(0);
            }
        }
        // z-psi of the slab lanes of the first and last chunk
        T zh0[2][R + 1], zh1[2][R + 1], ze0[2][R + 1], ze1[2][R + 1];
#pragma unroll
        for (int e = 0; e < 2; e++) {
#pragma unroll
            for (int r = 0; r <= R; r++) zh0[e][r] = zh1[e][r] = ze0[e][r] = ze1[e][r] = T(0);
            if (zs[e]) {
                const bool slab = zcq[e] >= 0;
                const unsigned zl = static_cast<unsigned>(slab ? zcq[e] : 0) * ES;
#pragma unroll
                for (int r = 0; r <= R; r++) {
                    const unsigned zb = zrow(i, jr[r]);
                    const T h0 = bld<T>(RZ, zl, zb + zr * zsb), h1 = bld<T>(RZ, zl, zb + (zr + 1) * zsb);
                    zh0[e][r] = slab ? h0 : T(0);
                    zh1[e][r] = slab ? h1 : T(0);
                    if (r >= 1) {
                        const T e0 = bld<T>(RZ, zl, zb + 4 * zsb), e1 = bld<T>(RZ, zl, zb + 5 * zsb);
                        ze0[e][r] = slab ? e0 : T(0);
                        ze1[e][r] = slab ? e1 : T(0);
                    }
                }
            }
        }
        unsigned tg[NCH][R + 1];
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int r = 0; r <= R; r++) {
                tg[c][r] = 0;
                if constexpr (TAGS)
                    if (r >= 1)
                        tg[c][r] = btag<TagT>(RT, static_cast<unsigned>(kc[c]) * sizeof(TagT),
                                              (static_cast<unsigned>(i) * static_cast<unsigned>(g.sx) +
                                               static_cast<unsigned>(jr[r]) * static_cast<unsigned>(sy)) *
                                                  sizeof(TagT));
            }

        // ---- H^{n+1/2}(i), rows j0-1 .. j0+R-1 (grid.c:70-89, cpml.c:185-215)
        const bool hx_i = i < g.nx + !g.px, hyz_i = i < g.nx;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int e = c == 0 ? 0 : 1;
            const bool zce = c == 0 || c == NCH - 1;
#pragma unroll
            for (int r = 0; r <= R; r++) {
                const int jc = jr[r];
                T eyk = from_upper(cy[c][r]);   // ey(i, j, k+1)
                T exk = from_upper(cx[c][r]);   // ex(i, j, k+1)
                const T uy = c + 1 < NCH ? lane_of(cy[c + 1 < NCH ? c + 1 : c][r], 0) : by[r];
                const T ux = c + 1 < NCH ? lane_of(cx[c + 1 < NCH ? c + 1 : c][r], 0) : bx[r];
                eyk = lane == WL - 1 ? uy : eyk;
                exk = lane == WL - 1 ? ux : exk;
                const int k = kc[c];
                const bool own = !first && r >= 1 && j0 - 1 + r < jend && col_own[c];
                const bool do_x = hx_i && jc < g.ny && k < g.nz;
                const bool do_y = r >= 1 && hyz_i && jc < g.ny + !g.py && k < g.nz;
                const bool do_z = hyz_i && jc < g.ny && k < g.nz + !g.pz;
                const T dAx = cz[c][r + 1] - cz[c][r];
                T dBx = eyk - cy[c][r];
                T dAy = exk - cx[c][r];
                const T dBy = nz[c][r] - cz[c][r];
                const T dAz = ny[c][r] - cy[c][r], dBz = cx[c][r + 1] - cx[c][r];
                if (zce && zs[e] && zon[e] && (do_x || do_y)) {
                    const bool slab = zcq[e] >= 0;
                    const T kh = zprof(e, 0), bh = zprof(e, 1), ah = zprof(e, 2);
                    T px_, py_;
                    if (do_x) dBx = zterm<T>(dBx, kh, bh, ah, slab, zh0[e][r], px_, dx);
                    if (do_y) dAy = zterm<T>(dAy, kh, bh, ah, slab, zh1[e][r], py_, dx);
                    if (slab && own) {
                        const unsigned zl = static_cast<unsigned>(zcq[e]) * ES, zb = zrow(i, jc);
                        if (do_x) bst(px_, RZ, zl, zb + zw * zsb);
                        if (do_y) bst(py_, RZ, zl, zb + (zw + 1) * zsb);
                    }
                }
                const T nhx = hx[c][r] - cb * (dAx - dBx);
                const T nhy = hy[c][r] - cb * (dAy - dBy);
                const T nhz = hz[c][r] - cb * (dAz - dBz);
                hx[c][r] = do_x ? nhx : hx[c][r];
                hy[c][r] = do_y ? nhy : hy[c][r];
                hz[c][r] = do_z ? nhz : hz[c][r];
                if (own) {
                    bst(hx[c][r], RB, vk[c], 3 * fsb + pb + rowb[r]);
                    bst(hy[c][r], RB, vk[c], 4 * fsb + pb + rowb[r]);
                    bst(hz[c][r], RB, vk[c], 5 * fsb + pb + rowb[r]);
                }
            }
        }
        // hz of column nz (z is not periodic: k = nz < nz + 1)
        if (bnd && !first) {
#pragma unroll
            for (int r = 1; r <= R; r++) {
                const int jc = jr[r];
                const bool do_z = hyz_i && jc < g.ny;
                const T dAz = nby[r] - by[r], dBz = bx[r + 1] - bx[r];
                const T nhz = bhz[r] - cb * (dAz - dBz);
                if (j0 - 1 + r < jend && lane == 0) bst(do_z ? nhz : bhz[r], RB, vb, 5 * fsb + pb + rowb[r]);
            }
        }

        // ---- sources and E^{n+1}(i), rows j0 .. j0+R-1 (materials.c:975-1058)
        if (!first) {
            const bool i_ok = g.px || i >= 1;
            const bool src = src_tile && a.src_box[0] <= i && a.src_box[1] >= i;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int e_ = c == 0 ? 0 : 1;
                const bool zce = c == 0 || c == NCH - 1;
                const int kk = kc[c];
#pragma unroll
                for (int r = 1; r <= R; r++) {
                    T hxk = from_lower(hx[c][r]);   // hx(i, j, k-1)
                    T hyk = from_lower(hy[c][r]);   // hy(i, j, k-1)
                    if (c > 0) {
                        const T lx = lane_of(hx[c > 0 ? c - 1 : 0][r], WL - 1);
                        const T ly = lane_of(hy[c > 0 ? c - 1 : 0][r], WL - 1);
                        hxk = lane == 0 ? lx : hxk;
                        hyk = lane == 0 ? ly : hyk;
                    }
                    const int jc = jr[r];
                    if (j0 - 1 + r < jend && col_own[c]) {
                        const bool in = i < g.nx && jc < g.ny && kk < g.nz;
                        const bool j_ok = g.py || jc >= 1, k_ok = g.pz || kk >= 1;
                        const bool on[3] = {in && j_ok && k_ok, in && i_ok && k_ok, in && i_ok && j_ok};
                        T ev[3] = {cx[c][r], cy[c][r], cz[c][r]};
                        if (src && src_lane[c]) {
#pragma unroll
                            for (int q = 0; q < 3; q++)
                                if ((a.src_pre >> q) & 1)
                                    ev[q] = nls == a.n_srcs
                                                ? apply_lsrc<T>(s_src, nls, ev[q], q, i, jc, kk, g.nx, g.ny, g.nz, 0)
                                                : apply_sources<T>(a.srcs, a.n_srcs, ev[q], q, i, jc, kk, g.nx, g.ny,
                                                                   g.nz, step, 0);
                        }
                        T dA[3] = {hz[c][r] - hz[c][r - 1], hx[c][r] - hxk, hy[c][r] - py[c][r]};
                        T dB[3] = {hy[c][r] - hyk, hz[c][r] - pz[c][r], hx[c][r] - hx[c][r - 1]};
                        if (zce && zs[e_] && zon[e_] && (on[0] || on[1])) {
                            const bool slab = zcq[e_] >= 0;
                            const T km = zprof(e_, 3), bm = zprof(e_, 4), am = zprof(e_, 5);
                            T px_, py_;
                            if (on[0]) dB[0] = zterm<T>(dB[0], km, bm, am, slab, ze0[e_][r], px_, dx);
                            if (on[1]) dA[1] = zterm<T>(dA[1], km, bm, am, slab, ze1[e_][r], py_, dx);
                            if (slab) {
                                const unsigned zl = static_cast<unsigned>(zcq[e_]) * ES, zb = zrow(i, jc);
                                if (on[0]) bst(px_, RZ, zl, zb + 4 * zsb);
                                if (on[1]) bst(py_, RZ, zl, zb + 5 * zsb);
                            }
                        }
                        if constexpr (!TAGS) {
#pragma unroll
                            for (int q = 0; q < 3; q++) {
                                const T n = T(1) * ev[q] + vac * (dA[q] - dB[q]);
                                ev[q] = on[q] ? n : ev[q];
                            }
                        } else {
                            // dispersive edges (split ADE: plain part here)
                            // and plain edges take the update; tensor edges
                            // keep their value for the tensor kernel, as in
                            // e_flagged3 (row mode runs only where no ADE
                            // update is left to the fused kernel)
                            const Coefs<T> kf = entry_coefs<T>(a.entries, s_ca, s_cb, s_fl, ncache, tg[c][r]);
                            const uint32_t fl = kf.fl & ~(RF_EDGE_ADE * 0x010101u);
#pragma unroll
                            for (int q = 0; q < 3; q++) {
                                const T n = kf.ca[q] * ev[q] + kf.cb[q] * (dA[q] - dB[q]);
                                const bool upd = on[q] && !(((fl >> (8 * q)) & 0xffu) & RF_EDGE_TENSOR);
                                ev[q] = upd ? n : ev[q];
                            }
                        }
                        if (src && src_lane[c]) {
#pragma unroll
                            for (int q = 0; q < 3; q++)
                                if ((a.src_post >> q) & 1)
                                    ev[q] = nls == a.n_srcs
                                                ? apply_lsrc<T>(s_src, nls, ev[q], q, i, jc, kk, g.nx, g.ny, g.nz, 1)
                                                : apply_sources<T>(a.srcs, a.n_srcs, ev[q], q, i, jc, kk, g.nx, g.ny,
                                                                   g.nz, step, 1);
                        }
                        bst(ev[0], RB, vk[c], pb + rowb[r]);
                        bst(ev[1], RB, vk[c], fsb + pb + rowb[r]);
                        bst(ev[2], RB, vk[c], 2 * fsb + pb + rowb[r]);
                    }
                }
            }
        }

#pragma unroll
        for (int c = 0; c < NCH; c++) {
#pragma unroll
            for (int r = 1; r <= R; r++) {
                py[c][r] = hy[c][r];
                pz[c][r] = hz[c][r];
            }
#pragma unroll
            for (int r = 0; r < R + 2; r++) {
                cx[c][r] = nx[c][r];
                cy[c][r] = ny[c][r];
                cz[c][r] = nz[c][r];
            }
        }
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
            bx[r] = nbx[r];
            if (r <= R) by[r] = nby[r];
        }
        i = first ? i_begin : i + 1;
    }
}

// Row-mode step: items are (row tile, x chunk), row tile fastest, taken
// from the XCD-major queue of k_leapfrog (independent wavefronts).
// (at most 168 VGPRs: three wavefronts per SIMD; four would spill)
template <typename TagT, int R, int NCH>
__global__ void __launch_bounds__(WL *WPB) __attribute__((amdgpu_waves_per_eu(3)))
k_leap_rows(rf_leap_t a) {
    constexpr bool TAGS = !std::is_same<TagT, NoTag>::value;
    extern __shared__ __align__(16) unsigned char s_raw[];
    const LeapLds<float> lv = leap_lds_view<float>(s_raw, a.n_srcs, a.n_adec, a.ncache);
    const long long step = *a.d_step;
    leap_stage<float, TAGS>(a, lv, step, threadIdx.x, blockDim.x);
    // z CPML profiles (kinv_h b_h a_h kinv_m b_m a_m) of the columns of
    // the first and last chunk: 1, 0, 0 inside the plain range
    __shared__ float s_zk[2 * 6 * WL];
    const int nz = a.g.n[2];
    for (int t = threadIdx.x; t < 2 * 6 * WL; t += blockDim.x) {
        const int l = t % WL, p = (t / WL) % 6, e = t / (6 * WL);
        const int q = e == 0 ? l : min(64 * (NCH - 1) + l, nz);
        const bool on = q < a.zplo || q >= a.zphi;
        const int qq = p < 3 ? min(q, nz - 1) : q;   // half-grid profiles have nz entries
        s_zk[t] = on ? static_cast<const float *>(a.zk[p])[qq] : (p % 3 == 0 ? 1.f : 0.f);
    }
    __syncthreads();
    const int lane = threadIdx.x & (WL - 1);
    const int xcd = blockIdx.x % N_XCD;
    if (blockIdx.x == 0 && threadIdx.x < N_XCD) a.ctr[(1 - a.qpar) * N_XCD + threadIdx.x] = 0u;
    const int item_lo = (int)(((long long)a.n_items * xcd) / N_XCD);
    const int item_hi = (int)(((long long)a.n_items * (xcd + 1)) / N_XCD);
    unsigned *my_ctr = a.ctr + a.qpar * N_XCD + xcd;
    for (;;) {
        unsigned ticket = 0;
        if (lane == 0) ticket = atomicAdd(my_ctr, 1u);
        const int item = item_lo + static_cast<int>(__builtin_amdgcn_readfirstlane(ticket));
        if (item >= item_hi) break;
        leap_row<TagT, R, NCH>(a, lv, s_zk, step, item % a.ntj, item / a.ntj, lane);
    }
}

// ------------------------------------------------------------- shell
// The nodes around the fused box (CPML slabs and their one-node margin)
// are updated one node per thread, H then E, with the expressions of the
// two-kernel path (yee_kernels.hip): H from the A buffers into B, then E
// from E^n (A) and H^{n+1/2} (B, written by the fused kernel and the shell
// H kernel) into B.  psi and the ADE state are in place: no other launch
// touches the slab memories.

__device__ __forceinline__ unsigned div_magic(unsigned n, unsigned m, unsigned s1, unsigned s2) {
    const unsigned t = __umulhi(m, n);
    return (t + ((n - t) >> s1)) >> s2;
}

// Node of flat shell index t (false past the end).  The box is chosen with
// compile-time indices so the argument struct is never indexed at run time;
// the index inside the box is split with multiply-high divisions.
__device__ __forceinline__ bool shell_node(const rf_shell_t &sh, long long t, int &i, int &j, int &k) {
    if (t >= sh.start[sh.n]) return false;
    // the box of the node: the boxes that start at or before it, less one
    int bx = 0;
#pragma unroll
    for (int b = 1; b < 6; b++) bx += (b < sh.n && t >= sh.start[b]) ? 1 : 0;
    long long s0;
    int l0, l1, l2;
    unsigned nk, nj, mk, mj, k1, k2, j1, j2;
    const int b0 = __builtin_amdgcn_readfirstlane(bx);
    if (__ballot(bx != b0) == 0) {
        // one box for every active lane (all but the wavefronts that
        // straddle two boxes): its fields are read once, as scalars
        s0 = sh.start[b0];
        l0 = sh.lo[b0][0];
        l1 = sh.lo[b0][1];
        l2 = sh.lo[b0][2];
        nj = static_cast<unsigned>(sh.hi[b0][1] - sh.lo[b0][1]);
        nk = static_cast<unsigned>(sh.hi[b0][2] - sh.lo[b0][2]);
        mk = sh.mk[b0];
        mj = sh.mj[b0];
        k1 = sh.sk[b0][0];
        k2 = sh.sk[b0][1];
        j1 = sh.sj[b0][0];
        j2 = sh.sj[b0][1];
    } else {
        s0 = 0;
        l0 = l1 = l2 = 0;
        nk = nj = 1;
        mk = mj = k1 = k2 = j1 = j2 = 0;
#pragma unroll
        for (int b = 0; b < 6; b++) {
            if (b == bx) {
                s0 = sh.start[b];
                l0 = sh.lo[b][0];
                l1 = sh.lo[b][1];
                l2 = sh.lo[b][2];
                nj = static_cast<unsigned>(sh.hi[b][1] - sh.lo[b][1]);
                nk = static_cast<unsigned>(sh.hi[b][2] - sh.lo[b][2]);
                mk = sh.mk[b];
                mj = sh.mj[b];
                k1 = sh.sk[b][0];
                k2 = sh.sk[b][1];
                j1 = sh.sj[b][0];
                j2 = sh.sj[b][1];
            }
        }
    }
    const unsigned r = static_cast<unsigned>(t - s0);
    const unsigned q = div_magic(r, mk, k1, k2);
    const unsigned p = div_magic(q, mj, j1, j2);
    k = l2 + static_cast<int>(r - q * nk);
    j = l1 + static_cast<int>(q - p * nj);
    i = l0 + static_cast<int>(p);
    return true;
}

// The z CPML term of a shell node (term_load, yee_device.h) with its psi in
// the z-psi store: read slot rd, write slot wr.
template <typename T>
__device__ __forceinline__ Term<T> zterm_load(const rf_cpml_t &c, const Geo &g, bool h, int i, int j, int k,
                                              const rf_zpsi_t &z, int rd, int wr) {
    Term<T> o;
    o.kinv = ((gmem_c<T> *)(h ? c.kinv_h[2] : c.kinv_m[2]))[k];
    o.b = o.a = o.psi = T(0);
    o.rd = o.wr = -1;
    const int cq = slab_coord(c, 2, k);
    if (cq >= 0) {
        const long long zo = ((long long)i * (g.ny + 1) + j) * 6 * z.zlen + cq;
        o.rd = zo + (long long)rd * z.zlen;
        o.wr = zo + (long long)wr * z.zlen;
        o.b = ((gmem_c<T> *)(h ? c.b_h[2] : c.b_m[2]))[k];
        o.a = ((gmem_c<T> *)(h ? c.a_h[2] : c.a_m[2]))[k];
        o.psi = ((gmem_c<T> *)z.zpsi)[o.rd];
    }
    return o;
}

// H^{n+1/2} of shell node t.
template <typename T, bool CPML>
__device__ __forceinline__ void shell_h_node(const rf_geom_t &gg, const rf_shell_t &sh, long long t,
                                             const T *__restrict__ ex_p, const T *__restrict__ ey_p,
                                             const T *__restrict__ ez_p, const T *__restrict__ hxa_p,
                                             const T *__restrict__ hya_p, const T *__restrict__ hza_p,
                                             T *__restrict__ hxb_p, T *__restrict__ hyb_p, T *__restrict__ hzb_p,
                                             T cb, T dx, const rf_cpml_t &cp, const rf_zpsi_t &zs) {
    gmem_c<T> *const ex = (gmem_c<T> *)ex_p, *const ey = (gmem_c<T> *)ey_p, *const ez = (gmem_c<T> *)ez_p;
    gmem_c<T> *const hxa = (gmem_c<T> *)hxa_p, *const hya = (gmem_c<T> *)hya_p, *const hza = (gmem_c<T> *)hza_p;
    gmem<T> *const hxb = (gmem<T> *)hxb_p, *const hyb = (gmem<T> *)hyb_p, *const hzb = (gmem<T> *)hzb_p;
    const Geo g = geo_of(gg);
    int i, j, k;
    if (!shell_node(sh, t, i, j, k)) return;
    const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;
    const bool do_x = (i < g.nx + !g.px) && j < g.ny && k < g.nz;
    const bool do_y = i < g.nx && (j < g.ny + !g.py) && k < g.nz;
    const bool do_z = i < g.nx && j < g.ny && (k < g.nz + !g.pz);
    // +1 neighbours with periodic wrap (grid.h:26-28)
    const long long di = (g.px && i + 1 == g.nx) ? -(long long)i * g.sx : g.sx;
    const long long dj = (g.py && j + 1 == g.ny) ? -(long long)j * g.sy : g.sy;
    const long long dk = (g.pz && k + 1 == g.nz) ? -(long long)k : 1;
    // every load of the node first, then the updates, then the stores (as
    // in shell_e_node)
    T hx = hxa[idx], hy = hya[idx], hz = hza[idx];
    const T exc = ex[idx], eyc = ey[idx], ezc = ez[idx];
    T ez_j = T(0), ey_k = T(0), ex_k = T(0), ez_i = T(0), ey_i = T(0), ex_j = T(0);
    if (do_x) {
        ez_j = ez[idx + dj];
        ey_k = ey[idx + dk];
    }
    if (do_y) {
        ex_k = ex[idx + dk];
        ez_i = ez[idx + di];
    }
    if (do_z) {
        ey_i = ey[idx + di];
        ex_j = ex[idx + dj];
    }
    Term<T> ta[3] = {}, tb[3] = {};
    if (CPML) {
        if (do_x) {
            ta[0] = term_load<T>(cp, g, true, 1, j, cp.psi_h[0][0], i, j, k);
            tb[0] = zs.zpsi ? zterm_load<T>(cp, g, true, i, j, k, zs, 2 * zs.zpar, 2 - 2 * zs.zpar)
                            : term_load<T>(cp, g, true, 2, k, cp.psi_h[0][1], i, j, k);
        }
        if (do_y) {
            ta[1] = zs.zpsi ? zterm_load<T>(cp, g, true, i, j, k, zs, 2 * zs.zpar + 1, 3 - 2 * zs.zpar)
                            : term_load<T>(cp, g, true, 2, k, cp.psi_h[1][0], i, j, k);
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
        }
        hx -= cb * (dA - dB);
    }
    if (do_y) {   // grid.c:78, cpml.c:198-202
        T dA = ex_k - exc;
        T dB = ez_i - ezc;
        if (CPML) {
            dA = term_apply<T>(ta[1], dA, dx);
            dB = term_apply<T>(tb[1], dB, dx);
        }
        hy -= cb * (dA - dB);
    }
    if (do_z) {   // grid.c:86, cpml.c:211-215
        T dA = ey_i - eyc;
        T dB = ex_j - exc;
        if (CPML) {
            dA = term_apply<T>(ta[2], dA, dx);
            dB = term_apply<T>(tb[2], dB, dx);
        }
        hz -= cb * (dA - dB);
    }
    if (CPML) {
        if (do_x) {
            term_store<T>(ta[0], cp.psi_h[0][0]);
            term_store<T>(tb[0], zs.zpsi ? zs.zpsi : cp.psi_h[0][1]);
        }
        if (do_y) {
            term_store<T>(ta[1], zs.zpsi ? zs.zpsi : cp.psi_h[1][0]);
            term_store<T>(tb[1], cp.psi_h[1][1]);
        }
        if (do_z) {
            term_store<T>(ta[2], cp.psi_h[2][0]);
            term_store<T>(tb[2], cp.psi_h[2][1]);
        }
    }
    hxb[idx] = hx;
    hyb[idx] = hy;
    hzb[idx] = hz;
}

template <typename T, bool CPML>
__global__ void __launch_bounds__(256)
k_shell_h(rf_geom_t gg, rf_shell_t sh, const T *__restrict__ ex, const T *__restrict__ ey,
          const T *__restrict__ ez, const T *__restrict__ hxa, const T *__restrict__ hya,
          const T *__restrict__ hza, T *__restrict__ hxb, T *__restrict__ hyb, T *__restrict__ hzb, T cb, T dx,
          rf_cpml_t cp, rf_zpsi_t zs) {
    // blocks renumbered per XCD (xcd_block), as in every shell sweep
    shell_h_node<T, CPML>(gg, sh, (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x, ex, ey, ez,
                          hxa, hya, hza, hxb, hyb, hzb, cb, dx, cp, zs);
}

// The shell H updates of many scenes: block b (renumbered per XCD)
// updates nodes of the job block_job[b], from the job's first block on.
template <typename T, bool CPML>
__global__ void __launch_bounds__(256)
k_shell_h_packed(const rf_shell_job_t *__restrict__ jobs, const unsigned short *__restrict__ block_job) {
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const rf_shell_job_t &jb = jobs[block_job[b]];
    const T *const *fa = reinterpret_cast<const T *const *>(jb.fa);
    T *const *hb = reinterpret_cast<T *const *>(jb.hb);
    shell_h_node<T, CPML>(jb.g, jb.sh, (long long)(b - jb.block0) * blockDim.x + threadIdx.x, fa[0], fa[1], fa[2],
                          fa[3], fa[4], fa[5], hb[0], hb[1], hb[2], static_cast<T>(jb.cb), static_cast<T>(jb.dx),
                          jb.cp, jb.zs);
}

// E update of one component on a shell node (e_component of the
// two-kernel path, returning the value).
template <typename T, typename TagT>
__device__ __forceinline__ T shell_e_value(int c, T e_old, T dA, T dB, const rf_mat_t &m, unsigned tag, int i,
                                           int j, int k) {
    using Entry = typename EntryOf<T>::type;
    using Ade = typename EntryOf<T>::ade;
    if constexpr (std::is_same<TagT, NoTag>::value) {
        return T(1) * e_old + static_cast<T>(m.vac_ca) * (dA - dB);
    } else {
        // the entry read straight from the table: with one tag for every
        // active lane (nearly every wavefront of the shell) through the
        // scalar cache, no block-wide staging and barrier ahead of the nodes
        typedef const __attribute__((address_space(1))) Entry g_entry;
        const g_entry *ents = (const g_entry *)m.entries;
        const unsigned t0 = __builtin_amdgcn_readfirstlane(tag);
        uint32_t fl;
        T ca, cb;
        int aid;
        if (__ballot(tag != t0) == 0) {
            const g_entry &en = ents[t0];
            fl = en.flags[c];
            ca = en.Ca[c];
            cb = en.Cb[c];
            aid = en.ade[c];
        } else {
            const g_entry &en = ents[tag];
            fl = en.flags[c];
            ca = en.Ca[c];
            cb = en.Cb[c];
            aid = en.ade[c];
        }
        if (fl & RF_EDGE_TENSOR) return e_old;
        if (fl & RF_EDGE_ADE) {
            const Ade &ad = static_cast<const Ade *>(m.ade)[aid];
            return ade_edge<T>(ad, ca, cb, e_old, dA, dB, static_cast<T *>(m.J), static_cast<T *>(m.P), 3 * ad.st,
                               ade_off(ad, c, i, j, k));
        }
        return ca * e_old + cb * (dA - dB);
    }
}

// Sources and E^{n+1} of shell node t.
template <typename T, typename TagT, bool CPML>
__device__ __forceinline__ void shell_e_node(const rf_geom_t &gg, const rf_shell_t &sh, long long t,
                                             const T *__restrict__ exa_p, const T *__restrict__ eya_p,
                                             const T *__restrict__ eza_p, const T *__restrict__ hx_p,
                                             const T *__restrict__ hy_p, const T *__restrict__ hz_p,
                                             T *__restrict__ exb_p, T *__restrict__ eyb_p, T *__restrict__ ezb_p,
                                             T dx, const rf_cpml_t &cp, const rf_zpsi_t &zs, const rf_mat_t &m,
                                             const rf_leap_src_t *srcs, int n_srcs, const long long *d_step) {
    gmem_c<T> *const exa = (gmem_c<T> *)exa_p, *const eya = (gmem_c<T> *)eya_p, *const eza = (gmem_c<T> *)eza_p;
    gmem_c<T> *const hx = (gmem_c<T> *)hx_p, *const hy = (gmem_c<T> *)hy_p, *const hz = (gmem_c<T> *)hz_p;
    gmem<T> *const exb = (gmem<T> *)exb_p, *const eyb = (gmem<T> *)eyb_p, *const ezb = (gmem<T> *)ezb_p;
    const Geo g = geo_of(gg);
    int i, j, k;
    if (!shell_node(sh, t, i, j, k)) return;
    const long long idx = (long long)i * g.sx + (long long)j * g.sy + k;
    // E ranges (grid.c:95-117): own axis from 0, the other axes from 1
    // unless periodic (PEC planes at 0 and n are never updated)
    const bool in = i < g.nx && j < g.ny && k < g.nz;
    const bool i_ok = g.px || i >= 1, j_ok = g.py || j >= 1, k_ok = g.pz || k >= 1;
    const bool do_x = in && j_ok && k_ok, do_y = in && i_ok && k_ok, do_z = in && i_ok && j_ok;
    // -1 neighbours with wrap at 0 (materials.c:808-810)
    const long long di = (i == 0) ? (long long)(g.nx - 1) * g.sx : -g.sx;
    const long long dj = (j == 0) ? (long long)(g.ny - 1) * g.sy : -g.sy;
    const long long dk = (k == 0) ? (long long)(g.nz - 1) : -1;
    const long long step = *(gmem_c<long long> *)d_step;
    T e0 = exa[idx], e1 = eya[idx], e2 = eza[idx];
    // the sources are applied only by wavefronts with a node under one
    const bool src_here = n_srcs && __ballot(sources_touch(srcs, n_srcs, i, j, k, g.nx, g.ny, g.nz)) != 0;
    if (src_here) {
        e0 = apply_sources<T>(srcs, n_srcs, e0, 0, i, j, k, g.nx, g.ny, g.nz, step, 0);
        e1 = apply_sources<T>(srcs, n_srcs, e1, 1, i, j, k, g.nx, g.ny, g.nz, step, 0);
        e2 = apply_sources<T>(srcs, n_srcs, e2, 2, i, j, k, g.nx, g.ny, g.nz, step, 0);
    }
    unsigned tag = 0;
    if constexpr (!std::is_same<TagT, NoTag>::value) tag = load_tag<TagT>(m.tag, idx);
    // every load of the node is issued first, then the updates, then the
    // psi stores: a psi store placed between them would keep the next
    // component's loads behind it (the psi arrays may alias as far as the
    // compiler knows)
    const T hxc = hx[idx], hyc = hy[idx], hzc = hz[idx];
    T hz_j = T(0), hy_k = T(0), hx_k = T(0), hz_i = T(0), hy_i = T(0), hx_j = T(0);
    if (do_x) {
        hz_j = hz[idx + dj];
        hy_k = hy[idx + dk];
    }
    if (do_y) {
        hx_k = hx[idx + dk];
        hz_i = hz[idx + di];
    }
    if (do_z) {
        hy_i = hy[idx + di];
        hx_j = hx[idx + dj];
    }
    Term<T> ta[3] = {}, tb[3] = {};
    if (CPML) {
        if (do_x) {
            ta[0] = term_load<T>(cp, g, false, 1, j, cp.psi_e[0][0], i, j, k);
            tb[0] = zs.zpsi ? zterm_load<T>(cp, g, false, i, j, k, zs, 4, 4)
                            : term_load<T>(cp, g, false, 2, k, cp.psi_e[0][1], i, j, k);
        }
        if (do_y) {
            ta[1] = zs.zpsi ? zterm_load<T>(cp, g, false, i, j, k, zs, 5, 5)
                            : term_load<T>(cp, g, false, 2, k, cp.psi_e[1][0], i, j, k);
            tb[1] = term_load<T>(cp, g, false, 0, i, cp.psi_e[1][1], i, j, k);
        }
        if (do_z) {
            ta[2] = term_load<T>(cp, g, false, 0, i, cp.psi_e[2][0], i, j, k);
            tb[2] = term_load<T>(cp, g, false, 1, j, cp.psi_e[2][1], i, j, k);
        }
    }
    if (do_x) {   // materials.c:975-1000
        T dA = hzc - hz_j;
        T dB = hyc - hy_k;
        if (CPML) {
            dA = term_apply<T>(ta[0], dA, dx);
            dB = term_apply<T>(tb[0], dB, dx);
        }
        e0 = shell_e_value<T, TagT>(0, e0, dA, dB, m, tag, i, j, k);
    }
    if (do_y) {   // materials.c:1015-1040
        T dA = hxc - hx_k;
        T dB = hzc - hz_i;
        if (CPML) {
            dA = term_apply<T>(ta[1], dA, dx);
            dB = term_apply<T>(tb[1], dB, dx);
        }
        e1 = shell_e_value<T, TagT>(1, e1, dA, dB, m, tag, i, j, k);
    }
    if (do_z) {   // materials.c:1050-1058
        T dA = hyc - hy_i;
        T dB = hxc - hx_j;
        if (CPML) {
            dA = term_apply<T>(ta[2], dA, dx);
            dB = term_apply<T>(tb[2], dB, dx);
        }
        e2 = shell_e_value<T, TagT>(2, e2, dA, dB, m, tag, i, j, k);
    }
    if (CPML) {
        if (do_x) {
            term_store<T>(ta[0], cp.psi_e[0][0]);
            term_store<T>(tb[0], zs.zpsi ? zs.zpsi : cp.psi_e[0][1]);
        }
        if (do_y) {
            term_store<T>(ta[1], zs.zpsi ? zs.zpsi : cp.psi_e[1][0]);
            term_store<T>(tb[1], cp.psi_e[1][1]);
        }
        if (do_z) {
            term_store<T>(ta[2], cp.psi_e[2][0]);
            term_store<T>(tb[2], cp.psi_e[2][1]);
        }
    }
    if (src_here) {
        e0 = apply_sources<T>(srcs, n_srcs, e0, 0, i, j, k, g.nx, g.ny, g.nz, step, 1);
        e1 = apply_sources<T>(srcs, n_srcs, e1, 1, i, j, k, g.nx, g.ny, g.nz, step, 1);
        e2 = apply_sources<T>(srcs, n_srcs, e2, 2, i, j, k, g.nx, g.ny, g.nz, step, 1);
    }
    exb[idx] = e0;
    eyb[idx] = e1;
    ezb[idx] = e2;
}

template <typename T, typename TagT, bool CPML>
__global__ void __launch_bounds__(256)
k_shell_e(rf_geom_t gg, rf_shell_t sh, const T *__restrict__ exa, const T *__restrict__ eya,
          const T *__restrict__ eza, const T *__restrict__ hx, const T *__restrict__ hy, const T *__restrict__ hz,
          T *__restrict__ exb, T *__restrict__ eyb, T *__restrict__ ezb, T dx, rf_cpml_t cp, rf_zpsi_t zs,
          rf_mat_t m, const rf_leap_src_t *srcs, int n_srcs, const long long *d_step) {
    shell_e_node<T, TagT, CPML>(gg, sh, (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x, exa,
                                eya, eza, hx, hy, hz, exb, eyb, ezb, dx, cp, zs, m, srcs, n_srcs, d_step);
}

// The shell E updates of many scenes (block b, renumbered per XCD: job
// block_job[b]).
template <typename T, typename TagT, bool CPML>
__global__ void __launch_bounds__(256)
k_shell_e_packed(const rf_shell_job_t *__restrict__ jobs, const unsigned short *__restrict__ block_job) {
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const rf_shell_job_t &jb = jobs[block_job[b]];
    const T *const *fa = reinterpret_cast<const T *const *>(jb.fa);
    const T *const *hb = reinterpret_cast<const T *const *>(jb.hb);
    T *const *eb = reinterpret_cast<T *const *>(jb.eb);
    shell_e_node<T, TagT, CPML>(jb.g, jb.sh, (long long)(b - jb.block0) * blockDim.x + threadIdx.x, fa[0], fa[1],
                                fa[2], hb[0], hb[1], hb[2], eb[0], eb[1], eb[2], static_cast<T>(jb.dx), jb.cp, jb.zs,
                                jb.m, jb.srcs, jb.n_srcs, jb.d_step);
}

template <typename T>
int shell_h_t(const rf_geom_t *g, const rf_shell_t *sh, void *const fa[6], void *const hb[3], double cb,
              double dx, const rf_cpml_t *cp, const rf_zpsi_t *z, hipStream_t st) {
    const long long n = sh->start[sh->n];
    const dim3 grd((unsigned)((n + 255) / 256));
    rf_cpml_t c = cp ? *cp : rf_cpml_t{};
    rf_zpsi_t zz = z ? *z : rf_zpsi_t{};
#define RF_SHELL_H(CP)                                                                                         \
    hipLaunchKernelGGL((k_shell_h<T, CP>), grd, dim3(256), 0, st, *g, *sh, static_cast<const T *>(fa[0]),      \
                       static_cast<const T *>(fa[1]), static_cast<const T *>(fa[2]),                           \
                       static_cast<const T *>(fa[3]), static_cast<const T *>(fa[4]),                           \
                       static_cast<const T *>(fa[5]), static_cast<T *>(hb[0]), static_cast<T *>(hb[1]),        \
                       static_cast<T *>(hb[2]), static_cast<T>(cb), static_cast<T>(dx), c, zz)
    if (cp) RF_SHELL_H(true);
    else RF_SHELL_H(false);
#undef RF_SHELL_H
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, typename TagT>
int shell_e_tag(const rf_geom_t *g, const rf_shell_t *sh, void *const ea[3], void *const hb[3], void *const eb[3],
                double dx, const rf_cpml_t *cp, const rf_zpsi_t *z, const rf_mat_t *m, const rf_leap_src_t *srcs,
                int n_srcs, const long long *d_step, hipStream_t st) {
    const long long n = sh->start[sh->n];
    const dim3 grd((unsigned)((n + 255) / 256));
    rf_cpml_t c = cp ? *cp : rf_cpml_t{};
    rf_zpsi_t zz = z ? *z : rf_zpsi_t{};
#define RF_SHELL_E(CP)                                                                                         \
    hipLaunchKernelGGL((k_shell_e<T, TagT, CP>), grd, dim3(256), 0, st, *g, *sh,                              \
                       static_cast<const T *>(ea[0]), static_cast<const T *>(ea[1]),                           \
                       static_cast<const T *>(ea[2]), static_cast<const T *>(hb[0]),                           \
                       static_cast<const T *>(hb[1]), static_cast<const T *>(hb[2]), static_cast<T *>(eb[0]),  \
                       static_cast<T *>(eb[1]), static_cast<T *>(eb[2]), static_cast<T>(dx), c, zz, *m, srcs,  \
                       n_srcs, d_step)
    if (cp) RF_SHELL_E(true);
    else RF_SHELL_E(false);
#undef RF_SHELL_E
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int shell_e_t(const rf_geom_t *g, const rf_shell_t *sh, void *const ea[3], void *const hb[3], void *const eb[3],
              double dx, const rf_cpml_t *cp, const rf_zpsi_t *z, const rf_mat_t *m, const rf_leap_src_t *srcs,
              int n_srcs, const long long *d_step, hipStream_t st) {
    if (!m->tag) return shell_e_tag<T, NoTag>(g, sh, ea, hb, eb, dx, cp, z, m, srcs, n_srcs, d_step, st);
    switch (m->tag_bytes) {
    case 1: return shell_e_tag<T, uint8_t>(g, sh, ea, hb, eb, dx, cp, z, m, srcs, n_srcs, d_step, st);
    case 2: return shell_e_tag<T, uint16_t>(g, sh, ea, hb, eb, dx, cp, z, m, srcs, n_srcs, d_step, st);
    case 4: return shell_e_tag<T, uint32_t>(g, sh, ea, hb, eb, dx, cp, z, m, srcs, n_srcs, d_step, st);
    }
    return -1;
}

// ------------------------------------------------------------ ADE pass
// With the split ADE step the fused kernel gives every dispersive edge of
// its box the first two terms of the trapezoidal update,
// t = Ca e^n + Cb (dA - dB), and stores t as E^{n+1}; this pass completes
// it, E^{n+1} = t - Cu sum_u, and advances J and P.  These are the
// operations of ade_edge (yee_device.h) in the same order and precision,
// so E, J and P are bitwise those of the one-pass update.
//
// The pass covers the part of the ADE state box that lies in the fused
// box, k fastest, ADE_NPT nodes per thread.  A thread reads its nodes'
// tags first and then only what their dispersive edges use: E^n, t and,
// per pole of the edge's ADE set, J and (Lorentz poles) P.  The box holds
// non-dispersive nodes too (gaps between layers, the dielectric stripes of
// a grating), whose lines are then not fetched at all.
constexpr int ADE_PASS_MAX_SETS = 64;   // ADE sets staged in LDS
constexpr int ADE_NPT = 1;              // nodes per thread of the ADE pass (2 and 4 measured slower)

template <typename T, int NS>
__device__ __forceinline__ T ade_finish(const typename EntryOf<T>::ade &a, T t, T e_old, const T jv[NS],
                                        const T pv[NS], T nj[NS], T np[NS]) {
    T u[NS];
    T sum_u = T(0);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (s < a.n_poles) {
            u[s] = a.has_p[s] ? a.kj[s] * jv[s] - a.kp[s] * pv[s] : a.kj[s] * jv[s];
            sum_u += u[s];
        }
    }
    const T e_new = t - a.Cu * sum_u;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (s < a.n_poles) {
            const T S = u[s] + a.v[s] * (e_new + e_old);
            nj[s] = S - jv[s];
            np[s] = a.has_p[s] ? pv[s] + a.hdt * S : pv[s];
        }
    }
    return e_new;
}

// Block blk of the ADE pass over the box of count nodes at (lo0, lo1, lo2)
// with j, k extents nj, nk (the whole block takes part: it stages LDS).
template <typename T, typename TagT, int NS>
__device__ __forceinline__ void ade_pass_block(const rf_geom_t &gg, int lo0, int lo1, int lo2, unsigned nj,
                                               unsigned nk, unsigned count, const T *__restrict__ exa,
                                               const T *__restrict__ eya, const T *__restrict__ eza,
                                               T *__restrict__ exb, T *__restrict__ eyb, T *__restrict__ ezb,
                                               const rf_mat_t &m, int ncache, unsigned blk, unsigned n_blocks) {
    using Entry = typename EntryOf<T>::type;
    using Ade = typename EntryOf<T>::ade;
    // every ADE set and, for the first ncache entries of the material
    // table, the flags and ADE set ids of the three components (one byte
    // each), staged in LDS
    extern __shared__ __align__(16) unsigned char s_raw[];
    Ade *s_ade = reinterpret_cast<Ade *>(s_raw);
    uint32_t *s_ent = reinterpret_cast<uint32_t *>(s_raw + (size_t)m.n_ade * sizeof(Ade));
    const Entry *ents = static_cast<const Entry *>(m.entries);
    for (int e = threadIdx.x; e < ncache; e += blockDim.x) {
        const Entry &en = ents[e];
        s_ent[2 * e] = en.flags[0] | (en.flags[1] << 8) | (en.flags[2] << 16);
        s_ent[2 * e + 1] = (static_cast<uint32_t>(en.ade[0]) & 0xffu) |
                           ((static_cast<uint32_t>(en.ade[1]) & 0xffu) << 8) |
                           ((static_cast<uint32_t>(en.ade[2]) & 0xffu) << 16);
    }
    const Ade *ades = static_cast<const Ade *>(m.ade);
    for (int e = threadIdx.x; e < m.n_ade; e += blockDim.x) s_ade[e] = ades[e];
    __syncthreads();

    // ADE_NPT nodes per thread, ADE_NPT * (gridDim.x * blockDim.x) >= count
    // apart: the tags of all of them are loaded, then the state of all
    // their dispersive edges, then they are updated, so each round trip
    // carries several nodes' loads
    const Geo g = geo_of(gg);
    const unsigned stride = n_blocks * blockDim.x;
    const unsigned t0 = blk * blockDim.x + threadIdx.x;
    // device memory, accessed as global memory (the packed pass reads
    // these pointers from its job table)
    gmem_c<T> *const J = (gmem_c<T> *)m.J;
    gmem_c<T> *const P = (gmem_c<T> *)m.P;
    long long idx[ADE_NPT];
    unsigned tag[ADE_NPT];
    bool node[ADE_NPT];
    int ii[ADE_NPT], jj[ADE_NPT], kk[ADE_NPT];
#pragma unroll
    for (int u = 0; u < ADE_NPT; u++) {
        const unsigned t = t0 + static_cast<unsigned>(u) * stride;
        node[u] = t < count;
        const unsigned tt = node[u] ? t : 0u;
        const unsigned q = tt / nk;
        kk[u] = lo2 + static_cast<int>(tt - q * nk);
        jj[u] = lo1 + static_cast<int>(q % nj);
        ii[u] = lo0 + static_cast<int>(q / nj);
        idx[u] = (long long)ii[u] * g.sx + (long long)jj[u] * g.sy + kk[u];
        tag[u] = node[u] ? load_tag<TagT>(m.tag, idx[u]) : 0u;
    }
    bool need[ADE_NPT][3];
    uint32_t ai[ADE_NPT];
    bool any = false;
#pragma unroll
    for (int u = 0; u < ADE_NPT; u++) {
        uint32_t fl;
        if (static_cast<int>(tag[u]) < ncache) {
            fl = s_ent[2 * tag[u]];
            ai[u] = s_ent[2 * tag[u] + 1];
        } else {
            const Entry &en = ents[tag[u]];
            fl = en.flags[0] | (en.flags[1] << 8) | (en.flags[2] << 16);
            ai[u] = (static_cast<uint32_t>(en.ade[0]) & 0xffu) | ((static_cast<uint32_t>(en.ade[1]) & 0xffu) << 8) |
                    ((static_cast<uint32_t>(en.ade[2]) & 0xffu) << 16);
        }
        // E ranges of the update (grid.c:95-117)
        const int i = ii[u], j = jj[u], k = kk[u];
        const bool in = node[u] && i < g.nx && j < g.ny && k < g.nz;
        const bool i_ok = g.px || i >= 1, j_ok = g.py || j >= 1, k_ok = g.pz || k >= 1;
        const bool on[3] = {in && j_ok && k_ok, in && i_ok && k_ok, in && i_ok && j_ok};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const uint32_t f = (fl >> (8 * c)) & 0xffu;
            need[u][c] = on[c] && (f & RF_EDGE_ADE) && !(f & RF_EDGE_TENSOR);
            any = any || need[u][c];
        }
    }
    if (!any) return;
    // the loads of all dispersive edges of the nodes, then the updates
    gmem_c<T> *ea_[3] = {(gmem_c<T> *)exa, (gmem_c<T> *)eya, (gmem_c<T> *)eza};
    gmem<T> *eb[3] = {(gmem<T> *)exb, (gmem<T> *)eyb, (gmem<T> *)ezb};
    T eo[ADE_NPT][3], et[ADE_NPT][3], jv[ADE_NPT][3][NS], pv[ADE_NPT][3][NS];
    long long off[ADE_NPT][3];
#pragma unroll
    for (int u = 0; u < ADE_NPT; u++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            eo[u][c] = et[u][c] = T(0);
            off[u][c] = 0;
#pragma unroll
            for (int s = 0; s < NS; s++) jv[u][c][s] = pv[u][c][s] = T(0);
            if (!need[u][c]) continue;
            const Ade &ad = s_ade[(ai[u] >> (8 * c)) & 0xff];
            off[u][c] = ade_off(ad, c, ii[u], jj[u], kk[u]);
            eo[u][c] = ea_[c][idx[u]];
            et[u][c] = eb[c][idx[u]];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                if (s >= ad.n_poles) continue;
                const long long o = 3 * s * ad.st + off[u][c];
                jv[u][c][s] = J[o];
                if (ad.has_p[s]) pv[u][c][s] = P[o];
            }
        }
    gmem<T> *const Jw = (gmem<T> *)m.J;
    gmem<T> *const Pw = (gmem<T> *)m.P;
#pragma unroll
    for (int u = 0; u < ADE_NPT; u++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (!need[u][c]) continue;
            const Ade &ad = s_ade[(ai[u] >> (8 * c)) & 0xff];
            T nj_[NS], np_[NS];
            eb[c][idx[u]] = ade_finish<T, NS>(ad, et[u][c], eo[u][c], jv[u][c], pv[u][c], nj_, np_);
            const int npl = ad.n_poles;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                if (s >= npl) continue;
                const long long o = 3 * s * ad.st + off[u][c];
                Jw[o] = nj_[s];
                if (ad.has_p[s]) Pw[o] = np_[s];
            }
        }
}

template <typename T, typename TagT, int NS>
__global__ void __launch_bounds__(256)
k_ade_pass(rf_geom_t gg, int lo0, int lo1, int lo2, unsigned nj, unsigned nk, unsigned count,
           const T *__restrict__ exa, const T *__restrict__ eya, const T *__restrict__ eza, T *__restrict__ exb,
           T *__restrict__ eyb, T *__restrict__ ezb, rf_mat_t m, int ncache, unsigned ade_blocks, rf_caps_t cs,
           int cap_ring, const long long *d_step) {
    // the first gridDim.x - ade_blocks blocks capture the plane monitors of
    // the step (their planes lie outside the ADE box, so their values are
    // final): one-value gathers, dispatched first so that they overlap the
    // streaming ADE blocks instead of trailing them; the ADE blocks are
    // renumbered per XCD among themselves (xcd_block)
    const unsigned cap_blocks = gridDim.x - ade_blocks;
    if (blockIdx.x < cap_blocks) {
        plane_capture_elem<T>(gg, cs, (long long)blockIdx.x * blockDim.x + threadIdx.x, cap_ring, *d_step);
        return;
    }
    const unsigned b = xcd_block(blockIdx.x - cap_blocks, ade_blocks);
    ade_pass_block<T, TagT, NS>(gg, lo0, lo1, lo2, nj, nk, count, exa, eya, eza, exb, eyb, ezb, m, ncache, b,
                                ade_blocks);
}

// The ADE passes of many scenes: block b (renumbered per XCD) runs block
// b - block0 of the job block_job[b].
template <typename T, typename TagT, int NS>
__global__ void __launch_bounds__(256)
k_ade_pass_packed(const rf_ade_job_t *__restrict__ jobs, const unsigned short *__restrict__ block_job) {
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const rf_ade_job_t &jb = jobs[block_job[b]];
    const unsigned nb = (jb.count + 255u) / 256u;
    ade_pass_block<T, TagT, NS>(jb.g, jb.lo[0], jb.lo[1], jb.lo[2], jb.nj, jb.nk, jb.count,
                                static_cast<const T *>(jb.ea[0]), static_cast<const T *>(jb.ea[1]),
                                static_cast<const T *>(jb.ea[2]), static_cast<T *>(jb.eb[0]),
                                static_cast<T *>(jb.eb[1]), static_cast<T *>(jb.eb[2]), jb.m, jb.ncache,
                                b - jb.block0, nb);
}

template <typename T, typename TagT>
int ade_pass_tag(const rf_geom_t *g, const int lo[3], const int hi[3], void *const ea[3], void *const eb[3],
                 const rf_mat_t *m, const rf_caps_t &cs, int cap_ring, const long long *d_step, hipStream_t st) {
    const unsigned nj = static_cast<unsigned>(hi[1] - lo[1]), nk = static_cast<unsigned>(hi[2] - lo[2]);
    const unsigned count = static_cast<unsigned>(hi[0] - lo[0]) * nj * nk;
    const unsigned ade_blocks = (count + 256u * ADE_NPT - 1) / (256u * ADE_NPT);
    long long cap_total = 0;
    for (int q = 0; q < cs.n; q++) cap_total += (long long)cs.c[q].na * cs.c[q].nb;
    const dim3 grd(ade_blocks + static_cast<unsigned>((cap_total + 255) / 256));
    using Ade = typename EntryOf<T>::ade;
    const int ncache = m->n_entries < 4096 ? m->n_entries : 4096;
    const size_t lds = (size_t)m->n_ade * sizeof(Ade) + (size_t)ncache * 8;
#define RF_ADE_PASS(NS)                                                                                        \
    hipLaunchKernelGGL((k_ade_pass<T, TagT, NS>), grd, dim3(256), lds, st, *g, lo[0], lo[1], lo[2], nj, nk,    \
                       count, static_cast<const T *>(ea[0]), static_cast<const T *>(ea[1]),                    \
                       static_cast<const T *>(ea[2]), static_cast<T *>(eb[0]), static_cast<T *>(eb[1]),        \
                       static_cast<T *>(eb[2]), *m, ncache, ade_blocks, cs, cap_ring, d_step)
    if (m->n_slots <= 1) RF_ADE_PASS(1);
    else if (m->n_slots == 2) RF_ADE_PASS(2);
    else if (m->n_slots <= 4) RF_ADE_PASS(4);
    else RF_ADE_PASS(RF_MAX_SLOTS);
#undef RF_ADE_PASS
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int ade_pass_t(const rf_geom_t *g, const int lo[3], const int hi[3], void *const ea[3], void *const eb[3],
               const rf_mat_t *m, const rf_caps_t &cs, int cap_ring, const long long *d_step, hipStream_t st) {
    switch (m->tag_bytes) {
    case 1: return ade_pass_tag<T, uint8_t>(g, lo, hi, ea, eb, m, cs, cap_ring, d_step, st);
    case 2: return ade_pass_tag<T, uint16_t>(g, lo, hi, ea, eb, m, cs, cap_ring, d_step, st);
    case 4: return ade_pass_tag<T, uint32_t>(g, lo, hi, ea, eb, m, cs, cap_ring, d_step, st);
    }
    return -1;
}

template <typename T>
size_t leap_lds(const rf_leap_t *a) {
    const size_t nls = a->n_srcs <= MAX_LSRC ? (size_t)a->n_srcs : 0;
    return nls * sizeof(LSrc) + (size_t)a->n_adec * sizeof(typename EntryOf<T>::ade) +
           (size_t)a->ncache * coef_bytes<T>();
}

// Wavefronts per block: WPB in independent mode, gz * gy cooperatively.
inline int block_waves(const rf_leap_t *a) { return a->gz > 0 ? a->gz * a->gy : WPB; }

template <typename T, typename TagT, int R>
int launch_tag(const rf_leap_t *a, unsigned blocks, hipStream_t st) {
    const size_t lds = leap_lds<T>(a);
    const unsigned threads = static_cast<unsigned>(WL * block_waves(a));
    if (a->gz > 0) {
        if (a->zc)
            hipLaunchKernelGGL((k_leapfrog<T, TagT, true, R, true>), dim3(blocks), dim3(threads), lds, st, *a);
        else
            hipLaunchKernelGGL((k_leapfrog<T, TagT, false, R, true>), dim3(blocks), dim3(threads), lds, st, *a);
    } else {
        if (a->zc)
            hipLaunchKernelGGL((k_leapfrog<T, TagT, true, R, false>), dim3(blocks), dim3(threads), lds, st, *a);
        else
            hipLaunchKernelGGL((k_leapfrog<T, TagT, false, R, false>), dim3(blocks), dim3(threads), lds, st, *a);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Wavefronts of the variant `a` selects that the device holds at once.
template <typename T, typename TagT, int R>
int capacity_tag(const rf_leap_t *a) {
    int nb = 0, cus = 0, dev = 0;
    const size_t lds = leap_lds<T>(a);
    const int bw = block_waves(a);
    const void *fn = a->gz > 0 ? (a->zc ? reinterpret_cast<const void *>(k_leapfrog<T, TagT, true, R, true>)
                                        : reinterpret_cast<const void *>(k_leapfrog<T, TagT, false, R, true>))
                               : (a->zc ? reinterpret_cast<const void *>(k_leapfrog<T, TagT, true, R, false>)
                                        : reinterpret_cast<const void *>(k_leapfrog<T, TagT, false, R, false>));
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WL * bw, lds);
    if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || nb <= 0)
        return 0;
    return nb * cus * bw;
}

template <typename T, int R>
int capacity_rows(const rf_leap_t *a) {
    if (!a->tag) return capacity_tag<T, NoTag, R>(a);
    switch (a->tag_bytes) {
    case 1: return capacity_tag<T, uint8_t, R>(a);
    case 2: return capacity_tag<T, uint16_t, R>(a);
    case 4: return capacity_tag<T, uint32_t, R>(a);
    }
    return 0;
}

int leap_capacity(int prec, const rf_leap_t *a) {
    if (prec == 0) return a->rows == 2 ? capacity_rows<double, 2>(a) : capacity_rows<double, 4>(a);
    return a->rows == 4 ? capacity_rows<float, 4>(a) : capacity_rows<float, 8>(a);
}

// Row mode: the kernel of a's tag width, rows and chunk count (fewer than
// four chunks: vacuum and one-byte tags only).
template <int R, int NCH>
const void *zrow_fn_nch(const rf_leap_t *a) {
    if (!a->tag) return reinterpret_cast<const void *>(k_leap_rows<NoTag, R, NCH>);
    if (a->tag_bytes == 1) return reinterpret_cast<const void *>(k_leap_rows<uint8_t, R, NCH>);
    if constexpr (NCH == ROW_NCH_MAX) {
        if (a->tag_bytes == 2) return reinterpret_cast<const void *>(k_leap_rows<uint16_t, R, NCH>);
        if (a->tag_bytes == 4) return reinterpret_cast<const void *>(k_leap_rows<uint32_t, R, NCH>);
    }
    return nullptr;
}

int zrow_chunks(const rf_leap_t *a) { return (a->g.n[2] + 63) / 64; }

const void *zrow_fn(const rf_leap_t *a) {
    if (a->rows != 1) return nullptr;
    switch (zrow_chunks(a)) {
    case 1: return zrow_fn_nch<1, 1>(a);
    case 2: return zrow_fn_nch<1, 2>(a);
    case 3: return zrow_fn_nch<1, 3>(a);
    case 4: return zrow_fn_nch<1, 4>(a);
    }
    return nullptr;
}

int zrow_capacity(const rf_leap_t *a) {
    const void *fn = zrow_fn(a);
    int nb = 0, cus = 0, dev = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WL * WPB, leap_lds<float>(a)) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || nb <= 0)
        return 0;
    return nb * cus * WPB;
}

int launch_zrow(const rf_leap_t *a, unsigned blocks, hipStream_t st) {
    const void *fn = zrow_fn(a);
    if (!fn) return -1;
    void *args[] = {const_cast<rf_leap_t *>(a)};
    if (hipLaunchKernel(fn, dim3(blocks), dim3(WL * WPB), args, leap_lds<float>(a), st) != hipSuccess) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, int R>
int launch_rows(const rf_leap_t *a, unsigned blocks, hipStream_t st) {
    if (!a->tag) return launch_tag<T, NoTag, R>(a, blocks, st);
    switch (a->tag_bytes) {
    case 1: return launch_tag<T, uint8_t, R>(a, blocks, st);
    case 2: return launch_tag<T, uint16_t, R>(a, blocks, st);
    case 4: return launch_tag<T, uint32_t, R>(a, blocks, st);
    }
    return -1;
}

// ------------------------------------------------ packed-scene launchers
template <typename T, typename TagT, int R>
const void *packed_fn(int zc) {
    return zc ? reinterpret_cast<const void *>(k_leapfrog_packed<T, TagT, true, R>)
              : reinterpret_cast<const void *>(k_leapfrog_packed<T, TagT, false, R>);
}

template <typename T, int R>
const void *packed_fn_rows(int tag_bytes, int zc) {
    switch (tag_bytes) {
    case 0: return packed_fn<T, NoTag, R>(zc);
    case 1: return packed_fn<T, uint8_t, R>(zc);
    case 2: return packed_fn<T, uint16_t, R>(zc);
    case 4: return packed_fn<T, uint32_t, R>(zc);
    }
    return nullptr;
}

const void *packed_kernel(int prec, const rf_leap_pack_t *pk) {
    if (prec == 0) {
        if (pk->rows == 2) return packed_fn_rows<double, 2>(pk->tag_bytes, pk->zc);
        if (pk->rows == 4) return packed_fn_rows<double, 4>(pk->tag_bytes, pk->zc);
    } else {
        if (pk->rows == 4) return packed_fn_rows<float, 4>(pk->tag_bytes, pk->zc);
        if (pk->rows == 8) return packed_fn_rows<float, 8>(pk->tag_bytes, pk->zc);
    }
    return nullptr;
}

template <typename T, typename TagT, bool CPML>
void shell_e_packed_launch(const rf_shell_job_t *jobs, const unsigned short *bj, int n_blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_shell_e_packed<T, TagT, CPML>), dim3(n_blocks), dim3(256), 0, st, jobs, bj);
}

template <typename T, bool CPML>
int shell_e_packed_tag(const rf_shell_job_t *jobs, const unsigned short *bj, int n_blocks, int tag_bytes,
                       hipStream_t st) {
    switch (tag_bytes) {
    case 0: shell_e_packed_launch<T, NoTag, CPML>(jobs, bj, n_blocks, st); break;
    case 1: shell_e_packed_launch<T, uint8_t, CPML>(jobs, bj, n_blocks, st); break;
    case 2: shell_e_packed_launch<T, uint16_t, CPML>(jobs, bj, n_blocks, st); break;
    case 4: shell_e_packed_launch<T, uint32_t, CPML>(jobs, bj, n_blocks, st); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

extern "C" const char *rf_leap_zrow_refusal(const rf_leap_t *a) {
    const int nz = a->g.n[2];
    if (!a->zc || ((a->g.per >> 2) & 1) || a->ib0[2] != 0 || a->ib1[2] != nz + 1)
        return "the box does not span every z column with the z CPML inside it";
    const int nch = zrow_chunks(a);
    if (nz < 1 || nch > ROW_NCH_MAX) return "nz is above 256";
    if (a->zplo > 64 || a->zlo > 64 || a->zphi < 64 * (nch - 1) || a->zhi < 64 * (nch - 1))
        return "a z CPML slab reaches past the first or last 64 columns";
    if (!zrow_fn(a)) return "no row-mode kernel for this tag width and nz";
    if (a->tag && a->n_slots > 0 && !a->ade_split) return "dispersive edges are updated in the fused kernel";
    if (nz == 64 * nch && a->n_srcs > 0 && a->src_box[5] >= nz) return "a source may touch column nz";
    if (a->coop || a->zsplit) return "cooperative blocks or a z split were requested";
    return nullptr;
}

extern "C" int rf_leap_plan(int prec, rf_leap_t *a, int rows, int chunk, int target_waves) {
    for (int q = 0; q < 3; q++) {
        const int top = ((a->g.per >> q) & 1) ? a->g.n[q] : a->g.n[q] + 1;
        if (a->ib0[q] < 0 || a->ib1[q] > top) return -1;
    }
    const long long planes = a->ib1[0] - a->ib0[0];
    const long long jn = a->ib1[1] - a->ib0[1];
    const long long kn = a->ib1[2] - a->ib0[2];
    if (a->kbrk < a->ib0[2] || a->kbrk > a->ib1[2]) return -1;
    if (rows <= 0) rows = 4;
    if (prec == 0 ? !(rows == 2 || rows == 4) : !(rows == 1 || rows == 4 || rows == 8)) return -1;
    a->rows = rows;
    a->zrow = prec == 1 && rows == 1;
    if (a->zrow && rf_leap_zrow_refusal(a)) return -1;
    // up to 40 KiB of LDS per block for the ADE sets (at most 16) and the
    // material table (plus the staged sources)
    const size_t ade_sz = prec == 0 ? sizeof(rf_ade64_t) : sizeof(rf_ade32_t);
    a->n_adec = a->tag ? (a->n_ade < 16 ? a->n_ade : 16) : 0;
    const int cap = (int)((40 * 1024 - a->n_adec * ade_sz) / (prec == 0 ? coef_bytes<double>() : coef_bytes<float>()));
    a->ncache = a->tag ? (a->n_entries < cap ? a->n_entries : cap) : 0;
    if (planes <= 0 || jn <= 0 || kn <= 0) {   /* empty box: nothing to launch */
        a->ntk = a->ntk0 = a->ntj = a->n_parts = a->n_items = 0;
        a->chunk = 1;
        return 0;
    }
    a->ntj = (int)((jn + rows - 1) / rows);
    if (a->zrow) {   // one tile of every z column per row tile, one part
        a->ntk = a->ntk0 = 1;
        a->n_parts = 1;
        a->ptk[0][0] = 0;
        a->ptk[0][1] = 1;
        a->ptk[0][2] = a->ptk[0][3] = 0;
        a->pzc[0] = 1;
        a->pgz[0] = a->pgy[0] = 0;
        a->gk = 0;
        int resident = target_waves > 0 ? target_waves : zrow_capacity(a);
        if (resident <= 0) resident = 4096;
        if (chunk <= 0) {
            // two items per resident wavefront: every wavefront of the
            // device has an item in both rounds (one round leaves more
            // planes at the tail; the 512 x 512 x 256 dispersive stack
            // measured 1109 us with one, 1074 us with two, 1082 us with four)
            long long chunks = (2LL * resident + a->ntj - 1) / a->ntj;
            if (chunks < 1) chunks = 1;
            chunk = (int)((planes + chunks - 1) / chunks);
            if (chunk < 8) chunk = 8;   // small boxes: each item recomputes a warm-up plane
        }
        if (chunk > planes) chunk = (int)planes;
        a->chunk = chunk;
        const long long items = (long long)a->ntj * ((planes + chunk - 1) / chunk);
        if (items > (1LL << 30)) return -1;
        a->pitems[0] = (int)items;
        long long blocks = (resident + WPB - 1) / WPB;
        if (blocks > (items + WPB - 1) / WPB) blocks = (items + WPB - 1) / WPB;
        a->pblocks[0] = (int)((blocks + N_XCD - 1) / N_XCD * N_XCD);
        return 0;
    }
    a->ntk0 = (a->kbrk - a->ib0[2] + KW - 1) / KW;
    a->ntk = a->ntk0 + (a->ib1[2] - a->kbrk + KW - 1) / KW;
    // z tiles whose halo lane or owned lanes lie in the z CPML (outside the
    // plain range) need the z CPML code; with the slabs at both ends of the
    // z range they are a prefix [0, zt0) and a suffix [zt1, ntk)
    int zt0 = 0, zt1 = a->ntk;
    if (a->zc && !a->zsplit) zt0 = zt1 = 0;   // one part with the z CPML code
    if (a->zc && a->zsplit) {
        const int nz = a->g.n[2];
        auto slab_tile = [&](int tk) {
            const int k0 = tk < a->ntk0 ? a->ib0[2] + tk * KW : a->kbrk + (tk - a->ntk0) * KW;
            const int kend = tk < a->ntk0 ? (k0 + KW < a->kbrk ? k0 + KW : a->kbrk) : a->ib1[2];
            for (int kt = k0 - 1; kt < kend; kt++) {
                const int kc = kt < 0 ? 0 : (kt > nz ? nz : kt);
                if (kc < a->zplo || kc >= a->zphi) return true;
            }
            return false;
        };
        while (zt0 < a->ntk && slab_tile(zt0)) zt0++;
        while (zt1 > zt0 && slab_tile(zt1 - 1)) zt1--;
        for (int tk = zt0; tk < zt1; tk++)
            if (slab_tile(tk)) zt0 = zt1 = 0;   // not of that shape: one part with the z CPML code
    }
    const long long tiles = (long long)a->ntk * a->ntj;
    int np = 0;
    if (zt1 > zt0) {   // tiles without z CPML lanes
        a->ptk[np][0] = zt0;
        a->ptk[np][1] = zt1 - zt0;
        a->ptk[np][2] = a->ptk[np][3] = 0;
        a->pzc[np++] = 0;
    }
    if (zt0 > 0 || zt1 < a->ntk) {
        a->ptk[np][0] = 0;
        a->ptk[np][1] = zt0;
        a->ptk[np][2] = zt1;
        a->ptk[np][3] = a->ntk - zt1;
        a->pzc[np++] = 1;
    }
    a->n_parts = np;
    // block shape of each part: cooperative blocks take coop & 15 (at most
    // WPB_MAX) neighbouring row tiles of one z tile
    if (a->coop < 0) return -1;
    a->gk = a->coop >> 4;
    for (int p = 0; p < np; p++) {
        if (a->coop == 0) {
            a->pgz[p] = a->pgy[p] = 0;
            continue;
        }
        const int gz = 1;
        int gy = a->coop & 15;
        if (gy < 1) gy = 1;
        if (gy > WPB_MAX / gz) gy = WPB_MAX / gz;
        if (gy > a->ntj) gy = a->ntj;
        a->pgz[p] = gz;
        a->pgy[p] = gy;
    }
    // resident wavefronts of each part's variant and block shape
    static long long cache_key[8];
    static int cache_val[8], cache_n = 0;
    const size_t lds = prec == 0 ? leap_lds<double>(a) : leap_lds<float>(a);
    int resident[2] = {0, 0};
    for (int p = 0; p < np; p++) {
        const int bw = a->pgz[p] > 0 ? a->pgz[p] * a->pgy[p] : 0;
        const long long key = ((long long)lds << 20) | (bw << 13) | (prec << 12) | (a->tag_bytes << 8) |
                              (a->pzc[p] << 7) | (rows << 3);
        for (int q = 0; q < cache_n; q++)
            if (cache_key[q] == key) resident[p] = cache_val[q];
        if (!resident[p]) {
            rf_leap_t v = *a;
            v.zc = a->pzc[p];
            v.gz = a->pgz[p];
            v.gy = a->pgy[p];
            resident[p] = leap_capacity(prec, &v);
            if (resident[p] <= 0) resident[p] = 4096;
            if (cache_n < 8) {
                cache_key[cache_n] = key;
                cache_val[cache_n++] = resident[p];
            }
        }
        if (target_waves > 0) resident[p] = target_waves;
    }
    if (chunk <= 0) {
        // about twelve items per resident wavefront, at least 8 planes
        // each: short items keep the wavefronts of an XCD on neighbouring
        // tiles at nearby planes (their halo rows meet in its L2) and leave
        // a short tail when the queue drains; each item recomputes one
        // warm-up plane
        long long chunks = (12LL * resident[0] + tiles - 1) / tiles;
        if (chunks < 1) chunks = 1;
        chunk = (int)((planes + chunks - 1) / chunks);
        if (chunk < 8) chunk = 8;
    }
    if (chunk > planes) chunk = (int)planes;
    a->chunk = chunk;
    const long long nch = (planes + chunk - 1) / chunk;
    for (int p = 0; p < np; p++) {
        const long long ntl = a->ptk[p][1] + a->ptk[p][3];
        long long items, bw;
        if (a->pgz[p] > 0) {   // groups of gz x gy tiles
            items = (ntl + a->pgz[p] - 1) / a->pgz[p] * ((a->ntj + a->pgy[p] - 1) / a->pgy[p]) * nch;
            bw = a->pgz[p] * a->pgy[p];
        } else {
            items = ntl * a->ntj * nch;
            bw = WPB;
        }
        if (items > (1LL << 30)) return -1;
        a->pitems[p] = (int)items;
        long long blocks = (resident[p] + bw - 1) / bw;
        const long long need = a->pgz[p] > 0 ? items : (items + bw - 1) / bw;
        if (blocks > need) blocks = need;
        a->pblocks[p] = (int)((blocks + N_XCD - 1) / N_XCD * N_XCD);
    }
    return 0;
}

extern "C" int rf_launch_leapfrog(int prec, const rf_leap_t *a, rf_stream_t *s) {
    if (a->n_parts == 0) return 0;
    if (a->n_parts < 0 || a->n_parts > 2 || a->ntk <= 0 || a->ntj <= 0 || a->chunk <= 0 || !a->ctr) return -1;
    if (a->n_srcs > 0 && !a->srcs) return -1;
    if (a->tag && !a->cold) return -1;
    if (a->zc && (!a->zpsi || a->zlen <= 0)) return -1;
    hipStream_t st = rf_raw_stream(s);
    for (int p = 0; p < a->n_parts; p++) {
        const int *r = a->ptk[p];
        if (a->pitems[p] <= 0 || a->pblocks[p] <= 0 || r[1] < 0 || r[3] < 0 || r[1] + r[3] <= 0 || r[0] < 0 ||
            r[2] < 0 || r[0] + r[1] > a->ntk || r[2] + r[3] > a->ntk || (a->pzc[p] && !a->zc))
            return -1;
        rf_leap_t b = *a;
        for (int q = 0; q < 4; q++) b.tkr[q] = r[q];
        b.zc = a->pzc[p];
        b.n_items = a->pitems[p];
        b.gz = a->pgz[p];
        b.gy = a->pgy[p];
        if (b.gz < 0 || b.gy < 0 || b.gz * b.gy > WPB_MAX || (b.gz > 0 && b.gy < 1)) return -1;
        b.ctr = a->ctr + 16 * p;
        const unsigned blocks = static_cast<unsigned>(a->pblocks[p]);
        int rc = -1;
        if (a->zrow) {
            if (prec != 1 || a->n_parts != 1 || b.gz != 0) return -1;
            rc = launch_zrow(&b, blocks, st);
        } else if (prec == 0) {
            switch (a->rows) {
            case 2: rc = launch_rows<double, 2>(&b, blocks, st); break;
            case 4: rc = launch_rows<double, 4>(&b, blocks, st); break;
            }
        } else {
            switch (a->rows) {
            case 4: rc = launch_rows<float, 4>(&b, blocks, st); break;
            case 8: rc = launch_rows<float, 8>(&b, blocks, st); break;
            }
        }
        if (rc) return rc;
    }
    return 0;
}

extern "C" int rf_launch_shell_h(int prec, const rf_geom_t *g, const rf_shell_t *sh, void *const fa[6],
                                 void *const hb[3], double cb, double dx, const rf_cpml_t *cpml,
                                 const rf_zpsi_t *z, rf_stream_t *s) {
    if (sh->n < 0 || sh->n > 6) return -1;
    if (sh->n == 0 || sh->start[sh->n] == 0) return 0;
    if (z && z->zpsi && !cpml) return -1;
    hipStream_t st = rf_raw_stream(s);
    return prec == 0 ? shell_h_t<double>(g, sh, fa, hb, cb, dx, cpml, z, st)
                     : shell_h_t<float>(g, sh, fa, hb, cb, dx, cpml, z, st);
}

extern "C" int rf_launch_ade_pass(int prec, const rf_geom_t *g, const int lo[3], const int hi[3], void *const ea[3],
                                  void *const eb[3], const rf_mat_t *mat, const rf_caps_t *caps, int cap_ring,
                                  const long long *d_step, rf_stream_t *s) {
    rf_caps_t cs;
    cs.n = 0;
    if (caps && caps->n > 0) {
        if (caps->n > RF_MAX_CAPS || cap_ring <= 0 || !d_step) return -1;
        cs = *caps;
        long long total = 0;
        for (int q = 0; q < cs.n; q++) {
            if (!cs.c[q].f || !cs.c[q].buf || cs.c[q].na <= 0 || cs.c[q].nb <= 0 || cs.c[q].axis < 0 ||
                cs.c[q].axis > 2)
                return -1;
            cs.c[q].start = total;
            total += (long long)cs.c[q].na * cs.c[q].nb;
        }
        if (total >= (1LL << 31)) return -1;
    }
    long long count = 1;
    for (int q = 0; q < 3; q++) {
        if (hi[q] < lo[q] || lo[q] < 0) return -1;
        count *= hi[q] - lo[q];
    }
    if (count == 0) return 0;
    if (count >= (1LL << 31) || !mat->tag || !mat->J || mat->n_slots < 1 || mat->n_slots > RF_MAX_SLOTS ||
        mat->n_ade < 1 || mat->n_ade > ADE_PASS_MAX_SETS)
        return -1;
    hipStream_t st = rf_raw_stream(s);
    return prec == 0 ? ade_pass_t<double>(g, lo, hi, ea, eb, mat, cs, cap_ring, d_step, st)
                     : ade_pass_t<float>(g, lo, hi, ea, eb, mat, cs, cap_ring, d_step, st);
}

extern "C" int rf_launch_shell_e(int prec, const rf_geom_t *g, const rf_shell_t *sh, void *const ea[3],
                                 void *const hb[3], void *const eb[3], double dx, const rf_cpml_t *cpml,
                                 const rf_zpsi_t *z, const rf_mat_t *mat, const rf_leap_src_t *srcs, int n_srcs,
                                 const long long *d_step, rf_stream_t *s) {
    if (sh->n < 0 || sh->n > 6) return -1;
    if (sh->n == 0 || sh->start[sh->n] == 0) return 0;
    if (n_srcs > 0 && !srcs) return -1;
    if (z && z->zpsi && !cpml) return -1;
    hipStream_t st = rf_raw_stream(s);
    return prec == 0 ? shell_e_t<double>(g, sh, ea, hb, eb, dx, cpml, z, mat, srcs, n_srcs, d_step, st)
                     : shell_e_t<float>(g, sh, ea, hb, eb, dx, cpml, z, mat, srcs, n_srcs, d_step, st);
}

extern "C" size_t rf_leap_lds_bytes(int prec, const rf_leap_t *a) {
    const size_t b = prec == 0 ? leap_lds<double>(a) : leap_lds<float>(a);
    return (b + 15) / 16 * 16;
}

extern "C" int rf_leap_pack_plan(int prec, rf_leap_pack_t *pk) {
    const void *fn = packed_kernel(prec, pk);
    if (!fn || pk->n_items <= 0 || pk->lds_wave < 0 || pk->lds_wave % 16) return -1;
    int nb = 0, cus = 0, dev = 0;
    const size_t lds = (size_t)pk->lds_wave * WPB;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WL * WPB, lds) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || nb <= 0)
        return -1;
    long long blocks = (long long)nb * cus;
    const long long need = (pk->n_items + WPB - 1) / WPB;
    if (blocks > need) blocks = need;
    pk->blocks = (int)((blocks + N_XCD - 1) / N_XCD * N_XCD);
    return 0;
}

extern "C" int rf_launch_leapfrog_packed(int prec, const rf_leap_pack_t *pk, rf_stream_t *s) {
    if (pk->n_items <= 0) return 0;
    const void *fn = packed_kernel(prec, pk);
    if (!fn || !pk->tab || !pk->item_base || !pk->item_scene || !pk->ctr || pk->blocks <= 0 ||
        pk->n_scenes <= 0 || pk->n_scenes > 65535 || pk->qpar < 0 || pk->qpar > 1)
        return -1;
    const size_t lds = (size_t)pk->lds_wave * WPB;
    void *args[] = {const_cast<rf_leap_t **>(&pk->tab), const_cast<int **>(&pk->item_base),
                    const_cast<unsigned short **>(&pk->item_scene), const_cast<int *>(&pk->n_items),
                    const_cast<unsigned **>(&pk->ctr), const_cast<int *>(&pk->qpar),
                    const_cast<int *>(&pk->lds_wave)};
    if (hipLaunchKernel(fn, dim3(pk->blocks), dim3(WL * WPB), args, lds, rf_raw_stream(s)) != hipSuccess) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_shell_h_packed(int prec, const rf_shell_job_t *jobs, const unsigned short *block_job,
                                        int n_blocks, int has_cpml, rf_stream_t *s) {
    if (n_blocks <= 0) return 0;
    if (!jobs || !block_job) return -1;
    hipStream_t st = rf_raw_stream(s);
    if (prec == 0) {
        if (has_cpml) hipLaunchKernelGGL((k_shell_h_packed<double, true>), dim3(n_blocks), dim3(256), 0, st, jobs, block_job);
        else hipLaunchKernelGGL((k_shell_h_packed<double, false>), dim3(n_blocks), dim3(256), 0, st, jobs, block_job);
    } else {
        if (has_cpml) hipLaunchKernelGGL((k_shell_h_packed<float, true>), dim3(n_blocks), dim3(256), 0, st, jobs, block_job);
        else hipLaunchKernelGGL((k_shell_h_packed<float, false>), dim3(n_blocks), dim3(256), 0, st, jobs, block_job);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_shell_e_packed(int prec, const rf_shell_job_t *jobs, const unsigned short *block_job,
                                        int n_blocks, int tag_bytes, int has_cpml, rf_stream_t *s) {
    if (n_blocks <= 0) return 0;
    if (!jobs || !block_job) return -1;
    hipStream_t st = rf_raw_stream(s);
    if (prec == 0)
        return has_cpml ? shell_e_packed_tag<double, true>(jobs, block_job, n_blocks, tag_bytes, st)
                        : shell_e_packed_tag<double, false>(jobs, block_job, n_blocks, tag_bytes, st);
    return has_cpml ? shell_e_packed_tag<float, true>(jobs, block_job, n_blocks, tag_bytes, st)
                    : shell_e_packed_tag<float, false>(jobs, block_job, n_blocks, tag_bytes, st);
}

// Pole-count class of the ADE pass variant: 1, 2, 4 or RF_MAX_SLOTS.
extern "C" int rf_ade_pass_class(int n_slots) {
    return n_slots <= 1 ? 1 : n_slots == 2 ? 2 : n_slots <= 4 ? 4 : RF_MAX_SLOTS;
}

extern "C" int rf_launch_ade_pass_packed(int prec, const rf_ade_job_t *jobs, const unsigned short *block_job,
                                         int n_blocks, int tag_bytes, int ns_class, size_t lds, rf_stream_t *s) {
    if (n_blocks <= 0) return 0;
    if (!jobs || !block_job) return -1;
    hipStream_t st = rf_raw_stream(s);
#define RF_ADE_PACKED(T, TT, NS) \
    hipLaunchKernelGGL((k_ade_pass_packed<T, TT, NS>), dim3(n_blocks), dim3(256), lds, st, jobs, block_job)
#define RF_ADE_PACKED_NS(T, TT)                  \
    switch (ns_class) {                          \
    case 1: RF_ADE_PACKED(T, TT, 1); break;      \
    case 2: RF_ADE_PACKED(T, TT, 2); break;      \
    case 4: RF_ADE_PACKED(T, TT, 4); break;      \
    default: RF_ADE_PACKED(T, TT, RF_MAX_SLOTS); \
    }
#define RF_ADE_PACKED_TAG(T)                                 \
    switch (tag_bytes) {                                     \
    case 1: RF_ADE_PACKED_NS(T, uint8_t); break;             \
    case 2: RF_ADE_PACKED_NS(T, uint16_t); break;            \
    case 4: RF_ADE_PACKED_NS(T, uint32_t); break;            \
    default: return -1;                                      \
    }
    if (prec == 0) {
        RF_ADE_PACKED_TAG(double)
    } else {
        RF_ADE_PACKED_TAG(float)
    }
#undef RF_ADE_PACKED_TAG
#undef RF_ADE_PACKED_NS
#undef RF_ADE_PACKED
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
