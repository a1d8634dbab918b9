// Scene painter kernels (include/fdtd_scene.h): node ids, the edges whose
// dual cell may cross an interface, and the sub-sampled classification of
// those edges.  Every test and every coefficient is evaluated in FP64 with
// the same operations in the same order as the host painting rules state
// them (built with -ffp-contract=off; sqrt and division are correctly
// rounded on the device as on the host), so the painted materials are the
// same bits whichever side computes them.  The host (scene.c) applies the
// results through the materials API in the host painter's order.
//
// Shapes sit in global memory (at most 250, read uniformly by a wavefront).
// Coordinates are in cell units, node (i,j,k) at (i,j,k).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "rf_device.h"
#include "rf_hip.h"

namespace {

// Point x inside shape s (sphere: |x - p| <= r; slab: inside the clip box
// and |n.(x - p)| <= half thickness; box: inside [lo, hi]).
__device__ bool shape_inside(const rf_paint_shape_t &s, double x0, double x1, double x2) {
    if (s.kind == 0) {
        const double d0 = x0 - s.p[0], d1 = x1 - s.p[1], d2 = x2 - s.p[2];
        return d0 * d0 + d1 * d1 + d2 * d2 <= s.radius * s.radius;
    }
    if (x0 < s.lo[0] || x0 > s.hi[0] || x1 < s.lo[1] || x1 > s.hi[1] || x2 < s.lo[2] || x2 > s.hi[2]) return false;
    if (s.kind == 2) return true;
    const double d = s.nrm[0] * (x0 - s.p[0]) + s.nrm[1] * (x1 - s.p[1]) + s.nrm[2] * (x2 - s.p[2]);
    return fabs(d) <= s.half_thickness;
}

// x within one cell of the shape's bounding box on every axis.
__device__ bool near_bbox(const rf_paint_shape_t &s, double x0, double x1, double x2) {
    return !(x0 < s.bb_lo[0] - 1.0 || x0 > s.bb_hi[0] + 1.0 || x1 < s.bb_lo[1] - 1.0 || x1 > s.bb_hi[1] + 1.0 ||
             x2 < s.bb_lo[2] - 1.0 || x2 > s.bb_hi[2] + 1.0);
}

// The unit cube centred at x may contain a boundary of the shape: the
// sphere's surface within one cell, a clip face within one cell, or a
// slab face within one cell along the normal.
__device__ bool shape_near_boundary(const rf_paint_shape_t &s, double x0, double x1, double x2) {
    if (!near_bbox(s, x0, x1, x2)) return false;
    if (s.kind == 0) {
        const double d0 = x0 - s.p[0], d1 = x1 - s.p[1], d2 = x2 - s.p[2];
        return fabs(sqrt(d0 * d0 + d1 * d1 + d2 * d2) - s.radius) <= 1.0;
    }
    const double x[3] = {x0, x1, x2};
    for (int a = 0; a < 3; a++)
        if (fabs(x[a] - s.lo[a]) <= 1.0 || fabs(x[a] - s.hi[a]) <= 1.0) return true;
    if (s.kind == 2) return false;
    const double d = s.nrm[0] * (x0 - s.p[0]) + s.nrm[1] * (x1 - s.p[1]) + s.nrm[2] * (x2 - s.p[2]);
    return fabs(fabs(d) - s.half_thickness) <= 1.0;
}

// Surface normal of the shape at x: the sphere's radial direction (+z at
// the centre), the slab's normal unless a clip face is strictly nearer, a
// box's nearest face (the first axis on ties).
__device__ void shape_normal(const rf_paint_shape_t &s, const double x[3], double n[3]) {
    if (s.kind == 0) {
        const double d[3] = {x[0] - s.p[0], x[1] - s.p[1], x[2] - s.p[2]};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r == 0.0) {
            n[0] = 0.0;
            n[1] = 0.0;
            n[2] = 1.0;
            return;
        }
        for (int a = 0; a < 3; a++) n[a] = d[a] / r;
        return;
    }
    double best;
    if (s.kind == 1) {
        const double d = s.nrm[0] * (x[0] - s.p[0]) + s.nrm[1] * (x[1] - s.p[1]) + s.nrm[2] * (x[2] - s.p[2]);
        best = fabs(fabs(d) - s.half_thickness);
        for (int a = 0; a < 3; a++) n[a] = s.nrm[a];
    } else {
        best = INFINITY;
    }
    for (int a = 0; a < 3; a++) {
        const double dl = fabs(x[a] - s.lo[a]), dh = fabs(x[a] - s.hi[a]);
        const double m = dl < dh ? dl : dh;
        if (m < best) {
            best = m;
            n[0] = n[1] = n[2] = 0.0;
            n[a] = 1.0;
        }
    }
}

// Node ids: the last shape containing each node of the ext box (255:
// none), and the number of (shape, node) containments into *painted.
__global__ void __launch_bounds__(256)
k_paint_nodes(const rf_paint_shape_t *__restrict__ sh, int ns, int e0, int e1, int e2, uint8_t *__restrict__ ids,
              unsigned long long *__restrict__ painted) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)e0 * e1 * e2;
    unsigned cnt = 0;
    if (t < total) {
        const int k = static_cast<int>(t % e2);
        const int j = static_cast<int>((t / e2) % e1);
        const int i = static_cast<int>(t / ((long long)e1 * e2));
        int id = 255;
        for (int q = 0; q < ns; q++)
            if (shape_inside(sh[q], (double)i, (double)j, (double)k)) {
                id = q;
                cnt++;
            }
        ids[t] = static_cast<uint8_t>(id);
    }
    // one atomic per wavefront
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(painted, (unsigned long long)cnt);
}

// Edge flags of the nodes of [0, n): bit c set when the E_c edge exists
// (every non-periodic axis other than c has index > 0) and its dual cell
// may contain a boundary of some shape.
__global__ void __launch_bounds__(256)
k_paint_flags(const rf_paint_shape_t *__restrict__ sh, int ns, int n0, int n1, int n2, unsigned per,
              uint8_t *__restrict__ flags) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n0 * n1 * n2) return;
    const int xi[3] = {static_cast<int>(t / ((long long)n1 * n2)), static_cast<int>((t / n2) % n1),
                       static_cast<int>(t % n2)};
    unsigned bits = 0;
    for (int c = 0; c < 3; c++) {
        bool valid = true;
        for (int a = 0; a < 3; a++)
            if (a != c && !((per >> a) & 1u) && xi[a] == 0) valid = false;
        if (!valid) continue;
        double x[3] = {(double)xi[0], (double)xi[1], (double)xi[2]};
        x[c] += 0.5;
        for (int q = 0; q < ns; q++)
            if (shape_near_boundary(sh[q], x[0], x[1], x[2])) {
                bits |= 1u << c;
                break;
            }
    }
    flags[t] = static_cast<uint8_t>(bits);
}

// Flagged edges per (i, j) row.
__global__ void __launch_bounds__(256)
k_paint_rows(const uint8_t *__restrict__ flags, int rows, int n2, int *__restrict__ count) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint8_t *f = flags + (long long)r * n2;
    int c = 0;
    for (int k = 0; k < n2; k++) c += __popc(f[k]);
    count[r] = c;
}

// The flagged edges of each row at the row's offset, k then c ascending:
// code ((i n1 + j) n2 + k) 3 + c.
__global__ void __launch_bounds__(256)
k_paint_list(const uint8_t *__restrict__ flags, int rows, int n2, const long long *__restrict__ off,
             long long *__restrict__ codes) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint8_t *f = flags + (long long)r * n2;
    long long o = off[r];
    for (int k = 0; k < n2; k++) {
        const unsigned b = f[k];
        for (int c = 0; c < 3; c++)
            if ((b >> c) & 1u) codes[o++] = ((long long)r * n2 + k) * 3 + c;
    }
}

constexpr int CL_WAVES = 4;   // wavefronts per block, one edge each

// Classification of the listed edges, one wavefront per edge: the shapes
// whose box (one cell wider) holds the edge are its candidates; each of
// the ss^3 sub-samples of the dual cell takes the last candidate that
// contains it (or the background), counted in LDS; then lane 0 picks the
// two most frequent materials (the lower index first on ties) and writes
// the edge's record: nothing, a scalar override where the cell is uniform
// and differs from the node's material, the scalar 1/M_cc with quantised
// fill and normal in the band near non-periodic faces, or the smoothed
// inverse-permittivity tensor M = (ln - lt) n n^T + lt I.
__global__ void __launch_bounds__(64 * CL_WAVES)
k_paint_classify(const rf_paint_shape_t *__restrict__ sh, int ns, const long long *__restrict__ codes,
                 long long n_edges, int n0, int n1, int n2, unsigned per, int ss, long long keep,
                 rf_paint_rec_t *__restrict__ recs) {
    __shared__ int s_cand[CL_WAVES][256];
    __shared__ unsigned s_hist[CL_WAVES][257];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n[3] = {n0, n1, n2};
    const int total = ss * ss * ss;
    for (long long base = (long long)blockIdx.x * CL_WAVES; base < n_edges;
         base += (long long)gridDim.x * CL_WAVES) {
        const long long e = base + w;
        const bool live = e < n_edges;
        int xi[3] = {0, 0, 0}, c = 0, nc = 0;
        double xe[3] = {0.0, 0.0, 0.0};
        if (live) {
            const long long code = codes[e];
            c = static_cast<int>(code % 3);
            const long long node = code / 3;
            xi[2] = static_cast<int>(node % n2);
            xi[1] = static_cast<int>((node / n2) % n1);
            xi[0] = static_cast<int>(node / ((long long)n1 * n2));
            for (int a = 0; a < 3; a++) xe[a] = (double)xi[a];
            xe[c] += 0.5;
            for (int q0 = 0; q0 < ns; q0 += 64) {
                const int q = q0 + lane;
                const bool hit = q < ns && near_bbox(sh[q], xe[0], xe[1], xe[2]);
                const unsigned long long mask = __ballot(hit);
                if (hit) s_cand[w][nc + __popcll(mask & ((1ull << lane) - 1ull))] = q;
                nc += __popcll(mask);
            }
            for (int t = lane; t <= nc; t += 64) s_hist[w][t] = 0u;
        }
        __syncthreads();
        if (live) {
            for (int smp = lane; smp < total; smp += 64) {
                const int a0 = smp / (ss * ss), a1 = (smp / ss) % ss, a2 = smp % ss;
                const double y0 = xe[0] + ((a0 + 0.5) / ss - 0.5);
                const double y1 = xe[1] + ((a1 + 0.5) / ss - 0.5);
                const double y2 = xe[2] + ((a2 + 0.5) / ss - 0.5);
                int slot = 0;
                for (int q = nc - 1; q >= 0; q--)
                    if (shape_inside(sh[s_cand[w][q]], y0, y1, y2)) {
                        slot = q + 1;
                        break;
                    }
                atomicAdd(&s_hist[w][slot], 1u);
            }
        }
        __syncthreads();
        if (live && lane == 0) {
            // slot 0 is the background (material -1), slot q + 1 candidate q;
            // slots ascend with the material index
            int s1 = -1, s2 = -1;
            for (int sl = 0; sl <= nc; sl++) {
                const unsigned cnt = s_hist[w][sl];
                if (!cnt) continue;
                if (s1 < 0 || cnt > s_hist[w][s1]) {
                    s2 = s1;
                    s1 = sl;
                } else if (s2 < 0 || cnt > s_hist[w][s2]) {
                    s2 = sl;
                }
            }
            const int m1 = s1 <= 0 ? -1 : s_cand[w][s1 - 1];
            // the node's own material: the last shape containing the node
            int nm = -1;
            for (int q = ns - 1; q >= 0; q--)
                if (shape_inside(sh[q], (double)xi[0], (double)xi[1], (double)xi[2])) {
                    nm = q;
                    break;
                }
            rf_paint_rec_t r;
            r.kind = 0;
            for (int v = 0; v < 6; v++) r.v[v] = 0.0;
            if (s2 < 0) {
                if (m1 != nm) {
                    r.kind = 1;
                    r.v[0] = m1 < 0 ? 1.0 : sh[m1].eps_r;
                }
            } else {
                const int m2 = s2 == 0 ? -1 : s_cand[w][s2 - 1];
                const int in = m1 > m2 ? m1 : m2, outm = m1 > m2 ? m2 : m1;
                const int s_in = m1 > m2 ? s1 : s2, s_out = m1 > m2 ? s2 : s1;
                const double e_in = sh[in].eps_r, e_out = outm < 0 ? 1.0 : sh[outm].eps_r;
                const int c_in = static_cast<int>(s_hist[w][s_in]), c_out = static_cast<int>(s_hist[w][s_out]);
                const double f = (double)c_in / (double)(c_in + c_out);
                double nv[3];
                shape_normal(sh[in], xe, nv);
                bool near_face = false;
                for (int a = 0; a < 3; a++)
                    if (!((per >> a) & 1u) && (xe[a] < (double)keep || xe[a] > (double)n[a] - (double)keep))
                        near_face = true;
                if (near_face) {
                    const double fq = floor(f * 64.0 + 0.5) / 64.0;
                    const double ncq = floor(nv[c] * nv[c] * 64.0 + 0.5) / 64.0;
                    const double lnq = (1.0 - fq) / e_out + fq / e_in;
                    const double ltq = 1.0 / ((1.0 - fq) * e_out + fq * e_in);
                    const double mcc = (lnq - ltq) * ncq + ltq;
                    r.kind = 2;
                    r.v[0] = 1.0 / mcc < 1.0 ? 1.0 : 1.0 / mcc;
                } else {
                    const double ln = (1.0 - f) / e_out + f / e_in;
                    const double lt = 1.0 / ((1.0 - f) * e_out + f * e_in);
                    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
                    r.kind = 3;
                    for (int v = 0; v < 6; v++)
                        r.v[v] = (ln - lt) * nv[ia[v]] * nv[ib[v]] + (ia[v] == ib[v] ? lt : 0.0);
                }
            }
            recs[e] = r;
        }
        __syncthreads();   // the wavefront's LDS is reused by the next edge
    }
}

}  // namespace

extern "C" int rf_launch_paint_nodes(const rf_paint_shape_t *sh, int ns, const int ext[3], uint8_t *ids,
                                     unsigned long long *painted, rf_stream_t *s) {
    const long long total = (long long)ext[0] * ext[1] * ext[2];
    if (ns < 0 || ns > 250 || (ns && !sh) || !ids || !painted || total <= 0 || total >= (1LL << 31)) return -1;
    hipLaunchKernelGGL(k_paint_nodes, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                       rf_raw_stream(s), sh, ns, ext[0], ext[1], ext[2], ids, painted);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_paint_flags(const rf_paint_shape_t *sh, int ns, const int n[3], unsigned per,
                                     uint8_t *flags, rf_stream_t *s) {
    const long long total = (long long)n[0] * n[1] * n[2];
    if (ns < 0 || ns > 250 || (ns && !sh) || !flags || total <= 0 || total >= (1LL << 31)) return -1;
    hipLaunchKernelGGL(k_paint_flags, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                       rf_raw_stream(s), sh, ns, n[0], n[1], n[2], per, flags);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_paint_rows(const uint8_t *flags, const int n[3], int *count, rf_stream_t *s) {
    const long long rows = (long long)n[0] * n[1];
    if (!flags || !count || rows <= 0 || rows >= (1LL << 31) || n[2] <= 0) return -1;
    hipLaunchKernelGGL(k_paint_rows, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0,
                       rf_raw_stream(s), flags, static_cast<int>(rows), n[2], count);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_paint_list(const uint8_t *flags, const int n[3], const long long *off, long long *codes,
                                    rf_stream_t *s) {
    const long long rows = (long long)n[0] * n[1];
    if (!flags || !off || !codes || rows <= 0 || rows >= (1LL << 31) || n[2] <= 0) return -1;
    hipLaunchKernelGGL(k_paint_list, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0,
                       rf_raw_stream(s), flags, static_cast<int>(rows), n[2], off, codes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int rf_launch_paint_classify(const rf_paint_shape_t *sh, int ns, const long long *codes,
                                        long long n_edges, const int n[3], unsigned per, int ss, long long keep,
                                        rf_paint_rec_t *recs, rf_stream_t *s) {
    if (n_edges <= 0) return 0;
    if (ns < 1 || ns > 250 || !sh || !codes || !recs || ss < 1 || ss > 16 || keep < 0) return -1;
    const long long want = (n_edges + CL_WAVES - 1) / CL_WAVES;
    const unsigned blocks = static_cast<unsigned>(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(k_paint_classify, dim3(blocks), dim3(64 * CL_WAVES), 0, rf_raw_stream(s), sh, ns, codes,
                       n_edges, n[0], n[1], n[2], per, ss, keep, recs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
