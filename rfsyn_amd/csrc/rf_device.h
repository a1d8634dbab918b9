/*
 * Thin C ABI between the C host code of librfsyn_fdtd and its HIP kernels.
 *
 * Everything here is plain C: device buffers are opaque pointers, kernel
 * arguments are plain structs passed by value.  The HIP side (*.hip) owns
 * the runtime calls; the host side (*.c) never includes HIP headers.
 */
#ifndef RFSYN_RF_DEVICE_H
#define RFSYN_RF_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ runtime */

typedef struct rf_stream rf_stream_t;      /* wraps one hipStream_t */
typedef struct rf_graph rf_graph_t;        /* wraps one hipGraphExec_t */

int rf_init(void);                          /* selects device 0; 0 on success */
int rf_malloc(void **p, size_t bytes);      /* zero-filled device memory */
void rf_free(void *p);
int rf_host_alloc(void **p, size_t bytes);  /* pinned host memory, zero-filled */
void rf_host_free(void *p);
int rf_h2d(void *dst, const void *src, size_t bytes, rf_stream_t *s);
int rf_d2h(void *dst, const void *src, size_t bytes, rf_stream_t *s);
int rf_d2d(void *dst, const void *src, size_t bytes, rf_stream_t *s);
int rf_memset(void *dst, int v, size_t bytes, rf_stream_t *s);
rf_stream_t *rf_stream_create(void);
void rf_stream_destroy(rf_stream_t *s);
int rf_sync(rf_stream_t *s);
int rf_graph_begin(rf_stream_t *s);
int rf_graph_end(rf_stream_t *s, rf_graph_t **out);
int rf_graph_launch(rf_graph_t *g, rf_stream_t *s);
void rf_graph_destroy(rf_graph_t *g);
const char *rf_error_string(void);          /* last runtime error text */
typedef struct rf_event rf_event_t;        /* wraps one hipEvent_t */
rf_event_t *rf_event_create(void);
void rf_event_destroy(rf_event_t *e);
int rf_event_record(rf_event_t *e, rf_stream_t *s);
int rf_event_elapsed_ms(rf_event_t *a, rf_event_t *b, float *ms);
/* An event for ordering only (no timing), and a wait of stream s on it:
 * inside a graph capture the pair forks and joins a second stream. */
rf_event_t *rf_event_create_sync(void);
int rf_stream_wait(rf_stream_t *s, rf_event_t *e);

/* ----------------------------------------------------------- geometry */

/* Padded reference layout: idx = i*sx + j*sy + k, sy = nz+1,
 * sx = (ny+1)*(nz+1). */
typedef struct {
    int n[3];           /* nx, ny, nz */
    int per;            /* periodic mask, x=1 y=2 z=4 */
    long long sx, sy;   /* plane and row strides in elements */
} rf_geom_t;

/* CPML on the device.  psi memories exist only in slabs: along axis a,
 * coordinate q is in the slab when q < lo[a] or q >= hi[a], and is stored
 * at compressed coordinate q (q < lo) or lo + q - hi (q >= hi).  A slab
 * array for derivative axis a has the padded extents of the grid with
 * axis a replaced by len[a] = lo[a] + (n[a] + 1 - hi[a]). */
typedef struct {
    const void *kinv_m[3], *b_m[3], *a_m[3];    /* n+1 entries per axis */
    const void *kinv_h[3], *b_h[3], *a_h[3];    /* n entries per axis */
    int lo[3], hi[3], len[3];
    /* plain range per axis: for q in [plo, phi) both kinv profiles are 1
     * and q lies outside the slab, so the stretched difference equals the
     * plain one (up to the sign of a zero) */
    int plo[3], phi[3];
    /* psi per component c and derivative s (axis (c+1+s)%3) */
    void *psi_h[3][2];
    void *psi_e[3][2];
} rf_cpml_t;

/* Material entry flags per component. */
#define RF_EDGE_TENSOR 1u       /* updated by the tensor kernel: skip */
#define RF_EDGE_ADE    2u       /* dispersive: ADE update */

/* Per-tag material entry (one per distinct combination of the three
 * component coefficients and flags), in the field precision. */
typedef struct {
    double Ca[3], Cb[3];
    uint32_t flags[3];
    int ade[3];                 /* ADE coefficient set per component, -1 none */
} rf_entry64_t;

typedef struct {
    float Ca[3], Cb[3];
    uint32_t flags[3];
    int ade[3];
} rf_entry32_t;

#define RF_MAX_SLOTS 8
/* Trapezoidal-ADE coefficients of one dispersive material. */
/* Trapezoidal-ADE coefficients of one dispersive material, and where its
 * state lives: the set's own node box [b0, b0+bn) (the nodes with an edge
 * of this set), J and P of pole s, component c, node (i,j,k) at element
 * base + (3 s + c) st + ((i-b0[0]) bn[1] + (j-b0[1])) bn[2] + (k-b0[2]),
 * poles s < n_poles only.  A box per set keeps each layer's state dense
 * (no lines of other layers, gaps or non-dispersive stripes in it). */
typedef struct {
    int n_poles;
    int has_p[RF_MAX_SLOTS];
    double kj[RF_MAX_SLOTS], kp[RF_MAX_SLOTS], v[RF_MAX_SLOTS];
    double Cu, hdt;
    long long base, st;
    int b0[3], bn[3];
} rf_ade64_t;

typedef struct {
    int n_poles;
    int has_p[RF_MAX_SLOTS];
    float kj[RF_MAX_SLOTS], kp[RF_MAX_SLOTS], v[RF_MAX_SLOTS];
    float Cu, hdt;
    long long base, st;
    int b0[3], bn[3];
} rf_ade32_t;

typedef struct {
    const void *tag;            /* per padded node; NULL = vacuum */
    int tag_bytes;              /* 1, 2 or 4 */
    const void *entries;        /* rf_entry64_t or rf_entry32_t */
    const void *ade;            /* rf_ade64_t or rf_ade32_t */
    /* union of the ADE sets' state boxes [b0, b0+bn) per axis (the nodes
     * with a dispersive edge); the state of each set is laid out in its
     * own box (rf_ade64_t / rf_ade32_t) */
    int b0[3], bn[3];
    int n_slots;                /* largest pole count of a set */
    int n_ade;                  /* ADE coefficient sets in `ade` */
    void *J, *P;                /* the sets' state regions, back to back */
    double vac_ca;              /* vacuum Cb = dt/(eps0 dx) when tag == NULL */
    int n_entries;              /* entries in the table */
} rf_mat_t;

/* One tensor edge: position, component, kind and neighbour weights.
 * Neighbours follow the order of the reference pair rule: for each other
 * component d in increasing order, the 4 d-edges at x + s_c e_c - s_d e_d
 * with the offset along the lower-numbered axis varying fastest; a zero
 * mask bit means "no tensor there". */
typedef struct {
    int i, j, k;
    int comp;
    int kind;                   /* 2 or 3 */
    uint32_t mask;              /* bit 4*slot + q: neighbour present */
    double diag, scale;
    double w[8];                /* [slot*4 + q], slot 0/1 = the two d's */
} rf_tedge64_t;

typedef struct {
    int i, j, k;
    int comp;
    int kind;
    uint32_t mask;
    float diag, scale;
    float w[8];
} rf_tedge32_t;

/* Sources and probes of the device-resident loop. */
typedef struct {
    long long idx;              /* padded index */
    int comp;                   /* 0..2 */
    int hard;
    const double *wave;         /* value per absolute step */
    int i, j, k;                /* node of idx */
} rf_point_src_t;

typedef struct {
    int comp, axis, index, hard;
    const double *wave;         /* value per absolute step */
} rf_plane_src_t;

/* One source of the fused step, in application order: the pre-E entries
 * (point sources in insertion order, then plane sources) and after them
 * the post-E plane sources. */
typedef struct {
    int plane;                  /* 0 = point node (i,j,k), 1 = plane axis/index */
    int comp, hard, after;
    int axis, index;            /* plane; index already wrapped on a periodic axis */
    int i, j, k;                /* point */
    const double *wave;         /* value per absolute step */
} rf_leap_src_t;

/* Data of the fused step that only its dispersive-edge path reads, kept
 * in device memory so the hot loop does not hold it in registers: the ADE
 * tables and state. */
typedef struct {
    const void *entries, *ade;
    void *J, *P;
} rf_leap_cold_t;

/* z columns a wavefront of the fused step owns (its 64 lanes minus the
 * two halo lanes). */
#define RF_LEAP_COLS 62

/* Fused leapfrog step: one kernel computes H^{n+1/2} and then E^{n+1}
 * from the "A" buffers (E^n, H^{n-1/2}) into the "B" buffers for the nodes
 * of the box [ib0, ib1), which must lie inside the CPML plain ranges with
 * a one-node margin (no CPML term is evaluated).  A wavefront owns `rows`
 * rows (y) by 62 columns (z) of the box and marches along x over `chunk`
 * planes, recomputing the one-cell H halo of its tile from the A buffers.
 * The ADE state is updated in place.  Every node of the box is written to
 * B (copies where a component is not updated).  The sources of the device
 * table `srcs` are applied in its order; src_box bounds the nodes they
 * touch (i0,i1,j0,j1,k0,k1 inclusive).  rf_leap_plan fills the tiling
 * fields (rows, chunk: 0 = automatic). */
typedef struct {
    rf_geom_t g;
    /* state blocks: the six arrays ex ey ez hx hy hz of state n (A, read)
     * and n+1 (B, written), `fstride` elements apart, 6 * fstride *
     * sizeof(T) < 2^32 bytes */
    const void *blk_a;
    void *blk_b;
    long long fstride;
    double cb, dx;
    int ib0[3], ib1[3];         /* owned node box */
    int kbrk;                   /* z tiles restart here (ib0[2] <= kbrk <= ib1[2]) */
    const void *tag;            /* material tags, NULL = vacuum */
    int tag_bytes, n_entries;
    const void *entries;
    const void *ade;            /* ADE coefficient sets (rf_ade32_t / rf_ade64_t) */
    int n_ade, n_slots;         /* sets; poles of the state arrays */
    int ade_split;              /* 1: dispersive edges get the plain part of
                                 * the update; rf_launch_ade_pass completes them */
    double vac_ca;
    const rf_leap_cold_t *cold; /* device */
    const rf_leap_src_t *srcs;  /* device table */
    int n_srcs;
    int src_box[6];
    int src_pre, src_post;      /* component bits with sources before / after the E update */
    const long long *d_step;
    /* CPML along z inside the box: plain range, psi slab [0,zlo) u [zhi,nz],
     * the z profiles of rf_cpml_t (kinv_h b_h a_h kinv_m b_m a_m) and the
     * z-psi store: for each (i, j) six slot rows of zlen values (compressed
     * k fastest), value (i, j, s, q) at ((i (ny+1) + j) 6 + s) zlen + q; H of
     * hx and hy in slots 2*zpar (read) and 2-2*zpar (written), then E of ex
     * and ey in slots 4 and 5 */
    int zc;
    int zplo, zphi, zlo, zhi, zlen;
    const void *zk[6];
    void *zpsi;
    long long zbytes;           /* size of the z-psi store (< 2^32) */
    int zpar;
    /* work queue: 4 x 8 zero-initialised device counters; launch part p
     * uses the set 2 p + qpar (alternate qpar from step to step) */
    unsigned *ctr;
    int qpar;
    /* tiling, filled by rf_leap_plan */
    int rows, chunk;            /* rows per item, x planes per item */
    int zrow;                   /* row mode: each item spans every z column */
    int ntk, ntk0, ntj;         /* tiles along z (ntk0 below kbrk) and y */
    int ncache, n_adec;         /* material entries and ADE sets staged in LDS */
    /* the step runs in n_parts launches over disjoint sets of z tiles:
     * part p takes the tiles ptk[p][0] + [0, ptk[p][1]) and
     * ptk[p][2] + [0, ptk[p][3]), with the z CPML code if pzc[p], in
     * pitems[p] work items on pblocks[p] blocks; with zsplit set (input)
     * the tiles with no lane in the z CPML form a part of their own that
     * runs the variant without it */
    int zsplit;
    /* coop (input): 0 = every wavefront takes its own work item; else
     * cooperative blocks: a block of gz x gy wavefronts takes a group of
     * gz = 1 z tile by gy = max(1, coop & 15) neighbouring row tiles of one
     * x chunk (gy <= 4; reduced to fit) and marches its planes together:
     * in lockstep (a barrier per plane) when gk = coop >> 4 is 0, else no
     * wavefront starts plane p before every one of the group has completed
     * plane p - gk */
    int coop;
    int n_parts;
    int ptk[2][4], pzc[2], pitems[2], pblocks[2], pgz[2], pgy[2];
    /* the part a launch runs (set by rf_launch_leapfrog); gz = 0:
     * independent mode */
    int tkr[4];
    int n_items;
    int gz, gy, gk;
} rf_leap_t;

/* Tiling of the fused step over a->ib0..ib1; rows: 2 or 4 (FP64), 4 or 8
 * (FP32), 0 = default; chunk 0 = about twelve work items per resident
 * wavefront, at least 8 planes each (target_waves > 0 overrides the
 * residency the device reports).  FP32 rows 1 selects row mode: an
 * item is one row of every z column (lane l of chunk c holds
 * column 64 c + l, ceil(nz / 64) chunks), for a box spanning z columns
 * 0..nz with the z CPML inside it, nz <= 256 (below 193: vacuum or
 * one-byte tags), both z CPML slabs within the first and last 64 columns,
 * dispersive edges left to the ADE pass, independent wavefronts, and (nz
 * a multiple of 64) no source on column nz; chunk 0 = about two items per
 * resident wavefront, at least 8 planes each.  0 on success, -1 for an
 * unsupported request. */
int rf_leap_plan(int prec, rf_leap_t *a, int rows, int chunk, int target_waves);
/* Why row mode cannot run on a's box and options (NULL: it can). */
const char *rf_leap_zrow_refusal(const rf_leap_t *a);
/* tag NULL = vacuum (vac_ca); an empty box launches nothing. */
int rf_launch_leapfrog(int prec, const rf_leap_t *a, rf_stream_t *s);

/* The nodes around the fused step's box: up to 6 disjoint boxes [lo, hi),
 * flattened with k fastest; box b holds flat indices [start[b],
 * start[b+1]). */
typedef struct {
    int n;
    int lo[6][3], hi[6][3];
    long long start[7];         /* each box holds fewer than 2^32 nodes */
    /* division by the box's k and j extents: q = (t + ((n - t) >> s1)) >> s2
     * with t = mulhi(m, n) (round-up multiply-high method) */
    unsigned mk[6], mj[6];
    unsigned sk[6][2], sj[6][2];
} rf_shell_t;

/* z-psi store of the fused step (see rf_leap_t): NULL = the z CPML terms
 * use the psi arrays of rf_cpml_t. */
typedef struct {
    void *zpsi;
    int zpar, zlen;
} rf_zpsi_t;

/* Shell H: H^{n+1/2} of the shell nodes from E^n and H^{n-1/2} of the A
 * buffers (fa: ex ey ez hx hy hz) into hb, psi_h in place (z terms in the
 * z-psi store when given). */
int rf_launch_shell_h(int prec, const rf_geom_t *g, const rf_shell_t *sh, void *const fa[6], void *const hb[3],
                      double cb, double dx, const rf_cpml_t *cpml, const rf_zpsi_t *z, rf_stream_t *s);
/* Shell E: sources and E^{n+1} of the shell nodes from E^n (ea) and
 * H^{n+1/2} (hb) into eb, psi_e and the ADE state in place. */
int rf_launch_shell_e(int prec, const rf_geom_t *g, const rf_shell_t *sh, void *const ea[3], void *const hb[3],
                      void *const eb[3], double dx, const rf_cpml_t *cpml, const rf_zpsi_t *z,
                      const rf_mat_t *mat, const rf_leap_src_t *srcs, int n_srcs, const long long *d_step,
                      rf_stream_t *s);

/* ------------------------------------------------------ packed scenes
 * The steps of many scenes in one launch per kernel kind.  Every scene's
 * arguments are a descriptor in a device table; the launch's work items
 * or blocks carry the scene they belong to. */

/* Fused steps of n_scenes scenes: tab[sc] is scene sc's rf_leap_t
 * (planned with one part, independent wavefronts, ncache and n_adec
 * within lds_wave bytes per wavefront), its work items are the global
 * items item_base[sc] .. item_base[sc] + tab[sc].n_items, and
 * item_scene[item] names the scene of each.  All scenes use the variant
 * (rows, tag_bytes, zc).  ctr: 2 x 8 zero-initialised counters, set qpar
 * this launch (the other set is zeroed by it). */
typedef struct {
    const rf_leap_t *tab;
    const int *item_base;
    const unsigned short *item_scene;
    int n_items, n_scenes;
    unsigned *ctr;
    int qpar;
    int rows, tag_bytes, zc;
    int lds_wave;
    int blocks;                 /* set by rf_leap_pack_plan */
} rf_leap_pack_t;
/* LDS bytes one wavefront needs to stage a's tables. */
size_t rf_leap_lds_bytes(int prec, const rf_leap_t *a);
/* pk->blocks for a launch of pk->n_items items (resident wavefronts of
 * the variant, at most one block per 4 items). */
int rf_leap_pack_plan(int prec, rf_leap_pack_t *pk);
int rf_launch_leapfrog_packed(int prec, const rf_leap_pack_t *pk, rf_stream_t *s);

/* The shell update of one scene in a packed launch: its arguments as in
 * rf_launch_shell_h / rf_launch_shell_e, and the first block of its
 * nodes in the launch (block0; its nodes take the following
 * ceil(nodes / 256) blocks). */
typedef struct {
    rf_geom_t g;
    rf_shell_t sh;
    const void *fa[6];          /* state n: ex ey ez hx hy hz (A) */
    void *hb[3];                /* H^{n+1/2} (B) */
    void *eb[3];                /* E^{n+1} (B) */
    double cb, dx;
    rf_cpml_t cp;               /* valid when the launch's has_cpml */
    rf_zpsi_t zs;               /* zpsi NULL: psi arrays of cp */
    rf_mat_t m;
    const rf_leap_src_t *srcs;
    int n_srcs;
    const long long *d_step;
    unsigned block0;
} rf_shell_job_t;
/* jobs, block_job: device; block_job[b] is the job of block b of n_blocks.
 * Every job uses the variant (tag_bytes, has_cpml). */
int rf_launch_shell_h_packed(int prec, const rf_shell_job_t *jobs, const unsigned short *block_job, int n_blocks,
                             int has_cpml, rf_stream_t *s);
int rf_launch_shell_e_packed(int prec, const rf_shell_job_t *jobs, const unsigned short *block_job, int n_blocks,
                             int tag_bytes, int has_cpml, rf_stream_t *s);

/* The split ADE passes of many scenes: per job the arguments of
 * rf_launch_ade_pass (box at lo with j, k extents nj, nk and count nodes,
 * E^n in ea, the plain part and then E^{n+1} in eb; no captures), ncache
 * = min(n_entries, 4096) entries staged in LDS, and the first block of
 * its ceil(count / 256) blocks in the launch.  Every job uses the variant
 * (tag_bytes, ns_class = rf_ade_pass_class(n_slots)); lds = the largest
 * n_ade * sizeof(rf_ade*_t) + 8 ncache of the jobs. */
typedef struct {
    rf_geom_t g;
    int lo[3];
    unsigned nj, nk, count;
    const void *ea[3];
    void *eb[3];
    rf_mat_t m;
    int ncache;
    unsigned block0;
} rf_ade_job_t;
int rf_ade_pass_class(int n_slots);
int rf_launch_ade_pass_packed(int prec, const rf_ade_job_t *jobs, const unsigned short *block_job, int n_blocks,
                              int tag_bytes, int ns_class, size_t lds, rf_stream_t *s);

/* Probe gathers of many scenes (rf_launch_probe_record per job, one block
 * each; every job's step counter advances). */
typedef struct {
    const void *f[6];
    const long long *idx;
    const int *comp;
    int n;
    double *ring;
    int ring_len;
    double *td;
    long long td_cap;
    long long *d_step;
} rf_probe_job_t;
int rf_launch_probe_record_packed(int prec, const rf_probe_job_t *jobs, int n_jobs, rf_stream_t *s);

/* Point-probe DFTs of many scenes over the last `count` steps
 * (rf_launch_probe_dft per job; jobs with n_freq 0 do nothing). */
typedef struct {
    const double *ring;
    int ring_len, n_probes;
    const int *freq_probe;
    int n_freq;
    const double *freq_w;
    double dt;
    const long long *d_step;
    double *acc_re, *acc_im, *cmp_re, *cmp_im;
    int comp_sum;
} rf_pdft_job_t;
int rf_launch_probe_dft_packed(const rf_pdft_job_t *jobs, int n_jobs, int max_freq, int count, rf_stream_t *s);

/* Plane-monitor captures in one launch: capture q copies the plane
 * `index` of `axis` of field array f (cells (ia, ib), the other two axes
 * in increasing order, ia < na, ib < nb) into slot step % ring_len of
 * its ring buf (the field's precision); start is set by the launcher. */
#define RF_MAX_CAPS 8
typedef struct {
    const void *f;
    void *buf;
    int axis, index, na, nb;
    long long start;
} rf_cap_t;
typedef struct {
    int n;
    rf_cap_t c[RF_MAX_CAPS];
} rf_caps_t;

/* ADE pass of the split step: for the nodes of [lo, hi) (inside the ADE
 * state box and the fused box), every dispersive edge whose E^{n+1} in eb
 * holds the plain part Ca e^n + Cb (dA - dB) of its update is completed
 * with E^n from ea, and J, P advance (ade_edge, yee_device.h).  At most
 * 64 ADE sets and 1 <= n_slots.  When caps is given (n > 0), extra blocks
 * of the same launch take the plane captures of step *d_step into rings
 * of cap_ring slots; every captured plane must lie outside [lo, hi), whose
 * values the pass is still writing. */
int rf_launch_ade_pass(int prec, const rf_geom_t *g, const int lo[3], const int hi[3], void *const ea[3],
                       void *const eb[3], const rf_mat_t *mat, const rf_caps_t *caps, int cap_ring,
                       const long long *d_step, rf_stream_t *s);

/* ----------------------------------------------------------- launches
 * `prec` selects the field precision: 0 = double, 1 = float.  All
 * launches go to stream s and return 0 or -1. */

int rf_launch_update_h(int prec, const rf_geom_t *g, void *const f[6], double cb, double dx,
                       const rf_cpml_t *cpml, rf_stream_t *s);
int rf_launch_update_e(int prec, const rf_geom_t *g, void *const f[6], double dx,
                       const rf_cpml_t *cpml, const rf_mat_t *mat, rf_stream_t *s);
/* rf_launch_update_e with the n_pts point sources pts (device) applied to
 * E^n first, as rf_launch_point_sources at the step *d_step would; every
 * source must sit on a node and component the E update writes (not a
 * boundary node it skips, not a tensor edge). */
int rf_launch_update_e_src(int prec, const rf_geom_t *g, void *const f[6], double dx, const rf_cpml_t *cpml,
                           const rf_mat_t *mat, const rf_point_src_t *pts, int n_pts, const long long *d_step,
                           rf_stream_t *s);
int rf_launch_tensor(int prec, const rf_geom_t *g, void *const f[6], const void *edges,
                     long long n_edges, rf_stream_t *s);

/* Point sources at step *d_step (before E). */
int rf_launch_point_sources(int prec, void *const f[6], const rf_point_src_t *srcs_dev, int n,
                            const long long *d_step, int step_offset, rf_stream_t *s);
/* Plane source over [0,n_a) x [0,n_b) of the plane; step index
 * *d_step + step_offset selects the waveform value. */
int rf_launch_plane_source(int prec, const rf_geom_t *g, void *const f[6], const rf_plane_src_t *p,
                           const long long *d_step, int step_offset, rf_stream_t *s);

/* Probe gather: ring[(step % ring_len) * n + p] = field value (as double);
 * td[p * td_cap + step] when step < td_cap.  Then increments *d_step. */
int rf_launch_probe_record(int prec, void *const f[6], const long long *idx_dev, const int *comp_dev,
                           int n, double *ring, int ring_len, double *td, long long td_cap,
                           long long *d_step, rf_stream_t *s);

/* Point-probe DFT over the last `count` ring entries, ending at step
 * *d_step - 1.  One accumulator per flattened (probe, frequency) pair t:
 * probe freq_probe[t], angular frequency freq_w[t] (rad/s):
 *   acc_re += v cos(w t_n) dt,  acc_im -= v sin(w t_n) dt,  t_n = (n+1) dt
 * (probe.c:87-91).  comp_sum != 0: Neumaier-compensated sums. */
int rf_launch_probe_dft(const double *ring, int ring_len, int n_probes, const int *freq_probe,
                        int n_freq, const double *freq_w, double dt, const long long *d_step,
                        int count, double *acc_re, double *acc_im, double *cmp_re, double *cmp_im,
                        int comp_sum, rf_stream_t *s);

/* Dataset output stage (output.hip): the mean power of n time-domain
 * samples as complex floats (re = (float)v, im = 0) summed in sample order
 * into *power (device); then the n I/Q samples into out (device, 2n
 * floats) with, when noise, the AWGN pair of draws 2i, 2i+1 of the PCG32
 * stream whose seeded state is state0 (increment inc) at per-axis
 * standard deviation std_f. */
int rf_launch_iq_power(const double *td, long long n, double *power, rf_stream_t *s);
int rf_launch_iq_noise(const double *td, long long n, int noise, float std_f, uint64_t state0, uint64_t inc,
                       float *out, rf_stream_t *s);
/* The output stage's device memory that is never allocated: a file-scope
 * array of the library, on the current device (set up on the first call
 * per device).  *power: one double; *iq: RF_IQ_SCRATCH_SAMPLES I/Q pairs.
 * 0 on success. */
#define RF_IQ_SCRATCH_SAMPLES 65536
#define RF_MAX_DEVICES 64
int rf_iq_scratch(double **power, float **iq);

int rf_launch_plane_captures(int prec, const rf_geom_t *g, const rf_caps_t *cs, int ring_len,
                             const long long *d_step, rf_stream_t *s);

/* Scene painter (scene_kernels.hip, include/fdtd_scene.h).  Shapes as the
 * device reads them: kind 0 sphere, 1 slab, 2 box; nrm the unit slab
 * normal; bb_lo/bb_hi the bounding box.  Records: kind 0 nothing, 1 scalar
 * override eps v[0], 2 scalar mixed eps v[0], 3 tensor v[0..5] (xx xy xz
 * yy yz zz). */
typedef struct {
    int kind;
    double eps_r;
    double p[3], radius, nrm[3], half_thickness, lo[3], hi[3];
    double bb_lo[3], bb_hi[3];
} rf_paint_shape_t;
typedef struct {
    int kind;
    double v[6];
} rf_paint_rec_t;
/* ids[(i ext1 + j) ext2 + k] = last shape containing node (i,j,k), 255
 * none; *painted += number of (shape, node) containments. */
int rf_launch_paint_nodes(const rf_paint_shape_t *sh, int ns, const int ext[3], uint8_t *ids,
                          unsigned long long *painted, rf_stream_t *s);
/* flags[(i n1 + j) n2 + k] bit c: the E_c edge exists and its dual cell may
 * contain a shape boundary. */
int rf_launch_paint_flags(const rf_paint_shape_t *sh, int ns, const int n[3], unsigned per, uint8_t *flags,
                          rf_stream_t *s);
/* count[i n1 + j]: flagged edges of row (i, j). */
int rf_launch_paint_rows(const uint8_t *flags, const int n[3], int *count, rf_stream_t *s);
/* codes[off[row] ...]: ((i n1 + j) n2 + k) 3 + c of the row's flagged
 * edges, k then c ascending. */
int rf_launch_paint_list(const uint8_t *flags, const int n[3], const long long *off, long long *codes,
                         rf_stream_t *s);
/* recs[e]: the painted coefficient of edge codes[e] (ss^3 sub-samples,
 * scalar band `keep` cells at non-periodic faces). */
int rf_launch_paint_classify(const rf_paint_shape_t *sh, int ns, const long long *codes, long long n_edges,
                             const int n[3], unsigned per, int ss, long long keep, rf_paint_rec_t *recs,
                             rf_stream_t *s);

/* Plane monitor DFT flush over `count` (<= RF_PDFT_TWS) buffered steps
 * ending at step *d_step - 1; accumulators [f][cell]; tw_scratch holds
 * 2 * nf * RF_PDFT_TWS doubles. */
#define RF_PDFT_TWS 128
int rf_launch_plane_dft(const void *buf, int buf_prec, long long cells, int ring_len, int nf,
                        const double *freq_w, double dt, const long long *d_step, int count,
                        double *tw_scratch, double *acc_re, double *acc_im, double *cmp_re,
                        double *cmp_im, int comp_sum, rf_stream_t *s);

/* Energy partial sums (double): out[0] = sum 0.5 eps E^2, out[1] = sum
 * 0.5 mu0 H^2 over i<nx, j<ny, k<nz, unscaled by dx^3.  eps of each edge
 * is dt/(dx Cb) from cb64 (per tag entry, per component) or vac_eps. */
int rf_launch_energy(int prec, const rf_geom_t *g, void *const f[6], const void *tag, int tag_bytes,
                     const double *eps_tab, double vac_eps, double mu0, double *partials,
                     int n_partials, double *out, rf_stream_t *s);
/* Energy audit of step n: E^n and H^{n-1/2} from fn (state n), H^{n+1/2}
 * from fp (state n+1); out[0..5) = sums over the cells of 0.5 eps E^2,
 * 0.5 mu0 (H-)^2, 0.5 mu0 (H+)^2, 0.5 mu0 ((H- + H+)/2)^2, 0.5 mu0 H- H+
 * (unscaled by dx^3; partials holds 5 * n_partials values). */
int rf_launch_audit(int prec, const rf_geom_t *g, void *const fn[6], void *const fp[6], const void *tag,
                    int tag_bytes, const double *eps_tab, double vac_eps, double mu0, double *partials,
                    int n_partials, double *out, rf_stream_t *s);
/* Adds to *acc the Poynting flux along +axis through the plane of E nodes
 * at `index` at time n+1/2 (unscaled by dx^2 dt): fn holds state n, fp
 * state n+1. */
int rf_launch_flux(int prec, const rf_geom_t *g, int axis, int index, void *const fn[6], void *const fp[6],
                   double *partials, int n_partials, double *acc, rf_stream_t *s);

/* *flag |= 1 when any of the n values at v (prec 0 double, 1 float) is a
 * NaN or an infinity; the caller clears *flag. */
int rf_launch_nonfinite(int prec, const void *v, long long n, unsigned *flag, rf_stream_t *s);

/* Precision conversion between the double host layout and device fields. */
int rf_launch_convert(void *dst, int dst_prec, const void *src, int src_prec, long long n,
                      rf_stream_t *s);

/* ------------------------------------------------ surrogate kernels
 * (sur_kernels.hip): BF16 operands as uint16_t bit patterns. */
#include <stdint.h>
/* C(m, n) = sum_k A(m, k) B(k, n) (+ bias[n]), A(m, k) = A[m sam + k sak],
 * B(k, n) = B[k sbk + n sbn], FP32 accumulation; splits > 1 partitions K
 * and sums the partials (part: splits M N floats) in split order. */
int rf_sur_gemm(int M, int N, int K, const uint16_t *A, long long sam, long long sak, const uint16_t *B,
                long long sbk, long long sbn, float *C, long long ldc, const float *bias, int splits, float *part,
                rf_stream_t *s);
int rf_sur_to_bf16(const float *x, uint16_t *y, long long n, rf_stream_t *s);
int rf_sur_gelu(const float *z, float *a, uint16_t *a16, long long n, rf_stream_t *s);
int rf_sur_gelu_bwd(const float *z, const float *da, float *dz, uint16_t *dz16, long long n, rf_stream_t *s);
int rf_sur_im2col(const uint16_t *in, uint16_t *col, int B, int S, int C, rf_stream_t *s);
int rf_sur_col2im(const float *dcol, float *din, int B, int S, int C, rf_stream_t *s);
int rf_sur_expand_y(const float *maps, uint16_t *vox, int B, int S, int C, rf_stream_t *s);
int rf_sur_add_scene(const float *U, const float *V, const float *bias, float *h, long long rows, int P, int N,
                     rf_stream_t *s);
int rf_sur_sum_queries(const float *dh, float *dU, uint16_t *dU16, long long scenes, int P, int N, rf_stream_t *s);
int rf_sur_mse(const float *y, const float *label, float *grad, uint16_t *grad16, long long n, double *partials,
               int n_partials, rf_stream_t *s);
int rf_sur_adam(float *w, uint16_t *w16, const float *g, float *m, float *v, long long n, float lr, float b1,
                float b2, float eps, float c1, float c2, rf_stream_t *s);

#ifdef __cplusplus
}
#endif

#endif
