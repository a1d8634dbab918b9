/*
 * Device-resident simulation handle (fdtd_native.h) and the one-call
 * entry point fdtd_simulate_echo (engine/fdtd/simulate.c:10-69).
 *
 * One step on the GPU stream, in the reference order (simulate.c:48-54,
 * validate_probe_dft.c:66-82):
 *   H kernel -> point/plane sources at t = n dt -> E kernel (+ tensor
 *   edges) -> post-E plane sources at t = (n+1) dt -> plane-monitor
 *   capture -> probe gather, which also advances the device step counter.
 * Every dft_block steps the buffered probe and plane values are folded
 * into the FP64 DFT accumulators.  A block of dft_block steps plus its
 * flush is captured once as a HIP graph and replayed, so the host issues
 * one launch per block.  Kernels read the step index from device memory,
 * which is what makes the captured block replayable.
 */
#include "host_internal.h"
#include "fdtd_native.h"
#include "simulate.h"
#include "source.h"
#include "constants.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    source_waveform_fn wf;
    const void *params;
    double amp;
    int after_e;          /* table holds amp*wf((n+1)dt) instead of amp*wf(n dt) */
    double *d_tab;        /* device, wave_cap entries */
} wave_t;

typedef struct {
    int comp;
    long long idx;
    size_t td_cap;
    size_t nf, f_off;
} sprobe_t;

typedef struct {
    int axis, index, comp[2], na, nb, nf;
    double *freqs;
    double *d_w;
    void *ring[2];
    double *acc_re[2], *acc_im[2], *cmp_re[2], *cmp_im[2];
} smon_t;

typedef struct {
    rf_plane_src_t d;
    int wave;             /* index into waves */
    int after_e;
} splane_t;

struct fdtd_sim {
    int prec;
    size_t es;
    size_t nx, ny, nz;
    double dx, dt;
    unsigned per;
    rf_geom_t geom;
    size_t cells;
    rf_stream_t *st;
    void *f[6];           /* current state: six arrays of one block */
    void *blk, *blk2;     /* the state blocks (f, and f2 of the fused path) */
    int has_cpml;
    rf_cpml_dev_t cp;
    int has_mat;
    rf_mat_dev_t mat;
    rf_mat_t vac;
    long long *d_step;
    size_t steps;
    int comp_sum;
    int K;
    int pending;
    int started;

    wave_t *waves; size_t n_waves;
    size_t wave_cap;
    unsigned wave_gen;    /* counts reallocations of the wave tables */
    int *pt_wave; source_t *pts; size_t n_pts;
    rf_point_src_t *d_pts;
    splane_t *planes; size_t n_planes;

    sprobe_t *probes; size_t n_probes;
    double *freqs_all; size_t n_freq_tot;
    long long *d_pidx; int *d_pcomp;
    double *d_ring, *d_td;
    size_t td_stride;
    int *d_fprobe; double *d_fw;
    double *d_acc_re, *d_acc_im, *d_cmp_re, *d_cmp_im;

    smon_t *mons; size_t n_mons;
    double *d_tw;

    double *d_energy;     /* partials + 2 results */
    int n_partials;
    double *d_aux;        /* 5 * n_partials partials + 5 results: audit and flux reductions */
    int (*flux)[2];       /* Poynting flux planes: axis, index */
    int n_flux;
    double *d_flux;       /* their time integrals (unscaled) */

    rf_graph_t *graphs[2];   /* captured block per buffer parity */
    size_t bytes;

    /* fused leapfrog path: state in f, next state written to f2, then swapped */
    int fused;
    int parity;
    void *f2[6];
    int leap_rows, leap_chunk, leap_waves;   /* tiling requests (0: automatic) */
    int plan_rows;           /* rows per item of the last planned fused step */
    unsigned *d_nonfinite;   /* flag of fdtd_sim_fields_finite */
    int leap_zsplit;         /* z CPML tiles in a launch of their own */
    int leap_coop;           /* cooperative blocks (rf_leap_t.coop) */
    rf_leap_src_t *d_srcs;   /* device source table of the fused step */
    rf_leap_cold_t *d_cold;  /* its cold data (ADE) */
    int ib0[3], ib1[3];      /* node box of the fused kernel */
    int kbrk;                /* its z tiles restart at the upper z slab */
    int zc, zpar;            /* z CPML inside the box; z-psi buffer parity */
    void *zpsi;              /* z-psi store of the box (and shell) nodes */
    unsigned *d_ctr;         /* work-queue counters of the fused kernel */
    int src_pre, src_post;   /* sourced components before / after E */
    rf_shell_t shell;        /* the nodes around it (CPML slabs + margin) */
    rf_stream_t *st2;        /* the shell's H update, beside the fused kernel */
    rf_event_t *fork, *join;
    int shell_overlap;       /* st2 in use (off for handles stepped in a batch) */
    int n_srcs, src_box[6];
    int post_in_kernel;      /* post-E plane sources inside the fused kernel */
    int ade_split_off;       /* keep the ADE update inside the fused kernel */
    rf_point_src_t *h_pts;   /* host copy of the point sources (device wave tables) */

    /* kernel timing with HIP events (profiling mode launches directly) */
    int profiling;
    rf_event_t **ev;
    size_t n_ev, ev_used;
    double prof_ms[3];       /* event slots: H / fused, E / ADE pass, shell */
    long prof_launches;
};

static int dalloc(fdtd_sim_t *s, void **p, size_t bytes) {
    if (rf_malloc(p, bytes) != 0) {
        rf_set_error("device allocation of %zu bytes failed: %s", bytes, rf_error_string());
        return -1;
    }
    s->bytes += bytes;
    return 0;
}

static int upload_fields(fdtd_sim_t *s, const fdtd_grid_t *g) {
    const double *src[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
    size_t bytes = s->cells * sizeof(double);
    void *tmp = NULL;
    if (s->prec && rf_malloc(&tmp, bytes) != 0) return -1;
    for (int q = 0; q < 6; q++) {
        int rc;
        if (s->prec == 0) {
            rc = rf_h2d(s->f[q], src[q], bytes, s->st);
        } else {
            rc = rf_h2d(tmp, src[q], bytes, s->st);
            if (!rc) rc = rf_launch_convert(s->f[q], 1, tmp, 0, (long long)s->cells, s->st);
        }
        if (!rc) rc = rf_sync(s->st);
        if (rc) { rf_free(tmp); return -1; }
    }
    rf_free(tmp);
    return 0;
}

fdtd_sim_t *fdtd_sim_create(const fdtd_grid_t *g, const cpml_t *pml, const materials_t *mat,
                            const fdtd_sim_options_t *opt) {
    if (!g) { rf_set_error("NULL grid"); return NULL; }
    if (pml && !rf_cpml_matches(pml, g)) { rf_set_error("CPML was built for another grid"); return NULL; }
    if (mat && !rf_materials_matches(mat, g)) { rf_set_error("materials were built for another grid"); return NULL; }
    fdtd_sim_options_t o = {FDTD_FP64, 0, 0};
    if (opt) o = *opt;
    if (o.precision != FDTD_FP64 && o.precision != FDTD_FP32) { rf_set_error("bad precision"); return NULL; }
    if (rf_init() != 0) { rf_set_error("no GPU: %s", rf_error_string()); return NULL; }
    fdtd_sim_t *s = calloc(1, sizeof *s);
    if (!s) return NULL;
    s->prec = o.precision == FDTD_FP32 ? 1 : 0;
    s->es = s->prec ? sizeof(float) : sizeof(double);
    s->nx = g->nx; s->ny = g->ny; s->nz = g->nz;
    s->dx = g->dx; s->dt = g->dt;
    s->per = g->periodic_axes;
    s->geom = rf_geom_of(g);
    s->cells = rf_cells(g);
    s->comp_sum = o.dft_compensated ? 1 : 0;
    /* 0: chosen at the first run (dft_block_of) */
    s->K = o.dft_block ? (int)o.dft_block : 0;
    if (s->K > RF_PDFT_TWS) s->K = RF_PDFT_TWS;
    s->st = rf_stream_create();
    if (!s->st) goto fail;
    if (dalloc(s, &s->blk, 6 * s->cells * s->es)) goto fail;
    for (int q = 0; q < 6; q++) s->f[q] = (char *)s->blk + q * s->cells * s->es;
    if (upload_fields(s, g)) goto fail;
    if (pml) {
        if (rf_cpml_build(pml, s->prec, &s->cp) != 0) goto fail;
        s->has_cpml = 1;
        s->bytes += s->cp.bytes;
    }
    if (mat) {
        if (rf_materials_build(mat, g, s->prec, &s->mat) != 0) goto fail;
        s->has_mat = 1;
        s->bytes += s->mat.bytes;
    }
    memset(&s->vac, 0, sizeof s->vac);
    s->vac.vac_ca = g->dt / (EPS0 * g->dx);
    if (dalloc(s, (void **)&s->d_step, sizeof(long long))) goto fail;
    s->n_partials = 1024;
    if (dalloc(s, (void **)&s->d_energy, (2 * (size_t)s->n_partials + 2) * sizeof(double))) goto fail;
    return s;
fail:
    if (!fdtd_last_error()[0]) rf_set_error("fdtd_sim_create: %s", rf_error_string());
    fdtd_sim_destroy(s);
    return NULL;
}

void fdtd_sim_destroy(fdtd_sim_t *s) {
    if (!s) return;
    if (s->st) rf_sync(s->st);
    rf_graph_destroy(s->graphs[0]);
    rf_graph_destroy(s->graphs[1]);
    rf_free(s->blk);
    rf_free(s->blk2);
    free(s->h_pts);
    if (s->has_cpml) rf_cpml_release(&s->cp);
    if (s->has_mat) rf_materials_release(&s->mat);
    for (size_t w = 0; w < s->n_waves; w++) rf_free(s->waves[w].d_tab);
    free(s->waves);
    free(s->pt_wave);
    free(s->pts);
    rf_free(s->d_pts);
    rf_free(s->d_srcs);
    rf_free(s->d_cold);
    rf_free(s->zpsi);
    rf_free(s->d_ctr);
    rf_free(s->d_nonfinite);
    free(s->planes);
    free(s->probes);
    free(s->freqs_all);
    rf_free(s->d_pidx); rf_free(s->d_pcomp); rf_free(s->d_ring); rf_free(s->d_td);
    rf_free(s->d_fprobe); rf_free(s->d_fw);
    rf_free(s->d_acc_re); rf_free(s->d_acc_im); rf_free(s->d_cmp_re); rf_free(s->d_cmp_im);
    for (size_t m = 0; m < s->n_mons; m++) {
        smon_t *mo = &s->mons[m];
        free(mo->freqs);
        rf_free(mo->d_w);
        for (int w = 0; w < 2; w++) {
            rf_free(mo->ring[w]);
            rf_free(mo->acc_re[w]); rf_free(mo->acc_im[w]);
            rf_free(mo->cmp_re[w]); rf_free(mo->cmp_im[w]);
        }
    }
    free(s->mons);
    rf_free(s->d_tw);
    rf_free(s->d_step);
    rf_free(s->d_energy);
    rf_free(s->d_aux);
    rf_free(s->d_flux);
    free(s->flux);
    for (size_t q = 0; q < s->n_ev; q++) rf_event_destroy(s->ev[q]);
    free(s->ev);
    if (s->st2) {
        rf_sync(s->st2);
        rf_stream_destroy(s->st2);
    }
    if (s->fork) rf_event_destroy(s->fork);
    if (s->join) rf_event_destroy(s->join);
    rf_stream_destroy(s->st);
    free(s);
}

static int add_wave(fdtd_sim_t *s, source_waveform_fn wf, const void *params, double amp, int after_e) {
    wave_t *w = realloc(s->waves, (s->n_waves + 1) * sizeof *w);
    if (!w) return -1;
    s->waves = w;
    wave_t nw = {wf, params, amp, after_e, NULL};
    s->waves[s->n_waves] = nw;
    return (int)s->n_waves++;
}

int fdtd_sim_add_point_source(fdtd_sim_t *s, const source_t *src) {
    if (!s || !src || !src->waveform) { rf_set_error("bad source"); return -1; }
    if (s->started) { rf_set_error("sources must be added before the first run"); return -1; }
    if (src->component > FIELD_EZ) { rf_set_error("sources drive E components only"); return -1; }
    if (src->i > s->nx || src->j > s->ny || src->k > s->nz) { rf_set_error("source outside the grid"); return -1; }
    int w = add_wave(s, src->waveform, src->params, src->amplitude, 0);
    if (w < 0) return -1;
    source_t *p = realloc(s->pts, (s->n_pts + 1) * sizeof *p);
    int *pw = realloc(s->pt_wave, (s->n_pts + 1) * sizeof *pw);
    if (p) s->pts = p;
    if (pw) s->pt_wave = pw;
    if (!p || !pw) return -1;
    s->pts[s->n_pts] = *src;
    s->pt_wave[s->n_pts] = w;
    s->n_pts++;
    return 0;
}

int fdtd_sim_add_plane_source(fdtd_sim_t *s, field_component_t component, int axis, size_t index,
                              source_mode_t mode, int after_e, double amplitude,
                              source_waveform_fn waveform, const void *params) {
    if (!s || component > FIELD_EZ || axis < 0 || axis > 2 || !waveform) { rf_set_error("bad plane source"); return -1; }
    if (s->started) { rf_set_error("sources must be added before the first run"); return -1; }
    size_t n[3] = {s->nx, s->ny, s->nz};
    if (index > n[axis]) { rf_set_error("plane outside the grid"); return -1; }
    int w = add_wave(s, waveform, params, amplitude, after_e ? 1 : 0);
    if (w < 0) return -1;
    splane_t *p = realloc(s->planes, (s->n_planes + 1) * sizeof *p);
    if (!p) return -1;
    s->planes = p;
    splane_t pl;
    memset(&pl, 0, sizeof pl);
    pl.d.comp = (int)component;
    pl.d.axis = axis;
    pl.d.index = (int)index;
    pl.d.hard = mode == SRC_HARD;
    pl.wave = w;
    pl.after_e = after_e ? 1 : 0;
    s->planes[s->n_planes++] = pl;
    return 0;
}

int fdtd_sim_add_probe(fdtd_sim_t *s, field_component_t component, size_t i, size_t j, size_t k,
                       size_t td_capacity, const double *freqs_hz, size_t n_freqs) {
    if (!s || component > FIELD_HZ || i > s->nx || j > s->ny || k > s->nz || (n_freqs && !freqs_hz)) {
        rf_set_error("bad probe");
        return -1;
    }
    if (s->started) { rf_set_error("probes must be added before the first run"); return -1; }
    sprobe_t *p = realloc(s->probes, (s->n_probes + 1) * sizeof *p);
    if (!p) return -1;
    s->probes = p;
    double *fa = realloc(s->freqs_all, (s->n_freq_tot + n_freqs + 1) * sizeof *fa);
    if (!fa) return -1;
    s->freqs_all = fa;
    memcpy(s->freqs_all + s->n_freq_tot, freqs_hz, n_freqs * sizeof(double));
    sprobe_t np;
    np.comp = (int)component;
    fdtd_grid_t tmp = {0};
    tmp.nx = s->nx; tmp.ny = s->ny; tmp.nz = s->nz; tmp.periodic_axes = s->per;
    np.idx = (long long)fdtd_index(&tmp, i, j, k);
    np.td_cap = td_capacity;
    np.nf = n_freqs;
    np.f_off = s->n_freq_tot;
    s->n_freq_tot += n_freqs;
    s->probes[s->n_probes] = np;
    return (int)s->n_probes++;
}

int fdtd_sim_add_plane_monitor(fdtd_sim_t *s, int axis, size_t index, const double *freqs_hz,
                               size_t n_freqs) {
    if (!s || axis < 0 || axis > 2 || !freqs_hz || !n_freqs) { rf_set_error("bad monitor"); return -1; }
    if (s->started) { rf_set_error("monitors must be added before the first run"); return -1; }
    size_t n[3] = {s->nx, s->ny, s->nz};
    if (index > n[axis]) { rf_set_error("monitor plane outside the grid"); return -1; }
    smon_t *nm = realloc(s->mons, (s->n_mons + 1) * sizeof *nm);
    if (!nm) return -1;
    s->mons = nm;
    smon_t *m = &s->mons[s->n_mons];
    memset(m, 0, sizeof *m);
    m->axis = axis;
    m->index = (int)index;
    int a = (axis + 1) % 3 < (axis + 2) % 3 ? (axis + 1) % 3 : (axis + 2) % 3;
    m->comp[0] = a;
    m->comp[1] = 3 - axis - a;
    m->na = (int)n[m->comp[0]];
    m->nb = (int)n[m->comp[1]];
    m->nf = (int)n_freqs;
    m->freqs = malloc(n_freqs * sizeof(double));
    if (!m->freqs) return -1;
    memcpy(m->freqs, freqs_hz, n_freqs * sizeof(double));
    s->n_mons++;
    return (int)s->n_mons - 1;
}

int fdtd_sim_add_flux_plane(fdtd_sim_t *s, int axis, size_t index) {
    if (!s || axis < 0 || axis > 2) { rf_set_error("bad flux plane"); return -1; }
    if (s->started) { rf_set_error("flux planes must be added before the first run"); return -1; }
    const size_t n[3] = {s->nx, s->ny, s->nz};
    if (index >= n[axis]) { rf_set_error("flux plane outside the grid"); return -1; }
    int (*nf)[2] = realloc(s->flux, (size_t)(s->n_flux + 1) * sizeof *nf);
    if (!nf) return -1;
    s->flux = nf;
    s->flux[s->n_flux][0] = axis;
    s->flux[s->n_flux][1] = (int)index;
    return s->n_flux++;
}

/* The fused leapfrog kernel is the default step; RFSYN_FUSED=0 in the
 * environment selects the separate H and E kernels instead. */
static int want_fused(const fdtd_sim_t *s) {
    const char *env = getenv("RFSYN_FUSED");
    if (env && env[0] == '0') return 0;
    /* the kernel addresses a state block and the z-psi store with 32-bit
     * byte offsets */
    const double zb = 6.0 * (double)(s->nx + 1) * (double)(s->ny + 1) * (s->nz + 1) * (double)s->es;
    return 6.0 * (double)s->cells * (double)s->es < 4294967296.0 && zb < 4294967296.0;
}

/* Multiply-high division by d >= 1 for 32-bit n (round-up method of
 * Granlund and Montgomery, "Division by invariant integers using
 * multiplication", 1994): l = ceil(log2 d), m = 2^32 (2^l - d) / d + 1,
 * n / d = (t + ((n - t) >> min(l,1))) >> max(l-1,0), t = mulhi(m, n). */
static void magic_u32(unsigned d, unsigned *m, unsigned s[2]) {
    unsigned l = 0;
    while ((1ull << l) < d) l++;
    *m = (unsigned)((((1ull << l) - d) << 32) / d + 1);
    s[0] = l < 1 ? l : 1;
    s[1] = l > 0 ? l - 1 : 0;
}

/* The fused kernel owns the nodes whose coordinate on every axis lies in
 * the CPML plain range with a one-node margin below it (so the halo it
 * recomputes is plain too); the rest (CPML slabs and margin) form up to 6
 * shell boxes updated node by node. */
static int plan_regions(fdtd_sim_t *s) {
    const size_t n[3] = {s->nx, s->ny, s->nz};
    int top[3];
    /* z CPML is evaluated inside the fused kernel (per lane) when z is a
     * non-periodic axis with a psi slab; x and y slabs go to the shell */
    s->zc = s->has_cpml && !((s->per >> 2) & 1) && s->cp.d.len[2] > 0;
    for (int a = 0; a < 3; a++) {
        top[a] = ((s->per >> a) & 1) ? (int)n[a] : (int)n[a] + 1;
        s->ib0[a] = 0;
        s->ib1[a] = top[a];
        if (s->has_cpml && !((s->per >> a) & 1) && !(a == 2 && s->zc)) {
            const int plo = s->cp.d.plo[a], phi = s->cp.d.phi[a];
            s->ib0[a] = plo == 0 ? 0 : plo + 1;
            s->ib1[a] = phi < top[a] ? phi : top[a];
            if (s->ib1[a] < s->ib0[a]) s->ib1[a] = s->ib0[a];
        }
    }
    rf_shell_t *sh = &s->shell;
    memset(sh, 0, sizeof *sh);
    s->kbrk = s->ib1[2];
    if (s->zc) {
        /* z tiles restart at the upper z slab, so the tiles below it have
         * no CPML lane, unless that takes more tiles than a plain split of
         * the z range (short z ranges) */
        int hz = s->cp.d.hi[2];
        s->kbrk = hz < s->ib0[2] ? s->ib0[2] : (hz > s->ib1[2] ? s->ib1[2] : hz);
        const int kw = RF_LEAP_COLS;
        const int plain = (s->ib1[2] - s->ib0[2] + kw - 1) / kw;
        const int split = (s->kbrk - s->ib0[2] + kw - 1) / kw + (s->ib1[2] - s->kbrk + kw - 1) / kw;
        if (split > plain) s->kbrk = s->ib1[2];
    }
    const int *b0 = s->ib0, *b1 = s->ib1;
    int box[6][2][3] = {
        {{0, 0, 0}, {b0[0], top[1], top[2]}},
        {{b1[0], 0, 0}, {top[0], top[1], top[2]}},
        {{b0[0], 0, 0}, {b1[0], b0[1], top[2]}},
        {{b0[0], b1[1], 0}, {b1[0], top[1], top[2]}},
        {{b0[0], b0[1], 0}, {b1[0], b1[1], b0[2]}},
        {{b0[0], b0[1], b1[2]}, {b1[0], b1[1], top[2]}},
    };
    long long at = 0;
    for (int q = 0; q < 6; q++) {
        long long cnt = 1;
        for (int a = 0; a < 3; a++) cnt *= box[q][1][a] > box[q][0][a] ? box[q][1][a] - box[q][0][a] : 0;
        if (!cnt) continue;
        if (cnt >= (1LL << 32)) return -1;
        for (int a = 0; a < 3; a++) {
            sh->lo[sh->n][a] = box[q][0][a];
            sh->hi[sh->n][a] = box[q][1][a];
        }
        sh->start[sh->n] = at;
        magic_u32((unsigned)(box[q][1][2] - box[q][0][2]), &sh->mk[sh->n], sh->sk[sh->n]);
        magic_u32((unsigned)(box[q][1][1] - box[q][0][1]), &sh->mj[sh->n], sh->sj[sh->n]);
        at += cnt;
        sh->n++;
    }
    sh->start[sh->n] = at;
    return 0;
}

/* Device arrays for probes and monitors, allocated once at the first run. */
/* The DFT block the handle runs with: the requested dft_block, else 32
 * steps, or 64 with plane monitors (their accumulators are read and
 * written once per block); even on the fused step. */
static int dft_block_of(const fdtd_sim_t *s, int fused) {
    int k = s->K ? s->K : (s->n_mons ? 64 : 32);
    if (fused && (k & 1)) k++;   /* a captured block must return to its buffer parity */
    return k;
}

static int finalize(fdtd_sim_t *s) {
    s->fused = want_fused(s);
    s->K = dft_block_of(s, s->fused);
    if (s->fused) {
        /* B starts as a copy of A so nodes no kernel writes agree */
        if (dalloc(s, &s->blk2, 6 * s->cells * s->es) || rf_d2d(s->blk2, s->blk, 6 * s->cells * s->es, s->st))
            return -1;
        for (int q = 0; q < 6; q++) s->f2[q] = (char *)s->blk2 + q * s->cells * s->es;
        if (rf_sync(s->st)) return -1;
        s->post_in_kernel = !(s->has_mat && s->mat.n_tensor);
        rf_leap_cold_t cold;
        memset(&cold, 0, sizeof cold);
        if (s->has_mat) {
            const rf_mat_t *m = &s->mat.d;
            cold.entries = m->entries;
            cold.ade = m->ade;
            cold.J = m->J;
            cold.P = m->P;
        }
        if (dalloc(s, (void **)&s->d_cold, sizeof cold) || rf_h2d(s->d_cold, &cold, sizeof cold, s->st) ||
            rf_sync(s->st))
            return -1;
        if (plan_regions(s)) {
            rf_set_error("a shell box of the fused step holds 2^32 nodes or more");
            return -1;
        }
        if (s->zc && dalloc(s, &s->zpsi, (s->nx + 1) * (s->ny + 1) * (size_t)s->cp.d.len[2] * 6 * s->es))
            return -1;
        if (dalloc(s, (void **)&s->d_ctr, 32 * sizeof(unsigned))) return -1;
        const char *e;
        s->leap_rows = (e = getenv("RFSYN_LEAP_ROWS")) ? atoi(e) : 0;
        s->leap_chunk = (e = getenv("RFSYN_LEAP_CHUNK")) ? atoi(e) : 0;
        s->leap_waves = (e = getenv("RFSYN_LEAP_WAVES")) ? atoi(e) : 0;
        s->leap_zsplit = (e = getenv("RFSYN_LEAP_ZSPLIT")) ? atoi(e) != 0 : 0;
        s->leap_coop = (e = getenv("RFSYN_LEAP_COOP")) ? atoi(e) : 0;
        if (s->leap_coop < 0 || s->leap_coop > 255) s->leap_coop = 0;
        s->ade_split_off = (e = getenv("RFSYN_ADE_SPLIT")) && atoi(e) == 0;
        /* the shell's H update reads state n only and writes H^{n+1/2} of
         * nodes outside the fused box, so it runs on a second stream beside
         * the fused kernel (RFSYN_SHELL_OVERLAP=0: after it, on one stream) */
        if (s->shell.n && !((e = getenv("RFSYN_SHELL_OVERLAP")) && atoi(e) == 0)) {
            s->st2 = rf_stream_create();
            s->fork = rf_event_create_sync();
            s->join = rf_event_create_sync();
            if (!s->st2 || !s->fork || !s->join) {
                rf_set_error("second stream for the shell: %s", rf_error_string());
                return -1;
            }
            s->shell_overlap = 1;
        }
    }
    if (s->n_flux) {
        if (!s->fused) {
            rf_set_error("flux planes need the fused step (state n is kept beside state n+1 there)");
            return -1;
        }
        if (dalloc(s, (void **)&s->d_flux, (size_t)s->n_flux * sizeof(double))) return -1;
    }
    if ((s->n_flux || s->fused) && !s->d_aux &&
        dalloc(s, (void **)&s->d_aux, (5 * (size_t)s->n_partials + 5) * sizeof(double)))
        return -1;
    int P = (int)s->n_probes;
    size_t cap = 0;
    for (size_t p = 0; p < s->n_probes; p++) if (s->probes[p].td_cap > cap) cap = s->probes[p].td_cap;
    s->td_stride = cap;
    if (P) {
        long long *idx = malloc(P * sizeof *idx);
        int *comp = malloc(P * sizeof *comp);
        if (!idx || !comp) { free(idx); free(comp); return -1; }
        for (int p = 0; p < P; p++) { idx[p] = s->probes[p].idx; comp[p] = s->probes[p].comp; }
        int rc = dalloc(s, (void **)&s->d_pidx, P * sizeof *idx) ||
                 dalloc(s, (void **)&s->d_pcomp, P * sizeof *comp) ||
                 rf_h2d(s->d_pidx, idx, P * sizeof *idx, s->st) ||
                 rf_h2d(s->d_pcomp, comp, P * sizeof *comp, s->st) || rf_sync(s->st);
        free(idx);
        free(comp);
        if (rc) return -1;
        if (dalloc(s, (void **)&s->d_ring, (size_t)s->K * P * sizeof(double))) return -1;
        if (cap && dalloc(s, (void **)&s->d_td, cap * P * sizeof(double))) return -1;
    }
    if (s->n_freq_tot) {
        size_t F = s->n_freq_tot;
        int *fp = malloc(F * sizeof *fp);
        double *fw = malloc(F * sizeof *fw);
        if (!fp || !fw) { free(fp); free(fw); return -1; }
        for (size_t p = 0; p < s->n_probes; p++)
            for (size_t q = 0; q < s->probes[p].nf; q++) {
                fp[s->probes[p].f_off + q] = (int)p;
                fw[s->probes[p].f_off + q] = 2.0 * M_PI * s->freqs_all[s->probes[p].f_off + q];
            }
        int rc = dalloc(s, (void **)&s->d_fprobe, F * sizeof *fp) ||
                 dalloc(s, (void **)&s->d_fw, F * sizeof *fw) ||
                 dalloc(s, (void **)&s->d_acc_re, F * sizeof(double)) ||
                 dalloc(s, (void **)&s->d_acc_im, F * sizeof(double)) ||
                 dalloc(s, (void **)&s->d_cmp_re, F * sizeof(double)) ||
                 dalloc(s, (void **)&s->d_cmp_im, F * sizeof(double)) ||
                 rf_h2d(s->d_fprobe, fp, F * sizeof *fp, s->st) ||
                 rf_h2d(s->d_fw, fw, F * sizeof *fw, s->st) || rf_sync(s->st);
        free(fp);
        free(fw);
        if (rc) return -1;
    }
    int max_nf = 0;
    for (size_t q = 0; q < s->n_mons; q++) {
        smon_t *m = &s->mons[q];
        if (m->nf > max_nf) max_nf = m->nf;
        size_t cells = (size_t)m->na * m->nb;
        double *fw = malloc(m->nf * sizeof *fw);
        if (!fw) return -1;
        for (int f = 0; f < m->nf; f++) fw[f] = 2.0 * M_PI * m->freqs[f];
        int rc = dalloc(s, (void **)&m->d_w, m->nf * sizeof *fw) ||
                 rf_h2d(m->d_w, fw, m->nf * sizeof *fw, s->st) || rf_sync(s->st);
        free(fw);
        if (rc) return -1;
        for (int w = 0; w < 2; w++) {
            size_t acc = (size_t)m->nf * cells * sizeof(double);
            if (dalloc(s, &m->ring[w], (size_t)s->K * cells * s->es) ||
                dalloc(s, (void **)&m->acc_re[w], acc) || dalloc(s, (void **)&m->acc_im[w], acc))
                return -1;
        }
    }
    if (max_nf && dalloc(s, (void **)&s->d_tw, (size_t)2 * max_nf * RF_PDFT_TWS * sizeof(double))) return -1;
    s->started = 1;
    return 0;
}

/* The fused step's source table in application order (point sources,
 * pre-E plane sources, then the post-E plane sources it applies itself)
 * and the bounding box of the nodes they touch. */
static int build_src_table(fdtd_sim_t *s) {
    size_t n = s->n_pts + s->n_planes;
    rf_leap_src_t *t = calloc(n ? n : 1, sizeof *t);
    if (!t) return -1;
    int m = 0;
    int box[6] = {INT_MAX, -1, INT_MAX, -1, INT_MAX, -1};
    const int nn[3] = {(int)s->nx, (int)s->ny, (int)s->nz};
    for (size_t p = 0; p < s->n_pts; p++) {
        rf_leap_src_t *e = &t[m++];
        e->plane = 0;
        e->comp = s->h_pts[p].comp;
        e->hard = s->h_pts[p].hard;
        e->after = 0;
        e->i = s->h_pts[p].i;
        e->j = s->h_pts[p].j;
        e->k = s->h_pts[p].k;
        e->wave = s->h_pts[p].wave;
        const int x[3] = {e->i, e->j, e->k};
        for (int a = 0; a < 3; a++) {
            if (x[a] < box[2 * a]) box[2 * a] = x[a];
            if (x[a] > box[2 * a + 1]) box[2 * a + 1] = x[a];
        }
    }
    for (int pass = 0; pass < 2; pass++)
        for (size_t p = 0; p < s->n_planes; p++) {
            const splane_t *pl = &s->planes[p];
            if ((pl->after_e != 0) != pass || (pass && !s->post_in_kernel)) continue;
            rf_leap_src_t *e = &t[m++];
            e->plane = 1;
            e->comp = pl->d.comp;
            e->hard = pl->d.hard;
            e->after = pass;
            e->axis = pl->d.axis;
            e->index = pl->d.index;
            if ((s->per >> e->axis) & 1) e->index %= nn[e->axis];
            e->wave = pl->d.wave;
            for (int a = 0; a < 3; a++) {
                const int lo = a == e->axis ? e->index : 0, hi = a == e->axis ? e->index : nn[a] - 1;
                if (lo < box[2 * a]) box[2 * a] = lo;
                if (hi > box[2 * a + 1]) box[2 * a + 1] = hi;
            }
        }
    s->src_pre = s->src_post = 0;
    for (int q = 0; q < m; q++) {
        if (t[q].after) s->src_post |= 1 << t[q].comp;
        else s->src_pre |= 1 << t[q].comp;
    }
    int rc = 0;
    if (m && !s->d_srcs) rc = dalloc(s, (void **)&s->d_srcs, n * sizeof *t);
    if (!rc && m) rc = rf_h2d(s->d_srcs, t, (size_t)m * sizeof *t, s->st) || rf_sync(s->st);
    free(t);
    if (rc) return -1;
    s->n_srcs = m;
    memcpy(s->src_box, box, sizeof box);
    return 0;
}

/* Waveform tables cover steps [0, need); grown geometrically. */
static int ensure_waves(fdtd_sim_t *s, size_t need) {
    if (need <= s->wave_cap) return 0;
    size_t cap = s->wave_cap ? s->wave_cap : 1024;
    while (cap < need) cap *= 2;
    double *host = malloc(cap * sizeof *host);
    if (!host) return -1;
    for (size_t w = 0; w < s->n_waves; w++) {
        wave_t *wv = &s->waves[w];
        for (size_t n = 0; n < cap; n++) {
            double t = wv->after_e ? (double)(n + 1) * s->dt : (double)n * s->dt;
            host[n] = wv->amp * wv->wf(t, wv->params);
        }
        double *d = NULL;
        if (rf_malloc((void **)&d, cap * sizeof(double)) != 0 ||
            rf_h2d(d, host, cap * sizeof(double), s->st) != 0 || rf_sync(s->st) != 0) {
            rf_free(d);
            free(host);
            return -1;
        }
        rf_free(wv->d_tab);
        wv->d_tab = d;
    }
    free(host);
    s->wave_cap = cap;
    s->wave_gen++;
    if (s->n_pts) {
        rf_point_src_t *hp = calloc(s->n_pts, sizeof *hp);
        if (!hp) return -1;
        fdtd_grid_t tmp = {0};
        tmp.nx = s->nx; tmp.ny = s->ny; tmp.nz = s->nz; tmp.periodic_axes = s->per;
        for (size_t p = 0; p < s->n_pts; p++) {
            hp[p].idx = (long long)fdtd_index(&tmp, s->pts[p].i, s->pts[p].j, s->pts[p].k);
            hp[p].comp = (int)s->pts[p].component;
            hp[p].hard = s->pts[p].mode == SRC_HARD;
            hp[p].i = (int)(hp[p].idx / s->geom.sx);
            hp[p].j = (int)((hp[p].idx % s->geom.sx) / s->geom.sy);
            hp[p].k = (int)(hp[p].idx % s->geom.sy);
            hp[p].wave = s->waves[s->pt_wave[p]].d_tab;
        }
        int rc = 0;
        if (!s->d_pts) rc = dalloc(s, (void **)&s->d_pts, s->n_pts * sizeof *hp);
        if (!rc) rc = rf_h2d(s->d_pts, hp, s->n_pts * sizeof *hp, s->st) || rf_sync(s->st);
        free(s->h_pts);
        s->h_pts = hp;
        if (rc) return -1;
    }
    for (size_t p = 0; p < s->n_planes; p++) s->planes[p].d.wave = s->waves[s->planes[p].wave].d_tab;
    if (s->fused && build_src_table(s)) return -1;
    /* sources pass their tables by value: captured blocks are stale */
    for (int q = 0; q < 2; q++) {
        rf_graph_destroy(s->graphs[q]);
        s->graphs[q] = NULL;
    }
    return 0;
}

static int mark(fdtd_sim_t *s, int q) {
    if (!s->profiling) return 0;
    return rf_event_record(s->ev[s->ev_used + (size_t)q], s->st);
}

/* Box of the split ADE pass: the ADE state box within the fused box.
 * Empty (0) when the ADE update stays in the fused kernel: no dispersive
 * material, or a source inside the fused kernel may touch a dispersive
 * node (its value enters the update before or after the ADE terms). */
static int ade_pass_box(const fdtd_sim_t *s, int lo[3], int hi[3]) {
    if (!s->has_mat || s->ade_split_off || s->mat.d.n_ade <= 0 || s->mat.d.n_ade > 64 || !s->mat.d.J ||
        s->mat.d.n_slots < 1)
        return 0;
    const rf_mat_t *m = &s->mat.d;
    int meets = s->n_srcs > 0;
    for (int a = 0; a < 3; a++) {
        lo[a] = m->b0[a] > s->ib0[a] ? m->b0[a] : s->ib0[a];
        hi[a] = m->b0[a] + m->bn[a] < s->ib1[a] ? m->b0[a] + m->bn[a] : s->ib1[a];
        if (hi[a] <= lo[a]) return 0;
        meets = meets && s->src_box[2 * a] < m->b0[a] + m->bn[a] && s->src_box[2 * a + 1] >= m->b0[a];
    }
    return !meets;
}

/* The fused step: the leapfrog kernel on the plain box (timed in event
 * slots 0-1), then the split ADE pass (slots 2-3) and the shell kernels
 * around the box (slots 4-5; with a second stream the shell's H update
 * runs beside the leapfrog and slots 4-5 time the wait for it and the
 * shell's E update). */
/* The fused kernel's arguments for the current buffers and the handle's
 * tiling requests; *m receives the materials it reads. */
static int fill_leap(fdtd_sim_t *s, rf_leap_t *out, const rf_mat_t **mout, int alo[3], int ahi[3]) {
    rf_leap_t a;
    memset(&a, 0, sizeof a);
    a.g = s->geom;
    a.blk_a = s->blk;
    a.blk_b = s->blk2;
    a.fstride = (long long)s->cells;
    a.cb = s->dt / (MU0 * s->dx);
    a.dx = s->dx;
    for (int q = 0; q < 3; q++) {
        a.ib0[q] = s->ib0[q];
        a.ib1[q] = s->ib1[q];
    }
    a.kbrk = s->kbrk;
    if (s->zc) {
        const rf_cpml_t *c = &s->cp.d;
        a.zc = 1;
        a.zplo = c->plo[2];
        a.zphi = c->phi[2];
        a.zlo = c->lo[2];
        a.zhi = c->hi[2];
        a.zlen = c->len[2];
        a.zk[0] = c->kinv_h[2];
        a.zk[1] = c->b_h[2];
        a.zk[2] = c->a_h[2];
        a.zk[3] = c->kinv_m[2];
        a.zk[4] = c->b_m[2];
        a.zk[5] = c->a_m[2];
        a.zpsi = s->zpsi;
        a.zbytes = (long long)((s->nx + 1) * (s->ny + 1) * (size_t)s->cp.d.len[2] * 6 * s->es);
        a.zpar = s->zpar;
    }
    const rf_mat_t *m = s->has_mat ? &s->mat.d : &s->vac;
    a.tag = m->tag;
    a.tag_bytes = m->tag_bytes;
    a.n_entries = m->n_entries;
    a.entries = m->entries;
    a.ade = m->ade;
    a.n_ade = m->n_ade;
    a.n_slots = m->n_slots;
    a.src_pre = s->src_pre;
    a.src_post = s->src_post;
    a.vac_ca = m->vac_ca;
    a.cold = s->d_cold;
    a.ctr = s->d_ctr;
    a.qpar = s->parity;
    a.srcs = s->d_srcs;
    a.n_srcs = s->n_srcs;
    memcpy(a.src_box, s->src_box, sizeof a.src_box);
    a.d_step = s->d_step;
    a.zsplit = s->leap_zsplit;
    a.coop = s->leap_coop;
    a.ade_split = ade_pass_box(s, alo, ahi);
    if (rf_leap_plan(s->prec, &a, s->leap_rows, s->leap_chunk, s->leap_waves)) {
        const char *why = s->prec == 1 && s->leap_rows == 1 ? rf_leap_zrow_refusal(&a) : NULL;
        rf_set_error("fused step: unsupported tiling (rows %d, chunk %d)%s%s", s->leap_rows, s->leap_chunk,
                     why ? ": " : "", why ? why : "");
        return -1;
    }
    s->plan_rows = a.rows;
    *out = a;
    *mout = m;
    return 0;
}

/* The plane captures of the step as extra blocks of the ADE pass: only
 * when nothing after the pass changes a field (no CPML shell, no tensor
 * edges, no post-E plane source outside the fused kernel), the monitors
 * fit one launch and no E plane they read crosses the pass box [lo, hi).
 * The captures read the new state in f2. */
static int caps_in_ade_pass(const fdtd_sim_t *s, const int lo[3], const int hi[3], rf_caps_t *cs) {
    if (!s->n_mons || 2 * s->n_mons > RF_MAX_CAPS || s->shell.n || (s->has_mat && s->mat.n_tensor)) return 0;
    for (size_t p = 0; p < s->n_planes; p++)
        if (s->planes[p].after_e && !s->post_in_kernel) return 0;
    const size_t n[3] = {s->nx, s->ny, s->nz};
    memset(cs, 0, sizeof *cs);
    for (size_t q = 0; q < s->n_mons; q++) {
        const smon_t *m = &s->mons[q];
        long long index = m->index;
        if ((s->per >> m->axis) & 1) index %= (long long)n[m->axis];
        for (int w = 0; w < 2; w++) {
            if (m->comp[w] < 3 && index >= lo[m->axis] && index < hi[m->axis]) return 0;
            rf_cap_t *c = &cs->c[cs->n++];
            c->f = s->f2[m->comp[w]];
            c->buf = m->ring[w];
            c->axis = m->axis;
            c->index = m->index;
            c->na = m->na;
            c->nb = m->nb;
        }
    }
    return 1;
}

/* The fused step's launches (A -> B), without the buffer swap.  With
 * caps_done given, the step's plane captures join the ADE pass where
 * caps_in_ade_pass allows it, and *caps_done says whether they did. */
static int launch_fused(fdtd_sim_t *s, int *caps_done) {
    rf_leap_t a;
    const rf_mat_t *m;
    int alo[3], ahi[3];
    rf_caps_t cs;
    if (caps_done) *caps_done = 0;
    if (fill_leap(s, &a, &m, alo, ahi)) return -1;
    const rf_cpml_t *cp = s->has_cpml ? &s->cp.d : NULL;
    rf_zpsi_t z = {s->zpsi, s->zpar, s->zc ? s->cp.d.len[2] : 0};
    const rf_zpsi_t *zp = s->zc ? &z : NULL;
    /* the shell's H update beside the fused kernel: it reads state n only
     * and writes H^{n+1/2} and the psi of nodes outside the fused box (z
     * psi: the write slot of its own columns, which the fused kernel only
     * reads from the read slot); the shell's E update waits for both */
    const int overlap = s->shell.n && s->shell_overlap;
    if (overlap && (rf_event_record(s->fork, s->st) || rf_stream_wait(s->st2, s->fork) ||
                    rf_launch_shell_h(s->prec, &s->geom, &s->shell, s->f, s->f2 + 3, a.cb, s->dx, cp, zp, s->st2) ||
                    rf_event_record(s->join, s->st2)))
        return -1;
    if (mark(s, 0) || rf_launch_leapfrog(s->prec, &a, s->st) || mark(s, 1) || mark(s, 2)) return -1;
    if (a.ade_split) {
        const int with_caps = caps_done && caps_in_ade_pass(s, alo, ahi, &cs);
        if (rf_launch_ade_pass(s->prec, &s->geom, alo, ahi, s->f, s->f2, m, with_caps ? &cs : NULL, s->K,
                               s->d_step, s->st))
            return -1;
        if (caps_done) *caps_done = with_caps;
    }
    if (mark(s, 3) || mark(s, 4)) return -1;
    if (s->shell.n) {
        if (overlap ? rf_stream_wait(s->st, s->join)
                    : rf_launch_shell_h(s->prec, &s->geom, &s->shell, s->f, s->f2 + 3, a.cb, s->dx, cp, zp, s->st))
            return -1;
        if (rf_launch_shell_e(s->prec, &s->geom, &s->shell, s->f, s->f2 + 3, s->f2, s->dx, cp, zp, m, s->d_srcs,
                              s->n_srcs, s->d_step, s->st))
            return -1;
    }
    return mark(s, 5);
}

static int enqueue_fused(fdtd_sim_t *s, int *caps_done) {
    if (launch_fused(s, caps_done)) return -1;
    /* B now holds state n+1 */
    for (int q = 0; q < 6; q++) {
        void *t = s->f[q];
        s->f[q] = s->f2[q];
        s->f2[q] = t;
    }
    void *tb = s->blk;
    s->blk = s->blk2;
    s->blk2 = tb;
    s->parity ^= 1;
    s->zpar ^= 1;
    return 0;
}


/* Whether the two-kernel step applies the point sources inside the E
 * update (rf_launch_update_e_src) instead of a kernel of their own: no
 * plane source between them (the sources keep their order), no tensor
 * edges (the E update leaves those to the tensor kernel), and every source
 * on a node and component the E update writes (grid.c:95-117 ranges). */
static int pts_in_update_e(const fdtd_sim_t *s) {
    if (!s->n_pts || !s->h_pts) return 0;
    if (s->has_mat && s->mat.n_tensor) return 0;
    for (size_t p = 0; p < s->n_planes; p++)
        if (!s->planes[p].after_e) return 0;
    const int px = s->per & 1, py = (s->per >> 1) & 1, pz = (s->per >> 2) & 1;
    for (size_t p = 0; p < s->n_pts; p++) {
        const rf_point_src_t *q = &s->h_pts[p];
        if (q->i < 0 || q->j < 0 || q->k < 0 || (size_t)q->i >= s->nx || (size_t)q->j >= s->ny ||
            (size_t)q->k >= s->nz)
            return 0;
        const int i_ok = px || q->i >= 1, j_ok = py || q->j >= 1, k_ok = pz || q->k >= 1;
        const int on = q->comp == 0 ? (j_ok && k_ok) : q->comp == 1 ? (i_ok && k_ok) : (i_ok && j_ok);
        if (!on) return 0;
    }
    return 1;
}

static int enqueue_step(fdtd_sim_t *s) {
    const rf_cpml_t *cp = s->has_cpml ? &s->cp.d : NULL;
    const rf_mat_t *mt = s->has_mat ? &s->mat.d : &s->vac;
    double cb = s->dt / (MU0 * s->dx);
    int caps_done = 0;
    if (s->fused) {
        if (enqueue_fused(s, &caps_done)) return -1;
        if (s->profiling) s->ev_used += 6;
        if (s->has_mat && s->mat.n_tensor &&
            rf_launch_tensor(s->prec, &s->geom, s->f, s->mat.tensor_edges, s->mat.n_tensor, s->st))
            return -1;
        /* with tensor edges the post-E sources follow the tensor update */
        for (size_t p = 0; p < s->n_planes && !s->post_in_kernel; p++)
            if (s->planes[p].after_e &&
                rf_launch_plane_source(s->prec, &s->geom, s->f, &s->planes[p].d, s->d_step, 0, s->st))
                return -1;
        /* flux at n+1/2 from state n (f2) and state n+1 (f) */
        for (int q = 0; q < s->n_flux; q++) {
            const size_t n[3] = {s->nx, s->ny, s->nz};
            const int a = s->flux[q][0];
            const size_t cells = n[(a + 1) % 3] * n[(a + 2) % 3];
            const int blocks = (int)((cells + 255) / 256 < (size_t)s->n_partials ? (cells + 255) / 256
                                                                                 : (size_t)s->n_partials);
            if (rf_launch_flux(s->prec, &s->geom, a, s->flux[q][1], s->f2, s->f, s->d_aux, blocks, s->d_flux + q,
                               s->st))
                return -1;
        }
    } else {
        if (mark(s, 0)) return -1;
        if (rf_launch_update_h(s->prec, &s->geom, s->f, cb, s->dx, cp, s->st)) return -1;
        if (mark(s, 1)) return -1;
        const int pts_in_e = pts_in_update_e(s);
        if (s->n_pts && !pts_in_e &&
            rf_launch_point_sources(s->prec, s->f, s->d_pts, (int)s->n_pts, s->d_step, 0, s->st))
            return -1;
        for (size_t p = 0; p < s->n_planes; p++)
            if (!s->planes[p].after_e &&
                rf_launch_plane_source(s->prec, &s->geom, s->f, &s->planes[p].d, s->d_step, 0, s->st))
                return -1;
        if (mark(s, 2)) return -1;
        if (rf_launch_update_e_src(s->prec, &s->geom, s->f, s->dx, cp, mt, pts_in_e ? s->d_pts : NULL,
                                   pts_in_e ? (int)s->n_pts : 0, s->d_step, s->st))
            return -1;
        if (mark(s, 3) || mark(s, 4) || mark(s, 5)) return -1;
        if (s->profiling) s->ev_used += 6;
        if (s->has_mat && s->mat.n_tensor &&
            rf_launch_tensor(s->prec, &s->geom, s->f, s->mat.tensor_edges, s->mat.n_tensor, s->st))
            return -1;
        for (size_t p = 0; p < s->n_planes; p++)
            if (s->planes[p].after_e &&
                rf_launch_plane_source(s->prec, &s->geom, s->f, &s->planes[p].d, s->d_step, 0, s->st))
                return -1;
    }
    /* the monitors' planes, up to RF_MAX_CAPS per launch */
    rf_caps_t cs;
    memset(&cs, 0, sizeof cs);
    for (size_t q = 0; q < s->n_mons && !caps_done; q++) {
        smon_t *m = &s->mons[q];
        for (int w = 0; w < 2; w++) {
            rf_cap_t *c = &cs.c[cs.n++];
            c->f = s->f[m->comp[w]];
            c->buf = m->ring[w];
            c->axis = m->axis;
            c->index = m->index;
            c->na = m->na;
            c->nb = m->nb;
            if (cs.n == RF_MAX_CAPS || (q + 1 == s->n_mons && w == 1)) {
                if (rf_launch_plane_captures(s->prec, &s->geom, &cs, s->K, s->d_step, s->st)) return -1;
                cs.n = 0;
            }
        }
    }
    return rf_launch_probe_record(s->prec, s->f, s->d_pidx, s->d_pcomp, (int)s->n_probes, s->d_ring, s->K,
                                  s->d_td, (long long)s->td_stride, s->d_step, s->st);
}

static int enqueue_flush(fdtd_sim_t *s, int count) {
    if (count <= 0) return 0;
    if (s->n_freq_tot &&
        rf_launch_probe_dft(s->d_ring, s->K, (int)s->n_probes, s->d_fprobe, (int)s->n_freq_tot, s->d_fw,
                            s->dt, s->d_step, count, s->d_acc_re, s->d_acc_im, s->d_cmp_re, s->d_cmp_im,
                            s->comp_sum, s->st))
        return -1;
    for (size_t q = 0; q < s->n_mons; q++) {
        smon_t *m = &s->mons[q];
        long long cells = (long long)m->na * m->nb;
        for (int w = 0; w < 2; w++)
            if (rf_launch_plane_dft(m->ring[w], s->prec, cells, s->K, m->nf, m->d_w, s->dt, s->d_step, count,
                                    s->d_tw, m->acc_re[w], m->acc_im[w], NULL, NULL, 0, s->st))
                return -1;
    }
    return 0;
}

static int capture_block(fdtd_sim_t *s) {
    if (rf_graph_begin(s->st)) return -1;
    int rc = 0;
    for (int q = 0; q < s->K && !rc; q++) rc = enqueue_step(s);
    if (!rc) rc = enqueue_flush(s, s->K);
    rf_graph_t *g = NULL;
    int rc2 = rf_graph_end(s->st, &g);
    if (rc || rc2) { rf_graph_destroy(g); return -1; }
    /* capture advanced the host-side buffer bookkeeping by K steps; K is
     * even on the fused path, so the parity is back where it started */
    s->graphs[s->parity] = g;
    return 0;
}

/* Folds the recorded event pairs into the totals (synchronises). */
static int prof_collect(fdtd_sim_t *s) {
    if (rf_sync(s->st)) return -1;
    for (size_t q = 0; q + 6 <= s->ev_used; q += 6) {
        for (int w = 0; w < 3; w++) {
            float ms = 0.f;
            if (rf_event_elapsed_ms(s->ev[q + 2 * w], s->ev[q + 2 * w + 1], &ms)) return -1;
            s->prof_ms[w] += ms;
        }
        s->prof_launches++;
    }
    s->ev_used = 0;
    return 0;
}

int fdtd_sim_set_profiling(fdtd_sim_t *s, int on) {
    if (!s) return -1;
    if (on && !s->ev) {
        s->n_ev = 6 * 256;
        s->ev = calloc(s->n_ev, sizeof *s->ev);
        if (!s->ev) return -1;
        for (size_t q = 0; q < s->n_ev; q++)
            if (!(s->ev[q] = rf_event_create())) return -1;
    }
    s->profiling = on ? 1 : 0;
    s->ev_used = 0;
    return 0;
}

int fdtd_sim_kernel_times(fdtd_sim_t *s, double *h_ms, double *e_ms, long *launches) {
    if (!s || !h_ms || !e_ms || !launches) return -1;
    *h_ms = s->prof_ms[0];
    *e_ms = s->prof_ms[1] + s->prof_ms[2];
    *launches = s->prof_launches;
    memset(s->prof_ms, 0, sizeof s->prof_ms);
    s->prof_launches = 0;
    return 0;
}

int fdtd_sim_kernel_times3(fdtd_sim_t *s, double ms[3], long *launches) {
    if (!s || !ms || !launches) return -1;
    for (int w = 0; w < 3; w++) ms[w] = s->prof_ms[w];
    *launches = s->prof_launches;
    memset(s->prof_ms, 0, sizeof s->prof_ms);
    s->prof_launches = 0;
    return 0;
}

/* Set-up before n_steps more steps: first-run allocation and the
 * waveform tables. */
static int run_prepare(fdtd_sim_t *s, size_t n_steps) {
    if (!s->started && finalize(s) != 0) {
        rf_set_error("fdtd_sim_run set-up: %s", rf_error_string());
        return -1;
    }
    if (ensure_waves(s, s->steps + n_steps) != 0) {
        rf_set_error("waveform tables: %s", rf_error_string());
        return -1;
    }
    return 0;
}

/* Enqueues at most `max` of the *left steps on the handle's stream (one
 * replayed graph block, or single steps when a block does not fit or
 * profiling is on) without waiting for them. */
static int run_enqueue(fdtd_sim_t *s, size_t *left, size_t max) {
    size_t budget = max;
    while (*left && budget) {
        if (s->profiling && s->ev_used + 6 > s->n_ev && prof_collect(s) != 0) return -1;
        if (!s->profiling && s->pending == 0 && *left >= (size_t)s->K && budget >= (size_t)s->K) {
            if (!s->graphs[s->parity] && capture_block(s) != 0) {
                /* a HIP error, else the message the failing step left */
                char inner[200];
                snprintf(inner, sizeof inner, "%s", rf_error_string()[0] ? rf_error_string() : fdtd_last_error());
                rf_set_error("graph capture: %s", inner);
                return -1;
            }
            if (rf_graph_launch(s->graphs[s->parity], s->st)) return -1;
            s->steps += (size_t)s->K;
            *left -= (size_t)s->K;
            budget -= (size_t)s->K;
            continue;
        }
        if (enqueue_step(s)) return -1;
        s->steps++;
        (*left)--;
        budget--;
        if (++s->pending == s->K) {
            if (enqueue_flush(s, s->K)) return -1;
            s->pending = 0;
        }
    }
    return 0;
}

/* Flushes the partial DFT block and waits for the handle's stream. */
static int run_finish(fdtd_sim_t *s) {
    if (s->pending) {
        if (enqueue_flush(s, s->pending)) return -1;
        s->pending = 0;
    }
    if (rf_sync(s->st)) {
        rf_set_error("fdtd_sim_run: %s", rf_error_string());
        return -1;
    }
    return s->profiling ? prof_collect(s) : 0;
}

int fdtd_sim_run(fdtd_sim_t *s, size_t n_steps) {
    if (!s) return -1;
    if (run_prepare(s, n_steps)) return -1;
    size_t left = n_steps;
    if (run_enqueue(s, &left, n_steps)) return -1;
    return run_finish(s);
}

/* Tiling autotuner: before the first step, times the fused kernel alone
 * (A -> B, `launches` launches per candidate) for each rows / chunk / z
 * split candidate and keeps the fastest.  The launches write only B, the
 * z-psi store, the in-kernel ADE state and the queue counters, which all
 * hold their initial values before the first step (B = A, the rest zero)
 * and are restored afterwards.  The tiling changes which wavefront updates
 * a node, never the arithmetic, so results do not depend on it. */
/* State before the first step again after timing launches: the queue
 * counters, B = A (a_pristine) or A = B, and the psi memories and ADE
 * state zero. */
static int restore_initial(fdtd_sim_t *s, int a_pristine) {
    const size_t fb = 6 * s->cells * s->es;
    s->parity = 0;
    s->zpar = 0;
    return rf_memset(s->d_ctr, 0, 32 * sizeof(unsigned), s->st) ||
           (a_pristine ? rf_d2d(s->blk2, s->blk, fb, s->st) : rf_d2d(s->blk, s->blk2, fb, s->st)) ||
           (s->zpsi && rf_memset(s->zpsi, 0, (s->nx + 1) * (s->ny + 1) * (size_t)s->cp.d.len[2] * 6 * s->es,
                                 s->st)) ||
           (s->has_cpml && s->cp.psi_sec_bytes &&
            (rf_memset(s->cp.psi_sec[0], 0, s->cp.psi_sec_bytes, s->st) ||
             rf_memset(s->cp.psi_sec[1], 0, s->cp.psi_sec_bytes, s->st))) ||
           (s->has_mat && s->mat.d.J &&
            (rf_memset(s->mat.d.J, 0, s->mat.ade_bytes, s->st) || rf_memset(s->mat.d.P, 0, s->mat.ade_bytes, s->st))) ||
           rf_sync(s->st);
}

/* Milliseconds of `launches` steps' field-update launches on one path (one
 * warm-up step first): the fused step's launches (leapfrog, ADE pass,
 * shell) or the two-kernel step's H and E updates.  The kernels the paths
 * share (sources, tensor, monitors, probes) are left out. */
static int time_path(fdtd_sim_t *s, int fused, int launches, rf_event_t *e0, rf_event_t *e1, float *ms) {
    const rf_cpml_t *cp = s->has_cpml ? &s->cp.d : NULL;
    const rf_mat_t *mt = s->has_mat ? &s->mat.d : &s->vac;
    const double cb = s->dt / (MU0 * s->dx);
    for (int w = 0; w <= launches; w++) {
        if (w == 1 && rf_event_record(e0, s->st)) return -1;
        if (fused) {
            if (launch_fused(s, NULL)) return -1;
            s->parity ^= 1;
            s->zpar ^= 1;
        } else if (rf_launch_update_h(s->prec, &s->geom, s->f, cb, s->dx, cp, s->st) ||
                   rf_launch_update_e(s->prec, &s->geom, s->f, s->dx, cp, mt, s->st)) {
            return -1;
        }
    }
    return rf_event_record(e1, s->st) || rf_sync(s->st) || rf_event_elapsed_ms(e0, e1, ms);
}

int fdtd_sim_autotune(fdtd_sim_t *s, int launches, int chosen[4]) {
    if (!s || launches < 1) return -1;
    if (s->steps) {
        rf_set_error("fdtd_sim_autotune must run before the first step");
        return -1;
    }
    if (run_prepare(s, 1)) return -1;
    if (!s->fused) return 0;   /* nothing to tune on the two-kernel path */
    rf_event_t *e0 = rf_event_create(), *e1 = rf_event_create();
    int rc = (e0 && e1) ? 0 : -1;
    /* FP32 rows 1 is row mode (independent wavefronts, no z split), tried
     * with its own chunk choice and skipped where the box does not allow it */
    const int rows32[3] = {4, 8, 1}, rows64[2] = {2, 4};
    const int chunks[5] = {0, 6, 10, 16, 32};
    const int *rows = s->prec == 0 ? rows64 : rows32;
    const int n_rows = s->prec == 0 ? 2 : 3;
    /* independent; cooperative blocks of 2 or 4 row tiles, in lockstep or
     * with a slack of 1 plane */
    const int coops[5] = {0, 2, 4, 16 + 2, 16 + 4};
    const int r0 = s->leap_rows, c0 = s->leap_chunk, z0 = s->leap_zsplit, g0 = s->leap_coop;
    int best[4] = {r0, c0, z0, g0};
    float best_ms = -1.f;
    for (int ri = 0; ri < n_rows && !rc; ri++)
        for (int ci = 0; ci < 5 && !rc; ci++)
            for (int zi = 0; zi < (s->zc ? 2 : 1) && !rc; zi++)
            for (int gi = 0; gi < 5 && !rc; gi++) {
                const int zrow = s->prec == 1 && rows[ri] == 1;
                if (zrow && (ci > 0 || zi > 0 || gi > 0)) continue;
                s->leap_rows = rows[ri];
                s->leap_chunk = chunks[ci];
                s->leap_zsplit = zi;
                s->leap_coop = coops[gi];
                rf_leap_t a;
                const rf_mat_t *m;
                int alo[3], ahi[3];
                if (zrow && fill_leap(s, &a, &m, alo, ahi)) continue;   /* not available for this box */
                for (int w = 0; w <= launches && !rc; w++) {   /* one warm-up launch */
                    if (w == 1 && rf_event_record(e0, s->st)) rc = -1;
                    if (!rc && fill_leap(s, &a, &m, alo, ahi)) rc = -1;
                    if (!rc && rf_launch_leapfrog(s->prec, &a, s->st)) rc = -1;
                    s->parity ^= 1;   /* alternate the counter sets as the step does */
                    s->zpar ^= 1;
                }
                float ms = 0.f;
                if (!rc && (rf_event_record(e1, s->st) || rf_sync(s->st) || rf_event_elapsed_ms(e0, e1, &ms)))
                    rc = -1;
                if (!rc && (best_ms < 0.f || ms < best_ms)) {
                    best_ms = ms;
                    best[0] = rows[ri];
                    best[1] = chunks[ci];
                    best[2] = zi;
                    best[3] = coops[gi];
                }
            }
    /* restore the state before the first step */
    if (!rc && restore_initial(s, 1)) rc = -1;
    s->leap_rows = rc ? r0 : best[0];
    s->leap_chunk = rc ? c0 : best[1];
    s->leap_zsplit = rc ? z0 : best[2];
    s->leap_coop = rc ? g0 : best[3];
    rf_event_destroy(e0);
    rf_event_destroy(e1);
    if (chosen) {
        chosen[0] = s->leap_rows;
        chosen[1] = s->leap_chunk;
        chosen[2] = s->leap_zsplit;
        chosen[3] = s->leap_coop;
    }
    return rc;
}

/* The fused step at its tiling against the two-kernel step, for a handle
 * run on its own: the two-kernel step wins on small grids whose fused box
 * holds few of the cells (the CPML shell is updated node by node, and a
 * lone small step is bound by launch latency).  Both give the same bits.
 * Kept fused with flux planes (they need the fused step's pair of states)
 * or when RFSYN_FUSED is set. */
int fdtd_sim_select_path(fdtd_sim_t *s, int launches, int *two_kernel) {
    if (!s || launches < 1) return -1;
    if (s->steps) {
        rf_set_error("fdtd_sim_select_path must run before the first step");
        return -1;
    }
    if (run_prepare(s, 1)) return -1;
    if (s->fused && s->n_flux == 0 && !getenv("RFSYN_FUSED")) {
        rf_event_t *e0 = rf_event_create(), *e1 = rf_event_create();
        int rc = (e0 && e1) ? 0 : -1;
        float t_fused = 0.f, t_two = 0.f;
        if (!rc && (time_path(s, 1, launches, e0, e1, &t_fused) || restore_initial(s, 1) ||
                    time_path(s, 0, launches, e0, e1, &t_two) || restore_initial(s, 0)))
            rc = -1;
        rf_event_destroy(e0);
        rf_event_destroy(e1);
        if (rc) return -1;
        if (t_two < 0.95f * t_fused) s->fused = 0;
    }
    if (two_kernel) *two_kernel = !s->fused;
    return 0;
}

static int batch_run(fdtd_batch_t *b, size_t n_steps);

int fdtd_sim_run_many(fdtd_sim_t *const *sims, size_t n_sims, size_t n_steps) {
    if (!sims && n_sims) return -1;
    for (size_t q = 0; q < n_sims; q++) {
        if (!sims[q]) return -1;
        for (size_t r = 0; r < q; r++)
            if (sims[r] == sims[q]) {
                rf_set_error("fdtd_sim_run_many: a handle appears twice");
                return -1;
            }
    }
    /* packed launches for all scenes when they allow it (a scene that
     * cannot be packed is found before any step is enqueued) */
    const char *pe = getenv("RFSYN_PACKED");
    if (n_sims > 1 && !(pe && pe[0] == '0')) {
        fdtd_batch_t *b = fdtd_batch_create(sims, n_sims);
        if (b) {
            const int rc = batch_run(b, n_steps);
            fdtd_batch_destroy(b);
            if (rc != -2) return rc ? -1 : 0;
        }
    }
    /* first-run set-up (finalize may enable the shell's second stream) */
    for (size_t q = 0; q < n_sims; q++)
        if (run_prepare(sims[q], n_steps)) return -1;
    /* scenes stepped together already overlap one another: a second stream
     * per handle only adds cross-stream waits between the few hardware
     * queues, so batched handles update their shell on their one stream
     * (graphs captured with the fork are dropped and captured again) */
    for (size_t q = 0; q < n_sims && n_sims > 1; q++) {
        fdtd_sim_t *s = sims[q];
        if (!s->shell_overlap) continue;
        if (rf_sync(s->st)) return -1;
        s->shell_overlap = 0;
        for (int p = 0; p < 2; p++) {
            rf_graph_destroy(s->graphs[p]);
            s->graphs[p] = NULL;
        }
    }
    /* one block of each handle in turn: each handle's stream runs on its
     * own, so small scenes fill the GPU together */
    size_t *left = malloc((n_sims ? n_sims : 1) * sizeof *left);
    if (!left) return -1;
    for (size_t q = 0; q < n_sims; q++) left[q] = n_steps;
    int rc = 0, busy = 1;
    while (!rc && busy) {
        busy = 0;
        for (size_t q = 0; q < n_sims && !rc; q++) {
            if (!left[q]) continue;
            rc = run_enqueue(sims[q], &left[q], (size_t)sims[q]->K);
            busy |= left[q] != 0;
        }
    }
    for (size_t q = 0; q < n_sims; q++) {
        const int r = run_finish(sims[q]);
        if (!rc) rc = r;
    }
    free(left);
    return rc;
}

/* ------------------------------------------------------------ batches */

/* Scenes stepped by packed launches (fdtd_batch_*): per step one fused
 * leapfrog launch over the boxes of every scene of a variant (rows per
 * tile, tag width, z CPML), one launch each for the shells' H and E
 * updates, one for all probe gathers, and per block of K steps one for
 * all point-probe DFTs.  A scene's other launches of the step (split ADE
 * pass, tensor edges, post-E plane sources, flux planes, plane-monitor
 * captures and DFTs) run in their place of the step's order.  Descriptors
 * of every scene for both buffer parities live in device tables, so a
 * block of K steps (K even) is one captured graph for the whole batch. */
#define BATCH_LDS_WAVE 8192     /* LDS bytes per wavefront of the packed leapfrog */

typedef struct {
    int rows, tag_bytes, zc;
    int n;                      /* scenes of the variant */
    int *scene;                 /* their batch indices */
    rf_leap_t *d_tab[2];
    int *d_base;
    unsigned short *d_item_scene;
    unsigned *d_ctr;
    rf_leap_pack_t pk;
} bgroup_t;

typedef struct {
    int tag_bytes, has_cpml;
    int n_blocks;
    rf_shell_job_t *d_jobs[2];
    unsigned short *d_block_job;
} bshell_t;

typedef struct {
    int tag_bytes, ns_class;
    int n_blocks;
    size_t lds;
    rf_ade_job_t *d_jobs[2];
    unsigned short *d_block_job;
} bade_t;

typedef struct {
    int ade;                    /* split ADE pass box */
    int alo[3], ahi[3];
    size_t expect;              /* the handle's steps after the batch's last run */
    unsigned wave_gen;          /* its wave tables when the tables were built */
    int parity;                 /* its buffer parity at phase 0 */
} bscene_t;

struct fdtd_batch {
    fdtd_sim_t **sims;
    size_t n;
    int prec, K;
    rf_stream_t *st;
    bscene_t *sc;
    bgroup_t *groups;
    int n_groups;
    bshell_t *shells;
    int n_shells;
    bade_t *ades;
    int n_ades;
    rf_probe_job_t *d_probe[2];
    rf_pdft_job_t *d_pdft;
    int max_freq;
    int built;
    int phase;                  /* steps through the batch since the tables were built, mod 2 */
    rf_graph_t *graphs[2];      /* captured block per phase */
    /* the shells' H update on a second stream beside the leapfrog launches
     * (it reads state n only and writes nodes outside the fused boxes) */
    rf_stream_t *st2;
    rf_event_t *fork, *join;
};

static void batch_drop_graphs(fdtd_batch_t *b) {
    for (int q = 0; q < 2; q++) {
        rf_graph_destroy(b->graphs[q]);
        b->graphs[q] = NULL;
    }
}

static void batch_free_tables(fdtd_batch_t *b) {
    batch_drop_graphs(b);
    for (int g = 0; g < b->n_groups; g++) {
        bgroup_t *gr = &b->groups[g];
        free(gr->scene);
        rf_free(gr->d_tab[0]);
        rf_free(gr->d_tab[1]);
        rf_free(gr->d_base);
        rf_free(gr->d_item_scene);
        rf_free(gr->d_ctr);
    }
    free(b->groups);
    b->groups = NULL;
    b->n_groups = 0;
    for (int q = 0; q < b->n_shells; q++) {
        rf_free(b->shells[q].d_jobs[0]);
        rf_free(b->shells[q].d_jobs[1]);
        rf_free(b->shells[q].d_block_job);
    }
    free(b->shells);
    b->shells = NULL;
    b->n_shells = 0;
    for (int q = 0; q < b->n_ades; q++) {
        rf_free(b->ades[q].d_jobs[0]);
        rf_free(b->ades[q].d_jobs[1]);
        rf_free(b->ades[q].d_block_job);
    }
    free(b->ades);
    b->ades = NULL;
    b->n_ades = 0;
    rf_free(b->d_probe[0]);
    rf_free(b->d_probe[1]);
    rf_free(b->d_pdft);
    b->d_probe[0] = b->d_probe[1] = NULL;
    b->d_pdft = NULL;
    b->built = 0;
}

void fdtd_batch_destroy(fdtd_batch_t *b) {
    if (!b) return;
    if (b->st) rf_sync(b->st);
    batch_free_tables(b);
    free(b->sc);
    free(b->sims);
    if (b->st2) {
        rf_sync(b->st2);
        rf_stream_destroy(b->st2);
    }
    if (b->fork) rf_event_destroy(b->fork);
    if (b->join) rf_event_destroy(b->join);
    rf_stream_destroy(b->st);
    free(b);
}

/* The host-side buffer swap of one fused step (enqueue_fused). */
static void swap_state(fdtd_sim_t *s) {
    for (int q = 0; q < 6; q++) {
        void *t = s->f[q];
        s->f[q] = s->f2[q];
        s->f2[q] = t;
    }
    void *tb = s->blk;
    s->blk = s->blk2;
    s->blk2 = tb;
    s->parity ^= 1;
    s->zpar ^= 1;
}

/* The fused kernel's descriptor of s for a packed launch: one part (no z
 * split), independent wavefronts, tables within BATCH_LDS_WAVE. */
static int packed_leap(fdtd_sim_t *s, rf_leap_t *a, int alo[3], int ahi[3], int *ade) {
    const int zs = s->leap_zsplit, co = s->leap_coop, rows = s->leap_rows;
    const rf_mat_t *m;
    s->leap_zsplit = 0;
    s->leap_coop = 0;
    if (s->prec == 1 && rows == 1) s->leap_rows = 0;   /* row mode has no packed variant */
    const int rc = fill_leap(s, a, &m, alo, ahi);
    s->leap_zsplit = zs;
    s->leap_coop = co;
    s->leap_rows = rows;
    if (rc) return -1;
    if (a->n_parts != 1) return -1;
    for (int q = 0; q < 4; q++) a->tkr[q] = a->ptk[0][q];
    a->zc = a->pzc[0];
    a->n_items = a->pitems[0];
    a->gz = a->gy = a->gk = 0;
    while (a->ncache > 0 && rf_leap_lds_bytes(s->prec, a) > BATCH_LDS_WAVE) a->ncache--;
    if (rf_leap_lds_bytes(s->prec, a) > BATCH_LDS_WAVE) return -1;
    *ade = a->ade_split;
    return 0;
}

static int h2d_new(void **dst, const void *src, size_t bytes, rf_stream_t *st) {
    if (rf_malloc(dst, bytes ? bytes : 16)) return -1;
    return bytes ? rf_h2d(*dst, src, bytes, st) : 0;
}

/* Builds the device tables from the handles' current state (phase 0). */
static int batch_build(fdtd_batch_t *b) {
    batch_free_tables(b);
    b->max_freq = 0;
    const size_t n = b->n;
    int rc = -1;
    rf_leap_t *tab[2] = {calloc(n, sizeof(rf_leap_t)), calloc(n, sizeof(rf_leap_t))};
    rf_shell_job_t *jobs[2] = {calloc(n, sizeof(rf_shell_job_t)), calloc(n, sizeof(rf_shell_job_t))};
    rf_probe_job_t *pj[2] = {calloc(n, sizeof(rf_probe_job_t)), calloc(n, sizeof(rf_probe_job_t))};
    rf_pdft_job_t *dj = calloc(n, sizeof(rf_pdft_job_t));
    int *gkey = calloc(n, sizeof(int));
    unsigned short *bj = NULL;
    int *base = NULL;
    unsigned short *isc = NULL;
    if (!tab[0] || !tab[1] || !jobs[0] || !jobs[1] || !pj[0] || !pj[1] || !dj || !gkey) goto out;
    for (size_t q = 0; q < n; q++) {
        fdtd_sim_t *s = b->sims[q];
        bscene_t *c = &b->sc[q];
        for (int ph = 0; ph < 2; ph++) {
            int alo[3], ahi[3], ade = 0;
            if (packed_leap(s, &tab[ph][q], alo, ahi, &ade)) {
                if (ph) swap_state(s);
                rf_set_error("batch: scene %zu: fused step not packable", q);
                goto out;
            }
            if (ph == 0) {
                c->ade = ade;
                memcpy(c->alo, alo, sizeof alo);
                memcpy(c->ahi, ahi, sizeof ahi);
            }
            rf_shell_job_t *j = &jobs[ph][q];
            j->g = s->geom;
            j->sh = s->shell;
            for (int f = 0; f < 6; f++) j->fa[f] = s->f[f];
            for (int f = 0; f < 3; f++) {
                j->hb[f] = s->f2[3 + f];
                j->eb[f] = s->f2[f];
            }
            j->cb = s->dt / (MU0 * s->dx);
            j->dx = s->dx;
            if (s->has_cpml) j->cp = s->cp.d;
            if (s->zc) {
                j->zs.zpsi = s->zpsi;
                j->zs.zpar = s->zpar;
                j->zs.zlen = s->cp.d.len[2];
            }
            j->m = s->has_mat ? s->mat.d : s->vac;
            j->srcs = s->d_srcs;
            j->n_srcs = s->n_srcs;
            j->d_step = s->d_step;
            /* the probe gather follows the swap */
            swap_state(s);
            rf_probe_job_t *p = &pj[ph][q];
            for (int f = 0; f < 6; f++) p->f[f] = s->f[f];
            p->idx = s->d_pidx;
            p->comp = s->d_pcomp;
            p->n = (int)s->n_probes;
            p->ring = s->d_ring;
            p->ring_len = s->K;
            p->td = s->d_td;
            p->td_cap = (long long)s->td_stride;
            p->d_step = s->d_step;
        }
        /* two swaps: back at phase 0 */
        c->expect = s->steps;
        c->wave_gen = s->wave_gen;
        c->parity = s->parity;
        rf_pdft_job_t *d = &dj[q];
        d->ring = s->d_ring;
        d->ring_len = s->K;
        d->n_probes = (int)s->n_probes;
        d->freq_probe = s->d_fprobe;
        d->n_freq = (int)s->n_freq_tot;
        d->freq_w = s->d_fw;
        d->dt = s->dt;
        d->d_step = s->d_step;
        d->acc_re = s->d_acc_re;
        d->acc_im = s->d_acc_im;
        d->cmp_re = s->d_cmp_re;
        d->cmp_im = s->d_cmp_im;
        d->comp_sum = s->comp_sum;
        if (d->n_freq > b->max_freq) b->max_freq = d->n_freq;
        const rf_leap_t *a = &tab[0][q];
        gkey[q] = (a->rows << 8) | ((a->tag ? a->tag_bytes : 0) << 1) | a->zc;
    }
    /* leapfrog groups, one per variant, scenes in batch order */
    b->groups = calloc(n, sizeof(bgroup_t));
    if (!b->groups) goto out;
    for (size_t q = 0; q < n; q++) {
        int g = 0;
        while (g < b->n_groups && (b->groups[g].rows << 8 | b->groups[g].tag_bytes << 1 | b->groups[g].zc) != gkey[q])
            g++;
        if (g == b->n_groups) {
            bgroup_t *gr = &b->groups[b->n_groups++];
            gr->rows = gkey[q] >> 8;
            gr->tag_bytes = (gkey[q] >> 1) & 0x7f;
            gr->zc = gkey[q] & 1;
            gr->scene = calloc(n, sizeof(int));
            if (!gr->scene) goto out;
        }
        b->groups[g].scene[b->groups[g].n++] = (int)q;
    }
    for (int g = 0; g < b->n_groups; g++) {
        bgroup_t *gr = &b->groups[g];
        long long items = 0;
        int lds = 0;
        for (int q = 0; q < gr->n; q++) {
            items += tab[0][gr->scene[q]].n_items;
            const int l = (int)rf_leap_lds_bytes(b->prec, &tab[0][gr->scene[q]]);
            if (l > lds) lds = l;
        }
        if (items <= 0 || items > INT_MAX) {
            rf_set_error("batch: %lld work items in one launch", items);
            goto out;
        }
        free(base);
        free(isc);
        base = malloc(gr->n * sizeof *base);
        isc = malloc((size_t)items * sizeof *isc);
        rf_leap_t *gt[2] = {malloc(gr->n * sizeof(rf_leap_t)), malloc(gr->n * sizeof(rf_leap_t))};
        if (!base || !isc || !gt[0] || !gt[1]) {
            free(gt[0]);
            free(gt[1]);
            goto out;
        }
        long long at = 0;
        for (int q = 0; q < gr->n; q++) {
            base[q] = (int)at;
            for (int i = 0; i < tab[0][gr->scene[q]].n_items; i++) isc[at + i] = (unsigned short)q;
            at += tab[0][gr->scene[q]].n_items;
            gt[0][q] = tab[0][gr->scene[q]];
            gt[1][q] = tab[1][gr->scene[q]];
        }
        int e = h2d_new((void **)&gr->d_tab[0], gt[0], gr->n * sizeof(rf_leap_t), b->st) ||
                h2d_new((void **)&gr->d_tab[1], gt[1], gr->n * sizeof(rf_leap_t), b->st) ||
                h2d_new((void **)&gr->d_base, base, gr->n * sizeof *base, b->st) ||
                h2d_new((void **)&gr->d_item_scene, isc, (size_t)items * sizeof *isc, b->st) ||
                rf_malloc((void **)&gr->d_ctr, 16 * sizeof(unsigned)) || rf_sync(b->st);
        free(gt[0]);
        free(gt[1]);
        if (e) goto out;
        rf_leap_pack_t *pk = &gr->pk;
        memset(pk, 0, sizeof *pk);
        pk->item_base = gr->d_base;
        pk->item_scene = gr->d_item_scene;
        pk->n_items = (int)items;
        pk->n_scenes = gr->n;
        pk->ctr = gr->d_ctr;
        pk->rows = gr->rows;
        pk->tag_bytes = gr->tag_bytes;
        pk->zc = gr->zc;
        pk->lds_wave = lds;
        if (rf_leap_pack_plan(b->prec, pk)) {
            rf_set_error("batch: packed leapfrog plan");
            goto out;
        }
    }
    /* shell groups by (tag width, CPML), blocks in batch order */
    b->shells = calloc(n, sizeof(bshell_t));
    if (!b->shells) goto out;
    for (size_t q = 0; q < n; q++) {
        fdtd_sim_t *s = b->sims[q];
        if (!s->shell.n || !s->shell.start[s->shell.n]) continue;
        const int tb = jobs[0][q].m.tag ? jobs[0][q].m.tag_bytes : 0;
        int g = 0;
        while (g < b->n_shells && !(b->shells[g].tag_bytes == tb && b->shells[g].has_cpml == s->has_cpml)) g++;
        if (g == b->n_shells) {
            b->shells[g].tag_bytes = tb;
            b->shells[g].has_cpml = s->has_cpml;
            b->n_shells++;
        }
    }
    for (int g = 0; g < b->n_shells; g++) {
        bshell_t *sh = &b->shells[g];
        rf_shell_job_t *gj[2] = {calloc(n, sizeof(rf_shell_job_t)), calloc(n, sizeof(rf_shell_job_t))};
        if (!gj[0] || !gj[1]) {
            free(gj[0]);
            free(gj[1]);
            goto out;
        }
        long long blocks = 0;
        int nj = 0;
        for (size_t q = 0; q < n; q++) {
            fdtd_sim_t *s = b->sims[q];
            if (!s->shell.n || !s->shell.start[s->shell.n]) continue;
            const int tb = jobs[0][q].m.tag ? jobs[0][q].m.tag_bytes : 0;
            if (tb != sh->tag_bytes || s->has_cpml != sh->has_cpml) continue;
            for (int ph = 0; ph < 2; ph++) {
                gj[ph][nj] = jobs[ph][q];
                gj[ph][nj].block0 = (unsigned)blocks;
            }
            blocks += (s->shell.start[s->shell.n] + 255) / 256;
            nj++;
        }
        if (blocks > INT_MAX || nj > 65535) {
            free(gj[0]);
            free(gj[1]);
            rf_set_error("batch: shell launch too large");
            goto out;
        }
        free(bj);
        bj = malloc((size_t)blocks * sizeof *bj);
        if (!bj) {
            free(gj[0]);
            free(gj[1]);
            goto out;
        }
        for (int q = 0; q < nj; q++) {
            const long long end = q + 1 < nj ? gj[0][q + 1].block0 : blocks;
            for (long long k = gj[0][q].block0; k < end; k++) bj[k] = (unsigned short)q;
        }
        sh->n_blocks = (int)blocks;
        int e = h2d_new((void **)&sh->d_jobs[0], gj[0], nj * sizeof(rf_shell_job_t), b->st) ||
                h2d_new((void **)&sh->d_jobs[1], gj[1], nj * sizeof(rf_shell_job_t), b->st) ||
                h2d_new((void **)&sh->d_block_job, bj, (size_t)blocks * sizeof *bj, b->st) || rf_sync(b->st);
        free(gj[0]);
        free(gj[1]);
        if (e) goto out;
    }
    /* split ADE passes by (tag width, pole class), blocks in batch order */
    b->ades = calloc(n, sizeof(bade_t));
    if (!b->ades) goto out;
    for (size_t q = 0; q < n; q++) {
        fdtd_sim_t *s = b->sims[q];
        if (!b->sc[q].ade) continue;
        const int tb = s->mat.d.tag_bytes, nc = rf_ade_pass_class(s->mat.d.n_slots);
        int g = 0;
        while (g < b->n_ades && !(b->ades[g].tag_bytes == tb && b->ades[g].ns_class == nc)) g++;
        if (g == b->n_ades) {
            b->ades[g].tag_bytes = tb;
            b->ades[g].ns_class = nc;
            b->n_ades++;
        }
    }
    for (int g = 0; g < b->n_ades; g++) {
        bade_t *ad = &b->ades[g];
        rf_ade_job_t *aj[2] = {calloc(n, sizeof(rf_ade_job_t)), calloc(n, sizeof(rf_ade_job_t))};
        if (!aj[0] || !aj[1]) {
            free(aj[0]);
            free(aj[1]);
            goto out;
        }
        long long blocks = 0;
        int nj = 0;
        for (size_t q = 0; q < n; q++) {
            fdtd_sim_t *s = b->sims[q];
            const bscene_t *c = &b->sc[q];
            if (!c->ade || s->mat.d.tag_bytes != ad->tag_bytes || rf_ade_pass_class(s->mat.d.n_slots) != ad->ns_class)
                continue;
            const long long cnt = (long long)(c->ahi[0] - c->alo[0]) * (c->ahi[1] - c->alo[1]) * (c->ahi[2] - c->alo[2]);
            if (cnt <= 0 || cnt >= (1LL << 31)) continue;
            for (int ph = 0; ph < 2; ph++) {
                rf_ade_job_t *j = &aj[ph][nj];
                j->g = s->geom;
                for (int a = 0; a < 3; a++) j->lo[a] = c->alo[a];
                j->nj = (unsigned)(c->ahi[1] - c->alo[1]);
                j->nk = (unsigned)(c->ahi[2] - c->alo[2]);
                j->count = (unsigned)cnt;
                for (int f = 0; f < 3; f++) {
                    j->ea[f] = s->f[f];
                    j->eb[f] = s->f2[f];
                }
                j->m = s->mat.d;
                j->ncache = s->mat.d.n_entries < 4096 ? s->mat.d.n_entries : 4096;
                j->block0 = (unsigned)blocks;
                swap_state(s);
            }
            const size_t ade_sz = s->prec == 0 ? sizeof(rf_ade64_t) : sizeof(rf_ade32_t);
            const size_t lds = (size_t)s->mat.d.n_ade * ade_sz + (size_t)aj[0][nj].ncache * 8;
            if (lds > ad->lds) ad->lds = lds;
            blocks += (cnt + 255) / 256;
            nj++;
        }
        if (blocks > INT_MAX || nj > 65535) {
            free(aj[0]);
            free(aj[1]);
            rf_set_error("batch: ADE launch too large");
            goto out;
        }
        free(bj);
        bj = malloc((size_t)(blocks ? blocks : 1) * sizeof *bj);
        if (!bj) {
            free(aj[0]);
            free(aj[1]);
            goto out;
        }
        for (int q = 0; q < nj; q++) {
            const long long end = q + 1 < nj ? aj[0][q + 1].block0 : blocks;
            for (long long k = aj[0][q].block0; k < end; k++) bj[k] = (unsigned short)q;
        }
        ad->n_blocks = (int)blocks;
        int e = h2d_new((void **)&ad->d_jobs[0], aj[0], nj * sizeof(rf_ade_job_t), b->st) ||
                h2d_new((void **)&ad->d_jobs[1], aj[1], nj * sizeof(rf_ade_job_t), b->st) ||
                h2d_new((void **)&ad->d_block_job, bj, (size_t)blocks * sizeof *bj, b->st) || rf_sync(b->st);
        free(aj[0]);
        free(aj[1]);
        if (e) goto out;
    }
    if (h2d_new((void **)&b->d_probe[0], pj[0], n * sizeof(rf_probe_job_t), b->st) ||
        h2d_new((void **)&b->d_probe[1], pj[1], n * sizeof(rf_probe_job_t), b->st) ||
        h2d_new((void **)&b->d_pdft, dj, n * sizeof(rf_pdft_job_t), b->st) || rf_sync(b->st))
        goto out;
    b->phase = 0;
    b->built = 1;
    rc = 0;
out:
    free(tab[0]);
    free(tab[1]);
    free(jobs[0]);
    free(jobs[1]);
    free(pj[0]);
    free(pj[1]);
    free(dj);
    free(gkey);
    free(bj);
    free(base);
    free(isc);
    if (rc) batch_free_tables(b);
    return rc;
}

/* The launches of one step of every scene at the batch's phase; the
 * handles' host state advances by one step (buffer swap). */
static int batch_enqueue_step(fdtd_batch_t *b) {
    const int ph = b->phase;
    const int overlap = b->st2 && b->n_shells;
    if (overlap) {
        if (rf_event_record(b->fork, b->st) || rf_stream_wait(b->st2, b->fork)) return -1;
        for (int g = 0; g < b->n_shells; g++) {
            const bshell_t *sh = &b->shells[g];
            if (rf_launch_shell_h_packed(b->prec, sh->d_jobs[ph], sh->d_block_job, sh->n_blocks, sh->has_cpml,
                                         b->st2))
                return -1;
        }
        if (rf_event_record(b->join, b->st2)) return -1;
    }
    for (int g = 0; g < b->n_groups; g++) {
        rf_leap_pack_t pk = b->groups[g].pk;
        pk.tab = b->groups[g].d_tab[ph];
        pk.qpar = ph;
        if (rf_launch_leapfrog_packed(b->prec, &pk, b->st)) return -1;
    }
    for (int g = 0; g < b->n_ades; g++) {
        const bade_t *ad = &b->ades[g];
        if (rf_launch_ade_pass_packed(b->prec, ad->d_jobs[ph], ad->d_block_job, ad->n_blocks, ad->tag_bytes,
                                      ad->ns_class, ad->lds, b->st))
            return -1;
    }
    for (int g = 0; g < b->n_shells && !overlap; g++) {
        const bshell_t *sh = &b->shells[g];
        if (rf_launch_shell_h_packed(b->prec, sh->d_jobs[ph], sh->d_block_job, sh->n_blocks, sh->has_cpml, b->st))
            return -1;
    }
    if (overlap && rf_stream_wait(b->st, b->join)) return -1;
    for (int g = 0; g < b->n_shells; g++) {
        const bshell_t *sh = &b->shells[g];
        if (rf_launch_shell_e_packed(b->prec, sh->d_jobs[ph], sh->d_block_job, sh->n_blocks, sh->tag_bytes,
                                     sh->has_cpml, b->st))
            return -1;
    }
    for (size_t q = 0; q < b->n; q++) {
        fdtd_sim_t *s = b->sims[q];
        swap_state(s);
        if (s->has_mat && s->mat.n_tensor &&
            rf_launch_tensor(s->prec, &s->geom, s->f, s->mat.tensor_edges, s->mat.n_tensor, b->st))
            return -1;
        for (size_t p = 0; p < s->n_planes && !s->post_in_kernel; p++)
            if (s->planes[p].after_e &&
                rf_launch_plane_source(s->prec, &s->geom, s->f, &s->planes[p].d, s->d_step, 0, b->st))
                return -1;
        for (int f = 0; f < s->n_flux; f++) {
            const size_t nn[3] = {s->nx, s->ny, s->nz};
            const int a = s->flux[f][0];
            const size_t cells = nn[(a + 1) % 3] * nn[(a + 2) % 3];
            const int blocks = (int)((cells + 255) / 256 < (size_t)s->n_partials ? (cells + 255) / 256
                                                                                  : (size_t)s->n_partials);
            if (rf_launch_flux(s->prec, &s->geom, a, s->flux[f][1], s->f2, s->f, s->d_aux, blocks, s->d_flux + f,
                               b->st))
                return -1;
        }
        rf_caps_t cs;
        memset(&cs, 0, sizeof cs);
        for (size_t m = 0; m < s->n_mons; m++) {
            smon_t *mo = &s->mons[m];
            for (int w = 0; w < 2; w++) {
                rf_cap_t *c = &cs.c[cs.n++];
                c->f = s->f[mo->comp[w]];
                c->buf = mo->ring[w];
                c->axis = mo->axis;
                c->index = mo->index;
                c->na = mo->na;
                c->nb = mo->nb;
                if (cs.n == RF_MAX_CAPS || (m + 1 == s->n_mons && w == 1)) {
                    if (rf_launch_plane_captures(s->prec, &s->geom, &cs, s->K, s->d_step, b->st)) return -1;
                    cs.n = 0;
                }
            }
        }
    }
    if (rf_launch_probe_record_packed(b->prec, b->d_probe[ph], (int)b->n, b->st)) return -1;
    b->phase ^= 1;
    return 0;
}

static int batch_enqueue_flush(fdtd_batch_t *b, int count) {
    if (count <= 0) return 0;
    if (rf_launch_probe_dft_packed(b->d_pdft, (int)b->n, b->max_freq, count, b->st)) return -1;
    for (size_t q = 0; q < b->n; q++) {
        fdtd_sim_t *s = b->sims[q];
        for (size_t m = 0; m < s->n_mons; m++) {
            smon_t *mo = &s->mons[m];
            const long long cells = (long long)mo->na * mo->nb;
            for (int w = 0; w < 2; w++)
                if (rf_launch_plane_dft(mo->ring[w], s->prec, cells, s->K, mo->nf, mo->d_w, s->dt, s->d_step, count,
                                        s->d_tw, mo->acc_re[w], mo->acc_im[w], NULL, NULL, 0, b->st))
                    return -1;
        }
    }
    return 0;
}

fdtd_batch_t *fdtd_batch_create(fdtd_sim_t *const *sims, size_t n_sims) {
    if (!sims || n_sims == 0 || n_sims > 65535) {
        rf_set_error("fdtd_batch_create: 1 to 65535 handles");
        return NULL;
    }
    for (size_t q = 0; q < n_sims; q++) {
        if (!sims[q]) {
            rf_set_error("fdtd_batch_create: NULL handle");
            return NULL;
        }
        for (size_t r = 0; r < q; r++)
            if (sims[r] == sims[q]) {
                rf_set_error("fdtd_batch_create: a handle appears twice");
                return NULL;
            }
        if (sims[q]->prec != sims[0]->prec ||
            (sims[q]->started ? sims[q]->K : dft_block_of(sims[q], 1)) !=
                (sims[0]->started ? sims[0]->K : dft_block_of(sims[0], 1))) {
            rf_set_error("fdtd_batch_create: the handles differ in precision or dft_block");
            return NULL;
        }
    }
    fdtd_batch_t *b = calloc(1, sizeof *b);
    if (!b) return NULL;
    b->sims = malloc(n_sims * sizeof *b->sims);
    b->sc = calloc(n_sims, sizeof *b->sc);
    b->st = rf_stream_create();
    if (!b->sims || !b->sc || !b->st) {
        fdtd_batch_destroy(b);
        return NULL;
    }
    memcpy(b->sims, sims, n_sims * sizeof *b->sims);
    const char *e = getenv("RFSYN_SHELL_OVERLAP");
    if (!(e && atoi(e) == 0)) {
        b->st2 = rf_stream_create();
        b->fork = rf_event_create_sync();
        b->join = rf_event_create_sync();
        if (!b->st2 || !b->fork || !b->join) {
            fdtd_batch_destroy(b);
            return NULL;
        }
    }
    b->n = n_sims;
    b->prec = sims[0]->prec;
    b->K = sims[0]->K;
    return b;
}

/* The scenes must be ready for packed steps: set up, on the fused step
 * (K is even there), not profiling, on a single stream.  -2 when a scene
 * cannot be packed, before any step is enqueued. */
static int batch_prepare(fdtd_batch_t *b, size_t n_steps) {
    for (size_t q = 0; q < b->n; q++) {
        fdtd_sim_t *s = b->sims[q];
        if (run_prepare(s, n_steps)) return -1;
        if (!s->fused || (s->K & 1) || s->profiling || s->K != b->sims[0]->K) {
            rf_set_error("fdtd_batch_run: scene %zu is not on the fused step, is profiling or has another "
                         "dft_block", q);
            return -2;
        }
    }
    b->K = b->sims[0]->K;
    for (size_t q = 0; q < b->n; q++) {
        fdtd_sim_t *s = b->sims[q];
        if (rf_sync(s->st)) return -1;
        if (s->pending && (enqueue_flush(s, s->pending) || rf_sync(s->st))) return -1;
        s->pending = 0;
        if (s->shell_overlap) {
            s->shell_overlap = 0;
            for (int p = 0; p < 2; p++) {
                rf_graph_destroy(s->graphs[p]);
                s->graphs[p] = NULL;
            }
        }
    }
    /* tables built from other buffers or wave tables, or a handle stepped
     * on its own since: build them again */
    int stale = !b->built;
    for (size_t q = 0; q < b->n && !stale; q++) {
        const fdtd_sim_t *s = b->sims[q];
        const bscene_t *c = &b->sc[q];
        stale = s->wave_gen != c->wave_gen || s->parity != (c->parity ^ b->phase) || s->steps != c->expect;
    }
    if (stale && batch_build(b)) return b->built ? -1 : -2;
    return 0;
}

static int batch_run(fdtd_batch_t *b, size_t n_steps) {
    const int pr = batch_prepare(b, n_steps);
    if (pr) return pr;
    size_t left = n_steps;
    int pending = 0;
    while (left) {
        if (pending == 0 && left >= (size_t)b->K) {
            rf_graph_t **g = &b->graphs[b->phase];
            if (!*g) {
                const int ph = b->phase;
                if (rf_graph_begin(b->st)) return -1;
                int rc = 0;
                for (int q = 0; q < b->K && !rc; q++) rc = batch_enqueue_step(b);
                if (!rc) rc = batch_enqueue_flush(b, b->K);
                const int rc2 = rf_graph_end(b->st, g);
                /* capture ran no kernel: K (even) swaps left every handle's
                 * state where it was */
                b->phase = ph;
                if (rc || rc2) {
                    rf_graph_destroy(*g);
                    *g = NULL;
                    rf_set_error("batch graph capture: %s", rf_error_string());
                    return -1;
                }
            }
            if (rf_graph_launch(*g, b->st)) return -1;
            for (size_t q = 0; q < b->n; q++) b->sims[q]->steps += (size_t)b->K;
            left -= (size_t)b->K;
            continue;
        }
        if (batch_enqueue_step(b)) return -1;
        for (size_t q = 0; q < b->n; q++) b->sims[q]->steps++;
        left--;
        if (++pending == b->K) {
            if (batch_enqueue_flush(b, b->K)) return -1;
            pending = 0;
        }
    }
    if (pending && batch_enqueue_flush(b, pending)) return -1;
    if (rf_sync(b->st)) {
        rf_set_error("fdtd_batch_run: %s", rf_error_string());
        return -1;
    }
    for (size_t q = 0; q < b->n; q++) b->sc[q].expect = b->sims[q]->steps;
    return 0;
}

int fdtd_batch_run(fdtd_batch_t *b, size_t n_steps) {
    if (!b) return -1;
    return batch_run(b, n_steps) ? -1 : 0;
}

size_t fdtd_batch_size(const fdtd_batch_t *b) { return b ? b->n : 0; }

/* The isfinite gates of the reference's validators (e.g.
 * validation_common.h:53), on the device: one scan of the state. */
int fdtd_sim_fields_finite(fdtd_sim_t *s) {
    if (!s) return -1;
    if (!s->d_nonfinite && dalloc(s, (void **)&s->d_nonfinite, sizeof(unsigned))) return -1;
    unsigned flag = 0;
    if (rf_memset(s->d_nonfinite, 0, sizeof(unsigned), s->st) ||
        rf_launch_nonfinite(s->prec, s->blk, 6 * (long long)s->cells, s->d_nonfinite, s->st) ||
        rf_d2h(&flag, s->d_nonfinite, sizeof(unsigned), s->st) || rf_sync(s->st)) {
        rf_set_error("fdtd_sim_fields_finite: %s", rf_error_string());
        return -1;
    }
    return flag ? 0 : 1;
}

int fdtd_batch_fields_finite(fdtd_batch_t *b, int *finite) {
    if (!b || !finite) return -1;
    if (rf_sync(b->st)) return -1;   /* the batch's steps are on its own stream */
    for (size_t q = 0; q < b->n; q++) {
        finite[q] = fdtd_sim_fields_finite(b->sims[q]);
        if (finite[q] < 0) return -1;
    }
    return 0;
}

int fdtd_batch_launches(const fdtd_batch_t *b, int *leapfrog, int *shell, int *ade) {
    if (!b || !b->built) return -1;
    if (leapfrog) *leapfrog = b->n_groups;
    if (shell) *shell = 2 * b->n_shells;
    if (ade) *ade = b->n_ades;
    return 0;
}

size_t fdtd_sim_steps_done(const fdtd_sim_t *s) { return s ? s->steps : 0; }

int fdtd_sim_dft_block(const fdtd_sim_t *s) {
    if (!s) return -1;
    return s->started ? s->K : dft_block_of(s, want_fused(s));
}

int fdtd_sim_leap_rows(const fdtd_sim_t *s) { return s ? s->plan_rows : -1; }

int fdtd_sim_streams(const fdtd_sim_t *s) {
    if (!s) return -1;
    return 1 + (s->st2 && s->shell_overlap && s->shell.n ? 1 : 0);
}

int fdtd_sim_fused(const fdtd_sim_t *s) { return s ? s->fused : -1; }

int fdtd_sim_fused_box(const fdtd_sim_t *s, int lo[3], int hi[3]) {
    if (!s || !lo || !hi || !s->started || !s->fused) return -1;
    for (int a = 0; a < 3; a++) {
        lo[a] = s->ib0[a];
        hi[a] = s->ib1[a];
    }
    return 0;
}

long fdtd_sim_probe_td(fdtd_sim_t *s, int probe, double *out, size_t n) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || (n && !out)) return -1;
    size_t have = s->steps < s->probes[probe].td_cap ? s->steps : s->probes[probe].td_cap;
    size_t c = n < have ? n : have;
    if (c && rf_d2h(out, s->d_td + (size_t)probe * s->td_stride, c * sizeof(double), s->st)) return -1;
    return (long)c;
}

/* Arguments of the output stage; -1 (with the error set) when they are
 * invalid or fewer than n samples are recorded. */
static int probe_iq_check(fdtd_sim_t *s, int probe, size_t n, const float *out_iq) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || (n && !out_iq)) {
        rf_set_error("bad probe or output");
        return -1;
    }
    size_t have = s->steps < s->probes[probe].td_cap ? s->steps : s->probes[probe].td_cap;
    if (n > have) {
        rf_set_error("probe %d holds %zu recorded samples, %zu asked", probe, have, n);
        return -1;
    }
    return 0;
}

/* The output stage with its device memory given: the mean power in
 * d_pow[0], the n I/Q samples in d_iq (2n floats). */
static int probe_iq_into(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                         uint64_t stream, float *out_iq, double *d_pow, float *d_iq) {
    const double *td = s->d_td + (size_t)probe * s->td_stride;
    int rc = 0;
    float std_f = 0.f;
    uint64_t state0 = 0, inc = (stream << 1u) | 1u;
    if (!rc && add_noise) {
        /* awgn.c:20-31: mean power in sample order (device), then the
         * standard deviation with the host's pow and sqrt */
        double power = 0.0;
        rc = rf_launch_iq_power(td, (long long)n, d_pow, s->st) || rf_d2h(&power, d_pow, sizeof power, s->st);
        if (!rc) {
            const double snr_linear = pow(10.0, (double)snr_db / 10.0);
            const double noise_power = (snr_linear > 0.0) ? power / snr_linear : 0.0;
            std_f = (float)sqrt(noise_power / 2.0);
            /* rng_seed (rng.c:3-9): the state before the first draw */
            state0 = inc;   /* the first step from state 0 */
            state0 += seed;
            state0 = state0 * 6364136223846793005ULL + inc;
        }
    }
    if (!rc) rc = rf_launch_iq_noise(td, (long long)n, add_noise != 0, std_f, state0, inc, d_iq, s->st) ||
                  rf_d2h(out_iq, d_iq, 2 * n * sizeof(float), s->st);
    return rc ? -1 : 0;
}

int fdtd_sim_probe_iq(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                      uint64_t stream, float *out_iq) {
    if (probe_iq_check(s, probe, n, out_iq)) return -1;
    if (n == 0) return 0;
    double *d_pow = NULL;
    float *d_iq = NULL;
    int rc = rf_malloc((void **)&d_pow, sizeof(double)) || rf_malloc((void **)&d_iq, 2 * n * sizeof(float));
    if (!rc) rc = probe_iq_into(s, probe, n, add_noise, snr_db, seed, stream, out_iq, d_pow, d_iq);
    rf_free(d_pow);
    rf_free(d_iq);
    return rc ? -1 : 0;
}

_Static_assert(FDTD_IQ_NOALLOC_MAX <= RF_IQ_SCRATCH_SAMPLES, "the scratch array holds the largest no-alloc call");

int fdtd_sim_probe_iq_noalloc(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                              uint64_t stream, float *out_iq) {
    if (probe_iq_check(s, probe, n, out_iq)) return -1;
    if (n > FDTD_IQ_NOALLOC_MAX) {
        rf_set_error("fdtd_sim_probe_iq_noalloc: %zu samples asked, at most %d", n, FDTD_IQ_NOALLOC_MAX);
        return -1;
    }
    if (n == 0) return 0;
    double *d_pow = NULL;
    float *d_iq = NULL;
    if (rf_iq_scratch(&d_pow, &d_iq)) {
        rf_set_error("fdtd_sim_probe_iq_noalloc: device scratch: %s", rf_error_string());
        return -1;
    }
    return probe_iq_into(s, probe, n, add_noise, snr_db, seed, stream, out_iq, d_pow, d_iq);
}

int fdtd_write_iq(const char *path, const float *iq, size_t n) {
    if (!path || (n && !iq)) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const size_t written = n ? fwrite(iq, 2 * sizeof(float), n, f) : 0;
    const int closed = fclose(f);
    return (written == n && closed == 0) ? 0 : -1;
}

int fdtd_write_transfer(const char *path, const double *re, const double *im, size_t n_probes, size_t n_freq) {
    if (!path || (n_freq && n_probes > SIZE_MAX / 16 / n_freq)) return -1;
    const size_t n = n_probes * n_freq;
    if (n && (!re || !im)) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    double buf[2 * 512];
    size_t done = 0;
    int ok = 1;
    while (ok && done < n) {
        const size_t m = n - done < 512 ? n - done : 512;
        for (size_t q = 0; q < m; q++) {
            buf[2 * q] = re[done + q];
            buf[2 * q + 1] = im[done + q];
        }
        ok = fwrite(buf, 2 * sizeof(double), m, f) == m;
        done += m;
    }
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

int fdtd_read_transfer(const char *path, double *re, double *im, size_t n_probes, size_t n_freq) {
    if (!path || (n_freq && n_probes > SIZE_MAX / 16 / n_freq)) return -1;
    const size_t n = n_probes * n_freq;
    if (n && (!re || !im)) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    double buf[2 * 512];
    size_t done = 0;
    int ok = 1;
    while (ok && done < n) {
        const size_t m = n - done < 512 ? n - done : 512;
        ok = fread(buf, 2 * sizeof(double), m, f) == m;
        for (size_t q = 0; ok && q < m; q++) {
            re[done + q] = buf[2 * q];
            im[done + q] = buf[2 * q + 1];
        }
        done += m;
    }
    /* the file holds exactly n values */
    if (ok) ok = fgetc(f) == EOF;
    fclose(f);
    return ok ? 0 : -1;
}

int fdtd_sim_probe_dft(fdtd_sim_t *s, int probe, double *re, double *im) {
    if (!s || probe < 0 || (size_t)probe >= s->n_probes || !re || !im) return -1;
    const sprobe_t *p = &s->probes[probe];
    if (!p->nf) return 0;
    if (!s->started) { memset(re, 0, p->nf * sizeof *re); memset(im, 0, p->nf * sizeof *im); return 0; }
    if (rf_d2h(re, s->d_acc_re + p->f_off, p->nf * sizeof(double), s->st) ||
        rf_d2h(im, s->d_acc_im + p->f_off, p->nf * sizeof(double), s->st))
        return -1;
    if (s->comp_sum) {
        double *c = malloc(p->nf * sizeof *c);
        if (!c) return -1;
        if (rf_d2h(c, s->d_cmp_re + p->f_off, p->nf * sizeof(double), s->st)) { free(c); return -1; }
        for (size_t q = 0; q < p->nf; q++) re[q] += c[q];
        if (rf_d2h(c, s->d_cmp_im + p->f_off, p->nf * sizeof(double), s->st)) { free(c); return -1; }
        for (size_t q = 0; q < p->nf; q++) im[q] += c[q];
        free(c);
    }
    return 0;
}

int fdtd_sim_plane_dft(fdtd_sim_t *s, int monitor, int which, double *re, double *im) {
    if (!s || monitor < 0 || (size_t)monitor >= s->n_mons || which < 0 || which > 1 || !re || !im) return -1;
    smon_t *m = &s->mons[monitor];
    size_t tot = (size_t)m->nf * m->na * m->nb;
    if (!s->started) { memset(re, 0, tot * sizeof *re); memset(im, 0, tot * sizeof *im); return 0; }
    if (rf_d2h(re, m->acc_re[which], tot * sizeof(double), s->st) ||
        rf_d2h(im, m->acc_im[which], tot * sizeof(double), s->st))
        return -1;
    return 0;
}

static int same_shape(const fdtd_sim_t *s, const fdtd_grid_t *g) {
    return s && g && s->nx == g->nx && s->ny == g->ny && s->nz == g->nz;
}

int fdtd_sim_download(fdtd_sim_t *s, fdtd_grid_t *g) {
    if (!same_shape(s, g)) return -1;
    double *dst[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
    size_t bytes = s->cells * sizeof(double);
    void *tmp = NULL;
    if (s->prec && rf_malloc(&tmp, bytes) != 0) return -1;
    for (int q = 0; q < 6; q++) {
        int rc;
        if (s->prec == 0) {
            rc = rf_d2h(dst[q], s->f[q], bytes, s->st);
        } else {
            rc = rf_launch_convert(tmp, 0, s->f[q], 1, (long long)s->cells, s->st);
            if (!rc) rc = rf_d2h(dst[q], tmp, bytes, s->st);
        }
        if (rc) { rf_free(tmp); return -1; }
    }
    rf_free(tmp);
    return 0;
}

int fdtd_sim_upload(fdtd_sim_t *s, const fdtd_grid_t *g) {
    if (!same_shape(s, g)) return -1;
    if (upload_fields(s, g)) return -1;
    /* the fused step leaves nodes it does not update (the padding of a
     * periodic axis) as they are in B: keep B equal to A there */
    if (s->fused && s->started) {
        for (int q = 0; q < 6; q++)
            if (rf_d2d(s->f2[q], s->f[q], s->cells * s->es, s->st)) return -1;
        if (rf_sync(s->st)) return -1;
    }
    return 0;
}

int fdtd_sim_energy(fdtd_sim_t *s, double *electric, double *magnetic) {
    if (!s || !electric || !magnetic) return -1;
    double vac_eps = s->dt / (s->dx * s->vac.vac_ca);
    const void *tag = s->has_mat ? s->mat.d.tag : NULL;
    int tb = s->has_mat ? s->mat.d.tag_bytes : 0;
    const double *eps = s->has_mat ? s->mat.eps_tab : NULL;
    double out[2];
    double *d_out = s->d_energy + 2 * (size_t)s->n_partials;
    if (rf_launch_energy(s->prec, &s->geom, s->f, tag, tb, eps, vac_eps, MU0, s->d_energy, s->n_partials,
                         d_out, s->st) ||
        rf_d2h(out, d_out, sizeof out, s->st))
        return -1;
    double vol = s->dx * s->dx * s->dx;
    *electric = out[0] * vol;
    *magnetic = out[1] * vol;
    return 0;
}

int fdtd_sim_energy_audit(fdtd_sim_t *s, double out[5], size_t *step) {
    if (!s || !out) return -1;
    if (!s->fused || s->steps == 0 || !s->d_aux) {
        rf_set_error("the energy audit needs the fused step and at least one completed step");
        return -1;
    }
    double vac_eps = s->dt / (s->dx * s->vac.vac_ca);
    const void *tag = s->has_mat ? s->mat.d.tag : NULL;
    int tb = s->has_mat ? s->mat.d.tag_bytes : 0;
    const double *eps = s->has_mat ? s->mat.eps_tab : NULL;
    double *d_out = s->d_aux + 5 * (size_t)s->n_partials;
    double h[5];
    /* state n = f2 (E^n, H^{n-1/2}), state n+1 = f (H^{n+1/2}) */
    if (rf_launch_audit(s->prec, &s->geom, s->f2, s->f, tag, tb, eps, vac_eps, MU0, s->d_aux, s->n_partials, d_out,
                        s->st) ||
        rf_d2h(h, d_out, sizeof h, s->st))
        return -1;
    const double vol = s->dx * s->dx * s->dx;
    for (int q = 0; q < 5; q++) out[q] = h[q] * vol;
    if (step) *step = s->steps - 1;
    return 0;
}

int fdtd_sim_flux(fdtd_sim_t *s, int plane, double *energy) {
    if (!s || !energy || plane < 0 || plane >= s->n_flux) return -1;
    if (!s->d_flux) {   /* nothing run yet */
        *energy = 0.0;
        return 0;
    }
    double v = 0.0;
    if (rf_d2h(&v, s->d_flux + plane, sizeof v, s->st)) return -1;
    *energy = v * s->dx * s->dx * s->dt;
    return 0;
}

size_t fdtd_sim_device_bytes(const fdtd_sim_t *s) { return s ? s->bytes : 0; }

/* ------------------------------------------------- fdtd_simulate_echo */

/* engine/fdtd/simulate.c:10-66 on the device-resident path: the handle
 * in `precision`, stepped n_steps; then either the probe series (out_td)
 * or the output stage's I/Q samples (out_iq, with `noise` if non-NULL). */
static int echo_run(const fdtd_sim_config_t *cfg, int precision, fdtd_material_setup_fn setup, void *context,
                    double *out_td, float *out_iq, const fdtd_echo_options_t *noise, double *sample_rate_out) {
    if (!cfg || (!out_td && !out_iq) || cfg->n_steps == 0) return -1;
    fdtd_grid_t *g = fdtd_grid_create(cfg->nx, cfg->ny, cfg->nz, cfg->dx, cfg->courant_safety);
    if (!g) return -1;
    int rc = -1;
    cpml_t *pml = NULL;
    materials_t *mat = NULL;
    fdtd_sim_t *s = NULL;
    if (cfg->src_i > g->nx || cfg->src_j > g->ny || cfg->src_k > g->nz ||
        cfg->probe_i > g->nx || cfg->probe_j > g->ny || cfg->probe_k > g->nz) goto out;
    pml = cpml_create(g, cpml_default_params(cfg->npml));
    if (!pml) goto out;
    mat = materials_create(g);
    if (!mat) goto out;
    if (setup && setup(g, mat, context) != 0) goto out;
    source_gaussian_deriv_params_t sp = {cfg->t0_over_dt * g->dt, cfg->tau_over_dt * g->dt};
    source_t src = source_make_point(cfg->src_i, cfg->src_j, cfg->src_k, cfg->src_component, SRC_SOFT, 1.0,
                                     source_waveform_gaussian_derivative, &sp);
    fdtd_sim_options_t opt = {precision, 0, 0};
    s = fdtd_sim_create(g, pml, mat, &opt);
    if (!s) goto out;
    /* source_apply ignores H components (source.c:45-49) */
    if (src.component <= FIELD_EZ && fdtd_sim_add_point_source(s, &src) != 0) goto out;
    if (fdtd_sim_add_probe(s, cfg->probe_component, cfg->probe_i, cfg->probe_j, cfg->probe_k, cfg->n_steps,
                           NULL, 0) != 0)
        goto out;
    if (fdtd_sim_run(s, cfg->n_steps) != 0) goto out;
    if (out_td && fdtd_sim_probe_td(s, 0, out_td, cfg->n_steps) != (long)cfg->n_steps) goto out;
    if (out_iq && fdtd_sim_probe_iq(s, 0, cfg->n_steps, noise && noise->add_noise, noise ? noise->snr_db : 0.f,
                                    noise ? noise->seed : 0, noise ? noise->stream : 0, out_iq) != 0)
        goto out;
    if (sample_rate_out) *sample_rate_out = 1.0 / g->dt;
    rc = 0;
out:
    fdtd_sim_destroy(s);
    materials_destroy(mat);
    cpml_destroy(pml);
    fdtd_grid_destroy(g);
    return rc;
}

int fdtd_simulate_echo_with_setup(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out,
                                  fdtd_material_setup_fn setup, void *context) {
    if (!out_td) return -1;
    return echo_run(cfg, FDTD_FP64, setup, context, out_td, NULL, NULL, sample_rate_out);
}

int fdtd_simulate_echo(const fdtd_sim_config_t *cfg, double *out_td, double *sample_rate_out) {
    return fdtd_simulate_echo_with_setup(cfg, out_td, sample_rate_out, NULL, NULL);
}

int fdtd_simulate_echo_iq(const fdtd_sim_config_t *cfg, const fdtd_echo_options_t *opt, float *out_iq,
                          double *sample_rate_out) {
    if (!out_iq || !opt || (opt->precision != FDTD_FP64 && opt->precision != FDTD_FP32)) return -1;
    return echo_run(cfg, opt->precision, NULL, NULL, NULL, out_iq, opt, sample_rate_out);
}
