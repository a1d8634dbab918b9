/*
 * Device-resident simulation handle (extension; no single reference
 * counterpart).
 *
 * The reference drives a run through the low-level calls in a host loop
 * (engine/fdtd/simulate.c:48-54; validate_probe_dft.c:66-82): H update,
 * source_apply at t = n dt, E update, then probes at t = (n+1) dt.  This
 * handle runs that same step order with every field, CPML memory, ADE
 * state, probe and DFT accumulator resident on the GPU, in FP64 (bitwise
 * the reference arithmetic) or FP32 fields with FP64 DFT accumulators.
 *
 * Lifetime: fdtd_sim_create snapshots the grid's current host fields, the
 * CPML profiles and the materials; later changes to those host objects are
 * not seen (use fdtd_sim_upload for fields).  All functions return 0 (or a
 * non-negative id) on success and -1 on invalid arguments or device errors.
 */
#ifndef RFSYN_FDTD_NATIVE_H
#define RFSYN_FDTD_NATIVE_H

#include <stddef.h>
#include <stdint.h>
#include "simulate.h"
#include "grid.h"
#include "cpml.h"
#include "materials.h"
#include "source.h"
#include "field_component.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdtd_sim fdtd_sim_t;

typedef enum {
    FDTD_FP64 = 0,   /* fields, CPML and ADE in double: the parity path */
    FDTD_FP32 = 1    /* fields, CPML and ADE in float; DFT and energy sums in double */
} fdtd_precision_t;

typedef struct {
    fdtd_precision_t precision;
    int dft_compensated;   /* point probes: 1 = Neumaier-compensated double
                              DFT sums, 0 = plain sums in the reference
                              order; plane monitors always use plain
                              double sums                               */
    size_t dft_block;      /* steps buffered between DFT flushes, at most
                              128 (0: 32, or 64 when the handle has plane
                              monitors); the results do not depend on it */
} fdtd_sim_options_t;

fdtd_sim_t *fdtd_sim_create(const fdtd_grid_t *g, const cpml_t *pml,
                            const materials_t *mat,
                            const fdtd_sim_options_t *opt);
void fdtd_sim_destroy(fdtd_sim_t *s);

/* Point source, applied as source_apply(src, g, n*dt) between the H and E
 * updates of step n.  The waveform is evaluated on the host. */
int fdtd_sim_add_point_source(fdtd_sim_t *s, const source_t *src);

/* Plane source on the E component `component` over the plane
 * {axis coordinate == index}, on every cell of that plane the E update
 * covers.  after_e = 0: applied like source_apply at t = n dt before the
 * E update; after_e = 1: added after the E update at t = (n+1) dt (the
 * injection of validation_common.h:16-18, 173). */
int fdtd_sim_add_plane_source(fdtd_sim_t *s, field_component_t component,
                              int axis, size_t index, source_mode_t mode,
                              int after_e, double amplitude,
                              source_waveform_fn waveform,
                              const void *params);

/* Point probe on any component, read after the E update of each step.
 * td_capacity: number of time-domain samples kept (0: none).
 * freqs_hz/n_freqs: running DFT at t = (n+1) dt, as probe_dft_record with
 * the caller of validate_probe_dft.c:77-80.  Returns the probe id. */
int fdtd_sim_add_probe(fdtd_sim_t *s, field_component_t component,
                       size_t i, size_t j, size_t k, size_t td_capacity,
                       const double *freqs_hz, size_t n_freqs);

/* Full-plane DFT monitor of the two E components tangential to the plane
 * {axis coordinate == index}, over [0,n_a) x [0,n_b) of the other axes in
 * increasing axis order, at t = (n+1) dt.  Returns the monitor id. */
int fdtd_sim_add_plane_monitor(fdtd_sim_t *s, int axis, size_t index,
                               const double *freqs_hz, size_t n_freqs);

/* Advance n_steps full steps. */
int fdtd_sim_run(fdtd_sim_t *s, size_t n_steps);

/* Batch of independent scenes (extension: the reference runs one scene
 * per process): advances every handle n_steps; returns when all are done.
 * Each handle ends in the state fdtd_sim_run would give it.  A handle may
 * appear once.  When every handle is on the fused step with the same
 * precision and dft_block, the scenes are stepped by packed launches (a
 * temporary fdtd_batch_t, below); otherwise one block of steps of each
 * handle is enqueued in turn on its own stream. */
int fdtd_sim_run_many(fdtd_sim_t *const *sims, size_t n_sims, size_t n_steps);

/* Packed batch (extension): many scenes of uneven sizes stepped together,
 * each kernel kind once per step for all scenes.  Per step, one fused
 * leapfrog launch covers the boxes of every scene whose tiling variant
 * matches (rows per tile, material-tag width, z CPML): the work items of
 * all scenes form one queue and each descriptor of a device table holds
 * one scene's buffers, CPML, materials and sources.  The CPML shells of
 * all scenes take one launch for H and one for E, all probe gathers one,
 * and per block of dft_block steps all point-probe DFTs one.  A block of
 * steps of the whole batch is one captured graph.  The split ADE passes
 * of all scenes take one launch per (tag width, pole count) variant.  A
 * scene's other launches (tensor edges, post-E plane sources, flux planes,
 * plane monitors) keep their place in its step.  Every handle
 * ends in the state fdtd_sim_run would give it, bit for bit.
 * fdtd_batch_create: 1..65535 distinct handles with equal precision and
 * dft_block (NULL otherwise); the handles must outlive the batch.
 * fdtd_batch_run: -1 if a handle is not on the fused step (see
 * fdtd_sim_select_path) or is profiling.  A handle stepped on its own
 * between runs is picked up where it is. */
typedef struct fdtd_batch fdtd_batch_t;
fdtd_batch_t *fdtd_batch_create(fdtd_sim_t *const *sims, size_t n_sims);
int fdtd_batch_run(fdtd_batch_t *b, size_t n_steps);
void fdtd_batch_destroy(fdtd_batch_t *b);
size_t fdtd_batch_size(const fdtd_batch_t *b);
/* Per-scene failure flags: finite[q] = fdtd_sim_fields_finite of scene q
 * (1 finite, 0 diverged) after the batch's queued steps.  0 on success. */
int fdtd_batch_fields_finite(fdtd_batch_t *b, int *finite);
/* Diagnostics after a run: packed leapfrog launches per step (one per
 * tiling variant), shell launches and split-ADE-pass launches per step
 * (any pointer may be NULL).  0, or -1 before the first run. */
int fdtd_batch_launches(const fdtd_batch_t *b, int *leapfrog, int *shell, int *ade);

/* Tiling autotuner (extension): before the first step, times the fused
 * step's kernel for each candidate tiling (rows per tile, planes per work
 * item, a separate launch for the tiles without z CPML lanes, independent
 * wavefronts or cooperative blocks of neighbouring tiles), `launches`
 * launches each, and keeps the fastest for this handle.  The tiling never
 * changes the results.  chosen (may be NULL) receives rows, planes per
 * item (0: automatic), the z split and the block mode (0: independent
 * wavefronts; else cooperative blocks of (mode & 15) row tiles of one z
 * tile, whose wavefronts stay within mode >> 4 planes of each other, 0 =
 * lockstep).  0 on success (also on the two-kernel path, where there is
 * nothing to tune). */
int fdtd_sim_autotune(fdtd_sim_t *s, int launches, int chosen[4]);

/* Path choice (extension) for a handle run on its own: before the first
 * step (after fdtd_sim_autotune, if used), times the fused step against
 * the two-kernel step (separate H and E sweeps), `launches` steps each,
 * and keeps the two-kernel step when it is faster by 5% or more (small
 * grids with CPML on most faces), unless the handle has flux planes or
 * RFSYN_FUSED is set.  Both paths give the same bits; the energy audit
 * needs the fused step.  Handles advanced together with
 * fdtd_sim_run_many are best left on the fused step, whose lower traffic
 * wins when many scenes share the device.  *two_kernel (may be NULL)
 * receives 1 when the two-kernel step was kept.  0 on success. */
int fdtd_sim_select_path(fdtd_sim_t *s, int launches, int *two_kernel);
size_t fdtd_sim_steps_done(const fdtd_sim_t *s);
/* Steps per DFT block the handle runs with (see dft_block); -1 for NULL. */
int fdtd_sim_dft_block(const fdtd_sim_t *s);
/* GPU failures recorded by the reference entry points that return nothing
 * (fdtd_update_h, fdtd_update_e, their _cpml and _materials variants).  By
 * default such a failure prints a message and aborts the process; with
 * RFSYN_GPU_ERRORS=record (read at each failure) the call instead records
 * the failure in fdtd_last_error(), counts it here and returns with the
 * grid's host fields not updated by that call. */
int fdtd_gpu_errors(void);
/* Whether every value of the handle's current state (E and H on every
 * node, the padding included) is finite: 1 yes, 0 no (a NaN or an
 * infinity: the run has diverged), -1 on error.  One device scan of the
 * state, about its size in HBM reads; waits for the handle's stream.  The
 * reference's validators make the same check on the host (isfinite gates,
 * e.g. validation_common.h:53). */
int fdtd_sim_fields_finite(fdtd_sim_t *s);
/* Rows per work item of the fused step's last planned tiling: 1 is row
 * mode (FP32: an item spans every z column), 0 before the fused step was
 * planned; -1 for NULL.  RFSYN_LEAP_ROWS, read at the handle's first
 * run, requests a value (fdtd_sim_autotune replaces it). */
int fdtd_sim_leap_rows(const fdtd_sim_t *s);
/* Diagnostics: streams the handle's next step uses (1, or 2 when the
 * CPML shell's H update runs on a second stream beside the fused kernel;
 * handles advanced together by fdtd_sim_run_many use one). */
int fdtd_sim_streams(const fdtd_sim_t *s);

/* Copies min(n, recorded) time-domain samples; returns the count or -1. */
long fdtd_sim_probe_td(fdtd_sim_t *s, int probe, double *out, size_t n);

/* Dataset output stage on the device (extension): the first n recorded
 * samples of time-domain probe `probe` as the FDTD stage's complex float
 * signal (cli/fdtd_stage.c:20-24: I = (float)v, Q = 0) and, when
 * add_noise, the AWGN stage of the reference on it (engine/noise/
 * awgn.c:16-44): per-axis standard deviation sqrt(P / 10^(snr_db/10) / 2)
 * from the signal's mean power P, and per sample two uniforms of
 * PCG32(seed, stream) (engine/core/rng.c; the CLI passes the example index
 * as the stream) into a float Box-Muller pair.  out_iq (host) receives 2n
 * floats, I and Q interleaved: the layout of the .iq files.  The power and
 * the uniforms are the reference's bits; the noise is within a few float
 * ulps of it (device logf, cosf, sinf).  -1 if fewer than n samples are
 * recorded. */
int fdtd_sim_probe_iq(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                      uint64_t stream, float *out_iq);
/* fdtd_sim_probe_iq without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: tight
 * serving loops, callers that forbid hipMalloc).  The same kernels, the
 * same results; the power and the I/Q samples are staged in a fixed
 * device array of the library (about 512 KiB per GPU, held until the
 * process ends) instead of buffers allocated per call, so n is at most
 * FDTD_IQ_NOALLOC_MAX (-1 beyond).  Every call on one GPU uses the same
 * array: calls that overlap on a GPU (from several threads) race. */
#define FDTD_IQ_NOALLOC_MAX 65536
int fdtd_sim_probe_iq_noalloc(fdtd_sim_t *s, int probe, size_t n, int add_noise, float snr_db, uint64_t seed,
                              uint64_t stream, float *out_iq);
/* The CLI's FDTD stage and, optionally, its AWGN stage in one call
 * (extension; cli/fdtd_stage.c:10-29 and settings.c:52-80): the echo of
 * fdtd_simulate_echo run on the device in `precision` (FP64 gives the
 * reference's samples), returned as the stage's I/Q signal (2 n_steps
 * floats) with, when add_noise, the AWGN stage on the device
 * (fdtd_sim_probe_iq; the CLI passes the example index as stream).
 * sample_rate_out (may be NULL) receives 1/dt.  0 on success, -1 on the
 * errors of fdtd_simulate_echo or a bad option. */
typedef struct {
    int precision;       /* FDTD_FP64 or FDTD_FP32 */
    int add_noise;
    float snr_db;
    uint64_t seed, stream;
} fdtd_echo_options_t;
int fdtd_simulate_echo_iq(const fdtd_sim_config_t *cfg, const fdtd_echo_options_t *opt, float *out_iq,
                          double *sample_rate_out);

/* Writes n I/Q samples (2n floats) to path as raw float32, I then Q, the
 * .iq file of the CLI (cli/main.c:70-85; little-endian on x86-64).  0 on
 * success, -1 on any failure. */
int fdtd_write_iq(const char *path, const float *iq, size_t n);

/* Dataset transfer-function file, the companion of the .iq files: the
 * n_probes x n_freq complex values (re[p n_freq + f], im[...]) as raw
 * float64 pairs, real then imaginary, probe-major, no header (the scene
 * JSON beside it records the shape and the frequencies; little-endian on
 * x86-64).  The reference has no such file (its CLI writes time-domain
 * .iq only); the learned surrogate's labels are these values.  Read
 * checks that the file holds exactly n_probes x n_freq values.  0 on
 * success, -1 on any failure. */
int fdtd_write_transfer(const char *path, const double *re, const double *im, size_t n_probes, size_t n_freq);
int fdtd_read_transfer(const char *path, double *re, double *im, size_t n_probes, size_t n_freq);
/* Accumulated DFT (acc_re, acc_im exactly as probe_dft_t holds them). */
int fdtd_sim_probe_dft(fdtd_sim_t *s, int probe, double *re, double *im);
/* Plane DFT of tangential component `which` (0: lower axis, 1: upper);
 * layout [f][a][b]. */
int fdtd_sim_plane_dft(fdtd_sim_t *s, int monitor, int which,
                       double *re, double *im);

/* Device fields <-> the grid's host arrays (converted to/from double). */
int fdtd_sim_download(fdtd_sim_t *s, fdtd_grid_t *g);
int fdtd_sim_upload(fdtd_sim_t *s, const fdtd_grid_t *g);

/* Electric and magnetic energy, 0.5 sum eps E^2 dV and 0.5 sum mu0 H^2 dV
 * over the cells each component updates, in double with a fixed
 * reduction order (eps from the scalar coefficient of each edge). */
int fdtd_sim_energy(fdtd_sim_t *s, double *electric, double *magnetic);

/* Energy audit of step n = steps done - 1 (audit_periodic_energy.c:43-71,
 * extension of the reference's host-side audit): from E^n, H^{n-1/2} and
 * H^{n+1/2}, out[0] = 0.5 sum eps (E^n)^2 dV, out[1] = 0.5 sum mu0
 * (H^{n-1/2})^2 dV, out[2] = 0.5 sum mu0 (H^{n+1/2})^2 dV, out[3] = 0.5 sum
 * mu0 ((H^{n-1/2} + H^{n+1/2})/2)^2 dV, out[4] = 0.5 sum mu0 H^{n-1/2}
 * H^{n+1/2} dV, over the cells fdtd_sim_energy sums.  out[0] + out[4] is
 * the modified energy that a closed lossless grid conserves to rounding.
 * Needs the fused step and at least one completed step; *step (may be
 * NULL) receives n. */
int fdtd_sim_energy_audit(fdtd_sim_t *s, double out[5], size_t *step);

/* Poynting flux plane (extension; the reference integrates E x H at single
 * probe points in validation_common.h): the plane of E nodes at `index`
 * (< n) along `axis`.  Each step n -> n+1 adds dt dx^2 times the sum over
 * the plane's cells of (0.5 (E^n + E^{n+1}))_t x (H^{n+1/2})_t along +axis,
 * with the tangential E at the nodes of the plane and the tangential H at
 * the same array index (half a cell above it): with this pairing the
 * change of the modified energy between two planes equals the difference
 * of their fluxes (the discrete Poynting theorem).  Add before the first
 * run; needs the fused step.  Returns the plane id or -1.
 * fdtd_sim_flux returns the time integral so far in joules. */
int fdtd_sim_add_flux_plane(fdtd_sim_t *s, int axis, size_t index);
int fdtd_sim_flux(fdtd_sim_t *s, int plane, double *energy);

/* Profiling mode: on != 0 launches every kernel directly (no graph
 * replay) and times the step's kernels with HIP events on the handle's
 * stream in three slots: the H kernel (fused step: the leapfrog
 * launches), the E kernel (fused step: the split ADE pass) and, on the
 * fused step, the shell kernels.  fdtd_sim_kernel_times returns the
 * accumulated milliseconds of slot 0 and of slots 1 + 2 and the number of
 * timed steps since the previous call, then resets them;
 * fdtd_sim_kernel_times3 returns the three slots. */
int fdtd_sim_set_profiling(fdtd_sim_t *s, int on);
int fdtd_sim_kernel_times(fdtd_sim_t *s, double *h_ms, double *e_ms, long *steps);
int fdtd_sim_kernel_times3(fdtd_sim_t *s, double ms[3], long *steps);

/* 1 when the run uses the fused leapfrog step (H, sources and E in one
 * pass over double-buffered fields on the box inside the x/y CPML, the
 * split ADE pass after it and node-wise shell kernels around the box),
 * 0 for separate H and E kernels; decided at the first run. */
int fdtd_sim_fused(const fdtd_sim_t *s);
/* The node box [lo, hi) per axis that the fused kernel owns (the CPML
 * plain region with a one-node margin); the nodes around it are updated
 * by the shell kernels.  Valid after the first run; 0 on success. */
int fdtd_sim_fused_box(const fdtd_sim_t *s, int lo[3], int hi[3]);

/* Bytes of device memory held by the handle. */
size_t fdtd_sim_device_bytes(const fdtd_sim_t *s);

/* Last error message of this thread ("" when none). */
const char *fdtd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
