/*
 * Private declarations shared by the C host files of librfsyn_fdtd.
 */
#ifndef RFSYN_HOST_INTERNAL_H
#define RFSYN_HOST_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "grid.h"
#include "cpml.h"
#include "materials.h"
#include "rf_device.h"

/* ------------------------------------------------------------ errors */

void rf_set_error(const char *fmt, ...);
/* For the void update calls of the reference API: a device failure cannot
 * be returned, so it is reported on stderr and the process aborts. */
/* A GPU failure inside a reference entry point that returns nothing
 * (fdtd_update_*): aborts with a message, or with RFSYN_GPU_ERRORS=record
 * records it (fdtd_last_error, fdtd_gpu_errors) and returns; RF_FAIL then
 * returns from the caller without updating. */
void rf_fatal(const char *what);
#define RF_FAIL(what) do { rf_fatal(what); return; } while (0)

/* ------------------------------------------------------- hash map */

/* Open-addressing map size_t key -> uint32 value (linear probing,
 * backward-shift deletion). */
typedef struct {
    size_t *keys;
    uint32_t *vals;
    size_t cap, count;
} rf_map_t;

int rf_map_put(rf_map_t *m, size_t key, uint32_t val);
int rf_map_get(const rf_map_t *m, size_t key, uint32_t *val);   /* 1 found */
int rf_map_del(rf_map_t *m, size_t key);
void rf_map_free(rf_map_t *m);

/* ------------------------------------------------------ device grid */

struct fdtd_grid_device {
    void *f[6];           /* FP64 device fields, reference layout */
    rf_stream_t *st;
};

rf_geom_t rf_geom_of(const fdtd_grid_t *g);
size_t rf_cells(const fdtd_grid_t *g);        /* (nx+1)(ny+1)(nz+1) */
/* Lazily allocates the FP64 device mirror; NULL on failure. */
struct fdtd_grid_device *rf_grid_dev(fdtd_grid_t *g);
/* Host <-> device copies of the fields named by mask (bit q = field q in
 * ex ey ez hx hy hz order). */
int rf_grid_upload(fdtd_grid_t *g, unsigned mask);
int rf_grid_download(fdtd_grid_t *g, unsigned mask);

/* -------------------------------------------------------- device cpml */

/* A device copy of a cpml_t's profiles plus zeroed psi slabs, in the
 * requested precision. */
typedef struct {
    rf_cpml_t d;
    void *owned[64];
    int n_owned;
    size_t bytes;
    void *psi_sec[2];           /* the E and H psi sections (all components) */
    size_t psi_sec_bytes;       /* bytes of each */
} rf_cpml_dev_t;

int rf_cpml_build(const cpml_t *p, int prec, rf_cpml_dev_t *out);
void rf_cpml_release(rf_cpml_dev_t *c);
/* Compat mirror kept inside the cpml_t (FP64). */
rf_cpml_dev_t *rf_cpml_dev(cpml_t *p);
int rf_cpml_matches(const cpml_t *p, const fdtd_grid_t *g);

/* --------------------------------------------------- device materials */

typedef struct {
    rf_mat_t d;
    void *tensor_edges;       /* rf_tedge64_t / rf_tedge32_t */
    long long n_tensor;
    double *eps_tab;          /* device, 3 per entry, FP64 */
    size_t n_entries;
    size_t ade_bytes;         /* bytes of each of d.J and d.P (0: none) */
    void *owned[16];
    int n_owned;
    size_t bytes;
} rf_mat_dev_t;

int rf_materials_build(const materials_t *m, const fdtd_grid_t *g, int prec, rf_mat_dev_t *out);
void rf_materials_release(rf_mat_dev_t *d);
/* Compat mirror kept inside materials_t (FP64), rebuilt when the host
 * description changed since the last build; ADE state persists. */
rf_mat_dev_t *rf_materials_dev(materials_t *m, const fdtd_grid_t *g);
int rf_materials_matches(const materials_t *m, const fdtd_grid_t *g);

#endif
