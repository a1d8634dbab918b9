/*
 * Learned surrogate (include/fdtd_surrogate.h): host orchestration of the
 * forward pass, the backward pass and the Adam step over the kernels of
 * sur_kernels.hip.  Every product runs on the matrix cores with BF16
 * operands; the activations' pre-GELU values, the gradients and the
 * master weights stay FP32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fdtd_surrogate.h"
#include "host_internal.h"

#define NL 4   /* convolution layers */

typedef struct {
    float *w, *g, *m, *v;   /* FP32 master, gradient, Adam moments */
    uint16_t *w16;          /* BF16 copy for the products (NULL for biases) */
    size_t n;
} param_t;

struct fdtd_sur {
    fdtd_sur_config_t c;
    rf_stream_t *st;
    /* parameters: conv W, b per layer; W1f W1c b1 W2 b2 W3 b3 */
    param_t p[2 * NL + 7];
    int n_par;
    size_t n_params;
    long adam_t;
    int S[NL + 1], C[NL + 1];   /* voxels per axis and channels at each stage */
    long long F;                /* features per scene */
    /* activations for max_batch scenes */
    float *maps, *coords32, *labels;
    uint16_t *vox, *coords16;
    uint16_t *col[NL];
    float *z[NL];
    uint16_t *a16[NL];
    float *U, *V, *h1z, *h2z, *out;
    uint16_t *h1, *h2;
    /* backward */
    float *dout, *dh, *dz, *dU, *dfeat, *dcol, *dact, *dzc;
    uint16_t *dout16, *dz16, *dU16, *dzc16;
    uint16_t *ones;
    long long n_ones;
    float *part;
    long long part_cap;
    double *d_loss;
    int n_loss;
    size_t bytes;
};

static int dev(fdtd_sur_t *m, void **p, size_t bytes) {
    if (rf_malloc(p, bytes)) return -1;
    m->bytes += bytes;
    return 0;
}

/* PCG32 (engine/core/rng.c:3-23): uniform in [0, 1). */
typedef struct { uint64_t state, inc; } pcg_t;
static uint32_t pcg_next(pcg_t *r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27), rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
static void pcg_seed(pcg_t *r, uint64_t seed, uint64_t stream) {
    r->state = 0;
    r->inc = (stream << 1) | 1u;
    pcg_next(r);
    r->state += seed;
    pcg_next(r);
}

static int add_param(fdtd_sur_t *m, size_t n, int matrix) {
    param_t *q = &m->p[m->n_par++];
    q->n = n;
    if (dev(m, (void **)&q->w, n * 4) || dev(m, (void **)&q->g, n * 4) || dev(m, (void **)&q->m, n * 4) ||
        dev(m, (void **)&q->v, n * 4))
        return -1;
    if (matrix && dev(m, (void **)&q->w16, n * 2)) return -1;
    m->n_params += n;
    return 0;
}

/* Copies host weights into the masters and refreshes the BF16 copies. */
static int upload_params(fdtd_sur_t *m, const float *host) {
    size_t off = 0;
    for (int k = 0; k < m->n_par; k++) {
        param_t *q = &m->p[k];
        if (rf_h2d(q->w, host + off, q->n * 4, m->st)) return -1;
        if (rf_memset(q->m, 0, q->n * 4, m->st) || rf_memset(q->v, 0, q->n * 4, m->st)) return -1;
        if (q->w16 && rf_sur_to_bf16(q->w, q->w16, (long long)q->n, m->st)) return -1;
        off += q->n;
    }
    m->adam_t = 0;
    return rf_sync(m->st);
}

static int valid(const fdtd_sur_config_t *c) {
    if (!c || c->grid < 16 || c->grid % 16 || c->in_channels < 1 || c->coords < 1 || c->width < 1 ||
        c->outputs < 1 || c->queries < 1 || c->max_batch < 1)
        return 0;
    for (int l = 0; l < NL; l++)
        if (c->channels[l] < 1) return 0;
    return 1;
}

void fdtd_sur_destroy(fdtd_sur_t *m) {
    if (!m) return;
    for (int k = 0; k < m->n_par; k++) {
        rf_free(m->p[k].w);
        rf_free(m->p[k].g);
        rf_free(m->p[k].m);
        rf_free(m->p[k].v);
        rf_free(m->p[k].w16);
    }
    void *bufs[] = {m->maps, m->coords32, m->labels, m->vox, m->coords16, m->U, m->V, m->h1z, m->h2z, m->out,
                    m->h1, m->h2, m->dout, m->dh, m->dz, m->dU, m->dfeat, m->dcol, m->dact, m->dzc, m->dout16,
                    m->dz16, m->dU16, m->dzc16, m->ones, m->part, m->d_loss};
    for (size_t q = 0; q < sizeof bufs / sizeof bufs[0]; q++) rf_free(bufs[q]);
    for (int l = 0; l < NL; l++) {
        rf_free(m->col[l]);
        rf_free(m->z[l]);
        rf_free(m->a16[l]);
    }
    rf_stream_destroy(m->st);
    free(m);
}

fdtd_sur_t *fdtd_sur_create(const fdtd_sur_config_t *cfg, uint64_t seed) {
    if (!valid(cfg)) {
        rf_set_error("invalid surrogate configuration");
        return NULL;
    }
    fdtd_sur_t *m = calloc(1, sizeof *m);
    if (!m) return NULL;
    m->c = *cfg;
    if (!(m->st = rf_stream_create())) goto fail;
    const long long B = cfg->max_batch, P = cfg->queries, W = cfg->width;
    m->S[0] = cfg->grid;
    m->C[0] = cfg->in_channels;
    for (int l = 0; l < NL; l++) {
        m->S[l + 1] = m->S[l] / 2;
        m->C[l + 1] = cfg->channels[l];
    }
    const long long s4 = m->S[NL];
    m->F = s4 * s4 * s4 * m->C[NL];
    for (int l = 0; l < NL; l++)
        if (add_param(m, (size_t)27 * m->C[l] * m->C[l + 1], 1) || add_param(m, (size_t)m->C[l + 1], 0)) goto fail;
    if (add_param(m, (size_t)m->F * W, 1) || add_param(m, (size_t)cfg->coords * W, 1) || add_param(m, (size_t)W, 0) ||
        add_param(m, (size_t)W * W, 1) || add_param(m, (size_t)W, 0) || add_param(m, (size_t)W * cfg->outputs, 1) ||
        add_param(m, (size_t)cfg->outputs, 0))
        goto fail;
    /* activations */
    long long S0 = m->S[0];
    if (dev(m, (void **)&m->maps, (size_t)(B * S0 * S0 * m->C[0]) * 4) ||
        dev(m, (void **)&m->vox, (size_t)(B * S0 * S0 * S0 * m->C[0]) * 2) ||
        dev(m, (void **)&m->coords32, (size_t)(B * P * cfg->coords) * 4) ||
        dev(m, (void **)&m->coords16, (size_t)(B * P * cfg->coords) * 2) ||
        dev(m, (void **)&m->labels, (size_t)(B * P * cfg->outputs) * 4))
        goto fail;
    long long max_col = 0, max_act = 0;
    for (int l = 0; l < NL; l++) {
        const long long so = m->S[l + 1], rows = B * so * so * so;
        const long long ncol = rows * 27 * m->C[l], nact = rows * m->C[l + 1];
        if (dev(m, (void **)&m->col[l], (size_t)ncol * 2) || dev(m, (void **)&m->z[l], (size_t)nact * 4) ||
            dev(m, (void **)&m->a16[l], (size_t)nact * 2))
            goto fail;
        if (l > 0 && ncol > max_col) max_col = ncol;
        if (nact > max_act) max_act = nact;
        const long long nin = B * (long long)m->S[l] * m->S[l] * m->S[l] * m->C[l];
        if (l > 0 && nin > max_act) max_act = nin;
    }
    const long long q = B * P;
    if (dev(m, (void **)&m->U, (size_t)(B * W) * 4) || dev(m, (void **)&m->V, (size_t)(q * W) * 4) ||
        dev(m, (void **)&m->h1z, (size_t)(q * W) * 4) || dev(m, (void **)&m->h2z, (size_t)(q * W) * 4) ||
        dev(m, (void **)&m->h1, (size_t)(q * W) * 2) || dev(m, (void **)&m->h2, (size_t)(q * W) * 2) ||
        dev(m, (void **)&m->out, (size_t)(q * cfg->outputs) * 4) ||
        dev(m, (void **)&m->dout, (size_t)(q * cfg->outputs) * 4) ||
        dev(m, (void **)&m->dout16, (size_t)(q * cfg->outputs) * 2) ||
        dev(m, (void **)&m->dh, (size_t)(q * W) * 4) || dev(m, (void **)&m->dz, (size_t)(q * W) * 4) ||
        dev(m, (void **)&m->dz16, (size_t)(q * W) * 2) || dev(m, (void **)&m->dU, (size_t)(B * W) * 4) ||
        dev(m, (void **)&m->dU16, (size_t)(B * W) * 2) || dev(m, (void **)&m->dfeat, (size_t)(B * m->F) * 4) ||
        dev(m, (void **)&m->dcol, (size_t)(max_col > 0 ? max_col : 1) * 4) ||
        dev(m, (void **)&m->dact, (size_t)max_act * 4) || dev(m, (void **)&m->dzc, (size_t)max_act * 4) ||
        dev(m, (void **)&m->dzc16, (size_t)max_act * 2))
        goto fail;
    /* a vector of BF16 ones for the bias gradients (column sums) */
    m->n_ones = max_act > q ? max_act : q;
    {
        float *ones32 = NULL;
        if (dev(m, (void **)&m->ones, (size_t)m->n_ones * 2) || rf_malloc((void **)&ones32, (size_t)m->n_ones * 4))
            goto fail;
        float *h = malloc((size_t)m->n_ones * 4);
        int rc = !h;
        if (!rc) {
            for (long long t = 0; t < m->n_ones; t++) h[t] = 1.f;
            rc = rf_h2d(ones32, h, (size_t)m->n_ones * 4, m->st) || rf_sur_to_bf16(ones32, m->ones, m->n_ones, m->st) ||
                 rf_sync(m->st);
        }
        free(h);
        rf_free(ones32);
        if (rc) goto fail;
    }
    m->part_cap = 32LL << 20;   /* split-K partials: 128 MiB */
    m->n_loss = 256;
    if (dev(m, (void **)&m->part, (size_t)m->part_cap * 4) ||
        dev(m, (void **)&m->d_loss, (size_t)m->n_loss * sizeof(double)))
        goto fail;
    /* He-uniform weights from PCG32, zero biases */
    {
        float *w = calloc(m->n_params, sizeof(float));
        if (!w) goto fail;
        pcg_t r;
        pcg_seed(&r, seed, 0);
        size_t off = 0;
        for (int k = 0; k < m->n_par; k++) {
            const param_t *pp = &m->p[k];
            if (pp->w16) {
                /* fan-in: rows of the [fan_in][fan_out] matrix */
                size_t fan_in;
                if (k < 2 * NL) fan_in = (size_t)27 * m->C[k / 2];
                else if (k == 2 * NL) fan_in = (size_t)m->F + cfg->coords;   /* W1f and W1c feed one layer */
                else if (k == 2 * NL + 1) fan_in = (size_t)m->F + cfg->coords;
                else fan_in = (size_t)W;
                const double lim = sqrt(6.0 / (double)fan_in);
                for (size_t t = 0; t < pp->n; t++)
                    w[off + t] = (float)((2.0 * ((pcg_next(&r) >> 8) / 16777216.0) - 1.0) * lim);
            }
            off += pp->n;
        }
        int rc = upload_params(m, w);
        free(w);
        if (rc) goto fail;
    }
    return m;
fail:
    fdtd_sur_destroy(m);
    return NULL;
}

size_t fdtd_sur_n_params(const fdtd_sur_t *m) { return m ? m->n_params : 0; }

int fdtd_sur_get_params(fdtd_sur_t *m, float *out) {
    if (!m || !out) return -1;
    size_t off = 0;
    for (int k = 0; k < m->n_par; k++) {
        if (rf_d2h(out + off, m->p[k].w, m->p[k].n * 4, m->st)) return -1;
        off += m->p[k].n;
    }
    return 0;
}

int fdtd_sur_set_params(fdtd_sur_t *m, const float *in) {
    if (!m || !in) return -1;
    return upload_params(m, in);
}

/* C[M][N] = A B with split-K chosen so that about 512 blocks run and the
 * partials fit the workspace (tiles as rf_sur_gemm shapes them: 256 x 16,
 * 128 x 32 or 128 x 64 by the width N). */
static int gemm(fdtd_sur_t *m, int M, int N, int K, const uint16_t *A, long long sam, long long sak,
                const uint16_t *B, long long sbk, long long sbn, float *C, const float *bias) {
    const int bm = N <= 16 ? 256 : 128, bn = N <= 16 ? 16 : (N <= 32 ? 32 : 64);
    const long long tiles = (long long)((M + bm - 1) / bm) * ((N + bn - 1) / bn);
    long long splits = 1;
    if (tiles < 512) splits = (512 + tiles - 1) / tiles;
    if (splits > K / 256) splits = K / 256;
    if (splits > 1024) splits = 1024;
    while (splits > 1 && splits * M * N > m->part_cap) splits /= 2;
    if (splits < 1) splits = 1;
    return rf_sur_gemm(M, N, K, A, sam, sak, B, sbk, sbn, C, N, bias, (int)splits, m->part, m->st);
}

static int upload_batch(fdtd_sur_t *m, const float *maps, const float *coords, const float *labels, int n) {
    const fdtd_sur_config_t *c = &m->c;
    const size_t S0 = (size_t)m->S[0], P = (size_t)c->queries;
    if (rf_h2d(m->maps, maps, (size_t)n * S0 * S0 * m->C[0] * 4, m->st) ||
        rf_h2d(m->coords32, coords, (size_t)n * P * c->coords * 4, m->st) ||
        (labels && rf_h2d(m->labels, labels, (size_t)n * P * c->outputs * 4, m->st)))
        return -1;
    return rf_sur_expand_y(m->maps, m->vox, n, (int)S0, m->C[0], m->st) ||
           rf_sur_to_bf16(m->coords32, m->coords16, (long long)n * P * c->coords, m->st);
}

static int forward(fdtd_sur_t *m, int n) {
    const fdtd_sur_config_t *c = &m->c;
    const uint16_t *in = m->vox;
    for (int l = 0; l < NL; l++) {
        const long long so = m->S[l + 1], rows = (long long)n * so * so * so;
        const int K = 27 * m->C[l], N = m->C[l + 1];
        if (rf_sur_im2col(in, m->col[l], n, m->S[l], m->C[l], m->st) ||
            gemm(m, (int)rows, N, K, m->col[l], K, 1, m->p[2 * l].w16, N, 1, m->z[l], m->p[2 * l + 1].w) ||
            rf_sur_gelu(m->z[l], NULL, m->a16[l], rows * N, m->st))
            return -1;
        in = m->a16[l];
    }
    const int W = c->width, P = c->queries, O = c->outputs;
    const long long q = (long long)n * P;
    const param_t *W1f = &m->p[2 * NL], *W1c = &m->p[2 * NL + 1], *b1 = &m->p[2 * NL + 2], *W2 = &m->p[2 * NL + 3],
                  *b2 = &m->p[2 * NL + 4], *W3 = &m->p[2 * NL + 5], *b3 = &m->p[2 * NL + 6];
    /* features: the last activation, F per scene */
    if (gemm(m, n, W, (int)m->F, m->a16[NL - 1], m->F, 1, W1f->w16, W, 1, m->U, NULL) ||
        gemm(m, (int)q, W, c->coords, m->coords16, c->coords, 1, W1c->w16, W, 1, m->V, NULL) ||
        rf_sur_add_scene(m->U, m->V, b1->w, m->h1z, q, P, W, m->st) ||
        rf_sur_gelu(m->h1z, NULL, m->h1, q * W, m->st) ||
        gemm(m, (int)q, W, W, m->h1, W, 1, W2->w16, W, 1, m->h2z, b2->w) ||
        rf_sur_gelu(m->h2z, NULL, m->h2, q * W, m->st) ||
        gemm(m, (int)q, O, W, m->h2, W, 1, W3->w16, O, 1, m->out, b3->w))
        return -1;
    return 0;
}

/* Gradient of the batch loss into every p[k].g; *loss on the host. */
static int backward(fdtd_sur_t *m, int n, double *loss) {
    const fdtd_sur_config_t *c = &m->c;
    const int W = c->width, P = c->queries, O = c->outputs;
    const long long q = (long long)n * P, nout = q * O;
    param_t *W1f = &m->p[2 * NL], *W1c = &m->p[2 * NL + 1], *b1 = &m->p[2 * NL + 2], *W2 = &m->p[2 * NL + 3],
            *b2 = &m->p[2 * NL + 4], *W3 = &m->p[2 * NL + 5], *b3 = &m->p[2 * NL + 6];
    if (rf_sur_mse(m->out, m->labels, m->dout, m->dout16, nout, m->d_loss, m->n_loss, m->st)) return -1;
    /* output layer */
    if (gemm(m, W, O, (int)q, m->h2, 1, W, m->dout16, O, 1, W3->g, NULL) ||
        gemm(m, 1, O, (int)q, m->ones, 0, 1, m->dout16, O, 1, b3->g, NULL) ||
        gemm(m, (int)q, W, O, m->dout16, O, 1, W3->w16, 1, O, m->dh, NULL) ||
        rf_sur_gelu_bwd(m->h2z, m->dh, m->dz, m->dz16, q * W, m->st))
        return -1;
    /* second layer */
    if (gemm(m, W, W, (int)q, m->h1, 1, W, m->dz16, W, 1, W2->g, NULL) ||
        gemm(m, 1, W, (int)q, m->ones, 0, 1, m->dz16, W, 1, b2->g, NULL) ||
        gemm(m, (int)q, W, W, m->dz16, W, 1, W2->w16, 1, W, m->dh, NULL) ||
        rf_sur_gelu_bwd(m->h1z, m->dh, m->dz, m->dz16, q * W, m->st))
        return -1;
    /* first layer: query part, scene part (sum over the scene's queries) */
    if (gemm(m, c->coords, W, (int)q, m->coords16, 1, c->coords, m->dz16, W, 1, W1c->g, NULL) ||
        gemm(m, 1, W, (int)q, m->ones, 0, 1, m->dz16, W, 1, b1->g, NULL) ||
        rf_sur_sum_queries(m->dz, m->dU, m->dU16, n, P, W, m->st) ||
        gemm(m, (int)m->F, W, n, m->a16[NL - 1], 1, m->F, m->dU16, W, 1, W1f->g, NULL) ||
        gemm(m, n, (int)m->F, W, m->dU16, W, 1, W1f->w16, 1, W, m->dfeat, NULL))
        return -1;
    /* convolutions, last to first: dact holds the gradient of layer l's
     * activation */
    const float *dact = m->dfeat;
    for (int l = NL - 1; l >= 0; l--) {
        const long long so = m->S[l + 1], rows = (long long)n * so * so * so;
        const int K = 27 * m->C[l], N = m->C[l + 1];
        if (rf_sur_gelu_bwd(m->z[l], dact, m->dzc, m->dzc16, rows * N, m->st) ||
            gemm(m, K, N, (int)rows, m->col[l], 1, K, m->dzc16, N, 1, m->p[2 * l].g, NULL) ||
            gemm(m, 1, N, (int)rows, m->ones, 0, 1, m->dzc16, N, 1, m->p[2 * l + 1].g, NULL))
            return -1;
        if (l == 0) break;
        if (gemm(m, (int)rows, K, N, m->dzc16, N, 1, m->p[2 * l].w16, 1, N, m->dcol, NULL) ||
            rf_sur_col2im(m->dcol, m->dact, n, m->S[l], m->C[l], m->st))
            return -1;
        dact = m->dact;
    }
    if (loss) {
        double part[256];
        if (rf_d2h(part, m->d_loss, (size_t)m->n_loss * sizeof(double), m->st)) return -1;
        double s = 0.0;
        for (int k = 0; k < m->n_loss; k++) s += part[k];
        *loss = s / (double)nout;
    }
    return 0;
}

static int check_n(const fdtd_sur_t *m, int n) { return m && n >= 1 && n <= m->c.max_batch; }

int fdtd_sur_predict(fdtd_sur_t *m, const float *maps, const float *coords, int n, float *out) {
    if (!check_n(m, n) || !maps || !coords || !out) return -1;
    if (upload_batch(m, maps, coords, NULL, n) || forward(m, n)) return -1;
    return rf_d2h(out, m->out, (size_t)n * m->c.queries * m->c.outputs * 4, m->st);
}

int fdtd_sur_gradient(fdtd_sur_t *m, const float *maps, const float *coords, const float *labels, int n, float *grad,
                      double *loss) {
    if (!check_n(m, n) || !maps || !coords || !labels) return -1;
    double l = 0.0;
    if (upload_batch(m, maps, coords, labels, n) || forward(m, n) || backward(m, n, &l)) return -1;
    if (loss) *loss = l;
    if (grad) {
        size_t off = 0;
        for (int k = 0; k < m->n_par; k++) {
            if (rf_d2h(grad + off, m->p[k].g, m->p[k].n * 4, m->st)) return -1;
            off += m->p[k].n;
        }
    }
    return 0;
}

int fdtd_sur_train_step(fdtd_sur_t *m, const float *maps, const float *coords, const float *labels, int n, float lr,
                        double *loss) {
    if (!check_n(m, n) || !maps || !coords || !labels) return -1;
    double l = 0.0;
    if (upload_batch(m, maps, coords, labels, n) || forward(m, n) || backward(m, n, &l)) return -1;
    m->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - pow(b1, (double)m->adam_t)), c2 = (float)(1.0 - pow(b2, (double)m->adam_t));
    for (int k = 0; k < m->n_par; k++) {
        param_t *q = &m->p[k];
        if (rf_sur_adam(q->w, q->w16, q->g, q->m, q->v, (long long)q->n, lr, b1, b2, 1e-8f, c1, c2, m->st)) return -1;
    }
    if (rf_sync(m->st)) return -1;
    if (loss) *loss = l;
    return 0;
}
