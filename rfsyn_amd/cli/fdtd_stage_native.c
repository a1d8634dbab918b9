/*
 * cli/fdtd_stage.c:10-43 on the device-resident path: the stage resizes
 * the signal to n_steps samples and fills them with the echo as complex
 * float samples (I = (float)v, Q = 0), now produced on the device by
 * fdtd_simulate_echo_iq, which also applies the AWGN stage there when the
 * options ask for it (the chain then has no separate awgn stage).
 */
#include "fdtd_stage_native.h"
#include "rf_signal.h"

#include <stdlib.h>

typedef struct {
    fdtd_sim_config_t cfg;
    fdtd_echo_options_t opt;
} native_stage_ctx_t;

static int native_stage_apply(void *ctx_v, signal_t *sig) {
    native_stage_ctx_t *ctx = (native_stage_ctx_t *)ctx_v;
    if (!ctx || !sig) return -1;
    if (signal_resize(sig, ctx->cfg.n_steps) != 0) return -1;
    double fs = 0.0;
    /* float complex is two floats, I then Q: the layout the device writes */
    if (fdtd_simulate_echo_iq(&ctx->cfg, &ctx->opt, (float *)sig->samples, &fs) != 0) return -1;
    sig->sample_rate_hz = fs;
    return 0;
}

static void native_stage_destroy(void *ctx_v) { free(ctx_v); }

transform_t fdtd_stage_native_create(const fdtd_sim_config_t *cfg, const fdtd_echo_options_t *opt) {
    if (!cfg) return (transform_t){NULL, NULL, NULL};
    native_stage_ctx_t *ctx = malloc(sizeof *ctx);
    if (!ctx) return (transform_t){NULL, NULL, NULL};
    ctx->cfg = *cfg;
    if (opt) {
        ctx->opt = *opt;
    } else {
        ctx->opt = (fdtd_echo_options_t){FDTD_FP64, 0, 0.f, 0, 0};
    }
    return (transform_t){native_stage_apply, native_stage_destroy, ctx};
}
