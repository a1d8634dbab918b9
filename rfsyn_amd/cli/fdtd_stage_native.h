/*
 * The CLI's FDTD stage on the MI355X core, as a drop-in for the
 * reference's cli/fdtd_stage.h: the same transform_t plugin
 * (engine/core/transform.h) built from the same fdtd_sim_config_t, plus
 * the options of the device-resident path (precision, and the AWGN stage
 * done on the device in the same call).  This file and its .c are
 * compiled into the CLI in place of cli/fdtd_stage.{h,c}, against the
 * CLI's engine/core headers and this repository's include/ headers, and
 * linked with librfsyn_fdtd.so.
 */
#ifndef RFSYN_FDTD_STAGE_NATIVE_H
#define RFSYN_FDTD_STAGE_NATIVE_H

#include "transform.h"
#include "fdtd_native.h"

/* cli/fdtd_stage.h:7 with the device options; opt NULL = FP64 without
 * noise, which gives the reference stage's samples. */
transform_t fdtd_stage_native_create(const fdtd_sim_config_t *cfg, const fdtd_echo_options_t *opt);

#endif
