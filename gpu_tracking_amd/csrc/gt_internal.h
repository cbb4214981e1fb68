// gt_internal.h -- shared between the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gtrack.h"

#include "gt_kparams.h"

// Kernel launchers (gt_kernels.hip).  All are asynchronous on `s`.
hipError_t gt_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
hipError_t gt_launch_bandpass(const void *frames, int dtype, float *proc, uint32_t nframes,
                              const GtKParams &p, hipStream_t s);
hipError_t gt_launch_localmax(const float *proc, uint32_t nframes, const GtKParams &p,
                              uint32_t *cand, uint32_t cand_cap, uint32_t *cand_count,
                              hipStream_t s);
hipError_t gt_launch_walk(const float *proc, const void *raw, int dtype, uint32_t nframes,
                          const GtKParams &p, const uint32_t *cand, uint32_t cand_cap,
                          const uint32_t *cand_count, gt_detection *out, uint32_t out_cap,
                          uint32_t *out_count, uint32_t frame0, hipStream_t s);
