// Host-side helpers shared between the library's host translation units
// (not part of the C ABI).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/sadmc.h"

#define SADMC_INTERNAL __attribute__((visibility("hidden")))

// The device a context runs on, and setting its error message (returns code).
SADMC_INTERNAL int ctx_device(const sadmc_ctx *c);
SADMC_INTERNAL int ctx_fail(sadmc_ctx *c, int code, const char *msg);
// The walker-set entropy reduction (plotting/parse-binning.py:150-171)
// behind sadmc_mean_entropy: scratch grown in the context when fixed_base is
// null, else the fixed_cap bytes of device memory at fixed_base.
SADMC_INTERNAL int mean_entropy_impl(sadmc_ctx *c, int32_t set, double e_a, double e_b, int64_t max_bins, double *e_first,
                      int64_t *nbins, double *S, int64_t *count, char *fixed_base, size_t fixed_cap);
