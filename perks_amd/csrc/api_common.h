// api_common.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

#include "j3d_common.h"
#include "perks_amd.h"

#define PERKS_HIP_TRY(expr)                                                                          \
    do {                                                                                             \
        hipError_t perks_e_ = (expr);                                                                \
        if (perks_e_ != hipSuccess) {                                                                \
            std::fprintf(stderr, "perks_amd: %s failed at %s:%d: %s\n", #expr, __FILE__, __LINE__,   \
                         hipGetErrorString(perks_e_));                                               \
            return PERKS_ERR_HIP;                                                                    \
        }                                                                                            \
    } while (0)

namespace perks {

// Resolves the coefficients of a run: the shape's reference weights, or the
// caller's.  Returns PERKS_ERR_INVALID_ARG for an unknown shape.
inline int resolve_coefs(const perks_stencil_opts* o, KCoefs* k, const ShapeDesc** sd)
{
    const ShapeDesc* s = shape_by_id(o->shape);
    if (!s) return PERKS_ERR_INVALID_ARG;
    for (int i = 0; i < PERKS_MAX_COEFS; ++i) k->c[i] = 0.0;
    if (o->use_ref_coefs) {
        for (int g = 0; g < s->ngroups; ++g) k->c[g] = s->g[g].coef;
        k->divisor = s->divisor;
    } else {
        for (int g = 0; g < s->ngroups; ++g) k->c[g] = o->coef[g];
        k->divisor = o->divisor;
    }
    if (sd) *sd = s;
    return PERKS_OK;
}

// Cached device facts (one GPU per process is the supported configuration;
// the cache is keyed by device id).
int device_caps_cached(perks_device_caps* caps);

}  // namespace perks
