// Material assignment + subpixel smoothing entry point.
#pragma once
#include "grid.hpp"

namespace mfd {
int matparams(mfd_ctx *ctx, const mfd_grid *g, int64_t nobj, const mfd_shape *shapes, const mfd_c64 *eps_t,
              const mfd_c64 *mu_t, mfd_c64 *eps_arr, mfd_c64 *mu_arr);
}
