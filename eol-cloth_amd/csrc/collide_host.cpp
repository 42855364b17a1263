// Proximity queries (CD / CD2): placeholder until the gfx950 collision kernels land.
#include "eolc_internal.h"

extern "C" int eolc_collide(eolc_ctx *c, double, int, const double *, const double *, int, const double *,
                            const double *, int, int, int, eolc_collision *, int, int *count)
{
    if (!c) return EOLC_ERR_INVALID;
    if (count) *count = 0;
    c->err = "collide: not available yet";
    return EOLC_ERR_INVALID;
}
