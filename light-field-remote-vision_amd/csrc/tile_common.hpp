// Shared pieces of the tile (binned splat) kernels.
#pragma once

#include "lfrv_internal.hpp"

namespace lfrv {

// Target tiling: tiles of tw x th pixels, ntx per row, ntiles in total.
struct TileGeom {
    int W, H, tw, th, ntx, ntiles;
};

// Per-tile entry lists (structure of arrays): entries start[t] .. start[t+1]-1
// of tile t; src = view * n + m; (tx, ty) = tau; w = omega of the source pixel.
struct TileEntries {
    const int *start;
    const int *src;
    const float *tx, *ty, *w;
    int active;  // 0: the tile kernels skip the entry loop
};

TileGeom tile_geom(const lfrv_ctx *ctx);
TileEntries tile_entries(const lfrv_ctx *ctx);
size_t tile_lds_bytes(const lfrv_ctx *ctx, int planes);

lfrv_status run_rhs_and_diag(lfrv_ctx *ctx, double alpha, double c0, double beta, bool use_views,
                             bool use_precond, const float *u0);
lfrv_status run_gauss_smooth(lfrv_ctx *ctx, const float *in, float *out, float *tmp);
lfrv_status run_view_stencils(lfrv_ctx *ctx);

}  // namespace lfrv
