// spatial.hpp -- packed spin-free active-space densities (spatial.cc).
#pragma once
#include <cstdint>

#include "layout.hpp"

namespace v2g {

int64_t spatial_d1_dim(const Layout &L);
int64_t spatial_d2_dim(const Layout &L);
// x: the primal vector on the host (at least up to the end of the D1b block)
void pack_spatial_density(const Layout &L, const double *x, double *d1, double *d2);

}  // namespace v2g
