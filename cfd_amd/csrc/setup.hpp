// setup.hpp -- host-side setup of the rough-channel substep: stretched y grid,
// immersed-boundary point lists, wavenumbers.  Native replacement of the
// parts of start.f90 that the hot path depends on.
#pragma once
#include <string>
#include <vector>

#include "cfd_amd.h"

namespace cfd {

// Staggered y grid.  u/w centres j = jlo_u .. jlo_u+nyu-1 (ghosts included),
// v faces j = jlo_v .. jlo_v+nyv-1 (walls included).  d2u/d2v hold the
// 3-point second-derivative coefficients (3 per point).
struct YGrid {
  int nyu = 0, nyv = 0, jlo_u = 0, jlo_v = 0;
  std::vector<double> yu, yv, dthdyu, dthdyv, d2u, d2v;
  double& YU(int j) { return yu[j - jlo_u]; }
  double& YV(int j) { return yv[j - jlo_v]; }
  double YU(int j) const { return yu[j - jlo_u]; }
  double YV(int j) const { return yv[j - jlo_v]; }
};

// One IB list entry as boundary_circ_rough writes it (1-based i,k; j plane).
struct IBPoint {
  int i, k, j, i2, k2, j2, i3, k3, j3;
  double lap, w2, w3;
};

bool build_ygrid(const cfd_params& p, YGrid& g, std::string& err);
// the first face j that is not strictly between its centres (yu(j) < yv(j) <
// yu(j+1)), INT_MIN for a grid that increases everywhere.  The reference's
// maps fold for some inputs (a canopy depth posth far from dsty times the tip
// spacing, posth = 0 with dsty > 0, steep polynomial maps)
int ygrid_fold(const YGrid& g);
// lists[grid][wall]: grid 0 = u/w centres, 1 = v faces; wall 0 bottom, 1 top.
bool build_iblists(const cfd_params& p, std::vector<IBPoint> lists[2][2], std::string& err);

}  // namespace cfd
