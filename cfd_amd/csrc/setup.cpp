// setup.cpp -- y grid and immersed-boundary lists for the rough channel.
//
// y grid: y_grid_canopy.f90:24-248 (polynomial channel map a x^4 + b x^2 + c
// between the post tips, quintic or two-layer septic map inside the canopy,
// 3-point non-uniform second-derivative coefficients).  ygrid = 1 swaps the
// tip-to-tip map for y = tanh(beta theta/q)/tanh(beta) and keeps the canopy
// extension.
//
// IB lists: boundary_circ_rough (geom_circ_rough_mixed_forcing_24pt.f90:1-622)
// and MergeSort3D (mergeSort.f90:81-129).  The integer outputs must agree bit
// for bit with the reference, so the floating-point expressions below keep the
// reference's operation order and its real-valued mod arithmetic.
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdio>
#include <string>

namespace cfd {
namespace {

// real**integer with the binary-exponentiation lowering of Fortran.
double powi(double x, int e) {
  double r = 1.0;
  for (;;) {
    if (e & 1) r *= x;
    e /= 2;
    if (e == 0) break;
    x *= x;
  }
  return r;
}

}  // namespace

bool build_ygrid(const cfd_params& p, YGrid& g, std::string& err) {
  const int dny = p.dsty, dny1 = p.dsty1, ny = p.ny;
  const int nn = ny - 1;  // points between tips minus one (y_grid_canopy.f90:31)
  g.jlo_u = -dny;
  g.jlo_v = -dny;
  g.nyu = ny + 2 * dny + 1;
  g.nyv = ny + 2 * dny;
  g.yu.assign(g.nyu, 0.0);
  g.dthdyu.assign(g.nyu, 0.0);
  g.d2u.assign(3 * g.nyu, 0.0);
  g.yv.assign(g.nyv, 0.0);
  g.dthdyv.assign(g.nyv, 0.0);
  g.d2v.assign(3 * g.nyv, 0.0);
  if (p.ygrid == 0 && p.ppp < 3) {
    err = "ppp must be >= 3";
    return false;
  }
  const double q = nn / 2.0;
  const int pp = p.ppp;
  double aaa = 0, bbb = 0, ccc = 0;
  if (p.ygrid == 0) {
    aaa = 15.0 * (p.dyq - 1.0) / ((8.0 * p.dyq + 7.0) * powi(q, pp));
    bbb = -30.0 * (p.dyq - 1.0) / ((8.0 * p.dyq + 7.0) * powi(q, pp - 2));
    ccc = 15.0 * p.dyq / ((8.0 * p.dyq + 7.0) * q);
  }
  const double tb = p.ytanh_beta;
  auto ymap = [&](double th) {
    if (p.ygrid == 0) return aaa * powi(th, pp) / 5.0 + bbb * powi(th, pp - 2) / 3.0 + ccc * th;
    return std::tanh(tb * th / q) / std::tanh(tb);
  };
  auto dthdy = [&](double th) {
    if (p.ygrid == 0) return 1.0 / (aaa * powi(th, pp - 1) + bbb * powi(th, pp - 3) + ccc);
    const double c = std::cosh(tb * th / q);
    return 1.0 / (tb / (q * std::tanh(tb)) / (c * c));
  };
  for (int j = 0; j <= nn - 1; ++j) {
    g.YV(j) = ymap(j - q);
    g.YU(j + 1) = ymap(j + 0.5 - q);
    g.dthdyv[j - g.jlo_v] = dthdy(j - q);
    g.dthdyu[j + 1 - g.jlo_u] = dthdy(j + 0.5 - q);
  }
  auto DU = [&](int j) -> double& { return g.dthdyu[j - g.jlo_u]; };
  auto DV = [&](int j) -> double& { return g.dthdyv[j - g.jlo_v]; };
  if (dny == 0) {
    g.YU(0) = ymap(-0.5 - q);
    DU(0) = dthdy(-0.5 - q);
    g.YV(nn) = ymap(nn - q);
    g.YU(nn + 1) = ymap(nn + 0.5 - q);
    DV(nn) = dthdy(nn - q);
    DU(nn + 1) = dthdy(nn + 0.5 - q);
  } else {
    const double tip = g.YV(1) - g.YV(0);
    const double posth = p.posth, posth1 = p.posth1;
    if (dny1 == 0 || dny1 == dny || posth1 == 0.0 || posth1 == posth) {
      // single quintic layer (y_grid_canopy.f90:101-117)
      const double a5 = -(15.0 * (posth - tip * dny)) / (7.0 * std::pow(double(dny), 5.0));
      const double a3 = (30.0 * (posth - tip * dny)) / (7.0 * std::pow(double(dny), 3.0));
      const double a1 = tip;
      const double x0 = -dny - 0.5;
      g.YU(-dny) = a5 * powi(x0, 5) / 5.0 + a3 * powi(x0, 3) / 3.0 + a1 * x0 - 1.0;
      DU(-dny) = 1.0 / (a5 * powi(x0, 4) + a3 * powi(x0, 2) + a1);
      for (int side = 0; side < 2; ++side) {
        const int ja = side ? 0 : -dny, jb = side ? dny : -1, sh = side ? nn : 0;
        const double off = side ? 1.0 : -1.0;
        for (int j = ja; j <= jb; ++j) {
          const double xj = j, xh = j + 0.5;
          g.YV(j + sh) = a5 * std::pow(xj, 5.0) / 5.0 + a3 * std::pow(xj, 3.0) / 3.0 + a1 * j + off;
          g.YU(j + sh + 1) = a5 * powi(xh, 5) / 5.0 + a3 * powi(xh, 3) / 3.0 + a1 * xh + off;
          DV(j + sh) = 1.0 / (a5 * std::pow(xj, 4.0) + a3 * std::pow(xj, 2.0) + a1);
          DU(j + sh + 1) = 1.0 / (a5 * powi(xh, 4) + a3 * powi(xh, 2) + a1);
        }
      }
    } else {
      // two layers: quintic above the turning point, septic below
      // (y_grid_canopy.f90:119-201)
      const double a5 = -(15.0 * (posth1 - tip * dny1)) / (7.0 * std::pow(double(dny1), 5.0));
      const double a3 = (30.0 * (posth1 - tip * dny1)) / (7.0 * std::pow(double(dny1), 3.0));
      const double a1 = tip;
      auto up = [&](int ja, int jb, int sh, double off) {
        for (int j = ja; j <= jb; ++j) {
          const double xj = j, xh = j + 0.5;
          g.YV(j + sh) = a5 * std::pow(xj, 5.0) / 5.0 + a3 * std::pow(xj, 3.0) / 3.0 + a1 * j + off;
          g.YU(j + sh + 1) = a5 * std::pow(xh, 5.0) / 5.0 + a3 * std::pow(xh, 3.0) / 3.0 + a1 * xh + off;
          DV(j + sh) = 1.0 / (a5 * std::pow(xj, 4.0) + a3 * std::pow(xj, 2.0) + a1);
          DU(j + sh + 1) = 1.0 / (a5 * std::pow(xh, 4.0) + a3 * std::pow(xh, 2.0) + a1);
        }
      };
      up(-dny, -1, 0, -1.0);
      up(0, dny1, nn, 1.0);
      const double gm = g.YV(-dny1 + 1) - g.YV(-dny1);
      const double D = dny, E = dny1;
      auto P = [](double x, double e) { return std::pow(x, e); };
      const double den = ((D - E) * (11.0 * P(D, 6.0) + 11.0 * P(D, 5.0) * E - 38.0 * P(D, 4.0) * P(E, 2.0) -
                                     38.0 * P(D, 3.0) * P(E, 3.0) + 67.0 * P(D, 2.0) * P(E, 4.0) +
                                     11.0 * D * P(E, 5.0) - 24.0 * P(E, 6.0)));
      const double b7 = -(70.0 * (posth - posth1 - D * gm + E * gm)) / den;
      const double b5 = (105.0 * (P(D, 2.0) + P(E, 2.0)) * (posth - posth1 - D * gm + E * gm)) / den;
      const double b3 = (210.0 * P(E, 2.0) * (P(D, 3.0) * gm - P(D, 2.0) * posth + P(D, 2.0) * posth1 -
                                              P(D, 2.0) * E * gm)) / den;
      const double b1 = (11.0 * P(D, 7.0) * gm - 11.0 * P(E, 7.0) * gm - 35.0 * P(E, 6.0) * posth +
                         35.0 * P(E, 6.0) * posth1 + 49.0 * P(D, 2.0) * P(E, 5.0) * gm -
                         49.0 * P(D, 5.0) * P(E, 2.0) * gm + 105.0 * P(D, 2.0) * P(E, 4.0) * posth -
                         105.0 * P(D, 2.0) * P(E, 4.0) * posth1) / den;
      const double b0 = -(-11.0 * gm * P(D, 7.0) * E + 11.0 * posth1 * P(D, 7.0) + 49.0 * gm * P(D, 5.0) * P(E, 3.0) -
                          49.0 * posth1 * P(D, 5.0) * P(E, 2.0) - 49.0 * gm * P(D, 3.0) * P(E, 5.0) +
                          105.0 * posth1 * P(D, 3.0) * P(E, 4.0) - 56.0 * posth * P(D, 2.0) * P(E, 5.0) +
                          11.0 * gm * D * P(E, 7.0) - 35.0 * posth1 * D * P(E, 6.0) + 24.0 * posth * P(E, 7.0)) / den;
      auto Ys = [&](double x, double c0) {
        return b7 * P(x, 7.0) / 7.0 + b5 * P(x, 5.0) / 5.0 + b3 * P(x, 3.0) / 3.0 + b1 * x + c0;
      };
      auto Ds = [&](double x) { return 1.0 / (b7 * P(x, 6.0) + b5 * P(x, 4.0) + b3 * P(x, 2.0) + b1); };
      for (int j = -dny; j <= -dny1 - 1; ++j) {
        g.YV(j) = Ys(j, b0) - 1.0;
        g.YU(j + 1) = Ys(j + 0.5, b0) - 1.0;
        DV(j) = Ds(j);
        DU(j + 1) = Ds(j + 0.5);
      }
      g.YU(-dny) = Ys(-dny - 0.5, b0) - 1.0;
      DU(-dny) = Ds(-dny - 0.5);
      for (int j = dny1 + 1; j <= dny; ++j) {
        g.YV(j + nn) = Ys(j, -b0) + 1.0;
        g.YU(j + nn + 1) = Ys(j + 0.5, -b0) + 1.0;
        DV(j + nn) = Ds(j);
        DU(j + nn + 1) = Ds(j + 0.5);
      }
    }
  }
  // second-derivative coefficients (y_grid_canopy.f90:217-248)
  auto c2 = [](std::vector<double>& d, int idx, double a, double b, double c) {
    d[3 * idx] = a;
    d[3 * idx + 1] = b;
    d[3 * idx + 2] = c;
  };
  auto interior = [&](std::vector<double>& d, int idx, double ym, double y0, double yp) {
    c2(d, idx, 2.0 / ((y0 - ym) * (yp - ym)), -2.0 / ((yp - y0) * (y0 - ym)), 2.0 / ((yp - y0) * (yp - ym)));
  };
  {
    const int j = -dny;
    const double hu = g.YU(j + 1) - g.YU(j), hv = g.YV(j + 1) - g.YV(j);
    c2(g.d2u, 0, 0.0, -2.0 / (hu * hu), 1.0 / (hu * hu));
    c2(g.d2v, 0, 0.0, -2.0 / (hv * hv), 1.0 / (hv * hv));
  }
  for (int j = -dny + 1; j <= nn + dny - 1; ++j) {
    interior(g.d2u, j - g.jlo_u, g.YU(j - 1), g.YU(j), g.YU(j + 1));
    interior(g.d2v, j - g.jlo_v, g.YV(j - 1), g.YV(j), g.YV(j + 1));
  }
  {
    const int j = nn + dny;
    interior(g.d2u, j - g.jlo_u, g.YU(j - 1), g.YU(j), g.YU(j + 1));
    const double hv = g.YV(j) - g.YV(j - 1);
    c2(g.d2v, j - g.jlo_v, 1.0 / (hv * hv), -2.0 / (hv * hv), 0.0);
    const int k = nn + dny + 1;
    const double hu = g.YU(k) - g.YU(k - 1);
    c2(g.d2u, k - g.jlo_u, 1.0 / (hu * hu), -2.0 / (hu * hu), 0.0);
  }
  return true;
}

int ygrid_fold(const YGrid& g) {
  for (int j = g.jlo_v; j < g.jlo_v + g.nyv; ++j)
    if (!(g.YU(j) < g.YV(j) && g.YV(j) < g.YU(j + 1))) return j;
  return INT_MIN;
}

// ---------------------------------------------------------------------------
// IB lists
// ---------------------------------------------------------------------------
namespace {

struct Circle {
  double r, ci, ck;
  double f(double x, double z) const { return (x - ci) * (x - ci) + (z - ck) * (z - ck) - r * r; }
};

// Interpolation stencil of a forcing point (lin_interp, :634-738): two fluid
// neighbours (x2,z2), (x3,z3) and weights reproducing a linear field that
// vanishes on the circle.
struct Stencil {
  int z2, x2, z3, x3;
  double w1, w2;
};

bool stencil(const Circle& c, double ie, double ke, Stencil& s) {
  if (ie == c.ci && ke == c.ck) return false;  // normal undefined at the centre
  const double aa = (ie >= c.ci) ? 1.0 : -1.0;
  const double zb = c.ck + (ke - c.ck) * c.r / std::sqrt((ke - c.ck) * (ke - c.ck) + (ie - c.ci) * (ie - c.ci));
  double xb = aa * std::sqrt(c.r * c.r - (zb - c.ck) * (zb - c.ck)) + c.ci;
  double nz = (zb - c.ck) / std::sqrt((zb - c.ck) * (zb - c.ck) + (xb - c.ci) * (xb - c.ci));
  const double nx = (xb - c.ci) / std::sqrt((zb - c.ck) * (zb - c.ck) + (xb - c.ci) * (xb - c.ci));
  if (ie == c.ci) {  // fore/aft of the circle: one neighbour along z
    xb = c.ci;
    nz = (zb - c.ck) / std::sqrt((zb - c.ck) * (zb - c.ck) + (xb - c.ci) * (xb - c.ci));
    const int sz = int(nz / std::fabs(nz));
    s.z2 = s.z3 = int(ke + sz);
    s.x2 = s.x3 = int(ie);
    s.w1 = (ke - zb) / (s.z2 - zb);
    s.w2 = 0.0;
    return true;
  }
  if (ke == c.ck) {  // sides: one neighbour along x
    const int sx = int(nx / std::fabs(nx));
    s.z2 = s.z3 = int(ke);
    s.x2 = s.x3 = int(ie + sx);
    s.w1 = (ie - xb) / (s.x2 - xb);
    s.w2 = 0.0;
    return true;
  }
  const int sz = int(nz / std::fabs(nz)), sx = int(nx / std::fabs(nx));
  if (std::fabs(ie - c.ci) == std::fabs(ke - c.ck)) {
    s.z2 = int(ke + sz); s.x2 = int(ie); s.z3 = int(ke); s.x3 = int(ie + sx);
  } else if (std::fabs(nx) < std::fabs(nz)) {
    s.z2 = int(ke + sz); s.x2 = int(ie); s.z3 = int(ke + sz); s.x3 = int(ie + sx);
  } else if (std::fabs(nx) > std::fabs(nz)) {
    s.z2 = int(ke); s.x2 = int(ie + sx); s.z3 = int(ke + sz); s.x3 = int(ie + sx);
  } else {
    return false;
  }
  const int x2 = s.x2, z2 = s.z2, x3 = s.x3, z3 = s.z3;
  const double den = xb * z2 - x2 * zb - xb * z3 + x3 * zb + double(x2 * z3) - double(x3 * z2);
  s.w1 = -(xb * z3 - x3 * zb - xb * ke + ie * zb + x3 * ke - ie * z3) / den;
  s.w2 = (xb * z2 - x2 * zb - xb * ke + ie * zb + x2 * ke - ie * z2) / den;
  return true;
}

// A point of the single-tile cross-section: position, two neighbours,
// weights and Laplacian coefficient (1 forcing, 0 solid).
struct TilePoint {
  int i, k, i2, k2, i3, k3;
  double w2, w3, lap;
};

// 1 + mod(v - 1 + d, d) evaluated in double precision, then truncated.
int wrap(int v, int d) { return int(1.0 + std::fmod(double(v) - 1.0 + d, double(d))); }

}  // namespace

bool build_iblists(const cfd_params& p, std::vector<IBPoint> lists[2][2], std::string& err) {
  for (int g = 0; g < 2; ++g)
    for (int w = 0; w < 2; ++w) lists[g][w].clear();
  if (p.geometry_type != 3) return true;
  if (p.ntilex == 0 || p.ntilez == 0) { err = "ntilex/ntilez equals 0"; return false; }
  if (p.ntilex % 2 != 0) { err = "ntilex is not even"; return false; }
  if (p.ngalx % p.ntilex != 0) { err = "nx not multiple of ntilex"; return false; }
  if (p.ngalz % p.ntilez != 0) { err = "nz not multiple of ntilez"; return false; }
  const int dnx = p.ngalx / p.ntilex, dnz = p.ngalz / p.ntilez;
  const Circle C{p.r, p.c_i, p.c_k};
  const int x1 = int(std::floor(C.ci - C.r - 1)), x2 = int(std::ceil(C.ci + C.r + 1));
  const int z1 = int(std::floor(C.ck - C.r - 1)), z2 = int(std::ceil(C.ck + C.r + 1));
  if (x1 < 1 || z1 < 1 || x2 > dnx || z2 > dnz) { err = "Circ too close to edge"; return false; }
  auto inside = [&](int i, int k) { return C.f(i, k); };

  std::vector<TilePoint> tile;
  auto solid = [&](int i, int k) {
    const int a = wrap(i, dnx), b = wrap(k, dnz);
    tile.push_back({a, b, a, b, a, b, 0.0, 0.0, 0.0});
  };
  auto forcing = [&](int i, int k, const Stencil& s) {
    tile.push_back({wrap(i, dnx), wrap(k, dnz), wrap(s.x2, dnx), wrap(s.z2, dnz), wrap(s.x3, dnx), wrap(s.z3, dnz),
                    s.w1, s.w2, 1.0});
  };
  // classify the cross-section of one post (:135-215)
  for (int k = z1 + 1; k <= z2 - 1; ++k) {
    for (int i = x1 + 1; i <= x2 - 1; ++i) {
      const double f = inside(i, k);
      if (f < 0.0) {
        const bool edge = inside(i - 1, k) > 0.0 || inside(i + 1, k) > 0.0 || inside(i, k + 1) > 0.0 ||
                          inside(i, k - 1) > 0.0;
        if (edge) {
          Stencil s;
          if (!stencil(C, i, k, s)) { err = "IB stencil: degenerate normal"; return false; }
          if (s.w1 < -1.0 || s.w2 < -1.0) {
            // too close to the surface: make it solid and force the outward neighbour
            solid(i, k);
            const int io = s.x2, ko = s.z2;
            Stencil so;
            if (!stencil(C, io, ko, so)) { err = "IB stencil: degenerate normal"; return false; }
            forcing(io, ko, so);
          } else {
            forcing(i, k, s);
          }
        } else {
          solid(i, k);
        }
      }
      if (f == 0.0) solid(i, k);
    }
  }
  if (p.extras24) {  // the four fixed points of the 24-point tile (:219-259)
    const int ex[4][2] = {{8, 16}, {16, 16}, {16, 8}, {8, 8}};
    for (auto& e : ex) {
      Stencil s;
      if (!stencil(C, e[0], e[1], s)) { err = "IB stencil: degenerate normal"; return false; }
      forcing(e[0], e[1], s);
    }
  }
  // planes of each stem; the plane at the post tip forces every point (Lap=1)
  const int dsty = p.dsty, ny = p.ny;
  struct Span { int j0, j1, jtip; };
  const Span span[2][2] = {
      {{-dsty + 1, 0, 0}, {ny, ny - 1 + dsty, ny}},          // u/w centres: bottom, top
      {{-dsty + 1, -1, -1}, {ny, ny - 2 + dsty, ny}}};       // v faces
  for (int g = 0; g < 2; ++g) {
    for (int w = 0; w < 2; ++w) {
      const Span s = span[g][w];
      std::vector<IBPoint> stem;
      // bottom wall lists the tip plane last, the top wall lists it first
      std::vector<int> planes;
      if (w == 0) {
        for (int j = s.j0; j <= s.j1; ++j) if (j != s.jtip) planes.push_back(j);
        planes.push_back(s.jtip);
      } else {
        planes.push_back(s.jtip);
        for (int j = s.j0; j <= s.j1; ++j) if (j != s.jtip) planes.push_back(j);
      }
      for (int j : planes)
        for (const TilePoint& t : tile)
          stem.push_back({t.i, t.k, j, t.i2, t.k2, j, t.i3, t.k3, j, j == s.jtip ? 1.0 : t.lap, t.w2, t.w3});
      // replicate over the tiles, every other tile row staggered in z (:462-577)
      std::vector<IBPoint>& out = lists[g][w];
      out.resize(stem.size() * p.ntilex * p.ntilez);
      const size_t ps = stem.size();
      for (int ix = 1; ix <= p.ntilex; ++ix) {
        const double stg = ((ix + 1) % 2) * dnz * p.stag;
        for (int iz = 1; iz <= p.ntilez; ++iz) {
          const size_t base = ps * (iz - 1) + ps * p.ntilez * (ix - 1);
          auto zk = [&](int k) {
            return int(1.0 + std::fmod(double(k + dnz * (iz - 1)) + stg - 1.0, 1.0 * p.ngalz));
          };
          for (size_t q = 0; q < ps; ++q) {
            IBPoint e = stem[q];
            e.i += dnx * (ix - 1); e.i2 += dnx * (ix - 1); e.i3 += dnx * (ix - 1);
            e.k = zk(e.k); e.k2 = zk(e.k2); e.k3 = zk(e.k3);
            out[base + q] = e;
          }
        }
      }
      // order by plane, then k, then i, keeping ties in list order
      std::stable_sort(out.begin(), out.end(), [](const IBPoint& a, const IBPoint& b) {
        if (a.j != b.j) return a.j < b.j;
        if (a.k != b.k) return a.k < b.k;
        return a.i < b.i;
      });
    }
  }
  return true;
}

}  // namespace cfd
