// kernels_spec.hip -- spectral-side operators of the substep: explicit RHS,
// nonlinear assembly, predictor, IB forcing in physical space, the batched
// per-(kx,kz) tridiagonal solves, the (0,0) flow-rate correction, the pressure
// Poisson solve with projection, and the CFL update.
//
// Spectral fields are f[plane][column] (column = k*Mx + i); one thread owns a
// column, so every y sweep reads one contiguous run of the row per wavefront.
#include "common.hpp"

namespace cfd {

namespace {

constexpr int kTB = 256;

struct Geo {
  int Mx, Ncol, nyu, nyv, nyp;
};

Geo geo(const Solver& s) { return Geo{s.Mx, s.Ncol, s.nyu, s.nyv, s.nyp}; }

// ---------------------------------------------------------------------------
// RHS0_u1/u3 (littleharsh.f90:613-699, laplacian_U 246-282):
//   du = (aRK/Re) L u - cRK Nu - dRK i k p   (+ -dRK mpgx on the (0,0) u mode)
// RHS0_u2 (littleharsh.f90:660-700, laplacian_V 284-320):
//   dv = (aRK/Re) L v - cRK Nv - dRK dthdyv (p(j+1) - p(j))
// grid (ceil(Ncol/kTB), nyu-2, 3); z: 0 u1, 1 u2, 2 u3
// ---------------------------------------------------------------------------
struct RhsArgs {
  const double2 *u1, *u2, *u3, *p, *n1, *n2, *n3;
  double2 *d1, *d2, *d3;
  const double *d2u, *d2v, *dthdyv, *k1x, *k2x, *k1z, *k2z;
  const DevScalars* sc;
  double C1, C2, dRK, ddthetavi;
  Geo g;
};

__global__ void __launch_bounds__(kTB) k_rhs0(RhsArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y + 1;
  const int f = blockIdx.z;
  if (col >= a.g.Ncol) return;
  const int i = col % a.g.Mx, k = col / a.g.Mx;
  const double k2x = a.k2x[i], k2z = a.k2z[k];
  const size_t N = a.g.Ncol;
  if (f == 1) {
    if (q > a.g.nyv - 2) return;
    const double* d = a.d2v + 3 * q;
    const double2 vm = a.u2[(q - 1) * N + col], v0 = a.u2[q * N + col], vp = a.u2[(q + 1) * N + col];
    const double c2 = k2x + k2z + d[1];
    const double2 Lv = make_double2(d[0] * vm.x + c2 * v0.x + d[2] * vp.x, d[0] * vm.y + c2 * v0.y + d[2] * vp.y);
    const double C3 = -a.dRK * a.ddthetavi * a.dthdyv[q];
    const double2 pp = a.p[q * N + col], pm = a.p[(q - 1) * N + col];
    const double2 nu = a.n2[q * N + col];
    a.d2[q * N + col] = make_double2(a.C1 * Lv.x + a.C2 * nu.x + C3 * (pp.x - pm.x),
                                     a.C1 * Lv.y + a.C2 * nu.y + C3 * (pp.y - pm.y));
    return;
  }
  const double2* u = f == 0 ? a.u1 : a.u3;
  const double2* nu = f == 0 ? a.n1 : a.n3;
  double2* du = f == 0 ? a.d1 : a.d3;
  const double* d = a.d2u + 3 * q;
  const double2 um = u[(q - 1) * N + col], u0 = u[q * N + col], up = u[(q + 1) * N + col];
  const double c2 = k2x + k2z + d[1];
  const double2 Lu = make_double2(d[0] * um.x + c2 * u0.x + d[2] * up.x, d[0] * um.y + c2 * u0.y + d[2] * up.y);
  const double c3 = -a.dRK * (f == 0 ? a.k1x[i] : a.k1z[k]);
  const double2 gp = imulk(c3, a.p[(q - 1) * N + col]);
  const double2 n = nu[q * N + col];
  double2 r = make_double2(a.C1 * Lu.x + a.C2 * n.x + gp.x, a.C1 * Lu.y + a.C2 * n.y + gp.y);
  if (f == 0 && col == 0) r.x += -a.dRK * a.sc->mpgx;
  du[q * N + col] = r;
}

// ---------------------------------------------------------------------------
// nonlinear assembly (FOU3D.f90:158-173, 232-244):
//   Nu1 = Nxz1 + dthdyu (uv(j) - uv(j-1)), Nu3 likewise with vw,
//   Nu2 = Nxz2 + dthdyv (vv(j+1) - vv(j));   du += -gRK Nu
// grid (ceil(Ncol/kTB), nyu-2, 3)
// ---------------------------------------------------------------------------
struct CombArgs {
  const void *nxz1, *nxz2, *nxz3, *fuv, *fvw, *fvv;  // TN: double2 (FP64) or float2 (mixed)
  double2 *n1, *n2, *n3, *d1, *d2, *d3;
  const double *dthdyu, *dthdyv;
  double C1, ddthetavi;
  Geo g;
};

__device__ __forceinline__ double2 as_d2(double2 v) { return v; }
__device__ __forceinline__ double2 as_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }

template <typename TN>
__global__ void __launch_bounds__(kTB) k_nlcombine(CombArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y + 1;
  const int f = blockIdx.z;
  if (col >= a.g.Ncol) return;
  const size_t N = a.g.Ncol, o = q * N + col;
  if (f == 1) {
    if (q > a.g.nyv - 2) return;
    const TN* fvv = (const TN*)a.fvv;
    const double2 vp = as_d2(fvv[(q + 1) * N + col]), v0 = as_d2(fvv[o]), x = as_d2(((const TN*)a.nxz2)[o]);
    const double dd = a.dthdyv[q];
    const double2 n = make_double2(x.x + (vp.x - v0.x) * dd * a.ddthetavi, x.y + (vp.y - v0.y) * dd * a.ddthetavi);
    a.n2[o] = n;
    const double2 d = a.d2[o];
    a.d2[o] = make_double2(d.x + a.C1 * n.x, d.y + a.C1 * n.y);
    return;
  }
  const TN* fl = (const TN*)(f == 0 ? a.fuv : a.fvw);
  const TN* nx = (const TN*)(f == 0 ? a.nxz1 : a.nxz3);
  double2* nu = f == 0 ? a.n1 : a.n3;
  double2* du = f == 0 ? a.d1 : a.d3;
  const double2 fp = as_d2(fl[o]), fm = as_d2(fl[(q - 1) * N + col]), x = as_d2(nx[o]);
  const double dd = a.dthdyu[q];
  const double2 n = make_double2(x.x + (fp.x - fm.x) * dd * a.ddthetavi, x.y + (fp.y - fm.y) * dd * a.ddthetavi);
  nu[o] = n;
  const double2 d = du[o];
  du[o] = make_double2(d.x + a.C1 * n.x, d.y + a.C1 * n.y);
}

// ---------------------------------------------------------------------------
// CFL (FOU3D.f90:177-183, 516-544; dtv start.f90:432)
// ---------------------------------------------------------------------------
__global__ void k_reset_max(DevScalars* sc) {
  sc->m1 = 0ull;
  sc->m2 = 0ull;
  sc->m3 = 0ull;
}

__global__ void k_cfl(DevScalars* sc, double alp, double bet, int N1h, int N2h, double dthetavi, double dtv,
                      double CFL) {
  const double m1 = __longlong_as_double((long long)sc->m1);
  const double m3 = __longlong_as_double((long long)sc->m3);
  const double m2 = fmax(1e-6, __longlong_as_double((long long)sc->m2));
  const double d1 = 1.0 / (alp * N1h * m1), d3 = 1.0 / (bet * N2h * m3), d2 = 1.0 / (m2 * dthetavi);
  double dtc = d1 < d3 ? d1 : d3;
  dtc = dtc < d2 ? dtc : d2;
  sc->dtc = dtc;
  sc->dt = (2.5 * dtv < CFL * dtc) ? 2.5 * dtv : CFL * dtc;
  sc->dti = 1.0 / sc->dt;
  sc->t = sc->t + sc->dt;
}

// ---------------------------------------------------------------------------
// predictor r = u + dt du, zero BC rows (solveU/solveV, littleharsh.f90:732-794)
// grid (ceil(Ncol/kTB), ny_g)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kTB) k_predictor(double2* du, const double2* u, int Ncol, int ny,
                                                   const DevScalars* sc) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (col >= Ncol) return;
  const size_t o = (size_t)q * Ncol + col;
  if (q == 0 || q == ny - 1) {
    du[o] = make_double2(0.0, 0.0);
    return;
  }
  const double dt = sc->dt;
  const double2 d = du[o], uu = u[o];
  du[o] = make_double2(uu.x + dt * d.x, uu.y + dt * d.y);
}

// ---------------------------------------------------------------------------
// IB direct forcing in physical space (immersed_boundaries_U/V,
// tridLU_3D_IB.f90:188-211, 248-271): one thread per distinct point; repeated
// entries of a point are applied in list order.
// ---------------------------------------------------------------------------
struct IBApplyArgs {
  const IBDev* L[3];
  const int* seg[3];
  int nseg[3];
  double* pr[3];
  const double* pu[3];
};

// grid (ceil(max nseg / 128), fields): blockIdx.y takes field f0 + y
__global__ void k_ib_apply(IBApplyArgs a, size_t plane_sz) {
  const int y = blockIdx.y;
  const IBDev* __restrict__ L = a.L[0];
  const int* __restrict__ seg = a.seg[0];
  int nseg = a.nseg[0];
  double* pr = a.pr[0];
  const double* pu = a.pu[0];
  if (y == 1) {
    L = a.L[1]; seg = a.seg[1]; nseg = a.nseg[1]; pr = a.pr[1]; pu = a.pu[1];
  } else if (y == 2) {
    L = a.L[2]; seg = a.seg[2]; nseg = a.nseg[2]; pr = a.pr[2]; pu = a.pu[2];
  }
  const int sgi = blockIdx.x * blockDim.x + threadIdx.x;
  if (sgi >= nseg) return;
  const int e0 = seg[sgi], e1 = seg[sgi + 1];
  const IBDev first = L[e0];
  double* r = pr + (size_t)first.plane * plane_sz;
  const double* u = pu + (size_t)first.plane * plane_sz;
  double v = r[first.o];
  for (int e = e0; e < e1; ++e) {
    const IBDev E = L[e];
    const double vf = E.w2 * u[E.o2] + E.w3 * u[E.o3];
    v = E.lap * (vf - u[E.o] + v);
  }
  r[first.o] = v;
}

// ---------------------------------------------------------------------------
// u/w/v implicit solves (LUsolU/LUsolV, tridLU_3D_IB.f90:50-160, with
// LU_buildU/V 349-413 and LU_dec 521-559), matrices built on the fly.
// grid (ceil(Ncol/64)), one thread per (kx,kz) column.
// ---------------------------------------------------------------------------
struct YArgs {
  double2* u;
  const double2* r;
  double* cs;
  const double* d2;
  const double *k2x, *k2z;
  const DevScalars* sc;
  double bRK, Re, g0, g1;  // BC coefficients of the first/last rows
  int Mx, Ncol, n;
};

__global__ void __launch_bounds__(64) k_ysolve(YArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.Ncol) return;
  const int i = col % a.Mx, k = col / a.Mx;
  const double k2x = a.k2x[i], k2z = a.k2z[k];
  const double beta = a.bRK * a.sc->dt / a.Re;
  const size_t N = a.Ncol;
  const int n = a.n;
  // row 0: (0, 1, g0)
  double d = 1.0;  // 1/a2(0) with a2(0) = 1
  double a3p = a.g0;
  double2 x = a.r[col];
  x = make_double2(x.x * d, x.y * d);
  a.u[col] = x;
  for (int q = 1; q < n; ++q) {
    double a1, a2, a3;
    if (q < n - 1) {
      const double* dd = a.d2 + 3 * q;
      a1 = -beta * dd[0];
      a2 = 1.0 - beta * (dd[1] + k2x + k2z);
      a3 = -beta * dd[2];
    } else {
      a1 = a.g1;
      a2 = 1.0;
      a3 = 0.0;
    }
    const double c = a3p * d;  // modified upper diagonal of row q-1
    a.cs[(q - 1) * N + col] = c;
    d = 1.0 / (a2 - a1 * c);
    const double2 rq = a.r[q * N + col];
    x = make_double2((rq.x - a1 * x.x) * d, (rq.y - a1 * x.y) * d);
    a.u[q * N + col] = x;
    a3p = a3;
  }
  for (int q = n - 2; q >= 0; --q) {
    const double c = a.cs[q * N + col];
    const double2 y = a.u[q * N + col];
    x = make_double2(y.x - c * x.x, y.y - c * x.y);
    a.u[q * N + col] = x;
  }
}

// ---------------------------------------------------------------------------
// (0,0) flow-rate correction (meanflow_ctU -> flowrate_corr, LUsol0,
// LU_build0, LU_dec0, flowrateRe/Im: littleharsh.f90:411-484, 522-537;
// tridLU_3D_IB.f90:320-346, 493-519, 605-622).  One thread: ~ny rows.
// ---------------------------------------------------------------------------
__global__ void k_meanflow(double2* u1, size_t Ncol, int jlo_u, int nn, const double* d2u, const double* yv,
                           int jlo_v, double* scr, DevScalars* sc, double bRK, double dRK, double Re, double g0,
                           double g1) {
  const int n = nn + 3;
  double *a1 = scr, *a2 = scr + n, *a3 = scr + 2 * n, *x = scr + 3 * n;
  const double dt = sc->dt;
  const double bdt = -dRK * dt;
  const double beta = bRK * dt / Re;
  x[0] = 0.0;
  for (int j = 1; j <= nn + 1; ++j) x[j] = bdt;
  x[nn + 2] = 0.0;
  a1[0] = 0.0;
  a2[0] = 1.0;
  a3[0] = g0;
  for (int j = 1; j <= nn + 1; ++j) {
    const double* d = d2u + 3 * (j - jlo_u);
    a1[j] = -beta * d[0];
    a2[j] = 1.0 - beta * d[1];
    a3[j] = -beta * d[2];
  }
  a1[nn + 2] = g1;
  a2[nn + 2] = 1.0;
  a3[nn + 2] = 0.0;
  a2[0] = 1.0 / a2[0];
  for (int j = 1; j < n; ++j) {
    a3[j - 1] = a3[j - 1] * a2[j - 1];
    a2[j] = 1.0 / (a2[j] - a1[j] * a3[j - 1]);
  }
  x[0] = x[0] * a2[0];
  for (int j = 1; j < n; ++j) x[j] = (x[j] - a1[j] * x[j - 1]) * a2[j];
  for (int j = n - 2; j >= 0; --j) x[j] = x[j] - a3[j] * x[j + 1];
  double Qx1 = 0.0, Qx0 = 0.0;
  for (int j = 1; j <= nn + 1; ++j) Qx1 = Qx1 + (yv[j - jlo_v] - yv[j - 1 - jlo_v]) * x[j];
  for (int j = 1; j <= nn + 1; ++j) Qx0 = Qx0 + (yv[j - jlo_v] - yv[j - 1 - jlo_v]) * u1[(size_t)(j - jlo_u) * Ncol].x;
  const double g = (sc->QxT - Qx0) / Qx1;
  sc->mpgx = sc->mpgx + g;
  for (int j = 0; j <= nn + 2; ++j) u1[(size_t)(j - jlo_u) * Ncol].x = u1[(size_t)(j - jlo_u) * Ncol].x + g * x[j];
  sc->dgx = g;
}

// flowrateIm of the (0,0) u column (meanflow_ctP, littleharsh.f90:504-520)
__global__ void k_flowrate(const double2* u1, size_t Ncol, int jlo_u, int nn, const double* yv, int jlo_v,
                           DevScalars* sc) {
  double Q = 0.0;
  for (int j = 1; j <= nn + 1; ++j) Q = Q + (yv[j - jlo_v] - yv[j - 1 - jlo_v]) * u1[(size_t)(j - jlo_u) * Ncol].x;
  sc->Qx = Q;
}

// ---------------------------------------------------------------------------
// solveP (littleharsh.f90:556-611): div, psi = div dti/(a+b), (0,0) pin,
// LUsolP with the LU of LU_buildP/LU_decP (tridLU_3D_IB.f90:284-318,
// 417-491, 562-603) recomputed in the same operation order, p += psi.
// grid (ceil(Ncol/64)), one thread per column.
// ---------------------------------------------------------------------------
struct PArgs {
  const double2 *u1, *u2, *u3;
  double2 *p, *psi, *div;
  double* cs;
  const double *plo, *pup, *pdg, *dthdyu, *k1x, *k2x, *k1z, *k2z;
  const DevScalars* sc;
  double apb, ddthetavi, pin;
  int Mx, Ncol, n;
};

__global__ void __launch_bounds__(64) k_solve_p(PArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.Ncol) return;
  const int i = col % a.Mx, k = col / a.Mx;
  const double kx = a.k1x[i], kz = a.k1z[k], k2x = a.k2x[i], k2z = a.k2z[k];
  const double dtirk = a.sc->dti / a.apb;
  const size_t N = a.Ncol;
  const int n = a.n;
  double d = 0.0, up = 0.0;
  double2 x = make_double2(0, 0);
  for (int q = 0; q < n; ++q) {
    // divergence at p row q: u planes q+1, v planes q+1 and q (littleharsh.f90:209-244)
    const double2 u = a.u1[(q + 1) * N + col], w = a.u3[(q + 1) * N + col];
    const double2 vp = a.u2[(q + 1) * N + col], vm = a.u2[q * N + col];
    const double2 du = imulk(kx, u), dw = imulk(kz, w);
    const double m = a.ddthetavi * a.dthdyu[q + 1];
    const double2 dv = make_double2(du.x + dw.x + m * (vp.x - vm.x), du.y + dw.y + m * (vp.y - vm.y));
    a.div[q * N + col] = dv;
    double2 rhs = make_double2(dtirk * dv.x, dtirk * dv.y);
    double lo = a.plo[q], dg = a.pdg[q] + k2x + k2z, uq = a.pup[q];
    if (col == 0 && q == 0) {
      rhs = make_double2(0.0, 0.0);
      lo = 0.0;
      dg = a.pin;
      uq = 0.0;
    }
    if (q == 0) {
      d = 1.0 / dg;
      x = make_double2(rhs.x * d, rhs.y * d);
    } else {
      const double c = up * d;
      a.cs[(q - 1) * N + col] = c;
      d = 1.0 / (dg - lo * c);
      x = make_double2((rhs.x - lo * x.x) * d, (rhs.y - lo * x.y) * d);
    }
    a.psi[q * N + col] = x;
    up = uq;
  }
  {
    const double2 pp = a.p[(n - 1) * N + col];
    a.p[(n - 1) * N + col] = make_double2(pp.x + x.x, pp.y + x.y);
  }
  for (int q = n - 2; q >= 0; --q) {
    const double c = a.cs[q * N + col];
    const double2 y = a.psi[q * N + col];
    x = make_double2(y.x - c * x.x, y.y - c * x.y);
    a.psi[q * N + col] = x;
    const double2 pp = a.p[q * N + col];
    a.p[q * N + col] = make_double2(pp.x + x.x, pp.y + x.y);
  }
}

// ---------------------------------------------------------------------------
// v_corr (littleharsh.f90:796-907): projection and wall/ghost BCs.
// grid (ceil(Ncol/kTB), nyu)
// ---------------------------------------------------------------------------
struct VArgs {
  double2 *u1, *u2, *u3;
  const double2* psi;
  const double *dthdyv, *k1x, *k1z;
  const DevScalars* sc;
  double apb, ddthetavi, g0, g1;
  Geo g;
};

__global__ void __launch_bounds__(kTB) k_vcorr(VArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (col >= a.g.Ncol) return;
  const int i = col % a.g.Mx, k = col / a.g.Mx;
  const size_t N = a.g.Ncol, o = q * N + col;
  const double dtrk = a.sc->dt * a.apb;
  const int nyu = a.g.nyu, nyv = a.g.nyv;
  // u, w interior rows with the ghost rows next to the walls
  if (q >= 1 && q <= nyu - 2) {
    const double2 ps = a.psi[(q - 1) * N + col];
    const double2 gx = imulk(a.k1x[i] * dtrk, ps), gz = imulk(a.k1z[k] * dtrk, ps);
    const double2 u = a.u1[o], w = a.u3[o];
    const double2 un = make_double2(u.x - gx.x, u.y - gx.y), wn = make_double2(w.x - gz.x, w.y - gz.y);
    a.u1[o] = un;
    a.u3[o] = wn;
    if (q == 1) {
      a.u1[col] = make_double2(-a.g0 * un.x, -a.g0 * un.y);
      a.u3[col] = make_double2(-a.g0 * wn.x, -a.g0 * wn.y);
    }
    if (q == nyu - 2) {
      a.u1[(nyu - 1) * N + col] = make_double2(-a.g1 * un.x, -a.g1 * un.y);
      a.u3[(nyu - 1) * N + col] = make_double2(-a.g1 * wn.x, -a.g1 * wn.y);
    }
  }
  if (q < nyv) {
    if (q == 0 || q == nyv - 1) {
      a.u2[o] = make_double2(0.0, 0.0);
    } else {
      const double2 pp = a.psi[q * N + col], pm = a.psi[(q - 1) * N + col];
      const double c = dtrk * a.ddthetavi * a.dthdyv[q];
      const double2 v = a.u2[o];
      a.u2[o] = make_double2(v.x - c * (pp.x - pm.x), v.y - c * (pp.y - pm.y));
    }
  }
}

// divergence on the p grid (littleharsh.f90:209-244); grid (ceil(Ncol/kTB), nyp)
__global__ void __launch_bounds__(kTB) k_divergence(const double2* u1, const double2* u2, const double2* u3,
                                                    double2* div, const double* dthdyu, const double* k1x,
                                                    const double* k1z, double ddthetavi, Geo g) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (col >= g.Ncol) return;
  const int i = col % g.Mx, k = col / g.Mx;
  const size_t N = g.Ncol;
  const double2 u = u1[(q + 1) * N + col], w = u3[(q + 1) * N + col];
  const double2 vp = u2[(q + 1) * N + col], vm = u2[q * N + col];
  const double2 du = imulk(k1x[i], u), dw = imulk(k1z[k], w);
  const double m = ddthetavi * dthdyu[q + 1];
  div[q * N + col] = make_double2(du.x + dw.x + m * (vp.x - vm.x), du.y + dw.y + m * (vp.y - vm.y));
}

}  // namespace

void launch_rhs0(Solver& s, int kRK) {
  RhsArgs a;
  a.u1 = s.F[CFD_U1]; a.u2 = s.F[CFD_U2]; a.u3 = s.F[CFD_U3]; a.p = s.F[CFD_P];
  a.n1 = s.F[CFD_NU1]; a.n2 = s.F[CFD_NU2]; a.n3 = s.F[CFD_NU3];
  a.d1 = s.F[CFD_DU1]; a.d2 = s.F[CFD_DU2]; a.d3 = s.F[CFD_DU3];
  a.d2u = s.d_d2u; a.d2v = s.d_d2v; a.dthdyv = s.d_dthdyv;
  a.k1x = s.d_k1x; a.k2x = s.d_k2x; a.k1z = s.d_k1z; a.k2z = s.d_k2z;
  a.sc = s.d_sc;
  a.C1 = s.aRK[kRK - 1] / s.P.Re;
  a.C2 = -s.cRK[kRK - 1];
  a.dRK = s.dRK[kRK - 1];
  a.ddthetavi = 1.0;
  a.g = geo(s);
  dim3 grid((s.Ncol + kTB - 1) / kTB, s.nyu - 2, 3);
  hipLaunchKernelGGL(k_rhs0, grid, dim3(kTB), 0, s.stream, a);
}

void launch_nlcombine(Solver& s, int kRK) {
  CombArgs a;
  a.nxz1 = s.nxz1; a.nxz2 = s.nxz2; a.nxz3 = s.nxz3; a.fuv = s.fuv; a.fvw = s.fvw; a.fvv = s.fvv;
  a.n1 = s.F[CFD_NU1]; a.n2 = s.F[CFD_NU2]; a.n3 = s.F[CFD_NU3];
  a.d1 = s.F[CFD_DU1]; a.d2 = s.F[CFD_DU2]; a.d3 = s.F[CFD_DU3];
  a.dthdyu = s.d_dthdyu; a.dthdyv = s.d_dthdyv;
  a.C1 = -s.gRK[kRK - 1];
  a.ddthetavi = 1.0;
  a.g = geo(s);
  dim3 grid((s.Ncol + kTB - 1) / kTB, s.nyu - 2, 3);
  if (s.mixed)
    hipLaunchKernelGGL(k_nlcombine<float2>, grid, dim3(kTB), 0, s.stream, a);
  else
    hipLaunchKernelGGL(k_nlcombine<double2>, grid, dim3(kTB), 0, s.stream, a);
}

void launch_reset_max(Solver& s) { hipLaunchKernelGGL(k_reset_max, dim3(1), dim3(1), 0, s.stream, s.d_sc); }

void launch_cfl(Solver& s) {
  hipLaunchKernelGGL(k_cfl, dim3(1), dim3(1), 0, s.stream, s.d_sc, s.P.alp, s.P.bet, s.N1 / 2, s.N2 / 2, 1.0, s.dtv,
                     s.P.CFL);
}

void launch_predictor(Solver& s, int field) {
  const int ny = field == 1 ? s.nyv : s.nyu;
  double2* du = s.F[CFD_DU1 + field];
  const double2* u = s.F[CFD_U1 + field];
  dim3 grid((s.Ncol + kTB - 1) / kTB, ny);
  hipLaunchKernelGGL(k_predictor, grid, dim3(kTB), 0, s.stream, du, u, s.Ncol, ny, (const DevScalars*)s.d_sc);
}

void launch_ib_apply(Solver& s, int f0, int f1) {
  IBApplyArgs a{};
  int m = 0;
  for (int f = f0; f < f1; ++f) {
    const int g = f == 1 ? 1 : 0, y = f - f0;
    a.L[y] = s.d_ibl[g];
    a.seg[y] = s.d_ibseg[g];
    a.nseg[y] = s.nibseg[g];
    a.pr[y] = s.pib_r[f];
    a.pu[y] = (const double*)s.pib_u[f];
    m = std::max(m, s.nibseg[g]);
  }
  if (m == 0) return;
  const size_t psz = (size_t)s.ngz * s.ngx;
  hipLaunchKernelGGL(k_ib_apply, dim3((m + 127) / 128, f1 - f0), dim3(128), 0, s.stream, a, psz);
}

void launch_ysolve(Solver& s, int field, int kRK) {
  YArgs a;
  a.u = s.F[CFD_U1 + field];
  a.r = s.F[CFD_DU1 + field];
  a.cs = s.cscr + (size_t)field * s.nyu * s.Ncol;
  const bool v = field == 1;
  a.d2 = v ? s.d_d2v : s.d_d2u;
  a.k2x = s.d_k2x;
  a.k2z = s.d_k2z;
  a.sc = s.d_sc;
  a.bRK = s.bRK[kRK - 1];
  a.Re = s.P.Re;
  a.g0 = v ? 0.0 : s.gw[0];
  a.g1 = v ? 0.0 : s.gw[1];
  a.Mx = s.Mx;
  a.Ncol = s.Ncol;
  a.n = v ? s.nyv : s.nyu;
  hipLaunchKernelGGL(k_ysolve, dim3((s.Ncol + 63) / 64), dim3(64), 0, s.stream, a);
}

void launch_meanflow(Solver& s, int kRK) {
  hipLaunchKernelGGL(k_meanflow, dim3(1), dim3(1), 0, s.stream, s.F[CFD_U1], (size_t)s.Ncol, s.jlo_u, s.nn, s.d_d2u,
                     s.d_yv, s.jlo_v, s.d_mf, s.d_sc, s.bRK[kRK - 1], s.dRK[kRK - 1], s.P.Re, s.gw[0], s.gw[1]);
}

// max |a| over n complex values (the modulus, as Fortran's abs of a complex)
// into *out, as the bit pattern of a non-negative double; *out starts at 0
__global__ void k_absmax(const double2* __restrict__ a, size_t n, unsigned long long* out) {
  double m = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, hypot(a[e].x, a[e].y));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

void launch_absmax(Solver& s, const double2* a, size_t n) {
  unsigned long long* out = &s.d_sc->amax;
  (void)hipMemsetAsync(out, 0, sizeof *out, s.stream);
  hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, s.stream, a, n, out);
}

// u1, u3 (nyu planes) and u2 (nyv planes) times sigma[column]: the same factor
// for every plane and component of a column, so the discrete divergence of
// each mode scales with it and the boundary rows keep their form
__global__ void k_mode_filter(double2* __restrict__ u1, double2* __restrict__ u2, double2* __restrict__ u3,
                              const double* __restrict__ sigma, int ncol, int nyu, int nyv) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nyu * ncol) return;
  const double f = sigma[e % ncol];
  u1[e] = make_double2(f * u1[e].x, f * u1[e].y);
  u3[e] = make_double2(f * u3[e].x, f * u3[e].y);
  if (e < (size_t)nyv * ncol) u2[e] = make_double2(f * u2[e].x, f * u2[e].y);
}

void launch_mode_filter(Solver& s) {
  const size_t n = (size_t)s.nyu * s.Ncol;
  hipLaunchKernelGGL(k_mode_filter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, s.F[CFD_U1],
                     s.F[CFD_U2], s.F[CFD_U3], (const double*)s.d_sigma, s.Ncol, s.nyu, s.nyv);
}

// The x/z part of the nonlinear term of plane q replaced by that of plane
// map[q] in the stacks a and b (b may be null) of nq planes: the nonlinear
// term as ops_in_planes leaves it on each rank (FOU3D.f90:1553-1613, every
// plane of a rank gets the derivative sums of the rank's last plane).  Every
// source plane maps to itself, so no plane is read after it is written.
template <typename C>
__global__ void k_plane_map(C* __restrict__ a, C* __restrict__ b, const int* __restrict__ map, int nq, int ncol) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)nq * ncol) return;
  const int q = (int)(e / ncol);
  const int src = map[q];
  if (src == q) return;
  const size_t o = (size_t)src * ncol + (e - (size_t)q * ncol);
  a[e] = a[o];
  if (b) b[e] = b[o];
}

template <typename C>
void plane_map_t(Solver& s) {
  const size_t nu = (size_t)s.nyu * s.Ncol, nv = (size_t)s.nyv * s.Ncol;
  hipLaunchKernelGGL(k_plane_map<C>, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, s.stream, (C*)s.nxz1,
                     (C*)s.nxz3, (const int*)s.d_umap, s.nyu, s.Ncol);
  hipLaunchKernelGGL(k_plane_map<C>, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s.stream, (C*)s.nxz2,
                     (C*)nullptr, (const int*)s.d_vmap, s.nyv, s.Ncol);
}

void launch_plane_map(Solver& s) {
  if (!s.compat_np) return;
  if (s.mixed)
    plane_map_t<float2>(s);
  else
    plane_map_t<double2>(s);
}

void launch_flowrate(Solver& s) {
  hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(1), 0, s.stream, (const double2*)s.F[CFD_U1], (size_t)s.Ncol, s.jlo_u,
                     s.nn, (const double*)s.d_yv, s.jlo_v, s.d_sc);
}

void launch_solve_p(Solver& s, int kRK) {
  PArgs a;
  a.u1 = s.F[CFD_U1]; a.u2 = s.F[CFD_U2]; a.u3 = s.F[CFD_U3];
  a.p = s.F[CFD_P]; a.psi = s.F[CFD_PSI]; a.div = s.F[CFD_DIV];
  a.cs = s.cscr;
  a.plo = s.d_plo; a.pup = s.d_pup; a.pdg = s.d_pdg; a.dthdyu = s.d_dthdyu;
  a.k1x = s.d_k1x; a.k2x = s.d_k2x; a.k1z = s.d_k1z; a.k2z = s.d_k2z;
  a.sc = s.d_sc;
  a.apb = s.aRK[kRK - 1] + s.bRK[kRK - 1];
  a.ddthetavi = 1.0;
  const double h = s.yg.YU(2) - s.yg.YU(1);
  a.pin = 1.0 / (h * h);
  a.Mx = s.Mx;
  a.Ncol = s.Ncol;
  a.n = s.nyp;
  hipLaunchKernelGGL(k_solve_p, dim3((s.Ncol + 63) / 64), dim3(64), 0, s.stream, a);
}

void launch_vcorr(Solver& s, int kRK) {
  VArgs a;
  a.u1 = s.F[CFD_U1]; a.u2 = s.F[CFD_U2]; a.u3 = s.F[CFD_U3];
  a.psi = s.F[CFD_PSI];
  a.dthdyv = s.d_dthdyv; a.k1x = s.d_k1x; a.k1z = s.d_k1z;
  a.sc = s.d_sc;
  a.apb = s.aRK[kRK - 1] + s.bRK[kRK - 1];
  a.ddthetavi = 1.0;
  a.g0 = s.gw[0];
  a.g1 = s.gw[1];
  a.g = geo(s);
  dim3 grid((s.Ncol + kTB - 1) / kTB, s.nyu);
  hipLaunchKernelGGL(k_vcorr, grid, dim3(kTB), 0, s.stream, a);
}

void launch_divergence(Solver& s) {
  dim3 grid((s.Ncol + kTB - 1) / kTB, s.nyp);
  hipLaunchKernelGGL(k_divergence, grid, dim3(kTB), 0, s.stream, (const double2*)s.F[CFD_U1],
                     (const double2*)s.F[CFD_U2], (const double2*)s.F[CFD_U3], s.F[CFD_DIV],
                     (const double*)s.d_dthdyu, (const double*)s.d_k1x, (const double*)s.d_k1z, 1.0, geo(s));
}

}  // namespace cfd
