// kernels_band.hip -- multi-band resolution: moving the long columns between
// the solver (all columns, the wall bands' resolution) and the central
// solver (long columns only, the central band's resolution), and the time
// step from both resolutions' CFL maxima.
//
// The reference keeps one column array per band (x(sband:eband),
// declaration.f90:140-150): band 2 holds the long columns over every plane,
// bands 1 and 3 the short columns over their own planes, and the central
// planes are transformed on the coarser Ngal(1:2,2) grid (modes_to_planes
// with bandPL, FOU3D.f90:808-906).  Here the fields keep every column over
// every plane; the central solver sees the long columns of u1, u2, u3 and
// returns the x/z derivative sums and y fluxes of the central planes.
#include "common.hpp"

namespace cfd {

namespace {

constexpr int kTB = 256;

struct GatherArgs {
  const double2* in[4];  // u1, u2, u3, p
  double2* out[4];       // the same fields of the central solver
  const int* lcol;       // per central column: the column here
  int Ncol, Ncolc;
  int ny[4];             // planes of each field
  int q0;
};

// grid (ceil(Ncolc / kTB), q1 - q0, nf): field f, plane q0 + y, central column c
__global__ void k_band_gather(GatherArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = a.q0 + blockIdx.y, f = blockIdx.z;
  if (c >= a.Ncolc || q >= a.ny[f]) return;
  a.out[f][(size_t)q * a.Ncolc + c] = a.in[f][(size_t)q * a.Ncol + a.lcol[c]];
}

struct ScatterArgs {
  void* out[6];           // nxz1, nxz3, fvv (u planes), nxz2, fuv, fvw (v planes)
  const void* in[6];      // the same arrays of the central solver
  const int* ccol;        // per column: central column or -1
  int Ncol, Ncolc;
  int u0, nu, v0, nv;     // central u planes [u0, u0 + nu), v planes [v0, v0 + nv)
};

// grid (ceil(Ncol / kTB), max(nu, nv), 6): array a, central plane t; C the
// nonlinear-term element (double2 FP64, float2 on the mixed path)
template <typename C>
__global__ void k_band_scatter(ScatterArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y, w = blockIdx.z;
  const bool onu = w < 3;
  if (col >= a.Ncol || t >= (onu ? a.nu : a.nv)) return;
  const int q = (onu ? a.u0 : a.v0) + t;
  const int c = a.ccol[col];
  // a short column has no value on a central plane (its band arrays stop
  // at the band interface): its fluxes there are zero
  C v;
  v.x = 0;
  v.y = 0;
  if (c >= 0) v = static_cast<const C*>(a.in[w])[(size_t)q * a.Ncolc + c];
  static_cast<C*>(a.out[w])[(size_t)q * a.Ncol + col] = v;
}

// dtc_calc (FOU3D.f90:516-544) on every rank with its band's mode counts and
// the MIN over the ranks (FOU3D.f90:177-183): the wall bands' maxima with N1,
// N2, the central band's with N1c, N2c; the v limit has no mode count
__global__ void k_cfl_bands(DevScalars* sc, const DevScalars* cc, double alp, double bet, int N1h, int N2h, int N1ch,
                            int N2ch, double dtv, double CFL) {
  const double m1 = __longlong_as_double((long long)sc->m1);
  const double m3 = __longlong_as_double((long long)sc->m3);
  const double c1 = __longlong_as_double((long long)cc->m1);
  const double c3 = __longlong_as_double((long long)cc->m3);
  const double m2 = fmax(1e-6, fmax(__longlong_as_double((long long)sc->m2), __longlong_as_double((long long)cc->m2)));
  const double d1 = 1.0 / (alp * N1h * m1), d3 = 1.0 / (bet * N2h * m3), d2 = 1.0 / m2;
  double dtc = d1 < d3 ? d1 : d3;
  dtc = dtc < d2 ? dtc : d2;
  const double e1 = 1.0 / (alp * N1ch * c1), e3 = 1.0 / (bet * N2ch * c3);
  const double dc = e1 < e3 ? e1 : e3;
  dtc = dtc < dc ? dtc : dc;
  sc->dtc = dtc;
  sc->dt = (2.5 * dtv < CFL * dtc) ? 2.5 * dtv : CFL * dtc;
  sc->dti = 1.0 / sc->dt;
  sc->t = sc->t + sc->dt;
}

}  // namespace

void launch_band_gather(Solver& s, int q0, int q1, bool with_p) {
  Solver& c = *s.mbc;
  GatherArgs a;
  const int fid[4] = {CFD_U1, CFD_U2, CFD_U3, CFD_P};
  for (int f = 0; f < 4; ++f) {
    a.in[f] = s.F[fid[f]];
    a.out[f] = c.F[fid[f]];
  }
  a.lcol = s.d_lcol;
  a.Ncol = s.Ncol;
  a.Ncolc = c.Ncol;
  a.ny[0] = a.ny[2] = s.nyu;
  a.ny[1] = s.nyv;
  a.ny[3] = s.nyp;
  a.q0 = q0;
  const dim3 grid((c.Ncol + kTB - 1) / kTB, q1 - q0, with_p ? 4 : 3);
  hipLaunchKernelGGL(k_band_gather, grid, dim3(kTB), 0, s.stream, a);
}

void launch_band_gather(Solver& s) {
  // the planes the central solver's transforms read
  launch_band_gather(s, s.gather_q[0], s.gather_q[1], false);
}

void launch_band_scatter(Solver& s) {
  Solver& c = *s.mbc;
  ScatterArgs a;
  void* out[6] = {s.nxz1, s.nxz3, s.fvv, s.nxz2, s.fuv, s.fvw};
  void* in[6] = {c.nxz1, c.nxz3, c.fvv, c.nxz2, c.fuv, c.fvw};
  for (int w = 0; w < 6; ++w) {
    a.out[w] = out[w];
    a.in[w] = in[w];
  }
  a.ccol = s.d_ccol;
  a.Ncol = s.Ncol;
  a.Ncolc = c.Ncol;
  // central planes: u N(4,1)+1 .. N(4,2), v N(3,1)+1 .. N(3,2)
  a.u0 = s.N4[1] + 1 - s.jlo_u;
  a.nu = s.N4[2] - s.N4[1];
  a.v0 = s.N3[1] + 1 - s.jlo_v;
  a.nv = s.N3[2] - s.N3[1];
  const dim3 grid((s.Ncol + kTB - 1) / kTB, a.nu > a.nv ? a.nu : a.nv, 6);
  if (s.mixed)
    hipLaunchKernelGGL(k_band_scatter<float2>, grid, dim3(kTB), 0, s.stream, a);
  else
    hipLaunchKernelGGL(k_band_scatter<double2>, grid, dim3(kTB), 0, s.stream, a);
}

void launch_cfl_bands(Solver& s) {
  hipLaunchKernelGGL(k_cfl_bands, dim3(1), dim3(1), 0, s.stream, s.d_sc, (const DevScalars*)s.mbc->d_sc, s.P.alp,
                     s.P.bet, s.N1 / 2, s.N2 / 2, s.N1c / 2, s.N2c / 2, s.dtv, s.P.CFL);
}

}  // namespace cfd
