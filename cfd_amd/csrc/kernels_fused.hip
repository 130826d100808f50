// kernels_fused.hip -- column-sweep kernels of the fused substep.
//
// One thread owns one (kx,kz) column and walks it in y.  Only Ncol threads
// exist per field (16 512 at 384x192x192, about three wavefronts per CU for
// the three velocity components), so each sweep loads its inputs CH rows at a
// time into registers before using them: the loads of a chunk are in flight
// together, which keeps enough bytes outstanding to stream from HBM.
//
//  k_ib_rhs     predictor r = u + dt du on the IB planes only, du built from
//               RHS0 (littleharsh.f90:613-730) and the nonlinear assembly
//               (FOU3D.f90:158-173, 232-244); the IB forcing then corrects it.
//  k_rhs_solve  the same du and r on every plane (the IB-corrected r taken
//               where the IB acts), Nu stored for the next substep, and the
//               Helmholtz solve (LU_buildU/V + LU_dec + LUsolU/V,
//               tridLU_3D_IB.f90:50-160, 349-413, 521-559) in the same sweep.
//  k_pressure   solveP (littleharsh.f90:556-611, LUsolP tridLU_3D_IB.f90:284-318)
//               and v_corr (littleharsh.f90:796-907) with the divergence of the
//               corrected field, in one forward/backward sweep.
#include <cstring>
#include <algorithm>

#include "common.hpp"

namespace cfd {

namespace {

constexpr int CH = 8;

struct FusedArgs {
  const double2* u[3];                  // u1, u2, u3 (old)
  double2* __restrict__ nu[3];          // Nu1..3: old in, new out
  const void* nxz[3];   // x/z part of the new nonlinear term (TN: double2 or float2)
  const void* flux[3];  // uv (v planes), vv (u planes), vw (v planes) (TN)
  const double2* __restrict__ p;
  double2* __restrict__ r[3];           // du1..3: IB-corrected r on IB planes
  double2* __restrict__ ysc[3];         // forward-sweep scratch
  double* __restrict__ cs[3];           // modified upper diagonals
  double2* uo[3];                       // new velocity (same arrays as u)
  const unsigned char* __restrict__ ibmask[3];  // 1 on IB planes
  const double *d2u, *d2v, *dthdyu, *dthdyv, *k1x, *k2x, *k1z, *k2z;
  const DevScalars* sc;
  double C1, C2, C1g, dRK, bRK, Re, ddthetavi, g0, g1;
  int Mx, Ncol, nyu, nyv;
  int ib_only;  // 1: k_ib_rhs mode
  // multi-band resolution: per column 1 = short (null: none); a short
  // column's rows t1..b3 of field f (plane indices of its grid) are its
  // band-1 top boundary row, rows between its bands and its band-3 bottom
  // boundary row; gw12 / gw31 their u-grid weights (start.f90:494-500)
  const unsigned char* shortc;
  int t1[3], b3[3];
  double gw12, gw31;
};

// pointers of one field, selected without indexing the kernel-argument
// arrays by a runtime value (that would place the arguments in scratch)
struct FieldPtrs {
  const double2* u;
  double2* nu;
  const void* nxz;
  const void* flux;
  double2* r;
  double2* ys;
  double* cs;
  double2* uo;
  const unsigned char* ib;
};

template <typename T>
__device__ __forceinline__ T pick(int f, T a0, T a1, T a2) {
  return f == 0 ? a0 : (f == 1 ? a1 : a2);
}

__device__ __forceinline__ FieldPtrs field_ptrs(const FusedArgs& a, int f) {
  FieldPtrs p;
  p.u = pick(f, a.u[0], a.u[1], a.u[2]);
  p.nu = pick(f, a.nu[0], a.nu[1], a.nu[2]);
  p.nxz = pick(f, a.nxz[0], a.nxz[1], a.nxz[2]);
  p.flux = pick(f, a.flux[0], a.flux[1], a.flux[2]);
  p.r = pick(f, a.r[0], a.r[1], a.r[2]);
  p.ys = pick(f, a.ysc[0], a.ysc[1], a.ysc[2]);
  p.cs = pick(f, a.cs[0], a.cs[1], a.cs[2]);
  p.uo = pick(f, a.uo[0], a.uo[1], a.uo[2]);
  p.ib = pick(f, a.ibmask[0], a.ibmask[1], a.ibmask[2]);
  return p;
}

// du and r of row q (interior) of field f for column col, from the loaded
// values; returns r, stores the new Nu through nnew
// d: the three d2 coefficients of row q, dd: dtheta/dy of row q (d2v and
// dthdyv for f = 1, d2u and dthdyu otherwise)
__device__ __forceinline__ double2 row_r(const FusedArgs& a, int f, int col, double k2x, double k2z, double kk,
                                         double dt, const double* d, double dd, double2 um, double2 u0, double2 up,
                                         double2 nold, double2 nx, double2 fa, double2 fb, double2 pa, double2 pb,
                                         double2& nnew) {
  struct {
    double2 um, u0, up, nold, nx, fa, fb, pa, pb;
  } v = {um, u0, up, nold, nx, fa, fb, pa, pb};
  const double c2 = k2x + k2z + d[1];
  const double2 L = make_double2(d[0] * v.um.x + c2 * v.u0.x + d[2] * v.up.x, d[0] * v.um.y + c2 * v.u0.y + d[2] * v.up.y);
  double2 du;
  if (f == 1) {
    const double C3 = -a.dRK * a.ddthetavi * dd;
    du = make_double2(a.C1 * L.x + a.C2 * v.nold.x + C3 * (v.pa.x - v.pb.x),
                      a.C1 * L.y + a.C2 * v.nold.y + C3 * (v.pa.y - v.pb.y));
  } else {
    const double2 gp = imulk(-a.dRK * kk, v.pa);
    du = make_double2(a.C1 * L.x + a.C2 * v.nold.x + gp.x, a.C1 * L.y + a.C2 * v.nold.y + gp.y);
    if (f == 0 && col == 0) du.x += -a.dRK * a.sc->mpgx;
  }
  nnew = make_double2(v.nx.x + (v.fa.x - v.fb.x) * dd * a.ddthetavi, v.nx.y + (v.fa.y - v.fb.y) * dd * a.ddthetavi);
  du = make_double2(du.x + a.C1g * nnew.x, du.y + a.C1g * nnew.y);
  return make_double2(v.u0.x + dt * du.x, v.u0.y + dt * du.y);
}

__device__ __forceinline__ double2 as_d2(double2 v) { return v; }
__device__ __forceinline__ double2 as_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }

// load the inputs of row q (interior) of field f; TN is the storage type of
// the nonlinear scratch (nxz, fluxes)
template <typename TN>
__device__ __forceinline__ void row_load(const FusedArgs& a, const FieldPtrs& F, int f, int q, size_t N, int col,
                                         double2& up, double2& nold, double2& nx, double2& fa, double2& fb,
                                         double2& pa, double2& pb) {
  up = F.u[(size_t)(q + 1) * N + col];
  // the old Nu enters with weight C2 = -cRK, zero at kRK = 1: its read is
  // skipped there (C2 * 0 is a zero, as C2 * Nu is)
  nold = a.C2 != 0.0 ? F.nu[(size_t)q * N + col] : make_double2(0.0, 0.0);
  const TN* nxz = (const TN*)F.nxz;
  const TN* flux = (const TN*)F.flux;
  nx = as_d2(nxz[(size_t)q * N + col]);
  if (f == 1) {
    fa = as_d2(flux[(size_t)(q + 1) * N + col]);
    fb = as_d2(flux[(size_t)q * N + col]);
    pa = a.p[(size_t)q * N + col];
    pb = a.p[(size_t)(q - 1) * N + col];
  } else {
    fa = as_d2(flux[(size_t)q * N + col]);
    fb = as_d2(flux[(size_t)(q - 1) * N + col]);
    pa = a.p[(size_t)(q - 1) * N + col];
    pb = pa;
  }
}

// grid (ceil(Ncol/64), 3)
template <typename TN>
__global__ void __launch_bounds__(64) k_rhs_solve(FusedArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (col >= a.Ncol) return;
  const int i = col % a.Mx, k = col / a.Mx;
  const double k2x = a.k2x[i], k2z = a.k2z[k];
  const double kk = (f == 0) ? a.k1x[i] : a.k1z[k];
  const size_t N = a.Ncol;
  const int n = (f == 1) ? a.nyv : a.nyu;
  const double dt = a.sc->dt;
  const double beta = a.bRK * dt / a.Re;
  const double* d2 = (f == 1) ? a.d2v : a.d2u;
  const double g0 = (f == 1) ? 0.0 : a.g0, g1 = (f == 1) ? 0.0 : a.g1;
  const FieldPtrs F = field_ptrs(a, f);
  const double2* u = F.u;
  double2* __restrict__ ys = F.ys;
  double* __restrict__ cs = F.cs;
  const unsigned char* __restrict__ ib = F.ib;
  // row 0: BC row (1, g0), r = 0
  double d = 1.0, a3p = g0;
  double2 x = make_double2(0.0, 0.0);
  ys[col] = x;
  double2 um = u[col], u0 = u[N + col];
  for (int q0 = 1; q0 < n - 1; q0 += CH) {
    double2 UP[CH], NO[CH], NX[CH], FA[CH], FB[CH], PA[CH], PB[CH], RB[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 + c;
      if (q < n - 1) {
        row_load<TN>(a, F, f, q, N, col, UP[c], NO[c], NX[c], FA[c], FB[c], PA[c], PB[c]);
        RB[c] = ib[q] ? F.r[(size_t)q * N + col] : make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 + c;
      if (q >= n - 1) break;
      double2 nnew;
      double2 rq = row_r(a, f, col, k2x, k2z, kk, dt, d2 + 3 * q, f == 1 ? a.dthdyv[q] : a.dthdyu[q], um, u0, UP[c],
                         NO[c], NX[c], FA[c], FB[c], PA[c], PB[c], nnew);
      F.nu[(size_t)q * N + col] = nnew;
      if (ib[q]) rq = RB[c];
      um = u0;
      u0 = UP[c];
      // Thomas forward step (reciprocal pivots, LU_dec order)
      const double* dq = d2 + 3 * q;
      const double a1 = -beta * dq[0];
      const double a2 = 1.0 - beta * (dq[1] + k2x + k2z);
      const double c_ = a3p * d;
      cs[(size_t)(q - 1) * N + col] = c_;
      d = 1.0 / (a2 - a1 * c_);
      x = make_double2((rq.x - a1 * x.x) * d, (rq.y - a1 * x.y) * d);
      ys[(size_t)q * N + col] = x;
      a3p = -beta * dq[2];
    }
  }
  // last row: BC row (g1, 1), r = 0
  {
    const int q = n - 1;
    const double c_ = a3p * d;
    cs[(size_t)(q - 1) * N + col] = c_;
    d = 1.0 / (1.0 - g1 * c_);
    x = make_double2((0.0 - g1 * x.x) * d, (0.0 - g1 * x.y) * d);
  }
  // back substitution
  double2* uo = F.uo;
  uo[(size_t)(n - 1) * N + col] = x;
  for (int q0 = n - 2; q0 >= 0; q0 -= CH) {
    double2 yv[CH];
    double cv[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 - c;
      if (q >= 0) {
        yv[c] = ys[(size_t)q * N + col];
        cv[c] = cs[(size_t)q * N + col];
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 - c;
      if (q < 0) break;
      x = make_double2(yv[c].x - cv[c] * x.x, yv[c].y - cv[c] * x.y);
      uo[(size_t)q * N + col] = x;
    }
  }
}

// ---------------------------------------------------------------------------
// Segmented column solve.  The same rows and Helmholtz system as k_rhs_solve,
// solved by substructuring so that a column is shared by S threads: rows are
// cut into segments of SEG rows, each followed by one separator row.  Thread
// (column, segment) builds r for its rows and eliminates its segment in
// registers, giving x_q = Y_q + G_q x_R + H_q x_L in terms of the separator
// values either side; each separator row then couples only its two
// neighbouring separators, a tridiagonal system of S-1 unknowns per column
// solved through LDS, after which every thread finishes its own rows.  No
// factor scratch goes to HBM and a column's 200-odd rows are walked by ~15
// threads instead of one.  Block (CB columns, SY >= S segments), grid
// (ceil(Ncol/CB), 3); rows past the end of the column are identity rows.
// ---------------------------------------------------------------------------
// one row of the Helmholtz system of field f: lower, diagonal and upper
// coefficients and right-hand side; rows past the column are identity rows
struct RowSys {
  double lo, dg, up;
  double2 r;
};

struct RowCtx {
  double k2x, k2z, kk, dt, beta, g0, g1;
  const double* md;          // LDS: per row the three d2 coefficients and dtheta/dy
  const unsigned char* ibl;  // LDS: 1 on the IB rows
  int n, col;
  size_t N;
  // a short column: rows t1..b3 are its band ends and the rows between
  // (t1 > n: a long column); gt1 / gb3 the band-end rows' coupling weights
  int t1, b3;
  double gt1, gb3;
};

// The same row in two halves, so that the loads of row q+1 can be issued
// before the arithmetic of row q: seg_fetch loads what row q reads from HBM
// (u(q+1) and, for interior rows, the old Nu, N_xz, the y fluxes and p),
// seg_build forms the row from them.
struct RowData {
  double2 un, nold, nx, fa, fb, pa, pb;
};

template <typename TN>
__device__ __forceinline__ RowData seg_fetch(const FusedArgs& a, const FieldPtrs& F, int f, const RowCtx& c, int q) {
  RowData d;
  d.un = make_double2(0.0, 0.0);
  if (q >= c.n) return d;
  if (q + 1 < c.n) d.un = F.u[(size_t)(q + 1) * c.N + c.col];
  if (q > 0 && q < c.n - 1) {
    double2 up;
    row_load<TN>(a, F, f, q, c.N, c.col, up, d.nold, d.nx, d.fa, d.fb, d.pa, d.pb);
  }
  return d;
}

__device__ __forceinline__ RowSys seg_build(const FusedArgs& a, const FieldPtrs& F, int f, const RowCtx& c, int q,
                                            const RowData& d, double2& um, double2& u0) {
  RowSys s;
  if (q >= c.n) {
    s.lo = 0.0; s.dg = 1.0; s.up = 0.0; s.r = make_double2(0.0, 0.0);
    return s;
  }
  if (q >= c.t1 && q <= c.b3) {
    // a short column: its band-1 top boundary row [gw(1,2), 1, 0], the rows
    // between its bands (identity, zero) and its band-3 bottom boundary row
    // [0, 1, gw(3,1)] (LU_buildU/V per band, tridLU_3D_IB.f90:349-413)
    s.lo = q == c.t1 ? c.gt1 : 0.0;
    s.dg = 1.0;
    s.up = q == c.b3 ? c.gb3 : 0.0;
    s.r = make_double2(0.0, 0.0);
  } else if (q == 0) {
    s.lo = 0.0; s.dg = 1.0; s.up = c.g0; s.r = make_double2(0.0, 0.0);
  } else if (q == c.n - 1) {
    s.lo = c.g1; s.dg = 1.0; s.up = 0.0; s.r = make_double2(0.0, 0.0);
  } else {
    double2 nnew;
    const double* dq = c.md + 4 * q;
    s.r = row_r(a, f, c.col, c.k2x, c.k2z, c.kk, c.dt, dq, dq[3], um, u0, d.un, d.nold, d.nx, d.fa, d.fb, d.pa, d.pb,
                nnew);
    F.nu[(size_t)q * c.N + c.col] = nnew;
    if (c.ibl[q]) s.r = F.r[(size_t)q * c.N + c.col];
    s.lo = -c.beta * dq[0];
    s.dg = 1.0 - c.beta * (dq[1] + c.k2x + c.k2z);
    s.up = -c.beta * dq[2];
  }
  um = u0;
  u0 = d.un;
  return s;
}

// Blocks are dealt to the 8 XCDs round-robin by block id; block b takes
// column tile 8 (slot / 3) + b % 8 and field slot % 3, slot = b / 8, so the
// three components of one tile run on one XCD at about the same time and the
// pressure rows they all read come from that XCD's L2 after the first.
// grid 24 ceil(ntiles / 8), ntiles = ceil(Ncol / CB)
// Block bx of the grid; k_rhs_seg runs it for one solver, k_rhs_seg_b for
// the members of a batch (blockDim.y may exceed a member's segment count).
template <int SEG, typename TN>
__device__ __forceinline__ void rhs_seg_body(const FusedArgs& a, int bx) {
  extern __shared__ double sm[];
  const int CB = blockDim.x, NT = blockDim.x * blockDim.y;
  const int cl = threadIdx.x, sg = threadIdx.y, me = sg * CB + cl;
  const int slot = bx >> 3;
  const int tile = (slot / 3) * 8 + (bx & 7), f = slot % 3;
  if (tile * CB >= a.Ncol) return;  // the whole block: no barrier is reached
  const int col = tile * CB + cl;
  const int n = (f == 1) ? a.nyv : a.nyu;
  const int S = (n + SEG + 1) / (SEG + 1);
  double2* Yb = reinterpret_cast<double2*>(sm);
  double2* Ye = Yb + NT;
  double2* Rs = Ye + NT;
  double2* Xs = Rs + NT;
  double* Gb = reinterpret_cast<double*>(Xs + NT);
  double* Hb = Gb + NT;
  double* Ge = Hb + NT;
  double* He = Ge + NT;
  double* As = He + NT;
  double* Ds = As + NT;
  double* Cs = Ds + NT;
  // per-row metadata of field f staged in LDS: the row loop then waits on
  // the field loads only
  double* md = Cs + NT;
  unsigned char* ibl = reinterpret_cast<unsigned char*>(md + 4 * n);
  const bool act = col < a.Ncol && sg < S;
  const size_t N = a.Ncol;
  const FieldPtrs F = field_ptrs(a, f);
  {
    const double* d2 = (f == 1) ? a.d2v : a.d2u;
    const double* dth = (f == 1) ? a.dthdyv : a.dthdyu;
    for (int q = me; q < n; q += NT) {
      md[4 * q] = d2[3 * q];
      md[4 * q + 1] = d2[3 * q + 1];
      md[4 * q + 2] = d2[3 * q + 2];
      md[4 * q + 3] = dth[q];
      ibl[q] = F.ib[q];
    }
  }
  __syncthreads();
  const double2* u = F.u;
  const int b = sg * (SEG + 1);
  double2 y[SEG];
  double g[SEG], h[SEG];
  if (act) {
    const int i = col % a.Mx, k = col / a.Mx;
    RowCtx c;
    c.k2x = a.k2x[i];
    c.k2z = a.k2z[k];
    c.kk = (f == 0) ? a.k1x[i] : a.k1z[k];
    c.dt = a.sc->dt;
    c.beta = a.bRK * c.dt / a.Re;
    c.md = md;
    c.ibl = ibl;
    c.g0 = (f == 1) ? 0.0 : a.g0;
    c.g1 = (f == 1) ? 0.0 : a.g1;
    c.n = n;
    c.col = col;
    c.N = N;
    const bool shc = a.shortc && a.shortc[col];
    c.t1 = shc ? (f == 0 ? a.t1[0] : (f == 1 ? a.t1[1] : a.t1[2])) : n + 1;
    c.b3 = shc ? (f == 0 ? a.b3[0] : (f == 1 ? a.b3[1] : a.b3[2])) : n + 1;
    c.gt1 = (f == 1) ? 0.0 : a.gw12;
    c.gb3 = (f == 1) ? 0.0 : a.gw31;
    double2 um = b >= 1 ? u[(size_t)(b - 1) * N + col] : make_double2(0.0, 0.0);
    double2 u0 = u[(size_t)b * N + col];
    // forward elimination inside the segment
    double2 yp = make_double2(0.0, 0.0);
    double gp = 0.0, hp = 1.0;
    RowData cur = seg_fetch<TN>(a, F, f, c, b);
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
      // rows b .. b+SEG-1, then the separator b+SEG (fetched for every
      // segment: past the column it loads nothing)
      const RowData nxt = seg_fetch<TN>(a, F, f, c, b + j + 1);
      const RowSys rs = seg_build(a, F, f, c, b + j, cur, um, u0);
      cur = nxt;
      const double inv = 1.0 / (rs.dg - rs.lo * gp);
      y[j] = make_double2((rs.r.x - rs.lo * yp.x) * inv, (rs.r.y - rs.lo * yp.y) * inv);
      g[j] = rs.up * inv;
      h[j] = -rs.lo * hp * inv;
      yp = y[j];
      gp = g[j];
      hp = h[j];
    }
    // separator row after the segment
    if (sg < S - 1) {
      const RowSys rs = seg_build(a, F, f, c, b + SEG, cur, um, u0);
      As[me] = rs.lo;
      Ds[me] = rs.dg;
      Cs[me] = rs.up;
      Rs[me] = rs.r;
    }
    // back substitution inside the segment: x_q = Y_q + G_q x_R + H_q x_L
    g[SEG - 1] = -g[SEG - 1];
#pragma unroll
    for (int j = SEG - 2; j >= 0; --j) {
      const double gj = g[j];
      y[j] = make_double2(y[j].x - gj * y[j + 1].x, y[j].y - gj * y[j + 1].y);
      h[j] = h[j] - gj * h[j + 1];
      g[j] = -gj * g[j + 1];
    }
    Yb[me] = y[0];
    Gb[me] = g[0];
    Hb[me] = h[0];
    Ye[me] = y[SEG - 1];
    Ge[me] = g[SEG - 1];
    He[me] = h[SEG - 1];
  }
  __syncthreads();
  // separators: tridiagonal in x_s, s = 0..S-2, one thread per column
  if (sg == 0 && col < a.Ncol && S > 1) {
    double cp = 0.0;
    double2 x = make_double2(0.0, 0.0);
    for (int s = 0; s < S - 1; ++s) {
      const int A = s * CB + cl, B = (s + 1) * CB + cl;
      const double as = As[A], cs = Cs[A];
      const double lo = as * He[A];
      const double dg = Ds[A] + as * Ge[A] + cs * Hb[B];
      const double2 rh = make_double2(Rs[A].x - as * Ye[A].x - cs * Yb[B].x, Rs[A].y - as * Ye[A].y - cs * Yb[B].y);
      const double den = 1.0 / (dg - lo * cp);
      x = make_double2((rh.x - lo * x.x) * den, (rh.y - lo * x.y) * den);
      cp = cs * Gb[B] * den;
      Xs[A] = x;
      Cs[A] = cp;  // reused as the modified upper coefficient
    }
    for (int s = S - 3; s >= 0; --s) {
      const int A = s * CB + cl;
      const double2 xn = Xs[A + CB];
      x = make_double2(Xs[A].x - Cs[A] * xn.x, Xs[A].y - Cs[A] * xn.y);
      Xs[A] = x;
    }
  }
  __syncthreads();
  if (act) {
    const double2 xl = sg > 0 ? Xs[me - CB] : make_double2(0.0, 0.0);
    const double2 xr = sg < S - 1 ? Xs[me] : make_double2(0.0, 0.0);
    double2* uo = F.uo;
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
      const int q = b + j;
      if (q < n)
        uo[(size_t)q * N + col] =
            make_double2(y[j].x + g[j] * xr.x + h[j] * xl.x, y[j].y + g[j] * xr.y + h[j] * xl.y);
    }
    if (sg < S - 1) uo[(size_t)(b + SEG) * N + col] = xr;
  }
}

template <int SEG, typename TN>
__global__ void __launch_bounds__(256) k_rhs_seg(FusedArgs a) {
  rhs_seg_body<SEG, TN>(a, blockIdx.x);
}

template <int SEG, typename TN>
__global__ void __launch_bounds__(256) k_rhs_seg_b(const FusedArgs* __restrict__ as, BMap bm) {
  int bx = 0, by = 0;
  const int m = bmap_member(bm, blockIdx.x, bx, by);
  if (m < 0) return;
  const FusedArgs a = as[m];  // a private copy, as the single-solver kernel's argument
  rhs_seg_body<SEG, TN>(a, bx);
}

// predictor on the IB planes only: grid (ceil(Ncol/kTB), max IB planes, 3)
template <typename TN>
__global__ void __launch_bounds__(256) k_ib_rhs(FusedArgs a, const int* __restrict__ planes0,
                                                const int* __restrict__ planes1, const int* __restrict__ planes2,
                                                int n0, int n1, int n2) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.z;
  const int b = blockIdx.y;
  const int nb = f == 0 ? n0 : f == 1 ? n1 : n2;
  if (col >= a.Ncol || b >= nb) return;
  const int q = (f == 0 ? planes0 : f == 1 ? planes1 : planes2)[b];
  const int i = col % a.Mx, k = col / a.Mx;
  const double kk = (f == 0) ? a.k1x[i] : a.k1z[k];
  const size_t N = a.Ncol;
  const FieldPtrs F = field_ptrs(a, f);
  const double2 um = F.u[(size_t)(q - 1) * N + col];
  const double2 u0 = F.u[(size_t)q * N + col];
  double2 up, nold, nx, fa, fb, pa, pb;
  row_load<TN>(a, F, f, q, N, col, up, nold, nx, fa, fb, pa, pb);
  double2 nnew;
  F.r[(size_t)q * N + col] =
      row_r(a, f, col, a.k2x[i], a.k2z[k], kk, a.sc->dt, (f == 1 ? a.d2v : a.d2u) + 3 * q,
            f == 1 ? a.dthdyv[q] : a.dthdyu[q], um, u0, up, nold, nx, fa, fb, pa, pb, nnew);
}

// ---------------------------------------------------------------------------
// pressure: divergence, psi, LU sweep, p += psi, projection, ghosts, walls,
// divergence of the corrected field.  grid (ceil(Ncol/64))
// ---------------------------------------------------------------------------
struct PressArgs {
  double2* __restrict__ u1;
  double2* __restrict__ u2;
  double2* __restrict__ u3;
  double2* __restrict__ p;
  double2* __restrict__ psi;
  double2* __restrict__ div;
  double* __restrict__ cs;
  const double *plo, *pup, *pdg, *dthdyu, *dthdyv, *k1x, *k2x, *k1z, *k2z;
  const DevScalars* sc;
  double apb, ddthetavi, pin, g0, g1;
  int Mx, Ncol, nyp;
  // multi-band resolution: per column 1 = short (null: none); a short
  // column's band 1 ends at p row P1 and its band 3 starts at p row P3
  const unsigned char* shortc;
  int P1, P3;
};

__global__ void __launch_bounds__(64) k_pressure(PressArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.Ncol) return;
  const int i = col % a.Mx, k = col / a.Mx;
  const double kx = a.k1x[i], kz = a.k1z[k], k2x = a.k2x[i], k2z = a.k2z[k];
  const double dtirk = a.sc->dti / a.apb;
  const double dtrk = a.sc->dt * a.apb;
  const size_t N = a.Ncol;
  const int n = a.nyp;
  double d = 0.0, up = 0.0;
  double2 x = make_double2(0, 0);
  double2 vm = a.u2[col];  // v at the wall face below p row 0
  for (int q0 = 0; q0 < n; q0 += CH) {
    double2 U[CH], W[CH], V[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 + c;
      if (q < n) {
        U[c] = a.u1[(size_t)(q + 1) * N + col];
        W[c] = a.u3[(size_t)(q + 1) * N + col];
        V[c] = a.u2[(size_t)(q + 1) * N + col];
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 + c;
      if (q >= n) break;
      const double2 du = imulk(kx, U[c]), dw = imulk(kz, W[c]);
      const double m = a.ddthetavi * a.dthdyu[q + 1];
      const double2 dv = make_double2(du.x + dw.x + m * (V[c].x - vm.x), du.y + dw.y + m * (V[c].y - vm.y));
      vm = V[c];
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
        const double c_ = up * d;
        a.cs[(size_t)(q - 1) * N + col] = c_;
        d = 1.0 / (dg - lo * c_);
        x = make_double2((rhs.x - lo * x.x) * d, (rhs.y - lo * x.y) * d);
      }
      a.psi[(size_t)q * N + col] = x;
      up = uq;
    }
  }
  // backward: psi, p, projection of the u/w row above each p row and of the
  // v face between p rows q and q+1; divergence of the corrected field one
  // row behind
  const double kxd = kx * dtrk, kzd = kz * dtrk;
  const int nyu = n + 2, nyv = n + 1;
  double2 pn = x;                            // psi(q+1)
  double2 un1 = make_double2(0, 0), wn1 = un1;  // corrected u, w of p row q+1
  double2 vn1 = make_double2(0.0, 0.0);      // corrected v at face q+2 (top wall: 0)
  a.u2[(size_t)(nyv - 1) * N + col] = vn1;
  for (int q0 = n - 1; q0 >= 0; q0 -= CH) {
    double2 Y[CH], P[CH], U[CH], W[CH], V[CH];
    double C[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 - c;
      if (q >= 0) {
        Y[c] = a.psi[(size_t)q * N + col];
        C[c] = (q < n - 1) ? a.cs[(size_t)q * N + col] : 0.0;
        P[c] = a.p[(size_t)q * N + col];
        U[c] = a.u1[(size_t)(q + 1) * N + col];
        W[c] = a.u3[(size_t)(q + 1) * N + col];
        V[c] = a.u2[(size_t)(q + 1) * N + col];
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int q = q0 - c;
      if (q < 0) break;
      double2 ps;
      if (q == n - 1) ps = Y[c];
      else ps = make_double2(Y[c].x - C[c] * pn.x, Y[c].y - C[c] * pn.y);
      a.psi[(size_t)q * N + col] = ps;
      a.p[(size_t)q * N + col] = make_double2(P[c].x + ps.x, P[c].y + ps.y);
      const double2 gx = imulk(kxd, ps), gz = imulk(kzd, ps);
      const double2 un = make_double2(U[c].x - gx.x, U[c].y - gx.y);
      const double2 wn = make_double2(W[c].x - gz.x, W[c].y - gz.y);
      a.u1[(size_t)(q + 1) * N + col] = un;
      a.u3[(size_t)(q + 1) * N + col] = wn;
      if (q == n - 1) {
        a.u1[(size_t)(nyu - 1) * N + col] = make_double2(-a.g1 * un.x, -a.g1 * un.y);
        a.u3[(size_t)(nyu - 1) * N + col] = make_double2(-a.g1 * wn.x, -a.g1 * wn.y);
      }
      // v face q+1 lies between p rows q and q+1
      double2 vn;
      if (q == n - 1) {
        vn = make_double2(0.0, 0.0);  // top wall face, set above
      } else {
        const double cc = dtrk * a.ddthetavi * a.dthdyv[q + 1];
        vn = make_double2(V[c].x - cc * (pn.x - ps.x), V[c].y - cc * (pn.y - ps.y));
        a.u2[(size_t)(q + 1) * N + col] = vn;
        // divergence of the corrected field at p row q+1
        const double2 du = imulk(kx, un1), dw = imulk(kz, wn1);
        const double m = a.ddthetavi * a.dthdyu[q + 2];
        a.div[(size_t)(q + 1) * N + col] =
            make_double2(du.x + dw.x + m * (vn1.x - vn.x), du.y + dw.y + m * (vn1.y - vn.y));
      }
      un1 = un;
      wn1 = wn;
      vn1 = vn;
      pn = ps;
    }
  }
  // bottom: wall face, ghosts, divergence at p row 0
  a.u2[col] = make_double2(0.0, 0.0);
  a.u1[col] = make_double2(-a.g0 * un1.x, -a.g0 * un1.y);
  a.u3[col] = make_double2(-a.g0 * wn1.x, -a.g0 * wn1.y);
  {
    const double2 du = imulk(kx, un1), dw = imulk(kz, wn1);
    const double m = a.ddthetavi * a.dthdyu[1];
    a.div[col] = make_double2(du.x + dw.x + m * (vn1.x - 0.0), du.y + dw.y + m * (vn1.y - 0.0));
  }
}

// Segmented pressure sweep: the system, outputs and order of k_pressure,
// solved by substructuring as in k_rhs_seg (SEG p rows per thread plus one
// separator row, separators through LDS).  Thread (column, segment) owns p
// rows b..last, the u/w rows above them and the v faces b..last; the face
// above its last row (the next segment's first face) is recomputed from that
// segment's LDS record to close the divergence of its last row, with the old
// v of that face read before the first barrier.
// CU: the first pass's u1 rows are kept for the projection too (the v faces
// always are); a register budget matter, chosen per segment length
template <int SEG, bool CU>
__device__ __forceinline__ void press_seg_body(const PressArgs& a, int bx) {
  extern __shared__ double sm[];
  const int CB = blockDim.x, NT = blockDim.x * blockDim.y;
  const int cl = threadIdx.x, sg = threadIdx.y, me = sg * CB + cl;
  const int col = bx * CB + cl;
  const int n = a.nyp;
  const int S = (n + SEG + 1) / (SEG + 1);
  double2* Yb = reinterpret_cast<double2*>(sm);
  double2* Ye = Yb + NT;
  double2* Rs = Ye + NT;
  double2* Xs = Rs + NT;
  double* Gb = reinterpret_cast<double*>(Xs + NT);
  double* Hb = Gb + NT;
  double* Ge = Hb + NT;
  double* He = Ge + NT;
  double* As = He + NT;
  double* Ds = As + NT;
  double* Cs = Ds + NT;
  // per-row metadata staged in LDS: plo, pdg, pup of p row q, dtheta/dy of
  // the u rows 0..n+1 and of the v faces 0..n
  double* lp = Cs + NT;
  double* ldu = lp + 3 * n;
  double* ldv = ldu + n + 2;
  for (int q = me; q < n + 2; q += NT) {
    if (q < n) {
      lp[3 * q] = a.plo[q];
      lp[3 * q + 1] = a.pdg[q];
      lp[3 * q + 2] = a.pup[q];
    }
    ldu[q] = a.dthdyu[q];
    if (q < n + 1) ldv[q] = a.dthdyv[q];
  }
  __syncthreads();
  const bool act = col < a.Ncol && sg < S;
  const size_t N = a.Ncol;
  const int b = sg * (SEG + 1);
  const int i = col % a.Mx, k = col / a.Mx;
  const bool shc = act && a.shortc && a.shortc[col];
  // v faces a short column keeps at zero: its band interfaces
  auto iface = [&](int face) { return shc && (face == a.P1 + 1 || face == a.P3); };
  // p rows between a short column's bands: no divergence there
  auto gap = [&](int q) { return shc && q > a.P1 && q < a.P3; };
  double kx = 0.0, kz = 0.0;
  double2 y[SEG];
  double g[SEG], h[SEG];
  double2 vnext = make_double2(0.0, 0.0);  // old v of the face above the separator
  // the rows' old v (face b+j) and, with CU, u1 (above p row b+j) from the
  // first pass, kept for the projection so that it does not read them again
  double2 cu[CU ? SEG + 1 : 1], cv[SEG + 1];
  if (act) {
    kx = a.k1x[i];
    kz = a.k1z[k];
    const double k2 = a.k2x[i] + a.k2z[k];
    const double dtirk = a.sc->dti / a.apb;
    double2 vq = a.u2[(size_t)b * N + col];  // old v at the face below row b
    cv[0] = vq;
    // row q: divergence of the old field as right-hand side; fetch issues
    // the loads of a row (u1, u3 above p row q and v of the face above it)
    // so that they can be in flight during the previous row's arithmetic
    struct Vel {
      double2 U, W, V1;
    };
    auto fetch = [&](int q) {
      Vel v;
      if (q < n) {
        v.U = a.u1[(size_t)(q + 1) * N + col];
        v.W = a.u3[(size_t)(q + 1) * N + col];
        v.V1 = a.u2[(size_t)(q + 1) * N + col];
      }
      return v;
    };
    auto row = [&](int q, const Vel& v, double& lo, double& dg, double& up, double2& rhs) {
      if (q >= n || (shc && q > a.P1 && q < a.P3)) {
        lo = 0.0; dg = 1.0; up = 0.0; rhs = make_double2(0.0, 0.0);
        if (q < n) vq = v.V1;  // the face chain goes on through the rows between the bands
        return;
      }
      const double2 U = v.U, W = v.W, V1 = v.V1;
      const double2 du = imulk(kx, U), dw = imulk(kz, W);
      const double m = a.ddthetavi * ldu[q + 1];
      const double2 dv = make_double2(du.x + dw.x + m * (V1.x - vq.x), du.y + dw.y + m * (V1.y - vq.y));
      vq = V1;
      rhs = make_double2(dtirk * dv.x, dtirk * dv.y);
      lo = lp[3 * q];
      dg = lp[3 * q + 1] + k2;
      up = lp[3 * q + 2];
      if (col == 0 && q == 0) {
        rhs = make_double2(0.0, 0.0);
        lo = 0.0;
        dg = a.pin;
        up = 0.0;
      }
      // a short column's band ends drop the flux through the interface face
      // (LU_buildP per band, tridLU_3D_IB.f90:417-491)
      if (shc && q == a.P1) {
        dg = -(a.ddthetavi * ldu[q + 1]) * (a.ddthetavi * ldv[q]) + k2;
        up = 0.0;
      }
      if (shc && q == a.P3) {
        dg = -(a.ddthetavi * ldu[q + 1]) * (a.ddthetavi * ldv[q + 1]) + k2;
        lo = 0.0;
      }
    };
    double2 yp = make_double2(0.0, 0.0);
    double gp = 0.0, hp = 1.0;
    Vel cur = fetch(b);
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
      double lo, dg, up;
      double2 r;
      const Vel nxt = fetch(b + j + 1);
      if constexpr (CU) cu[j] = cur.U;
      cv[j + 1] = cur.V1;
      row(b + j, cur, lo, dg, up, r);
      cur = nxt;
      const double inv = 1.0 / (dg - lo * gp);
      y[j] = make_double2((r.x - lo * yp.x) * inv, (r.y - lo * yp.y) * inv);
      g[j] = up * inv;
      h[j] = -lo * hp * inv;
      yp = y[j];
      gp = g[j];
      hp = h[j];
    }
    if (sg < S - 1) {
      double lo, dg, up;
      double2 r;
      if constexpr (CU) cu[SEG] = cur.U;
      row(b + SEG, cur, lo, dg, up, r);
      vnext = vq;  // old v of face b+SEG+1
      As[me] = lo;
      Ds[me] = dg;
      Cs[me] = up;
      Rs[me] = r;
    }
    g[SEG - 1] = -g[SEG - 1];
#pragma unroll
    for (int j = SEG - 2; j >= 0; --j) {
      const double gj = g[j];
      y[j] = make_double2(y[j].x - gj * y[j + 1].x, y[j].y - gj * y[j + 1].y);
      h[j] = h[j] - gj * h[j + 1];
      g[j] = -gj * g[j + 1];
    }
    Yb[me] = y[0];
    Gb[me] = g[0];
    Hb[me] = h[0];
    Ye[me] = y[SEG - 1];
    Ge[me] = g[SEG - 1];
    He[me] = h[SEG - 1];
  }
  __syncthreads();
  if (sg == 0 && col < a.Ncol && S > 1) {
    double cp = 0.0;
    double2 x = make_double2(0.0, 0.0);
    for (int s = 0; s < S - 1; ++s) {
      const int A = s * CB + cl, B = (s + 1) * CB + cl;
      const double as = As[A], cs = Cs[A];
      const double lo = as * He[A];
      const double dg = Ds[A] + as * Ge[A] + cs * Hb[B];
      const double2 rh = make_double2(Rs[A].x - as * Ye[A].x - cs * Yb[B].x, Rs[A].y - as * Ye[A].y - cs * Yb[B].y);
      const double den = 1.0 / (dg - lo * cp);
      x = make_double2((rh.x - lo * x.x) * den, (rh.y - lo * x.y) * den);
      cp = cs * Gb[B] * den;
      Xs[A] = x;
      Cs[A] = cp;
    }
    for (int s = S - 3; s >= 0; --s) {
      const int A = s * CB + cl;
      const double2 xn = Xs[A + CB];
      x = make_double2(Xs[A].x - Cs[A] * xn.x, Xs[A].y - Cs[A] * xn.y);
      Xs[A] = x;
    }
  }
  __syncthreads();
  if (!act || b >= n) return;
  const double2 xl = sg > 0 ? Xs[me - CB] : make_double2(0.0, 0.0);
  const double2 xr = sg < S - 1 ? Xs[me] : make_double2(0.0, 0.0);
  const double dtrk = a.sc->dt * a.apb;
  const double kxd = kx * dtrk, kzd = kz * dtrk;
  const int nyu = n + 2;
  double2 psm = xl;                             // psi of the row below
  double2 unp = make_double2(0.0, 0.0), wnp = unp, vprev = unp;  // corrected u, w, v of the row below
  // psi, p, projection of the u/w row above p row q and of face q, and the
  // divergence of p row q-1 once both its faces are known
  auto proc = [&](int j, double2 ps) {
    const int q = b + j;
    a.psi[(size_t)q * N + col] = ps;
    const double2 P = a.p[(size_t)q * N + col];
    a.p[(size_t)q * N + col] = make_double2(P.x + ps.x, P.y + ps.y);
    double2 U;
    if constexpr (CU)
      U = cu[j];
    else
      U = a.u1[(size_t)(q + 1) * N + col];
    const double2 W = a.u3[(size_t)(q + 1) * N + col];
    const double2 gx = imulk(kxd, ps), gz = imulk(kzd, ps);
    const double2 un = make_double2(U.x - gx.x, U.y - gx.y);
    const double2 wn = make_double2(W.x - gz.x, W.y - gz.y);
    a.u1[(size_t)(q + 1) * N + col] = un;
    a.u3[(size_t)(q + 1) * N + col] = wn;
    if (q == 0) {
      a.u1[col] = make_double2(-a.g0 * un.x, -a.g0 * un.y);
      a.u3[col] = make_double2(-a.g0 * wn.x, -a.g0 * wn.y);
    }
    if (q == n - 1) {
      a.u1[(size_t)(nyu - 1) * N + col] = make_double2(-a.g1 * un.x, -a.g1 * un.y);
      a.u3[(size_t)(nyu - 1) * N + col] = make_double2(-a.g1 * wn.x, -a.g1 * wn.y);
    }
    double2 vf = make_double2(0.0, 0.0);  // bottom wall face, a short column's interface faces
    if (q > 0 && !iface(q)) {
      const double2 V = cv[j];
      const double cc = dtrk * a.ddthetavi * ldv[q];
      vf = make_double2(V.x - cc * (ps.x - psm.x), V.y - cc * (ps.y - psm.y));
    }
    a.u2[(size_t)q * N + col] = vf;
    if (q > b) {
      const double2 du = imulk(kx, unp), dw = imulk(kz, wnp);
      const double m = a.ddthetavi * ldu[q];
      a.div[(size_t)(q - 1) * N + col] =
          gap(q - 1) ? make_double2(0.0, 0.0)
                     : make_double2(du.x + dw.x + m * (vf.x - vprev.x), du.y + dw.y + m * (vf.y - vprev.y));
    }
    unp = un;
    wnp = wn;
    vprev = vf;
    psm = ps;
  };
#pragma unroll
  for (int j = 0; j < SEG; ++j) {
    const int q = b + j;
    if (q < n)
      proc(j, make_double2(y[j].x + g[j] * xr.x + h[j] * xl.x, y[j].y + g[j] * xr.y + h[j] * xl.y));
  }
  int last;
  if (sg < S - 1) {
    last = b + SEG;
    proc(SEG, xr);
  } else {
    last = n - 1;
  }
  // face above the last row: the top wall, or the next segment's first face
  double2 vf = make_double2(0.0, 0.0);
  if (last == n - 1) {
    a.u2[(size_t)n * N + col] = vf;
  } else if (iface(last + 1)) {
    // the next segment writes this face (zero)
  } else {
    const int B = me + CB;
    const double2 xr2 = sg + 1 < S - 1 ? Xs[B] : make_double2(0.0, 0.0);
    const double2 ps1 = make_double2(Yb[B].x + Gb[B] * xr2.x + Hb[B] * xr.x, Yb[B].y + Gb[B] * xr2.y + Hb[B] * xr.y);
    const double cc = dtrk * a.ddthetavi * ldv[last + 1];
    vf = make_double2(vnext.x - cc * (ps1.x - psm.x), vnext.y - cc * (ps1.y - psm.y));
  }
  const double2 du = imulk(kx, unp), dw = imulk(kz, wnp);
  const double m = a.ddthetavi * ldu[last + 1];
  a.div[(size_t)last * N + col] = gap(last) ? make_double2(0.0, 0.0)
                                            : make_double2(du.x + dw.x + m * (vf.x - vprev.x),
                                                           du.y + dw.y + m * (vf.y - vprev.y));
}

template <int SEG, bool CU>
__global__ void __launch_bounds__(256) k_press_seg(PressArgs a) {
  press_seg_body<SEG, CU>(a, blockIdx.x);
}

template <int SEG, bool CU>
__global__ void __launch_bounds__(256) k_press_seg_b(const PressArgs* __restrict__ as, BMap bm) {
  int bx = 0, by = 0;
  const int m = bmap_member(bm, blockIdx.x, bx, by);
  if (m < 0) return;
  const PressArgs a = as[m];
  press_seg_body<SEG, CU>(a, bx);
}

// (0,0) flow-rate correction staged through LDS: the column, metrics and
// scratch live in LDS; one thread runs the sequential recurrences in the
// reference order (flowrate_corr, littleharsh.f90:452-484) while a thread of
// the second wave forms the flow-rate sum of the old profile.  The
// recurrences walk the column in chunks of kMfChunk rows: the chunk's
// coefficients are read from LDS into registers together (their loads
// overlap), the chunk is eliminated in registers, and only the carried
// values link one chunk to the next.
constexpr int kMfChunk = 8;

__global__ void __launch_bounds__(256) k_meanflow_lds(double2* __restrict__ u1, size_t Ncol, int jlo_u, int nn,
                                                      const double* __restrict__ d2u, const double* __restrict__ yv,
                                                      int jlo_v, DevScalars* sc, double bRK, double dRK, double Re,
                                                      double g0, double g1) {
  extern __shared__ double sm[];
  const int n = nn + 3;
  double *a1 = sm, *a2 = sm + n, *a3 = sm + 2 * n, *x = sm + 3 * n, *dy = sm + 4 * n, *ur = sm + 5 * n;
  const double dt = sc->dt;
  const double beta = bRK * dt / Re;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    ur[j] = u1[(size_t)(j - jlo_u) * Ncol].x;
    dy[j] = (j >= 1 && j <= nn + 1) ? (yv[j - jlo_v] - yv[j - 1 - jlo_v]) : 0.0;
    if (j == 0) {
      a1[j] = 0.0; a2[j] = 1.0; a3[j] = g0; x[j] = 0.0;
    } else if (j == nn + 2) {
      a1[j] = g1; a2[j] = 1.0; a3[j] = 0.0; x[j] = 0.0;
    } else {
      const double* d = d2u + 3 * (j - jlo_u);
      a1[j] = -beta * d[0];
      a2[j] = 1.0 - beta * d[1];
      a3[j] = -beta * d[2];
      x[j] = -dRK * dt;
    }
  }
  __syncthreads();
  __shared__ double gsh, q0sh;
  if (threadIdx.x == 64) {
    double Qx0 = 0.0;
    for (int j = 1; j <= nn + 1; ++j) Qx0 = Qx0 + dy[j] * ur[j];
    q0sh = Qx0;
  }
  if (threadIdx.x == 0) {
    constexpr int C = kMfChunk;
    // factorisation: a2 <- 1 / pivot, a3 <- a3 / pivot of the row above
    double p2 = 1.0 / a2[0], p3 = a3[0];  // row j-1 carried in registers
    a2[0] = p2;
    int j = 1;
    for (; j + C <= n; j += C) {
      double c1[C], c2[C], c3[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c1[q] = a1[j + q];
        c2[q] = a2[j + q];
        c3[q] = a3[j + q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const double s3 = p3 * p2;
        a3[j + q - 1] = s3;
        p2 = 1.0 / (c2[q] - c1[q] * s3);
        a2[j + q] = p2;
        p3 = c3[q];
      }
    }
    for (; j < n; ++j) {
      const double c1 = a1[j], c2 = a2[j], c3 = a3[j];
      const double s3 = p3 * p2;
      a3[j - 1] = s3;
      p2 = 1.0 / (c2 - c1 * s3);
      a2[j] = p2;
      p3 = c3;
    }
    // forward substitution
    double xp = x[0] * a2[0];
    x[0] = xp;
    for (j = 1; j + C <= n; j += C) {
      double c1[C], c2[C], cx[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c1[q] = a1[j + q];
        c2[q] = a2[j + q];
        cx[q] = x[j + q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        xp = (cx[q] - c1[q] * xp) * c2[q];
        x[j + q] = xp;
      }
    }
    for (; j < n; ++j) {
      xp = (x[j] - a1[j] * xp) * a2[j];
      x[j] = xp;
    }
    // back substitution
    for (j = n - 2; j - C + 1 >= 0; j -= C) {
      double c3[C], cx[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c3[q] = a3[j - q];
        cx[q] = x[j - q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        xp = cx[q] - c3[q] * xp;
        x[j - q] = xp;
      }
    }
    for (; j >= 0; --j) {
      xp = x[j] - a3[j] * xp;
      x[j] = xp;
    }
    double Qx1 = 0.0;
    for (int j = 1; j <= nn + 1; ++j) Qx1 = Qx1 + dy[j] * x[j];
    gsh = Qx1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double g = (sc->QxT - q0sh) / gsh;
    sc->mpgx = sc->mpgx + g;
    sc->dgx = g;
    gsh = g;
  }
  __syncthreads();
  const double g = gsh;
  for (int j = threadIdx.x; j < n; j += blockDim.x) u1[(size_t)(j - jlo_u) * Ncol].x = ur[j] + g * x[j];
}


// The flow-rate correction split in two (littleharsh.f90:411-537): u11 =
// (I - beta D^2)^-1 (-dRK dt) and Q(u11) depend on dt and the RK
// coefficients only, so k_mf_u11 can run beside the column solves; k_mf_apply
// then takes Q(u), g = (QxT - Q(u)) / Q(u11), mpgx += g and u += g u11 for the
// (0,0) column.  The same operations in the same order as k_meanflow_lds.
// u11[0..n-1] the solution, u11[n] = Q(u11); one block.
__global__ void __launch_bounds__(256) k_mf_u11(const double* __restrict__ d2u, const double* __restrict__ yv,
                                                int jlo_u, int jlo_v, int nn, const DevScalars* sc, double bRK,
                                                double dRK, double Re, double g0, double g1, double* __restrict__ u11) {
  extern __shared__ double sm[];
  const int n = nn + 3;
  double *a1 = sm, *a2 = sm + n, *a3 = sm + 2 * n, *x = sm + 3 * n, *dy = sm + 4 * n;
  const double dt = sc->dt;
  const double beta = bRK * dt / Re;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    dy[j] = (j >= 1 && j <= nn + 1) ? (yv[j - jlo_v] - yv[j - 1 - jlo_v]) : 0.0;
    if (j == 0) {
      a1[j] = 0.0; a2[j] = 1.0; a3[j] = g0; x[j] = 0.0;
    } else if (j == nn + 2) {
      a1[j] = g1; a2[j] = 1.0; a3[j] = 0.0; x[j] = 0.0;
    } else {
      const double* d = d2u + 3 * (j - jlo_u);
      a1[j] = -beta * d[0];
      a2[j] = 1.0 - beta * d[1];
      a3[j] = -beta * d[2];
      x[j] = -dRK * dt;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    constexpr int C = kMfChunk;
    double p2 = 1.0 / a2[0], p3 = a3[0];
    a2[0] = p2;
    int j = 1;
    for (; j + C <= n; j += C) {
      double c1[C], c2[C], c3[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c1[q] = a1[j + q];
        c2[q] = a2[j + q];
        c3[q] = a3[j + q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const double s3 = p3 * p2;
        a3[j + q - 1] = s3;
        p2 = 1.0 / (c2[q] - c1[q] * s3);
        a2[j + q] = p2;
        p3 = c3[q];
      }
    }
    for (; j < n; ++j) {
      const double c1 = a1[j], c2 = a2[j], c3 = a3[j];
      const double s3 = p3 * p2;
      a3[j - 1] = s3;
      p2 = 1.0 / (c2 - c1 * s3);
      a2[j] = p2;
      p3 = c3;
    }
    double xp = x[0] * a2[0];
    x[0] = xp;
    for (j = 1; j + C <= n; j += C) {
      double c1[C], c2[C], cx[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c1[q] = a1[j + q];
        c2[q] = a2[j + q];
        cx[q] = x[j + q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        xp = (cx[q] - c1[q] * xp) * c2[q];
        x[j + q] = xp;
      }
    }
    for (; j < n; ++j) {
      xp = (x[j] - a1[j] * xp) * a2[j];
      x[j] = xp;
    }
    for (j = n - 2; j - C + 1 >= 0; j -= C) {
      double c3[C], cx[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c3[q] = a3[j - q];
        cx[q] = x[j - q];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        xp = cx[q] - c3[q] * xp;
        x[j - q] = xp;
      }
    }
    for (; j >= 0; --j) {
      xp = x[j] - a3[j] * xp;
      x[j] = xp;
    }
    double Qx1 = 0.0;
    for (int jj = 1; jj <= nn + 1; ++jj) Qx1 = Qx1 + dy[jj] * x[jj];
    u11[n] = Qx1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) u11[j] = x[j];
}

__global__ void __launch_bounds__(256) k_mf_apply(double2* __restrict__ u1, size_t Ncol, int jlo_u, int nn,
                                                  const double* __restrict__ yv, int jlo_v, DevScalars* sc,
                                                  const double* __restrict__ u11) {
  extern __shared__ double sm[];
  const int n = nn + 3;
  double *ur = sm, *dy = sm + n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    ur[j] = u1[(size_t)(j - jlo_u) * Ncol].x;
    dy[j] = (j >= 1 && j <= nn + 1) ? (yv[j - jlo_v] - yv[j - 1 - jlo_v]) : 0.0;
  }
  __syncthreads();
  __shared__ double gsh;
  if (threadIdx.x == 0) {
    double Qx0 = 0.0;
    for (int j = 1; j <= nn + 1; ++j) Qx0 = Qx0 + dy[j] * ur[j];
    const double g = (sc->QxT - Qx0) / u11[n];
    sc->mpgx = sc->mpgx + g;
    sc->dgx = g;
    gsh = g;
  }
  __syncthreads();
  const double g = gsh;
  for (int j = threadIdx.x; j < n; j += blockDim.x) u1[(size_t)(j - jlo_u) * Ncol].x = ur[j] + g * u11[j];
}

}  // namespace

void fused_args(Solver& s, int kRK, void* out) {
  FusedArgs& a = *static_cast<FusedArgs*>(out);
  for (int f = 0; f < 3; ++f) {
    a.u[f] = s.F[CFD_U1 + f];
    a.uo[f] = s.F[CFD_U1 + f];
    a.nu[f] = s.F[CFD_NU1 + f];
    a.r[f] = s.F[CFD_DU1 + f];
    a.ysc[f] = s.ysc[f];
    a.cs[f] = s.cscr + (size_t)f * s.nyu * s.Ncol;
    a.ibmask[f] = s.d_ibmask[f];
  }
  a.nxz[0] = s.nxz1; a.nxz[1] = s.nxz2; a.nxz[2] = s.nxz3;
  a.flux[0] = s.fuv; a.flux[1] = s.fvv; a.flux[2] = s.fvw;
  a.p = s.F[CFD_P];
  a.d2u = s.d_d2u; a.d2v = s.d_d2v; a.dthdyu = s.d_dthdyu; a.dthdyv = s.d_dthdyv;
  a.k1x = s.d_k1x; a.k2x = s.d_k2x; a.k1z = s.d_k1z; a.k2z = s.d_k2z;
  a.sc = s.d_sc;
  a.C1 = s.aRK[kRK - 1] / s.P.Re;
  a.C2 = -s.cRK[kRK - 1];
  a.C1g = -s.gRK[kRK - 1];
  a.dRK = s.dRK[kRK - 1];
  a.bRK = s.bRK[kRK - 1];
  a.Re = s.P.Re;
  a.ddthetavi = 1.0;
  a.g0 = s.gw[0];
  a.g1 = s.gw[1];
  a.Mx = s.Mx;
  a.Ncol = s.Ncol;
  a.nyu = s.nyu;
  a.nyv = s.nyv;
  a.ib_only = 0;
  a.shortc = s.mb ? s.d_short : nullptr;
  // short columns' band ends in plane indices: u rows N(4,1)+1 and N(4,2),
  // v rows N(3,1)+1 and N(3,2)
  a.t1[0] = a.t1[2] = s.N4[1] + 1 - s.jlo_u;
  a.b3[0] = a.b3[2] = s.N4[2] - s.jlo_u;
  a.t1[1] = s.N3[1] + 1 - s.jlo_v;
  a.b3[1] = s.N3[2] - s.jlo_v;
  a.gw12 = s.gw12;
  a.gw31 = s.gw31;
}

void launch_ib_rhs(Solver& s, int kRK) {
  const int n0 = s.nibp[0], n1 = s.nibp[1], n2 = s.nibp[2];
  const int nmax = std::max(n0, std::max(n1, n2));
  if (nmax == 0) return;
  FusedArgs a;
  fused_args(s, kRK, &a);
  dim3 grid((s.Ncol + 255) / 256, nmax, 3);
  if (s.mixed)
    hipLaunchKernelGGL(k_ib_rhs<float2>, grid, dim3(256), 0, s.stream, a, (const int*)s.d_ibq[0],
                       (const int*)s.d_ibq[1], (const int*)s.d_ibq[2], n0, n1, n2);
  else
    hipLaunchKernelGGL(k_ib_rhs<double2>, grid, dim3(256), 0, s.stream, a, (const int*)s.d_ibq[0],
                       (const int*)s.d_ibq[1], (const int*)s.d_ibq[2], n0, n1, n2);
}

// The segmented solves are instantiated for these rows per segment; the
// solver's seg_rows picks one, solve_threads caps the block.
template <int SEG>
void rhs_seg_launch(Solver& s, const FusedArgs& a, int sy) {
  int cb = 1;
  while (cb * 2 * sy <= s.solve_threads) cb *= 2;
  // separator system (15 doubles per thread), then per row 4 doubles of
  // metadata and the IB flag (nyu rows: the longer of the u and v columns)
  const size_t lds = sizeof(double) * (15 * (size_t)cb * sy + 4 * (size_t)s.nyu) + (size_t)s.nyu;
  const int ntiles = (s.Ncol + cb - 1) / cb;
  const dim3 grid(24 * ((ntiles + 7) / 8)), block(cb, sy);
  if (s.mixed)
    hipLaunchKernelGGL((k_rhs_seg<SEG, float2>), grid, block, lds, s.stream, a);
  else
    hipLaunchKernelGGL((k_rhs_seg<SEG, double2>), grid, block, lds, s.stream, a);
}

template <int SEG>
void press_seg_launch(Solver& s, const PressArgs& a, int sy) {
  int cb = 1;
  while (cb * 2 * sy <= s.solve_threads) cb *= 2;
  // separator system (15 doubles per thread), then 5 nyp + 3 doubles of
  // per-row metadata
  const size_t lds = sizeof(double) * (15 * (size_t)cb * sy + 5 * (size_t)s.nyp + 3);
  // the u1 rows fit the register budget (two waves per SIMD) up to 9 rows
  hipLaunchKernelGGL((k_press_seg<SEG, (SEG <= 9)>), dim3((s.Ncol + cb - 1) / cb), dim3(cb, sy), lds, s.stream, a);
}

bool seg_rows_supported(int r) { return r == 5 || r == 9 || r == 13 || r == 21; }

void launch_rhs_solve(Solver& s, int kRK) {
  FusedArgs a;
  fused_args(s, kRK, &a);
  const int R = s.seg_rows;
  const int sy = (s.nyu + R + 1) / (R + 1);  // segments of the longest column (S of k_rhs_seg)
  if (sy <= 256) {
    switch (R) {
      case 5: rhs_seg_launch<5>(s, a, sy); break;
      case 9: rhs_seg_launch<9>(s, a, sy); break;
      case 21: rhs_seg_launch<21>(s, a, sy); break;
      default: rhs_seg_launch<13>(s, a, sy); break;
    }
  } else if (s.mixed) {
    hipLaunchKernelGGL(k_rhs_solve<float2>, dim3((s.Ncol + 63) / 64, 3), dim3(64), 0, s.stream, a);
  } else {
    hipLaunchKernelGGL(k_rhs_solve<double2>, dim3((s.Ncol + 63) / 64, 3), dim3(64), 0, s.stream, a);
  }
}

namespace {
void press_args(Solver& s, int kRK, PressArgs& a) {
  a.u1 = s.F[CFD_U1]; a.u2 = s.F[CFD_U2]; a.u3 = s.F[CFD_U3];
  a.p = s.F[CFD_P]; a.psi = s.F[CFD_PSI]; a.div = s.F[CFD_DIV];
  a.cs = s.cscr;
  a.plo = s.d_plo; a.pup = s.d_pup; a.pdg = s.d_pdg; a.dthdyu = s.d_dthdyu; a.dthdyv = s.d_dthdyv;
  a.k1x = s.d_k1x; a.k2x = s.d_k2x; a.k1z = s.d_k1z; a.k2z = s.d_k2z;
  a.sc = s.d_sc;
  a.apb = s.aRK[kRK - 1] + s.bRK[kRK - 1];
  a.ddthetavi = 1.0;
  const double h = s.yg.YU(2) - s.yg.YU(1);
  a.pin = 1.0 / (h * h);
  a.g0 = s.gw[0];
  a.g1 = s.gw[1];
  a.Mx = s.Mx;
  a.Ncol = s.Ncol;
  a.nyp = s.nyp;
  a.shortc = s.mb ? s.d_short : nullptr;
  a.P1 = s.N4[1] - s.jlo_p;      // p row N(4,1): the end of band 1
  a.P3 = s.N4[2] + 1 - s.jlo_p;  // p row N(4,2)+1: the start of band 3
}
}  // namespace

void launch_pressure(Solver& s, int kRK) {
  PressArgs a;
  press_args(s, kRK, a);
  const int R = s.seg_rows;
  const int sy = (s.nyp + R + 1) / (R + 1);  // S of k_press_seg
  if (sy <= 256) {
    switch (R) {
      case 5: press_seg_launch<5>(s, a, sy); break;
      case 9: press_seg_launch<9>(s, a, sy); break;
      case 21: press_seg_launch<21>(s, a, sy); break;
      default: press_seg_launch<13>(s, a, sy); break;
    }
  } else {
    hipLaunchKernelGGL(k_pressure, dim3((s.Ncol + 63) / 64), dim3(64), 0, s.stream, a);
  }
}

void launch_mf_u11(Solver& s, int kRK, hipStream_t st) {
  const int n = s.nn + 3;
  hipLaunchKernelGGL(k_mf_u11, dim3(1), dim3(256), sizeof(double) * 5 * n, st, (const double*)s.d_d2u,
                     (const double*)s.d_yv, s.jlo_u, s.jlo_v, s.nn, (const DevScalars*)s.d_sc, s.bRK[kRK - 1],
                     s.dRK[kRK - 1], s.P.Re, s.gw[0], s.gw[1], s.d_u11);
}

void launch_mf_apply(Solver& s) {
  const int n = s.nn + 3;
  hipLaunchKernelGGL(k_mf_apply, dim3(1), dim3(256), sizeof(double) * 2 * n, s.stream, s.F[CFD_U1], (size_t)s.Ncol,
                     s.jlo_u, s.nn, (const double*)s.d_yv, s.jlo_v, s.d_sc, (const double*)s.d_u11);
}

void launch_meanflow_lds(Solver& s, int kRK) {
  const int n = s.nn + 3;
  const size_t lds = sizeof(double) * 6 * n;
  hipLaunchKernelGGL(k_meanflow_lds, dim3(1), dim3(256), lds, s.stream, s.F[CFD_U1], (size_t)s.Ncol, s.jlo_u, s.nn,
                     (const double*)s.d_d2u, (const double*)s.d_yv, s.jlo_v, s.d_sc, s.bRK[kRK - 1],
                     s.dRK[kRK - 1], s.P.Re, s.gw[0], s.gw[1]);
}

// ---------------------------------------------------------------------------
// batched launches of the column sweeps (ensembles)
// ---------------------------------------------------------------------------
namespace {
int seg_index(int R) { return R == 5 ? 0 : R == 9 ? 1 : R == 21 ? 3 : 2; }
constexpr int kSegOf[4] = {5, 9, 13, 21};
}  // namespace

bool batch_part_fused(Solver& s, int cls, int kRK, BatchPart& part, std::string& err) {
  part = BatchPart();
  if (cls != BC_RHS && cls != BC_PRESS) {
    err = "not a column-sweep class";
    return false;
  }
  const int R = seg_rows_supported(s.seg_rows) ? s.seg_rows : 13;
  const int n = cls == BC_RHS ? s.nyu : s.nyp;
  const int sy = (n + R + 1) / (R + 1);
  if (sy > 256) {
    err = "the column sweep of this solver is not segmented";
    return false;
  }
  int cb = 1;
  while (cb * 2 * sy <= s.solve_threads) cb *= 2;
  part.block = dim3(cb, sy);
  part.lds_thread = sizeof(double) * 15;
  if (cls == BC_RHS) {
    FusedArgs a;
    fused_args(s, kRK, &a);
    const int ntiles = (s.Ncol + cb - 1) / cb;
    part.gx = 24 * ((ntiles + 7) / 8);
    part.lds_fixed = sizeof(double) * 4 * (size_t)s.nyu + (size_t)s.nyu;
    part.args.resize(sizeof(a));
    std::memcpy(part.args.data(), &a, sizeof(a));
  } else {
    PressArgs a;
    press_args(s, kRK, a);
    part.gx = (s.Ncol + cb - 1) / cb;
    part.lds_fixed = sizeof(double) * (5 * (size_t)s.nyp + 3);
    part.args.resize(sizeof(a));
    std::memcpy(part.args.data(), &a, sizeof(a));
  }
  part.gy = 1;
  // variant: rows per segment, precision (column sweeps of u only), threads per column tile
  part.key = (seg_index(R) * 2 + (cls == BC_RHS && s.mixed ? 1 : 0)) * 512 + cb;
  return true;
}

void batch_launch_fused(hipStream_t st, const BatchLaunch& b) {
  const int seg = kSegOf[(b.key / 512) / 2];
  const bool f32 = ((b.key / 512) & 1) != 0;
  if (b.cls == BC_RHS) {
    const FusedArgs* as = static_cast<const FusedArgs*>(b.dargs);
#define C(S)                                                                                         \
  case S:                                                                                            \
    if (f32)                                                                                         \
      hipLaunchKernelGGL((k_rhs_seg_b<S, float2>), b.grid, b.block, b.lds, st, as, b.map);          \
    else                                                                                             \
      hipLaunchKernelGGL((k_rhs_seg_b<S, double2>), b.grid, b.block, b.lds, st, as, b.map);         \
    return;
    switch (seg) {
      C(5) C(9) C(13) C(21)
      default: return;
    }
#undef C
  }
  if (b.cls == BC_PRESS) {
    const PressArgs* as = static_cast<const PressArgs*>(b.dargs);
#define C(S)                                                                                      \
  case S:                                                                                         \
    hipLaunchKernelGGL((k_press_seg_b<S, (S <= 9)>), b.grid, b.block, b.lds, st, as, b.map);      \
    return;
    switch (seg) {
      C(5) C(9) C(13) C(21)
      default: return;
    }
#undef C
  }
}

}  // namespace cfd
