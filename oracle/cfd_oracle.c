/*
 * cfd_oracle.c -- CPU FP64 restatement of the reference RK3/CN substep of the
 * rough-channel DNS (Mel222/CFD).  TEST INFRASTRUCTURE ONLY: the checker for
 * the HIP product path and the timed CPU baseline ("port").  Nothing in the
 * product links it.
 *
 * Layout follows the reference: spectral fields are complex(8) f(j,column),
 * y fastest, columns ordered k = 1..N2 outer, i = 0..N1/2 inner
 * (start.f90:1467-1481).  Physical planes are xPL(Ngalx+2, Ngalz)
 * (declaration.f90:79-84, start.f90:566-577).
 */
#include "cfd_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* complex FFT: recursive mixed-radix decimation in time, factors 4,2,3,5,...
   (the factor order of cffti1, cft_buff.f:144).  Generic radix butterfly.   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, nf;
  int fac[64];       /* p1, m1, p2, m2, ... */
  cplx* tw;          /* tw[j] = exp(-2 pi i j / n) */
} fplan;

static void fplan_init(fplan* P, int n) {
  P->n = n;
  P->tw = (cplx*)malloc(sizeof(cplx) * (n > 0 ? n : 1));
  for (int j = 0; j < n; ++j) {
    /* reduce the angle to the first octant for an accurate table */
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
    P->tw[j] = (double)cosl(a) + I * (double)sinl(a);
  }
  int m = n, nf = 0;
  const int tryh[4] = {4, 2, 3, 5};
  for (int t = 0; m > 1;) {
    int p = (t < 4) ? tryh[t] : (2 * t - 3); /* 4,2,3,5 then odd numbers 5,7,9,... */
    if (m % p == 0) {
      m /= p;
      P->fac[2 * nf] = p;
      P->fac[2 * nf + 1] = m;
      ++nf;
    } else {
      ++t;
    }
  }
  P->nf = nf;
}
static void fplan_free(fplan* P) { free(P->tw); P->tw = NULL; }

static void fft_work(cplx* out, const cplx* in, int fstride, int istride,
                     const int* fac, const fplan* P, int sign) {
  const int p = fac[0], m = fac[1];
  if (m == 1) {
    for (int q = 0; q < p; ++q) out[q] = in[(size_t)q * fstride * istride];
  } else {
    for (int q = 0; q < p; ++q)
      fft_work(out + (size_t)q * m, in + (size_t)q * fstride * istride, fstride * p,
               istride, fac + 2, P, sign);
  }
  const int n = P->n;
  cplx s[16];
  cplx* big = NULL;
  cplx* scr = s;
  if (p > 16) { big = (cplx*)malloc(sizeof(cplx) * p); scr = big; }
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) scr[q] = out[k + q * m];
    for (int u = 0; u < p; ++u) {
      const int idx = k + u * m;
      cplx acc = scr[0];
      int t = 0;
      for (int q = 1; q < p; ++q) {
        t += fstride * idx;
        t %= n;
        cplx w = P->tw[t];
        if (sign > 0) w = conj(w);
        acc += scr[q] * w;
      }
      out[idx] = acc;
    }
  }
  free(big);
}

/* out = sum_k in[k] exp(sign 2 pi i jk/n), unnormalised; in and out distinct */
static void cfft(const fplan* P, const cplx* in, cplx* out, int sign) {
  if (P->n == 1) { out[0] = in[0]; return; }
  fft_work(out, in, 1, 1, P->fac, P, sign);
}

/* ------------------------------------------------------------------------ */
/* reference-convention 1-D transforms                                       */
/* ------------------------------------------------------------------------ */
/* complex: forward (iopt<0) e^{-i}, scaled by (1/n); inverse e^{+i},
   unscaled (cft_buff.f:38-60) */
static void cft_line(const fplan* P, double* c, int ise, int iopt, cplx* w0, cplx* w1) {
  const int n = P->n;
  for (int j = 0; j < n; ++j) w0[j] = c[(size_t)j * ise] + I * c[(size_t)j * ise + 1];
  if (iopt < 0) {
    cfft(P, w0, w1, -1);
    const double dn1 = 1.0 / n;
    for (int j = 0; j < n; ++j) {
      c[(size_t)j * ise] = dn1 * creal(w1[j]);
      c[(size_t)j * ise + 1] = dn1 * cimag(w1[j]);
    }
  } else {
    cfft(P, w0, w1, +1);
    for (int j = 0; j < n; ++j) {
      c[(size_t)j * ise] = creal(w1[j]);
      c[(size_t)j * ise + 1] = cimag(w1[j]);
    }
  }
}

/* real: row c[0..n+1]; forward packs c(2k+1)+i c(2k+2) = (1/n) sum x e^{-i},
   with c(2) = c(n+2) = 0; inverse reads Re/Im pairs, ignores the imaginary
   parts of the k=0 and k=n/2 terms, and writes n reals (rft_buff.f:57-83).
   Uses a half-length complex transform (h = n/2). */
static void rft_line(const fplan* Ph, const cplx* twn, double* c, int iopt, cplx* w0, cplx* w1) {
  const int h = Ph->n, n = 2 * h;
  if (iopt < 0) {
    for (int m = 0; m < h; ++m) w0[m] = c[2 * m] + I * c[2 * m + 1];
    cfft(Ph, w0, w1, -1);
    const double dn1 = 1.0 / n;
    for (int k = 0; k <= h; ++k) {
      const cplx zk = w1[k % h];
      const cplx zc = conj(w1[(h - k) % h]);
      const cplx fe = 0.5 * (zk + zc);
      const cplx fo = -0.5 * I * (zk - zc);
      const cplx X = fe + twn[k] * fo;
      c[2 * k] = dn1 * creal(X);
      c[2 * k + 1] = dn1 * cimag(X);
    }
    c[1] = 0.0;
    c[n + 1] = 0.0;
  } else {
    /* X_k for k=0..h with the imaginary parts of X_0 and X_h dropped */
    for (int k = 0; k < h; ++k) {
      cplx xk = c[2 * k] + I * c[2 * k + 1];
      cplx xh = c[2 * (h - k)] + I * c[2 * (h - k) + 1];
      if (k == 0) { xk = creal(xk); xh = creal(xh); }
      const cplx xc = conj(xh);
      w0[k] = (xk + xc) + I * conj(twn[k]) * (xk - xc);
    }
    cfft(Ph, w0, w1, +1);
    for (int m = 0; m < h; ++m) {
      c[2 * m] = creal(w1[m]);
      c[2 * m + 1] = cimag(w1[m]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* state                                                                     */
/* ------------------------------------------------------------------------ */
struct or_state {
  or_params P;
  int ngx, ngz, ny, dsty, N1, N2, Mx, Ncol;
  int jlo_u, jhi_u, nyu, jlo_v, jhi_v, nyv, jlo_p, jhi_p, nyp;
  int nn; /* ny-2 after y_grid_canopy (y_grid_canopy.f90:252) */
  double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i;
  double gw[2], gwi[2];
  double *k1x, *k2x, *k1z, *k2z;
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  double dtv, ddthetavi, dthetavi;
  double t, dt, dti, mpgx, dgx, QxT, Qx, dtc;
  cplx* F[12];           /* u1 u2 u3 p Nu1 Nu2 Nu3 du1 du2 du3 psi div */
  cplx *u1i, *u2i, *u3i, *uvf, *vvc, *vwf;
  double* DG;
  int nib[2][2];
  int* ibl[2][2];
  double* ibw[2][2];
  fplan pz, px2; /* Ngz complex; Ngx/2 complex for the real transform */
  cplx* twx;     /* exp(-2 pi i k/Ngx), k=0..Ngx/2 */
  struct or_stats* st; /* statistics accumulators, created by or_stats_init */
  /* the as-shipped last-plane nonlinear term of np ranks (0: per plane) */
  int cnp;
  int* clim;             /* planelim(grid, 1:2, 0:np-1), u then v: [(g*2+e)*np + r] */
  cplx *xz1, *xz3, *xz2; /* x/z derivative sums of every plane */
  /* multi-band resolution (start.f90:89-111, 1420-1557): mb = 1 when the
     central band is transformed on the coarser Ngal(1:2,2) grid */
  int mb, ngx2, ngz2, N1c, N2c;
  int N3[4], N4[4];         /* band limits N(3,0:3), N(4,0:3) (y_grid_canopy.f90:250-275) */
  unsigned char* shortc;    /* per column: 1 = a short column (wall bands only) */
  double gw12, gw31, gwi12, gwi31; /* gridweighting(_interp) at the band interfaces */
  fplan pz2, px22;          /* central-band transforms */
  cplx* twx2;
};

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }

#define YU(s, j) ((s)->yu[(j) - (s)->jlo_u])
#define YV(s, j) ((s)->yv[(j) - (s)->jlo_v])
#define DTHU(s, j) ((s)->dthdyu[(j) - (s)->jlo_u])
#define DTHV(s, j) ((s)->dthdyv[(j) - (s)->jlo_v])
#define DYU2(s, m, j) ((s)->dyu2i[3 * ((j) - (s)->jlo_u) + (m) - 1])
#define DYV2(s, m, j) ((s)->dyv2i[3 * ((j) - (s)->jlo_v) + (m) - 1])

/* field element f(j, col) of grid g (0 u, 1 v, 2 p) */
static inline size_t fu(const or_state* s, int col, int j) { return (size_t)col * s->nyu + (j - s->jlo_u); }
static inline size_t fv(const or_state* s, int col, int j) { return (size_t)col * s->nyv + (j - s->jlo_v); }
static inline size_t fp(const or_state* s, int col, int j) { return (size_t)col * s->nyp + (j - s->jlo_p); }
/* band segments of a column and the gridweighting at their ends (below) */
static int col_segs(const or_state* s, int col, int g, int lo[2], int hi[2]);
static double gw_lo(const or_state* s, int lo);
static double gw_hi(const or_state* s, int hi);
static double gwi_lo(const or_state* s, int lo);
static double gwi_hi(const or_state* s, int hi);

/* x**e for an integer exponent e >= 0, by binary exponentiation (the
   usual lowering of Fortran real**integer) */
static double ipow(double x, int e) {
  double r = 1.0;
  for (;;) {
    if (e & 1) r *= x;
    e /= 2;
    if (e == 0) break;
    x *= x;
  }
  return r;
}

/* ------------------------------------------------------------------------ */
/* y grid: y_grid_canopy.f90:24-299 (ygrid 0); tanh variant (ygrid 1)        */
/* ------------------------------------------------------------------------ */
static int build_ygrid(or_state* s, char* err, int errlen) {
  const or_params* P = &s->P;
  const int dny = P->dsty, dny1 = P->dsty1;
  const int nn = s->ny - 1; /* nn = N(3,nband)+1 = ny-1 (y_grid_canopy.f90:31) */
  const double q = nn / 2.0;
  const int pp = P->ppp;
  const double dyq = P->dyq;
  double aaa = 0, bbb = 0, ccc = 0;
  if (P->ygrid == 0) {
    aaa = 15.0 * (dyq - 1.0) / ((8.0 * dyq + 7.0) * ipow(q, pp));
    bbb = -30.0 * (dyq - 1.0) / ((8.0 * dyq + 7.0) * ipow(q, pp - 2));
    ccc = 15.0 * dyq / ((8.0 * dyq + 7.0) * q);
  }
  const double tb = P->ytanh_beta;
  /* channel map theta -> y and its derivative */
#define YCH(th) (P->ygrid == 0 ? (aaa * ipow((th), pp) / 5.0 + bbb * ipow((th), pp - 2) / 3.0 + ccc * (th)) \
                               : tanh(tb * (th) / q) / tanh(tb))
#define DCH(th) (P->ygrid == 0 ? (1.0 / (aaa * ipow((th), pp - 1) + bbb * ipow((th), pp - 3) + ccc)) \
                               : (1.0 / (tb / (q * tanh(tb)) / (cosh(tb * (th) / q) * cosh(tb * (th) / q)))))
  if (P->ygrid == 0 && pp < 3) { snprintf(err, errlen, "ppp must be >= 3"); return 1; }
  /* arrays: yu(-dny:nn+dny+1), yv(-dny:nn+dny) */
  for (int j = 0; j <= nn - 1; ++j) {
    YV(s, j) = YCH(j - q);
    YU(s, j + 1) = YCH(j + 0.5 - q);
    DTHV(s, j) = DCH(j - q);
    DTHU(s, j + 1) = DCH(j + 0.5 - q);
  }
  if (dny == 0) {
    YU(s, 0) = YCH(-0.5 - q);
    DTHU(s, 0) = DCH(-0.5 - q);
    const int j = nn;
    YV(s, j) = YCH(j - q);
    YU(s, j + 1) = YCH(j + 0.5 - q);
    DTHV(s, j) = DCH(j - q);
    DTHU(s, j + 1) = DCH(j + 0.5 - q);
  } else {
    const double grid_tip = YV(s, 1) - YV(s, 0);
    const double posth = P->posth, posth1 = P->posth1;
    if (dny1 == 0 || dny1 == dny || posth1 == 0.0 || posth1 == posth) {
      const double a5 = -(15.0 * (posth - grid_tip * dny)) / (7.0 * pow((double)dny, 5.0));
      const double a3 = (30.0 * (posth - grid_tip * dny)) / (7.0 * pow((double)dny, 3.0));
      const double a1 = grid_tip;
      const double x0 = -dny - 0.5;
      YU(s, -dny) = a5 * ipow(x0, 5) / 5.0 + a3 * ipow(x0, 3) / 3.0 + a1 * x0 - 1.0;
      DTHU(s, -dny) = 1.0 / (a5 * ipow(x0, 4) + a3 * ipow(x0, 2) + a1);
      for (int j = -dny; j <= -1; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j) = a5 * pow(xj, 5.0) / 5.0 + a3 * pow(xj, 3.0) / 3.0 + a1 * j - 1.0;
        YU(s, j + 1) = a5 * ipow(xh, 5) / 5.0 + a3 * ipow(xh, 3) / 3.0 + a1 * xh - 1.0;
        DTHV(s, j) = 1.0 / (a5 * pow(xj, 4.0) + a3 * pow(xj, 2.0) + a1);
        DTHU(s, j + 1) = 1.0 / (a5 * ipow(xh, 4) + a3 * ipow(xh, 2) + a1);
      }
      for (int j = 0; j <= dny; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j + nn) = a5 * pow(xj, 5.0) / 5.0 + a3 * pow(xj, 3.0) / 3.0 + a1 * j + 1.0;
        YU(s, j + nn + 1) = a5 * ipow(xh, 5) / 5.0 + a3 * ipow(xh, 3) / 3.0 + a1 * xh + 1.0;
        DTHV(s, j + nn) = 1.0 / (a5 * pow(xj, 4.0) + a3 * pow(xj, 2.0) + a1);
        DTHU(s, j + nn + 1) = 1.0 / (a5 * ipow(xh, 4) + a3 * ipow(xh, 2) + a1);
      }
    } else {
      /* two-layer canopy grid (y_grid_canopy.f90:119-201) */
      const double a5 = -(15.0 * (posth1 - grid_tip * dny1)) / (7.0 * pow((double)dny1, 5.0));
      const double a3 = (30.0 * (posth1 - grid_tip * dny1)) / (7.0 * pow((double)dny1, 3.0));
      const double a1 = grid_tip;
      for (int j = -dny; j <= -1; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j) = a5 * pow(xj, 5.0) / 5.0 + a3 * pow(xj, 3.0) / 3.0 + a1 * j - 1.0;
        YU(s, j + 1) = a5 * pow(xh, 5.0) / 5.0 + a3 * pow(xh, 3.0) / 3.0 + a1 * xh - 1.0;
        DTHV(s, j) = 1.0 / (a5 * pow(xj, 4.0) + a3 * pow(xj, 2.0) + a1);
        DTHU(s, j + 1) = 1.0 / (a5 * pow(xh, 4.0) + a3 * pow(xh, 2.0) + a1);
      }
      for (int j = 0; j <= dny1; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j + nn) = a5 * pow(xj, 5.0) / 5.0 + a3 * pow(xj, 3.0) / 3.0 + a1 * j + 1.0;
        YU(s, j + nn + 1) = a5 * pow(xh, 5.0) / 5.0 + a3 * pow(xh, 3.0) / 3.0 + a1 * xh + 1.0;
        DTHV(s, j + nn) = 1.0 / (a5 * pow(xj, 4.0) + a3 * pow(xj, 2.0) + a1);
        DTHU(s, j + nn + 1) = 1.0 / (a5 * pow(xh, 4.0) + a3 * pow(xh, 2.0) + a1);
      }
      const double gm = YV(s, -dny1 + 1) - YV(s, -dny1);
      const double D = (double)dny, E = (double)dny1;
      const double den = ((D - E) * (11.0 * pow(D, 6.0) + 11.0 * pow(D, 5.0) * E - 38.0 * pow(D, 4.0) * pow(E, 2.0) -
                                     38.0 * pow(D, 3.0) * pow(E, 3.0) + 67.0 * pow(D, 2.0) * pow(E, 4.0) +
                                     11.0 * D * pow(E, 5.0) - 24.0 * pow(E, 6.0)));
      const double b7 = -(70.0 * (posth - posth1 - D * gm + E * gm)) / den;
      const double b5 = (105.0 * (pow(D, 2.0) + pow(E, 2.0)) * (posth - posth1 - D * gm + E * gm)) / den;
      const double b3 = (210.0 * pow(E, 2.0) * (pow(D, 3.0) * gm - pow(D, 2.0) * posth + pow(D, 2.0) * posth1 -
                                                pow(D, 2.0) * E * gm)) / den;
      const double b1 = (11.0 * pow(D, 7.0) * gm - 11.0 * pow(E, 7.0) * gm - 35.0 * pow(E, 6.0) * posth +
                         35.0 * pow(E, 6.0) * posth1 + 49.0 * pow(D, 2.0) * pow(E, 5.0) * gm -
                         49.0 * pow(D, 5.0) * pow(E, 2.0) * gm + 105.0 * pow(D, 2.0) * pow(E, 4.0) * posth -
                         105.0 * pow(D, 2.0) * pow(E, 4.0) * posth1) / den;
      const double b0 = -(-11.0 * gm * pow(D, 7.0) * E + 11.0 * posth1 * pow(D, 7.0) + 49.0 * gm * pow(D, 5.0) * pow(E, 3.0) -
                          49.0 * posth1 * pow(D, 5.0) * pow(E, 2.0) - 49.0 * gm * pow(D, 3.0) * pow(E, 5.0) +
                          105.0 * posth1 * pow(D, 3.0) * pow(E, 4.0) - 56.0 * posth * pow(D, 2.0) * pow(E, 5.0) +
                          11.0 * gm * D * pow(E, 7.0) - 35.0 * posth1 * D * pow(E, 6.0) + 24.0 * posth * pow(E, 7.0)) / den;
      for (int j = -dny; j <= -dny1 - 1; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j) = b7 * pow(xj, 7.0) / 7.0 + b5 * pow(xj, 5.0) / 5.0 + b3 * pow(xj, 3.0) / 3.0 + b1 * j + b0 - 1.0;
        YU(s, j + 1) = b7 * pow(xh, 7.0) / 7.0 + b5 * pow(xh, 5.0) / 5.0 + b3 * pow(xh, 3.0) / 3.0 + b1 * xh + b0 - 1.0;
        DTHV(s, j) = 1.0 / (b7 * pow(xj, 6.0) + b5 * pow(xj, 4.0) + b3 * pow(xj, 2.0) + b1);
        DTHU(s, j + 1) = 1.0 / (b7 * pow(xh, 6.0) + b5 * pow(xh, 4.0) + b3 * pow(xh, 2.0) + b1);
      }
      {
        const double xm = -dny - 0.5;
        YU(s, -dny) = b7 * pow(xm, 7.0) / 7.0 + b5 * pow(xm, 5.0) / 5.0 + b3 * pow(xm, 3.0) / 3.0 + b1 * xm + b0 - 1.0;
        DTHU(s, -dny) = 1.0 / (b7 * pow(xm, 6.0) + b5 * pow(xm, 4.0) + b3 * pow(xm, 2.0) + b1);
      }
      for (int j = dny1 + 1; j <= dny; ++j) {
        const double xj = j, xh = j + 0.5;
        YV(s, j + nn) = b7 * pow(xj, 7.0) / 7.0 + b5 * pow(xj, 5.0) / 5.0 + b3 * pow(xj, 3.0) / 3.0 + b1 * j - b0 + 1.0;
        YU(s, j + nn + 1) = b7 * pow(xh, 7.0) / 7.0 + b5 * pow(xh, 5.0) / 5.0 + b3 * pow(xh, 3.0) / 3.0 + b1 * xh - b0 + 1.0;
        DTHV(s, j + nn) = 1.0 / (b7 * pow(xj, 6.0) + b5 * pow(xj, 4.0) + b3 * pow(xj, 2.0) + b1);
        DTHU(s, j + nn + 1) = 1.0 / (b7 * pow(xh, 6.0) + b5 * pow(xh, 4.0) + b3 * pow(xh, 2.0) + b1);
      }
    }
  }
#undef YCH
#undef DCH
  /* 3-point second-derivative coefficients (y_grid_canopy.f90:217-248) */
  int j = -dny;
  DYU2(s, 1, j) = 0.0;
  DYU2(s, 2, j) = -2.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j + 1) - YU(s, j)));
  DYU2(s, 3, j) = 1.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j + 1) - YU(s, j)));
  DYV2(s, 1, j) = 0.0;
  DYV2(s, 2, j) = -2.0 / ((YV(s, j + 1) - YV(s, j)) * (YV(s, j + 1) - YV(s, j)));
  DYV2(s, 3, j) = 1.0 / ((YV(s, j + 1) - YV(s, j)) * (YV(s, j + 1) - YV(s, j)));
  for (j = -dny + 1; j <= nn + dny - 1; ++j) {
    DYU2(s, 1, j) = 2.0 / ((YU(s, j) - YU(s, j - 1)) * (YU(s, j + 1) - YU(s, j - 1)));
    DYU2(s, 2, j) = -2.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j) - YU(s, j - 1)));
    DYU2(s, 3, j) = 2.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j + 1) - YU(s, j - 1)));
    DYV2(s, 1, j) = 2.0 / ((YV(s, j) - YV(s, j - 1)) * (YV(s, j + 1) - YV(s, j - 1)));
    DYV2(s, 2, j) = -2.0 / ((YV(s, j + 1) - YV(s, j)) * (YV(s, j) - YV(s, j - 1)));
    DYV2(s, 3, j) = 2.0 / ((YV(s, j + 1) - YV(s, j)) * (YV(s, j + 1) - YV(s, j - 1)));
  }
  j = nn + dny;
  DYU2(s, 1, j) = 2.0 / ((YU(s, j) - YU(s, j - 1)) * (YU(s, j + 1) - YU(s, j - 1)));
  DYU2(s, 2, j) = -2.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j) - YU(s, j - 1)));
  DYU2(s, 3, j) = 2.0 / ((YU(s, j + 1) - YU(s, j)) * (YU(s, j + 1) - YU(s, j - 1)));
  DYV2(s, 1, j) = 1.0 / ((YV(s, j) - YV(s, j - 1)) * (YV(s, j) - YV(s, j - 1)));
  DYV2(s, 2, j) = -2.0 / ((YV(s, j) - YV(s, j - 1)) * (YV(s, j) - YV(s, j - 1)));
  DYV2(s, 3, j) = 0.0;
  j = nn + dny + 1;
  DYU2(s, 1, j) = 1.0 / ((YU(s, j) - YU(s, j - 1)) * (YU(s, j) - YU(s, j - 1)));
  DYU2(s, 2, j) = -2.0 / ((YU(s, j) - YU(s, j - 1)) * (YU(s, j) - YU(s, j - 1)));
  DYU2(s, 3, j) = 0.0;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* IB lists: boundary_circ_rough (geom_circ_rough_mixed_forcing_24pt.f90)    */
/* ------------------------------------------------------------------------ */
static double circ(double x, double z, double r, double ci, double ck) { /* :624-632 */
  return (x - ci) * (x - ci) + (z - ck) * (z - ck) - r * r;
}

/* lin_interp (:634-738); interpp = (z2, x2, z3, x3) */
static int lin_interp(double ie, double ke, int* interpp, double* wc, double r, double ci, double ck) {
  double aa = (ie >= ci) ? 1.0 : -1.0;
  double zb = ck + (ke - ck) * r / sqrt((ke - ck) * (ke - ck) + (ie - ci) * (ie - ci));
  double xb = aa * sqrt(r * r - (zb - ck) * (zb - ck)) + ci;
  double nz = (zb - ck) / sqrt((zb - ck) * (zb - ck) + (xb - ci) * (xb - ci));
  double nx = (xb - ci) / sqrt((zb - ck) * (zb - ck) + (xb - ci) * (xb - ci));
  int z2, x2, z3, x3, nzs, nxs;
  if (ie == ci && ke == ck) return 1; /* forcing point at the centre: normal undefined */
  if (ie == ci) {
    xb = ci;
    nz = (zb - ck) / sqrt((zb - ck) * (zb - ck) + (xb - ci) * (xb - ci));
    nzs = (int)(nz / fabs(nz));
    z2 = (int)(ke + nzs);
    x2 = (int)ie;
    z3 = (int)(ke + nzs);
    x3 = (int)ie;
    wc[0] = (ke - zb) / (z2 - zb);
    wc[1] = 0.0;
  } else if (ke == ck) {
    nxs = (int)(nx / fabs(nx));
    z2 = (int)ke;
    x2 = (int)(ie + nxs);
    z3 = (int)ke;
    x3 = (int)(ie + nxs);
    wc[0] = (ie - xb) / (x2 - xb);
    wc[1] = 0.0;
  } else {
    nzs = (int)(nz / fabs(nz));
    nxs = (int)(nx / fabs(nx));
    if (fabs(ie - ci) == fabs(ke - ck)) {
      z2 = (int)(ke + nzs); x2 = (int)ie; z3 = (int)ke; x3 = (int)(ie + nxs);
    } else if (fabs(nx) < fabs(nz)) {
      z2 = (int)(ke + nzs); x2 = (int)ie; z3 = (int)(ke + nzs); x3 = (int)(ie + nxs);
    } else if (fabs(nx) > fabs(nz)) {
      z2 = (int)ke; x2 = (int)(ie + nxs); z3 = (int)(ke + nzs); x3 = (int)(ie + nxs);
    } else {
      return 1;
    }
    const double den = xb * z2 - x2 * zb - xb * z3 + x3 * zb + (double)(x2 * z3) - (double)(x3 * z2);
    wc[0] = -(xb * z3 - x3 * zb - xb * ke + ie * zb + x3 * ke - ie * z3) / den;
    wc[1] = (xb * z2 - x2 * zb - xb * ke + ie * zb + x2 * ke - ie * z2) / den;
  }
  interpp[0] = z2; interpp[1] = x2; interpp[2] = z3; interpp[3] = x3;
  return 0;
}

typedef struct { int l[6]; double w[5]; } xzpt;

static int wrapi(int v, int d) { /* 1+mod(v-1d0+d, 1d0*d) for integer v */
  double a = (double)v - 1.0 + d;
  return (int)(1.0 + fmod(a, (double)d));
}
static void add_forcing(xzpt* e, int ie, int ke, const int* ip, const double* wc, int dnx, int dnz) {
  e->l[0] = wrapi(ie, dnx); e->l[1] = wrapi(ke, dnz);
  e->l[2] = wrapi(ip[1], dnx); e->l[3] = wrapi(ip[0], dnz);
  e->l[4] = wrapi(ip[3], dnx); e->l[5] = wrapi(ip[2], dnz);
  e->w[0] = wc[0]; e->w[1] = wc[1]; e->w[2] = 0; e->w[3] = 0; e->w[4] = 1.0;
}
static void add_solid(xzpt* e, int ie, int ke, int dnx, int dnz) {
  e->l[0] = e->l[2] = e->l[4] = wrapi(ie, dnx);
  e->l[1] = e->l[3] = e->l[5] = wrapi(ke, dnz);
  for (int q = 0; q < 5; ++q) e->w[q] = 0.0;
}

/* stable merge sort of indices by metric (MergeSort3D, mergeSort.f90:81-129) */
static void msort(long* key, int* idx, int n, long* tk, int* ti) {
  if (n < 2) return;
  int mid = n / 2;
  msort(key, idx, mid, tk, ti);
  msort(key + mid, idx + mid, n - mid, tk, ti);
  int a = 0, b = mid, c = 0;
  while (a < mid && b < n) {
    if (key[a] <= key[b]) { tk[c] = key[a]; ti[c++] = idx[a++]; }
    else { tk[c] = key[b]; ti[c++] = idx[b++]; }
  }
  while (a < mid) { tk[c] = key[a]; ti[c++] = idx[a++]; }
  while (b < n) { tk[c] = key[b]; ti[c++] = idx[b++]; }
  memcpy(key, tk, sizeof(long) * n);
  memcpy(idx, ti, sizeof(int) * n);
}

static void sort_list(int* l9, double* w3, int n) {
  if (n <= 0) return;
  int imin = l9[0], imax = l9[0], kmin = l9[1], kmax = l9[1];
  for (int q = 0; q < n; ++q) {
    if (l9[9 * q] < imin) imin = l9[9 * q];
    if (l9[9 * q] > imax) imax = l9[9 * q];
    if (l9[9 * q + 1] < kmin) kmin = l9[9 * q + 1];
    if (l9[9 * q + 1] > kmax) kmax = l9[9 * q + 1];
  }
  const long si = imax - imin + 1, sk = kmax - kmin + 1;
  long* key = (long*)malloc(sizeof(long) * n * 2);
  int* idx = (int*)malloc(sizeof(int) * n * 2);
  for (int q = 0; q < n; ++q) {
    key[q] = si * sk * l9[9 * q + 2] + si * l9[9 * q + 1] + l9[9 * q];
    idx[q] = q;
  }
  msort(key, idx, n, key + n, idx + n);
  int* l2 = (int*)malloc(sizeof(int) * 9 * n);
  double* w2 = (double*)malloc(sizeof(double) * 3 * n);
  for (int q = 0; q < n; ++q) {
    memcpy(l2 + 9 * q, l9 + 9 * idx[q], 9 * sizeof(int));
    memcpy(w2 + 3 * q, w3 + 3 * idx[q], 3 * sizeof(double));
  }
  memcpy(l9, l2, sizeof(int) * 9 * n);
  memcpy(w3, w2, sizeof(double) * 3 * n);
  free(key); free(idx); free(l2); free(w2);
}

static int build_ib(or_state* s, char* err, int errlen) {
  const or_params* P = &s->P;
  for (int g = 0; g < 2; ++g)
    for (int w = 0; w < 2; ++w) { s->nib[g][w] = 0; s->ibl[g][w] = NULL; s->ibw[g][w] = NULL; }
  if (P->geometry_type != 3) return 0;
  const int dsty = P->dsty, ny = s->ny;
  /* y extents (:53-62) with Ngal(4,3) = ny-1+dsty and Ngal(3,3) = ny-2+dsty */
  const int nyu11 = -dsty + 1, nyu21 = 0, nyv11 = -dsty + 1, nyv21 = -1;
  const int nyu12 = (ny - 1 + dsty) - dsty + 1, nyu22 = ny - 1 + dsty;
  const int nyv12 = (ny - 2 + dsty) + 1 - dsty + 1, nyv22 = ny - 2 + dsty;
  if (P->ntilex == 0 || P->ntilez == 0) { snprintf(err, errlen, "ntilex/ntilez equals 0"); return 1; }
  if (P->ntilex % 2 != 0) { snprintf(err, errlen, "ntilex is not even"); return 1; }
  if (s->ngx % P->ntilex != 0) { snprintf(err, errlen, "nx not multiple of ntilex"); return 1; }
  if (s->ngz % P->ntilez != 0) { snprintf(err, errlen, "nz not multiple of ntilez"); return 1; }
  const int dnx = s->ngx / P->ntilex, dnz = s->ngz / P->ntilez;
  const double r = P->r, ci = P->c_i, ck = P->c_k;
  const int dx1 = (int)floor(ci - r - 1), dx2 = (int)ceil(ci + r + 1);
  const int dz1 = (int)floor(ck - r - 1), dz2 = (int)ceil(ck + r + 1);
  if (dx1 < 1 || dz1 < 1 || dx2 > dnx || dz2 > dnz) { snprintf(err, errlen, "Circ too close to edge"); return 1; }
  const int W = dx2 - dx1 + 1;
  double* tc = (double*)malloc(sizeof(double) * W * (dz2 - dz1 + 1));
#define TC(i, k) tc[((k) - dz1) * W + ((i) - dx1)]
  for (int k = dz1; k <= dz2; ++k)
    for (int i = dx1; i <= dx2; ++i) TC(i, k) = circ((double)i, (double)k, r, ci, ck);
  int cap = 2 * W * (dz2 - dz1 + 1) + 8, nxz = 0;
  xzpt* xz = (xzpt*)calloc(cap, sizeof(xzpt));
  int ip[4];
  double wc[2];
  for (int k = dz1 + 1; k <= dz2 - 1; ++k) {
    for (int i = dx1 + 1; i <= dx2 - 1; ++i) {
      if (TC(i, k) < 0.0) {
        const int rig = TC(i - 1, k) > 0.0, lef = TC(i + 1, k) > 0.0;
        const int bot = TC(i, k + 1) > 0.0, top = TC(i, k - 1) > 0.0;
        if (rig || lef || bot || top) {
          if (lin_interp(i, k, ip, wc, r, ci, ck)) { snprintf(err, errlen, "lin_interp: degenerate normal"); free(tc); free(xz); return 1; }
          if (wc[0] < -1.0 || wc[1] < -1.0) {
            add_solid(&xz[nxz++], i, k, dnx, dnz);
            const int io = ip[1], ko = ip[0];
            if (lin_interp(io, ko, ip, wc, r, ci, ck)) { snprintf(err, errlen, "lin_interp: degenerate normal"); free(tc); free(xz); return 1; }
            add_forcing(&xz[nxz++], io, ko, ip, wc, dnx, dnz);
          } else {
            add_forcing(&xz[nxz++], i, k, ip, wc, dnx, dnz);
          }
        } else {
          add_solid(&xz[nxz++], i, k, dnx, dnz);
        }
      }
      if (TC(i, k) == 0.0) add_solid(&xz[nxz++], i, k, dnx, dnz);
    }
  }
  if (P->extras24) { /* (:219-259) */
    const int ex[4][2] = {{8, 16}, {16, 16}, {16, 8}, {8, 8}};
    for (int q = 0; q < 4; ++q) {
      if (lin_interp(ex[q][0], ex[q][1], ip, wc, r, ci, ck)) { snprintf(err, errlen, "lin_interp: degenerate normal"); free(tc); free(xz); return 1; }
      add_forcing(&xz[nxz++], ex[q][0], ex[q][1], ip, wc, dnx, dnz);
    }
  }
#undef TC
  free(tc);
  /* stems (:272-459): grid g, wall w; plane list and which plane gets Lap=1 */
  for (int g = 0; g < 2; ++g) {
    for (int w = 0; w < 2; ++w) {
      int j0, j1, jf;
      if (g == 0) { j0 = w ? nyu12 : nyu11; j1 = w ? nyu22 : nyu21; jf = w ? nyu12 : nyu21; }
      else { j0 = w ? nyv12 : nyv11; j1 = w ? nyv22 : nyv21; jf = w ? nyv12 : nyv21; }
      const int nplanes = j1 - j0 + 1;
      const int pst = nplanes * nxz;
      const int ntot = P->ntilex * P->ntilez * pst;
      int* l9 = (int*)malloc(sizeof(int) * 9 * (ntot > 0 ? ntot : 1));
      double* w3 = (double*)malloc(sizeof(double) * 3 * (ntot > 0 ? ntot : 1));
      /* first stem: bottom writes the ordinary planes first and the Lap=1
         plane last; top writes the Lap=1 plane first */
      int il = 0;
      for (int pass = 0; pass < 2; ++pass) {
        for (int j = j0; j <= j1; ++j) {
          const int isf = (j == jf);
          if ((w == 0 && pass == 0 && isf) || (w == 0 && pass == 1 && !isf)) continue;
          if ((w == 1 && pass == 0 && !isf) || (w == 1 && pass == 1 && isf)) continue;
          for (int q = 0; q < nxz; ++q) {
            int* L = l9 + 9 * il;
            L[0] = xz[q].l[0]; L[1] = xz[q].l[1]; L[2] = j;
            L[3] = xz[q].l[2]; L[4] = xz[q].l[3]; L[5] = j;
            L[6] = xz[q].l[4]; L[7] = xz[q].l[5]; L[8] = j;
            w3[3 * il] = isf ? 1.0 : xz[q].w[4];
            w3[3 * il + 1] = xz[q].w[0];
            w3[3 * il + 2] = xz[q].w[1];
            ++il;
          }
        }
      }
      /* tile replication with stagger (:462-577) */
      for (int ix = 1; ix <= P->ntilex; ++ix) {
        for (int iz = 1; iz <= P->ntilez; ++iz) {
          const int sh = pst * (iz - 1) + pst * P->ntilez * (ix - 1);
          const double stg = ((ix + 1) % 2) * dnz * P->stag;
          for (int q = 0; q < pst; ++q) {
            int* D = l9 + 9 * (q + sh);
            const int* S = l9 + 9 * q;
            for (int c = 0; c < 3; ++c) {
              const int ii = S[3 * c], kk = S[3 * c + 1], jj = S[3 * c + 2];
              D[3 * c] = ii + dnx * (ix - 1);
              D[3 * c + 1] = (int)(1.0 + fmod((double)(kk + dnz * (iz - 1)) + stg - 1.0, 1.0 * s->ngz));
              D[3 * c + 2] = jj;
            }
            if (sh) memcpy(w3 + 3 * (q + sh), w3 + 3 * q, 3 * sizeof(double));
          }
        }
      }
      sort_list(l9, w3, ntot);
      s->nib[g][w] = ntot;
      s->ibl[g][w] = l9;
      s->ibw[g][w] = w3;
    }
  }
  free(xz);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* create / destroy                                                          */
/* ------------------------------------------------------------------------ */
static void lu_build_p(or_state* s);
static int central_plane(const or_state* s, int g, int j);

or_state* or_create(const or_params* p, char* err, int errlen) {
  char dummy[8];
  if (!err) { err = dummy; errlen = sizeof dummy; }
  err[0] = 0;
  if (p->ngalx < 4 || p->ngalz < 4 || p->ny < 4 || p->dsty < 0) { snprintf(err, errlen, "bad sizes"); return NULL; }
  if (p->ngalx % 2 || p->ngalz % 2) { snprintf(err, errlen, "Ngal must be even"); return NULL; }
  or_state* s = (or_state*)calloc(1, sizeof(or_state));
  s->P = *p;
  s->ngx = p->ngalx; s->ngz = p->ngalz; s->ny = p->ny; s->dsty = p->dsty;
  s->N1 = 2 * (s->ngx / 3); s->N2 = 2 * (s->ngz / 3);      /* start.f90:108-111 */
  s->Mx = s->N1 / 2 + 1; s->Ncol = s->Mx * s->N2;          /* start.f90:1423 */
  s->jlo_u = -p->dsty; s->jhi_u = p->ny + p->dsty; s->nyu = s->jhi_u - s->jlo_u + 1;
  s->jlo_v = -p->dsty; s->jhi_v = p->ny - 1 + p->dsty; s->nyv = s->jhi_v - s->jlo_v + 1;
  s->jlo_p = -p->dsty + 1; s->jhi_p = p->ny - 1 + p->dsty; s->nyp = s->jhi_p - s->jlo_p + 1;
  s->nn = p->ny - 2;
  s->yu = calloc(s->nyu, sizeof(double)); s->dthdyu = calloc(s->nyu, sizeof(double));
  s->dyu2i = calloc(3 * s->nyu, sizeof(double));
  s->yv = calloc(s->nyv, sizeof(double)); s->dthdyv = calloc(s->nyv, sizeof(double));
  s->dyv2i = calloc(3 * s->nyv, sizeof(double));
  if (build_ygrid(s, err, errlen) || build_ib(s, err, errlen)) { or_destroy(s); return NULL; }
  s->ddthetavi = 1.0; s->dthetavi = 1.0;
  /* gridweighting(midband,.), gridweighting_interp(midband,.) (start.f90:494-508) */
  const int N40 = -p->dsty, N30 = -p->dsty, N43 = p->ny - 1 + p->dsty, N33 = p->ny - 2 + p->dsty;
  s->gw[0] = -(YV(s, N30) - YU(s, N40)) / (YV(s, N30) - YU(s, N40 + 1));
  s->gw[1] = (YU(s, N43 + 1) - YV(s, N33 + 1)) / (YV(s, N33 + 1) - YU(s, N43));
  s->gwi[0] = (YU(s, N40) - YU(s, N40 + 1)) / (YV(s, N30) - YU(s, N40 + 1));
  s->gwi[1] = (YU(s, N43 + 1) - YU(s, N43)) / (YV(s, N33 + 1) - YU(s, N43));
  /* wavenumbers (def_k, start.f90:965-986); k1 stored as its imaginary part */
  s->k1x = calloc(s->Mx, sizeof(double)); s->k2x = calloc(s->Mx, sizeof(double));
  s->k1z = calloc(s->N2, sizeof(double)); s->k2z = calloc(s->N2, sizeof(double));
  for (int i = 0; i < s->Mx; ++i) { s->k1x[i] = p->alp * i; s->k2x[i] = -((p->alp * i) * (p->alp * i)); }
  for (int k = 1; k <= s->N2; ++k) {
    double kz = (k <= s->N2 / 2) ? p->bet * (k - 1) : -(p->bet * (s->N2 - k + 1));
    s->k1z[k - 1] = kz;
    s->k2z[k - 1] = -(kz * kz);
  }
  /* RK3 (start.f90:395-409) */
  const double a[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0}, g[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double c[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0}, d[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
  for (int q = 0; q < 3; ++q) { s->aRK[q] = a[q]; s->bRK[q] = a[q]; s->gRK[q] = g[q]; s->cRK[q] = c[q]; s->dRK[q] = d[q]; }
  /* dtv (start.f90:432) uses k2F_z(N2/2), i.e. kz = N2/2-1 */
  s->dtv = p->Re / (-s->k2x[s->N1 / 2] - s->k2z[s->N2 / 2 - 1] +
                    4.0 / ((YV(s, N30 + 1) - YV(s, N30)) * (YV(s, N30 + 1) - YV(s, N30))));
  s->mpgx = p->mpgx;
  const size_t su = (size_t)s->Ncol * s->nyu, sv = (size_t)s->Ncol * s->nyv, sp = (size_t)s->Ncol * s->nyp;
  const size_t sz[12] = {su, sv, su, sp, su, sv, su, su, sv, su, sp, sp};
  for (int q = 0; q < 12; ++q) s->F[q] = (cplx*)calloc(sz[q], sizeof(cplx));
  s->u1i = calloc(sv, sizeof(cplx)); s->u2i = calloc(su, sizeof(cplx)); s->u3i = calloc(sv, sizeof(cplx));
  s->uvf = calloc(sv, sizeof(cplx)); s->vvc = calloc(su, sizeof(cplx)); s->vwf = calloc(sv, sizeof(cplx));
  s->DG = calloc(3 * sp, sizeof(double));
  fplan_init(&s->pz, s->ngz);
  fplan_init(&s->px2, s->ngx / 2);
  s->twx = malloc(sizeof(cplx) * (s->ngx / 2 + 1));
  for (int k = 0; k <= s->ngx / 2; ++k) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / (long double)s->ngx;
    s->twx[k] = (double)cosl(ang) + I * (double)sinl(ang);
  }
  /* band table y limits (y_grid_canopy.f90:250-275), the bottom band h_ny
     above the lower tips (start.f90:199-226: 0.25 for smooth walls, two post
     spacings for canopies) */
  {
    const double h = p->geometry_type == 3 ? 2.0 * (2.0 * 3.14159265358979323846 / p->bet) / p->ntilez : 0.25;
    const double target = YU(s, 0) + h;
    int j = -p->dsty;
    while (j < s->jhi_u && YU(s, j) < target) ++j;
    s->N3[0] = -p->dsty; s->N3[1] = j; s->N3[2] = (p->ny - 2) - j; s->N3[3] = p->ny - 2 + p->dsty;
    s->N4[0] = -p->dsty; s->N4[1] = j + 1; s->N4[2] = s->N3[2]; s->N4[3] = s->N3[3] + 1;
  }
  s->ngx2 = p->ngalx2 > 0 ? p->ngalx2 : s->ngx;
  s->ngz2 = p->ngalz2 > 0 ? p->ngalz2 : s->ngz;
  s->N1c = 2 * (s->ngx2 / 3); s->N2c = 2 * (s->ngz2 / 3);
  s->mb = (s->ngx2 != s->ngx || s->ngz2 != s->ngz);
  s->shortc = (unsigned char*)calloc(s->Ncol, 1);
  if (s->mb) {
    if (s->ngx2 > s->ngx || s->ngz2 > s->ngz || s->ngx2 < 4 || s->ngz2 < 4 || s->ngx2 % 2 || s->ngz2 % 2) {
      snprintf(err, errlen, "central band resolution must be even, at least 4 and at most the wall bands'");
      or_destroy(s); return NULL;
    }
    /* each band needs its own rows: the short columns' interface ghost rows
       N(4,1)+1 and N(4,2) must differ */
    if (!(s->N3[0] < s->N3[1] && s->N3[1] + 2 < s->N3[2] && s->N3[2] < s->N3[3])) {
      snprintf(err, errlen, "band limits leave no central band");
      or_destroy(s); return NULL;
    }
    for (int g = 0; g < 2; ++g)
      for (int w = 0; w < 2; ++w)
        for (int q = 0; q < s->nib[g][w]; ++q) {
          const int j = s->ibl[g][w][9 * q + 2];
          if (central_plane(s, g, j)) { snprintf(err, errlen, "forcing planes in the central band"); or_destroy(s); return NULL; }
        }
    /* long columns: i <= N(1,2)/2 and kz in the central band's N(2,2) modes
       (proc_lims_columns, start.f90:1424-1501) */
    for (int col = 0; col < s->Ncol; ++col) {
      const int i = col % s->Mx, k = col / s->Mx;
      const int kin = k < s->N2c / 2 || k >= s->N2 - s->N2c / 2;
      s->shortc[col] = !(i <= s->N1c / 2 && kin);
    }
    s->gw12 = (YU(s, s->N4[1] + 1) - YV(s, s->N3[1] + 1)) / (YV(s, s->N3[1] + 1) - YU(s, s->N4[1]));
    s->gw31 = -(YV(s, s->N3[2]) - YU(s, s->N4[2])) / (YV(s, s->N3[2]) - YU(s, s->N4[2] + 1));
    s->gwi12 = (YU(s, s->N4[1] + 1) - YU(s, s->N4[1])) / (YV(s, s->N3[1] + 1) - YU(s, s->N4[1]));
    s->gwi31 = (YU(s, s->N4[2]) - YU(s, s->N4[2] + 1)) / (YV(s, s->N3[2]) - YU(s, s->N4[2] + 1));
    fplan_init(&s->pz2, s->ngz2);
    fplan_init(&s->px22, s->ngx2 / 2);
    s->twx2 = malloc(sizeof(cplx) * (s->ngx2 / 2 + 1));
    for (int k = 0; k <= s->ngx2 / 2; ++k) {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * k / (long double)s->ngx2;
      s->twx2[k] = (double)cosl(ang) + I * (double)sinl(ang);
    }
  }
  lu_build_p(s);
  return s;
}

void or_band_info(const or_state* s, int* b, unsigned char* shortc) {
  b[0] = s->mb; b[1] = s->N1c; b[2] = s->N2c;
  for (int q = 0; q < 4; ++q) { b[3 + q] = s->N3[q]; b[7 + q] = s->N4[q]; }
  if (shortc) memcpy(shortc, s->shortc, s->Ncol);
}

static void or_stats_free(or_state* s);

void or_destroy(or_state* s) {
  if (!s) return;
  free(s->yu); free(s->yv); free(s->dthdyu); free(s->dthdyv); free(s->dyu2i); free(s->dyv2i);
  free(s->k1x); free(s->k2x); free(s->k1z); free(s->k2z);
  for (int q = 0; q < 12; ++q) free(s->F[q]);
  free(s->u1i); free(s->u2i); free(s->u3i); free(s->uvf); free(s->vvc); free(s->vwf);
  free(s->DG);
  for (int g = 0; g < 2; ++g) for (int w = 0; w < 2; ++w) { free(s->ibl[g][w]); free(s->ibw[g][w]); }
  if (s->pz.tw) fplan_free(&s->pz);
  if (s->px2.tw) fplan_free(&s->px2);
  free(s->twx);
  or_stats_free(s);
  free(s->clim); free(s->xz1); free(s->xz3); free(s->xz2);
  free(s->shortc);
  if (s->pz2.tw) fplan_free(&s->pz2);
  if (s->px22.tw) fplan_free(&s->px22);
  free(s->twx2);
  free(s);
}

void or_dims(const or_state* s, int* d) {
  d[0] = s->N1; d[1] = s->N2; d[2] = s->Mx; d[3] = s->Ncol; d[4] = s->nyu; d[5] = s->nyv; d[6] = s->nyp;
  d[7] = s->jlo_u; d[8] = s->jlo_v; d[9] = s->jlo_p; d[10] = s->ngx; d[11] = s->ngz; d[12] = s->ny; d[13] = s->dsty;
  d[14] = s->nib[0][0] + s->nib[0][1]; d[15] = s->nib[1][0] + s->nib[1][1];
}
void or_get_ygrid(const or_state* s, double* yu, double* yv, double* a, double* b, double* c, double* d) {
  memcpy(yu, s->yu, sizeof(double) * s->nyu); memcpy(yv, s->yv, sizeof(double) * s->nyv);
  memcpy(a, s->dthdyu, sizeof(double) * s->nyu); memcpy(b, s->dthdyv, sizeof(double) * s->nyv);
  memcpy(c, s->dyu2i, sizeof(double) * 3 * s->nyu); memcpy(d, s->dyv2i, sizeof(double) * 3 * s->nyv);
}
void or_get_metrics(const or_state* s, double* gw, double* gwi, double* dtv) {
  gw[0] = s->gw[0]; gw[1] = s->gw[1]; gwi[0] = s->gwi[0]; gwi[1] = s->gwi[1]; *dtv = s->dtv;
}
int or_get_iblist(const or_state* s, int grid, int wall, int* l9, double* w3) {
  const int n = s->nib[grid][wall];
  if (l9 && n) memcpy(l9, s->ibl[grid][wall], sizeof(int) * 9 * n);
  if (w3 && n) memcpy(w3, s->ibw[grid][wall], sizeof(double) * 3 * n);
  return n;
}
void or_get_dg(const or_state* s, double* dg) { memcpy(dg, s->DG, sizeof(double) * 3 * (size_t)s->Ncol * s->nyp); }

static size_t fsize(const or_state* s, int which) {
  static const int grid[12] = {0, 1, 0, 2, 0, 1, 0, 0, 1, 0, 2, 2};
  const int g = grid[which];
  return (size_t)s->Ncol * (g == 0 ? s->nyu : g == 1 ? s->nyv : s->nyp);
}
int or_set_field(or_state* s, int w, const double* src) {
  if (w < 0 || w > 11) return 1;
  memcpy(s->F[w], src, sizeof(cplx) * fsize(s, w));
  return 0;
}
int or_get_field(const or_state* s, int w, double* dst) {
  if (w < 0 || w > 11) return 1;
  memcpy(dst, s->F[w], sizeof(cplx) * fsize(s, w));
  return 0;
}
void or_set_scalars(or_state* s, const double* v) {
  s->t = v[0]; s->dt = v[1]; s->dti = v[2]; s->mpgx = v[3]; s->dgx = v[4]; s->QxT = v[5]; s->Qx = v[6]; s->dtc = v[7];
}
void or_get_scalars(const or_state* s, double* v) {
  v[0] = s->t; v[1] = s->dt; v[2] = s->dti; v[3] = s->mpgx; v[4] = s->dgx; v[5] = s->QxT; v[6] = s->Qx; v[7] = s->dtc;
}

/* ------------------------------------------------------------------------ */
/* spectral y operators                                                      */
/* ------------------------------------------------------------------------ */
#define U1 (s->F[0])
#define U2 (s->F[1])
#define U3 (s->F[2])
#define PP (s->F[3])
#define NU1 (s->F[4])
#define NU2 (s->F[5])
#define NU3 (s->F[6])
#define DU1 (s->F[7])
#define DU2 (s->F[8])
#define DU3 (s->F[9])
#define PSI (s->F[10])
#define DIV (s->F[11])

static inline cplx imul(double k, cplx u) { /* (0,k)*u as Fortran complex product */
  return (0.0 * creal(u) - k * cimag(u)) + I * (0.0 * cimag(u) + k * creal(u));
}

/* RHS0_u1 / RHS0_u3 with laplacian_U (littleharsh.f90:246-282, 613-699) */
static void rhs0_u(or_state* s, cplx* du, const cplx* u, const cplx* Nu, int comp, int kRK) {
  const double C1 = s->aRK[kRK - 1] / s->P.Re, C2 = -s->cRK[kRK - 1];
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    const double k2x = s->k2x[i], k2z = s->k2z[k];
    const double kk = (comp == 1) ? s->k1x[i] : s->k1z[k];
    const double c3 = -s->dRK[kRK - 1] * kk; /* C3 = -dRK*k1F (imaginary) */
    int lo[2], hi[2];
    const int ns = col_segs(s, col, 0, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) {
        const cplx Lu = DYU2(s, 1, j) * u[fu(s, col, j - 1)] + (k2x + k2z + DYU2(s, 2, j)) * u[fu(s, col, j)] +
                        DYU2(s, 3, j) * u[fu(s, col, j + 1)];
        du[fu(s, col, j)] = C1 * Lu + C2 * Nu[fu(s, col, j)] + imul(c3, PP[fp(s, col, j)]);
      }
  }
  if (comp == 1) {
    const double C = -s->dRK[kRK - 1] * s->mpgx;
    for (int j = s->jlo_u + 1; j <= s->jhi_u - 1; ++j) du[fu(s, 0, j)] += C;
  }
}

/* RHS0_u2 with laplacian_V (littleharsh.f90:284-320, 660-700) */
static void rhs0_v(or_state* s, int kRK) {
  const double C1 = s->aRK[kRK - 1] / s->P.Re, C2 = -s->cRK[kRK - 1];
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    const double k2x = s->k2x[i], k2z = s->k2z[k];
    int lo[2], hi[2];
    const int ns = col_segs(s, col, 1, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) {
        const double C3 = -s->dRK[kRK - 1] * s->ddthetavi * DTHV(s, j);
        const cplx Lv = DYV2(s, 1, j) * U2[fv(s, col, j - 1)] + (k2x + k2z + DYV2(s, 2, j)) * U2[fv(s, col, j)] +
                        DYV2(s, 3, j) * U2[fv(s, col, j + 1)];
        DU2[fv(s, col, j)] = C1 * Lv + C2 * NU2[fv(s, col, j)] + C3 * (PP[fp(s, col, j + 1)] - PP[fp(s, col, j)]);
      }
  }
}

void or_rhs0(or_state* s, int kRK) {
  rhs0_u(s, DU1, U1, NU1, 1, kRK);
  rhs0_v(s, kRK);
  rhs0_u(s, DU3, U3, NU3, 3, kRK);
}

/* ------------------------------------------------------------------------ */
/* plane transforms                                                          */
/* ------------------------------------------------------------------------ */
typedef struct { cplx *w0, *w1; } scratch;

/* the resolution a plane is transformed at: the wall bands' (fine) or, with
   multi-band resolution, the central band's */
typedef struct {
  int ngx, ngz, N1, N2, coarse;
  const fplan *pz, *px2;
  const cplx* twx;
} pview;
static pview view_of(const or_state* s, int coarse) {
  pview v;
  if (coarse) { v.ngx = s->ngx2; v.ngz = s->ngz2; v.N1 = s->N1c; v.N2 = s->N2c; v.pz = &s->pz2; v.px2 = &s->px22; v.twx = s->twx2; }
  else { v.ngx = s->ngx; v.ngz = s->ngz; v.N1 = s->N1; v.N2 = s->N2; v.pz = &s->pz; v.px2 = &s->px2; v.twx = s->twx; }
  v.coarse = coarse;
  return v;
}
/* plane j of the u (g 0), v (g 1) or p (g 2) grid lies in the central band
   (planes N(4,1)+1..N(4,2), N(3,1)+1..N(3,2), N(4,1)..N(4,2)+1: the plane
   ranges of the ranks of band 2 in proc_lims_planes, start.f90:1650-1951) */
static int central_plane(const or_state* s, int g, int j) {
  if (!s->mb) return 0;
  if (g == 2) return j >= s->N4[1] && j <= s->N4[2] + 1;
  return g == 0 ? (j > s->N4[1] && j <= s->N4[2]) : (j > s->N3[1] && j <= s->N3[2]);
}
/* slot of reference mode k (fine column numbering, 0-based) in the view's
   z transform (modes_to_planes with dk, start.f90:1561-1573) */
static inline int vslot(const or_state* s, const pview* v, int k) { return (k < s->N2 / 2) ? k : k - s->N2 + v->ngz; }

/* modes_to_planes: column values of plane j into xPL (zero elsewhere)
   (FOU3D.f90:825-842; k slot = k - dk, start.f90:1561-1573); a central
   plane takes the long columns only */
static void to_plane_v(const or_state* s, const pview* v, const cplx* f, int ny_g, int jlo_g, int j, double* P) {
  const int igal = v->ngx + 2;
  memset(P, 0, sizeof(double) * igal * v->ngz);
  for (int col = 0; col < s->Ncol; ++col) {
    if (v->coarse && s->shortc[col]) continue;
    const int i = col % s->Mx, k = col / s->Mx; /* k 0-based */
    const int slot = vslot(s, v, k);
    const cplx val = f[(size_t)col * ny_g + (j - jlo_g)];
    P[slot * igal + 2 * i] = creal(val);
    P[slot * igal + 2 * i + 1] = cimag(val);
  }
}
static inline int slot_of(const or_state* s, int k) { return (k < s->N2 / 2) ? k : k - s->N2 + s->ngz; }

/* four_to_phys_u (zero_nyq=1) / four_to_phys_du (FOU3D.f90:546-592) */
static void four_to_phys_v(const pview* v, double* P, int zero_nyq, scratch* w) {
  const int igal = v->ngx + 2;
  if (zero_nyq) memset(P + (v->ngz / 2) * igal, 0, sizeof(double) * igal);
  for (int i = 0; i < (v->N1 + 2) / 2; ++i) cft_line(v->pz, P + 2 * i, igal, +1, w->w0, w->w1);
  for (int k = 0; k < v->ngz; ++k) rft_line(v->px2, v->twx, P + k * igal, +1, w->w0, w->w1);
}
/* phys_to_four_du (FOU3D.f90:1619-1639) */
static void phys_to_four_v(const pview* v, double* P, scratch* w) {
  const int igal = v->ngx + 2;
  for (int k = 0; k < v->ngz; ++k) rft_line(v->px2, v->twx, P + k * igal, -1, w->w0, w->w1);
  for (int i = 0; i < (v->N1 + 2) / 2; ++i) cft_line(v->pz, P + 2 * i, igal, -1, w->w0, w->w1);
}
static inline cplx pl_v(const pview* v, const double* P, int i, int slot) {
  const int igal = v->ngx + 2;
  return P[slot * igal + 2 * i] + I * P[slot * igal + 2 * i + 1];
}
/* the same at the wall bands' resolution */
static void to_plane(const or_state* s, const cplx* f, int ny_g, int jlo_g, int j, double* P) {
  const pview v = view_of(s, 0);
  to_plane_v(s, &v, f, ny_g, jlo_g, j, P);
}
static void four_to_phys(const or_state* s, double* P, int zero_nyq, scratch* w) {
  const pview v = view_of(s, 0);
  four_to_phys_v(&v, P, zero_nyq, w);
}
static void phys_to_four(const or_state* s, double* P, scratch* w) {
  const pview v = view_of(s, 0);
  phys_to_four_v(&v, P, w);
}
static inline cplx pl(const or_state* s, const double* P, int i, int slot) {
  const pview v = view_of(s, 0);
  return pl_v(&v, P, i, slot);
}

/* rows of column col on grid g (0 u, 1 v, 2 p), inclusive, boundary rows
   included: one segment for a long column, two for a short one, the wall
   bands' jlim (start.f90:1536-1557); returns the segment count */
static int col_segs(const or_state* s, int col, int g, int lo[2], int hi[2]) {
  const int jl = g == 0 ? s->jlo_u : (g == 1 ? s->jlo_v : s->jlo_p);
  const int jh = g == 0 ? s->jhi_u : (g == 1 ? s->jhi_v : s->jhi_p);
  lo[0] = jl;
  if (!s->mb || !s->shortc[col]) {
    hi[0] = jh;
    return 1;
  }
  hi[0] = g == 0 ? s->N4[1] + 1 : (g == 1 ? s->N3[1] + 1 : s->N4[1]);
  lo[1] = g == 0 ? s->N4[2] : (g == 1 ? s->N3[2] : s->N4[2] + 1);
  hi[1] = jh;
  return 2;
}
/* gridweighting of the u-grid boundary row at the start (lo) or end of a
   segment: the wall's, or the band interface's (start.f90:494-500) */
static double gw_lo(const or_state* s, int lo) { return lo == s->jlo_u ? s->gw[0] : s->gw31; }
static double gw_hi(const or_state* s, int hi) { return hi == s->jhi_u ? s->gw[1] : s->gw12; }
static double gwi_lo(const or_state* s, int lo) { return lo == s->jlo_u ? s->gwi[0] : s->gwi31; }
static double gwi_hi(const or_state* s, int hi) { return hi == s->jhi_u ? s->gwi[1] : s->gwi12; }

/* interp_u (FOU3D.f90:261-281): u/w centres -> v faces, all faces */
static void interp_u(or_state* s, cplx* ui, const cplx* u) {
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    int lo[2], hi[2];
    const int ns = col_segs(s, col, 1, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g]; j <= hi[g]; ++j)
        ui[fv(s, col, j)] = ((YV(s, j) - YU(s, j)) * u[fu(s, col, j + 1)] + (YU(s, j + 1) - YV(s, j)) * u[fu(s, col, j)]) /
                            (YU(s, j + 1) - YU(s, j));
  }
}
/* interp_v (FOU3D.f90:283-308): v faces -> centres, ghosts extrapolated */
static void interp_v(or_state* s, cplx* ui, const cplx* v) {
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    int lo[2], hi[2], vlo[2], vhi[2];
    const int ns = col_segs(s, col, 0, lo, hi);
    col_segs(s, col, 1, vlo, vhi);
    for (int g = 0; g < ns; ++g) {
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j)
        ui[fu(s, col, j)] = ((YU(s, j) - YV(s, j - 1)) * v[fv(s, col, j)] + (YV(s, j) - YU(s, j)) * v[fv(s, col, j - 1)]) /
                            (YV(s, j) - YV(s, j - 1));
      const cplx a = ui[fu(s, col, lo[g] + 1)], b = ui[fu(s, col, hi[g] - 1)];
      ui[fu(s, col, lo[g])] = gwi_lo(s, lo[g]) * (v[fv(s, col, vlo[g])] - a) + a;
      ui[fu(s, col, hi[g])] = gwi_hi(s, hi[g]) * (v[fv(s, col, vhi[g])] - b) + b;
    }
  }
}

static double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* the as-shipped nonlinear term of a run on np ranks                        */
/* ------------------------------------------------------------------------ */
/* proc_lims_planes for one resolution in every band (start.f90:1652-1866):
   procs per band from the load balance, then planelim per rank; u grid over
   N(4,.), v grid over N(3,.).  lim[(g*2+e)*np + r], g 0 u, 1 v, e 0 from,
   1 to.  1 when the balance cannot run (np < 3), 3 when the band limits are
   not strictly increasing. */
#define NBT(r, b) (N20[4 * (b) + (r) - 1])
int or_plane_partition(const int* N20, int ngalx, int ngalz, int np, int* lim) {
  if (np < 3) return 1;
  /* band limits not strictly increasing: an empty or negative band, ranges
     that overlap */
  for (int row = 3; row <= 4; ++row)
    if (!(NBT(row, 0) < NBT(row, 1) && NBT(row, 1) < NBT(row, 2) && NBT(row, 2) < NBT(row, 3))) return 3;
  enum { bot = 1, mid = 2, top = 3, nband = 3 };
  double load_band[4] = {0}, i_load[4] = {0}, maxload;
  int procs[4] = {0}, procs_b[4], nplanes, iplanes, rem;
  for (int iband = 1; iband <= nband; ++iband)
    load_band[iband] = (double)(ngalx + 2) * ngalz * (NBT(4, iband) - NBT(4, iband - 1));
#define BALANCE()                                                                         \
  do {                                                                                    \
    rem = 0;                                                                              \
    nplanes = NBT(4, bot) - NBT(4, 0);                                                    \
    procs[bot] = (int)ceil(nplanes * 1.0 / iplanes);                                      \
    procs[top] = procs[bot];                                                              \
    rem = rem + procs[bot] + procs[top];                                                  \
    i_load[bot] = load_band[bot] / nplanes * iplanes;                                     \
    i_load[top] = load_band[top] / nplanes * iplanes;                                     \
    nplanes = NBT(4, mid) - NBT(4, bot);                                                  \
    procs[mid] = np - rem;                                                                \
    if (procs[mid] > 0) i_load[mid] = load_band[mid] / nplanes * ceil(nplanes * 1.0 / procs[mid]); \
  } while (0)
#define MAXLOAD() fmax(i_load[1], fmax(i_load[2], i_load[3]))
  iplanes = 1; /* one rank per fine plane first */
  BALANCE();
  while (procs[mid] <= 0) { iplanes = iplanes + 1; BALANCE(); }
  maxload = 2.0 * MAXLOAD();
  while (MAXLOAD() < maxload) { iplanes = iplanes + 1; maxload = MAXLOAD(); BALANCE(); }
  iplanes = iplanes - 1;
  BALANCE();
#undef BALANCE
#undef MAXLOAD
  for (int g = 0; g < 2; ++g) {
    const int row = g == 0 ? 4 : 3;
    for (int b = 0; b <= nband; ++b) procs_b[b] = procs[b];
    for (int iband = 1; iband <= nband; ++iband) {
      nplanes = NBT(row, iband) - NBT(row, iband - 1);
      iplanes = nplanes / procs_b[iband];
      rem = nplanes - iplanes * procs_b[iband];
      for (int iproc = 1; iproc <= procs_b[iband]; ++iproc) {
        const int r = iproc - 1 + procs_b[iband - 1];
        int from, to;
        if (iproc <= rem) { from = (iplanes + 1) * (iproc - 1) + 1; to = (iplanes + 1) * iproc; }
        else { from = iplanes * (iproc - rem - 1) + (iplanes + 1) * rem + 1; to = iplanes * (iproc - rem) + (iplanes + 1) * rem; }
        lim[(g * 2) * np + r] = from + NBT(row, iband - 1);
        lim[(g * 2 + 1) * np + r] = to + NBT(row, iband - 1);
      }
      procs_b[iband] = procs_b[iband - 1] + procs_b[iband];
    }
  }
  return 0;
}
#undef NBT

/* ops_in_planes' assembly (FOU3D.f90:1553-1613): on rank r the derivative
   scratch holds the sums of the last plane of limPL_incw (planelim, the first
   rank one plane lower, the last one plane higher: start.f90:1866-1869), and
   every plane of limPL_excw (= planelim) receives them */
static void lastplane_assign(or_state* s) {
  const int np = s->cnp;
  for (int g = 0; g < 2; ++g)
    for (int r = 0; r < np; ++r) {
      const int from = s->clim[(g * 2) * np + r], to = s->clim[(g * 2 + 1) * np + r];
      const int inc_from = from - (r == 0), inc_to = to + (r == np - 1);
      if (inc_from > inc_to) continue;
      const int last = inc_to;
#pragma omp parallel for schedule(static)
      for (int col = 0; col < s->Ncol; ++col)
        for (int j = from; j <= to; ++j) {
          if (g == 0) {
            NU1[fu(s, col, j)] = s->xz1[fu(s, col, last)];
            NU3[fu(s, col, j)] = s->xz3[fu(s, col, last)];
          } else {
            NU2[fv(s, col, j)] = s->xz2[fv(s, col, last)];
          }
        }
    }
}

int or_set_compat_lastplane(or_state* s, const int* N20, int np) {
  free(s->clim); free(s->xz1); free(s->xz3); free(s->xz2);
  s->clim = NULL; s->xz1 = s->xz3 = s->xz2 = NULL;
  s->cnp = 0;
  if (np == 0) return 0;
  if (s->mb) return 4; /* the partition here is for one resolution in every band */
  int* lim = (int*)malloc(sizeof(int) * 4 * (size_t)np);
  if (or_plane_partition(N20, s->ngx, s->ngz, np, lim)) { free(lim); return 1; }
  for (int g = 0; g < 2; ++g) {
    const int lo = g == 0 ? s->jlo_u : s->jlo_v, hi = g == 0 ? s->jhi_u : s->jhi_v;
    for (int r = 0; r < np; ++r) {
      const int from = lim[(g * 2) * np + r], to = lim[(g * 2 + 1) * np + r];
      if (from <= to && (from < lo + 1 || to > hi - 1)) { free(lim); return 2; }
    }
  }
  s->clim = lim;
  s->xz1 = (cplx*)calloc((size_t)s->Ncol * s->nyu, sizeof(cplx));
  s->xz3 = (cplx*)calloc((size_t)s->Ncol * s->nyu, sizeof(cplx));
  s->xz2 = (cplx*)calloc((size_t)s->Ncol * s->nyv, sizeof(cplx));
  s->cnp = np;
  return 0;
}

/* nonlinear (FOU3D.f90:41-259) with ops_in_planes (1529-1617) done per plane */
void or_nonlinear(or_state* s, int kRK) {
  interp_u(s, s->u1i, U1);
  interp_v(s, s->u2i, U2);
  interp_u(s, s->u3i, U3);
  const size_t psz = (size_t)(s->ngx + 2) * s->ngz;
  /* CFL maxima per resolution: [0] the wall bands, [1] the central band */
  double m1f = 0.0, m3f = 0.0, m1c = 0.0, m3c = 0.0, m2 = 1e-6;
  const int nU = s->nyu, nV = s->nyv;
  /* u planes: uu, uw, vv, ww */
#pragma omp parallel reduction(max : m1f, m3f, m1c, m3c)
  {
    double* B = (double*)malloc(sizeof(double) * psz * 7);
    scratch w = {malloc(sizeof(cplx) * (s->ngx + s->ngz + 8)), malloc(sizeof(cplx) * (s->ngx + s->ngz + 8))};
#pragma omp for schedule(dynamic, 1)
    for (int jj = 0; jj < nU; ++jj) {
      const int j = s->jlo_u + jj;
      const int cen = central_plane(s, 0, j);
      const pview v = view_of(s, cen);
      const int igal = v.ngx + 2, ngz = v.ngz, ngx = v.ngx;
      const size_t vsz = (size_t)igal * ngz;
      double *Pu = B, *Pv = B + vsz, *Pw = B + 2 * vsz, *Quu = B + 3 * vsz, *Quw = B + 4 * vsz, *Qvv = B + 5 * vsz, *Qww = B + 6 * vsz;
      to_plane_v(s, &v, U1, nU, s->jlo_u, j, Pu);
      to_plane_v(s, &v, s->u2i, nU, s->jlo_u, j, Pv);
      to_plane_v(s, &v, U3, nU, s->jlo_u, j, Pw);
      four_to_phys_v(&v, Pu, 1, &w);
      four_to_phys_v(&v, Pv, 1, &w);
      four_to_phys_v(&v, Pw, 1, &w);
      memset(Quu, 0, sizeof(double) * vsz * 4);
      double m1 = 0.0, m3 = 0.0;
      for (int k = 0; k < ngz; ++k)
        for (int i = 0; i < ngx; ++i) {
          const size_t o = (size_t)k * igal + i;
          Quu[o] = Pu[o] * Pu[o];
          Quw[o] = Pu[o] * Pw[o];
          Qvv[o] = Pv[o] * Pv[o];
          Qww[o] = Pw[o] * Pw[o];
          m1 = dmax(m1, fabs(Pu[o]));
          m3 = dmax(m3, fabs(Pw[o]));
        }
      if (cen) { m1c = dmax(m1c, m1); m3c = dmax(m3c, m3); }
      else { m1f = dmax(m1f, m1); m3f = dmax(m3f, m3); }
      phys_to_four_v(&v, Quu, &w);
      phys_to_four_v(&v, Quw, &w);
      phys_to_four_v(&v, Qvv, &w);
      phys_to_four_v(&v, Qww, &w);
      const int interior = (j > s->jlo_u && j < s->jhi_u);
      for (int col = 0; col < s->Ncol; ++col) {
        if (cen && s->shortc[col]) { /* no short column on a central plane: its flux rows there stay zero */
          s->vvc[fu(s, col, j)] = 0.0;
          continue;
        }
        const int i = col % s->Mx, k = col / s->Mx, sl = vslot(s, &v, k);
        s->vvc[fu(s, col, j)] = pl_v(&v, Qvv, i, sl);
        if (s->cnp) { /* every plane's sums, walls included; assigned per rank below */
          const int zero = (sl == ngz - v.N2 / 2);
          s->xz1[fu(s, col, j)] = zero ? 0.0 : imul(s->k1x[i], pl_v(&v, Quu, i, sl)) + imul(s->k1z[k], pl_v(&v, Quw, i, sl));
          s->xz3[fu(s, col, j)] = zero ? 0.0 : imul(s->k1x[i], pl_v(&v, Quw, i, sl)) + imul(s->k1z[k], pl_v(&v, Qww, i, sl));
        } else if (interior) {
          const int zero = (sl == ngz - v.N2 / 2); /* der_x/der_z zero band (FOU3D.f90:329-333) */
          cplx a = 0.0, b = 0.0;
          if (!zero) {
            a = imul(s->k1x[i], pl_v(&v, Quu, i, sl)) + imul(s->k1z[k], pl_v(&v, Quw, i, sl));
            b = imul(s->k1x[i], pl_v(&v, Quw, i, sl)) + imul(s->k1z[k], pl_v(&v, Qww, i, sl));
          }
          NU1[fu(s, col, j)] = a;
          NU3[fu(s, col, j)] = b;
        }
      }
    }
    free(B); free(w.w0); free(w.w1);
  }
  /* v planes: uv, vw */
#pragma omp parallel reduction(max : m2)
  {
    double* B = (double*)malloc(sizeof(double) * psz * 5);
    scratch w = {malloc(sizeof(cplx) * (s->ngx + s->ngz + 8)), malloc(sizeof(cplx) * (s->ngx + s->ngz + 8))};
#pragma omp for schedule(dynamic, 1)
    for (int jj = 0; jj < nV; ++jj) {
      const int j = s->jlo_v + jj;
      const int cen = central_plane(s, 1, j);
      const pview v = view_of(s, cen);
      const int igal = v.ngx + 2, ngz = v.ngz, ngx = v.ngx;
      const size_t vsz = (size_t)igal * ngz;
      double *Pu = B, *Pv = B + vsz, *Pw = B + 2 * vsz, *Quv = B + 3 * vsz, *Qvw = B + 4 * vsz;
      to_plane_v(s, &v, s->u1i, nV, s->jlo_v, j, Pu);
      to_plane_v(s, &v, U2, nV, s->jlo_v, j, Pv);
      to_plane_v(s, &v, s->u3i, nV, s->jlo_v, j, Pw);
      four_to_phys_v(&v, Pu, 1, &w);
      four_to_phys_v(&v, Pv, 1, &w);
      four_to_phys_v(&v, Pw, 1, &w);
      memset(Quv, 0, sizeof(double) * vsz * 2);
      const int interior = (j > s->jlo_v && j < s->jhi_v);
      for (int k = 0; k < ngz; ++k)
        for (int i = 0; i < ngx; ++i) {
          const size_t o = (size_t)k * igal + i;
          Quv[o] = Pu[o] * Pv[o];
          Qvw[o] = Pw[o] * Pv[o];
          if (interior) m2 = dmax(m2, fabs(Pv[o]) * DTHV(s, j)); /* dtc_calc (FOU3D.f90:530-536) */
        }
      phys_to_four_v(&v, Quv, &w);
      phys_to_four_v(&v, Qvw, &w);
      for (int col = 0; col < s->Ncol; ++col) {
        if (cen && s->shortc[col]) {
          s->uvf[fv(s, col, j)] = 0.0;
          s->vwf[fv(s, col, j)] = 0.0;
          continue;
        }
        const int i = col % s->Mx, k = col / s->Mx, sl = vslot(s, &v, k);
        s->uvf[fv(s, col, j)] = pl_v(&v, Quv, i, sl);
        s->vwf[fv(s, col, j)] = pl_v(&v, Qvw, i, sl);
        if (s->cnp) {
          const int zero = (sl == ngz - v.N2 / 2);
          s->xz2[fv(s, col, j)] = zero ? 0.0 : imul(s->k1x[i], pl_v(&v, Quv, i, sl)) + imul(s->k1z[k], pl_v(&v, Qvw, i, sl));
        } else if (interior) {
          const int zero = (sl == ngz - v.N2 / 2);
          NU2[fv(s, col, j)] = zero ? 0.0 : imul(s->k1x[i], pl_v(&v, Quv, i, sl)) + imul(s->k1z[k], pl_v(&v, Qvw, i, sl));
        }
      }
    }
    free(B); free(w.w0); free(w.w1);
  }
  if (s->cnp) lastplane_assign(s);
  /* y derivatives and assembly (FOU3D.f90:158-173, 453-493), per band */
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    int lo[2], hi[2];
    int ns = col_segs(s, col, 0, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) {
        NU1[fu(s, col, j)] += (s->uvf[fv(s, col, j)] - s->uvf[fv(s, col, j - 1)]) * DTHU(s, j) * s->ddthetavi;
        NU3[fu(s, col, j)] += (s->vwf[fv(s, col, j)] - s->vwf[fv(s, col, j - 1)]) * DTHU(s, j) * s->ddthetavi;
      }
    ns = col_segs(s, col, 1, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j)
        NU2[fv(s, col, j)] += (s->vvc[fu(s, col, j + 1)] - s->vvc[fu(s, col, j)]) * DTHV(s, j) * s->ddthetavi;
  }
  /* CFL (FOU3D.f90:177-183, 516-544): each band's limit with its own mode
     counts, the smallest over the bands (the MIN all-reduce over the ranks,
     each rank in one band) */
  if (kRK == 1) {
    const double dtl1 = 1.0 / (s->P.alp * (s->N1 / 2) * m1f);
    const double dtl3 = 1.0 / (s->P.bet * (s->N2 / 2) * m3f);
    const double dtl2 = 1.0 / (m2 * s->dthetavi);
    double dtc = dtl1 < dtl3 ? dtl1 : dtl3;
    dtc = dtc < dtl2 ? dtc : dtl2;
    if (s->mb) {
      const double c1 = 1.0 / (s->P.alp * (s->N1c / 2) * m1c), c3 = 1.0 / (s->P.bet * (s->N2c / 2) * m3c);
      const double dc = c1 < c3 ? c1 : c3;
      dtc = dtc < dc ? dtc : dc;
    }
    s->dtc = dtc;
    s->dt = (2.5 * s->dtv < s->P.CFL * dtc) ? 2.5 * s->dtv : s->P.CFL * dtc;
    s->dti = 1.0 / s->dt;
    s->t = s->t + s->dt;
  }
  /* du += -gRK*Nu (FOU3D.f90:232-244) */
  const double C1 = -s->gRK[kRK - 1];
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    int lo[2], hi[2];
    int ns = col_segs(s, col, 0, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) {
        DU1[fu(s, col, j)] += C1 * NU1[fu(s, col, j)];
        DU3[fu(s, col, j)] += C1 * NU3[fu(s, col, j)];
      }
    ns = col_segs(s, col, 1, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) DU2[fv(s, col, j)] += C1 * NU2[fv(s, col, j)];
  }
}

/* ------------------------------------------------------------------------ */
/* immersed boundaries (tridLU_3D_IB.f90:163-281)                            */
/* ------------------------------------------------------------------------ */
static void immersed(or_state* s, cplx* rhs, const cplx* u, int grid) {
  if (s->nib[grid][0] + s->nib[grid][1] == 0) return;
  const int ny_g = grid == 0 ? s->nyu : s->nyv, jlo = grid == 0 ? s->jlo_u : s->jlo_v;
  const int jhi = grid == 0 ? s->jhi_u : s->jhi_v;
  const int igal = s->ngx + 2;
  const size_t psz = (size_t)igal * s->ngz;
  for (int wall = 0; wall < 2; ++wall) {
    const int n = s->nib[grid][wall];
    const int* L = s->ibl[grid][wall];
    const double* W = s->ibw[grid][wall];
    if (!n) continue;
    int jmin = L[2], jmax = L[2];
    for (int q = 0; q < n; ++q) { if (L[9 * q + 2] < jmin) jmin = L[9 * q + 2]; if (L[9 * q + 2] > jmax) jmax = L[9 * q + 2]; }
    if (jmin <= jlo) jmin = jlo + 1;
    if (jmax >= jhi) jmax = jhi - 1;
#pragma omp parallel
    {
      double* R = (double*)malloc(sizeof(double) * psz * 2);
      double* Uu = R + psz;
      scratch w = {malloc(sizeof(cplx) * (s->ngx + s->ngz + 8)), malloc(sizeof(cplx) * (s->ngx + s->ngz + 8))};
#pragma omp for schedule(dynamic, 1)
      for (int j = jmin; j <= jmax; ++j) {
        to_plane(s, rhs, ny_g, jlo, j, R);
        to_plane(s, u, ny_g, jlo, j, Uu);
        four_to_phys(s, R, 0, &w);
        four_to_phys(s, Uu, 0, &w);
        for (int q = 0; q < n; ++q) {
          const int* e = L + 9 * q;
          if (e[2] != j) continue;
          const double Lap = W[3 * q], w2 = W[3 * q + 1], w3 = W[3 * q + 2];
          const double vf = w2 * Uu[(e[4] - 1) * igal + e[3] - 1] + w3 * Uu[(e[7] - 1) * igal + e[6] - 1];
          double* r = &R[(e[1] - 1) * igal + e[0] - 1];
          *r = Lap * (vf - Uu[(e[1] - 1) * igal + e[0] - 1] + *r);
        }
        phys_to_four(s, R, &w);
        for (int col = 0; col < s->Ncol; ++col) {
          const int i = col % s->Mx, k = col / s->Mx;
          rhs[(size_t)col * ny_g + (j - jlo)] = pl(s, R, i, slot_of(s, k));
        }
      }
      free(R); free(w.w0); free(w.w1);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* tridiagonal solves                                                        */
/* ------------------------------------------------------------------------ */
/* LU_dec (tridLU_3D_IB.f90:521-559) */
static void lu_dec(double* a1, double* a2, double* a3, int n) {
  a2[0] = 1.0 / a2[0];
  for (int j = 1; j < n; ++j) {
    a3[j - 1] = a3[j - 1] * a2[j - 1];
    a2[j] = 1.0 / (a2[j] - a1[j] * a3[j - 1]);
  }
}
static void lu_sol(const double* a1, const double* a2, const double* a3, cplx* x, const cplx* r, int n) {
  x[0] = r[0] * a2[0];
  for (int j = 1; j < n; ++j) x[j] = (r[j] - a1[j] * x[j - 1]) * a2[j];
  for (int j = n - 2; j >= 0; --j) x[j] = x[j] - a3[j] * x[j + 1];
}

/* solveU / solveV predictor (littleharsh.f90:732-794) + LUsolU/LUsolV
   (tridLU_3D_IB.f90:50-160) with LU_buildU/LU_buildV (349-413) */
static void predictor(or_state* s, cplx* du, const cplx* u, int grid) {
  const int ny_g = grid == 0 ? s->nyu : s->nyv, jlo = grid == 0 ? s->jlo_u : s->jlo_v;
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    cplx* d = du + (size_t)col * ny_g;
    const cplx* uu = u + (size_t)col * ny_g;
    int lo[2], hi[2];
    const int ns = col_segs(s, col, grid, lo, hi);
    for (int g = 0; g < ns; ++g) {
      const int a = lo[g] - jlo, b = hi[g] - jlo;
      d[a] = 0.0;
      for (int q = a + 1; q < b; ++q) d[q] = uu[q] + s->dt * d[q];
      d[b] = 0.0;
    }
  }
}

static void lusol_uv(or_state* s, cplx* u, const cplx* rhs, int grid, int kRK) {
  const int n = grid == 0 ? s->nyu : s->nyv, jlo = grid == 0 ? s->jlo_u : s->jlo_v;
  const double beta = s->bRK[kRK - 1] * s->dt / s->P.Re;
#pragma omp parallel
  {
    double* a = (double*)malloc(sizeof(double) * 3 * n);
    double *a1 = a, *a2 = a + n, *a3 = a + 2 * n;
#pragma omp for schedule(static)
    for (int col = 0; col < s->Ncol; ++col) {
      const int i = col % s->Mx, k = col / s->Mx;
      const double k2x = s->k2x[i], k2z = s->k2z[k];
      int lo[2], hi[2];
      const int ns = col_segs(s, col, grid, lo, hi);
      for (int g = 0; g < ns; ++g) {
        /* one band's system, boundary rows at its ends (LU_buildU/V with the
           band's gridweighting, tridLU_3D_IB.f90:349-413) */
        const int q0 = lo[g] - jlo, m = hi[g] - lo[g] + 1;
        a1[0] = 0.0; a2[0] = 1.0; a3[0] = grid == 0 ? gw_lo(s, lo[g]) : 0.0;
        for (int q = 1; q < m - 1; ++q) {
          const int j = lo[g] + q;
          const double d1 = grid == 0 ? DYU2(s, 1, j) : DYV2(s, 1, j);
          const double d2 = grid == 0 ? DYU2(s, 2, j) : DYV2(s, 2, j);
          const double d3 = grid == 0 ? DYU2(s, 3, j) : DYV2(s, 3, j);
          a1[q] = -beta * d1;
          a2[q] = 1.0 - beta * (d2 + k2x + k2z);
          a3[q] = -beta * d3;
        }
        a1[m - 1] = grid == 0 ? gw_hi(s, hi[g]) : 0.0; a2[m - 1] = 1.0; a3[m - 1] = 0.0;
        lu_dec(a1, a2, a3, m);
        lu_sol(a1, a2, a3, u + (size_t)col * n + q0, rhs + (size_t)col * n + q0, m);
      }
    }
    free(a);
  }
}

void or_solve_u(or_state* s, int comp, int kRK) {
  cplx* u = comp == 1 ? U1 : U3;
  cplx* du = comp == 1 ? DU1 : DU3;
  predictor(s, du, u, 0);
  immersed(s, du, u, 0);
  lusol_uv(s, u, du, 0, kRK);
}
void or_solve_v(or_state* s, int kRK) {
  predictor(s, DU2, U2, 1);
  immersed(s, DU2, U2, 1);
  lusol_uv(s, U2, DU2, 1, kRK);
}

/* flowrateIm / flowrateRe (littleharsh.f90:411-450): tip to tip */
static double flowrate_c(const or_state* s, const cplx* u0) {
  double Q = 0.0;
  for (int j = 1; j <= s->nn + 1; ++j) Q = Q + (YV(s, j) - YV(s, j - 1)) * creal(u0[j - s->jlo_u]);
  return Q;
}
double or_flowrate(const or_state* s) { return flowrate_c(s, U1); }

/* meanflow_ctP (littleharsh.f90:504-520): with a constant mean pressure
   gradient the flow rate of the (0,0) column of u1 is only recorded */
void or_meanflow_ctP(or_state* s) { s->Qx = flowrate_c(s, U1); }

/* meanflow_ctU -> flowrate_corr with LUsol0/LU_build0/LU_dec0
   (littleharsh.f90:452-484, 522-537; tridLU_3D_IB.f90:320-346, 493-519, 605-622) */
void or_meanflow_ctU(or_state* s, int kRK) {
  const int nn = s->nn, n = nn + 3;
  double* a = (double*)malloc(sizeof(double) * 4 * n);
  double *a1 = a, *a2 = a + n, *a3 = a + 2 * n, *x = a + 3 * n;
  const double bdt = -s->dRK[kRK - 1] * s->dt;
  x[0] = 0.0;
  for (int j = 1; j <= nn + 1; ++j) x[j] = bdt;
  x[nn + 2] = 0.0;
  const double beta = s->bRK[kRK - 1] * s->dt / s->P.Re;
  a1[0] = 0.0; a2[0] = 1.0; a3[0] = s->gw[0];
  for (int j = 1; j <= nn + 1; ++j) {
    a1[j] = -beta * DYU2(s, 1, j);
    a2[j] = 1.0 - beta * DYU2(s, 2, j);
    a3[j] = -beta * DYU2(s, 3, j);
  }
  a1[nn + 2] = s->gw[1]; a2[nn + 2] = 1.0; a3[nn + 2] = 0.0;
  lu_dec(a1, a2, a3, n);
  x[0] = x[0] * a2[0];
  for (int j = 1; j < n; ++j) x[j] = (x[j] - a1[j] * x[j - 1]) * a2[j];
  for (int j = n - 2; j >= 0; --j) x[j] = x[j] - a3[j] * x[j + 1];
  double Qx1 = 0.0;
  for (int j = 1; j <= nn + 1; ++j) Qx1 = Qx1 + (YV(s, j) - YV(s, j - 1)) * x[j];
  const double Qx0 = flowrate_c(s, U1);
  const double g = (s->QxT - Qx0) / Qx1;
  s->mpgx = s->mpgx + g;
  for (int j = 0; j <= nn + 2; ++j) U1[j - s->jlo_u] = U1[j - s->jlo_u] + g * x[j];
  s->dgx = g;
  free(a);
}

/* divergence (littleharsh.f90:209-244) */
void or_divergence(or_state* s) {
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    int lo[2], hi[2];
    const int ns = col_segs(s, col, 2, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g]; j <= hi[g]; ++j)
        DIV[fp(s, col, j)] = imul(s->k1x[i], U1[fu(s, col, j)]) + imul(s->k1z[k], U3[fu(s, col, j)]) +
                             s->ddthetavi * DTHU(s, j) * (U2[fv(s, col, j)] - U2[fv(s, col, j - 1)]);
  }
}

/* LU_buildP + LU_decP (tridLU_3D_IB.f90:417-491, 562-603), once at start;
   per band of a column, the first and last rows without the flux through
   the band's end faces.  Rows of a short column between its bands are
   identity rows. */
static void lu_build_p(or_state* s) {
  const int n = s->nyp;
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    const double k2x = s->k2x[i], k2z = s->k2z[k];
    double* dg = s->DG + (size_t)3 * n * col; /* dg[3*q + m] */
    for (int q = 0; q < 3 * n; ++q) dg[q] = 0.0;
    for (int q = 0; q < n; ++q) dg[3 * q + 1] = 1.0;
    int lo[2], hi[2];
    const int nseg = col_segs(s, col, 2, lo, hi);
    for (int g = 0; g < nseg; ++g) {
      const int ns = lo[g], ne = hi[g], q0 = ns - s->jlo_p, m = ne - ns + 1;
      double* d = dg + 3 * q0;
#define DVB(j) (-s->ddthetavi * DTHU(s, (j) + 1))
#define GVB(j) (-s->ddthetavi * DTHV(s, (j)))
#define DVT(j) (s->ddthetavi * DTHU(s, (j)))
#define GVT(j) (s->ddthetavi * DTHV(s, (j) - 1))
      d[3 * 0 + 1] = (DVT(ns) * GVB(ns)) + k2x + k2z;
      d[3 * 0 + 2] = (DVT(ns) * GVT(ns + 1));
      d[3 * (m - 1) + 1] = (DVB(ne - 1) * GVT(ne)) + k2x + k2z;
      d[3 * (m - 1) + 0] = (DVB(ne - 1) * GVB(ne - 1));
      for (int j = ns + 1; j <= ne - 1; ++j) {
        const int q = j - ns;
        d[3 * q + 1] = (DVB(j - 1) * GVT(j)) + (DVT(j) * GVB(j)) + k2x + k2z;
        d[3 * q + 0] = (DVB(j - 1) * GVB(j - 1));
        d[3 * q + 2] = (DVT(j) * GVT(j + 1));
      }
#undef DVB
#undef GVB
#undef DVT
#undef GVT
      if (col == 0) { /* (0,0) pin (tridLU_3D_IB.f90:468-474) */
        d[0] = 0.0;
        d[1] = 1.0 / ((YU(s, 2) - YU(s, 1)) * (YU(s, 2) - YU(s, 1)));
        d[2] = 0.0;
      }
      d[1] = 1.0 / d[1];
      for (int q = 1; q < m; ++q) {
        d[3 * (q - 1) + 2] = d[3 * (q - 1) + 2] * d[3 * (q - 1) + 1];
        d[3 * q + 1] = 1.0 / (d[3 * q + 1] - d[3 * q + 0] * d[3 * (q - 1) + 2]);
      }
    }
  }
}

/* solveP (littleharsh.f90:556-611) + LUsolP (tridLU_3D_IB.f90:284-318), per band */
void or_solve_p(or_state* s, int kRK) {
  const double dtirk = s->dti / (s->aRK[kRK - 1] + s->bRK[kRK - 1]);
  or_divergence(s);
  const int n = s->nyp;
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    int lo[2], hi[2];
    const int nseg = col_segs(s, col, 2, lo, hi);
    for (int g = 0; g < nseg; ++g) {
      const int q0 = lo[g] - s->jlo_p, m = hi[g] - lo[g] + 1;
      cplx* x = PSI + (size_t)col * n + q0;
      const cplx* dv = DIV + (size_t)col * n + q0;
      for (int q = 0; q < m; ++q) x[q] = dtirk * dv[q];
      if (col == 0) x[0] = 0.0;
      const double* dg = s->DG + (size_t)3 * n * col + 3 * q0;
      x[0] = x[0] * dg[1];
      for (int q = 1; q < m; ++q) x[q] = (x[q] - dg[3 * q] * x[q - 1]) * dg[3 * q + 1];
      for (int q = m - 2; q >= 0; --q) x[q] = x[q] - dg[3 * q + 2] * x[q + 1];
      cplx* pp = PP + (size_t)col * n + q0;
      for (int q = 0; q < m; ++q) pp[q] = pp[q] + x[q];
    }
  }
}

/* v_corr (littleharsh.f90:796-907) */
void or_vcorr(or_state* s, int kRK) {
  const double dtrk = s->dt * (s->aRK[kRK - 1] + s->bRK[kRK - 1]);
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    const cplx kx = imul(s->k1x[i], dtrk), kz = imul(s->k1z[k], dtrk);
    int lo[2], hi[2];
    int ns = col_segs(s, col, 0, lo, hi);
    for (int g = 0; g < ns; ++g)
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j) {
        U1[fu(s, col, j)] = U1[fu(s, col, j)] - kx * PSI[fp(s, col, j)];
        U3[fu(s, col, j)] = U3[fu(s, col, j)] - kz * PSI[fp(s, col, j)];
      }
    /* ghosts at the walls only: the bottom of bands 1, 2 and the top of
       bands 2, 3; a short column's interface ghosts keep their solve values */
    U1[fu(s, col, s->jlo_u)] = -s->gw[0] * U1[fu(s, col, s->jlo_u + 1)];
    U1[fu(s, col, s->jhi_u)] = -s->gw[1] * U1[fu(s, col, s->jhi_u - 1)];
    U3[fu(s, col, s->jlo_u)] = -s->gw[0] * U3[fu(s, col, s->jlo_u + 1)];
    U3[fu(s, col, s->jhi_u)] = -s->gw[1] * U3[fu(s, col, s->jhi_u - 1)];
    ns = col_segs(s, col, 1, lo, hi);
    for (int g = 0; g < ns; ++g) {
      for (int j = lo[g] + 1; j <= hi[g] - 1; ++j)
        U2[fv(s, col, j)] = U2[fv(s, col, j)] - dtrk * s->ddthetavi * DTHV(s, j) * (PSI[fp(s, col, j + 1)] - PSI[fp(s, col, j)]);
      U2[fv(s, col, lo[g])] = 0.0;
      U2[fv(s, col, hi[g])] = 0.0;
    }
  }
  or_divergence(s);
}

/* one RK substep (littleharsh.f90:163-189) */
void or_substep(or_state* s, int kRK) {
  or_rhs0(s, kRK);
  or_nonlinear(s, kRK);
  or_solve_u(s, 1, kRK);
  or_solve_v(s, kRK);
  or_solve_u(s, 3, kRK);
  if (s->P.flag_ctpress == 0) or_meanflow_ctU(s, kRK);
  or_solve_p(s, kRK);
  or_vcorr(s, kRK);
}

/* time loop with the RK loop restored (D1) and meanflow_ctP (192-196) */
void or_step(or_state* s, int nsteps) {
#ifdef _OPENMP
  if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
  for (int it = 0; it < nsteps; ++it) {
    for (int kRK = 1; kRK <= 3; ++kRK) or_substep(s, kRK);
    if (s->P.flag_ctpress == 1) or_meanflow_ctP(s);
  }
}

/* ------------------------------------------------------------------------ */
/* FFT known-answer hooks                                                    */
/* ------------------------------------------------------------------------ */
void or_cft(int n, double* c, int ise, int isa, int m, int iopt) {
  fplan P;
  fplan_init(&P, n);
  cplx* w0 = malloc(sizeof(cplx) * n);
  cplx* w1 = malloc(sizeof(cplx) * n);
  for (int v = 0; v < m; ++v) cft_line(&P, c + (size_t)v * isa, ise, iopt, w0, w1);
  free(w0); free(w1);
  fplan_free(&P);
}
void or_rft(int n, double* c, int isa, int m, int iopt) {
  fplan P;
  fplan_init(&P, n / 2);
  cplx* tw = malloc(sizeof(cplx) * (n / 2 + 1));
  for (int k = 0; k <= n / 2; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * k / (long double)n;
    tw[k] = (double)cosl(a) + I * (double)sinl(a);
  }
  cplx* w0 = malloc(sizeof(cplx) * n);
  cplx* w1 = malloc(sizeof(cplx) * n);
  for (int v = 0; v < m; ++v) rft_line(&P, tw, c + (size_t)v * isa, iopt, w0, w1);
  free(w0); free(w1); free(tw);
  fplan_free(&P);
}

/* ------------------------------------------------------------------------ */
/* statistics and spectra (stats.f90:1-253, spectra.f90:1-59), accumulated  */
/* as the reference does at kRK = 1 of a sampled step (FOU3D.f90:107-110,   */
/* 185-221).  One spectral resolution: the spectra's midband columns are    */
/* all columns and its y ranges (jlim(.,.,midband), start.f90:1538-1557)     */
/* all planes of each grid.  The plane sets of the wall bands are those of  */
/* the reference's plane partition with walls (limPL_incw,                   */
/* start.f90:1759-1866): u planes N(4,0)..N(4,1) and N(4,2)+1..N(4,3)+1,     */
/* v planes N(3,0)..N(3,1) and N(3,2)+1..N(3,3)+1, p planes                  */
/* N(4,0)+1..N(4,1)-1 and N(4,2)+2..N(4,3).                                  */
/* ------------------------------------------------------------------------ */
enum { ST_U, ST_U2, ST_V, ST_V2, ST_W, ST_W2, ST_P, ST_P2, ST_UV, ST_UW, ST_VW, ST_WX, ST_WX2, ST_N };
static const int st_grid[ST_N] = {0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1};

struct or_stats {
  int N[20];
  int dnx, dnz, nyc[3], istat;
  int bot[3][2], top[3][2], jref_bot[3], jref_top[3]; /* band plane ranges; j1 = j - jref_bot, jref_top - j */
  double* prof[ST_N];                    /* Um(j) ... over the planes of the grid */
  double* cb[ST_N];                      /* conddown sums (dnx, dnz, 0:nyc-1), i fastest */
  double* ct[ST_N];                      /* condup sums, same layout */
  double* sp[5];                         /* spU spV spW spUV spP: sp(j, column), y fastest */
};

static void or_stats_free(or_state* s) {
  struct or_stats* t = s->st;
  if (!t) return;
  for (int q = 0; q < ST_N; ++q) { free(t->prof[q]); free(t->cb[q]); free(t->ct[q]); }
  for (int q = 0; q < 5; ++q) free(t->sp[q]);
  free(t);
  s->st = NULL;
}

static int ny_grid(const or_state* s, int g) { return g == 0 ? s->nyu : g == 1 ? s->nyv : s->nyp; }
static int jlo_grid(const or_state* s, int g) { return g == 0 ? s->jlo_u : g == 1 ? s->jlo_v : s->jlo_p; }

/* N20: the band table N(1:4, 0:4) (start.f90:89-111, y_grid_canopy.f90:250-275) */
int or_stats_init(or_state* s, const int* N20) {
  or_stats_free(s);
  if (s->P.ntilex <= 0 || s->P.ntilez <= 0) return 1;
  struct or_stats* t = (struct or_stats*)calloc(1, sizeof *t);
  memcpy(t->N, N20, sizeof t->N);
#define NB(r, b) (t->N[4 * (b) + (r) - 1])
  /* wall bands exist only for band limits in order (the plane partition of
     proc_lims_planes, start.f90:1759-1866, counts N(r,b) - N(r,b-1) planes) */
  for (int r = 3; r <= 4; ++r)
    if (!(NB(r, 0) <= NB(r, 1) && NB(r, 1) <= NB(r, 2) && NB(r, 2) <= NB(r, 3))) {
      free(t);
      return 2;
    }
  t->dnx = s->ngx / s->P.ntilex; /* start.f90:456-457 */
  t->dnz = s->ngz / s->P.ntilez;
  t->bot[0][0] = NB(4, 0); t->bot[0][1] = NB(4, 1); t->top[0][0] = NB(4, 2) + 1; t->top[0][1] = NB(4, 3) + 1;
  t->bot[1][0] = NB(3, 0); t->bot[1][1] = NB(3, 1); t->top[1][0] = NB(3, 2) + 1; t->top[1][1] = NB(3, 3) + 1;
  t->bot[2][0] = NB(4, 0) + 1; t->bot[2][1] = NB(4, 1) - 1; t->top[2][0] = NB(4, 2) + 2; t->top[2][1] = NB(4, 3);
  /* Ny(grid, 0) and Ny(grid, nband) + 1 (start.f90:928-933; stats.f90:137, 186) */
  t->jref_bot[0] = NB(4, 0); t->jref_top[0] = NB(4, 3) + 1;
  t->jref_bot[1] = NB(3, 0); t->jref_top[1] = NB(3, 3) + 1;
  t->jref_bot[2] = NB(4, 0) + 1; t->jref_top[2] = NB(4, 3);
#undef NB
  for (int g = 0; g < 3; ++g) {
    t->nyc[g] = t->bot[g][1] - t->bot[g][0] + 1; /* 0:Ny(grid,1)-Ny(grid,0) (start.f90:2723-2728) */
    if (t->nyc[g] < 1 || t->jref_top[g] - t->top[g][0] >= t->nyc[g] || t->jref_top[g] - t->top[g][1] < 0) {
      free(t);
      return 2;
    }
  }
  const size_t cell = (size_t)t->dnx * t->dnz;
  for (int q = 0; q < ST_N; ++q) {
    const int g = st_grid[q];
    t->prof[q] = calloc(ny_grid(s, g), sizeof(double));
    t->cb[q] = calloc(cell * t->nyc[g], sizeof(double));
    t->ct[q] = calloc(cell * t->nyc[g], sizeof(double));
  }
  const int spg[5] = {0, 1, 0, 0, 2};
  for (int q = 0; q < 5; ++q) t->sp[q] = calloc((size_t)s->Ncol * ny_grid(s, spg[q]), sizeof(double));
  s->st = t;
  return 0;
}

void or_stats_dims(const or_state* s, int* d) {
  const struct or_stats* t = s->st;
  if (!t) { memset(d, 0, sizeof(int) * 6); return; }
  d[0] = t->dnx; d[1] = t->dnz; d[2] = t->nyc[0]; d[3] = t->nyc[1]; d[4] = t->nyc[2]; d[5] = t->istat;
}

/* part 0: profile of quantity q (13, write order of write_stats); 1: bottom
   conditional sums; 2: top conditional sums; 3: spectrum q (0 U, 1 V, 2 W,
   3 UV, 4 P), sp(j, column) */
int or_stats_get(const or_state* s, int q, int part, double* out) {
  const struct or_stats* t = s->st;
  if (!t) return 1;
  const size_t cell = (size_t)t->dnx * t->dnz;
  if (part == 3) {
    if (q < 0 || q > 4) return 1;
    const int spg[5] = {0, 1, 0, 0, 2};
    memcpy(out, t->sp[q], sizeof(double) * s->Ncol * ny_grid(s, spg[q]));
    return 0;
  }
  if (q < 0 || q >= ST_N || part < 0 || part > 2) return 1;
  const int g = st_grid[q];
  if (part == 0) memcpy(out, t->prof[q], sizeof(double) * ny_grid(s, g));
  else memcpy(out, part == 1 ? t->cb[q] : t->ct[q], sizeof(double) * cell * t->nyc[g]);
  return 0;
}

void or_stats_reset(or_state* s) {
  if (!s->st) return;
  int N[20];
  memcpy(N, s->st->N, sizeof N);
  or_stats_init(s, N);
}

/* plane sums over the Ngalx x Ngalz points, i fastest (stats.f90:16-88) */
static void sum1(double* m, double* m2, const double* x, int igal, int nx, int nz) {
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) *m = *m + x[k * igal + i];
  if (m2)
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i) *m2 = *m2 + x[k * igal + i] * x[k * igal + i];
}
static void sum2(double* m, const double* x1, const double* x2, int igal, int nx, int nz) {
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) *m = *m + x1[k * igal + i] * x2[k * igal + i];
}

/* conddown1/condup1 (stats.f90:110-140, 170-205): tile sums at distance j1
   from the wall; sign applies to the first moment */
static void cond1(const struct or_stats* t, double* ym, double* y2m, const double* x, int igal, int j1, double sgn,
                  int ntx, int ntz) {
  const size_t o = (size_t)j1 * t->dnx * t->dnz;
  for (int k = 0; k < t->dnz; ++k)
    for (int i = 0; i < t->dnx; ++i)
      for (int iz = 0; iz < ntz; ++iz)
        for (int ix = 0; ix < ntx; ++ix) {
          const double v = x[(k + t->dnz * iz) * igal + i + t->dnx * ix];
          const size_t c = o + (size_t)k * t->dnx + i;
          ym[c] = ym[c] + (sgn == 1.0 ? v : v * sgn);
          y2m[c] = y2m[c] + v * v;
        }
}
/* conddown2/condup2 (stats.f90:142-168, 207-253) */
static void cond2(const struct or_stats* t, double* ym, const double* x1, const double* x2, int igal, int j1,
                  double s1, double s2, int top, int ntx, int ntz) {
  const size_t o = (size_t)j1 * t->dnx * t->dnz;
  for (int k = 0; k < t->dnz; ++k)
    for (int i = 0; i < t->dnx; ++i)
      for (int iz = 0; iz < ntz; ++iz)
        for (int ix = 0; ix < ntx; ++ix) {
          const size_t e = (size_t)(k + t->dnz * iz) * igal + i + t->dnx * ix;
          const size_t c = o + (size_t)k * t->dnx + i;
          ym[c] = ym[c] + (top ? x1[e] * s1 * x2[e] * s2 : x1[e] * x2[e]);
        }
}

/* band of plane j of grid g: 0 none, 1 bottom (j1 in *j1), 2 top */
static int band_of(const struct or_stats* t, int g, int j, int* j1) {
  if (j >= t->bot[g][0] && j <= t->bot[g][1]) { *j1 = j - t->jref_bot[g]; return 1; }
  if (j >= t->top[g][0] && j <= t->top[g][1]) { *j1 = t->jref_top[g] - j; return 2; }
  return 0;
}

/* one sample: spectra (spectra.f90:1-59) then stats (stats.f90:1-108) of
   the current fields and pressure */
int or_stats(or_state* s) {
  struct or_stats* t = s->st;
  if (!t) return 1;
#ifdef _OPENMP
  if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
  interp_v(s, s->u2i, U2); /* u2_itp (FOU3D.f90:80) */
  interp_u(s, s->u3i, U3); /* u3_itp (FOU3D.f90:81) */
  /* buildsp / buildspUV (spectra.f90:17-59): the midband's columns, with a
     coarser central band the long columns only */
#pragma omp parallel for schedule(static)
  for (int col = 0; col < s->Ncol; ++col) {
    if (s->mb && s->shortc[col]) continue;
    for (int j = s->jlo_u; j <= s->jhi_u; ++j) {
      const size_t e = fu(s, col, j);
      t->sp[0][e] = t->sp[0][e] + creal(U1[e]) * creal(U1[e]) + cimag(U1[e]) * cimag(U1[e]);
      t->sp[2][e] = t->sp[2][e] + creal(U3[e]) * creal(U3[e]) + cimag(U3[e]) * cimag(U3[e]);
      t->sp[3][e] = t->sp[3][e] + creal(U1[e]) * creal(s->u2i[e]) + cimag(U1[e]) * cimag(s->u2i[e]);
    }
    for (int j = s->jlo_v; j <= s->jhi_v; ++j) {
      const size_t e = fv(s, col, j);
      t->sp[1][e] = t->sp[1][e] + creal(U2[e]) * creal(U2[e]) + cimag(U2[e]) * cimag(U2[e]);
    }
    for (int j = s->jlo_p; j <= s->jhi_p; ++j) {
      const size_t e = fp(s, col, j);
      t->sp[4][e] = t->sp[4][e] + creal(PP[e]) * creal(PP[e]) + cimag(PP[e]) * cimag(PP[e]);
    }
  }
  const int igal = s->ngx + 2, ngz = s->ngz, ngx = s->ngx;
  const size_t psz = (size_t)igal * ngz;
  const int ntx = s->P.ntilex, ntz = s->P.ntilez;
  /* u planes (ops_in_planes four_to_phys_u, FOU3D.f90:1556-1558) */
#pragma omp parallel
  {
    double* B = (double*)malloc(sizeof(double) * psz * 3);
    scratch w = {malloc(sizeof(cplx) * (ngx + ngz + 8)), malloc(sizeof(cplx) * (ngx + ngz + 8))};
#pragma omp for schedule(dynamic, 1)
    for (int jj = 0; jj < s->nyu; ++jj) {
      const int j = s->jlo_u + jj;
      double *Pu = B, *Pv = B + psz, *Pw = B + 2 * psz;
      /* a central plane on the central band's grid, the long columns only */
      const pview v = view_of(s, central_plane(s, 0, j));
      const int vig = v.ngx + 2;
      to_plane_v(s, &v, U1, s->nyu, s->jlo_u, j, Pu);
      to_plane_v(s, &v, s->u2i, s->nyu, s->jlo_u, j, Pv);
      to_plane_v(s, &v, U3, s->nyu, s->jlo_u, j, Pw);
      four_to_phys_v(&v, Pu, 1, &w);
      four_to_phys_v(&v, Pv, 1, &w);
      four_to_phys_v(&v, Pw, 1, &w);
      sum1(&t->prof[ST_U][jj], &t->prof[ST_U2][jj], Pu, vig, v.ngx, v.ngz);
      sum1(&t->prof[ST_W][jj], &t->prof[ST_W2][jj], Pw, vig, v.ngx, v.ngz);
      sum2(&t->prof[ST_UV][jj], Pu, Pv, vig, v.ngx, v.ngz);
      sum2(&t->prof[ST_UW][jj], Pu, Pw, vig, v.ngx, v.ngz);
      int j1;
      const int b = band_of(t, 0, j, &j1);
      if (b) {
        double** A = b == 1 ? t->cb : t->ct;
        const int top = b == 2;
        cond1(t, A[ST_U], A[ST_U2], Pu, igal, j1, 1.0, ntx, ntz);
        cond1(t, A[ST_W], A[ST_W2], Pw, igal, j1, 1.0, ntx, ntz);
        cond2(t, A[ST_UV], Pu, Pv, igal, j1, 1.0, -1.0, top, ntx, ntz);
        cond2(t, A[ST_UW], Pu, Pw, igal, j1, 1.0, 1.0, top, ntx, ntz);
      }
    }
    free(B); free(w.w0); free(w.w1);
  }
  /* v planes: u2, u3_itp and omega_x (FOU3D.f90:187-208) */
#pragma omp parallel
  {
    double* B = (double*)malloc(sizeof(double) * psz * 4);
    scratch w = {malloc(sizeof(cplx) * (ngx + ngz + 8)), malloc(sizeof(cplx) * (ngx + ngz + 8))};
    cplx* d = (cplx*)malloc(sizeof(cplx) * s->Ncol);
#pragma omp for schedule(dynamic, 1)
    for (int jj = 0; jj < s->nyv; ++jj) {
      const int j = s->jlo_v + jj;
      double *Pv = B, *Pw = B + psz, *Wx = B + 2 * psz, *D = B + 3 * psz;
      const int cen = central_plane(s, 1, j);
      const pview v = view_of(s, cen);
      const int vig = v.ngx + 2, vmx = v.N1 / 2 + 1;
      const size_t vsz = (size_t)vig * v.ngz;
      to_plane_v(s, &v, U2, s->nyv, s->jlo_v, j, Pv);
      /* der_z of the u2 plane (FOU3D.f90:373-413): modes 0..N1/2 of the
         slots below N2/2 and above Ngz-N2/2, zero elsewhere (the slot of
         kz = -N2/2 included); k1z by the mode's index in the wall bands'
         numbering */
      memset(Wx, 0, sizeof(double) * vsz);
      for (int sl = 0; sl < v.ngz; ++sl) {
        int k;
        if (sl < v.N2 / 2) k = sl;
        else if (sl >= v.ngz - v.N2 / 2 + 1) k = sl - v.ngz + s->N2;
        else continue;
        for (int i = 0; i < vmx; ++i) {
          const cplx z = imul(s->k1z[k], pl_v(&v, Pv, i, sl));
          Wx[sl * vig + 2 * i] = creal(z);
          Wx[sl * vig + 2 * i + 1] = cimag(z);
        }
      }
      /* der_yv_h_wx (FOU3D.f90:495-514) */
      for (int col = 0; col < s->Ncol; ++col)
        d[col] = (U3[fu(s, col, j + 1)] - U3[fu(s, col, j)]) * DTHV(s, j) * s->ddthetavi;
      /* modes_to_planes of the columns into a zeroed plane (a central plane:
         the long columns); only the u2 placement wrote other slots, and those
         were zero */
      memset(D, 0, sizeof(double) * vsz);
      for (int col = 0; col < s->Ncol; ++col) {
        if (cen && s->shortc[col]) continue;
        const int i = col % s->Mx, k = col / s->Mx, sl = vslot(s, &v, k);
        D[sl * vig + 2 * i] = creal(d[col]);
        D[sl * vig + 2 * i + 1] = cimag(d[col]);
      }
      for (size_t e = 0; e < vsz; ++e) Wx[e] = -Wx[e] + D[e];
      four_to_phys_v(&v, Wx, 0, &w);
      to_plane_v(s, &v, s->u3i, s->nyv, s->jlo_v, j, Pw);
      four_to_phys_v(&v, Pv, 1, &w);
      four_to_phys_v(&v, Pw, 1, &w);
      sum1(&t->prof[ST_V][jj], &t->prof[ST_V2][jj], Pv, vig, v.ngx, v.ngz);
      /* wx sums run over i = 1..N1, k = 1..N2 of the physical plane
         (stats.f90:56-62), the band's N1, N2 */
      sum1(&t->prof[ST_WX][jj], &t->prof[ST_WX2][jj], Wx, vig, v.N1, v.N2);
      sum2(&t->prof[ST_VW][jj], Pv, Pw, vig, v.ngx, v.ngz);
      int j1;
      const int b = band_of(t, 1, j, &j1);
      if (b) {
        double** A = b == 1 ? t->cb : t->ct;
        const int top = b == 2;
        cond1(t, A[ST_V], A[ST_V2], Pv, igal, j1, top ? -1.0 : 1.0, ntx, ntz);
        cond1(t, A[ST_WX], A[ST_WX2], Wx, igal, j1, 1.0, ntx, ntz);
        cond2(t, A[ST_VW], Pv, Pw, igal, j1, -1.0, 1.0, top, ntx, ntz);
      }
    }
    free(B); free(d); free(w.w0); free(w.w1);
  }
  /* p planes (FOU3D.f90:210-215, four_to_phys_du) */
#pragma omp parallel
  {
    double* Pp = (double*)malloc(sizeof(double) * psz);
    scratch w = {malloc(sizeof(cplx) * (ngx + ngz + 8)), malloc(sizeof(cplx) * (ngx + ngz + 8))};
#pragma omp for schedule(dynamic, 1)
    for (int jj = 0; jj < s->nyp; ++jj) {
      const int j = s->jlo_p + jj;
      const pview v = view_of(s, central_plane(s, 2, j));
      to_plane_v(s, &v, PP, s->nyp, s->jlo_p, j, Pp);
      four_to_phys_v(&v, Pp, 0, &w);
      sum1(&t->prof[ST_P][jj], &t->prof[ST_P2][jj], Pp, v.ngx + 2, v.ngx, v.ngz);
      int j1;
      const int b = band_of(t, 2, j, &j1);
      if (b) cond1(t, b == 1 ? t->cb[ST_P] : t->ct[ST_P], b == 1 ? t->cb[ST_P2] : t->ct[ST_P2], Pp, igal, j1, 1.0,
                   ntx, ntz);
    }
    free(Pp); free(w.w0); free(w.w1);
  }
  t->istat += 1;
  return 0;
}
