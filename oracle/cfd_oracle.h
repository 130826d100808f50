/*
 * cfd_oracle.h -- CPU FP64 restatement of the reference RK3/CN substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker / the timed CPU
 * baseline, never as the thing measured or shipped.
 *
 * Every routine cites the reference file:line it restates (reference tree:
 * Mel222/CFD, Fortran).  Deliberate deviations are listed here:
 *   D1  the RK loop is restored (kRK = 1..3, littleharsh.f90:159-163).
 *   D2  the x/z nonlinear derivatives are assembled per plane
 *       (intent of FOU3D.f90:1553-1613), not from the last plane.
 *   D8  y-grid: the polynomial grid of y_grid_canopy.f90 (ygrid = 0) and a
 *       tanh tip-to-tip grid (ygrid = 1) with the same canopy extension.
 * Multi-band resolution (start.f90:89-111, 1420-1557): with ngalx2/ngalz2
 * below ngalx/ngalz the central y band (N(4,1) < j <= N(4,2) on the u grid)
 * is transformed on the coarser Ngal(1:2,2) grid and holds only the "long"
 * columns (the N(1:2,2) modes, which span every plane); the other "short"
 * columns live only in the two wall bands, each band a closed system with
 * its own boundary rows at the band interface (gridweighting(1,2),
 * gridweighting(3,1), start.f90:494-508).  Fields keep one array per grid
 * over all fine columns and all planes; a short column's rows between its
 * two bands are not part of the system and stay as they are (zero).
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* integers first, then doubles: the ctypes mirror in tests/ depends on it */
  int ngalx, ngalz, ny;        /* Ngal(1,1), Ngal(2,1), ny (input.in:15-19) */
  int dsty, dsty1, ppp;        /* canopy points in y; polynomial exponent */
  int ygrid;                   /* 0 polynomial, 1 tanh */
  int geometry_type;           /* 0 smooth, 3 canopies (start.f90:201-226) */
  int ntilex, ntilez;
  int extras24;                /* geom_circ_rough_mixed_forcing_24pt.f90:219-259 */
  int flag_ctpress;            /* 0 constant flow rate, 1 constant mpgx */
  int ngalx2, ngalz2;          /* Ngal(1,2), Ngal(2,2) of the central band; 0: as band 1 */
  double Re, alp, bet, dyq;
  double posth, posth1, ytanh_beta;
  double CFL, mpgx;
  double r, c_i, c_k, stag;    /* geom_circ_rough.txt */
} or_params;

typedef struct or_state or_state;

or_state* or_create(const or_params* p, char* err, int errlen);
void      or_destroy(or_state* s);

/* dims[0..15]: N1 N2 Mx Ncol nyu nyv nyp jlo_u jlo_v jlo_p ngx ngz ny dsty nib_u nib_v
   (the wall bands' resolution; or_band_info gives the rest) */
void or_dims(const or_state* s, int* dims);
/* y grid, index j - jlo: yu/dthdyu/dyu2i(3,.) over nyu, yv/dthdyv/dyv2i over nyv */
void or_get_ygrid(const or_state* s, double* yu, double* yv, double* dthdyu,
                  double* dthdyv, double* dyu2i, double* dyv2i);
/* gw[0..1] gridweighting(mid,1..2); gwi[0..1] gridweighting_interp(mid,1..2); sc: dtv */
void or_get_metrics(const or_state* s, double* gw, double* gwi, double* dtv);
/* IB list of one grid (0 = u/w, 1 = v) and wall (0 bottom, 1 top): returns n;
   fills l9[9*n] (i,k,j,i2,k2,j2,i3,k3,j3; column-major like the reference)
   and w3[3*n] (Lap, w2, w3) when the pointers are non-null */
int  or_get_iblist(const or_state* s, int grid, int wall, int* l9, double* w3);
/* pressure LU (3, nyp, Ncol) after LU_decP */
void or_get_dg(const or_state* s, double* dg);
/* multi-band description: b[0] 1 when the central band is coarser, b[1..2]
   N1, N2 of the central band, b[3..10] N(3,0:3), N(4,0:3) (the band table's
   y limits); shortc[Ncol] (may be null) 1 for the wall-band-only columns */
void or_band_info(const or_state* s, int* b, unsigned char* shortc);

/* fields in reference column layout: complex(8) f(j, column), y fastest,
   interleaved (re, im).  which: 0 u1 1 u2 2 u3 3 p 4 Nu1 5 Nu2 6 Nu3
   7 du1 8 du2 9 du3 10 psi 11 div */
int  or_set_field(or_state* s, int which, const double* src);
int  or_get_field(const or_state* s, int which, double* dst);
/* scalars: t dt dti mpgx dgx QxT Qx dtc */
void or_set_scalars(or_state* s, const double* v);
void or_get_scalars(const or_state* s, double* v);

/* operators (reference names) */
void or_rhs0(or_state* s, int kRK);                 /* RHS0_u1/u2/u3 */
void or_nonlinear(or_state* s, int kRK);            /* nonlinear (NL part) */
/* the as-shipped last-plane nonlinear term of np >= 3 ranks (0: per plane);
   N20 the band table N(1:4, 0:4) */
int  or_set_compat_lastplane(or_state* s, const int* N20, int np);
/* proc_lims_planes (start.f90:1652-1866): planelim of the u and v grids */
int  or_plane_partition(const int* N20, int ngalx, int ngalz, int np, int* lim);
void or_solve_u(or_state* s, int comp, int kRK);    /* solveU(u1|u3) incl. IB */
void or_solve_v(or_state* s, int kRK);              /* solveV(u2) incl. IB */
void or_meanflow_ctU(or_state* s, int kRK);
void or_solve_p(or_state* s, int kRK);
void or_vcorr(or_state* s, int kRK);
void or_divergence(or_state* s);
double or_flowrate(const or_state* s);              /* flowrateIm on (0,0) u1 */
void or_meanflow_ctP(or_state* s);                 /* meanflow_ctP: Qx = flowrateIm */
void or_substep(or_state* s, int kRK);
void or_step(or_state* s, int nsteps);
void or_set_threads(int n);

/* statistics and spectra (stats.f90, spectra.f90): or_stats_init with the
   band table N(1:4,0:4) (20 ints) allocates zeroed accumulators (returns 1
   for ntilex/ntilez <= 0, 2 for band ranges that do not fit); or_stats adds
   one sample of the current fields; or_stats_dims gives dnx, dnz, the wall
   band plane counts of the u, v, p grids and istat; or_stats_get copies one
   accumulator (part 0 profile, 1 bottom / 2 top conditional sums
   (dnx, dnz, j1), 3 spectrum sp(j, column) of U V W UV P); or_stats_reset
   zeroes everything.  With a coarser central band its planes are summed on
   the central band's grid (the long columns only), and the spectra cover the
   long columns (the midband's, spectra.f90:1-59); the short columns' spectra
   stay zero. */
int  or_stats_init(or_state* s, const int* N20);
int  or_stats(or_state* s);
void or_stats_dims(const or_state* s, int* d);
int  or_stats_get(const or_state* s, int q, int part, double* out);
void or_stats_reset(or_state* s);

/* FFT known-answer hooks with the reference calling conventions
   (cft_buff.f:1-65, rft_buff.f:24-86) */
void or_cft(int n, double* c, int ise, int isa, int m, int iopt);
void or_rft(int n, double* c, int isa, int m, int iopt);

#ifdef __cplusplus
}
#endif
#endif
