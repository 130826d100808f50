/*
 * cfd_amd.h -- C ABI of the MI355X-native RK3/Crank-Nicolson substep of the
 * Fourier(x,z) / finite-difference(y) rough-channel DNS.
 *
 * The reference (Mel222/CFD, Fortran + MPI) has no plugin API: its hot path
 * is a set of external subroutines called from the main loop
 * (littleharsh.f90:163-189) that share state through `module declaration`.
 * Each entry point below replaces one of those subroutines; the state they
 * share lives behind a cfd_handle on one GPU.  A Fortran caller binds them
 * with ISO_C_BINDING (see cfd_amd/fortran/cfd_amd_mod.f90).
 *
 * Conventions
 *  - Every function returns 0 on success, >0 on error (CFD_E*); the message
 *    is available from cfd_last_error(handle) (handle may be NULL for
 *    cfd_create failures).  The Fortran wrapper turns a nonzero status into
 *    `error stop`, in place of the reference's `stop`.
 *  - Spectral fields cross the boundary in the reference layout
 *    complex(8) f(jlim1:jlim2, column): y fastest, columns ordered
 *    k = 1..N2 outer and i = 0..N1/2 inner (start.f90:1467-1481), as
 *    interleaved (re, im) doubles.  Copies are explicit and synchronous.
 *  - Handles are not thread-safe; several handles may coexist.
 */
#ifndef CFD_AMD_H
#define CFD_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_EINVAL 1  /* invalid argument / configuration */
#define CFD_EHIP 2    /* HIP runtime error */
#define CFD_ENOMEM 3  /* device allocation failed */

typedef struct cfd_solver* cfd_handle;

/* Parameters of input.in (start.f90:62-227) and geom_circ_rough.txt
   (geom_circ_rough_mixed_forcing_24pt.f90:113-123).  Integers first, then
   doubles.  ngalx/ngalz are the wall bands' resolution (Ngalx1, Ngalz1);
   ngalx2/ngalz2 the central band's (Ngalx2, Ngalz2, input.in:16, 18), 0 for
   the same resolution in every band.  With a coarser central band (multi-
   band resolution, start.f90:89-111, 1420-1557) the fields keep every
   column of the wall bands over every plane (see cfd_set_field), and the
   path runs through cfd_substep / cfd_step (FP64 or mixed). */
typedef struct cfd_params {
  int ngalx, ngalz;     /* physical grid points in x and z (Ngalx1, Ngalz1)      */
  int ny;               /* y points tip to tip, y(0)=-1 .. y(ny-1)=1 (input.in:19) */
  int dsty, dsty1;      /* canopy points in y; points in the upper canopy layer   */
  int ppp;              /* polynomial exponent of the y map (input.in:21)         */
  int ygrid;            /* 0 = polynomial (y_grid_canopy.f90), 1 = tanh          */
  int geometry_type;    /* 0 smooth, 3 canopies                                   */
  int ntilex, ntilez;   /* post tiles per wall                                    */
  int extras24;         /* 1 = the 4 fixed forcing points of the 24pt variant     */
  int flag_ctpress;     /* 0 constant flow rate, 1 constant mean pressure grad.  */
  int precision;        /* 0 = FP64 everywhere; 1 = mixed: FP32 transforms,      */
                        /* products and nonlinear scratch, FP64 fields, solves, */
                        /* projection and flow rate (four-step FFT lengths only) */
  int device;           /* HIP device ordinal                                     */
  int ngalx2, ngalz2;   /* central band's x and z points (0: as ngalx, ngalz)   */
  double Re, alp, bet;  /* bulk Reynolds number; Lx = 2 pi/alp, Lz = 2 pi/bet     */
  double dyq;           /* dymax/dymin of the polynomial map                      */
  double posth, posth1; /* canopy depth (outer units); turning point             */
  double ytanh_beta;    /* stretching of the tanh map                             */
  double CFL, mpgx;     /* CFL number; mean pressure gradient (flag_ctpress=1)    */
  double r, c_i, c_k, stag; /* post radius, centre (grid units in a tile), stagger */
} cfd_params;

/* field ids for cfd_set_field / cfd_get_field */
enum {
  CFD_U1 = 0, CFD_U2, CFD_U3, CFD_P, CFD_NU1, CFD_NU2, CFD_NU3,
  CFD_DU1, CFD_DU2, CFD_DU3, CFD_PSI, CFD_DIV, CFD_NFIELDS
};
/* scalar ids for cfd_set_scalars / cfd_get_scalars (8 doubles) */
enum { CFD_T = 0, CFD_DT, CFD_DTI, CFD_MPGX, CFD_DGX, CFD_QXT, CFD_QX, CFD_DTC, CFD_NSCALARS };

/* Setup: replaces start (start.f90:36-649): y_grid_canopy, def_k,
   proc_lims_columns (one device owns every column), getbounds /
   boundary_circ_rough (IB lists), gridweighting, LU_buildP. */
int cfd_create(const cfd_params* p, cfd_handle* out);

/* Launch shapes (no counterpart in the reference, whose CPU loops have
   none).  fft: 0 = the four-step register FFT kernels where the lengths
   have them, 1 = the LDS Stockham kernels (FP64 path only); fixed at
   creation.  seg_rows: rows per segment of the segmented y solves, 5, 9, 13
   or 21; solve_threads: threads per block of those solves, 64, 128 or 256.
   Every choice gives the same solution to rounding.  cfd_create uses fft 0
   (1 when the environment sets CFD_AMD_FFT=stockham), seg_rows 13 and
   solve_threads 256. */
typedef struct cfd_tuning {
  int fft, seg_rows, solve_threads;
} cfd_tuning;
int cfd_create_tuned(const cfd_params* p, const cfd_tuning* t, cfd_handle* out);
int cfd_get_tuning(cfd_handle h, cfd_tuning* t);
int cfd_set_tuning(cfd_handle h, const cfd_tuning* t); /* fft must not change */

/* Host-only setup products (no GPU needed): the y grid and IB lists that
   cfd_create builds, for a caller that keeps its own copy or checks them.
   Array sizes: nyu = ny+2*dsty+1, nyv = ny+2*dsty; see cfd_get_ygrid and
   cfd_get_iblist. */
int cfd_build_ygrid(const cfd_params* p, double* yu, double* yv, double* dthdyu,
                    double* dthdyv, double* dyu2i, double* dyv2i);
int cfd_build_iblist(const cfd_params* p, int grid, int wall, int* l9, double* w3, int* n);
int cfd_destroy(cfd_handle h);
const char* cfd_last_error(cfd_handle h);

/* dims[16]: N1 N2 Mx Ncol nyu nyv nyp jlo_u jlo_v jlo_p ngx ngz ny dsty nib_u nib_v */
int cfd_dims(cfd_handle h, int* dims);
/* y grid (y_grid_canopy.f90): yu, dthdyu, dyu2i(3,.) over nyu points from
   j = jlo_u; yv, dthdyv, dyv2i over nyv points from jlo_v */
int cfd_get_ygrid(cfd_handle h, double* yu, double* yv, double* dthdyu,
                  double* dthdyv, double* dyu2i, double* dyv2i);
/* IB list of one grid (0 = u/w, 1 = v) and wall (0 bottom, 1 top) as written
   by boundary_circ_rough: l9 (9,n) = i,k,j,i2,k2,j2,i3,k3,j3 and w3 (3,n) =
   Lap, w2, w3.  *n receives the count; l9/w3 may be NULL to query it. */
int cfd_get_iblist(cfd_handle h, int grid, int wall, int* l9, double* w3, int* n);

/* Fields as (Re, Im) pairs f(plane, column), the plane index fastest:
   src[2 * (c * ny_g + q)], column c = k * Mx + i (kx index i, kz index k in
   column order), plane q = j - jlo of the field's grid (u/w, v or p).  With
   a coarser central band a short column (i > N1c/2, or kz outside the
   central band's N2c modes; proc_lims_columns, start.f90:1424-1501) has no
   rows between its two wall bands: u/w planes N(4,1)+2 .. N(4,2)-1, v planes
   N(3,1)+2 .. N(3,2)-1, p planes N(4,1)+1 .. N(4,2) (the band table of
   start.f90:1420-1557) are read as zero and returned as zero there. */
int cfd_set_field(cfd_handle h, int which, const double* src);
int cfd_get_field(cfd_handle h, int which, double* dst);
int cfd_set_scalars(cfd_handle h, const double* v);
int cfd_get_scalars(cfd_handle h, double* v);

/* Hot-path operators.  kRK = 1, 2, 3.  One at a time they exist for one
   resolution in every band; with a coarser central band only cfd_substep
   and cfd_step run (the others return CFD_EINVAL). */
int cfd_rhs0(cfd_handle h, int kRK);             /* RHS0_u1/u2/u3   littleharsh.f90:613-730 */
int cfd_nonlinear(cfd_handle h, int kRK);        /* nonlinear       FOU3D.f90:41-259        */
int cfd_solve_u(cfd_handle h, int comp, int kRK);/* solveU(u1|u3)   littleharsh.f90:732-762 */
int cfd_solve_v(cfd_handle h, int kRK);          /* solveV(u2)      littleharsh.f90:764-794 */
int cfd_meanflow_ctU(cfd_handle h, int kRK);     /* meanflow_ctU    littleharsh.f90:522-537 */
int cfd_solve_p(cfd_handle h, int kRK);          /* solveP          littleharsh.f90:556-611 */
int cfd_vcorr(cfd_handle h, int kRK);            /* v_corr          littleharsh.f90:796-907 */
int cfd_substep(cfd_handle h, int kRK);          /* body of the RK loop, littleharsh.f90:163-189 */
/* nsteps time steps of 3 substeps each (the time loop of littleharsh.f90:159-200
   with the RK loop restored), plus meanflow_ctP when flag_ctpress = 1.
   Asynchronous: returns once the work is queued on the handle's stream. */
int cfd_step(cfd_handle h, int nsteps);
int cfd_meanflow_ctP(cfd_handle h);              /* meanflow_ctP    littleharsh.f90:504-520 */
/* The record of the history file (FOU3D.f90:2504-2534), v[11] = t, dtv,
   dtc, dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip: err = max |div| over the p
   planes (error, littleharsh.f90:322-349), Qx the tip-to-tip flow rate of
   the (0,0) column of u1 (flowrateIm; stored as the Qx scalar, as record_out
   does), Umax its largest real part between the tips (maxvel,
   littleharsh.f90:486-502), Uslip its value interpolated to the lower tip
   face yv(0).  Synchronous. */
int cfd_hist(cfd_handle h, double* v);
/* Modal filter of the velocity after every RK step (no counterpart in the
   reference; off by default): every plane of column (i, kz) of u1, u2, u3
   is multiplied by sigma = exp(-alpha ((eta - eta_c)/(1 - eta_c))^order)
   where eta = max(|i|/(N1/2), |kz|/(N2/2)) > eta_c, 1 elsewhere.  One factor
   per column keeps each mode divergence-free.  A dissipation for coarse,
   under-resolved grids; alpha = 0 switches it off. */
int cfd_set_mode_filter(cfd_handle h, double alpha, int order, double eta_c);
/* The reference's as-shipped nonlinear term for a run on np ranks
   (ops_in_planes, FOU3D.f90:1553-1613): the x/z derivative scratch is
   overwritten plane by plane over each rank's planes with walls and then
   copied to all the rank's planes without walls, so every plane of a rank
   takes the x/z terms of its last plane (the top rank's: the top ghost u
   plane, the top wall v plane); the y terms stay per plane.  np >= 3 (the
   reference's partition needs three ranks); np = 0 restores the per-plane
   term, the default.  Kept across substeps. */
int cfd_set_compat_lastplane(cfd_handle h, int np);
/* The reference's plane partition over np >= 3 ranks (start.f90:1652-1866,
   one resolution in every band): lim[(g * 2 + e) * np + r] = planelim(grid,
   e + 1, r), the first (e = 0) and last (e = 1) plane of rank r without
   walls, g = 0 the u grid, 1 the v grid, in the reference's y indices.
   Host-only. */
int cfd_plane_partition(const cfd_params* p, int np, int* lim);
/* ---------------------------------------------------------------------------
 * Field files in the reference layout: Fortran sequential unformatted, 4-byte
 * little-endian record markers (record_out, FOU3D.f90:2172-2345; u_to_buff /
 * buff_to_u, FOU3D.f90:2598-2654; read_in, start.f90:2251-2700):
 *   record 1  t, Re, alp, bet, mpgx (real8), nband, iter (int4), 88 int4 zeros
 *   record 2  N(4, 0:nband+1) (int4, column-major)
 *   record 3  y(j0 .. j0+ny_g-1), dthetavi, dthdy(same range)   (real8)
 *   then one record per plane j ascending:
 *             j, id, nx, nz (int4), y(j) (real8), buff(nx, nz) (real8)
 * buff holds (Re, Im) of kx = 0..N1/2 fastest (nx = N1+2 reals), then the
 * nz = N2 z-modes in column order (kz >= 0, then the negative modes).
 * ------------------------------------------------------------------------- */
typedef struct {
  double t, Re, alp, bet, mpgx;
  int nband, iter;
  int N[20]; /* N(1:4, 0:nband+1) for nband = 3 */
} cfd_file_header;

/* Write one field file.  y, dthdy: the grid array of the field, ny_arr values
 * for j = jy0 .. jy0+ny_arr-1 (record 3; the p file carries the u/w centres);
 * planes: ny_g planes of nx*nz reals numbered j0 .. j0+ny_g-1, each record
 * with its y(j).  Host-only. */
int cfd_field_file_write(const char* path, const cfd_file_header* hd, int id, double dthetavi, const double* y,
                         const double* dthdy, int ny_arr, int jy0, int ny_g, int j0, int nx, int nz,
                         const double* planes);
/* Read one field file into ny_g planes of nx*nz reals (zeroed first).  A file
 * plane of other size is copied as buff_to_u copies it into a plane of ngz
 * z-slots (FOU3D.f90:2631-2654): x reals up to min(nx, nx_file), the first
 * min(nz_file, ngz)/2 z-modes from the bottom and the same number of negative
 * modes from the top, then kept where they fall inside the nz target modes;
 * *mismatch counts such planes.  Host-only. */
int cfd_field_file_read(const char* path, cfd_file_header* hd, int ny_g, int nx, int nz, int ngz, double* planes,
                        int* mismatch);

/* four_to_phys (four_to_phys.f90:61-120, resize 328-348): the physical values
 * of every plane of a field, phys[plane][z][x] (Ngalz x Ngalx per plane): the
 * N1 x N2 modes zero-padded to Ngalx x Ngalz without the kz = -N2/2 mode,
 * inverse z then x transforms, unscaled (cft_buff.f / rft_buff.f inverse
 * conventions). */
int cfd_four_to_phys(cfd_handle h, int which, double* phys);

/* No counterpart in the reference: the variant of cfd_four_to_phys for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Planes [plane0,
 * plane0 + nplanes) of field `which` are written to phys, a DEVICE buffer of
 * nplanes x Ngalz x Ngalx doubles on the handle's device, asynchronously on
 * the handle's stream (cfd_sync before reading it); the values are those of
 * cfd_four_to_phys for the same planes.  No allocation: the half-transformed
 * planes and job tables live in a fixed 16 MiB per-GPU scratch, so nplanes x
 * Ngalz x (N1/2 + 1) x 16 bytes (plus under 100 bytes per plane of job tables) must
 * fit in it (e.g. 40 planes of a 384 x 192 grid); larger requests fail with
 * CFD_EINVAL and a message giving the limit.  Every call on one GPU shares
 * the scratch: calls that overlap on one GPU race. */
int cfd_four_to_phys_planes(cfd_handle h, int which, int plane0, int nplanes, double* phys);

/* The band table N(1:4, 0:4) that record_out writes (start.f90:89-111;
 * y_grid_canopy.f90:250-275): x/z mode counts per band and the band limits
 * in y, the bottom band reaching h_ny above the lower tips (0.25 for smooth
 * walls, two post spacings for posts, start.f90:199-226).  Host-only. */
int cfd_band_table(const cfd_params* p, int* N20);

/* record_out (FOU3D.f90:2172-2345): write u1, u2, u3, p of the handle as
 * dirout/<f>_<N1>x<N2>x<ny>_t<int(10 t)>.dat (fields %04d, time %05d) with
 * the current t and mpgx and the given iteration number. */
int cfd_record_out(cfd_handle h, const char* dirout, int iter);
/* mblock_ini + read_in (start.f90:2061-2097, 2251-2700) with the flag_init
 * branches of start.f90:1974-2009: reads dirin/u1<tail>, u2<tail>, u3<tail>,
 * p<tail>, tail = fnameimb without its first two characters.  flag_init 1:
 * t = 0, *iter0 = 0; flag_init 2: t and *iter0 from the u1 header.  QxT is set
 * to the flow rate of the read u1 unless flag_ctpress = 1; mpgx is left as
 * it is (the reference does not restore it).  A different nband stops with
 * CFD_EINVAL; a different alp, bet or N is reported on stderr and read on,
 * as the reference warns and reads on. */
int cfd_read_in(cfd_handle h, const char* dirin, const char* fnameimb, int flag_init, int* iter0);

int cfd_sync(cfd_handle h);
/* device memory held by the handle, bytes (no counterpart in the reference) */
int cfd_device_bytes(cfd_handle h, double* bytes);

/* Statistics and spectra (stats.f90, spectra.f90).
 *
 * cfd_stats adds one sample of the current state (the reference's
 * nonlinear with flagst, FOU3D.f90:107-110 and 185-221): the spectra
 * sp(j, column) += |u|^2 of u1, u2, u3, p and Re(u1 conj u2_itp) over every
 * plane and column (spectra.f90:17-59); the plane sums of the physical u1,
 * u1^2, u2, u2^2, u3, u3^2, p, p^2, u1 u2_itp, u1 u3, u2 u3_itp and
 * omega_x, omega_x^2 (the omega_x sums over i <= N1, k <= N2 of the plane,
 * stats.f90:56-62); and the tile-conditional sums of the same over the
 * planes of the two wall bands (conddown / condup, stats.f90:110-253),
 * tiles of dnx = Ngalx/ntilex by dnz = Ngalz/ntilez points.
 *
 * cfd_write_stats writes dirout/spU_<tag> ... dirout/spP_<tag> (recvwrspec,
 * spectra.f90:132-182) and dirout + "stats_<tag>" (write_stats,
 * stats.f90:355-392; no separator, as the reference concatenates them),
 * tag = <N1>x<N2>x<ny>_<int(10 t)>.dat (%04d, %05d), with the current t and
 * mpgx in the headers, then zeroes every accumulator and istat.
 *
 * cfd_stats_request(h, flags, dirout, iter): bit 0 sample, bit 1 write; the
 * next nonlinear term with kRK = 1 (in cfd_nonlinear, cfd_substep or
 * cfd_step) serves it where the reference's nonlinear does: the sample of
 * the fields at the start of the substep, the files after the time-step
 * update, with the new t and the start-of-substep mpgx in the headers and
 * the start-of-substep t in the names.
 *
 * cfd_stats_dims: d[6] = dnx, dnz, wall-band plane counts of the u, v and p
 * grids (the third extent of the conditional records), istat.
 *
 * With a coarser central band (Ngalx2/Ngalz2) the central planes (u
 * N(4,1)+1..N(4,2), v N(3,1)+1..N(3,2), p N(4,1)..N(4,2)+1) are summed on
 * the Ngalx2 x Ngalz2 grid from the long columns, the stats file's Ngal
 * table carries Ngalx2, Ngalz2 in band 2, and the spectra cover the long
 * columns (the midband's) in N1c/2+1 x N2c/2+1 modes (spectra.f90:122-170).
 *
 * All four return CFD_EINVAL when ntilex or ntilez is below 1 or the band
 * table's limits are not in order (no wall bands to condition on). */
int cfd_stats(cfd_handle h);
int cfd_write_stats(cfd_handle h, const char* dirout, int iter);
int cfd_stats_request(cfd_handle h, int flags, const char* dirout, int iter);
int cfd_stats_dims(cfd_handle h, int* d);

/* Kernel-class timing with HIP events on the handle's stream.  classes:
   0 z-inverse FFT, 1 x-pass FFT+products, 2 z-forward FFT, 3 spectral
   elementwise, 4 y-solves (u,v,w), 5 pressure solve + correction, 6 IB.
   ms[c] = accumulated milliseconds, cnt[c] = launches since enable. */
int cfd_profile(cfd_handle h, int enable);
int cfd_profile_read(cfd_handle h, double* ms, int* cnt);


/* ------------------------------------------------------------------------
 * Learned closure network (no counterpart in the reference: the learned
 * coarse-grid correction of the nonlinear term).  A 3D CNN over a volume of
 * ny planes x nz x nx points (nx a multiple of 32, nz of 8), periodic in x
 * and z, zero beyond the planes: four 3x3x3 convolutions 4 -> 32 -> 32 -> 32
 * -> 3 channels with GELU after the first three, BF16 weights and
 * activations, FP32 accumulation and FP32 master weights trained by Adam.
 * Parameters in cfd_cnn_nparams() floats: per layer W[t][ci][co] (t = 9 (dy+1)
 * + 3 (dz+1) + (dx+1)) then b[co].  Volumes on the host are point-major,
 * point = (y nz + z) nx + x: input [p][4] (u, v, w, mask), target and
 * output [p][3].  The dataset holds nslots (input, target) pairs on the
 * device, filled by cfd_closure_snapshot.
 * ------------------------------------------------------------------------ */
typedef struct cfd_cnn_s* cfd_cnn;
int cfd_cnn_nparams(void);
int cfd_cnn_create(int device, int ny, int nz, int nx, cfd_cnn* out);
int cfd_cnn_destroy(cfd_cnn c);
const char* cfd_cnn_last_error(cfd_cnn c);
int cfd_cnn_set_params(cfd_cnn c, const float* w);
int cfd_cnn_get_params(cfd_cnn c, float* w);
int cfd_cnn_get_grads(cfd_cnn c, float* g);
int cfd_cnn_set_input(cfd_cnn c, const float* x);
int cfd_cnn_set_target(cfd_cnn c, const float* t);
int cfd_cnn_forward(cfd_cnn c);
int cfd_cnn_get_output(cfd_cnn c, float* y);
/* gradients of the mean squared error of the last forward pass against the
 * target; loss may be null */
int cfd_cnn_backward(cfd_cnn c, double* loss);
int cfd_cnn_adam(cfd_cnn c, float lr, float b1, float b2, float eps);
int cfd_cnn_dataset(cfd_cnn c, int nslots);
int cfd_cnn_train_slot(cfd_cnn c, int slot, float lr, float b1, float b2, float eps, double* loss);
int cfd_cnn_eval_slot(cfd_cnn c, int slot, double* loss);
int cfd_cnn_load_slot(cfd_cnn c, int slot);
int cfd_cnn_device_bytes(cfd_cnn c, double* bytes);
/* normalisation of the closure data: s6 = u1, u2, u3 input scales, then the
 * three correction scales (all > 0; set by the first cfd_closure_snapshot
 * when unset) */
int cfd_cnn_set_scales(cfd_cnn c, const float* s6);
int cfd_cnn_get_scales(cfd_cnn c, float* s6);

/* The closure on the solver side (no counterpart in the reference).
 *
 * cfd_closure_snapshot(fine, coarse, c, slot, rms): one training sample
 * into dataset slot `slot`: the fine state box-filtered onto the coarse
 * modes (u_c = Hx(kx) Hz(kz) u_f, H(k) = sinc(k dx_c / 2), dx_c the coarse
 * spacing) is the input, filt(N(u_f)) - N_c(u_c) the target (N the nonlinear
 * term of each solver, FOU3D.f90:41-259), both as physical planes of the
 * coarse grid scaled by the network's scales.  The solvers share the y grid;
 * the coarse solver's fields are overwritten.  rms[6] (may be null) returns
 * the root mean squares of u1, u2, u3 and the three targets.
 *
 * cfd_closure_attach(h, c, every): from now on every nonlinear term of h
 * (cfd_nonlinear, cfd_substep, cfd_step) adds the network's correction: the
 * network evaluates the state in the nonlinear term of kRK = 1 of every
 * `every`-th step (and in the first one after attaching) and the other
 * substeps reuse that correction; c = NULL detaches.  The (0,0) column of
 * the u1 and u3 corrections loses its mean over y weighted by the plane
 * widths, so the correction exerts no net force on the flow.  Both paths:
 * the correction is formed in FP64 and, on the mixed path (precision 1),
 * added to its FP32 nonlinear-term storage; the network volume must be
 * nyu x Ngalz x Ngalx of h.
 *
 * cfd_closure_phys(h, which, out): the physical planes of the last sample
 * or application: which 0..2 the velocities u1, u2, u3, 3..5 the
 * corrections (planes of the component's grid x Ngalz x Ngalx). */
int cfd_closure_snapshot(cfd_handle fine, cfd_handle coarse, cfd_cnn c, int slot, double* rms);
int cfd_closure_attach(cfd_handle h, cfd_cnn c, int every);
/* backscatter clipping of the attached correction (default off): where a
 * component's correction c and velocity u have u c < 0 (the correction, which
 * enters the momentum equation as -c, would feed energy into the resolved
 * field) that component of the correction is zero; the velocity of a point
 * is the physical value the network reads there. */
int cfd_closure_clip(cfd_handle h, int clip);
int cfd_closure_phys(cfd_handle h, int which, double* out);

/* Ensembles (no counterpart in the reference: many independent channels,
 * e.g. of uneven grids, stepped together).  cfd_ensemble_create captures,
 * for each kRK, one HIP graph that runs the substep of every member, each on
 * its own stream (after kRK = 3, the flow-rate integral of members with
 * flag_ctpress = 1, as cfd_step does); cfd_ensemble_substep / _step launch
 * the graphs on the ensemble's stream.  Members keep their state and
 * handles: synchronise the ensemble (cfd_ensemble_sync) before reading or
 * driving a member directly.  Every member must have run one substep
 * eagerly before (launch attributes are set on first use), with profiling
 * off, no closure attached and no statistics request pending; the graphs
 * hold kernel nodes only. */
typedef struct cfd_ensemble_s* cfd_ensemble;
int cfd_ensemble_create(cfd_handle* members, int n, cfd_ensemble* out);
/* The same substeps with batched kernels: the z-inverse, x-pass and
 * z-forward of the nonlinear term and the two fused column sweeps run as one
 * launch per kernel variant (transform length, precision, segment shape)
 * over a table of the members (their arguments and block ranges, in device
 * memory owned by the ensemble), so the members' blocks of a kernel class
 * share a launch; each member's own kernels (CFL, IB correction, flow-rate
 * correction, filter) run between them.  Each kRK graph is built node by
 * node: a batch depends only on the latest work of its own members.  Every
 * member must use the four-step transforms and the segmented column sweeps
 * (single resolution, no last-plane replay); otherwise CFD_EINVAL names the
 * member.  Results are bit-identical to cfd_ensemble_create's graphs and to
 * eager stepping. */
int cfd_ensemble_create_batched(cfd_handle* members, int n, cfd_ensemble* out);
int cfd_ensemble_destroy(cfd_ensemble e);
const char* cfd_ensemble_last_error(cfd_ensemble e);
int cfd_ensemble_substep(cfd_ensemble e, int kRK);
int cfd_ensemble_step(cfd_ensemble e, int nsteps);
int cfd_ensemble_sync(cfd_ensemble e);

#ifdef __cplusplus
}
#endif
#endif
