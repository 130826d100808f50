// api.cpp -- C ABI of the rough-channel substep (include/cfd_amd.h): setup,
// device state, layout conversion at the boundary and operator sequencing.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "cnn.hpp"
#include "common.hpp"
#include "io.hpp"

using namespace cfd;

struct cfd_solver : public Solver {
  double alloc_bytes = 0.0;  // device bytes held by this handle
  // event pool for kernel-class timing
  std::vector<hipEvent_t> evpool;
  std::vector<int> evclass;
  size_t evn = 0;
  // per-field offsets into the IB job tables
  int ibz_off[4] = {0, 0, 0, 0}, ibx_off[4] = {0, 0, 0, 0};
  std::vector<void*> allocs;
  // the interpolating nonlinear z-inverse jobs, per plane: u2 on u planes,
  // u3 on v planes (reused by the statistics)
  std::vector<ZInvJob> nl_u2i, nl_u3i;
  // statistics and spectra (stats.f90, spectra.f90), set up on first use
  struct Stats {
    bool ready = false;
    int N[20] = {};
    int dnx = 0, dnz = 0, nyc[3] = {0, 0, 0}, istat = 0, chunk = 0;
    int bot[3][2] = {}, top[3][2] = {}, jref_bot[3] = {}, jref_top[3] = {};
    double *prof[13] = {}, *cb[13] = {}, *ct[13] = {}, *sp[5] = {};
    double2* Zc[3] = {};
    double* Pc[3] = {};
    double2* wxs = nullptr;
    ZInvJob *d_zj[3] = {}, *d_ju2 = nullptr;
    XJob* d_xj[3] = {};
    StatSumTask* d_sum[3] = {};
    StatCondTask* d_cond[3] = {};
    int nsum[3] = {0, 0, 0}, ncond[3] = {0, 0, 0};
    // sample / write requested for the next substep with kRK = 1
    int pending = 0, pending_iter = 0;
    std::string pending_dir;
  } st;
  // learned closure: buffers between the solver and the network volume
  struct Closure {
    bool ready = false;
    cfd_cnn_s* cnn = nullptr;  // attached network (applied in every nonlinear term)
    double2* Z = nullptr;      // one field of half-transformed planes
    double* Pu[3] = {};        // physical u1 (u planes), u2 (v planes), u3 (u planes)
    double* Pc[3] = {};        // physical corrections of the three components
    double2* Cs[3] = {};       // spectral corrections
    unsigned char* mask = nullptr;  // 1 at the forcing points of the u grid
    ZInvJob* d_zj[6] = {};     // z-inverse jobs: u1, u2, u3 (fields), C1, C2, C3
    XJob *d_xi[6] = {}, *d_xf[3] = {};
    ZFwdJob* d_zf[3] = {};
    int nzj[6] = {}, nxi[6] = {}, nxf[3] = {};
    double* acc = nullptr;     // 6 sums of squares of a sample
    int every = 1;             // the correction is recomputed at kRK = 1 of every `every`-th step
    int count = 0;             // steps since attaching
    bool have = false;         // Cs holds a correction
    bool clip = false;         // backscatter clipping of the correction
  } cl;
};

namespace {

thread_local std::string g_err;

int fail(cfd_solver* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  g_err = msg;
  return code;
}

#define HIPCHK(s, x)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail((s), CFD_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(cfd_solver* s, T** p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) return fail(s, CFD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->alloc_bytes += (double)n * sizeof(T);
  hipMemset(q, 0, n * sizeof(T));
  s->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return 0;
}

template <typename T>
int upload(cfd_solver* s, T** p, const std::vector<T>& h) {
  if (int rc = dalloc(s, p, h.size())) return rc;
  if (!h.empty()) HIPCHK(s, hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

bool make_plan(int L, FftPlan& P) {
  P.L = L;
  P.nr = 0;
  int m = L;
  const int tryr[4] = {4, 2, 3, 5};
  for (int t = 0; t < 4 && m > 1;) {
    if (m % tryr[t] == 0) {
      if (P.nr >= kMaxRadix) return false;
      P.rad[P.nr++] = tryr[t];
      m /= tryr[t];
    } else {
      ++t;
    }
  }
  return m == 1;
}

std::vector<double2> twiddles(int L) {
  std::vector<double2> w(L);
  for (int m = 0; m < L; ++m) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / (long double)L;
    w[m] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return w;
}

std::vector<float2> twiddles_f(int L) {
  std::vector<float2> w(L);
  for (int m = 0; m < L; ++m) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / (long double)L;
    w[m] = make_float2((float)cosl(a), (float)sinl(a));
  }
  return w;
}

// untyped device buffer of n elements of esz bytes, zeroed
int balloc(cfd_solver* s, void** p, size_t n, size_t esz) {
  char* q = nullptr;
  if (int rc = dalloc(s, &q, (n ? n : 1) * esz)) return rc;
  *p = q;
  return 0;
}

int grid_of(int which) {
  static const int g[CFD_NFIELDS] = {0, 1, 0, 2, 0, 1, 0, 0, 1, 0, 2, 2};
  return g[which];
}

int ny_of(const Solver& s, int which) {
  const int g = grid_of(which);
  return g == 0 ? s.nyu : g == 1 ? s.nyv : s.nyp;
}

// kernel-class timing
struct Timed {
  cfd_solver* s;
  int c;
  Timed(cfd_solver* s_, int c_) : s(s_), c(c_) {
    if (!s->prof) return;
    if (s->evn + 2 > s->evpool.size()) {
      // drain the pool
      hipStreamSynchronize(s->stream);
      for (size_t q = 0; q + 1 < s->evn; q += 2) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, s->evpool[q], s->evpool[q + 1]);
        s->prof_ms[s->evclass[q / 2]] += ms;
        s->prof_cnt[s->evclass[q / 2]] += 1;
      }
      s->evn = 0;
    }
    hipEventRecord(s->evpool[s->evn], s->stream);
  }
  ~Timed() {
    if (!s->prof) return;
    hipEventRecord(s->evpool[s->evn + 1], s->stream);
    s->evclass[s->evn / 2] = c;
    s->evn += 2;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// create / destroy
// ---------------------------------------------------------------------------
namespace {
int tuning_check(const cfd_tuning& t, std::string& err) {
  if (t.fft != 0 && t.fft != 1) err = "tuning: fft must be 0 (four-step where it applies) or 1 (Stockham)";
  else if (!seg_rows_supported(t.seg_rows)) err = "tuning: seg_rows must be 5, 9, 13 or 21";
  else if (t.solve_threads != 64 && t.solve_threads != 128 && t.solve_threads != 256)
    err = "tuning: solve_threads must be 64, 128 or 256";
  else return CFD_OK;
  return CFD_EINVAL;
}
int create_impl(const cfd_params* pp, const cfd_tuning* tp, cfd_handle* out);
int setup_bands(cfd_solver* s);
}  // namespace

extern "C" int cfd_create(const cfd_params* pp, cfd_handle* out) { return create_impl(pp, nullptr, out); }

extern "C" int cfd_create_tuned(const cfd_params* pp, const cfd_tuning* t, cfd_handle* out) {
  if (!t) return fail(nullptr, CFD_EINVAL, "null tuning");
  std::string err;
  if (tuning_check(*t, err)) return fail(nullptr, CFD_EINVAL, err);
  return create_impl(pp, t, out);
}

extern "C" int cfd_get_tuning(cfd_handle h, cfd_tuning* t) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !t) return fail(s, CFD_EINVAL, "null argument");
  t->fft = s->fft_stockham ? 1 : 0;
  t->seg_rows = s->seg_rows;
  t->solve_threads = s->solve_threads;
  return CFD_OK;
}

extern "C" int cfd_set_tuning(cfd_handle h, const cfd_tuning* t) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !t) return fail(s, CFD_EINVAL, "null argument");
  std::string err;
  if (tuning_check(*t, err)) return fail(s, CFD_EINVAL, err);
  if ((t->fft == 1) != s->fft_stockham) return fail(s, CFD_EINVAL, "tuning: the FFT engine is fixed at creation");
  if (s->mb && (s->nyu + t->seg_rows + 1) / (t->seg_rows + 1) > 256)
    return fail(s, CFD_EINVAL, "tuning: multi-band resolution needs the segmented solves (more rows per segment)");
  s->seg_rows = t->seg_rows;
  s->solve_threads = t->solve_threads;
  return CFD_OK;
}

namespace {
int create_impl(const cfd_params* pp, const cfd_tuning* tp, cfd_handle* out) {
  if (!pp || !out) return fail(nullptr, CFD_EINVAL, "null argument");
  *out = nullptr;
  const cfd_params& p = *pp;
  if (p.ngalx < 6 || p.ngalz < 6 || p.ny < 4 || p.dsty < 0) return fail(nullptr, CFD_EINVAL, "grid sizes out of range");
  if (p.ngalx % 2 || p.ngalz % 2) return fail(nullptr, CFD_EINVAL, "Ngalx and Ngalz must be even");
  if (p.ngalx > 1024 || p.ngalz > 768)
    return fail(nullptr, CFD_EINVAL, "Ngalx <= 1024 and Ngalz <= 768 (LDS-resident transforms)");
  if (p.precision != 0 && p.precision != 1)
    return fail(nullptr, CFD_EINVAL, "precision must be 0 (FP64) or 1 (mixed: FP32 nonlinear term)");
  if (p.precision == 1 && !(cfd::fft4_x_supported(p.ngalx, true) && cfd::fft4_z_supported(p.ngalz, true)))
    return fail(nullptr, CFD_EINVAL, "the mixed path needs Ngalx and Ngalz with four-step FFT kernels");
  if (p.Re <= 0 || p.alp <= 0 || p.bet <= 0) return fail(nullptr, CFD_EINVAL, "Re, alp, bet must be positive");
  cfd_solver* s = new (std::nothrow) cfd_solver();
  if (!s) return fail(nullptr, CFD_ENOMEM, "host allocation failed");
  auto bail = [&](int rc) {
    g_err = s->err;
    cfd_destroy(s);
    return rc;
  };
  s->P = p;
  s->dev = p.device;
  if (hipSetDevice(p.device) != hipSuccess) {
    s->err = "hipSetDevice failed";
    return bail(CFD_EHIP);
  }
  // the device address of the no-allocation scratch, once per device
  if (int rc = noalloc_init(p.device, s->err)) return bail(rc);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
    s->err = "hipStreamCreate failed";
    return bail(CFD_EHIP);
  }
  s->ngx = p.ngalx;
  s->ngz = p.ngalz;
  s->ny = p.ny;
  s->dsty = p.dsty;
  s->N1 = 2 * (p.ngalx / 3);  // start.f90:108-111
  s->N2 = 2 * (p.ngalz / 3);
  s->Mx = s->N1 / 2 + 1;
  s->Ncol = s->Mx * s->N2;
  s->nn = p.ny - 2;
  // the central band's resolution (start.f90:89-111): multi-band when coarser
  s->ngx2 = p.ngalx2 > 0 ? p.ngalx2 : s->ngx;
  s->ngz2 = p.ngalz2 > 0 ? p.ngalz2 : s->ngz;
  s->N1c = 2 * (s->ngx2 / 3);
  s->N2c = 2 * (s->ngz2 / 3);
  s->mb = s->ngx2 != s->ngx || s->ngz2 != s->ngz;
  std::string err;
  if (!build_ygrid(p, s->yg, err) || !build_iblists(p, s->ib, err)) {
    s->err = err;
    return bail(CFD_EINVAL);
  }
  // a folded grid has no solution to converge to: refuse it (the reference
  // runs on it)
  if (const int jf = ygrid_fold(s->yg); jf != INT_MIN) {
    s->err = "the y grid folds back at face j = " + std::to_string(jf) +
             ": yu(j) < yv(j) < yu(j+1) fails (canopy depth posth does not fit dsty, or the map is not monotonic)";
    return bail(CFD_EINVAL);
  }
  s->nyu = s->yg.nyu;
  s->nyv = s->yg.nyv;
  s->nyp = s->nyu - 2;
  s->nl_x[1] = s->nl_u[1] = s->nyu;  // every plane through the nonlinear transforms
  s->nl_v[1] = s->nyv;
  s->jlo_u = s->yg.jlo_u;
  s->jlo_v = s->yg.jlo_v;
  s->jlo_p = s->jlo_u + 1;
  const YGrid& g = s->yg;
  // ghost weights (start.f90:494-508), band 2 = the whole column here
  const int N40 = -p.dsty, N30 = -p.dsty, N43 = p.ny - 1 + p.dsty, N33 = p.ny - 2 + p.dsty;
  s->gw[0] = -(g.YV(N30) - g.YU(N40)) / (g.YV(N30) - g.YU(N40 + 1));
  s->gw[1] = (g.YU(N43 + 1) - g.YV(N33 + 1)) / (g.YV(N33 + 1) - g.YU(N43));
  s->gwi[0] = (g.YU(N40) - g.YU(N40 + 1)) / (g.YV(N30) - g.YU(N40 + 1));
  s->gwi[1] = (g.YU(N43 + 1) - g.YU(N43)) / (g.YV(N33 + 1) - g.YU(N43));
  // wavenumbers (def_k, start.f90:965-986); first derivative stored as i*k1
  s->k1x.resize(s->Mx);
  s->k2x.resize(s->Mx);
  s->k1z.resize(s->N2);
  s->k2z.resize(s->N2);
  for (int i = 0; i < s->Mx; ++i) {
    s->k1x[i] = p.alp * i;
    s->k2x[i] = -((p.alp * i) * (p.alp * i));
  }
  for (int k = 0; k < s->N2; ++k) {
    const double kz = (k < s->N2 / 2) ? p.bet * k : -(p.bet * (s->N2 - k));
    s->k1z[k] = kz;
    s->k2z[k] = -(kz * kz);
  }
  // RK3 (start.f90:395-409)
  const double a[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0}, gg[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double c[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0}, d[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
  for (int q = 0; q < 3; ++q) {
    s->aRK[q] = a[q];
    s->bRK[q] = a[q];
    s->gRK[q] = gg[q];
    s->cRK[q] = c[q];
    s->dRK[q] = d[q];
  }
  const double h0 = g.YV(N30 + 1) - g.YV(N30);
  s->dtv = p.Re / (-s->k2x[s->N1 / 2] - s->k2z[s->N2 / 2 - 1] + 4.0 / (h0 * h0));  // start.f90:432

  // FFT plans
  if (!make_plan(s->ngz, s->pz) || !make_plan(s->ngx, s->px)) {
    s->err = "Ngalx and Ngalz must factor into 2, 3 and 5";
    return bail(CFD_EINVAL);
  }
  if (tp) {
    s->fft_stockham = tp->fft == 1;
    s->seg_rows = tp->seg_rows;
    s->solve_threads = tp->solve_threads;
  } else {
    const char* e = getenv("CFD_AMD_FFT");
    s->fft_stockham = e && std::string(e) == "stockham";
  }
  s->mixed = p.precision == 1;
  if (s->mixed && tp && tp->fft == 1) {
    s->err = "the mixed path runs the four-step FFT kernels only";
    return bail(CFD_EINVAL);
  }
  if (s->mixed) s->fft_stockham = false;
  int rc;
  if ((rc = upload(s, &s->d_twz, twiddles(s->ngz))) || (rc = upload(s, &s->d_twx, twiddles(s->ngx))) ||
      (rc = upload(s, &s->d_twzf, twiddles_f(s->ngz))) || (rc = upload(s, &s->d_twxf, twiddles_f(s->ngx))) ||
      (rc = dalloc(s, &s->d_zrow, (size_t)s->ngx)))
    return bail(rc);

  // metrics
  std::vector<double> plo(s->nyp), pup(s->nyp), pdg(s->nyp);
  {
    // pressure operator D_y G_y rows (LU_buildP, tridLU_3D_IB.f90:417-466)
    const int ns = s->jlo_p, ne = s->jlo_p + s->nyp - 1;
    auto DU = [&](int j) { return g.dthdyu[j - g.jlo_u]; };
    auto DV = [&](int j) { return g.dthdyv[j - g.jlo_v]; };
    auto Dvb = [&](int j) { return -1.0 * DU(j + 1); };
    auto Gvb = [&](int j) { return -1.0 * DV(j); };
    auto Dvt = [&](int j) { return 1.0 * DU(j); };
    auto Gvt = [&](int j) { return 1.0 * DV(j - 1); };
    for (int j = ns; j <= ne; ++j) {
      const int q = j - ns;
      if (j == ns) {
        pdg[q] = Dvt(ns) * Gvb(ns);
        pup[q] = Dvt(ns) * Gvt(ns + 1);
        plo[q] = 0.0;
      } else if (j == ne) {
        pdg[q] = Dvb(ne - 1) * Gvt(ne);
        plo[q] = Dvb(ne - 1) * Gvb(ne - 1);
        pup[q] = 0.0;
      } else {
        pdg[q] = (Dvb(j - 1) * Gvt(j)) + (Dvt(j) * Gvb(j));
        plo[q] = Dvb(j - 1) * Gvb(j - 1);
        pup[q] = Dvt(j) * Gvt(j + 1);
      }
    }
  }
  if ((rc = upload(s, &s->d_dthdyu, g.dthdyu)) || (rc = upload(s, &s->d_dthdyv, g.dthdyv)) ||
      (rc = upload(s, &s->d_d2u, g.d2u)) || (rc = upload(s, &s->d_d2v, g.d2v)) || (rc = upload(s, &s->d_yv, g.yv)) ||
      (rc = upload(s, &s->d_k1x, s->k1x)) || (rc = upload(s, &s->d_k2x, s->k2x)) ||
      (rc = upload(s, &s->d_k1z, s->k1z)) || (rc = upload(s, &s->d_k2z, s->k2z)) || (rc = upload(s, &s->d_plo, plo)) ||
      (rc = upload(s, &s->d_pup, pup)) || (rc = upload(s, &s->d_pdg, pdg)) ||
      (rc = dalloc(s, &s->d_mf, (size_t)4 * (s->ny + 2))) || (rc = dalloc(s, &s->d_sc, 1)) ||
      (rc = dalloc(s, &s->d_u11, (size_t)s->nn + 4)))
    return bail(rc);

  // fields and scratch
  const size_t N = s->Ncol;
  for (int f = 0; f < CFD_NFIELDS; ++f)
    if ((rc = dalloc(s, &s->F[f], (size_t)ny_of(*s, f) * N))) return bail(rc);
  const size_t esz = s->mixed ? sizeof(float2) : sizeof(double2);  // nonlinear-term storage
  if ((rc = balloc(s, &s->nxz1, s->nyu * N, esz)) || (rc = balloc(s, &s->nxz3, s->nyu * N, esz)) ||
      (rc = balloc(s, &s->fvv, s->nyu * N, esz)) || (rc = balloc(s, &s->nxz2, s->nyv * N, esz)) ||
      (rc = balloc(s, &s->fuv, s->nyv * N, esz)) || (rc = balloc(s, &s->fvw, s->nyv * N, esz)) ||
      (rc = dalloc(s, &s->cscr, (size_t)3 * s->nyu * N)))
    return bail(rc);
  for (int f = 0; f < 3; ++f)
    if ((rc = dalloc(s, &s->ysc[f], (size_t)s->nyu * N))) return bail(rc);
  // Z rows padded to a multiple of 128 bytes when the four-step kernels own
  // all three nonlinear passes (the Stockham fallback reads rows of Mx)
  {
    const bool four = (s->mixed || !s->fft_stockham) && cfd::fft4_z_supported(s->ngz, s->mixed) &&
                      cfd::fft4_x_supported(s->ngx, s->mixed) && s->N2 == 2 * (s->ngz / 3) &&
                      s->N1 == 2 * (s->ngx / 3);
    const int per = (int)(128 / esz);
    s->zs = four ? (s->Mx + per - 1) / per * per : s->Mx;
  }
  const size_t zp = (size_t)s->ngz * s->zs;
  const int zplanes[12] = {s->nyu, s->nyu, s->nyu, s->nyv, s->nyv, s->nyv,
                           s->nyu, s->nyu, s->nyu, s->nyu, s->nyv, s->nyv};
  for (int q = 0; q < 12; ++q)
    if ((rc = balloc(s, &s->Z[q], zplanes[q] * zp, esz))) return bail(rc);

  // nonlinear z-inverse jobs (interp_u / interp_v folded in, FOU3D.f90:261-308)
  {
    std::vector<ZInvJob> J;
    auto job = [&](const double2* src, void* dst, int pd, int mode, int pa, int pb, int pc, double wa, double wb,
                   double gw) {
      ZInvJob z{};
      z.src = src;
      z.dst = dst;
      z.pa = pa;
      z.pb = pb;
      z.pc = pc;
      z.pd = pd;
      z.mode = mode;
      z.wa = wa;
      z.wb = wb;
      z.g = gw;
      J.push_back(z);
    };
    auto vw = [&](int j, double& wa, double& wb) {  // centre j from faces j (wa) and j-1 (wb)
      const double den = g.YV(j) - g.YV(j - 1);
      wa = (g.YU(j) - g.YV(j - 1)) / den;
      wb = (g.YV(j) - g.YU(j)) / den;
    };
    // ordered by y-plane: the jobs of u-plane q and v-plane q read the
    // velocity planes q-1..q+1, so neighbouring jobs share their sources
    for (int q = 0; q < std::max(s->nyu, s->nyv); ++q) {
      if (q < s->nyu) {
        const int j = s->jlo_u + q;
        job(s->F[CFD_U1], s->Z[0], q, 0, q, 0, 0, 0, 0, 0);
        job(s->F[CFD_U3], s->Z[2], q, 0, q, 0, 0, 0, 0, 0);
        double wa, wb;
        if (q == 0) {
          vw(j + 1, wa, wb);
          job(s->F[CFD_U2], s->Z[1], q, 2, 1, 0, 0, wa, wb, s->gwi[0]);
        } else if (q == s->nyu - 1) {
          vw(j - 1, wa, wb);
          job(s->F[CFD_U2], s->Z[1], q, 2, s->nyv - 1, s->nyv - 2, s->nyv - 1, wa, wb, s->gwi[1]);
        } else {
          vw(j, wa, wb);
          job(s->F[CFD_U2], s->Z[1], q, 1, q, q - 1, 0, wa, wb, 0);
        }
      }
      if (q < s->nyv) {
        const int j = s->jlo_v + q;
        const double den = g.YU(j + 1) - g.YU(j);
        const double wa = (g.YV(j) - g.YU(j)) / den, wb = (g.YU(j + 1) - g.YV(j)) / den;
        job(s->F[CFD_U1], s->Z[3], q, 1, q + 1, q, 0, wa, wb, 0);
        job(s->F[CFD_U2], s->Z[4], q, 0, q, 0, 0, 0, 0, 0);
        job(s->F[CFD_U3], s->Z[5], q, 1, q + 1, q, 0, wa, wb, 0);
      }
    }
    s->n_nljobs = (int)J.size();
    if ((rc = upload(s, &s->d_nljobs, J))) return bail(rc);
    s->h_nljobs = J;
    for (const ZInvJob& z : J) {
      if (z.dst == s->Z[1]) s->nl_u2i.push_back(z);
      if (z.dst == s->Z[5]) s->nl_u3i.push_back(z);
    }
  }

  // IB lists and transform jobs
  for (int gr = 0; gr < 2; ++gr) {
    std::vector<IBPoint> all = s->ib[gr][0];
    all.insert(all.end(), s->ib[gr][1].begin(), s->ib[gr][1].end());
    const int jlo = gr == 0 ? s->jlo_u : s->jlo_v, ny_g = gr == 0 ? s->nyu : s->nyv;
    std::vector<int>& pl = s->ibplanes[gr];
    for (const IBPoint& e : all) pl.push_back(e.j);
    std::sort(pl.begin(), pl.end());
    pl.erase(std::unique(pl.begin(), pl.end()), pl.end());
    std::vector<IBDev> dl;
    std::vector<int> seg;
    for (size_t q = 0; q < all.size(); ++q) {
      const IBPoint& e = all[q];
      auto inr = [&](int i, int k) { return i >= 1 && i <= s->ngx && k >= 1 && k <= s->ngz; };
      if (e.j2 != e.j || e.j3 != e.j || !inr(e.i, e.k) || !inr(e.i2, e.k2) || !inr(e.i3, e.k3) || e.j - jlo < 1 ||
          e.j - jlo > ny_g - 2) {
        s->err = "IB list entry outside the grid";
        return bail(CFD_EINVAL);
      }
      IBDev d{};
      d.o = (e.k - 1) * s->ngx + (e.i - 1);
      d.o2 = (e.k2 - 1) * s->ngx + (e.i2 - 1);
      d.o3 = (e.k3 - 1) * s->ngx + (e.i3 - 1);
      d.plane = int(std::lower_bound(pl.begin(), pl.end(), e.j) - pl.begin());
      d.lap = e.lap;
      d.w2 = e.w2;
      d.w3 = e.w3;
      if (q == 0 || e.i != all[q - 1].i || e.k != all[q - 1].k || e.j != all[q - 1].j) seg.push_back((int)q);
      dl.push_back(d);
    }
    seg.push_back((int)all.size());
    s->nibseg[gr] = (int)seg.size() - 1;
    if ((rc = upload(s, &s->d_ibl[gr], dl)) || (rc = upload(s, &s->d_ibseg[gr], seg))) return bail(rc);
  }
  {
    std::vector<ZInvJob> zj;
    std::vector<XJob> xi, xf;
    std::vector<ZFwdJob> zf;
    for (int f = 0; f < 3; ++f) {
      const int gr = f == 1 ? 1 : 0;
      const int jlo = gr == 0 ? s->jlo_u : s->jlo_v;
      const int nb = (int)s->ibplanes[gr].size();
      s->nibp[f] = nb;
      if ((rc = dalloc(s, &s->zib_r[f], (size_t)nb * zp)) || (rc = dalloc(s, &s->zib_u[f], (size_t)nb * zp)) ||
          (rc = dalloc(s, &s->pib_r[f], (size_t)nb * s->ngz * s->ngx)) ||
          (rc = dalloc(s, &s->pib_u[f], (size_t)nb * s->ngz * s->ngx)))
        return bail(rc);
      s->ibz_off[f] = (int)zj.size() / 2;  // jobs come in (r, u) pairs
      s->ibx_off[f] = (int)xi.size();
      for (int b = 0; b < nb; ++b) {
        const int plane = s->ibplanes[gr][b] - jlo;
        ZInvJob r{}, u{};
        r.src = s->F[CFD_DU1 + f];
        r.dst = s->zib_r[f];
        r.pa = plane;
        r.pd = b;
        u.src = s->F[CFD_U1 + f];
        u.dst = s->zib_u[f];
        u.pa = plane;
        u.pd = b;
        zj.push_back(r);
        zj.push_back(u);
        XJob x{};
        x.za = s->zib_r[f];
        x.zb = s->zib_u[f];
        x.pa = s->pib_r[f];
        x.pb = s->pib_u[f];
        x.plane = b;
        xi.push_back(x);
        XJob y{};
        y.pa = s->pib_r[f];
        y.zo = s->zib_r[f];
        y.plane = b;
        xf.push_back(y);
        ZFwdJob z{};
        z.src = s->zib_r[f];
        z.dst = s->F[CFD_DU1 + f];
        z.ps = b;
        z.pd = plane;
        zf.push_back(z);
      }
    }
    s->ibz_off[3] = (int)zj.size() / 2;
    s->ibx_off[3] = (int)xi.size();
    s->n_ibz = (int)zj.size();
    s->n_ibxi = (int)xi.size();
    s->n_ibxf = (int)xf.size();
    s->n_ibzf = (int)zf.size();
    if ((rc = upload(s, &s->d_ibz, zj)) || (rc = upload(s, &s->d_ibxi, xi)) || (rc = upload(s, &s->d_ibxf, xf)) ||
        (rc = upload(s, &s->d_ibzf, zf)))
      return bail(rc);
  }
  // IB plane masks and plane-index lists of the fused sweeps
  for (int f = 0; f < 3; ++f) {
    const int gr = f == 1 ? 1 : 0;
    const int jlo = gr == 0 ? s->jlo_u : s->jlo_v;
    std::vector<unsigned char> mask(s->nyu, 0);
    std::vector<int> q;
    for (int j : s->ibplanes[gr]) {
      mask[j - jlo] = 1;
      q.push_back(j - jlo);
    }
    if ((rc = upload(s, &s->d_ibmask[f], mask)) || (rc = upload(s, &s->d_ibq[f], q))) return bail(rc);
  }
  // multi-band resolution: short columns, the central solver, CFL plane sets
  if (s->mb)
    if ((rc = setup_bands(s))) return bail(rc);
  // scalars
  {
    DevScalars h{};
    h.mpgx = p.mpgx;
    HIPCHK(s, hipMemcpy(s->d_sc, &h, sizeof h, hipMemcpyHostToDevice));
  }
  s->evpool.resize(4096);
  s->evclass.resize(2048);
  for (auto& e : s->evpool)
    if (hipEventCreate(&e) != hipSuccess) {
      s->err = "hipEventCreate failed";
      return bail(CFD_EHIP);
    }
  HIPCHK(s, hipDeviceSynchronize());
  *out = s;
  return CFD_OK;
}

// Multi-band resolution (start.f90:89-111, 1420-1557): the band limits, the
// short columns (the wall bands' modes outside the central band's N1c x N2c,
// proc_lims_columns start.f90:1424-1501), the interface boundary weights
// (start.f90:494-500), and the central solver: the same y grid at the central
// band's x/z resolution, on this solver's stream, whose nonlinear transforms
// give the central planes' products over the long columns.
int setup_bands(cfd_solver* s) {
  const cfd_params& p = s->P;
  if (s->ngx2 > s->ngx || s->ngz2 > s->ngz || s->ngx2 % 2 || s->ngz2 % 2 || s->ngx2 < 6 || s->ngz2 < 6)
    return fail(s, CFD_EINVAL, "the central band's Ngalx2, Ngalz2 must be even, at least 6 and at most Ngalx, Ngalz");
  int N[20];
  band_table(p, s->yg, N);
  for (int b = 0; b < 4; ++b) {
    s->N3[b] = N[4 * b + 2];
    s->N4[b] = N[4 * b + 3];
  }
  // each wall band keeps its own rows: the short columns' interface ghost
  // rows N(4,1)+1 and N(4,2) must differ
  if (!(s->N3[0] < s->N3[1] && s->N3[1] + 2 < s->N3[2] && s->N3[2] < s->N3[3]))
    return fail(s, CFD_EINVAL, "the band limits leave no central band for a second resolution");
  for (int g = 0; g < 2; ++g)
    for (int j : s->ibplanes[g])
      if (g == 0 ? (j > s->N4[1] && j <= s->N4[2]) : (j > s->N3[1] && j <= s->N3[2]))
        return fail(s, CFD_EINVAL, "immersed-boundary planes in the central band");
  const int R = s->seg_rows;
  if ((s->nyu + R + 1) / (R + 1) > 256) return fail(s, CFD_EINVAL, "multi-band resolution: too many planes");
  const YGrid& g = s->yg;
  s->gw12 = (g.YU(s->N4[1] + 1) - g.YV(s->N3[1] + 1)) / (g.YV(s->N3[1] + 1) - g.YU(s->N4[1]));
  s->gw31 = -(g.YV(s->N3[2]) - g.YU(s->N4[2])) / (g.YV(s->N3[2]) - g.YU(s->N4[2] + 1));
  // column sets
  s->h_short.assign(s->Ncol, 0);
  std::vector<int> ccol(s->Ncol, -1), lcol;
  const int Mxc = s->N1c / 2 + 1;
  for (int col = 0; col < s->Ncol; ++col) {
    const int i = col % s->Mx, k = col / s->Mx;
    const bool kin = k < s->N2c / 2 || k >= s->N2 - s->N2c / 2;
    if (i <= s->N1c / 2 && kin) {
      const int kc = k < s->N2c / 2 ? k : k - s->N2 + s->N2c;  // column order of the central solver
      ccol[col] = kc * Mxc + i;
    } else {
      s->h_short[col] = 1;
    }
  }
  lcol.assign((size_t)Mxc * s->N2c, -1);
  for (int col = 0; col < s->Ncol; ++col)
    if (ccol[col] >= 0) lcol[ccol[col]] = col;
  // CFL plane sets: the wall bands' planes here, the central planes there
  std::vector<unsigned char> fu(s->nyu), fv(s->nyv), cu(s->nyu), cv(s->nyv);
  for (int q = 0; q < s->nyu; ++q) {
    const int j = s->jlo_u + q;
    cu[q] = j > s->N4[1] && j <= s->N4[2];
    fu[q] = !cu[q];
  }
  for (int q = 0; q < s->nyv; ++q) {
    const int j = s->jlo_v + q;
    cv[q] = j > s->N3[1] && j <= s->N3[2];
    fv[q] = !cv[q];
  }
  int rc;
  // The planes split between the two resolutions.  Central u planes [u0,
  // u1) and v planes [v0, v1) take their products from the central solver:
  // this solver's z-forward leaves them out, and its x-pass the planes [x0,
  // x1) where both the u and the v plane are central (plane N(3,1)+1 keeps
  // its fine u products); the central solver's x-pass covers [c0, c1), the
  // planes with a central u or v plane, its z-forward only the central
  // planes, and its fields take the long columns of the planes the
  // interpolations of those read, [c0 - 1, c1 + 1).  Each z-inverse job list
  // keeps the jobs of its x-pass planes.
  const int u0 = s->N4[1] + 1 - s->jlo_u, u1 = s->N4[2] + 1 - s->jlo_u;
  const int v0 = s->N3[1] + 1 - s->jlo_v, v1 = s->N3[2] + 1 - s->jlo_v;
  const int x0 = std::max(u0, v0), x1 = std::max(x0, std::min(u1, v1));
  const int c0 = std::min(u0, v0), c1 = std::max(u1, v1);
  auto set_win = [](int* w, int lo, int hi, int s0, int s1) {
    w[0] = lo;
    w[1] = hi;
    w[2] = s0;
    w[3] = s1;
  };
  set_win(s->nl_x, 0, s->nyu, x0, x1);
  set_win(s->nl_u, 0, s->nyu, u0, u1);
  set_win(s->nl_v, 0, s->nyv, v0, v1);
  s->gather_q[0] = std::max(c0 - 1, 0);
  s->gather_q[1] = std::min(c1 + 1, s->nyu);
  auto keep_jobs = [&](cfd_solver* t, int lo, int hi, int s0, int s1) {
    std::vector<ZInvJob> J;
    for (const ZInvJob& z : t->h_nljobs)
      if (z.pd >= lo && z.pd < hi && (z.pd < s0 || z.pd >= s1)) J.push_back(z);
    t->n_nljobs = (int)J.size();
    return upload(t, &t->d_nljobs, J);
  };
  if ((rc = keep_jobs(s, 0, s->nyu, x0, x1))) return rc;
  if ((rc = upload(s, &s->d_short, s->h_short)) || (rc = upload(s, &s->d_ccol, ccol)) ||
      (rc = upload(s, &s->d_lcol, lcol)) || (rc = upload(s, &s->d_cfu, fu)) || (rc = upload(s, &s->d_cfv, fv)))
    return rc;
  // the central solver
  cfd_params pc = p;
  pc.ngalx = s->ngx2;
  pc.ngalz = s->ngz2;
  pc.ngalx2 = pc.ngalz2 = 0;
  pc.geometry_type = 0;  // its IB planes are none of its business
  const cfd_tuning t{s->fft_stockham ? 1 : 0, s->seg_rows, s->solve_threads};
  cfd_handle hc = nullptr;
  if ((rc = create_impl(&pc, &t, &hc))) return fail(s, rc, std::string("central band solver: ") + g_err);
  cfd_solver* c = static_cast<cfd_solver*>(hc);
  s->mbc = c;
  // its band table (statistics) is this solver's: same y grid, same
  // geometry type (its IB lists, built at creation, stay empty)
  c->P.geometry_type = p.geometry_type;
  if (c->nyu != s->nyu || c->nyv != s->nyv || c->Ncol != Mxc * s->N2c)
    return fail(s, CFD_EINVAL, "central band solver: planes or columns differ");
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(c->stream));
  HIPCHK(s, hipStreamDestroy(c->stream));
  c->stream = s->stream;
  c->borrowed_stream = true;
  if ((rc = upload(c, &c->d_cfu, cu)) || (rc = upload(c, &c->d_cfv, cv)) || (rc = keep_jobs(c, c0, c1, 0, 0)))
    return fail(s, rc, std::string("central band solver: ") + c->err);
  set_win(c->nl_x, c0, c1, 0, 0);
  set_win(c->nl_u, u0, u1, 0, 0);
  set_win(c->nl_v, v0, v1, 0, 0);
  s->alloc_bytes += c->alloc_bytes;
  return CFD_OK;
}
}  // namespace

extern "C" int cfd_build_ygrid(const cfd_params* p, double* yu, double* yv, double* du, double* dv, double* d2u,
                               double* d2v) {
  if (!p) return fail(nullptr, CFD_EINVAL, "null argument");
  YGrid g;
  std::string err;
  if (!build_ygrid(*p, g, err)) return fail(nullptr, CFD_EINVAL, err);
  if (yu) std::memcpy(yu, g.yu.data(), sizeof(double) * g.nyu);
  if (yv) std::memcpy(yv, g.yv.data(), sizeof(double) * g.nyv);
  if (du) std::memcpy(du, g.dthdyu.data(), sizeof(double) * g.nyu);
  if (dv) std::memcpy(dv, g.dthdyv.data(), sizeof(double) * g.nyv);
  if (d2u) std::memcpy(d2u, g.d2u.data(), sizeof(double) * 3 * g.nyu);
  if (d2v) std::memcpy(d2v, g.d2v.data(), sizeof(double) * 3 * g.nyv);
  return CFD_OK;
}

extern "C" int cfd_build_iblist(const cfd_params* p, int grid, int wall, int* l9, double* w3, int* n) {
  if (!p || grid < 0 || grid > 1 || wall < 0 || wall > 1) return fail(nullptr, CFD_EINVAL, "bad argument");
  std::vector<IBPoint> L[2][2];
  std::string err;
  if (!build_iblists(*p, L, err)) return fail(nullptr, CFD_EINVAL, err);
  const std::vector<IBPoint>& v = L[grid][wall];
  if (n) *n = (int)v.size();
  for (size_t q = 0; q < v.size(); ++q) {
    const IBPoint& e = v[q];
    if (l9) {
      const int a[9] = {e.i, e.k, e.j, e.i2, e.k2, e.j2, e.i3, e.k3, e.j3};
      std::memcpy(l9 + 9 * q, a, sizeof a);
    }
    if (w3) {
      w3[3 * q] = e.lap;
      w3[3 * q + 1] = e.w2;
      w3[3 * q + 2] = e.w3;
    }
  }
  return CFD_OK;
}

extern "C" int cfd_destroy(cfd_handle h) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return CFD_OK;
  hipSetDevice(s->dev);
  if (s->mbc) {  // the central solver runs on this stream: it goes first
    if (s->stream) hipStreamSynchronize(s->stream);
    cfd_destroy(static_cast<cfd_solver*>(s->mbc));
    s->mbc = nullptr;
  }
  // both streams drain before any buffer goes: the side stream's u11 solve
  // may still be writing when a substep returned between fork and join
  if (s->stream) hipStreamSynchronize(s->stream);
  if (s->side) hipStreamSynchronize(s->side);
  for (void* q : s->allocs) hipFree(q);
  for (auto& e : s->evpool)
    if (e) hipEventDestroy(e);
  if (s->ev_fork) hipEventDestroy(s->ev_fork);
  if (s->ev_join) hipEventDestroy(s->ev_join);
  if (s->side) hipStreamDestroy(s->side);
  if (s->stream && !s->borrowed_stream) hipStreamDestroy(s->stream);
  delete s;
  return CFD_OK;
}

extern "C" const char* cfd_last_error(cfd_handle h) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  return s ? s->err.c_str() : g_err.c_str();
}

extern "C" int cfd_dims(cfd_handle h, int* d) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !d) return fail(s, CFD_EINVAL, "null argument");
  const int v[16] = {s->N1, s->N2, s->Mx, s->Ncol, s->nyu, s->nyv, s->nyp, s->jlo_u,
                     s->jlo_v, s->jlo_p, s->ngx, s->ngz, s->ny, s->dsty,
                     (int)(s->ib[0][0].size() + s->ib[0][1].size()), (int)(s->ib[1][0].size() + s->ib[1][1].size())};
  std::memcpy(d, v, sizeof v);
  return CFD_OK;
}

extern "C" int cfd_get_ygrid(cfd_handle h, double* yu, double* yv, double* du, double* dv, double* d2u, double* d2v) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  const YGrid& g = s->yg;
  if (yu) std::memcpy(yu, g.yu.data(), sizeof(double) * g.nyu);
  if (yv) std::memcpy(yv, g.yv.data(), sizeof(double) * g.nyv);
  if (du) std::memcpy(du, g.dthdyu.data(), sizeof(double) * g.nyu);
  if (dv) std::memcpy(dv, g.dthdyv.data(), sizeof(double) * g.nyv);
  if (d2u) std::memcpy(d2u, g.d2u.data(), sizeof(double) * 3 * g.nyu);
  if (d2v) std::memcpy(d2v, g.d2v.data(), sizeof(double) * 3 * g.nyv);
  return CFD_OK;
}

extern "C" int cfd_get_iblist(cfd_handle h, int grid, int wall, int* l9, double* w3, int* n) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || grid < 0 || grid > 1 || wall < 0 || wall > 1) return fail(s, CFD_EINVAL, "bad IB list selector");
  const std::vector<IBPoint>& L = s->ib[grid][wall];
  if (n) *n = (int)L.size();
  for (size_t q = 0; q < L.size(); ++q) {
    const IBPoint& e = L[q];
    if (l9) {
      const int v[9] = {e.i, e.k, e.j, e.i2, e.k2, e.j2, e.i3, e.k3, e.j3};
      std::memcpy(l9 + 9 * q, v, sizeof v);
    }
    if (w3) {
      w3[3 * q] = e.lap;
      w3[3 * q + 1] = e.w2;
      w3[3 * q + 2] = e.w3;
    }
  }
  return CFD_OK;
}

namespace {
// multi-band resolution: the planes [g0, g1) of field `which` that a short
// column does not have (the rows between its wall bands); none when g0 == g1
void gap_planes(const Solver& s, int which, int& g0, int& g1) {
  g0 = g1 = 0;
  if (!s.mb) return;
  const int g = grid_of(which);
  if (g == 0) {
    g0 = s.N4[1] + 2 - s.jlo_u;
    g1 = s.N4[2] - s.jlo_u;
  } else if (g == 1) {
    g0 = s.N3[1] + 2 - s.jlo_v;
    g1 = s.N3[2] - s.jlo_v;
  } else {
    g0 = s.N4[1] + 1 - s.jlo_p;
    g1 = s.N4[2] + 1 - s.jlo_p;
  }
}
void zero_gaps(const Solver& s, int which, std::vector<double2>& t) {
  int g0, g1;
  gap_planes(s, which, g0, g1);
  const size_t N = s.Ncol;
  for (int q = g0; q < g1; ++q)
    for (size_t c = 0; c < N; ++c)
      if (s.h_short[c]) t[q * N + c] = make_double2(0.0, 0.0);
}
}  // namespace

// reference column layout f(j, column) <-> device f[plane][column]
extern "C" int cfd_set_field(cfd_handle h, int which, const double* src) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !src || which < 0 || which >= CFD_NFIELDS) return fail(s, CFD_EINVAL, "bad field argument");
  const int ny = ny_of(*s, which);
  const size_t N = s->Ncol;
  std::vector<double2> t((size_t)ny * N);
  for (size_t c = 0; c < N; ++c)
    for (int q = 0; q < ny; ++q) t[q * N + c] = make_double2(src[2 * (c * ny + q)], src[2 * (c * ny + q) + 1]);
  zero_gaps(*s, which, t);
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(s->F[which], t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
  return CFD_OK;
}

extern "C" int cfd_get_field(cfd_handle h, int which, double* dst) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !dst || which < 0 || which >= CFD_NFIELDS) return fail(s, CFD_EINVAL, "bad field argument");
  const int ny = ny_of(*s, which);
  const size_t N = s->Ncol;
  std::vector<double2> t((size_t)ny * N);
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(t.data(), s->F[which], t.size() * sizeof(double2), hipMemcpyDeviceToHost));
  zero_gaps(*s, which, t);
  for (size_t c = 0; c < N; ++c)
    for (int q = 0; q < ny; ++q) {
      dst[2 * (c * ny + q)] = t[q * N + c].x;
      dst[2 * (c * ny + q) + 1] = t[q * N + c].y;
    }
  return CFD_OK;
}

extern "C" int cfd_set_scalars(cfd_handle h, const double* v) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !v) return fail(s, CFD_EINVAL, "null argument");
  DevScalars d{};
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  d.t = v[CFD_T];
  d.dt = v[CFD_DT];
  d.dti = v[CFD_DTI];
  d.mpgx = v[CFD_MPGX];
  d.dgx = v[CFD_DGX];
  d.QxT = v[CFD_QXT];
  d.Qx = v[CFD_QX];
  d.dtc = v[CFD_DTC];
  HIPCHK(s, hipMemcpy(s->d_sc, &d, sizeof d, hipMemcpyHostToDevice));
  return CFD_OK;
}

extern "C" int cfd_get_scalars(cfd_handle h, double* v) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !v) return fail(s, CFD_EINVAL, "null argument");
  DevScalars d{};
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  const double o[CFD_NSCALARS] = {d.t, d.dt, d.dti, d.mpgx, d.dgx, d.QxT, d.Qx, d.dtc};
  std::memcpy(v, o, sizeof o);
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
namespace {

int check(cfd_solver* s, int kRK) {
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  if (kRK < 1 || kRK > 3) return fail(s, CFD_EINVAL, "kRK must be 1, 2 or 3");
  if (hipSetDevice(s->dev) != hipSuccess) return fail(s, CFD_EHIP, "hipSetDevice failed");
  return CFD_OK;
}

// entry points that exist for one resolution in every band
int one_band(cfd_solver* s, const char* what) {
  if (s && s->mb)
    return fail(s, CFD_EINVAL, std::string(what) + ": not available with a coarser central band (multi-band runs "
                                                   "through cfd_substep / cfd_step)");
  return CFD_OK;
}

int launched(cfd_solver* s) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, CFD_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// statistics and spectra (stats.f90, spectra.f90)
//
// One sample adds, for the current fields and pressure, the spectra of the
// midband columns (here every column and every plane) and the plane and
// tile-conditional statistics of the physical planes.  The physical planes
// come from the solver's own FP64 z- and x-inverse kernels (the nonlinear
// term's transforms: the same mode placement, the kz = -N2/2 line kept),
// streamed in chunks of kStatChunk planes so the scratch stays small at any
// grid size.  Wall bands: the plane sets of the reference's partition with
// walls (limPL_incw, start.f90:1759-1866); the conditional sums of the two
// walls are kept apart and added when written, as write_stats adds them.
// ---------------------------------------------------------------------------
constexpr int kStatChunk = 64;  // planes per chunk (even: x-inverse jobs pair planes)
enum { SQ_U, SQ_U2, SQ_V, SQ_V2, SQ_W, SQ_W2, SQ_P, SQ_P2, SQ_UV, SQ_UW, SQ_VW, SQ_WX, SQ_WX2, SQ_N };
const int kStatGrid[SQ_N] = {0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1};
const int kSpecGrid[5] = {0, 1, 0, 0, 2};  // spU spV spW spUV spP

int ny_grid(const Solver& s, int g) { return g == 0 ? s.nyu : g == 1 ? s.nyv : s.nyp; }
int jlo_grid(const Solver& s, int g) { return g == 0 ? s.jlo_u : g == 1 ? s.jlo_v : s.jlo_p; }

int stats_setup(cfd_solver* s) {
  auto& S = s->st;
  if (S.ready) return CFD_OK;
  const cfd_params& p = s->P;
  // a central solver (multi-band) sums profiles only: no tiles of its own
  const bool central = s->borrowed_stream;
  if (!central && (p.ntilex < 1 || p.ntilez < 1 || s->ngx / p.ntilex < 1 || s->ngz / p.ntilez < 1))
    return fail(s, CFD_EINVAL, "statistics: ntilex and ntilez must give tiles of at least one point");
  band_table(p, s->yg, S.N);
  auto NB = [&](int r, int b) { return S.N[4 * b + r - 1]; };
  for (int r = 3; r <= 4; ++r)
    if (!(NB(r, 0) <= NB(r, 1) && NB(r, 1) <= NB(r, 2) && NB(r, 2) <= NB(r, 3)))
      return fail(s, CFD_EINVAL,
                  "statistics: the band limits N(r,0:3) are not in order (the wall bands overlap), so the "
                  "reference's plane partition has no wall bands to condition on");
  S.dnx = central ? 1 : s->ngx / p.ntilex;  // start.f90:456-457
  S.dnz = central ? 1 : s->ngz / p.ntilez;
  S.bot[0][0] = NB(4, 0), S.bot[0][1] = NB(4, 1), S.top[0][0] = NB(4, 2) + 1, S.top[0][1] = NB(4, 3) + 1;
  S.bot[1][0] = NB(3, 0), S.bot[1][1] = NB(3, 1), S.top[1][0] = NB(3, 2) + 1, S.top[1][1] = NB(3, 3) + 1;
  S.bot[2][0] = NB(4, 0) + 1, S.bot[2][1] = NB(4, 1) - 1, S.top[2][0] = NB(4, 2) + 2, S.top[2][1] = NB(4, 3);
  // Ny(grid, 0) and Ny(grid, nband) + 1 (start.f90:928-933)
  S.jref_bot[0] = NB(4, 0), S.jref_top[0] = NB(4, 3) + 1;
  S.jref_bot[1] = NB(3, 0), S.jref_top[1] = NB(3, 3) + 1;
  S.jref_bot[2] = NB(4, 0) + 1, S.jref_top[2] = NB(4, 3);
  for (int g = 0; g < 3; ++g) {
    S.nyc[g] = S.bot[g][1] - S.bot[g][0] + 1;  // 0:Ny(grid,1)-Ny(grid,0)
    const int jlo = jlo_grid(*s, g), jhi = jlo + ny_grid(*s, g) - 1;
    if (S.nyc[g] < 1 || S.bot[g][0] < jlo || S.top[g][1] > jhi || S.jref_top[g] - S.top[g][0] >= S.nyc[g] ||
        S.jref_top[g] - S.top[g][1] < 0)
      return fail(s, CFD_EINVAL, "statistics: wall band ranges do not fit the grid");
  }
  int rc;
  const size_t cell = (size_t)S.dnx * S.dnz;
  for (int q = 0; q < SQ_N; ++q) {
    const int g = kStatGrid[q];
    if ((rc = dalloc(s, &S.prof[q], ny_grid(*s, g))) || (rc = dalloc(s, &S.cb[q], cell * S.nyc[g])) ||
        (rc = dalloc(s, &S.ct[q], cell * S.nyc[g])))
      return rc;
  }
  for (int q = 0; q < 5; ++q)
    if ((rc = dalloc(s, &S.sp[q], (size_t)s->Ncol * ny_grid(*s, kSpecGrid[q])))) return rc;
  // chunk buffers
  S.chunk = kStatChunk;
  const size_t zpl = (size_t)s->ngz * s->Mx, ppl = (size_t)s->ngz * s->ngx;
  for (int f = 0; f < 3; ++f)
    if ((rc = dalloc(s, &S.Zc[f], zpl * S.chunk)) || (rc = dalloc(s, &S.Pc[f], ppl * S.chunk))) return rc;
  if ((rc = dalloc(s, &S.wxs, (size_t)s->Ncol * S.chunk))) return rc;
  if ((rc = upload(s, &S.d_ju2, s->nl_u2i))) return rc;
  // z-inverse jobs, chunk by chunk (nf per plane), and x-inverse jobs
  // pairing planes t, t+1 of one field (a last odd plane pairs with itself)
  for (int g = 0; g < 3; ++g) {
    const int ny = ny_grid(*s, g), nf = g == 2 ? 1 : 3;
    std::vector<ZInvJob> zj;
    std::vector<XJob> xj;
    for (int q0 = 0; q0 < ny; q0 += S.chunk) {
      const int len = std::min(S.chunk, ny - q0);
      for (int t = 0; t < len; ++t) {
        const int q = q0 + t;
        for (int f = 0; f < nf; ++f) {
          ZInvJob z{};
          z.pa = q;
          z.pd = t;
          z.mode = 0;
          if (g == 0) {
            if (f == 1) z = s->nl_u2i[q];  // u2 interpolated to the u plane
            else z.src = s->F[f == 0 ? CFD_U1 : CFD_U3];
          } else if (g == 1) {
            if (f == 0) z.src = s->F[CFD_U2];
            else if (f == 1) z = s->nl_u3i[q];  // u3 interpolated to the v plane
            else {
              z.src = S.wxs;  // omega_x of the chunk
              z.pa = t;
            }
          } else {
            z.src = s->F[CFD_P];
          }
          z.pd = t;
          z.flags = 0;
          z.dst = S.Zc[f];
          zj.push_back(z);
        }
      }
      for (int f = 0; f < nf; ++f)
        for (int t = 0; t < len; t += 2) {
          const bool pair = t + 1 < len;
          XJob x{};
          x.za = S.Zc[f];
          x.zb = pair ? S.Zc[f] + zpl : S.Zc[f];
          x.pa = S.Pc[f];
          x.pb = pair ? S.Pc[f] + ppl : S.Pc[f];
          x.plane = t;
          xj.push_back(x);
        }
    }
    if ((rc = upload(s, &S.d_zj[g], zj)) || (rc = upload(s, &S.d_xj[g], xj))) return rc;
  }
  // reduction tasks per grid (stats.f90:16-108)
  double** P = S.Pc;
  const int gx = s->ngx, gz = s->ngz;
  std::vector<StatSumTask> su = {{P[0], nullptr, S.prof[SQ_U], S.prof[SQ_U2], gx, gz},
                                 {P[2], nullptr, S.prof[SQ_W], S.prof[SQ_W2], gx, gz},
                                 {P[0], P[1], S.prof[SQ_UV], nullptr, gx, gz},
                                 {P[0], P[2], S.prof[SQ_UW], nullptr, gx, gz}};
  // the omega_x sums run over i = 1..N1, k = 1..N2 of the physical plane (stats.f90:56-62)
  std::vector<StatSumTask> sv = {{P[0], nullptr, S.prof[SQ_V], S.prof[SQ_V2], gx, gz},
                                 {P[0], P[1], S.prof[SQ_VW], nullptr, gx, gz},
                                 {P[2], nullptr, S.prof[SQ_WX], S.prof[SQ_WX2], s->N1, s->N2}};
  std::vector<StatSumTask> sp = {{P[0], nullptr, S.prof[SQ_P], S.prof[SQ_P2], gx, gz}};
  auto ct = [&](int g, const double* a, const double* b, int q1, int q2, double sa, double sb) {
    StatCondTask c{};
    c.a = a;
    c.b = b;
    c.mb = S.cb[q1];
    c.mt = S.ct[q1];
    c.m2b = q2 >= 0 ? S.cb[q2] : nullptr;
    c.m2t = q2 >= 0 ? S.ct[q2] : nullptr;
    c.sa = sa;
    c.sb = sb;
    c.bot0 = S.bot[g][0];
    c.bot1 = S.bot[g][1];
    c.top0 = S.top[g][0];
    c.top1 = S.top[g][1];
    c.jref_bot = S.jref_bot[g];
    c.jref_top = S.jref_top[g];
    return c;
  };
  // condup signs (stats.f90:98-105)
  std::vector<StatCondTask> cu = {ct(0, P[0], nullptr, SQ_U, SQ_U2, 1.0, 1.0), ct(0, P[2], nullptr, SQ_W, SQ_W2, 1.0, 1.0),
                                  ct(0, P[0], P[1], SQ_UV, -1, 1.0, -1.0), ct(0, P[0], P[2], SQ_UW, -1, 1.0, 1.0)};
  std::vector<StatCondTask> cv = {ct(1, P[0], nullptr, SQ_V, SQ_V2, -1.0, 1.0),
                                  ct(1, P[2], nullptr, SQ_WX, SQ_WX2, 1.0, 1.0),
                                  ct(1, P[0], P[1], SQ_VW, -1, -1.0, 1.0)};
  std::vector<StatCondTask> cp = {ct(2, P[0], nullptr, SQ_P, SQ_P2, 1.0, 1.0)};
  const std::vector<StatSumTask>* sums[3] = {&su, &sv, &sp};
  const std::vector<StatCondTask>* conds[3] = {&cu, &cv, &cp};
  for (int g = 0; g < 3; ++g) {
    if ((rc = upload(s, &S.d_sum[g], *sums[g])) || (rc = upload(s, &S.d_cond[g], *conds[g]))) return rc;
    S.nsum[g] = (int)sums[g]->size();
    S.ncond[g] = (int)conds[g]->size();
  }
  S.ready = true;
  return CFD_OK;
}

// one sample of spectra and statistics, queued on the solver's stream.  With
// a coarser central band the central solver samples the long columns of
// every plane too; its profiles give the central planes' (stats_write).
int stats_sample(cfd_solver* s) {
  if (int rc = stats_setup(s)) return rc;
  if (s->mb) {
    launch_band_gather(*s, 0, std::max(s->nyu, std::max(s->nyv, s->nyp)), true);
    if (int rc = stats_sample(static_cast<cfd_solver*>(s->mbc))) return fail(s, rc, s->mbc->err);
  }
  // a central solver gives profiles only: the spectra and the wall bands'
  // conditional sums are its parent's
  const bool central = s->borrowed_stream;
  auto& S = s->st;
  StatsSpecArgs A{};
  A.u1 = s->F[CFD_U1];
  A.u2 = s->F[CFD_U2];
  A.u3 = s->F[CFD_U3];
  A.p = s->F[CFD_P];
  A.ju2 = S.d_ju2;
  A.spU = S.sp[0];
  A.spV = S.sp[1];
  A.spW = S.sp[2];
  A.spUV = S.sp[3];
  A.spP = S.sp[4];
  A.Ncol = s->Ncol;
  A.nyu = s->nyu;
  A.nyv = s->nyv;
  A.nyp = s->nyp;
  if (!central) launch_spectra(*s, A);
  for (int g = 0; g < 3; ++g) {
    const int ny = ny_grid(*s, g), nf = g == 2 ? 1 : 3;
    for (int q0 = 0; q0 < ny; q0 += S.chunk) {
      const int len = std::min(S.chunk, ny - q0);
      if (g == 1) {
        StatsWxArgs W{s->F[CFD_U2], s->F[CFD_U3], s->d_dthdyv, s->d_k1z, S.wxs, q0, s->Ncol, s->Mx, s->N2};
        launch_omega_x(*s, W, len);
      }
      launch_zinv(*s, S.d_zj[g] + (size_t)q0 * nf, len * nf, false);
      launch_xpass_inv2(*s, S.d_xj[g] + (size_t)nf * (q0 / 2), nf * ((len + 1) / 2));
      launch_stat_sums(*s, S.d_sum[g], S.nsum[g], len, q0);
      if (!central) launch_stat_cond(*s, S.d_cond[g], S.ncond[g], len, S.dnx, S.dnz, jlo_grid(*s, g) + q0);
    }
  }
  S.istat += 1;
  return launched(s);
}

// zero the accumulators (stats.f90:396-426, spectra.f90:124-128)
int stats_zero(cfd_solver* s) {
  auto& S = s->st;
  const size_t cell = (size_t)S.dnx * S.dnz;
  for (int q = 0; q < SQ_N; ++q) {
    const int g = kStatGrid[q];
    HIPCHK(s, hipMemsetAsync(S.prof[q], 0, ny_grid(*s, g) * sizeof(double), s->stream));
    HIPCHK(s, hipMemsetAsync(S.cb[q], 0, cell * S.nyc[g] * sizeof(double), s->stream));
    HIPCHK(s, hipMemsetAsync(S.ct[q], 0, cell * S.nyc[g] * sizeof(double), s->stream));
  }
  for (int q = 0; q < 5; ++q)
    HIPCHK(s, hipMemsetAsync(S.sp[q], 0, (size_t)s->Ncol * ny_grid(*s, kSpecGrid[q]) * sizeof(double), s->stream));
  S.istat = 0;
  return CFD_OK;
}

// write_spect + write_stats (spectra.f90:62-182, stats.f90:255-442), then
// zero the accumulators.  With a coarser central band the central planes'
// profiles are the central solver's (their sums run over the ngx2 x ngz2
// grid) and the spectra cover the long columns, folded into the central
// band's N1c/2+1 x N2c/2+1 modes (buffSP of the midband, spectra.f90:
// 122-170).  The header carries t_hdr and mpgx_hdr; the file names end in
// _<N1>x<N2>x<ny>_<int(10 t_name)>.dat, the stats file directly after dirout
// ('stats_', no separator, stats.f90:357) and the spectra after dirout/
// ('/spU_' ..., spectra.f90:99-108).
int stats_write(cfd_solver* s, const char* dirout, int iter, double t_hdr, double mpgx_hdr, double t_name) {
  if (int rc = stats_setup(s)) return rc;
  cfd_solver* c = static_cast<cfd_solver*>(s->mbc);
  if (s->mb)
    if (int rc = stats_setup(c)) return fail(s, rc, c->err);
  auto& S = s->st;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const size_t cell = (size_t)S.dnx * S.dnz;
  std::vector<double> prof[SQ_N], cond[SQ_N];
  for (int q = 0; q < SQ_N; ++q) {
    const int g = kStatGrid[q];
    prof[q].resize(ny_grid(*s, g));
    HIPCHK(s, hipMemcpy(prof[q].data(), S.prof[q], prof[q].size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> b(cell * S.nyc[g]), t(cell * S.nyc[g]);
    HIPCHK(s, hipMemcpy(b.data(), S.cb[q], b.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(s, hipMemcpy(t.data(), S.ct[q], t.size() * sizeof(double), hipMemcpyDeviceToHost));
    cond[q].resize(b.size());
    for (size_t e = 0; e < b.size(); ++e) cond[q][e] = b[e] + t[e];  // xC = xC + buffC (recvXC)
    if (s->mb) {
      // the central planes: u N(4,1)+1..N(4,2), v N(3,1)+1..N(3,2), p
      // N(4,1)..N(4,2)+1 (the plane ranges of band 2's ranks,
      // start.f90:1650-1951)
      const int j0 = g == 0 ? s->N4[1] + 1 : g == 1 ? s->N3[1] + 1 : s->N4[1];
      const int j1 = g == 0 ? s->N4[2] : g == 1 ? s->N3[2] : s->N4[2] + 1;
      const int q0 = j0 - jlo_grid(*s, g), n = j1 - j0 + 1;
      if (q0 < 0 || q0 + n > (int)prof[q].size()) return fail(s, CFD_EINVAL, "statistics: central planes off the grid");
      HIPCHK(s, hipMemcpy(prof[q].data() + q0, c->st.prof[q] + q0, n * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  cfd_file_header hd{};
  hd.t = t_hdr;
  hd.Re = s->P.Re;
  hd.alp = s->P.alp;
  hd.bet = s->P.bet;
  hd.mpgx = mpgx_hdr;
  hd.nband = 3;
  hd.iter = iter;
  std::memcpy(hd.N, S.N, sizeof hd.N);
  int Ngal[20];
  std::memcpy(Ngal, S.N, sizeof Ngal);
  for (int b = 1; b <= 3; ++b) {  // Ngal(1:2, band) = the physical sizes (start.f90:92-100)
    Ngal[4 * b] = s->mb && b == 2 ? s->ngx2 : s->ngx;
    Ngal[4 * b + 1] = s->mb && b == 2 ? s->ngz2 : s->ngz;
  }
  char tag[64];
  snprintf(tag, sizeof tag, "%04dx%04dx%04d_%05d.dat", s->N1, s->N2, s->ny, (int)(10.0 * t_name));
  // spectra: fold -kz onto kz (k > N2/2 -> N2 + 2 - k with the wall bands'
  // N2, so kz = -N2/2, or -N2c/2 of a coarser central band, lands in the last
  // row), double every kx > 0 (recvwrspec, spectra.f90:129-170); the rows and
  // columns are the midband's (N1c/2+1 x N2c/2+1 with a coarser central
  // band, whose short columns have no spectra)
  const char* spname[5] = {"spU", "spV", "spW", "spUV", "spP"};
  const int mx = s->mb ? s->N1c / 2 + 1 : s->Mx, nzh = (s->mb ? s->N2c : s->N2) / 2 + 1;
  for (int q = 0; q < 5; ++q) {
    const int ny = ny_grid(*s, kSpecGrid[q]);
    std::vector<double> raw((size_t)s->Ncol * ny), buf((size_t)mx * nzh * ny, 0.0);
    HIPCHK(s, hipMemcpy(raw.data(), S.sp[q], raw.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int pass = 0; pass < 2; ++pass)  // assigned from kz >= 0, then the folded -kz added
      for (int k = 0; k < s->N2; ++k) {
        const bool neg = k >= s->N2 / 2;
        if (neg != (pass == 1)) continue;
        const int kk = neg ? s->N2 - k : k;  // 0-based row of buffSP(., 1:N2/2+1, .)
        if (neg ? kk > nzh - 1 : k >= nzh - 1) continue;  // a short column's kz (coarser central band)
        for (int jq = 0; jq < ny; ++jq)
          for (int i = 0; i < mx; ++i) {
            double& d = buf[((size_t)jq * nzh + kk) * mx + i];
            const double v = raw[(size_t)jq * s->Ncol + (size_t)k * s->Mx + i];
            d = pass == 0 ? v : d + v;
          }
      }
    for (size_t e = 0; e < buf.size(); ++e)
      if (e % mx != 0) buf[e] = 2.0 * buf[e];
    const std::string path = std::string(dirout) + "/" + spname[q] + "_" + tag;
    const std::string e = write_spectrum_file(path.c_str(), hd, s->yg, S.istat, buf);
    if (!e.empty()) return fail(s, CFD_EINVAL, "write_spect: " + e);
  }
  const std::string path = std::string(dirout) + "stats_" + tag;
  const std::string e = write_stats_file(path.c_str(), hd, Ngal, s->yg, S.istat, prof, cond);
  if (!e.empty()) return fail(s, CFD_EINVAL, "write_stats: " + e);
  if (s->mb)
    if (int rc = stats_zero(c)) return fail(s, rc, c->err);
  return stats_zero(s);
}

int read_scalars(cfd_solver* s, DevScalars& d) {
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// learned closure: the coarse solver's side (the network is cnn.cpp)
//
// The network sees the volume of the u planes (nyu x Ngz x Ngx): channel 0
// u1 and channel 2 u3 of u plane y, channel 1 u2 of v plane y (zero on the
// last plane), channel 3 the IB mask (the forcing points of the u grid); it
// returns the three components of a correction of the nonlinear term on the
// same planes (components 1 and 3 on the interior u planes, 2 on the interior
// v planes), which the solver transforms to the spectral modes and adds to
// the x/z part of the nonlinear term before the y-derivative assembly.
// ---------------------------------------------------------------------------
int closure_setup(cfd_solver* s) {
  auto& C = s->cl;
  if (C.ready) return CFD_OK;
  if (!cnn_dims_ok(s->nyu, s->ngz, s->ngx))
    return fail(s, CFD_EINVAL, "closure: Ngalx must be a multiple of 32 and Ngalz of 8");
  const size_t zpl = (size_t)s->ngz * s->Mx, ppl = (size_t)s->ngz * s->ngx;
  const int nyf[3] = {s->nyu, s->nyv, s->nyu};
  int rc;
  if ((rc = dalloc(s, &C.Z, zpl * s->nyu)) || (rc = dalloc(s, &C.mask, ppl * s->nyu)) || (rc = dalloc(s, &C.acc, 6)))
    return rc;
  for (int f = 0; f < 3; ++f)
    if ((rc = dalloc(s, &C.Pu[f], ppl * nyf[f])) || (rc = dalloc(s, &C.Pc[f], ppl * nyf[f])) ||
        (rc = dalloc(s, &C.Cs[f], (size_t)s->Ncol * nyf[f])))
      return rc;
  // mask: the forcing points of both walls of the u grid
  {
    std::vector<unsigned char> m(ppl * s->nyu, 0);
    for (int w = 0; w < 2; ++w)
      for (const IBPoint& e : s->ib[0][w]) {
        const int q = e.j - s->jlo_u;
        if (q >= 0 && q < s->nyu && e.i >= 1 && e.i <= s->ngx && e.k >= 1 && e.k <= s->ngz)
          m[(size_t)q * ppl + (size_t)(e.k - 1) * s->ngx + (e.i - 1)] = 1;
      }
    HIPCHK(s, hipMemcpy(C.mask, m.data(), m.size(), hipMemcpyHostToDevice));
  }
  // jobs: z-inverse of a field (mode 0, all planes) into Z; x-inverse pairs
  // of planes of Z into a physical buffer; x-forward of a physical buffer
  // into Z; z-forward of Z into a spectral buffer
  for (int f = 0; f < 6; ++f) {
    const int g = f % 3, ny = nyf[g];
    const double2* src = f < 3 ? s->F[g == 0 ? CFD_U1 : g == 1 ? CFD_U2 : CFD_U3] : C.Cs[g];
    double* dst = f < 3 ? C.Pu[g] : C.Pc[g];
    std::vector<ZInvJob> zj(ny);
    for (int q = 0; q < ny; ++q) {
      ZInvJob z{};
      z.src = src;
      z.dst = C.Z;
      z.pa = q;
      z.pd = q;
      zj[q] = z;
    }
    std::vector<XJob> xj;
    for (int q = 0; q < ny; q += 2) {
      const bool pair = q + 1 < ny;
      XJob x{};
      x.za = C.Z;
      x.zb = pair ? C.Z + zpl : C.Z;
      x.pa = dst;
      x.pb = pair ? dst + ppl : dst;
      x.plane = q;
      xj.push_back(x);
    }
    if ((rc = upload(s, &C.d_zj[f], zj)) || (rc = upload(s, &C.d_xi[f], xj))) return rc;
    C.nzj[f] = (int)zj.size();
    C.nxi[f] = (int)xj.size();
  }
  for (int g = 0; g < 3; ++g) {
    std::vector<XJob> xf;
    std::vector<ZFwdJob> zf;
    for (int q = 0; q < nyf[g]; ++q) {
      XJob x{};
      x.pa = C.Pc[g];
      x.zo = C.Z;
      x.plane = q;
      xf.push_back(x);
      ZFwdJob z{};
      z.src = C.Z;
      z.dst = C.Cs[g];
      z.ps = q;
      z.pd = q;
      zf.push_back(z);
    }
    if ((rc = upload(s, &C.d_xf[g], xf)) || (rc = upload(s, &C.d_zf[g], zf))) return rc;
    C.nxf[g] = (int)xf.size();
  }
  C.ready = true;
  return CFD_OK;
}

// the correction (FP64) into the interior planes of the x/z part of the
// nonlinear term, stored in the precision of the path
void add_correction(cfd_solver* s, void* nxz, const double2* c, int ny) {
  if (s->mixed)
    launch_add_planes(s->stream, (float2*)nxz, c, s->Ncol, 1, ny - 1);
  else
    launch_add_planes(s->stream, (double2*)nxz, c, s->Ncol, 1, ny - 1);
}

// physical planes of job set f (0..2 velocity, 3..5 corrections)
void closure_to_phys(cfd_solver* s, int f) {
  auto& C = s->cl;
  launch_zinv(*s, C.d_zj[f], C.nzj[f], false);
  launch_xpass_inv2(*s, C.d_xi[f], C.nxi[f]);
}

// the network's correction in the x/z part of the nonlinear term (nxz1..3,
// interior planes): recomputed from the current state at kRK = 1 of every
// `every`-th step (and when none is held), otherwise the held one
int closure_apply(cfd_solver* s, int kRK) {
  auto& C = s->cl;
  cfd_cnn_s* c = C.cnn;
  bool fresh = !C.have;
  if (kRK == 1) {
    fresh = fresh || C.count % C.every == 0;
    C.count += 1;
  }
  void* nxz[3] = {s->nxz1, s->nxz2, s->nxz3};
  const int nyf[3] = {s->nyu, s->nyv, s->nyu};
  if (!fresh) {
    for (int g = 0; g < 3; ++g) add_correction(s, nxz[g], C.Cs[g], nyf[g]);
    return CFD_OK;
  }
  for (int f = 0; f < 3; ++f) closure_to_phys(s, f);
  launch_pack(s->stream, C.Pu[0], C.Pu[1], C.Pu[2], C.mask, s->nyu, s->nyv, s->ngz, s->ngx, c->scale, 1, c->x);
  hipStream_t keep = c->stream;
  c->stream = s->stream;  // the network runs on the solver's stream here
  const int rc = cnn_forward_dev(c, c->x);
  c->stream = keep;
  if (rc) return fail(s, rc, std::string("closure: ") + c->err);
  launch_unpack(s->stream, c->y, s->nyu, s->nyv, s->ngz, s->ngx, c->scale + 3, C.Pc[0], C.Pc[1], C.Pc[2],
                C.clip ? C.Pu[0] : nullptr, C.clip ? C.Pu[1] : nullptr, C.clip ? C.Pu[2] : nullptr);
  C.have = true;
  for (int g = 0; g < 3; ++g) {
    launch_xpass_fwd2(*s, C.d_xf[g], C.nxf[g]);
    launch_zfwd_plain(*s, C.d_zf[g], C.nxf[g]);
    // no net streamwise or spanwise force from the network (u and v planes
    // share their first index, so v face q - 1 lies below u plane q)
    if (g != 1) launch_zero_net_force(s->stream, C.Cs[g], s->Ncol, 1, nyf[g] - 1, s->d_yv);
    add_correction(s, nxz[g], C.Cs[g], nyf[g]);
  }
  return CFD_OK;
}

// A statistics request (cfd_stats_request) is served by the next nonlinear
// term with kRK = 1, inside cfd_nonlinear or cfd_substep, where the
// reference's nonlinear serves flagst / flagwr (FOU3D.f90:56-69, 107-110,
// 177-226): the sample is of the fields at the start of the substep, and the
// files are written after the time-step update, with the new t and the mpgx
// of the start of the substep in the header and the name taken from the t of
// the start.
struct StatsReq {
  int flags = 0;
  DevScalars d0{};
};

int stats_req_begin(cfd_solver* s, int kRK, StatsReq& r) {
  r.flags = kRK == 1 ? s->st.pending : 0;
  if (!r.flags) return CFD_OK;
  s->st.pending = 0;
  if (r.flags & 2)
    if (int rc = read_scalars(s, r.d0)) return rc;
  if (r.flags & 1) return stats_sample(s);
  return CFD_OK;
}

// after the time-step update; a failed write leaves the step to complete
int stats_req_end(cfd_solver* s, const StatsReq& r) {
  if (!(r.flags & 2)) return CFD_OK;
  DevScalars d1{};
  if (int rc = read_scalars(s, d1)) return rc;
  return stats_write(s, s->st.pending_dir.c_str(), s->st.pending_iter, d1.t, r.d0.mpgx, r.d0.t);
}

// The nonlinear term's transforms (nonlinear, littleharsh.f90:520-700): every
// plane at the solver's resolution; with multi-band resolution the central
// solver transforms the long columns of the central planes at the central
// band's resolution and its products replace this solver's on those planes
// (short columns carry none there), and dt takes both resolutions' limits.
void nl_transforms(cfd_solver* s, int kRK) {
  if (kRK == 1) launch_reset_max(*s);
  if (s->mb) {
    cfd_solver* c = static_cast<cfd_solver*>(s->mbc);
    if (kRK == 1) launch_reset_max(*c);
    { Timed t(s, KC_ELEM); launch_band_gather(*s); }
    { Timed t(s, KC_ZINV); launch_zinv(*c, c->d_nljobs, c->n_nljobs, true); }
    { Timed t(s, KC_XPASS); launch_xpass_nl(*c, kRK == 1); }
    { Timed t(s, KC_ZFWD); launch_zfwd_nl(*c); }
  }
  { Timed t(s, KC_ZINV); launch_zinv(*s, s->d_nljobs, s->n_nljobs, true); }
  { Timed t(s, KC_XPASS); launch_xpass_nl(*s, kRK == 1); }
  {
    Timed t(s, KC_ZFWD);
    launch_zfwd_nl(*s);
    launch_plane_map(*s);
  }
  if (s->mb) {
    { Timed t(s, KC_ELEM); launch_band_scatter(*s); }
    if (kRK == 1) launch_cfl_bands(*s);
  } else if (kRK == 1) {
    launch_cfl(*s);
  }
}

int op_nonlinear(cfd_solver* s, int kRK) {
  StatsReq req;
  if (int rc = stats_req_begin(s, kRK, req)) return rc;
  nl_transforms(s, kRK);
  const int wrc = stats_req_end(s, req);
  if (s->cl.cnn)
    if (int rc = closure_apply(s, kRK)) return rc;
  { Timed t(s, KC_ELEM); launch_nlcombine(*s, kRK); }
  return wrc;
}

// IB correction of the predictor of fields [f0, f1) (0 u1, 1 u2, 2 u3)
void op_ib(cfd_solver* s, int f0, int f1) {
  const int z0 = s->ibz_off[f0], z1 = s->ibz_off[f1];
  const int x0 = s->ibx_off[f0], x1 = s->ibx_off[f1];
  if (x1 <= x0) return;
  Timed t(s, KC_IB);
  launch_zinv(*s, s->d_ibz + 2 * z0, 2 * (z1 - z0));
  launch_xpass_inv2(*s, s->d_ibxi + x0, x1 - x0);
  launch_ib_apply(*s, f0, f1);
  launch_xpass_fwd2(*s, s->d_ibxf + x0, x1 - x0);
  launch_zfwd_plain(*s, s->d_ibzf + x0, x1 - x0);
}

void op_solve(cfd_solver* s, int f, int kRK) {
  { Timed t(s, KC_ELEM); launch_predictor(*s, f); }
  op_ib(s, f, f + 1);
  { Timed t(s, KC_YSOLVE); launch_ysolve(*s, f, kRK); }
}

// The substep with fused column sweeps: the nonlinear transforms, the
// predictor on the IB planes and its IB correction, then one sweep per
// column that builds RHS0 + nonlinear assembly + predictor and solves the
// Helmholtz system, the (0,0) flow-rate correction, and one sweep per column
// for solveP + v_corr.  Same arithmetic as the operator-by-operator path;
// du holds r only on the IB planes afterwards.
// A pending statistics request is served as in op_nonlinear.
int op_substep(cfd_solver* s, int kRK) {
  StatsReq req;
  if (int rc = stats_req_begin(s, kRK, req)) return rc;
  nl_transforms(s, kRK);
  // a failed write (say, a missing directory) still completes the substep
  const int wrc = stats_req_end(s, req);
  if (s->cl.cnn)
    if (int rc = closure_apply(s, kRK)) return rc;
  // the u11 solve of the flow-rate correction beside the IB and column
  // solves (it needs dt only); serial while the stream is being captured
  bool side = false;
  if (s->P.flag_ctpress == 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(s, hipStreamIsCapturing(s->stream, &cs));
    side = cs == hipStreamCaptureStatusNone;
    if (side) {
      HIPCHK(s, hipEventRecord(s->ev_fork, s->stream));
      HIPCHK(s, hipStreamWaitEvent(s->side, s->ev_fork, 0));
      launch_mf_u11(*s, kRK, s->side);
      HIPCHK(s, hipEventRecord(s->ev_join, s->side));
    }
  }
  {
    Timed t(s, KC_IB);
    launch_ib_rhs(*s, kRK);
  }
  op_ib(s, 0, 3);
  { Timed t(s, KC_YSOLVE); launch_rhs_solve(*s, kRK); }
  if (s->P.flag_ctpress == 0) {
    // the join comes before the class timer starts, so the elementwise class
    // times k_mf_apply (and the serial u11 solve when captured), not the wait
    if (side) HIPCHK(s, hipStreamWaitEvent(s->stream, s->ev_join, 0));
    Timed t(s, KC_ELEM);
    if (!side) launch_mf_u11(*s, kRK, s->stream);
    launch_mf_apply(*s);
  }
  { Timed t(s, KC_PRESS); launch_pressure(*s, kRK); }
  if (kRK == 3 && s->filt) launch_mode_filter(*s);
  ++s->nsub;
  return wrc;
}

}  // namespace

extern "C" int cfd_rhs0(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_rhs0")) return rc;
  launch_rhs0(*s, kRK);
  return launched(s);
}
extern "C" int cfd_nonlinear(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_nonlinear")) return rc;
  if (int rc = op_nonlinear(s, kRK)) return rc;
  return launched(s);
}
extern "C" int cfd_solve_u(cfd_handle h, int comp, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_solve_u")) return rc;
  if (comp != 1 && comp != 3) return fail(s, CFD_EINVAL, "comp must be 1 or 3");
  op_solve(s, comp - 1, kRK);
  return launched(s);
}
extern "C" int cfd_solve_v(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_solve_v")) return rc;
  op_solve(s, 1, kRK);
  return launched(s);
}
extern "C" int cfd_meanflow_ctU(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_meanflow_ctU")) return rc;
  launch_meanflow_lds(*s, kRK);
  return launched(s);
}
extern "C" int cfd_solve_p(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_solve_p")) return rc;
  launch_solve_p(*s, kRK);
  return launched(s);
}
extern "C" int cfd_vcorr(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = one_band(s, "cfd_vcorr")) return rc;
  launch_vcorr(*s, kRK);
  launch_divergence(*s);
  if (kRK == 3 && s->filt) launch_mode_filter(*s);
  return launched(s);
}
extern "C" int cfd_substep(cfd_handle h, int kRK) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, kRK)) return rc;
  if (int rc = op_substep(s, kRK)) return rc;
  return launched(s);
}
extern "C" int cfd_step(cfd_handle h, int nsteps) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (nsteps < 0) return fail(s, CFD_EINVAL, "nsteps must be >= 0");
  for (int it = 0; it < nsteps; ++it) {
    for (int kRK = 1; kRK <= 3; ++kRK)
      if (int rc = op_substep(s, kRK)) return rc;
    if (s->P.flag_ctpress == 1) launch_flowrate(*s);
    if (int rc = launched(s)) return rc;
  }
  return CFD_OK;
}
extern "C" int cfd_meanflow_ctP(cfd_handle h) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  launch_flowrate(*s);
  return launched(s);
}

// The quantities record_out prints and appends to the history file
// (FOU3D.f90:2504-2534): Uslip, the flow rate (flowrateIm, which also sets
// Qx), Umax (maxvel, littleharsh.f90:486-502) from the (0,0) column of u1,
// and err, the largest |div| over the p planes (error, littleharsh.f90:322-349).
extern "C" int cfd_hist(cfd_handle h, double* v) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !v) return fail(s, CFD_EINVAL, "null argument");
  HIPCHK(s, hipSetDevice(s->dev));
  launch_absmax(*s, s->F[CFD_DIV], (size_t)s->nyp * s->Ncol);
  std::vector<double2> col(s->nyu);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy2D(col.data(), sizeof(double2), s->F[CFD_U1], s->Ncol * sizeof(double2), sizeof(double2),
                        s->nyu, hipMemcpyDeviceToHost));
  DevScalars d{};
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  const YGrid& g = s->yg;
  auto U = [&](int j) { return col[j - s->jlo_u].x; };
  const int nn = s->nn;
  const double Uslip = (U(1) * (g.YV(0) - g.YU(0)) + U(0) * (g.YU(1) - g.YV(0))) / (g.YU(1) - g.YU(0));
  double Q = 0.0;
  for (int j = 1; j <= nn + 1; ++j) Q = Q + (g.YV(j) - g.YV(j - 1)) * U(j);
  double Umax = 0.0;
  for (int j = 0; j <= nn + 2; ++j) Umax = std::max(Umax, U(j));
  double err;
  std::memcpy(&err, &d.amax, sizeof err);
  d.Qx = Q;
  HIPCHK(s, hipMemcpy(&s->d_sc->Qx, &d.Qx, sizeof d.Qx, hipMemcpyHostToDevice));
  const double out[11] = {d.t, s->dtv, d.dtc, d.dt, err, Q, d.QxT, d.mpgx, d.dgx, Umax, Uslip};
  std::memcpy(v, out, sizeof out);
  return CFD_OK;
}

// modal filter (no counterpart in the reference): sigma = 1 for eta <=
// eta_c, exp(-alpha ((eta - eta_c) / (1 - eta_c))^order) above, with eta =
// max(|i| / (N1/2), |kz index| / (N2/2)) of the column; alpha = 0 switches
// it off
extern "C" int cfd_set_mode_filter(cfd_handle h, double alpha, int order, double eta_c) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  if (int rc = one_band(s, "cfd_set_mode_filter")) return rc;
  if (!(alpha >= 0.0) || order < 1 || !(eta_c >= 0.0 && eta_c < 1.0))
    return fail(s, CFD_EINVAL, "mode filter: alpha >= 0, order >= 1, 0 <= eta_c < 1");
  if (alpha == 0.0) {
    s->filt = false;
    return CFD_OK;
  }
  std::vector<double> sig(s->Ncol);
  for (int k = 0; k < s->N2; ++k) {
    const int kzi = k < s->N2 / 2 ? k : s->N2 - k;
    for (int i = 0; i < s->Mx; ++i) {
      const double eta = std::max((double)i / (s->N1 / 2), (double)kzi / (s->N2 / 2));
      sig[(size_t)k * s->Mx + i] = eta > eta_c ? std::exp(-alpha * std::pow((eta - eta_c) / (1.0 - eta_c), order)) : 1.0;
    }
  }
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (!s->d_sigma)
    if (int rc = dalloc(s, &s->d_sigma, sig.size())) return rc;
  HIPCHK(s, hipMemcpy(s->d_sigma, sig.data(), sig.size() * sizeof(double), hipMemcpyHostToDevice));
  s->filt = true;
  return CFD_OK;
}

// the reference's as-shipped nonlinear term for np ranks (no counterpart as
// a switch in the reference, which always runs it): in ops_in_planes
// (FOU3D.f90:1553-1613) the derivative scratch is overwritten plane by plane
// over a rank's planes with walls (limPL_incw) and then copied to all of its
// planes without walls (limPL_excw), so every plane takes the x/z terms of
// the rank's last plane; the top rank's last planes are the top ghost u
// plane and the top wall v plane.  np = 0 restores the per-plane term.
extern "C" int cfd_set_compat_lastplane(cfd_handle h, int np) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  if (int rc = one_band(s, "cfd_set_compat_lastplane")) return rc;
  if (np == 0) {
    s->compat_np = 0;
    return CFD_OK;
  }
  if (np < 3) return fail(s, CFD_EINVAL, "last-plane replay: the reference's plane partition needs np >= 3");
  int N[20];
  band_table(s->P, s->yg, N);
  std::vector<int> lim(4 * (size_t)np);
  if (!plane_partition(s->P, N, np, lim.data()))
    return fail(s, CFD_EINVAL, "last-plane replay: no plane partition for this np");
  std::vector<int> um(s->nyu), vm(s->nyv);
  for (int q = 0; q < s->nyu; ++q) um[q] = q;
  for (int q = 0; q < s->nyv; ++q) vm[q] = q;
  for (int g = 0; g < 2; ++g) {
    std::vector<int>& m = g == 0 ? um : vm;
    const int jlo = g == 0 ? s->jlo_u : s->jlo_v, ny = (int)m.size();
    for (int r = 0; r < np; ++r) {
      const int from = lim[(g * 2 + 0) * np + r], to = lim[(g * 2 + 1) * np + r];
      // limPL_incw: the first rank's range starts one plane lower, the last
      // rank's ends one plane higher (start.f90:1866-1869)
      const int last = to + (r == np - 1 ? 1 : 0);
      if (from > to) continue;
      if (from - jlo < 1 || to - jlo > ny - 2 || last - jlo > ny - 1)
        return fail(s, CFD_EINVAL, "last-plane replay: partition outside the grid");
      for (int j = from; j <= to; ++j) m[j - jlo] = last - jlo;
    }
    for (int q = 0; q < ny; ++q)
      if (m[m[q]] != m[q]) return fail(s, CFD_EINVAL, "last-plane replay: a source plane is also replaced");
  }
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (!s->d_umap) {
    if (int rc = dalloc(s, &s->d_umap, um.size())) return rc;
    if (int rc = dalloc(s, &s->d_vmap, vm.size())) return rc;
  }
  HIPCHK(s, hipMemcpy(s->d_umap, um.data(), um.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(s, hipMemcpy(s->d_vmap, vm.data(), vm.size() * sizeof(int), hipMemcpyHostToDevice));
  s->compat_np = np;
  return CFD_OK;
}

extern "C" int cfd_sync(cfd_handle h) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return CFD_OK;
}

extern "C" int cfd_profile(cfd_handle h, int enable) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  s->prof = enable != 0;
  s->evn = 0;
  for (int c = 0; c < KC_N; ++c) {
    s->prof_ms[c] = 0;
    s->prof_cnt[c] = 0;
  }
  return CFD_OK;
}

extern "C" int cfd_profile_read(cfd_handle h, double* ms, int* cnt) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (size_t q = 0; q + 1 < s->evn; q += 2) {
    float m = 0.f;
    HIPCHK(s, hipEventElapsedTime(&m, s->evpool[q], s->evpool[q + 1]));
    s->prof_ms[s->evclass[q / 2]] += m;
    s->prof_cnt[s->evclass[q / 2]] += 1;
  }
  s->evn = 0;
  for (int c = 0; c < KC_N; ++c) {
    if (ms) ms[c] = s->prof_ms[c];
    if (cnt) cnt[c] = s->prof_cnt[c];
  }
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// field files (record_out / read_in)
// ---------------------------------------------------------------------------
namespace {

struct FileField {
  int which, id, grid;  // field, record id, 0 = u/w centres, 1 = v faces, 2 = p centres
  const char* name;
};
const FileField kFileFields[4] = {{CFD_U1, 1, 0, "u1"}, {CFD_U2, 2, 1, "u2"}, {CFD_U3, 3, 0, "u3"}, {CFD_P, 4, 2, "p"}};

// the planes of grid g that the central band's ranks write (limPL_incw of
// the midband ranks, start.f90:1758-1866): u N(4,1)+1 .. N(4,2), v N(3,1)+1
// .. N(3,2), p N(4,1) .. N(4,2)+1, at N(1,2)+2 x N(2,2)
BandPlanes central_planes(const Solver& s, int g) {
  BandPlanes b;
  if (!s.mb) return b;
  if (g == 0) {
    b.q0 = s.N4[1] + 1 - s.jlo_u;
    b.q1 = s.N4[2] + 1 - s.jlo_u;
  } else if (g == 1) {
    b.q0 = s.N3[1] + 1 - s.jlo_v;
    b.q1 = s.N3[2] + 1 - s.jlo_v;
  } else {
    b.q0 = s.N4[1] - s.jlo_p;
    b.q1 = s.N4[2] + 2 - s.jlo_p;
  }
  b.nx = s.N1c + 2;
  b.nz = s.N2c;
  return b;
}

}  // namespace

extern "C" int cfd_record_out(cfd_handle h, const char* dirout, int iter) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !dirout) return fail(s, CFD_EINVAL, "null argument");
  DevScalars d{};
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  cfd_file_header hd{};
  hd.t = d.t;
  hd.Re = s->P.Re;
  hd.alp = s->P.alp;
  hd.bet = s->P.bet;
  hd.mpgx = d.mpgx;
  hd.nband = 3;
  hd.iter = iter;
  band_table(s->P, s->yg, hd.N);
  char tail[96];
  snprintf(tail, sizeof tail, "_%04d", s->N1);
  std::string name_tail = tail;
  snprintf(tail, sizeof tail, "x%04dx%04d_t%05d.dat", s->N2, s->ny, (int)(10.0 * d.t));
  name_tail += tail;
  const int nx = 2 * s->Mx, nz = s->N2;
  for (const FileField& ff : kFileFields) {
    const int ny_g = ny_of(*s, ff.which);
    std::vector<double> planes((size_t)ny_g * nx * nz);
    HIPCHK(s, hipMemcpy(planes.data(), s->F[ff.which], planes.size() * sizeof(double), hipMemcpyDeviceToHost));
    // u1, u3, p carry the u/w centres and their metric, u2 the v faces
    // (FOU3D.f90:2242-2340); dthetavi is never assigned in the reference
    // (declaration.f90:72), so its zero-initialised module value is written
    const bool v = ff.grid == 1;
    const std::vector<double>& y = v ? s->yg.yv : s->yg.yu;
    const std::vector<double>& dy = v ? s->yg.dthdyv : s->yg.dthdyu;
    const int jlo = ff.grid == 0 ? s->jlo_u : (ff.grid == 1 ? s->jlo_v : s->jlo_p);
    const std::string path = std::string(dirout) + "/" + ff.name + name_tail;
    const std::string e = write_field_file(path.c_str(), hd, ff.id, 0.0, y.data(), dy.data(), (int)y.size(),
                                           v ? s->jlo_v : s->jlo_u, ny_g, jlo, nx, nz, planes.data(),
                                           central_planes(*s, ff.grid));
    if (!e.empty()) return fail(s, CFD_EINVAL, e);
  }
  return CFD_OK;
}

extern "C" int cfd_read_in(cfd_handle h, const char* dirin, const char* fnameimb, int flag_init, int* iter0) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !dirin || !fnameimb) return fail(s, CFD_EINVAL, "null argument");
  if (flag_init != 1 && flag_init != 2) return fail(s, CFD_EINVAL, "flag_init must be 1 or 2 to read fields");
  const std::string name = fnameimb;
  const std::string tail = name.size() > 2 ? name.substr(2) : std::string();  // start.f90:2274
  int N[20];
  band_table(s->P, s->yg, N);
  const int nx = 2 * s->Mx, nz = s->N2;
  cfd_file_header h1{};
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (const FileField& ff : kFileFields) {
    const int ny_g = ny_of(*s, ff.which);
    std::vector<double> planes((size_t)ny_g * nx * nz);
    cfd_file_header hd{};
    int mism = 0;
    const std::string path = std::string(dirin) + "/" + ff.name + tail;
    const std::string e =
        read_field_file(path.c_str(), hd, ny_g, nx, nz, s->ngz, planes.data(), mism, central_planes(*s, ff.grid));
    if (!e.empty()) return fail(s, CFD_EINVAL, e);
    if (ff.which == CFD_U1) {
      h1 = hd;
      if (hd.nband != 3) return fail(s, CFD_EINVAL, "wrong startfile? nband in " + path);
      if (hd.alp != s->P.alp || hd.bet != s->P.bet)
        fprintf(stderr, "wrong startfile? alp,alp_old= %g %g bet,bet_old= %g %g\n", s->P.alp, hd.alp, s->P.bet,
                hd.bet);
      bool diff = false;  // start.f90:2305-2321: N(1:3, bands) and N(3,0), N(3,nband)
      for (int b = 1; b <= 3; ++b)
        for (int r = 0; r < 3; ++r) diff = diff || N[b * 4 + r] != hd.N[b * 4 + r];
      diff = diff || N[2] != hd.N[2] || N[3 * 4 + 2] != hd.N[3 * 4 + 2];
      if (diff) fprintf(stderr, "WARNING startfile size (%s)\n", path.c_str());
    }
    if (mism) fprintf(stderr, "WARNING!: unexpected size of %d planes in %s\n", mism, path.c_str());
    HIPCHK(s, hipMemcpy(s->F[ff.which], planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  // flag_init branches (start.f90:1974-2009): time, iteration, target flow rate
  DevScalars d{};
  HIPCHK(s, hipMemcpy(&d, s->d_sc, sizeof d, hipMemcpyDeviceToHost));
  d.t = flag_init == 2 ? h1.t : 0.0;
  if (iter0) *iter0 = flag_init == 2 ? h1.iter : 0;
  if (s->P.flag_ctpress != 1) {
    // flowrateIm of the (0,0) column, tip to tip (littleharsh.f90:411-428)
    std::vector<double2> col(s->nyu);
    HIPCHK(s, hipMemcpy2D(col.data(), sizeof(double2), s->F[CFD_U1], s->Ncol * sizeof(double2), sizeof(double2),
                          s->nyu, hipMemcpyDeviceToHost));
    double Q = 0.0;
    for (int j = 1; j <= s->nn + 1; ++j) Q = Q + (s->yg.YV(j) - s->yg.YV(j - 1)) * col[j - s->jlo_u].x;
    d.QxT = Q;
  }
  HIPCHK(s, hipMemcpy(s->d_sc, &d, sizeof d, hipMemcpyHostToDevice));
  return CFD_OK;
}

extern "C" int cfd_device_bytes(cfd_handle h, double* bytes) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !bytes) return fail(s, CFD_EINVAL, "null argument");
  *bytes = s->alloc_bytes;
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// physical fields (four_to_phys.f90:61-120): every plane of a field zero-padded
// from N to Ngal modes without the kz = -N2/2 mode (resize, four_to_phys.f90:
// 328-348), inverse z then x transforms, unscaled
// ---------------------------------------------------------------------------
extern "C" int cfd_four_to_phys(cfd_handle h, int which, double* phys) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !phys || which < 0 || which >= CFD_NFIELDS) return fail(s, CFD_EINVAL, "bad field argument");
  if (int rc = one_band(s, "cfd_four_to_phys")) return rc;
  const int ny = ny_of(*s, which);
  const size_t zpl = (size_t)s->ngz * s->Mx, ppl = (size_t)s->ngz * s->ngx;
  HIPCHK(s, hipSetDevice(s->dev));
  double2* z = nullptr;
  double* p = nullptr;
  ZInvJob* dz = nullptr;
  XJob* dx = nullptr;
  auto cleanup = [&]() {
    hipFree(z);
    hipFree(p);
    hipFree(dz);
    hipFree(dx);
  };
  std::vector<ZInvJob> zj(ny);
  for (int q = 0; q < ny; ++q) {
    ZInvJob j{};
    j.src = s->F[which];
    j.dst = nullptr;  // set below
    j.pa = q;
    j.pd = q;
    j.mode = 0;
    j.flags = 1;
    zj[q] = j;
  }
  // x transforms pair planes (q, q+1); a last odd plane pairs with itself
  const int nx = (ny + 1) / 2;
  std::vector<XJob> xj(nx);
  if (hipMalloc(&z, zpl * ny * sizeof(double2)) != hipSuccess || hipMalloc(&p, ppl * ny * sizeof(double)) != hipSuccess ||
      hipMalloc(&dz, zj.size() * sizeof(ZInvJob)) != hipSuccess || hipMalloc(&dx, xj.size() * sizeof(XJob)) != hipSuccess) {
    cleanup();
    return fail(s, CFD_ENOMEM, "four_to_phys: device buffers");
  }
  for (auto& j : zj) j.dst = z;
  for (int m = 0; m < nx; ++m) {
    const int q = 2 * m;
    const bool pair = q + 1 < ny;
    XJob j{};
    j.za = z;
    j.zb = pair ? z + zpl : z;
    j.pa = p;
    j.pb = pair ? p + ppl : p;
    j.plane = q;
    xj[m] = j;
  }
  if (hipMemcpy(dz, zj.data(), zj.size() * sizeof(ZInvJob), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dx, xj.data(), xj.size() * sizeof(XJob), hipMemcpyHostToDevice) != hipSuccess) {
    cleanup();
    return fail(s, CFD_EHIP, "four_to_phys: job upload");
  }
  launch_zinv(*s, dz, ny, false);
  launch_xpass_inv2(*s, dx, nx);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(phys, p, ppl * ny * sizeof(double), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(s, CFD_EHIP, std::string("four_to_phys: ") + hipGetErrorString(e));
  return CFD_OK;
}

// the same transforms without allocating: planes [plane0, plane0 + nplanes)
// into the caller's device buffer, asynchronously on the handle's stream, the
// half-transformed planes and job tables in the per-GPU scratch of noalloc.hip
extern "C" int cfd_four_to_phys_planes(cfd_handle h, int which, int plane0, int nplanes, double* phys) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s || !phys || which < 0 || which >= CFD_NFIELDS) return fail(s, CFD_EINVAL, "bad field argument");
  if (int rc = one_band(s, "cfd_four_to_phys_planes")) return rc;
  const int ny = ny_of(*s, which);
  if (plane0 < 0 || nplanes < 1 || plane0 + nplanes > ny)
    return fail(s, CFD_EINVAL, "planes [" + std::to_string(plane0) + ", " + std::to_string(plane0 + nplanes) +
                                   ") outside the field's " + std::to_string(ny) + " planes");
  HIPCHK(s, hipSetDevice(s->dev));
  std::string err;
  if (int rc = four_to_phys_noalloc(*s, which, plane0, nplanes, phys, err)) return fail(s, rc, err);
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// statistics and spectra
// ---------------------------------------------------------------------------
extern "C" int cfd_stats(cfd_handle h) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  return stats_sample(s);
}

extern "C" int cfd_write_stats(cfd_handle h, const char* dirout, int iter) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (!dirout) return fail(s, CFD_EINVAL, "null dirout");
  DevScalars d{};
  if (int rc = read_scalars(s, d)) return rc;
  return stats_write(s, dirout, iter, d.t, d.mpgx, d.t);
}

extern "C" int cfd_stats_request(cfd_handle h, int flags, const char* dirout, int iter) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (flags < 0 || flags > 3) return fail(s, CFD_EINVAL, "flags must be 0..3 (bit 0 sample, bit 1 write)");
  if ((flags & 2) && !dirout) return fail(s, CFD_EINVAL, "null dirout");
  if (flags)
    if (int rc = stats_setup(s)) return rc;
  s->st.pending = flags;
  s->st.pending_dir = (flags & 2) ? dirout : "";
  s->st.pending_iter = iter;
  return CFD_OK;
}

extern "C" int cfd_stats_dims(cfd_handle h, int* d) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (!d) return fail(s, CFD_EINVAL, "null argument");
  if (int rc = stats_setup(s)) return rc;
  const auto& S = s->st;
  d[0] = S.dnx;
  d[1] = S.dnz;
  d[2] = S.nyc[0];
  d[3] = S.nyc[1];
  d[4] = S.nyc[2];
  d[5] = S.istat;
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// learned closure
// ---------------------------------------------------------------------------
extern "C" int cfd_closure_attach(cfd_handle h, cfd_cnn c, int every) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (int rc = one_band(s, "cfd_closure_attach")) return rc;
  if (!c) {
    s->cl.cnn = nullptr;
    s->cl.have = false;
    return CFD_OK;
  }
  if (every < 1) return fail(s, CFD_EINVAL, "closure: every must be >= 1");
  if (c->NY != s->nyu || c->NZ != s->ngz || c->NX != s->ngx)
    return fail(s, CFD_EINVAL, "closure: the network volume must be nyu x Ngalz x Ngalx of this solver");
  for (int k = 0; k < 6; ++k)
    if (!(c->scale[k] > 0.0f)) return fail(s, CFD_EINVAL, "closure: the network has no normalisation scales");
  if (int rc = closure_setup(s)) return rc;
  s->cl.cnn = c;
  s->cl.every = every;
  s->cl.count = 0;
  s->cl.have = false;
  return CFD_OK;
}

extern "C" int cfd_closure_clip(cfd_handle h, int clip) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (!s) return fail(s, CFD_EINVAL, "null handle");
  s->cl.clip = clip != 0;
  return CFD_OK;
}

// one training sample: the box-filtered state of the fine solver on the
// coarse grid (input) and the correction filt(N(u)) - N_coarse(filt(u))
// (target), into dataset slot `slot`.  The coarse solver's fields are
// overwritten; rms[6] returns the root mean square of u1, u2, u3 and of the
// three corrections; the first sample of a network without scales sets them.
extern "C" int cfd_closure_snapshot(cfd_handle hf, cfd_handle hc, cfd_cnn c, int slot, double* rms) {
  cfd_solver* f = static_cast<cfd_solver*>(hf);
  cfd_solver* s = static_cast<cfd_solver*>(hc);
  if (int rc = check(f, 2)) return rc;
  if (int rc = check(s, 2)) return rc;
  if (int rc = one_band(f, "cfd_closure_snapshot")) return rc;
  if (int rc = one_band(s, "cfd_closure_snapshot")) return rc;
  if (!c || slot < 0 || slot >= c->nslots) return fail(s, CFD_EINVAL, "closure: bad network or slot");
  if (f->ny != s->ny || f->dsty != s->dsty || f->P.ygrid != s->P.ygrid || s->N1 > f->N1 || s->N2 > f->N2)
    return fail(s, CFD_EINVAL, "closure: the coarse solver must share the y grid and have fewer x/z modes");
  if (c->NY != s->nyu || c->NZ != s->ngz || c->NX != s->ngx)
    return fail(s, CFD_EINVAL, "closure: the network volume must be nyu x Ngalz x Ngalx of the coarse solver");
  if (f->mixed) return fail(s, CFD_EINVAL, "closure: samples come from the FP64 path");
  if (int rc = closure_setup(s)) return rc;
  auto& C = s->cl;
  cfd_cnn_s* keep = C.cnn;
  C.cnn = nullptr;  // the coarse nonlinear term below is the plain one
  // fine nonlinear term (kRK = 2: no time-step update), then on the coarse side
  if (int rc = op_nonlinear(f, 2)) return rc;
  HIPCHK(s, hipStreamSynchronize(f->stream));
  const int nyf[3] = {s->nyu, s->nyv, s->nyu};
  const int uf[3] = {CFD_U1, CFD_U2, CFD_U3}, nf[3] = {CFD_NU1, CFD_NU2, CFD_NU3};
  for (int g = 0; g < 3; ++g)
    launch_filter(s->stream, f->F[uf[g]], nullptr, s->F[uf[g]], f->Mx, f->N2, s->Mx, s->N2, s->ngx, s->ngz, 1.0, 0.0,
                  nyf[g]);
  launch_filter(s->stream, f->F[CFD_P], nullptr, s->F[CFD_P], f->Mx, f->N2, s->Mx, s->N2, s->ngx, s->ngz, 1.0, 0.0,
                s->nyp);
  const int rcn = op_nonlinear(s, 2);
  C.cnn = keep;
  if (rcn) return rcn;
  for (int g = 0; g < 3; ++g) {
    launch_filter(s->stream, f->F[nf[g]], s->F[nf[g]], C.Cs[g], f->Mx, f->N2, s->Mx, s->N2, s->ngx, s->ngz, 1.0, -1.0,
                  nyf[g]);
    // the nonlinear term is defined on the interior planes only
    HIPCHK(s, hipMemsetAsync(C.Cs[g], 0, (size_t)s->Ncol * sizeof(double2), s->stream));
    HIPCHK(s, hipMemsetAsync(C.Cs[g] + (size_t)(nyf[g] - 1) * s->Ncol, 0, (size_t)s->Ncol * sizeof(double2),
                             s->stream));
  }
  for (int k = 0; k < 6; ++k) closure_to_phys(s, k);
  // root mean squares
  const size_t ppl = (size_t)s->ngz * s->ngx;
  HIPCHK(s, hipMemsetAsync(C.acc, 0, 6 * sizeof(double), s->stream));
  for (int k = 0; k < 6; ++k) launch_sumsq(s->stream, k < 3 ? C.Pu[k] : C.Pc[k - 3], ppl * nyf[k % 3], C.acc + k);
  double h[6];
  HIPCHK(s, hipMemcpyAsync(h, C.acc, sizeof h, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (int k = 0; k < 6; ++k) h[k] = std::sqrt(h[k] / (double)(ppl * nyf[k % 3]));
  if (rms)
    for (int k = 0; k < 6; ++k) rms[k] = h[k];
  if (!(c->scale[0] > 0.0f))
    for (int k = 0; k < 6; ++k) c->scale[k] = h[k] > 0.0 ? (float)h[k] : 1.0f;
  __bf16* xs = c->dx + c->npts * 4 * (size_t)slot;
  __bf16* ts = c->dt + c->npts * 4 * (size_t)slot;
  launch_pack(s->stream, C.Pu[0], C.Pu[1], C.Pu[2], C.mask, s->nyu, s->nyv, s->ngz, s->ngx, c->scale, 1, xs);
  launch_pack(s->stream, C.Pc[0], C.Pc[1], C.Pc[2], nullptr, s->nyu, s->nyv, s->ngz, s->ngx, c->scale + 3, 0, ts);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return launched(s);
}

// physical values of the last sample's or the last application's buffers:
// which 0..2 the velocities u1, u2, u3, 3..5 the corrections (planes of the
// component's grid x Ngalz x Ngalx)
extern "C" int cfd_closure_phys(cfd_handle h, int which, double* out) {
  cfd_solver* s = static_cast<cfd_solver*>(h);
  if (int rc = check(s, 1)) return rc;
  if (int rc = one_band(s, "cfd_closure_phys")) return rc;
  if (!out || which < 0 || which > 5 || !s->cl.ready) return fail(s, CFD_EINVAL, "closure: nothing to read");
  const int g = which % 3, ny = g == 1 ? s->nyv : s->nyu;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(out, which < 3 ? s->cl.Pu[g] : s->cl.Pc[g], (size_t)ny * s->ngz * s->ngx * sizeof(double),
                      hipMemcpyDeviceToHost));
  return CFD_OK;
}

// ---------------------------------------------------------------------------
// ensembles: independent solvers stepped by one HIP graph per RK substep
// ---------------------------------------------------------------------------
struct cfd_ensemble_s {
  int dev = 0;
  std::vector<cfd_solver*> m;
  hipStream_t root = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t g[3] = {};
  hipGraphExec_t ge[3] = {};
  std::string err;
  // batched kernels: per kRK and kernel class the launches over the member
  // table, whose arguments live in dargs
  bool batched = false;
  void* dargs = nullptr;
  std::vector<BatchLaunch> bl[3][BC_N];
};

namespace {
thread_local std::string g_ens_err;

void ens_free(cfd_ensemble_s* e) {
  if (e->dargs) hipFree(e->dargs);
  for (int k = 0; k < 3; ++k) {
    if (e->ge[k]) hipGraphExecDestroy(e->ge[k]);
    if (e->g[k]) hipGraphDestroy(e->g[k]);
  }
  for (hipEvent_t v : e->join)
    if (v) hipEventDestroy(v);
  if (e->fork) hipEventDestroy(e->fork);
  if (e->root) hipStreamDestroy(e->root);
  delete e;
}

int ens_fail(cfd_ensemble_s* e, int rc, const std::string& msg) {
  g_ens_err = msg;
  if (e) e->err = msg;
  return rc;
}
}  // namespace

namespace {
// The member tables of the batched kernels: for every kRK and class, each
// member's part (its arguments and block count), grouped by kernel variant
// and block width into launches of at most kMaxBatch members; the arguments
// of all launches go to one device buffer.
int ens_build_batches(cfd_ensemble_s* e, std::string& msg) {
  struct Pending {
    BatchLaunch b;
    size_t off;  // argument offset in the blob
  };
  std::vector<char> blob;
  std::vector<Pending> pend[3][BC_N];
  for (int k = 0; k < 3; ++k) {
    for (int cls = 0; cls < BC_N; ++cls) {
      std::vector<BatchPart> parts(e->m.size());
      for (size_t i = 0; i < e->m.size(); ++i) {
        std::string err;
        const bool fft = cls == BC_ZINV_NL || cls == BC_XPASS || cls == BC_ZFWD;
        const bool ok = fft ? batch_part_fft(*e->m[i], cls, k + 1, parts[i], err)
                            : batch_part_fused(*e->m[i], cls, k + 1, parts[i], err);
        if (!ok) {
          msg = "ensemble: member " + std::to_string(i) + ": " + err;
          return CFD_EINVAL;
        }
      }
      std::vector<bool> done(parts.size(), false);
      for (size_t i = 0; i < parts.size(); ++i) {
        if (done[i]) continue;
        // members of the same variant and block width, in member order
        std::vector<size_t> grp;
        for (size_t j = i; j < parts.size(); ++j)
          if (!done[j] && parts[j].key == parts[i].key && parts[j].block.x == parts[i].block.x &&
              grp.size() < (size_t)kMaxBatch) {
            grp.push_back(j);
            done[j] = true;
          }
        Pending pd;
        BatchLaunch& b = pd.b;
        b.cls = cls;
        b.key = parts[i].key;
        b.lead = (int)i;
        b.tw = parts[i].tw;
        unsigned by = 1;
        size_t fixed = 0;
        for (size_t j : grp) {
          by = std::max(by, parts[j].block.y);
          fixed = std::max(fixed, parts[j].lds_fixed);
        }
        b.block = dim3(parts[i].block.x, by, 1);
        b.lds = fixed + parts[i].lds_thread * (size_t)b.block.x * by;
        const size_t asz = parts[i].args.size();
        blob.resize((blob.size() + 255) / 256 * 256);
        pd.off = blob.size();
        int nb = 0;
        b.map.n = 0;
        for (size_t j : grp) {
          const int cnt = parts[j].gx * parts[j].gy;
          if (cnt <= 0) continue;
          const int t = b.map.n++;
          b.map.off[t] = nb;
          b.map.cnt[t] = cnt;
          b.map.gx[t] = parts[j].gx;
          b.members.push_back((int)j);
          nb += (cnt + 7) / 8 * 8;
          blob.insert(blob.end(), parts[j].args.begin(), parts[j].args.end());
          (void)asz;
        }
        if (b.map.n == 0) continue;
        b.grid = dim3(nb);
        pend[k][cls].push_back(pd);
      }
    }
  }
  if (blob.empty()) return CFD_OK;
  if (hipMalloc(&e->dargs, blob.size()) != hipSuccess) {
    msg = "ensemble: device memory for the member tables";
    return CFD_ENOMEM;
  }
  if (hipMemcpy(e->dargs, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
    msg = "ensemble: copy of the member tables";
    return CFD_EHIP;
  }
  for (int k = 0; k < 3; ++k)
    for (int cls = 0; cls < BC_N; ++cls)
      for (Pending& pd : pend[k][cls]) {
        pd.b.dargs = static_cast<const char*>(e->dargs) + pd.off;
        e->bl[k][cls].push_back(pd.b);
      }
  return CFD_OK;
}

// The batched substep kRK as an explicit graph.  Each batch launch and each
// stretch of one member's own kernels (on the member's stream; the u11 solve
// of the flow-rate correction on its side stream) is captured alone and
// copied in as kernel nodes whose dependencies are the nodes that last
// worked on its members, so a batch waits only for its own members and
// members of different batches overlap as in the per-member graphs.  Order
// per member as in op_substep (and, after kRK = 3 with flag_ctpress = 1, the
// flow-rate integral, as cfd_step does).
hipError_t ens_batched_graph(cfd_ensemble_s* e, int kRK, hipGraph_t* out) {
  const int k = kRK - 1, n = (int)e->m.size();
  hipGraph_t G = nullptr;
  hipError_t err = hipGraphCreate(&G, 0);
  std::vector<hipGraphNode_t> last(n, nullptr), u11(n, nullptr);
  auto deps_of = [&](const std::vector<int>& ms) {
    std::vector<hipGraphNode_t> d;
    for (int i : ms)
      if (last[i] && std::find(d.begin(), d.end(), last[i]) == d.end()) d.push_back(last[i]);
    return d;
  };
  // f's launches on stream cs, captured alone, then copied into G as kernel
  // nodes (a chain: one stream) after deps; returns the chain's last node
  // (deps unchanged when f launches nothing).  Kernel nodes, not a child
  // graph: the runtime runs a child graph as one unit on one queue.
  auto add = [&](hipStream_t cs, std::vector<hipGraphNode_t> deps, auto f) -> hipGraphNode_t {
    if (err != hipSuccess) return nullptr;
    err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) return nullptr;
    f();
    const hipError_t le = hipGetLastError();
    hipGraph_t sub = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &sub);
    err = le != hipSuccess ? le : ec;
    if (err == hipSuccess && !sub) err = hipErrorInvalidValue;
    hipGraphNode_t tail = deps.size() == 1 ? deps[0] : nullptr;
    size_t nn = 0, ne = 0;
    if (err == hipSuccess) err = hipGraphGetNodes(sub, nullptr, &nn);
    if (err == hipSuccess) err = hipGraphGetEdges(sub, nullptr, nullptr, &ne);
    std::vector<hipGraphNode_t> nodes(nn), from(ne), to(ne), copy(nn, nullptr);
    if (err == hipSuccess && nn) err = hipGraphGetNodes(sub, nodes.data(), &nn);
    if (err == hipSuccess && ne) err = hipGraphGetEdges(sub, from.data(), to.data(), &ne);
    // the capture of one stream is a chain: walk it from its root
    std::vector<int> npred(nn, 0), next(nn, -1);
    auto idx = [&](hipGraphNode_t v) { return int(std::find(nodes.begin(), nodes.end(), v) - nodes.begin()); };
    for (size_t q = 0; q < ne && err == hipSuccess; ++q) {
      const int a = idx(from[q]), b = idx(to[q]);
      if (next[a] >= 0 || ++npred[b] > 1) err = hipErrorInvalidValue;  // not a chain
      next[a] = b;
    }
    int v = -1;
    for (size_t q = 0; q < nn; ++q)
      if (!npred[q]) v = v < 0 ? int(q) : -2;
    if (err == hipSuccess && nn && v < 0) err = hipErrorInvalidValue;
    std::vector<hipGraphNode_t> pre = deps;
    for (; err == hipSuccess && nn && v >= 0; v = next[v]) {
      hipGraphNodeType t;
      err = hipGraphNodeGetType(nodes[v], &t);
      if (err == hipSuccess && t != hipGraphNodeTypeKernel) err = hipErrorInvalidValue;
      hipKernelNodeParams kp{};
      if (err == hipSuccess) err = hipGraphKernelNodeGetParams(nodes[v], &kp);
      hipGraphNode_t nd = nullptr;
      if (err == hipSuccess) err = hipGraphAddKernelNode(&nd, G, pre.data(), pre.size(), &kp);
      pre.assign(1, nd);
      tail = nd;
    }
    if (err == hipSuccess && !nn && deps.size() > 1) err = hipGraphAddEmptyNode(&tail, G, deps.data(), deps.size());
    if (sub) hipGraphDestroy(sub);
    return tail;
  };
  auto batches = [&](int cls) {
    for (const BatchLaunch& b : e->bl[k][cls]) {
      hipGraphNode_t nd = add(e->root, deps_of(b.members), [&]() {
        if (cls == BC_RHS || cls == BC_PRESS) batch_launch_fused(e->root, b);
        else batch_launch_fft(e->root, b);
      });
      for (int i : b.members) last[i] = nd;
    }
  };
  auto each = [&](auto has, auto f) {
    for (int i = 0; i < n; ++i) {
      cfd_solver* s = e->m[i];
      if (!has(s)) continue;
      last[i] = add(s->stream, deps_of({i}), [&]() { f(s); });
    }
  };
  auto all = [](cfd_solver*) { return true; };
  if (kRK == 1) each(all, [&](cfd_solver* s) { launch_reset_max(*s); });
  batches(BC_ZINV_NL);
  batches(BC_XPASS);
  batches(BC_ZFWD);
  if (kRK == 1) each(all, [&](cfd_solver* s) { launch_cfl(*s); });
  // the u11 solve needs dt only: beside the IB correction and the sweeps
  for (int i = 0; i < n; ++i) {
    cfd_solver* s = e->m[i];
    if (s->P.flag_ctpress == 0)
      u11[i] = add(s->side, deps_of({i}), [&]() { launch_mf_u11(*s, kRK, s->side); });
  }
  each(all, [&](cfd_solver* s) {
    launch_ib_rhs(*s, kRK);
    op_ib(s, 0, 3);
  });
  batches(BC_RHS);
  for (int i = 0; i < n; ++i) {
    cfd_solver* s = e->m[i];
    if (s->P.flag_ctpress != 0) continue;
    std::vector<hipGraphNode_t> d = deps_of({i});
    if (u11[i]) d.push_back(u11[i]);
    last[i] = add(s->stream, d, [&]() { launch_mf_apply(*s); });
  }
  batches(BC_PRESS);
  if (kRK == 3)
    each([](cfd_solver* s) { return s->filt || s->P.flag_ctpress == 1; },
         [&](cfd_solver* s) {
           if (s->filt) launch_mode_filter(*s);
           if (s->P.flag_ctpress == 1) launch_flowrate(*s);
         });
  for (cfd_solver* s : e->m) ++s->nsub;
  if (err != hipSuccess && G) {
    hipGraphDestroy(G);
    G = nullptr;
  }
  *out = G;
  return err;
}

// Capture: the root stream records an event that every member's stream
// waits on; each member's substep kRK (and, after kRK = 3 of a member with a
// prescribed pressure gradient, its flow-rate integral, as cfd_step does)
// is launched on its own stream; the root waits on an event of each member.
// Batched: the graphs of ens_batched_graph instead.
// The graph holds kernel nodes only.  Members must have run at least one
// substep eagerly (so every kernel's launch attributes are set), have
// profiling off, no closure attached and no statistics request pending.
int ens_create(cfd_handle* hs, int n, bool batched, cfd_ensemble* out) {
  if (!hs || n < 1 || !out) return ens_fail(nullptr, CFD_EINVAL, "ensemble: bad arguments");
  *out = nullptr;
  cfd_ensemble_s* e = new cfd_ensemble_s();
  e->batched = batched;
  auto bail = [&](int rc, const std::string& msg) {
    g_ens_err = msg;
    ens_free(e);
    return rc;
  };
  for (int i = 0; i < n; ++i) {
    cfd_solver* s = static_cast<cfd_solver*>(hs[i]);
    if (!s) return bail(CFD_EINVAL, "ensemble: null member");
    if (i && s->dev != e->dev) return bail(CFD_EINVAL, "ensemble: members on different devices");
    if (s->prof || s->cl.cnn || s->st.pending)
      return bail(CFD_EINVAL, "ensemble: a member has profiling, a closure or a statistics request");
    if (!s->nsub) return bail(CFD_EINVAL, "ensemble: every member must run one substep before capture");
    for (cfd_solver* q : e->m)
      if (q == s) return bail(CFD_EINVAL, "ensemble: a member appears twice");
    if (!i) e->dev = s->dev;
    e->m.push_back(s);
  }
  if (hipSetDevice(e->dev) != hipSuccess || hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
    return bail(CFD_EHIP, "ensemble: stream or event");
  e->join.assign(n, nullptr);
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) != hipSuccess)
      return bail(CFD_EHIP, "ensemble: event");
  // nothing of the members may be in flight while capturing
  for (cfd_solver* s : e->m)
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(CFD_EHIP, "ensemble: member stream");
  if (batched) {
    std::string msg;
    if (int rc = ens_build_batches(e, msg)) return bail(rc, msg);
  }
  for (int k = 0; k < 3; ++k) {
    const int kRK = k + 1;
    if (batched) {
      hipGraph_t g = nullptr;
      const hipError_t err = ens_batched_graph(e, kRK, &g);
      e->g[k] = g;
      if (err != hipSuccess || !g)
        return bail(CFD_EHIP, std::string("ensemble: graph of batched substep ") + std::to_string(kRK) + ": " +
                                  hipGetErrorString(err));
      if (hipGraphInstantiate(&e->ge[k], g, nullptr, nullptr, 0) != hipSuccess)
        return bail(CFD_EHIP, "ensemble: graph instantiation");
      continue;
    }
    if (hipStreamBeginCapture(e->root, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return bail(CFD_EHIP, "ensemble: begin capture");
    hipError_t err = hipEventRecord(e->fork, e->root);
    int rc = CFD_OK;
    for (int i = 0; i < n && err == hipSuccess && rc == CFD_OK; ++i) {
      cfd_solver* s = e->m[i];
      err = hipStreamWaitEvent(s->stream, e->fork, 0);
      if (err != hipSuccess) break;
      rc = op_substep(s, kRK);
      if (kRK == 3 && s->P.flag_ctpress == 1) launch_flowrate(*s);
      if (err == hipSuccess) err = hipEventRecord(e->join[i], s->stream);
      if (err == hipSuccess) err = hipStreamWaitEvent(e->root, e->join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->root, &g);
    e->g[k] = g;
    if (err != hipSuccess || ec != hipSuccess || rc != CFD_OK || !g)
      return bail(CFD_EHIP, std::string("ensemble: capture of substep ") + std::to_string(kRK) + ": " +
                                hipGetErrorString(err != hipSuccess ? err : ec));
    if (hipGraphInstantiate(&e->ge[k], g, nullptr, nullptr, 0) != hipSuccess)
      return bail(CFD_EHIP, "ensemble: graph instantiation");
  }
  *out = e;
  return CFD_OK;
}
}  // namespace

extern "C" int cfd_ensemble_create(cfd_handle* hs, int n, cfd_ensemble* out) { return ens_create(hs, n, false, out); }

extern "C" int cfd_ensemble_create_batched(cfd_handle* hs, int n, cfd_ensemble* out) {
  return ens_create(hs, n, true, out);
}

extern "C" int cfd_ensemble_destroy(cfd_ensemble e) {
  if (!e) return CFD_OK;
  hipSetDevice(e->dev);
  if (e->root) hipStreamSynchronize(e->root);
  ens_free(e);
  return CFD_OK;
}

extern "C" const char* cfd_ensemble_last_error(cfd_ensemble e) { return e ? e->err.c_str() : g_ens_err.c_str(); }

extern "C" int cfd_ensemble_substep(cfd_ensemble e, int kRK) {
  if (!e || kRK < 1 || kRK > 3) return ens_fail(e, CFD_EINVAL, "ensemble: kRK must be 1, 2 or 3");
  hipSetDevice(e->dev);
  const hipError_t err = hipGraphLaunch(e->ge[kRK - 1], e->root);
  if (err != hipSuccess) return ens_fail(e, CFD_EHIP, std::string("ensemble: graph launch: ") + hipGetErrorString(err));
  return CFD_OK;
}

extern "C" int cfd_ensemble_step(cfd_ensemble e, int nsteps) {
  if (!e || nsteps < 0) return ens_fail(e, CFD_EINVAL, "ensemble: bad arguments");
  for (int it = 0; it < nsteps; ++it)
    for (int kRK = 1; kRK <= 3; ++kRK)
      if (int rc = cfd_ensemble_substep(e, kRK)) return rc;
  return CFD_OK;
}

extern "C" int cfd_ensemble_sync(cfd_ensemble e) {
  if (!e) return ens_fail(e, CFD_EINVAL, "null ensemble");
  hipSetDevice(e->dev);
  const hipError_t err = hipStreamSynchronize(e->root);
  if (err != hipSuccess) return ens_fail(e, CFD_EHIP, std::string("ensemble: ") + hipGetErrorString(err));
  return CFD_OK;
}
