// io.cpp -- field files in the reference layout (host code).
//
// The files are Fortran sequential unformatted: each record is its payload
// framed by a 4-byte little-endian length before and after.  Layout of the
// records: record_out (FOU3D.f90:2172-2345) and read_in (start.f90:2251-2700),
// summarised in cfd_amd.h.  The plane buffer of u_to_buff (FOU3D.f90:2598-2627)
// is the spectral plane with x-modes fastest and the z-modes in column order,
// which is also the device layout of one plane, so planes move as one block.
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "io.hpp"

namespace cfd {

namespace {

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// a record under construction
struct Rec {
  std::vector<unsigned char> b;
  template <typename T>
  void put(const T* p, size_t n) {
    const unsigned char* c = reinterpret_cast<const unsigned char*>(p);
    b.insert(b.end(), c, c + n * sizeof(T));
  }
  template <typename T>
  void put(T v) {
    put(&v, 1);
  }
};

bool write_rec(FILE* f, const Rec& r) {
  const uint32_t n = (uint32_t)r.b.size();
  return fwrite(&n, 4, 1, f) == 1 && (n == 0 || fwrite(r.b.data(), 1, n, f) == n) && fwrite(&n, 4, 1, f) == 1;
}

// read one record's payload; false at end of file or on a broken frame
bool read_rec(FILE* f, std::vector<unsigned char>& b) {
  uint32_t n = 0, m = 0;
  if (fread(&n, 4, 1, f) != 1) return false;
  b.resize(n);
  if (n && fread(b.data(), 1, n, f) != n) return false;
  if (fread(&m, 4, 1, f) != 1 || m != n) return false;
  return true;
}

template <typename T>
T get(const std::vector<unsigned char>& b, size_t off) {
  T v;
  std::memcpy(&v, b.data() + off, sizeof(T));
  return v;
}

Rec header_rec(const cfd_file_header& hd) {
  Rec r1;
  r1.put(hd.t);
  r1.put(hd.Re);
  r1.put(hd.alp);
  r1.put(hd.bet);
  r1.put(hd.mpgx);
  r1.put((int32_t)hd.nband);
  r1.put((int32_t)hd.iter);
  const int32_t zeros[88] = {};
  r1.put(zeros, 88);
  return r1;
}

Rec int20_rec(const int* v) {
  Rec r;
  for (int q = 0; q < 20; ++q) r.put((int32_t)v[q]);
  return r;
}

Rec grid_rec(const std::vector<double>& y, double dtheta, const std::vector<double>& dthdy) {
  Rec r;
  r.put(y.data(), y.size());
  r.put(dtheta);
  r.put(dthdy.data(), dthdy.size());
  return r;
}

}  // namespace

std::string write_field_file(const char* path, const cfd_file_header& hd, int id, double dthetavi, const double* y,
                             const double* dthdy, int ny_arr, int jy0, int ny_g, int j0, int nx, int nz,
                             const double* planes, const BandPlanes& bp) {
  if (j0 < jy0 || j0 + ny_g > jy0 + ny_arr) return "plane range outside the grid array";
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return std::string("cannot open ") + path + " for writing";
  const Rec r1 = header_rec(hd), r2 = int20_rec(hd.N);
  Rec r3;
  r3.put(y, (size_t)ny_arr);
  r3.put(dthetavi);
  r3.put(dthdy, (size_t)ny_arr);
  if (!write_rec(F.f, r1) || !write_rec(F.f, r2) || !write_rec(F.f, r3)) return std::string("write error on ") + path;
  const size_t pl = (size_t)nx * nz;
  std::vector<double> cb;
  for (int q = 0; q < ny_g; ++q) {
    Rec r;
    const bool cen = q >= bp.q0 && q < bp.q1;
    r.put((int32_t)(j0 + q));
    r.put((int32_t)id);
    r.put((int32_t)(cen ? bp.nx : nx));
    r.put((int32_t)(cen ? bp.nz : nz));
    r.put(y[j0 + q - jy0]);
    if (cen) {
      // the central band's modes of the plane: kx reals 1..bp.nx, the
      // bp.nz/2 lowest z modes and the bp.nz/2 most negative ones kept
      cb.assign((size_t)bp.nx * bp.nz, 0.0);
      const double* u = planes + q * pl;
      for (int kc = 0; kc < bp.nz; ++kc) {
        const int k = kc < bp.nz / 2 ? kc : kc - bp.nz + nz;
        for (int i = 0; i < bp.nx; ++i) cb[(size_t)kc * bp.nx + i] = u[(size_t)k * nx + i];
      }
      r.put(cb.data(), cb.size());
    } else {
      r.put(planes + q * pl, pl);
    }
    if (!write_rec(F.f, r)) return std::string("write error on ") + path;
  }
  return "";
}

std::string read_field_file(const char* path, cfd_file_header& hd, int ny_g, int nx, int nz, int ngz, double* planes,
                            int& mismatch, const BandPlanes& bp) {
  mismatch = 0;
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return std::string("cannot open ") + path;
  std::vector<unsigned char> b;
  if (!read_rec(F.f, b) || b.size() < 5 * 8 + 2 * 4) return std::string("bad header record in ") + path;
  hd.t = get<double>(b, 0);
  hd.Re = get<double>(b, 8);
  hd.alp = get<double>(b, 16);
  hd.bet = get<double>(b, 24);
  hd.mpgx = get<double>(b, 32);
  hd.nband = get<int32_t>(b, 40);
  hd.iter = get<int32_t>(b, 44);
  if (!read_rec(F.f, b)) return std::string("bad N record in ") + path;
  for (int q = 0; q < 20; ++q) hd.N[q] = (size_t)(q + 1) * 4 <= b.size() ? get<int32_t>(b, (size_t)q * 4) : 0;
  if (!read_rec(F.f, b)) return std::string("bad grid record in ") + path;
  const size_t pl = (size_t)nx * nz;
  std::memset(planes, 0, sizeof(double) * pl * ny_g);
  for (int q = 0; q < ny_g; ++q) {
    if (!read_rec(F.f, b) || b.size() < 4 * 4 + 8) return std::string("missing plane record in ") + path;
    const int nxf = get<int32_t>(b, 8), nzf = get<int32_t>(b, 12);
    if (nxf < 0 || nzf < 0 || b.size() != 4 * 4 + 8 + (size_t)nxf * nzf * 8)
      return std::string("bad plane record in ") + path;
    const bool cen = q >= bp.q0 && q < bp.q1;
    if (nxf != (cen ? bp.nx : nx) || nzf != (cen ? bp.nz : nz)) ++mismatch;
    const size_t d0 = 4 * 4 + 8;
    double* u = planes + q * pl;
    // buff_to_u into a plane of ngz z-slots, then the slots that carry one of
    // the nz target modes (1..nz/2 at the bottom, the rest at the top)
    const int kh = std::min(nzf / 2, ngz / 2), ix = std::min(nxf, nx);
    auto copy = [&](int kf, int slot) {
      int kt = -1;
      if (slot <= nz / 2) kt = slot;
      else if (slot >= ngz - (nz - nz / 2) + 1) kt = slot - ngz + nz;
      if (kt < 1 || kt > nz) return;
      for (int i = 0; i < ix; ++i)
        u[(size_t)(kt - 1) * nx + i] = get<double>(b, d0 + ((size_t)(kf - 1) * nxf + i) * 8);
    };
    for (int kf = 1; kf <= kh; ++kf) copy(kf, kf);
    for (int kf = nzf - kh + 1; kf <= nzf; ++kf) copy(kf, kf + ngz - nzf);
  }
  return "";
}

std::string write_stats_file(const char* path, const cfd_file_header& hd, const int Ngal[20], const YGrid& g,
                             int istat, const std::vector<double> prof[13], const std::vector<double> cond[13]) {
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return std::string("cannot open ") + path + " for writing";
  Rec ri;
  ri.put((int32_t)istat);
  // dthetai = 1/dtheta = 1 (start.f90:681-682)
  bool ok = write_rec(F.f, header_rec(hd)) && write_rec(F.f, int20_rec(Ngal)) &&
            write_rec(F.f, grid_rec(g.yu, 1.0, g.dthdyu)) && write_rec(F.f, grid_rec(g.yv, 1.0, g.dthdyv)) &&
            write_rec(F.f, ri);
  for (int q = 0; q < 13 && ok; ++q) {
    Rec r;
    r.put(prof[q].data(), prof[q].size());
    ok = write_rec(F.f, r);
  }
  for (int q = 0; q < 13 && ok; ++q) {
    Rec r;
    r.put(cond[q].data(), cond[q].size());
    ok = write_rec(F.f, r);
  }
  return ok ? "" : std::string("write error on ") + path;
}

std::string write_spectrum_file(const char* path, const cfd_file_header& hd, const YGrid& g, int istat,
                                const std::vector<double>& sp) {
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return std::string("cannot open ") + path + " for writing";
  Rec ri, rs;
  ri.put((int32_t)istat);
  rs.put(sp.data(), sp.size());
  const bool ok = write_rec(F.f, header_rec(hd)) && write_rec(F.f, int20_rec(hd.N)) &&
                  write_rec(F.f, grid_rec(g.yu, 1.0, g.dthdyu)) && write_rec(F.f, grid_rec(g.yv, 1.0, g.dthdyv)) &&
                  write_rec(F.f, ri) && write_rec(F.f, rs);
  return ok ? "" : std::string("write error on ") + path;
}

void band_table(const cfd_params& p, const YGrid& g, int N[20]) {
  // N(1:4, 0:nband+1), nband = 3 (start.f90:89-111, y_grid_canopy.f90:250-275)
  const int N1 = 2 * (p.ngalx / 3), N2 = 2 * (p.ngalz / 3), dny = p.dsty, nn = p.ny - 2;
  // the central band's modes (Ngal(1:2, 2), 0: as the wall bands')
  const int N1c = p.ngalx2 > 0 ? 2 * (p.ngalx2 / 3) : N1, N2c = p.ngalz2 > 0 ? 2 * (p.ngalz2 / 3) : N2;
  // height of the fine bottom band (start.f90:199-226 for the geometry types built here)
  const double pi = 3.14159265358979323846;
  const double h = p.geometry_type == 3 ? 2.0 * (2.0 * pi / p.bet) / p.ntilez : 0.25;
  const double target = g.YU(0) + h;
  int j = -dny;
  while (j < g.jlo_u + g.nyu - 1 && g.YU(j) < target) ++j;
  const int N31 = j, N32 = nn - j, N33 = nn + dny;
  const int tab[5][4] = {{0, 0, -dny, -dny}, {N1, N2, N31, N31 + 1}, {N1c, N2c, N32, N32}, {N1, N2, N33, N33 + 1},
                         {-2, 0, 0, 0}};
  for (int b = 0; b < 5; ++b)
    for (int r = 0; r < 4; ++r) N[b * 4 + r] = tab[b][r];
}

// The reference's plane partition over np ranks (start.f90:1652-1866),
// each band's load from its own Ngal(1:2, iband): ranks per band from its load balance, then contiguous ranges of
// planes per rank, the ranks with one plane more first.  lim[(g * 2 + e) *
// np + r] is planelim(grid, e + 1, r) for g 0 the u grid (band limits N(4,
// .)), 1 the v grid (N(3, .)); walls excluded.  False when np < 3 (the
// balance then gives the central band no rank and never ends) and when the
// band limits are not strictly increasing (tall posts on a short grid: a
// band with no or a negative number of planes, whose ranges overlap their
// neighbours').
bool plane_partition(const cfd_params& p, const int N[20], int np, int* lim) {
  if (np < 3) return false;
  auto NB = [&](int r, int b) { return N[b * 4 + r - 1]; };
  for (int r = 3; r <= 4; ++r)
    if (!(NB(r, 0) < NB(r, 1) && NB(r, 1) < NB(r, 2) && NB(r, 2) < NB(r, 3))) return false;
  const int bot = 1, mid = 2, top = 3;
  double load[4] = {0, 0, 0, 0}, il[4] = {0, 0, 0, 0};
  for (int b = 1; b <= 3; ++b) {
    const int gx = b == 2 && p.ngalx2 > 0 ? p.ngalx2 : p.ngalx, gz = b == 2 && p.ngalz2 > 0 ? p.ngalz2 : p.ngalz;
    load[b] = (double)(gx + 2) * gz * (NB(4, b) - NB(4, b - 1));
  }
  auto iceil = [](double x) { return (int)std::ceil(x); };
  int procs[4] = {0, 0, 0, 0};
  // ranks of the two fine bands with iplanes planes each, the rest to the
  // central band, and the per-rank loads
  auto assign = [&](int iplanes) {
    int nplanes = NB(4, bot) - NB(4, 0);
    procs[bot] = iceil(nplanes * 1.0 / iplanes);
    procs[top] = procs[bot];
    const int rem = procs[bot] + procs[top];
    il[bot] = load[bot] / nplanes * iplanes;
    il[top] = load[top] / nplanes * iplanes;
    nplanes = NB(4, mid) - NB(4, bot);
    procs[mid] = np - rem;
    if (procs[mid] > 0) il[mid] = load[mid] / nplanes * iceil(nplanes * 1.0 / procs[mid]);
  };
  auto maxl = [&]() { return std::max(il[1], std::max(il[2], il[3])); };
  int iplanes = 1;
  assign(1);
  while (procs[mid] <= 0) assign(++iplanes);
  double maxload = 2.0 * maxl();
  while (maxl() < maxload) {
    maxload = maxl();
    assign(++iplanes);
  }
  assign(--iplanes);
  // contiguous ranges per band: u grid over N(4, .), v grid over N(3, .)
  for (int g = 0; g < 2; ++g) {
    const int r = g == 0 ? 4 : 3;
    int first = 0;  // ranks of the bands below
    for (int b = 1; b <= 3; ++b) {
      const int nplanes = NB(r, b) - NB(r, b - 1), nr = procs[b];
      const int ip = nplanes / nr, rem = nplanes - ip * nr;
      for (int k = 1; k <= nr; ++k) {
        const int from = k <= rem ? (ip + 1) * (k - 1) + 1 : ip * (k - rem - 1) + (ip + 1) * rem + 1;
        const int to = k <= rem ? (ip + 1) * k : ip * (k - rem) + (ip + 1) * rem;
        lim[(g * 2 + 0) * np + first + k - 1] = from + NB(r, b - 1);
        lim[(g * 2 + 1) * np + first + k - 1] = to + NB(r, b - 1);
      }
      first += nr;
    }
  }
  return true;
}

}  // namespace cfd

extern "C" int cfd_plane_partition(const cfd_params* p, int np, int* lim) {
  if (!p || !lim || np < 3) return CFD_EINVAL;
  cfd::YGrid g;
  std::string err;
  if (!cfd::build_ygrid(*p, g, err)) return CFD_EINVAL;
  int N[20];
  cfd::band_table(*p, g, N);
  return cfd::plane_partition(*p, N, np, lim) ? CFD_OK : CFD_EINVAL;
}

extern "C" int cfd_field_file_write(const char* path, const cfd_file_header* hd, int id, double dthetavi,
                                    const double* y, const double* dthdy, int ny_arr, int jy0, int ny_g, int j0,
                                    int nx, int nz, const double* planes) {
  if (!path || !hd || !y || !dthdy || !planes || ny_arr <= 0 || ny_g <= 0 || nx <= 0 || nz <= 0) return CFD_EINVAL;
  const std::string e =
      cfd::write_field_file(path, *hd, id, dthetavi, y, dthdy, ny_arr, jy0, ny_g, j0, nx, nz, planes);
  if (!e.empty()) {
    fprintf(stderr, "cfd_field_file_write: %s\n", e.c_str());
    return CFD_EINVAL;
  }
  return CFD_OK;
}

extern "C" int cfd_field_file_read(const char* path, cfd_file_header* hd, int ny_g, int nx, int nz, int ngz,
                                   double* planes, int* mismatch) {
  if (!path || !hd || !planes || ny_g <= 0 || nx <= 0 || nz <= 0 || ngz < nz) return CFD_EINVAL;
  int mm = 0;
  const std::string e = cfd::read_field_file(path, *hd, ny_g, nx, nz, ngz, planes, mm);
  if (mismatch) *mismatch = mm;
  if (!e.empty()) {
    fprintf(stderr, "cfd_field_file_read: %s\n", e.c_str());
    return CFD_EINVAL;
  }
  return CFD_OK;
}

extern "C" int cfd_band_table(const cfd_params* p, int* N20) {
  if (!p || !N20) return CFD_EINVAL;
  cfd::YGrid g;
  std::string err;
  if (!cfd::build_ygrid(*p, g, err)) {
    fprintf(stderr, "cfd_band_table: %s\n", err.c_str());
    return CFD_EINVAL;
  }
  cfd::band_table(*p, g, N20);
  return CFD_OK;
}
