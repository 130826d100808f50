// io.hpp -- field files in the reference layout (host code, io.cpp).
#pragma once
#include <string>
#include <vector>

#include "cfd_amd.h"
#include "setup.hpp"

namespace cfd {

// planes [q0, q1) of a field file that carry the central band's nx x nz
// modes (multi-band resolution: record_out writes each plane at its band's
// N(1, band)+2 x N(2, band), FOU3D.f90:2172-2345); q0 == q1: none
struct BandPlanes {
  int q0 = 0, q1 = 0, nx = 0, nz = 0;
};
// one field file; "" on success, else the error message
std::string write_field_file(const char* path, const cfd_file_header& hd, int id, double dthetavi, const double* y,
                             const double* dthdy, int ny_arr, int jy0, int ny_g, int j0, int nx, int nz,
                             const double* planes, const BandPlanes& bp = BandPlanes());
// planes of a size other than nx x nz (bp's on its planes) are counted in
// mismatch and embedded as buff_to_u embeds them
std::string read_field_file(const char* path, cfd_file_header& hd, int ny_g, int nx, int nz, int ngz, double* planes,
                            int& mismatch, const BandPlanes& bp = BandPlanes());
// write_stats (stats.f90:255-442): header, Ngal(4,0:4), yu/dthetai/dthdyu,
// yv/dthetai/dthdyv, istat, the 13 profiles, the 13 conditional arrays
std::string write_stats_file(const char* path, const cfd_file_header& hd, const int Ngal[20], const YGrid& g,
                             int istat, const std::vector<double> prof[13], const std::vector<double> cond[13]);
// recvwrspec (spectra.f90:132-182): header, N, the two grid records, istat,
// buffSP(0:N1/2, 1:N2/2+1, planes)
std::string write_spectrum_file(const char* path, const cfd_file_header& hd, const YGrid& g, int istat,
                                const std::vector<double>& sp);
// N(1:4, 0:4) of the reference for this configuration
void band_table(const cfd_params& p, const YGrid& g, int N[20]);
bool plane_partition(const cfd_params& p, const int N[20], int np, int* lim);

}  // namespace cfd
