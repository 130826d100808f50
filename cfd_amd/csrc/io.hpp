// io.hpp -- field files in the reference layout (host code, io.cpp).
#pragma once
#include <string>

#include "cfd_amd.h"
#include "setup.hpp"

namespace cfd {

// one field file; "" on success, else the error message
std::string write_field_file(const char* path, const cfd_file_header& hd, int id, double dthetavi, const double* y,
                             const double* dthdy, int ny_arr, int jy0, int ny_g, int j0, int nx, int nz,
                             const double* planes);
std::string read_field_file(const char* path, cfd_file_header& hd, int ny_g, int nx, int nz, int ngz, double* planes,
                            int& mismatch);
// N(1:4, 0:4) of the reference for this configuration (uniform resolution)
void band_table(const cfd_params& p, const YGrid& g, int N[20]);

}  // namespace cfd
