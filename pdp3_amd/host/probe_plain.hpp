// Plain-text probe writer (ProbePlain, include/probePlain.h,
// src/probePlain.cpp:9-224): the same directory layout, file names and text
// format as the reference, so its post-processing tools read the output.
//
//   <root>/<component>[/<specie>]/frame_<r0>:<z0>_<r1>:<z1>/<file>.dat
//   <root>/<component>[/<specie>]/col_<z>/..., row_<r>/..., dot_<r>_<z>/...,
//   mpframe_<r0>:<z0>_<r1>:<z1>/...
//
// Values are appended with the default ostream format (6 significant digits)
// followed by one blank, row after row.
#pragma once

#include <string>
#include <vector>

#include "parameters.hpp"

namespace pdp3 {

class ProbePlain {
 public:
  ProbePlain(const std::string &root, const ProbeSpec &spec);
  // type 0..3: `data` is a row-major grid with `row_stride` entries per row
  void write(const std::string &name, const double *data, size_t row_stride) const;
  // type 4: particles inside the frame, src/probePlain.cpp:180-223
  void mpwrite(const std::string &name, double dr, double dz, const std::vector<double> &r,
               const std::vector<double> &z, const std::vector<double> &vr, const std::vector<double> &vphi,
               const std::vector<double> &vz) const;
  const std::string &path() const { return path_; }

  ProbeSpec spec;
  int step_number = 0;

 private:
  std::string path_;
  std::string data_path(const std::string &name) const { return path_ + "/" + name + ".dat"; }
};

// mkdir -p
void make_directory(const std::string &path);

}  // namespace pdp3
