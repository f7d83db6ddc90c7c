#include "probe_plain.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <fstream>
#include <stdexcept>

namespace pdp3 {

void make_directory(const std::string &path) {
  std::string cur;
  size_t pos = 0;
  while (pos != std::string::npos) {
    pos = path.find('/', pos + 1);
    cur = path.substr(0, pos);
    if (cur.empty()) continue;
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
      throw std::runtime_error("can not create directory ``" + cur + "''");
  }
}

ProbePlain::ProbePlain(const std::string &root, const ProbeSpec &s) : spec(s) {
  std::string comp = s.specie == "none" ? s.component : s.component + "/" + s.specie;
  const std::string base = root + "/" + comp + "/";
  switch (s.type) {
    case 0:
      path_ = base + "frame_" + std::to_string(s.r_start) + ":" + std::to_string(s.z_start) + "_" +
              std::to_string(s.r_end) + ":" + std::to_string(s.z_end);
      break;
    case 1:
      path_ = base + "col_" + std::to_string(s.z_start);
      break;
    case 2:
      path_ = base + "row_" + std::to_string(s.r_start);
      break;
    case 3:
      path_ = base + "dot_" + std::to_string(s.r_start) + "_" + std::to_string(s.z_start);
      break;
    case 4:
      path_ = base + "mpframe_" + std::to_string(s.r_start) + ":" + std::to_string(s.z_start) + "_" +
              std::to_string(s.r_end) + ":" + std::to_string(s.z_end);
      break;
    default:
      throw std::runtime_error("unknown probe type");
  }
  make_directory(path_);
}

void ProbePlain::write(const std::string &name, const double *data, size_t row_stride) const {
  std::ofstream out(data_path(name), std::ios::app);
  if (!out) throw std::runtime_error("can not write ``" + data_path(name) + "''");
  auto v = [&](int i, int k) { return data[(size_t)i * row_stride + (size_t)k]; };
  switch (spec.type) {
    case 0:  // write_frame, src/probePlain.cpp:83-101
      for (int i = spec.r_start; i < spec.r_end; i++)
        for (int k = spec.z_start; k < spec.z_end; k++) out << v(i, k) << " ";
      break;
    case 1:  // write_col: one z column over r_start .. r_end-1
      for (int i = spec.r_start; i < spec.r_end; i++) out << v(i, spec.z_start) << " ";
      break;
    case 2:  // write_row
      for (int k = spec.z_start; k < spec.z_end; k++) out << v(spec.r_start, k) << " ";
      break;
    case 3:  // write_dot
      out << v(spec.r_start, spec.z_start) << " ";
      break;
    default:
      break;
  }
}

void ProbePlain::mpwrite(const std::string &name, double dr, double dz, const std::vector<double> &r,
                         const std::vector<double> &z, const std::vector<double> &vr,
                         const std::vector<double> &vphi, const std::vector<double> &vz) const {
  std::ofstream fr(data_path(name + "_vel_r"), std::ios::app), fphi(data_path(name + "_vel_phi"), std::ios::app),
      fz(data_path(name + "_vel_z"), std::ios::app), fpr(data_path(name + "_pos_r"), std::ios::app),
      fpz(data_path(name + "_pos_z"), std::ios::app);
  for (size_t j = 0; j < r.size(); j++) {
    // cell numbers as unsigned, src/probePlain.cpp:209-212
    unsigned int ri = (unsigned int)((int)std::ceil(r[j] / dr) - 1);
    unsigned int zk = (unsigned int)((int)std::ceil(z[j] / dz) - 1);
    if (ri >= (unsigned)spec.r_start && ri <= (unsigned)spec.r_end && zk >= (unsigned)spec.z_start &&
        zk <= (unsigned)spec.z_end) {
      fr << vr[j] << " ";
      fphi << vphi[j] << " ";
      fz << vz[j] << " ";
      fpr << r[j] << " ";
      fpz << z[j] << " ";
    }
  }
}

}  // namespace pdp3
