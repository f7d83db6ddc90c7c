// parameters.xml, read as the reference's Parameters class does
// (include/parameters.h:17-140, src/parameters.cpp:6-318), plus one optional
// <device> section for the choices the reference makes at compile time:
//
//   <device>
//     <precision>fp64</precision>      <!-- fp64 (plain form) | fp32 (cell-relative form) -->
//     <pusher>boris</pusher>           <!-- boris | vay (PUSHER_* of config.h) -->
//     <sort_interval>0</sort_interval> <!-- cell sort of plasma species every N steps, 0 = never -->
//   </device>
#pragma once

#include <string>
#include <vector>

namespace pdp3 {

// struct probe, include/parameters.h:17-27
struct ProbeSpec {
  std::string component;
  std::string specie;  // "none" when absent
  int type;            // 0 frame, 1 col, 2 row, 3 dot, 4 mpframe
  int r_start, r_end, z_start, z_end;
  int schedule;
};

// struct particle_specie, include/parameters.h:29-37
struct SpecieSpec {
  std::string name;
  unsigned int mass;
  int charge;
  double number_macro;
  double left_density, right_density;
  double temperature;
};

class Parameters {
 public:
  // throws ConfigError: code 78 when the file can not be read (the
  // reference's exit(78), src/parameters.cpp:13-17), 1 for missing or invalid
  // entries (its abort(), src/parameters.cpp:45-58)
  explicit Parameters(const std::string &xml_file_name);

  bool debug = false;
  bool use_hdf5 = false;
  // <geometry>
  double r_size = 0, z_size = 0;
  int n_grid_r = 0, n_grid_z = 0;
  // <pml>
  double pml_comparative_l_1 = 0, pml_comparative_l_2 = 0, pml_comparative_l_3 = 0;
  double pml_sigma_1 = 0, pml_sigma_2 = 0;
  // <time>
  double start_time = 0, relaxation_time = 0, current_time = 0, end_time = 0, step_interval = 0;
  // <particles>
  std::vector<SpecieSpec> particle_species;
  // <particle_beam>
  std::string beam_name;
  int beam_particle_charge = 0;
  unsigned int beam_particle_mass = 0;
  unsigned int beam_number_bunches = 0;
  double beam_bunches_distance = 0, beam_initial_velocity = 0;
  unsigned int bunch_number_macro = 0;
  double bunch_lenght = 0, bunch_radius = 0, bunch_density = 0;
  // <boundary_maxwell_conditions>, <boundary_conditions>
  double boundary_maxwell_e_phi_upper = 0, boundary_maxwell_e_phi_left = 0, boundary_maxwell_e_phi_right = 0;
  int boundary_conditions = 0;
  // <file_save_parameters>
  std::string dump_result_path, dump_data_root;
  unsigned int dump_frames_per_file = 1;
  bool dump_compress = false;
  int dump_compress_level = 0;
  std::vector<ProbeSpec> probes;
  // <device>
  int precision = 0;  // PDP3_PREC_*
  int pusher = 0;     // PDP3_PUSHER_*
  int sort_interval = 0;
};

struct ConfigError : std::exception {
  int code;
  std::string msg;
  ConfigError(int c, std::string m) : code(c), msg(std::move(m)) {}
  const char *what() const noexcept override { return msg.c_str(); }
};

}  // namespace pdp3
