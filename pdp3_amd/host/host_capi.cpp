// extern "C" wrappers of the host side (include/pdp3_host.h)
#include "../../include/pdp3_host.h"

#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include "../../include/pdp3_hip.h"
#include "load_init_param.hpp"
#include "loaders.hpp"

namespace {
thread_local std::string g_error;

int set_error(int code, const std::string &msg) {
  g_error = msg;
  return code;
}
}  // namespace

extern "C" {

double pdp3_host_random_reverse(double vel, int power) { return pdp3::loaders::random_reverse(vel, power); }

int pdp3_host_load_cylindrical(int n, double r_size, double z_size, double dr, double dz, double n1, double n2,
                               double left_plasma_boundary, double *r, double *z, double *w) {
  if (n <= 0 || !r || !z || !w) return PDP3_ERR_BAD_ARG;
  pdp3::loaders::load_cylindrical(n, r_size, z_size, dr, dz, n1, n2, left_plasma_boundary, r, z, w);
  return PDP3_OK;
}

int pdp3_host_maxwell_juttner(int n, double temperature_ev, double *energies) {
  if (n < 0 || (n > 0 && !energies) || !(temperature_ev > 0)) return PDP3_ERR_BAD_ARG;
  pdp3::loaders::maxwell_juttner(n, temperature_ev, energies);
  return PDP3_OK;
}

int pdp3_host_velocity_distribution(int n, double mass_kg, double temperature_ev, double *vr, double *vphi,
                                    double *vz) {
  if (n < 0 || (n > 0 && (!vr || !vphi || !vz)) || !(mass_kg > 0) || !(temperature_ev > 0)) return PDP3_ERR_BAD_ARG;
  pdp3::loaders::velocity_distribution(n, mass_kg, temperature_ev, vr, vphi, vz);
  return PDP3_OK;
}

int pdp3_host_bunch_positions(int n, double radius, double dr, double dz, double duration, double velocity,
                              double density, double *r, double *w) {
  if (n <= 0 || !r || !w) return PDP3_ERR_BAD_ARG;
  pdp3::loaders::bunch_positions(n, radius, dr, dz, duration, velocity, density, r, w);
  return PDP3_OK;
}

int pdp3_host_bunch_inject_draws(int n, double *u) {
  if (n < 0 || (n > 0 && !u)) return PDP3_ERR_BAD_ARG;
  pdp3::loaders::bunch_inject_draws(n, u);
  return PDP3_OK;
}

int pdp3_host_mb_table(double *out) {
  if (!out) return PDP3_ERR_BAD_ARG;
  std::memcpy(out, pdp3::loaders::mb_table(), 1000 * sizeof(double));
  return PDP3_OK;
}

int pdp3_host_mj_table(double *out) {
  if (!out) return PDP3_ERR_BAD_ARG;
  std::memcpy(out, pdp3::loaders::mj_table(), 1000 * sizeof(double));
  return PDP3_OK;
}

int pdp3_host_parameters_dump(const char *xml_file_name, char *buf, size_t buf_size) {
  if (!xml_file_name || !buf || buf_size == 0) return PDP3_ERR_BAD_ARG;
  try {
    pdp3::Parameters p(xml_file_name);
    std::ostringstream o;
    o.precision(17);
    o << "debug=" << p.debug << "\nuse_hdf5=" << p.use_hdf5 << "\nr_size=" << p.r_size << "\nz_size=" << p.z_size
      << "\nn_grid_r=" << p.n_grid_r << "\nn_grid_z=" << p.n_grid_z << "\npml_l1=" << p.pml_comparative_l_1
      << "\npml_l2=" << p.pml_comparative_l_2 << "\npml_l3=" << p.pml_comparative_l_3
      << "\npml_sigma1=" << p.pml_sigma_1 << "\npml_sigma2=" << p.pml_sigma_2 << "\nstart_time=" << p.start_time
      << "\nrelaxation_time=" << p.relaxation_time << "\ncurrent_time=" << p.current_time
      << "\nend_time=" << p.end_time << "\nstep_interval=" << p.step_interval;
    for (size_t i = 0; i < p.particle_species.size(); i++) {
      const auto &s = p.particle_species[i];
      o << "\nspecies" << i << "=" << s.name << "," << s.charge << "," << s.mass << "," << s.number_macro << ","
        << s.left_density << "," << s.right_density << "," << s.temperature;
    }
    o << "\nbeam=" << p.beam_name << "," << p.beam_particle_charge << "," << p.beam_particle_mass << ","
      << p.beam_number_bunches << "," << p.beam_bunches_distance << "," << p.beam_initial_velocity << ","
      << p.bunch_number_macro << "," << p.bunch_lenght << "," << p.bunch_radius << "," << p.bunch_density
      << "\nresult_path=" << p.dump_result_path << "\ndata_root=" << p.dump_data_root
      << "\nframes_per_file=" << p.dump_frames_per_file << "\ncompress=" << p.dump_compress;
    for (size_t i = 0; i < p.probes.size(); i++) {
      const auto &q = p.probes[i];
      o << "\nprobe" << i << "=" << q.type << "," << q.component << "," << q.specie << "," << q.r_start << ","
        << q.z_start << "," << q.r_end << "," << q.z_end << "," << q.schedule;
    }
    o << "\nprecision=" << p.precision << "\npusher=" << p.pusher << "\nsort_interval=" << p.sort_interval << "\n";
    const std::string s = o.str();
    if (s.size() + 1 > buf_size) return set_error(PDP3_ERR_BAD_ARG, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return PDP3_OK;
  } catch (const pdp3::ConfigError &e) {
    return set_error(e.code, e.what());
  } catch (const std::exception &e) {
    return set_error(1, e.what());
  }
}

int pdp3_host_run(const char *xml_file_name, int quiet, int *steps) {
  if (!xml_file_name) return PDP3_ERR_BAD_ARG;
  try {
    pdp3::LoadInitParam lip(xml_file_name);
    lip.quiet = quiet != 0;
    int n = lip.run();
    if (steps) *steps = n;
    return PDP3_OK;
  } catch (const pdp3::ConfigError &e) {
    return set_error(e.code, e.what());
  } catch (const pdp3::Error &e) {
    return set_error(e.code, e.what());
  } catch (const std::exception &e) {
    return set_error(1, e.what());
  }
}

const char *pdp3_host_last_error(void) { return g_error.c_str(); }

}  // extern "C"
