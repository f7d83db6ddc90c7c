#include "load_init_param.hpp"

#include <chrono>
#include <iomanip>
#include <iostream>
#include <map>

namespace pdp3 {

namespace {
std::chrono::steady_clock::time_point g_start = std::chrono::steady_clock::now();

std::string duration_since_start() {
  long s = (long)std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - g_start).count();
  char buf[64];
  std::snprintf(buf, sizeof buf, "%ldh %ldm %lds", s / 3600, (s / 60) % 60, s % 60);
  return buf;
}

int grid_component(const std::string &c) {
  static const std::map<std::string, int> m = {{"E_r", PDP3_E_R},     {"E_phi", PDP3_E_PHI}, {"E_z", PDP3_E_Z},
                                               {"H_r", PDP3_H_R},     {"H_phi", PDP3_H_PHI}, {"H_z", PDP3_H_Z},
                                               {"J_r", PDP3_J_R},     {"J_phi", PDP3_J_PHI}, {"J_z", PDP3_J_Z},
                                               {"rho_beam", PDP3_RHO_BUNCH}};
  auto it = m.find(c);
  return it == m.end() ? -1 : it->second;
}
}  // namespace

LoadInitParam::LoadInitParam(const std::string &xml_file_name) : params(xml_file_name) {
  if (params.use_hdf5)
    throw ConfigError(1, "ERROR! PDP3 build without HDF5 support. Please, update your configfile to use plaintext "
                         "``<use_hdf5>false</use_hdf5>''");
  geom = std::make_unique<Geometry>(params.r_size, params.z_size, params.n_grid_r, params.n_grid_z);
  time = std::make_unique<Time>(params.current_time, params.start_time, params.relaxation_time, params.end_time,
                                params.step_interval);
  device = std::make_unique<Device>(*geom, *time, params.pusher, params.precision);
  device->check(pdp3_set_sort_interval(device->ctx(), params.sort_interval));
  geom->set_epsilon();
  // Parameters::init_pml, src/parameters.cpp:246-247
  if (params.pml_comparative_l_1 != 0 || params.pml_comparative_l_2 != 0 || params.pml_comparative_l_3 != 0)
    geom->set_pml(params.pml_comparative_l_1, params.pml_comparative_l_2, params.pml_comparative_l_3,
                  params.pml_sigma_1, params.pml_sigma_2);

  // probes, src/loadInitParam.cpp:65-83
  const std::string root = params.dump_result_path + "/" + params.dump_data_root;
  if (params.dump_compress) std::cerr << "WARNING! compression is not supported by plaintext backend" << std::endl;
  for (const ProbeSpec &s : params.probes) c_probes.emplace_back(root, s);

  init_fields();
  init_particles();
  init_beam();
}

// src/loadInitParam.cpp:250-259
void LoadInitParam::init_fields() {
  efield = std::make_unique<EField>(geom.get());
  hfield = std::make_unique<HField>(geom.get());
  efield->boundary_conditions();
  efield->set_homogeneous_efield(0., 0., 0.);
  hfield->set_homogeneous_h(0., 0., 0.);
  c_current = std::make_unique<Current>(geom.get());
  c_rho_bunch = std::make_unique<ChargeDensity>(geom.get());
  temperature = std::make_unique<TemperatureClassic>(geom.get());
  density = std::make_unique<ParticlesDensity>(geom.get());
}

// src/loadInitParam.cpp:138-166
void LoadInitParam::init_particles() {
  for (const SpecieSpec &sp : params.particle_species) {
    auto p = std::make_unique<Particles>(sp.name, sp.charge, sp.mass, (int)sp.number_macro, geom.get());
    p->load_cylindrical_spatial_distribution(sp.left_density, sp.right_density, 0);
    p->velocity_distribution(sp.temperature);
    p_list.part_list.push_back(p.get());
    species.push_back(std::move(p));
  }
}

// the bunches of LoadInitParam::init_beam (src/loadInitParam.cpp:168-248):
// created here with their numbers; each joins the loop at its injection time
void LoadInitParam::init_beam() {
  if (params.bunch_number_macro == 0 || params.beam_initial_velocity == 0) return;
  const double duration = params.bunch_lenght / params.beam_initial_velocity;
  const double hole_duration = params.beam_bunches_distance / params.beam_initial_velocity;
  for (unsigned int n = 0; n < params.beam_number_bunches; n++) {
    auto b = std::make_unique<Bunch>(params.beam_name + "#" + std::to_string(n), params.beam_particle_charge,
                                     params.beam_particle_mass, (int)params.bunch_number_macro, geom.get(), duration,
                                     params.bunch_radius, params.bunch_density, params.beam_initial_velocity, (int)n,
                                     hole_duration);
    p_list.part_list.push_back(b.get());
    species.push_back(std::move(b));
  }
}

// src/loadInitParam.cpp:261-317
void LoadInitParam::print_data(const ProbePlain &w, double step_time) {
  if (quiet) return;
  if (printed_step_ % 20 == 0)
    std::cout << std::endl
              << std::left << std::setw(8) << "Step" << std::setw(13) << "Saved Frame" << std::setw(25)
              << "Dumping Probe Name" << std::setw(21) << "Shape" << std::setw(18) << "Model Time (sec)"
              << std::setw(21) << "Simulation Duration" << std::endl;
  static const char *kinds[] = {"frame", "column", "row", "dot", "mpframe"};
  std::string name = std::string(kinds[w.spec.type]) + "/" + w.spec.component;
  if (w.spec.specie != "none") name += "/" + w.spec.specie;
  const ProbeSpec &s = w.spec;
  std::string shape;
  switch (s.type) {
    case 1:
      shape = "[" + std::to_string(s.z_start) + "]";
      break;
    case 2:
      shape = "[" + std::to_string(s.r_start) + "]";
      break;
    case 3:
      shape = "[" + std::to_string(s.r_start) + "," + std::to_string(s.z_start) + "]";
      break;
    default:
      shape = "[" + std::to_string(s.r_start) + "," + std::to_string(s.z_start) + "," + std::to_string(s.r_end) +
              "," + std::to_string(s.z_end) + "]";
  }
  std::cout << std::left << std::setw(8) << w.step_number * s.schedule << std::setw(13) << w.step_number
            << std::setw(25) << name << std::setw(21) << shape << std::setw(18) << step_time << std::setw(18)
            << duration_since_start() << std::endl;
  ++printed_step_;
}

// LoadInitParam::dump_data, src/loadInitParam.cpp:319-407.  The bunch density
// is written as the loop left it (one charge weighting per step, which the
// reference's own smoke goldens hold).
void LoadInitParam::dump_data(double step_time) {
  std::map<int, std::vector<double>> grids;  // downloads shared by the probes of one step
  for (ProbePlain &w : c_probes) {
    if ((int)(step_time / time->delta_t) % w.spec.schedule != 0) continue;
    const std::string file_name = std::to_string(w.step_number / (int)params.dump_frames_per_file);
    print_data(w, step_time);
    const int comp = grid_component(w.spec.component);
    if (w.spec.type == 4) {
      for (auto &p : species)
        if (p->name == w.spec.component) {
          p->download();
          w.mpwrite(std::to_string(w.step_number), geom->dr, geom->dz, p->pos_r, p->pos_z, p->vel_r, p->vel_phi,
                    p->vel_z);
        }
    } else if (comp >= 0) {
      auto it = grids.find(comp);
      if (it == grids.end()) {
        std::vector<double> g((size_t)geom->n_grid_r * (size_t)geom->n_grid_z);
        device->check(pdp3_grid_download(device->ctx(), comp, g.data()));
        it = grids.emplace(comp, std::move(g)).first;
      }
      w.write(file_name, it->second.data(), (size_t)geom->n_grid_z);
    } else if (w.spec.component == "T") {
      // src/loadInitParam.cpp:369-379, on the device (no particle download)
      for (auto &p : species)
        if (p->name == w.spec.specie) {
          temperature->reset();
          temperature->calc_t(p.get());
          temperature->normalize(p.get());
          w.write(file_name, temperature->t.data(), (size_t)geom->n_grid_z);
        }
    } else if (w.spec.component == "density") {
      // src/loadInitParam.cpp:381-390, on the device
      for (auto &p : species)
        if (p->name == w.spec.specie) {
          density->reset();
          density->calc_density(p.get());
          w.write(file_name, density->density.data(), (size_t)geom->n_grid_z);
        }
    } else {
      std::cerr << "WARNING! probe for component ``" << w.spec.component << "'' does not exist. Skipping"
                << std::endl;
    }
    ++w.step_number;
  }
}

// LoadInitParam::run, src/loadInitParam.cpp:455-589
int LoadInitParam::run() {
  if (!quiet) std::cout << std::endl << "Launch Simulation" << std::endl << std::endl;
  int steps = 0;
  while (time->current_time < time->end_time) {
    const double step_time = time->current_time;
    device->step(*time, 1);  // init_beam, H, J reset, push + deposit, E, bunch rho, t += dt
    dump_data(step_time);
    steps++;
  }
  if (!quiet) std::cout << std::endl << "Simulation Completed" << std::endl << std::endl;
  return steps;
}

}  // namespace pdp3
