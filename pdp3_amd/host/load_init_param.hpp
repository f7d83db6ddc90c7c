// The drop-in driver: LoadInitParam (include/loadInitParam.h,
// src/loadInitParam.cpp) over the MI355X device.  Reads parameters.xml,
// builds the geometry, fields, plasma species (TESTMODE loaders) and the
// bunches of the beam, then runs the loop of LoadInitParam::run
// (src/loadInitParam.cpp:455-589) one fused device step at a time and writes
// the probes after each step.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "parameters.hpp"
#include "pdp3.hpp"
#include "probe_plain.hpp"

namespace pdp3 {

class LoadInitParam {
 public:
  explicit LoadInitParam(const std::string &xml_file_name);
  // the main loop; returns the number of steps run
  int run();

  Parameters params;
  std::unique_ptr<Geometry> geom;
  std::unique_ptr<Time> time;
  std::unique_ptr<Device> device;
  std::unique_ptr<EField> efield;
  std::unique_ptr<HField> hfield;
  std::unique_ptr<Current> c_current;
  std::unique_ptr<ChargeDensity> c_rho_bunch;
  std::unique_ptr<TemperatureClassic> temperature;  // the "T" probes (device)
  std::unique_ptr<ParticlesDensity> density;        // the "density" probes (device)
  std::vector<std::unique_ptr<Particles>> species;  // plasma species, then bunches
  ParticlesList p_list;
  std::vector<ProbePlain> c_probes;
  bool quiet = false;

 private:
  void init_fields();
  void init_particles();
  void init_beam();
  void dump_data(double step_time);
  void print_data(const ProbePlain &w, double step_time);
  int printed_step_ = 0;
};

}  // namespace pdp3
