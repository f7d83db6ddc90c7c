#include "parameters.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "../../include/pdp3_hip.h"
#include "xml.hpp"

namespace pdp3 {

namespace {

using xml::Element;

const Element *child(const Element *e, const char *name) {
  const Element *c = e->first_child(name);
  if (!c)
    throw ConfigError(1, std::string("ERROR: Can not get child, named ``") + name + "'' from XML element ``" + e->name +
                             "''. Seems, configuration file is corrupted, or inconsistent");
  return c;
}
const char *attr(const Element *e, const char *name, bool required = true) {
  const char *a = e->attribute(name);
  if (!a && required)
    throw ConfigError(1, std::string("ERROR: Can not get attribute, named ``") + name + "'' from XML element ``" +
                             e->name + "''. Seems, configuration file is corrupted, or inconsistent");
  return a;
}
// the reference converts with atoi / atof (src/parameters.cpp), which read
// a leading number and ignore the rest ("1e5" -> 1 through atoi)
int as_int(const Element *e) { return std::atoi(e->text.c_str()); }
double as_double(const Element *e) { return std::atof(e->text.c_str()); }
// lib::to_bool, src/lib.cpp:36-43: "true" in any letter case is true,
// anything else false
bool as_bool(const Element *e) {
  std::string s = e->text;
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s == "true";
}

}  // namespace

Parameters::Parameters(const std::string &xml_file_name) {
  xml::Document doc;
  try {
    doc = xml::load_file(xml_file_name);
  } catch (const std::exception &ex) {
    throw ConfigError(78, std::string("ERROR: Can not read configuration file ``") + xml_file_name + "``: " + ex.what());
  }
  const Element *root = doc.root.get();
  if (root->name != "initial_parameters")
    throw ConfigError(1, "ERROR: the root element of ``" + xml_file_name + "'' is not <initial_parameters>");

  debug = as_bool(child(root, "debug"));
  use_hdf5 = as_bool(child(root, "use_hdf5"));

  // init_geometry, src/parameters.cpp:214-232
  const Element *g = child(root, "geometry");
  r_size = as_double(child(g, "r_size"));
  z_size = as_double(child(g, "z_size"));
  n_grid_r = as_int(child(g, debug ? "debug_n_grid_r" : "n_grid_r"));
  n_grid_z = as_int(child(g, debug ? "debug_n_grid_z" : "n_grid_z"));

  // init_pml, src/parameters.cpp:234-255
  const Element *p = child(root, "pml");
  pml_comparative_l_1 = as_double(child(p, "comparative_l_1"));
  pml_comparative_l_2 = as_double(child(p, "comparative_l_2"));
  pml_comparative_l_3 = as_double(child(p, "comparative_l_3"));
  pml_sigma_1 = as_double(child(p, "sigma_1"));
  pml_sigma_2 = as_double(child(p, "sigma_2"));

  // init_time, src/parameters.cpp:257-270
  const Element *t = child(root, "time");
  start_time = as_double(child(t, "start_time"));
  relaxation_time = as_double(child(t, "relaxation_time"));
  current_time = as_double(child(t, "current_time"));
  end_time = as_double(child(t, "end_time"));
  step_interval = as_double(child(t, "step_interval"));

  // init_particles, src/parameters.cpp:77-104
  for (const Element *s : child(root, "particles")->children_named("particle_specie")) {
    SpecieSpec sp;
    sp.name = attr(s, "name");
    sp.charge = as_int(child(s, "charge"));
    sp.mass = (unsigned int)as_int(child(s, "mass"));
    sp.number_macro = as_double(child(s, debug ? "debug_number_macro" : "number_macro"));
    sp.left_density = as_double(child(s, "left_density"));
    sp.right_density = as_double(child(s, "right_density"));
    sp.temperature = as_double(child(s, "temperature"));
    particle_species.push_back(sp);
  }

  // init_beam, src/parameters.cpp:176-199
  const Element *b = child(root, "particle_beam");
  const char *bn = b->attribute("name");
  beam_name = bn ? bn : "";
  beam_particle_charge = as_int(child(b, "charge"));
  beam_particle_mass = (unsigned int)as_int(child(b, "mass"));
  beam_number_bunches = (unsigned int)as_int(child(b, "number_bunches"));
  beam_bunches_distance = as_double(child(b, "bunches_distance"));
  beam_initial_velocity = as_double(child(b, "initial_velocity"));
  bunch_number_macro = (unsigned int)as_double(child(b, debug ? "debug_bunch_number_macro" : "bunch_number_macro"));
  bunch_lenght = as_double(child(b, "bunch_length"));
  bunch_radius = as_double(child(b, "bunch_radius"));
  bunch_density = as_double(child(b, "bunch_density"));

  // init_boundary, src/parameters.cpp:201-212 (optional here: the
  // reference's constructor never calls it)
  if (const Element *bm = root->first_child("boundary_maxwell_conditions")) {
    boundary_maxwell_e_phi_upper = as_double(child(bm, "e_fi_upper"));
    boundary_maxwell_e_phi_left = as_double(child(bm, "e_fi_left"));
    boundary_maxwell_e_phi_right = as_double(child(bm, "e_fi_right"));
  }
  if (const Element *bc = root->first_child("boundary_conditions")) boundary_conditions = as_int(bc);

  // init_data_dump_parameters, src/parameters.cpp:272-318 (non-LEGACY)
  const Element *f = child(root, "file_save_parameters");
  dump_result_path = child(f, "result_path")->text;
  dump_data_root = child(f, debug ? "debug_data_root" : "data_root")->text;
  dump_frames_per_file = (unsigned int)as_int(child(f, "frames_per_file"));
  dump_compress = as_bool(child(f, "compress"));
  dump_compress_level = as_int(child(f, "compress_level"));
  if (dump_frames_per_file == 0) throw ConfigError(1, "ERROR: frames_per_file must be positive");

  // init_probes, src/parameters.cpp:106-174
  for (const Element *pe : child(f, "probes")->children_named("probe")) {
    ProbeSpec pr;
    pr.component = attr(pe, "component");
    const char *sp = attr(pe, "specie", false);
    pr.specie = sp ? sp : "none";
    std::string type = attr(pe, "type");
    if (type == "frame" || type == "mpframe") {
      pr.type = type == "frame" ? 0 : 4;
      pr.r_start = std::atoi(attr(pe, "r_start"));
      pr.r_end = std::atoi(attr(pe, "r_end"));
      pr.z_start = std::atoi(attr(pe, "z_start"));
      pr.z_end = std::atoi(attr(pe, "z_end"));
    } else if (type == "col") {
      pr.type = 1;
      pr.r_start = 0;
      pr.r_end = n_grid_r;
      pr.z_start = std::atoi(attr(pe, "z"));
      pr.z_end = -1;
    } else if (type == "row") {
      pr.type = 2;
      pr.r_start = std::atoi(attr(pe, "r"));
      pr.r_end = -1;
      pr.z_start = 0;
      pr.z_end = n_grid_z;
    } else if (type == "dot") {
      pr.type = 3;
      pr.r_start = std::atoi(attr(pe, "r"));
      pr.r_end = -1;
      pr.z_start = std::atoi(attr(pe, "z"));
      pr.z_end = -1;
    } else {
      throw ConfigError(1, "ERROR! probe type ``" + type + "'' is not supported");
    }
    pr.schedule = std::atoi(attr(pe, "schedule"));
    if (pr.schedule <= 0) throw ConfigError(1, "ERROR! probe schedule must be positive");
    probes.push_back(pr);
  }

  // <device> (optional)
  if (const Element *d = root->first_child("device")) {
    if (const Element *e = d->first_child("precision")) {
      if (e->text == "fp64")
        precision = PDP3_PREC_FP64;
      else if (e->text == "fp32")
        precision = PDP3_PREC_FP32;
      else
        throw ConfigError(1, "ERROR! <precision> must be fp64 or fp32, not ``" + e->text + "''");
    }
    if (const Element *e = d->first_child("pusher")) {
      if (e->text == "boris")
        pusher = PDP3_PUSHER_BORIS;
      else if (e->text == "vay")
        pusher = PDP3_PUSHER_VAY;
      else
        throw ConfigError(1, "ERROR! <pusher> must be boris or vay, not ``" + e->text + "''");
    }
    if (const Element *e = d->first_child("sort_interval")) sort_interval = as_int(e);
  }
}

}  // namespace pdp3
