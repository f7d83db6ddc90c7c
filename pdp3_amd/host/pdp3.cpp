#include "pdp3.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "loaders.hpp"

namespace pdp3 {

namespace {
constexpr double kLightSpeed2 = 3.0e8 * 3.0e8;  // include/constant.h:18, pow(c, 2)

Device &device_of(Geometry *g) {
  if (!g || !g->device) throw Error(PDP3_ERR_STATE, "no pdp3::Device is attached to this Geometry");
  return *g->device;
}
}  // namespace

// ---------------------------------------------------------------- Geometry

Geometry::Geometry(double rs, double zs, int ngr, int ngz)
    : r_size(rs), z_size(zs), n_grid_r(ngr), n_grid_z(ngz) {
  if (ngr < 2 || ngz < 2) throw Error(PDP3_ERR_BAD_ARG, "Geometry needs at least 2 nodes per direction");
  dr = r_size / (n_grid_r - 1);
  dz = z_size / (n_grid_z - 1);
  epsilon.assign((size_t)ngr * (size_t)ngz, 1.0);
  sigma.assign((size_t)ngr * (size_t)ngz, 0.0);
}

void Geometry::set_epsilon() {
  std::fill(epsilon.begin(), epsilon.end(), 1.0);
  if (device) device->check(pdp3_grid_upload(device->ctx(), PDP3_EPSILON, epsilon.data()));
}

void Geometry::set_pml(double l1, double l2, double l3, double sigma1, double sigma2) {
  Device &d = device_of(this);
  d.check(pdp3_geometry_set_pml(d.ctx(), l1, l2, l3, sigma1, sigma2));
  download_sigma();
}

void Geometry::download_sigma() {
  Device &d = device_of(this);
  d.check(pdp3_grid_download(d.ctx(), PDP3_SIGMA, sigma.data()));
}

// ------------------------------------------------------------------ Device

Device::Device(Geometry &g, Time &t, int pusher, int prec) : geom(g), time(t), precision(prec) {
  int st = pdp3_ctx_create(g.n_grid_r, g.n_grid_z, g.r_size, g.z_size, t.current_time, t.start_time,
                           t.relaxation_time, t.end_time, t.delta_t, pusher, prec, &ctx_);
  if (st) throw Error(st, std::string("pdp3_ctx_create failed (status ") + std::to_string(st) + ")");
  g.device = this;
}

Device::~Device() {
  if (geom.device == this) geom.device = nullptr;
  pdp3_ctx_destroy(ctx_);
}

void Device::check(int status) const {
  if (status != PDP3_OK) {
    const char *msg = pdp3_last_error(ctx_);
    throw Error(status, std::string("pdp3: ") + (msg && *msg ? msg : "error") + " (status " +
                            std::to_string(status) + ")");
  }
}

void Device::push_time(const Time &t) const { check(pdp3_time_set(ctx_, t.current_time)); }

void Device::pull_time(Time &t) const {
  double ct = 0, dt = 0;
  check(pdp3_time_get(ctx_, &ct, &dt));
  t.current_time = ct;
}

void Device::step(Time &t, int nsteps) {
  push_time(t);
  check(pdp3_step(ctx_, nsteps));
  pull_time(t);
}

// --------------------------------------------------------------- GridArray

std::vector<double> &GridArray::download() {
  Device &d = device_of(geom_);
  host.resize((size_t)geom_->n_grid_r * (size_t)geom_->n_grid_z);
  d.check(pdp3_grid_download(d.ctx(), comp_, host.data()));
  return host;
}

void GridArray::upload() {
  Device &d = device_of(geom_);
  if (host.size() != (size_t)geom_->n_grid_r * (size_t)geom_->n_grid_z)
    throw Error(PDP3_ERR_BAD_ARG, "host mirror has the wrong size");
  d.check(pdp3_grid_upload(d.ctx(), comp_, host.data()));
}

// ------------------------------------------------------------------ fields

Field::Field(Geometry *g, int cr, int cphi, int cz) : geom1(g), field_r(g, cr), field_phi(g, cphi), field_z(g, cz) {
  device_of(g);
}

EField::EField(Geometry *g) : Field(g, PDP3_E_R, PDP3_E_PHI, PDP3_E_Z) {}

void EField::calc_field(HField *, Time *t, Current *) {
  Device &d = device_of(geom1);
  d.push_time(*t);
  d.check(pdp3_efield_calc(d.ctx()));
}

void EField::set_homogeneous_efield(double er, double ephi, double ez) {
  Device &d = device_of(geom1);
  d.check(pdp3_efield_set_homogeneous(d.ctx(), er, ephi, ez));
}

void EField::boundary_conditions() {
  Device &d = device_of(geom1);
  d.check(pdp3_efield_boundary_conditions(d.ctx()));
}

std::vector<double> EField::get_field(double radius, double longitude) {
  Device &d = device_of(geom1);
  std::vector<double> out(3);
  d.check(pdp3_efield_get_field(d.ctx(), &radius, &longitude, 1, out.data()));
  return out;
}

HField::HField(Geometry *g)
    : Field(g, PDP3_H_R, PDP3_H_PHI, PDP3_H_Z),
      field_r_half_time(g, PDP3_HH_R),
      field_phi_half_time(g, PDP3_HH_PHI),
      field_z_half_time(g, PDP3_HH_Z) {}

void HField::calc_field(EField *, Time *t) {
  Device &d = device_of(geom1);
  d.push_time(*t);
  d.check(pdp3_hfield_calc(d.ctx()));
}

void HField::set_homogeneous_h(double hr, double hphi, double hz) {
  Device &d = device_of(geom1);
  d.check(pdp3_hfield_set_homogeneous(d.ctx(), hr, hphi, hz));
}

std::vector<double> HField::get_field(double radius, double longitude) {
  Device &d = device_of(geom1);
  std::vector<double> out(3);
  d.check(pdp3_hfield_get_field(d.ctx(), &radius, &longitude, 1, out.data()));
  return out;
}

Current::Current(Geometry *g) : geom1(g), j_r(g, PDP3_J_R), j_phi(g, PDP3_J_PHI), j_z(g, PDP3_J_Z) { device_of(g); }

void Current::reset_j() {
  Device &d = device_of(geom1);
  d.check(pdp3_current_reset(d.ctx()));
}

ChargeDensity::ChargeDensity(Geometry *g) : geom1(g), rho(g, PDP3_RHO_BUNCH) { device_of(g); }

void ChargeDensity::reset_rho() {
  Device &d = device_of(geom1);
  d.check(pdp3_rho_bunch_reset(d.ctx()));
}

void BoundaryMaxwellConditions::specify_initial_field(Geometry *g, double upper, double left, double right) {
  Device &d = device_of(g);
  d.check(pdp3_boundary_specify_initial_field(d.ctx(), upper, left, right));
}

// --------------------------------------------------------------- particles

Particles::Particles(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom)
    : name(p_name), geom1(geom) {
  Device &d = device_of(geom);
  charge_e = (int)p_charge;
  init_const_mass = p_mass;
  charge = p_charge * 1.6e-19;  // include/constant.h:16
  mass = p_mass * 9.1e-31;      // include/constant.h:14
  number = (unsigned int)p_number;
  d.check(pdp3_species_add(d.ctx(), charge_e, p_mass, p_number, &index_));
  // every particle starts alive, src/particles.cpp:47-53
  is_alive.assign(number, 1);
  d.check(pdp3_species_upload(d.ctx(), index_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              is_alive.data()));
}

void Particles::load_cylindrical_spatial_distribution(double n1, double n2, double left_plasma_boundary) {
  const int n = (int)number;
  pos_r.resize(n);
  pos_z.resize(n);
  weight.resize(n);
  loaders::load_cylindrical(n, geom1->r_size, geom1->z_size, geom1->dr, geom1->dz, n1, n2, left_plasma_boundary,
                            pos_r.data(), pos_z.data(), weight.data());
  Device &d = device_of(geom1);
  d.check(pdp3_species_upload(d.ctx(), index_, pos_r.data(), pos_z.data(), nullptr, nullptr, nullptr, weight.data(),
                              nullptr));
}

void Particles::velocity_distribution(double tempr_ev) {
  const int n = (int)number;
  vel_r.resize(n);
  vel_phi.resize(n);
  vel_z.resize(n);
  loaders::velocity_distribution(n, mass, tempr_ev, vel_r.data(), vel_phi.data(), vel_z.data());
  Device &d = device_of(geom1);
  d.check(pdp3_species_upload(d.ctx(), index_, nullptr, nullptr, vel_r.data(), vel_phi.data(), vel_z.data(), nullptr,
                              nullptr));
}

void Particles::step_v(EField *, HField *, Time *t) {
  Device &d = device_of(geom1);
  d.push_time(*t);
  d.check(pdp3_particles_step_v(d.ctx(), index_));
}

void Particles::full_current_distribution(Current *, Time *t) {
  Device &d = device_of(geom1);
  d.push_time(*t);
  d.check(pdp3_particles_full_current_distribution(d.ctx(), index_));
}

void Particles::charge_weighting(ChargeDensity *) {
  Device &d = device_of(geom1);
  d.check(pdp3_particles_charge_weighting(d.ctx(), index_));
}

void Particles::download() {
  Device &d = device_of(geom1);
  int n = 0;
  d.check(pdp3_species_count(d.ctx(), index_, &n));
  number = (unsigned int)n;
  pos_r.resize(n);
  pos_z.resize(n);
  vel_r.resize(n);
  vel_phi.resize(n);
  vel_z.resize(n);
  weight.resize(n);
  is_alive.resize(n);
  d.check(pdp3_species_download(d.ctx(), index_, pos_r.data(), pos_z.data(), vel_r.data(), vel_phi.data(),
                                vel_z.data(), weight.data(), is_alive.data()));
}

void Particles::upload() {
  const size_t n = number;
  if (pos_r.size() != n || pos_z.size() != n || vel_r.size() != n || vel_phi.size() != n || vel_z.size() != n ||
      weight.size() != n || is_alive.size() != n)
    throw Error(PDP3_ERR_BAD_ARG, "Particles::upload: every host mirror must hold `number` entries");
  Device &d = device_of(geom1);
  d.check(pdp3_species_upload(d.ctx(), index_, pos_r.data(), pos_z.data(), vel_r.data(), vel_phi.data(),
                              vel_z.data(), weight.data(), is_alive.data()));
}

Bunch::Bunch(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom,
             double b_duration, double b_radius, double b_density, double b_init_velocity, int b_number,
             double h_duration)
    : Particles(geom),
      duration(b_duration),
      hole_duration(h_duration),
      density(b_density),
      velocity(b_init_velocity),
      radius(b_radius),
      bunch_number(b_number) {
  name = p_name;
  charge_e = (int)p_charge;
  init_const_mass = p_mass;
  charge = p_charge * 1.6e-19;
  mass = p_mass * 9.1e-31;
  number = (unsigned int)p_number;
  std::vector<double> r(number), w(number), u(number);
  loaders::bunch_positions(p_number, radius, geom->dr, geom->dz, duration, velocity, density, r.data(), w.data());
  loaders::bunch_inject_draws(p_number, u.data());
  add(r, w, u);
}

Bunch::Bunch(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom,
             double b_duration, double b_radius, double b_density, double b_init_velocity, int b_number,
             double h_duration, const std::vector<double> &r, const std::vector<double> &w,
             const std::vector<double> &inject_u)
    : Particles(geom),
      duration(b_duration),
      hole_duration(h_duration),
      density(b_density),
      velocity(b_init_velocity),
      radius(b_radius),
      bunch_number(b_number) {
  name = p_name;
  charge_e = (int)p_charge;
  init_const_mass = p_mass;
  charge = p_charge * 1.6e-19;
  mass = p_mass * 9.1e-31;
  number = (unsigned int)p_number;
  if (r.size() != number || w.size() != number || inject_u.size() != number)
    throw Error(PDP3_ERR_BAD_ARG, "Bunch: r, w and inject_u must hold p_number entries");
  add(r, w, inject_u);
}

void Bunch::add(const std::vector<double> &r, const std::vector<double> &w, const std::vector<double> &u) {
  bunch_ = true;
  pos_r = r;
  weight = w;
  Device &d = device_of(geom1);
  d.check(pdp3_bunch_add(d.ctx(), charge_e, init_const_mass, (int)number, r.data(), w.data(), u.data(), duration,
                         hole_duration, velocity, bunch_number, &index_));
}

void Bunch::bunch_inject(Time *t) {
  Device &d = device_of(geom1);
  d.push_time(*t);
  d.check(pdp3_bunch_inject(d.ctx(), index_));
}

void ParticlesList::step_v(EField *e, HField *h, Time *t) {
  for (Particles *p : part_list) p->step_v(e, h, t);
}

void ParticlesList::full_current_distribution(Current *c, Time *t) {
  for (Particles *p : part_list) p->full_current_distribution(c, t);
}

// ------------------------------------------------------------- diagnostics

TemperatureClassic::TemperatureClassic(Geometry *g) : geom(g), t((size_t)g->n_grid_r * (size_t)g->n_grid_z, 0.0) {}

void TemperatureClassic::reset() { std::fill(t.begin(), t.end(), 0.0); }

void TemperatureClassic::calc_t(Particles *prtls) {
  Device &d = device_of(geom);
  d.check(pdp3_temperature(d.ctx(), prtls->index(), t.data()));
}

ParticlesDensity::ParticlesDensity(Geometry *g)
    : geom(g), density((size_t)g->n_grid_r * (size_t)g->n_grid_z, 0.0) {}

void ParticlesDensity::reset() { std::fill(density.begin(), density.end(), 0.0); }

void ParticlesDensity::calc_density(Particles *prtls) {
  Device &d = device_of(geom);
  d.check(pdp3_density(d.ctx(), prtls->index(), density.data()));
}

}  // namespace pdp3
