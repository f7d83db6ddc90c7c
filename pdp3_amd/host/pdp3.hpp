// C++ class surface of the reference's r-z PIC loop (cosmonaut-ok/pdp3),
// backed by the MI355X C ABI (include/pdp3_hip.h).
//
// The classes keep the reference's names, constructors and per-step methods
// so that the reference's driver code ports line for line:
//
//   Geometry geom(r_size, z_size, n_grid_r, n_grid_z);   // include/geometry.h:20-24
//   Time time(ct, st, rt, et, dt);                       // include/pdp3Time.h:13
//   Device dev(geom, time, PDP3_PUSHER_BORIS, PDP3_PREC_FP64);   // the one new line
//   geom.set_pml(l1, l2, l3, s1, s2);
//   EField efield(&geom); HField hfield(&geom);          // include/eField.h, hField.h
//   Current current(&geom); ChargeDensity rho_bunch(&geom);
//   ParticlesList plist; plist.part_list.push_back(new Particles(...));
//   hfield.calc_field(&efield, &time); current.reset_j(); plist.step_v(&efield, &hfield, &time); ...
//
// Field, current and particle data live in HBM; the public arrays of the
// reference (double** field_r, pos, vel, ...) become host mirrors refreshed
// by download() and pushed back by upload().  Errors are reported as
// pdp3::Error exceptions carrying the C ABI status code.  Time is owned by
// the device: Time::current_time is pushed before and read back after every
// call that depends on it.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pdp3_hip.h"

namespace pdp3 {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class Device;

// Time, include/pdp3Time.h:6-14
class Time {
 public:
  double current_time, start_time, relaxation_time, end_time, delta_t;
  Time(double ct, double st, double rt, double et, double dt)
      : current_time(ct), start_time(st), relaxation_time(rt), end_time(et), delta_t(dt) {}
};

// Geometry, include/geometry.h:5-28.  epsilon / sigma are host mirrors
// (download_sigma refreshes sigma after set_pml).
class Geometry {
 public:
  double r_size, z_size;
  int n_grid_r, n_grid_z;
  double dr, dz;
  std::vector<double> epsilon, sigma;  // n_grid_r * n_grid_z, row-major
  Device *device = nullptr;            // set by Device

  // src/geometry.cpp:8-31: dr = r_size / (n_grid_r - 1), dz likewise
  Geometry(double rs, double zs, int ngr, int ngz);
  double get_dr() const { return dr; }
  double get_dz() const { return dz; }
  // epsilon = 1 everywhere, src/geometry.cpp:33-39 (the device starts so)
  void set_epsilon();
  // src/geometry.cpp:56-107
  void set_pml(double comparative_l_1, double comparative_l_2, double comparative_l_3, double sigma1,
               double sigma2);
  void download_sigma();
};

// owns the GPU context of one simulation (geometry + time + pusher + form)
class Device {
 public:
  Device(Geometry &geom, Time &time, int pusher = PDP3_PUSHER_BORIS, int precision = PDP3_PREC_FP64);
  ~Device();
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;
  pdp3_ctx *ctx() const { return ctx_; }
  // throws Error for a non-zero status
  void check(int status) const;
  void push_time(const Time &t) const;
  void pull_time(Time &t) const;
  // fused loop body (init_beam, H, J, push+deposit, E, bunch rho, t += dt)
  void step(Time &t, int nsteps = 1);
  Geometry &geom;
  Time &time;
  int precision;

 private:
  pdp3_ctx *ctx_ = nullptr;
};

// a grid component with its host mirror
class GridArray {
 public:
  GridArray(Geometry *g, int comp) : geom_(g), comp_(comp) {}
  std::vector<double> &download();
  void upload();
  std::vector<double> host;  // mirror, row-major n_grid_r x n_grid_z
  double at(int i, int k) const { return host[(size_t)i * (size_t)geom_->n_grid_z + (size_t)k]; }
  double &at(int i, int k) { return host[(size_t)i * (size_t)geom_->n_grid_z + (size_t)k]; }

 private:
  Geometry *geom_;
  int comp_;
};

class Current;
class HField;

// Field, include/field.h:5-26
class Field {
 public:
  Geometry *geom1;
  GridArray field_r, field_phi, field_z;
  Field(Geometry *g, int cr, int cphi, int cz);
};

// EField, include/eField.h:17-43
class EField : public Field {
 public:
  explicit EField(Geometry *geom1_t);
  void calc_field(HField *h_field1, Time *time1, Current *current1);
  void set_homogeneous_efield(double E_r, double E_phi, double E_z);
  void boundary_conditions();
  // src/eField.cpp:204-316 (the reference returns a new double[3])
  std::vector<double> get_field(double radius, double longitude);
};

// HField, include/hField.h:11-33
class HField : public Field {
 public:
  GridArray field_r_half_time, field_phi_half_time, field_z_half_time;
  explicit HField(Geometry *geom1_t);
  void calc_field(EField *e_field1, Time *time1);
  void set_homogeneous_h(double H_r, double H_phi, double H_z);
  // src/hField.cpp:159-267, on the half-time field
  std::vector<double> get_field(double radius, double longitude);
};

// Current, include/current.h:6-29
class Current {
 public:
  Geometry *geom1;
  GridArray j_r, j_phi, j_z;
  explicit Current(Geometry *geom1_t);
  void reset_j();
  std::vector<double> &get_j_r() { return j_r.download(); }
  std::vector<double> &get_j_phi() { return j_phi.download(); }
  std::vector<double> &get_j_z() { return j_z.download(); }
};

// ChargeDensity, include/chargeDensity.h:5-18.  The device holds one
// density grid, the bunch density (LoadInitParam::c_rho_bunch).
class ChargeDensity {
 public:
  Geometry *geom1;
  GridArray rho;
  explicit ChargeDensity(Geometry *geom1_t);
  void reset_rho();
  std::vector<double> &get_rho() { return rho.download(); }
};

// Particles, include/particles.h:34-165, with SoA host mirrors
class Particles {
 public:
  std::string name;
  double charge;  // per real particle, C (charge_e * 1.6e-19)
  double mass;    // per real particle, kg (mass_e * 9.1e-31)
  int charge_e;
  double init_const_mass;  // mass in electron masses
  unsigned int number;
  Geometry *geom1;
  std::vector<double> pos_r, pos_z, vel_r, vel_phi, vel_z, weight;
  std::vector<uint8_t> is_alive;

  // src/particles.cpp:23-55: charge and mass in electron units
  Particles(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom);
  virtual ~Particles() = default;
  // TESTMODE loaders, src/particles.cpp:228-331 (computed on the host,
  // uploaded to the device)
  void load_cylindrical_spatial_distribution(double n1, double n2, double left_plasma_boundary);
  void velocity_distribution(double tempr_ev);
  // src/particles.cpp:1296-1316
  void step_v(EField *e_fld, HField *h_fld, Time *t);
  // src/particles.cpp:1367-1373
  void full_current_distribution(Current *current, Time *t);
  // src/particles.cpp:101-226 (into the device's bunch density)
  void charge_weighting(ChargeDensity *ro1);
  // host mirrors <-> device (upload needs all mirrors sized `number`)
  void download();
  void upload();
  int index() const { return index_; }
  bool is_bunch() const { return bunch_; }

 protected:
  explicit Particles(Geometry *geom) : geom1(geom) {}
  int index_ = -1;
  bool bunch_ = false;
};

// Bunch, include/bunch.h:4-36.  The bunch joins the device's species list at
// the first step with t >= bunch_number * (duration + hole_duration)
// (src/loadInitParam.cpp:214-241).
class Bunch : public Particles {
 public:
  double duration, hole_duration, density, velocity, radius;
  int bunch_number;
  // the reference's constructor with its TESTMODE draws, src/bunch.cpp:6-80
  Bunch(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom,
        double b_duration, double b_radius, double b_density, double b_init_velocity, int b_number,
        double h_duration);
  // the same with the caller's radial positions, weights and injection draws
  Bunch(const std::string &p_name, double p_charge, double p_mass, int p_number, Geometry *geom,
        double b_duration, double b_radius, double b_density, double b_init_velocity, int b_number,
        double h_duration, const std::vector<double> &r, const std::vector<double> &w,
        const std::vector<double> &inject_u);
  // src/bunch.cpp:87-124
  void bunch_inject(Time *time);

 private:
  void add(const std::vector<double> &r, const std::vector<double> &w, const std::vector<double> &u);
};

// ParticlesList, include/particlesList.h:9-26
class ParticlesList {
 public:
  std::vector<Particles *> part_list;
  void step_v(EField *e_fld, HField *h_fld, Time *t);
  void full_current_distribution(Current *current, Time *t);
};

// BoundaryMaxwellConditions, include/boundaryMaxwellConditions.h:6-16
class BoundaryMaxwellConditions {
 public:
  EField *e_fld;
  explicit BoundaryMaxwellConditions(EField *e_fld_t) : e_fld(e_fld_t) {}
  void specify_initial_field(Geometry *cyl_geom, double E_fi_upper, double E_fi_left, double E_fi_right);
};

// TemperatureClassic, include/temperatureClassic.h: the per-cell sums and the
// normalisation run on the device over the resident particle state
// (pdp3_temperature); t is the host mirror, n_grid_r x n_grid_z row-major
// (row n_grid_r-1 and column n_grid_z-1 are 0, empty cells NaN).
class TemperatureClassic {
 public:
  Geometry *geom;
  std::vector<double> t;
  explicit TemperatureClassic(Geometry *g);
  void reset();
  // src/temperatureClassic.cpp:57-103 and :127-158 in one device pass
  void calc_t(Particles *prtls);
  // the device's calc_t already normalised: kept for the reference's call sequence
  void normalize(Particles *prtls) {}
};

// ParticlesDensity, include/particlesDensity.h: calc_density on the device
// (pdp3_density); density is the host mirror, n_grid_r x n_grid_z row-major.
class ParticlesDensity {
 public:
  Geometry *geom;
  std::vector<double> density;
  explicit ParticlesDensity(Geometry *g);
  void reset();
  // src/particlesDensity.cpp:40-225
  void calc_density(Particles *prtls);
};

}  // namespace pdp3
