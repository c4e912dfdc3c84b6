// Host-side objects of the LJ + Ewald path, with the reference's class names,
// CASL commands and virtual interface, running on the caviar_amd C ABI.
//
// Reference interfaces mirrored here (paths under the reference tree):
//   Domain                      include/caviar/objects/domain.h, objects/domain.cpp:206-266
//   Atom_data (basic)           include/caviar/objects/atom_data.h, objects/atom_data.cpp:78-306
//   Neighborlist (basic)        include/caviar/objects/neighborlist.h:35-186,
//                               objects/neighborlist.cpp:66-88,193-251
//   Force_field                 include/caviar/objects/force_field.h:33-61
//   force_field::Lj_cell_list   objects/force_field/lj_cell_list.cpp:33-83
//   force_field::Lj             objects/force_field/lj.cpp:42-339
//   force_field::Electrostatic_short_range  objects/force_field/electrostatic_short_range.cpp:39-131
//   force_field::Electrostatic_ewald_r  objects/force_field/electrostatic_ewald_r.cpp:31-110
//   force_field::Electrostatic_ewald_k  objects/force_field/electrostatic_ewald_k.cpp:31-159
//   force_field::Electrostatic_ewald_slab_correction
//                               objects/force_field/electrostatic_ewald_slab_correction.cpp:30-131
//   Md_simulator (basic)        objects/md_simulator.cpp:58-233,382-617
//
// The particle arrays belong to Atom_data, as in the reference; here they live
// on the GPU inside one cav_ctx that Atom_data owns, and the host vectors below
// are mirrors refreshed by Atom_data::sync_from_device().  Force fields and the
// neighbour list of one Atom_data share its context.  Md_simulator steps the
// whole set in one fused call (cav_step); the per-object calls
// (Neighborlist::build, Force_field::calculate_acceleration, ...) serve callers
// that drive the objects themselves.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "caviar_amd.h"
#include "casl.hpp"

namespace caviar_amd {

struct Vector {
  double x = 0.0, y = 0.0, z = 0.0;
};

// Throws caviar_amd::Error with cav_last_error() when a C-ABI call fails.
void check(cav_ctx *ctx, int status, const char *what);

class Objects;

// What every object can reach, as the reference's Pointers base gives each
// object its object_container (include/caviar/utility/pointers.h).
struct Pointers {
  Objects *objects = nullptr;
};

class Domain : public Pointers {
 public:
  bool read(casl::Parser *parser);
  void generate();
  Vector lower_global, upper_global;
  int boundary_condition[3] = {0, 0, 0};
  bool generated = false;
};

class Atom_data : public Pointers {
 public:
  Atom_data();
  ~Atom_data();
  Atom_data(const Atom_data &) = delete;
  Atom_data &operator=(const Atom_data &) = delete;

  bool read(casl::Parser *parser);
  void verify_settings();
  bool add_atom(unsigned id, unsigned type, const Vector &pos, const Vector &vel);
  void add_xyz_data_file(casl::Parser *parser);
  // add_random_velocity SEED AMP (atom_data.cpp:597-615): libstdc++'s mt19937
  // and uniform_real_distribution, then the centre-of-mass velocity removed.
  void add_random_velocity(unsigned seed, double amplitude);
  // kinetic_energy() now; temperature() and pressure() as last finalized
  // (kinetic_energy.cpp:140-292).
  double kinetic_energy();
  double temperature();
  double pressure();
  void reset_owned_acceleration();
  void reset_virial();
  double virial();
  // Number of atom types: one past the largest type in use or with a mass.
  int num_types() const;

  // The device context holding this atom data (created on first use).
  cav_ctx *device();
  // Push box, types and atoms to the device (before setup).
  void upload();
  // Refresh the host mirrors of the owned arrays from the device.
  void sync_from_device();
  // Ghost positions and the owned index each ghost images.
  void ghosts(std::vector<Vector> &pos, std::vector<int> &src);

  struct Owned {
    std::vector<unsigned> id, type;
    std::vector<int> molecule_index;
    std::vector<Vector> position, velocity, acceleration;
  } atom_struct_owned;
  struct TypeParams {
    std::vector<double> mass, mass_inv, charge, radius;
  } atom_type_params;
  double ghost_cutoff = 0.0;
  double k_b = -1.0;               // atom_data.cpp:54
  int n_r_df = -1;                 // from the boundary conditions (atom_data.cpp:322-350)
  bool temperature_process = true; // atom_data.h:735
  bool pressure_process = false;
  Domain *domain = nullptr;
  int device_index = 0;
  int precision = CAV_FP64;

 private:
  cav_ctx *ctx_ = nullptr;
};

class Neighborlist : public Pointers {
 public:
  bool read(casl::Parser *parser);
  void verify_settings();
  // Neighborlist::init: push the settings into the context (before setup).
  void init();
  // Neighborlist::build(update): rebuild test + list; returns whether it rebuilt.
  bool build(bool update);
  // The reference's neighlist: for each owned i, the owned j > i and every
  // ghost (as n + g) within cutoff_extra.  Order within a row is the build's.
  std::vector<std::vector<unsigned>> neighlist();

  double cutoff = 0.0;
  double cutoff_extra_coef = 1.12347;  // neighborlist.h:134
  double dt = 0.0;                     // accepted, unused (as in the reference)
  bool make_verlet_list_from_cell_list = false;
  bool build_cell_list = false;
  bool all_atom_test = false;
  bool rebuild_test = false;
  Atom_data *atom_data = nullptr;
  Domain *domain = nullptr;
};

class Force_field : public Pointers {
 public:
  virtual ~Force_field() = default;
  virtual bool read(casl::Parser *parser) = 0;
  virtual void verify_settings() = 0;
  // Push this field's settings into the context (before setup).
  virtual void init() = 0;
  // Adds this field's accelerations (and virial) into Atom_data.
  virtual void calculate_acceleration();
  virtual double energy();
  virtual double potential(const Vector &r);
  virtual Vector field(const Vector &r);
  // Which cav_force_field bit this object is.
  virtual int mask() const = 0;
  virtual const char *type_name() const = 0;

  double cutoff = 0.0;
  Atom_data *atom_data = nullptr;
  Domain *domain = nullptr;
  Neighborlist *neighborlist = nullptr;
};

namespace force_field {

class Lj_cell_list : public Force_field {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  double energy() override;
  int mask() const override { return CAV_FF_LJ; }
  const char *type_name() const override { return "lj_cell_list"; }
  // Row-ragged like the reference's std::vector<std::vector<double>>; entries
  // never set read as 0.
  std::vector<std::vector<double>> epsilon, sigma;
};

// force_field lj: the LJ over the Verlet list (lj.cpp:42-483).  The matrices
// are kept as read (row-ragged, unset entries 0; lambda_s / lambda_e unset 1);
// the C ABI resolves them as Lj::verify_settings does (by-atom mixing, WCA
// cutoffs, lambda scaling).
class Lj : public Force_field {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  int mask() const override { return CAV_FF_LJ_LIST; }
  const char *type_name() const override { return "lj"; }
  std::vector<std::vector<double>> epsilon, sigma, cutoff_list, lambda_s, lambda_e;
  std::vector<double> epsilon_atom, sigma_atom;
  bool input_by_array = false, input_by_atom = false;
  bool cutoff_list_activated = false, wca = false;
  bool lambda_s_is_set = false, lambda_e_is_set = false;
  bool ignore_intra_molecule = false;
  bool jump_fix = false, monitor_jump = false, make_off_diagonal_vectors = false;
  double jump_tol = 1e-6;  // lj.cpp:33
};

// force_field electrostatic_short_range: shifted Coulomb within the cutoff
// over the Verlet list plus a uniform external field
// (electrostatic_short_range.cpp:39-258).
class Electrostatic_short_range : public Force_field {
 public:
  Electrostatic_short_range() { cutoff = -1.0; }  // electrostatic_short_range.cpp:34
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  int mask() const override { return CAV_FF_SHORT_RANGE; }
  const char *type_name() const override { return "electrostatic_short_range"; }
  double k_electrostatic = 1.0, beta = 1.0;
  Vector external_field;
  bool lambda_is_set = false;
  std::vector<std::vector<double>> lambda;  // missing entries read as 1
};

class Electrostatic_ewald_r : public Force_field {
 public:
  Electrostatic_ewald_r() { cutoff = 0.0; }
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  double energy() override;
  double potential(const Vector &r) override;
  Vector field(const Vector &r) override;
  int mask() const override { return CAV_FF_EWALD_R; }
  const char *type_name() const override { return "electrostatic_ewald_r"; }
  double k_electrostatic = 1.0;
  double alpha = 1.0;
  bool lambda_is_set = false;
  std::vector<std::vector<double>> lambda;  // missing entries read as 1
};

class Electrostatic_ewald_k : public Force_field {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  double energy() override;
  double self_energy();
  double potential(const Vector &r) override;
  Vector field(const Vector &r) override;
  int mask() const override { return CAV_FF_EWALD_K; }
  const char *type_name() const override { return "electrostatic_ewald_k"; }
  // S(k) in the reference's k order (potential_k_coef_cmplx), re/im interleaved.
  std::vector<double> potential_k_coef();

  double k_electrostatic = 1.0;
  double alpha = -1.0;  // no default in the reference: must be set
  int kx_max = 1, ky_max = 1, kz_max = 1;
  bool dipole = false;
  double epsilon_dipole = 1.0;
  // set by slab_normal_axis or slab_lz; axis z and thickness 1 by default
  // (electrostatic_ewald_k.cpp:31-45)
  bool slab_geometry = false;
  int slab_normal_axis = 2;
  double slab_lz = 1.0;
};

// force_field electrostatic_ewald_slab_correction: the ELC term beside Ewald
// in slab geometry (electrostatic_ewald_slab_correction.cpp:30-181).
class Electrostatic_ewald_slab_correction : public Force_field {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  double energy() override;
  double potential(const Vector &r) override;
  Vector field(const Vector &r) override;
  int mask() const override { return CAV_FF_ELC; }
  const char *type_name() const override { return "electrostatic_ewald_slab_correction"; }
  // electrostatic_ewald_slab_correction.cpp:31-42
  int slab_normal_axis = 2;
  double k_electrostatic = 1.0;
  int kx_max = 1, ky_max = 1;
  double lz = 1.0;
};

}  // namespace force_field

// Velocity-rescaling constraints of the velocity-Verlet step
// (md_simulator.cpp:619-633), the reference's constraint objects.
class Constraint : public Pointers {
 public:
  virtual ~Constraint() = default;
  virtual bool read(casl::Parser *parser) = 0;
  virtual void verify_settings() = 0;
  // Push the settings into the context (before setup).
  virtual void init() = 0;
  virtual const char *type_name() const = 0;
  Atom_data *atom_data = nullptr;
};

namespace constraint {

// constraint berendsen (constraint/berendsen.cpp:25-146; berendsen.h:50-60)
class Berendsen : public Constraint {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  const char *type_name() const override { return "berendsen"; }
  double temperature = 0.0, coupling = -1.0, dt = -1.0;
  int step = 1;
  double lambda_min = 0.8, lambda_max = 1.25;
};

// constraint nve (constraint/nve.cpp:25-178)
class Nve : public Constraint {
 public:
  bool read(casl::Parser *parser) override;
  void verify_settings() override;
  void init() override;
  const char *type_name() const override { return "nve"; }
  double energy_per_dof = -1.0, energy_tot = -1.0, temperature = -1.0, kb = -1.0, kbt = -1.0;
  int step = 1;
};

}  // namespace constraint

namespace writer {

// writer atom_data, serial output (writer/atom_data.cpp, writer/atom_data/
// dump_xyz_serial.cpp:33-71, dump_energy.cpp): xyz frames, kinetic energy,
// temperature and pressure, every *_step steps, as "<name>.xyz" / ".txt" with
// the stream's default formatting; xyz_precision (an addition, default the
// stream's 6 digits) writes the xyz frames at up to 17 significant digits.
class Atom_data : public Pointers {
 public:
  ~Atom_data();
  bool read(casl::Parser *parser);
  void write(long step, double time);
  // whether write() has anything to write at this step
  bool wants(long step) const;
  caviar_amd::Atom_data *atom_data = nullptr;
  long xyz_step = 100, energy_step = 100, temperature_step = 100, pressure_step = 100;
  bool output_xyz = false, output_energy = false, output_temperature = false, output_pressure = false;
  std::string file_name_xyz = "o_xyz", file_name_energy = "o_energy",
              file_name_temperature = "o_temperature", file_name_pressure = "o_pressure";
  bool xyz_output_id = false, xyz_output_velocity = false, xyz_output_acceleration = false;
  int xyz_precision = 6;
  std::string dir;  // where relative file names go (the script's folder is not used: the run's)

 private:
  struct Files;
  Files *files_ = nullptr;
  void open();
};

}  // namespace writer

class Md_simulator : public Pointers {
 public:
  bool read(casl::Parser *parser);
  void verify_settings();
  // Md_simulator::initialize + setup (md_simulator.cpp:303,504-549).
  void setup();
  // One velocity-Verlet step (md_simulator.cpp:382-617).
  void step();
  // Md_simulator::run: steps from initial_step to final_step.
  void run();

  Atom_data *atom_data = nullptr;
  std::vector<Force_field *> force_field;
  std::vector<Neighborlist *> neighborlist;
  std::vector<Constraint *> constraint;
  std::vector<writer::Atom_data *> writer;
  double dt = 0.001;
  double time = 0.0;  // dt * initial_step at setup, then += dt per step
  long initial_step = 0, final_step = 0, current_step = 0;
  bool initialized = false;
  // Read scripts without the GPU: run / step only count their steps.
  static bool dry_run;
  long steps_requested = 0;

 private:
  void advance(long nsteps);
};

// Named objects of a script (the reference's Object_container,
// interpreter/object_container.cpp); owns them.
class Objects {
 public:
  std::map<std::string, std::unique_ptr<Domain>> domain;
  std::map<std::string, std::unique_ptr<Atom_data>> atom_data;
  std::map<std::string, std::unique_ptr<Neighborlist>> neighborlist;
  std::map<std::string, std::unique_ptr<Force_field>> force_field;
  std::map<std::string, std::unique_ptr<Md_simulator>> md_simulator;
  std::map<std::string, std::unique_ptr<Constraint>> constraint;
  std::map<std::string, std::unique_ptr<writer::Atom_data>> writer;
  std::vector<std::string> order;  // creation order of every name

  bool has(const std::string &name) const;
  Domain *find_domain(const std::string &name, casl::Parser *p);
  Atom_data *find_atom_data(const std::string &name, casl::Parser *p);
  Neighborlist *find_neighborlist(const std::string &name, casl::Parser *p);
  Force_field *find_force_field(const std::string &name, casl::Parser *p);
  Constraint *find_constraint(const std::string &name, casl::Parser *p);
  writer::Atom_data *find_writer(const std::string &name, casl::Parser *p);
};

}  // namespace caviar_amd
