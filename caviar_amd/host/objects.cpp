#include "objects.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <random>

namespace caviar_amd {

void check(cav_ctx *ctx, int status, const char *what)
{
  if (status == CAV_OK) return;
  std::string msg = std::string(what) + ": ";
  msg += ctx ? cav_last_error(ctx) : "invalid context";
  throw Error(msg);
}

// ----------------------------------------------------------------- Objects --

bool Objects::has(const std::string &name) const
{
  return std::find(order.begin(), order.end(), name) != order.end();
}

template <typename M>
static typename M::mapped_type::element_type *find_in(M &m, const std::string &name, const char *kind,
                                                      casl::Parser *p)
{
  auto it = m.find(name);
  if (it == m.end()) {
    const std::string msg = std::string("Undefined ") + kind + " NAME: " + name;
    if (p) p->fail(msg);
    throw Error(msg);
  }
  return it->second.get();
}

Domain *Objects::find_domain(const std::string &n, casl::Parser *p) { return find_in(domain, n, "domain", p); }
Atom_data *Objects::find_atom_data(const std::string &n, casl::Parser *p)
{
  return find_in(atom_data, n, "atom_data", p);
}
Neighborlist *Objects::find_neighborlist(const std::string &n, casl::Parser *p)
{
  return find_in(neighborlist, n, "neighborlist", p);
}
Force_field *Objects::find_force_field(const std::string &n, casl::Parser *p)
{
  return find_in(force_field, n, "force_field", p);
}
Constraint *Objects::find_constraint(const std::string &n, casl::Parser *p)
{
  return find_in(constraint, n, "constraint", p);
}
writer::Atom_data *Objects::find_writer(const std::string &n, casl::Parser *p)
{
  return find_in(writer, n, "writer", p);
}

// Reads "keyword value ..." pairs until the end of the line, as every
// reference read() does with GET_A_TOKEN_FOR_CREATION.  `handle` returns false
// for an unknown keyword.  A handler that consumes the rest of the line itself
// sets `done` (the reference's read() returning early).
template <typename F>
static bool read_commands(casl::Parser *p, F handle)
{
  bool done = false;
  while (!done) {
    casl::Token t = p->value();
    if (t.kind == casl::Kind::eol) return true;
    if (t.kind == casl::Kind::eof) return false;
    if (t.kind != casl::Kind::identifier) p->fail("a command name was expected");
    if (!handle(t.text, done)) p->fail("Undefined variable or command: " + t.text);
  }
  return true;
}

// One element of a row-ragged matrix, GET_A_STDVECTOR_STDVECTOR_REAL_ELEMENT
// (interpreter/object_handler/preprocessors.h:568-590): rows grow on demand and
// new entries take `fill`.
static double read_matrix_element(casl::Parser *p, std::vector<std::vector<double>> &m, double fill,
                                  const char *what)
{
  const long i = p->integer(std::string(what) + " row");
  const long j = p->integer(std::string(what) + " column");
  const double v = p->real(what);
  if (i < 0 || j < 0) p->fail(std::string(what) + ": negative type index");
  if ((long)m.size() < i + 1) m.resize(i + 1);
  if ((long)m[i].size() < j + 1) m[i].resize(j + 1, fill);
  m[i][j] = v;
  return v;
}

// Dense nt x nt row-major copy of a ragged matrix; entries never set take `fill`.
static std::vector<double> dense(const std::vector<std::vector<double>> &m, int nt, double fill)
{
  std::vector<double> out((size_t)nt * nt, fill);
  for (int i = 0; i < nt && i < (int)m.size(); ++i)
    for (int j = 0; j < nt && j < (int)m[i].size(); ++j) out[(size_t)i * nt + j] = m[i][j];
  return out;
}

// ------------------------------------------------------------------ Domain --

bool Domain::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "lower_global.x" || t == "xmin")
      lower_global.x = p->real("xmin");
    else if (t == "upper_global.x" || t == "xmax")
      upper_global.x = p->real("xmax");
    else if (t == "lower_global.y" || t == "ymin")
      lower_global.y = p->real("ymin");
    else if (t == "upper_global.y" || t == "ymax")
      upper_global.y = p->real("ymax");
    else if (t == "lower_global.z" || t == "zmin")
      lower_global.z = p->real("zmin");
    else if (t == "upper_global.z" || t == "zmax")
      upper_global.z = p->real("zmax");
    else if (t == "boundary_condition" || t == "bc") {
      for (int d = 0; d < 3; ++d) boundary_condition[d] = (int)p->integer("boundary_condition");
    } else if (t == "generate")
      generate();
    else
      return false;
    return true;
  });
}

void Domain::generate()
{
  if (!(upper_global.x > lower_global.x && upper_global.y > lower_global.y &&
        upper_global.z > lower_global.z))
    throw Error("domain: upper bounds must exceed lower bounds");
  generated = true;
}

// --------------------------------------------------------------- Atom_data --

Atom_data::Atom_data() = default;

Atom_data::~Atom_data()
{
  if (ctx_) cav_destroy(ctx_);
}

cav_ctx *Atom_data::device()
{
  if (!ctx_) {
    check(nullptr, cav_create(device_index, &ctx_), "cav_create");
    check(ctx_, cav_set_precision(ctx_, precision), "cav_set_precision");
  }
  return ctx_;
}

bool Atom_data::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "ghost_cutoff") {
      ghost_cutoff = p->real("ghost_cutoff");
      if (ghost_cutoff < 0.0) p->fail("ghost_cutoff have to non-negative.");
    } else if (t == "k_b") {
      k_b = p->real("k_b");
      if (k_b < 0.0) p->fail("k_b have to non-negative.");
    } else if (t == "n_r_df") {
      n_r_df = (int)p->integer("n_r_df");
    } else if (t == "temperature_process") {
      temperature_process = p->integer("temperature_process") != 0;
    } else if (t == "pressure_process") {
      pressure_process = p->integer("pressure_process") != 0;
    } else if (t == "add_random_velocity") {
      // atom_data.cpp:280-288: seed, amplitude (its absolute value), then the
      // rest of the line is not read
      const long seed = p->integer("seed");
      double amp = p->real("amplitude");
      if (amp < 0) amp = -amp;
      add_random_velocity((unsigned)seed, amp);
      p->end_of_line();
      done = true;
    } else if (t == "add_type_mass") {
      const long i = p->integer("type");
      const double m = p->real("mass");
      if (i < 0) p->fail("negative type index");
      if ((long)atom_type_params.mass.size() < i + 1) {
        atom_type_params.mass.resize(i + 1, 0.0);
        atom_type_params.mass_inv.resize(i + 1, 0.0);
      }
      atom_type_params.mass[i] = m;
      atom_type_params.mass_inv[i] = m == 0.0 ? 0.0 : 1.0 / m;  // atom_data.cpp:236-239
    } else if (t == "add_type_charge") {
      const long i = p->integer("type");
      const double c = p->real("charge");
      if (i < 0) p->fail("negative type index");
      if ((long)atom_type_params.charge.size() < i + 1) atom_type_params.charge.resize(i + 1, 0.0);
      atom_type_params.charge[i] = c;
    } else if (t == "add_type_radius") {
      const long i = p->integer("type");
      const double r = p->real("radius");
      if (i < 0) p->fail("negative type index");
      if ((long)atom_type_params.radius.size() < i + 1) atom_type_params.radius.resize(i + 1, 0.0);
      atom_type_params.radius[i] = r;
    } else if (t == "set_domain" || t == "domain") {
      domain = objects->find_domain(p->name("domain"), p);
    } else if (t == "add_xyz_data_file") {
      add_xyz_data_file(p);  // reads to the end of the line (atom_data.cpp:251-255)
      done = true;
    } else
      return false;
    return true;
  });
}

bool Atom_data::add_atom(unsigned id, unsigned type, const Vector &pos, const Vector &vel)
{
  auto &o = atom_struct_owned;
  o.id.push_back(id);
  o.type.push_back(type);
  o.position.push_back(pos);
  o.velocity.push_back(vel);
  o.acceleration.push_back(Vector{});
  o.molecule_index.push_back(-1);  // atom_data.cpp:509
  return true;
}

// add_xyz_data_file [file_name "F"] [read_velocity] (atom_data/add_xyz_data_file.cpp):
// line 1 the atom count, line 2 a comment, then "type x y z [vx vy vz]" per atom;
// ids continue from the atoms already held.
void Atom_data::add_xyz_data_file(casl::Parser *p)
{
  std::string file;
  bool read_velocity = false;
  for (;;) {
    casl::Token t = p->value();
    if (t.kind == casl::Kind::eol || t.kind == casl::Kind::eof) break;
    if (t.kind != casl::Kind::identifier) p->fail("add_xyz_data_file: unexpected token");
    if (t.text == "file_name")
      file = p->string_literal("file_name");
    else if (t.text == "read_velocity")
      read_velocity = true;
    else
      p->fail("add_xyz_data_file: unsupported option " + t.text);
  }
  if (file.empty()) p->fail("add_xyz_data_file: file_name missing");
  const std::string full = file[0] == '/' ? file : p->dir() + "/" + file;
  casl::Variables none;
  casl::Parser f = casl::Parser::from_file(full, &none);
  casl::Token tn = f.value();
  if (tn.kind != casl::Kind::int_number) f.fail("Expected integer literal");
  const long n = tn.ivalue;
  if (n == 0) f.fail("no atom found in the xyz file. Something is wrong.");
  f.end_of_line();
  f.end_of_line();  // comment line
  auto lit_real = [&](void) {
    casl::Token t = f.value();
    if (t.kind == casl::Kind::real_number) return t.rvalue;
    if (t.kind == casl::Kind::int_number) return (double)t.ivalue;
    f.fail("Expected real number literal");
  };
  for (long i = 0; i < n; ++i) {
    casl::Token tt = f.value();
    if (tt.kind != casl::Kind::int_number || tt.ivalue < 0) f.fail("Expected integer literal");
    Vector pos, vel;
    pos.x = lit_real();
    pos.y = lit_real();
    pos.z = lit_real();
    if (read_velocity) {
      vel.x = lit_real();
      vel.y = lit_real();
      vel.z = lit_real();
    }
    f.end_of_line();
    add_atom((unsigned)atom_struct_owned.id.size(), (unsigned)tt.ivalue, pos, vel);
  }
}

void Atom_data::add_random_velocity(unsigned seed, double amplitude)
{
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> dist(-amplitude, amplitude);
  auto &o = atom_struct_owned;
  for (auto &v : o.velocity) {
    v.x = dist(mt);
    v.y = dist(mt);
    v.z = dist(mt);
  }
  // owned_velocity_cm (kinetic_energy.cpp:413-462): sum v m / sum m
  double f[3] = {0, 0, 0}, msum = 0;
  const auto &mass = atom_type_params.mass;
  for (size_t i = 0; i < o.velocity.size(); ++i) {
    const double m = o.type[i] < mass.size() ? mass[o.type[i]] : 0.0;
    msum += m;
    f[0] += o.velocity[i].x * m;
    f[1] += o.velocity[i].y * m;
    f[2] += o.velocity[i].z * m;
  }
  const Vector cm{f[0] / msum, f[1] / msum, f[2] / msum};
  for (size_t i = 0; i < o.velocity.size(); ++i)
    if (o.type[i] < mass.size() && mass[o.type[i]] != 0.0) {
      o.velocity[i].x -= cm.x;
      o.velocity[i].y -= cm.y;
      o.velocity[i].z -= cm.z;
    }
}

double Atom_data::kinetic_energy()
{
  double ke = 0.0;
  check(device(), cav_thermo(device(), &ke, nullptr, nullptr), "cav_thermo");
  return ke;
}

double Atom_data::temperature()
{
  double t = 0.0;
  check(device(), cav_thermo(device(), nullptr, &t, nullptr), "cav_thermo");
  return t;
}

double Atom_data::pressure()
{
  double p = 0.0;
  check(device(), cav_thermo(device(), nullptr, nullptr, &p), "cav_thermo");
  return p;
}

void Atom_data::verify_settings()
{
  if (!domain) throw Error("atom_data: domain = nullptr");
  if (atom_struct_owned.position.empty()) throw Error("atom_data: no atoms");
  unsigned tmax = 0;
  for (unsigned t : atom_struct_owned.type) tmax = std::max(tmax, t);
  if (atom_type_params.mass.size() < tmax + 1)
    throw Error("atom_data: no mass given for type " + std::to_string(atom_type_params.mass.size()));
}

int Atom_data::num_types() const
{
  unsigned tmax = 0;
  for (unsigned t : atom_struct_owned.type) tmax = std::max(tmax, t);
  return (int)std::max<size_t>(tmax + 1, atom_type_params.mass.size());
}

void Atom_data::upload()
{
  verify_settings();
  cav_ctx *c = device();
  const double lo[3] = {domain->lower_global.x, domain->lower_global.y, domain->lower_global.z};
  const double hi[3] = {domain->upper_global.x, domain->upper_global.y, domain->upper_global.z};
  check(c, cav_set_box(c, lo, hi, domain->boundary_condition), "cav_set_box");
  check(c, cav_set_ghost_cutoff(c, ghost_cutoff), "cav_set_ghost_cutoff");
  const int nt = num_types();
  std::vector<double> mass(nt, 0.0), charge(nt, 0.0);
  for (int t = 0; t < nt; ++t) {
    if (t < (int)atom_type_params.mass.size()) mass[t] = atom_type_params.mass[t];
    if (t < (int)atom_type_params.charge.size()) charge[t] = atom_type_params.charge[t];
  }
  check(c, cav_set_types(c, nt, mass.data(), charge.data()), "cav_set_types");
  if (temperature_process && k_b < 0) throw Error("atom_data: k_b is not set.");
  check(c, cav_set_thermo(c, k_b, n_r_df, temperature_process ? 1 : 0, pressure_process ? 1 : 0),
        "cav_set_thermo");
  const auto &o = atom_struct_owned;
  const int n = (int)o.position.size();
  std::vector<int> id(n), type(n), mol(n);
  std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    id[i] = (int)o.id[i];
    type[i] = (int)o.type[i];
    mol[i] = o.molecule_index[i];
    pos[3 * i] = o.position[i].x;
    pos[3 * i + 1] = o.position[i].y;
    pos[3 * i + 2] = o.position[i].z;
    vel[3 * i] = o.velocity[i].x;
    vel[3 * i + 1] = o.velocity[i].y;
    vel[3 * i + 2] = o.velocity[i].z;
  }
  check(c, cav_upload_atoms(c, n, id.data(), type.data(), mol.data(), pos.data(), vel.data()),
        "cav_upload_atoms");
}

static void download_vec(cav_ctx *c, int which, std::vector<Vector> &v)
{
  std::vector<double> buf(3 * v.size());
  check(c, cav_download(c, which, buf.data()), "cav_download");
  for (size_t i = 0; i < v.size(); ++i) v[i] = Vector{buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]};
}

void Atom_data::sync_from_device()
{
  cav_ctx *c = device();
  download_vec(c, CAV_POS, atom_struct_owned.position);
  download_vec(c, CAV_VEL, atom_struct_owned.velocity);
  download_vec(c, CAV_ACC, atom_struct_owned.acceleration);
}

void Atom_data::ghosts(std::vector<Vector> &pos, std::vector<int> &src)
{
  cav_ctx *c = device();
  int g = 0;
  check(c, cav_num_ghosts(c, &g), "cav_num_ghosts");
  std::vector<double> buf(3 * (size_t)g);
  src.assign(g, 0);
  if (g) check(c, cav_get_ghosts(c, buf.data(), src.data()), "cav_get_ghosts");
  pos.resize(g);
  for (int i = 0; i < g; ++i) pos[i] = Vector{buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]};
}

void Atom_data::reset_owned_acceleration()
{
  check(ctx_, cav_reset_acceleration(device()), "cav_reset_acceleration");
}

void Atom_data::reset_virial() { check(ctx_, cav_reset_virial(device()), "cav_reset_virial"); }

double Atom_data::virial()
{
  double v = 0.0;
  check(ctx_, cav_virial(device(), &v), "cav_virial");
  return v;
}

// ------------------------------------------------------------ Neighborlist --

bool Neighborlist::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      domain = objects->find_domain(p->name("domain"), p);
    else if (t == "cutoff") {
      cutoff = p->real("cutoff");
      if (!(cutoff > 0.0)) p->fail("cutoff have to be positive.");
    } else if (t == "cutoff_extra_coef") {
      cutoff_extra_coef = p->real("cutoff_extra_coef");
      if (cutoff_extra_coef < 1.0) p->fail("cutoff_extra_coef have to be equal or larger than 1.0");
    } else if (t == "dt")
      dt = p->real("dt");
    else if (t == "make_verlet_list_from_cell_list")
      make_verlet_list_from_cell_list = true;
    else if (t == "build_cell_list")
      build_cell_list = true;
    else if (t == "all_atom_test")
      all_atom_test = true;
    else if (t == "rebuild_test")
      rebuild_test = true;
    else
      return false;
    return true;
  });
}

void Neighborlist::verify_settings()
{
  if (!atom_data) throw Error("neighborlist: atom_data = nullptr");
  if (!(cutoff > 0.0)) throw Error("neighborlist: cutoff not set");
}

void Neighborlist::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  check(c, cav_set_neighbor(c, cutoff, cutoff_extra_coef, make_verlet_list_from_cell_list,
                            build_cell_list, all_atom_test, rebuild_test),
        "cav_set_neighbor");
}

bool Neighborlist::build(bool update)
{
  cav_ctx *c = atom_data->device();
  int rebuilt = 0;
  check(c, cav_neighbor_build(c, update ? 1 : 0, &rebuilt), "cav_neighbor_build");
  return rebuilt != 0;
}

std::vector<std::vector<unsigned>> Neighborlist::neighlist()
{
  cav_ctx *c = atom_data->device();
  int n = 0;
  check(c, cav_num_atoms(c, &n), "cav_num_atoms");
  long np = 0;
  check(c, cav_num_pairs(c, &np), "cav_num_pairs");
  std::vector<int> ij(2 * (size_t)np);
  if (np) check(c, cav_get_pairs(c, ij.data(), np, &np), "cav_get_pairs");
  std::vector<std::vector<unsigned>> out(n);
  for (long k = 0; k < np; ++k) out[ij[2 * k]].push_back((unsigned)ij[2 * k + 1]);
  return out;
}

// ------------------------------------------------------------- Force_field --

static cav_ctx *ff_ctx(Force_field *f)
{
  if (!f->atom_data) throw Error(std::string(f->type_name()) + ": atom_data = nullptr");
  return f->atom_data->device();
}

void Force_field::calculate_acceleration()
{
  cav_ctx *c = ff_ctx(this);
  check(c, cav_forces(c, mask()), "cav_forces");
}

double Force_field::energy()
{
  throw Error(std::string(type_name()) + ": The energy calculation of this force_field is not implemented");
}

double Force_field::potential(const Vector &)
{
  throw Error(std::string(type_name()) + ": The potential calculation of this force_field is not implemented");
}

Vector Force_field::field(const Vector &)
{
  throw Error(std::string(type_name()) + ": The field calculation of this force_field is not implemented");
}

static void point_query(Force_field *f, int which, const Vector &r, Vector *e, double *phi)
{
  cav_ctx *c = ff_ctx(f);
  const double pt[3] = {r.x, r.y, r.z};
  double fe[3], pot;
  check(c, cav_point_field(c, which, 1, pt, fe, &pot), "cav_point_field");
  if (e) *e = Vector{fe[0], fe[1], fe[2]};
  if (phi) *phi = pot;
}

namespace force_field {

bool Lj_cell_list::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "cutoff") {
      cutoff = p->real("cutoff");
      if (cutoff < 0.0) p->fail("Force field cutoff have to non-negative.");
    } else if (t == "epsilon") {
      if (read_matrix_element(p, epsilon, 0.0, "epsilon") < 0) p->fail("Epsilon have to be non-negative.");
    } else if (t == "sigma") {
      if (read_matrix_element(p, sigma, 0.0, "sigma") < 0) p->fail("Sigma have to be non-negative.");
    } else if (t == "set_neighborlist" || t == "neighborlist")
      neighborlist = objects->find_neighborlist(p->name("neighborlist"), p);
    else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else
      return false;
    return true;
  });
}

void Lj_cell_list::verify_settings()
{
  if (!atom_data) throw Error("lj_cell_list: atom_data = nullptr");
  if (!neighborlist) throw Error("lj_cell_list: neighborlist = nullptr");
}

void Lj_cell_list::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  const int nt = atom_data->num_types();
  const auto e = dense(epsilon, nt, 0.0), s = dense(sigma, nt, 0.0);
  check(c, cav_set_lj(c, cutoff, e.data(), s.data()), "cav_set_lj");
}

// The reference's Lj_cell_list has no energy; this is the energy whose
// gradient is its shifted force (zero at the cutoff).
double Lj_cell_list::energy()
{
  cav_ctx *c = ff_ctx(this);
  double e = 0.0;
  check(c, cav_energy(c, &e, nullptr, nullptr), "cav_energy");
  return e;
}

bool Lj::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "cutoff") {
      cutoff = p->real("cutoff");
      if (cutoff < 0.0) p->fail("Force field cutoff have to non-negative.");
    } else if (t == "cutoff_list") {
      if (read_matrix_element(p, cutoff_list, 0.0, "cutoff_list") < 0) p->fail("Epsilon have to be non-negative.");
      cutoff_list_activated = true;
    } else if (t == "epsilon") {
      if (read_matrix_element(p, epsilon, 0.0, "epsilon") < 0) p->fail("Epsilon have to be non-negative.");
      input_by_array = true;
    } else if (t == "sigma") {
      if (read_matrix_element(p, sigma, 0.0, "sigma") < 0) p->fail("Sigma have to be non-negative.");
      input_by_array = true;
    } else if (t == "lambda_s") {
      if (read_matrix_element(p, lambda_s, 1.0, "lambda_s") < 0) p->fail("Sigma have to be non-negative.");
      lambda_s_is_set = true;
    } else if (t == "lambda_e") {
      if (read_matrix_element(p, lambda_e, 1.0, "lambda_e") < 0) p->fail("Sigma have to be non-negative.");
      lambda_e_is_set = true;
    } else if (t == "epsilon_atom" || t == "sigma_atom") {
      // GET_A_STDVECTOR_REAL_ELEMENT (preprocessors.h:592-600): grows with 0
      auto &v = t == "epsilon_atom" ? epsilon_atom : sigma_atom;
      const long i = p->integer(t + " index");
      const double x = p->real(t);
      if (i < 0) p->fail(t + ": negative type index");
      if ((long)v.size() < i + 1) v.resize(i + 1, 0.0);
      v[i] = x;
      input_by_atom = true;
    } else if (t == "set_neighborlist" || t == "neighborlist")
      neighborlist = objects->find_neighborlist(p->name("neighborlist"), p);
    else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      domain = objects->find_domain(p->name("domain"), p);
    else if (t == "jump_tol") {
      jump_fix = true;
      jump_tol = p->real("jump_tol");
    } else if (t == "monitor_jump")
      monitor_jump = true;
    else if (t == "ignore_intra_molecule")
      ignore_intra_molecule = true;
    else if (t == "wca") {
      wca = true;
      cutoff_list_activated = true;
    } else if (t == "make_off_diagonal_vectors")
      make_off_diagonal_vectors = true;
    else
      return false;
    return true;
  });
}

void Lj::verify_settings()
{
  if (!atom_data) throw Error("lj: atom_data = nullptr");
  if (!neighborlist) throw Error("lj: neighborlist = nullptr");
  if (input_by_atom && input_by_array) throw Error("lj: (input_by_atom && input_by_array)");
  if (!input_by_atom && !input_by_array) throw Error("lj: no epsilon / sigma given");
}

void Lj::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  const int nt = atom_data->num_types();
  std::vector<double> e, s, ea, sa;
  if (input_by_array) {
    e = dense(epsilon, nt, 0.0);
    s = dense(sigma, nt, 0.0);
  } else {
    ea.assign(nt, 0.0);
    sa.assign(nt, 0.0);
    for (int i = 0; i < nt && i < (int)epsilon_atom.size(); ++i) ea[i] = epsilon_atom[i];
    for (int i = 0; i < nt && i < (int)sigma_atom.size(); ++i) sa[i] = sigma_atom[i];
  }
  const auto cl = dense(cutoff_list, nt, 0.0);
  const auto ls = dense(lambda_s, nt, 1.0), le = dense(lambda_e, nt, 1.0);
  check(c,
        cav_set_lj_list(c, cutoff, input_by_array ? e.data() : nullptr, input_by_array ? s.data() : nullptr,
                        input_by_atom ? ea.data() : nullptr, input_by_atom ? sa.data() : nullptr,
                        cutoff_list_activated && !wca ? cl.data() : nullptr, wca ? 1 : 0,
                        lambda_s_is_set ? ls.data() : nullptr, lambda_e_is_set ? le.data() : nullptr,
                        ignore_intra_molecule ? 1 : 0, jump_fix ? 1 : 0, jump_tol),
        "cav_set_lj_list");
}

bool Electrostatic_short_range::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "cutoff") {
      cutoff = p->real("cutoff");
      if (cutoff < 0.0) p->fail("Force field cutoff have to non-negative.");
    } else if (t == "external_field") {
      external_field.x = p->real("external_field");
      external_field.y = p->real("external_field");
      external_field.z = p->real("external_field");
    } else if (t == "k_electrostatic") {
      k_electrostatic = p->real("k_electrostatic");
      if (k_electrostatic < 0) p->fail("k_electrostatic has to be non-negative.");
    } else if (t == "lambda") {
      if (read_matrix_element(p, lambda, 1.0, "lambda") < 0) p->fail("Sigma have to be non-negative.");
      lambda_is_set = true;
    } else if (t == "beta")
      beta = p->real("beta");
    else if (t == "set_neighborlist" || t == "neighborlist")
      neighborlist = objects->find_neighborlist(p->name("neighborlist"), p);
    else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      domain = objects->find_domain(p->name("domain"), p);
    else
      return false;
    return true;
  });
}

void Electrostatic_short_range::verify_settings()
{
  if (!atom_data) throw Error("electrostatic_short_range: atom_data = nullptr");
  if (!neighborlist) throw Error("electrostatic_short_range: neighborlist = nullptr");
  if (cutoff < 0.0) throw Error("electrostatic_short_range: Force field cutoff have to non-negative.");
}

void Electrostatic_short_range::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  const int nt = atom_data->num_types();
  const auto lam = dense(lambda, nt, 1.0);
  const double ext[3] = {external_field.x, external_field.y, external_field.z};
  check(c, cav_set_short_range(c, cutoff, k_electrostatic, beta, ext, lambda_is_set ? lam.data() : nullptr),
        "cav_set_short_range");
}

bool Electrostatic_ewald_r::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "cutoff") {
      cutoff = p->real("cutoff");
      if (cutoff < 0.0) p->fail("Force field cutoff have to non-negative.");
    } else if (t == "k_electrostatic") {
      k_electrostatic = p->real("k_electrostatic");
      if (k_electrostatic < 0) p->fail("k_electrostatic has to be non-negative.");
    } else if (t == "alpha") {
      alpha = p->real("alpha");
      if (alpha < 0.0) p->fail("alpha have to non-negative.");
    } else if (t == "lambda") {
      if (read_matrix_element(p, lambda, 1.0, "lambda") < 0) p->fail("lambda has to be non-negative.");
      lambda_is_set = true;
    } else if (t == "set_neighborlist" || t == "neighborlist")
      neighborlist = objects->find_neighborlist(p->name("neighborlist"), p);
    else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else
      return false;
    return true;
  });
}

void Electrostatic_ewald_r::verify_settings()
{
  if (!atom_data) throw Error("electrostatic_ewald_r: atom_data = nullptr");
  if (!neighborlist) throw Error("electrostatic_ewald_r: neighborlist = nullptr");
}

void Electrostatic_ewald_r::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  const int nt = atom_data->num_types();
  // verify_settings of the reference pads lambda with 1 (electrostatic_ewald_r.cpp:99-110)
  const auto lam = dense(lambda, nt, 1.0);
  check(c, cav_set_ewald_r(c, alpha, k_electrostatic, lambda_is_set ? lam.data() : nullptr),
        "cav_set_ewald_r");
}

double Electrostatic_ewald_r::energy()
{
  cav_ctx *c = ff_ctx(this);
  double e = 0.0;
  check(c, cav_energy(c, nullptr, &e, nullptr), "cav_energy");
  return e;
}

double Electrostatic_ewald_r::potential(const Vector &r)
{
  double phi = 0.0;
  point_query(this, 0, r, nullptr, &phi);
  return phi;
}

Vector Electrostatic_ewald_r::field(const Vector &r)
{
  Vector e;
  point_query(this, 0, r, &e, nullptr);
  return e;
}

bool Electrostatic_ewald_k::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "k_max") {
      const long k = p->integer("k_max");
      if (k < 0) p->fail("'k_max' have to non-negative.");
      kx_max = ky_max = kz_max = (int)k;
    } else if (t == "kx_max" || t == "ky_max" || t == "kz_max") {
      const long k = p->integer(t);
      if (k < 0) p->fail("'" + t + "' have to non-negative.");
      (t == "kx_max" ? kx_max : t == "ky_max" ? ky_max : kz_max) = (int)k;
    } else if (t == "slab_normal_axis") {
      casl::Token a = p->value();
      if (a.kind != casl::Kind::identifier || (a.text != "x" && a.text != "y" && a.text != "z"))
        p->fail("expected 'x' or 'y' or 'z' for axis.");
      slab_geometry = true;
      slab_normal_axis = a.text == "x" ? 0 : a.text == "y" ? 1 : 2;
    } else if (t == "k_electrostatic") {
      k_electrostatic = p->real("k_electrostatic");
      if (k_electrostatic < 0) p->fail("'k_electrostatic' has to be non-negative.");
    } else if (t == "epsilon_dipole") {
      epsilon_dipole = p->real("epsilon_dipole");
      if (epsilon_dipole < 0) p->fail("'epsilon_dipole' has to be non-negative.");
      dipole = true;
    } else if (t == "slab_lz" || t == "slab_thickness") {
      slab_lz = p->real("slab_lz");
      if (slab_lz < 0.0) p->fail("slab_lz have to non-negative.");
      slab_geometry = true;
    } else if (t == "alpha") {
      alpha = p->real("alpha");
      if (alpha < 0.0) p->fail("alpha have to non-negative.");
    } else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      domain = objects->find_domain(p->name("domain"), p);
    else
      return false;
    return true;
  });
}

void Electrostatic_ewald_k::verify_settings()
{
  if (!atom_data) throw Error("electrostatic_ewald_k: atom_data = nullptr");
  if (!domain) throw Error("electrostatic_ewald_k: domain = nullptr");
  if (alpha < 0.0) throw Error("electrostatic_ewald_k: alpha is not set");
}

void Electrostatic_ewald_k::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  check(c, cav_set_ewald_k(c, alpha, k_electrostatic, kx_max, ky_max, kz_max, dipole ? 1 : 0,
                           epsilon_dipole, slab_geometry ? slab_normal_axis : -1, slab_lz),
        "cav_set_ewald_k");
}

double Electrostatic_ewald_k::energy()
{
  cav_ctx *c = ff_ctx(this);
  double e = 0.0;
  check(c, cav_energy(c, nullptr, nullptr, &e), "cav_energy");
  return e;
}

// -k_e alpha / sqrt(pi) sum_j q_j^2 over the owned atoms in order
// (electrostatic_ewald_k/energy.cpp:31-50).
double Electrostatic_ewald_k::self_energy()
{
  const auto &o = atom_data->atom_struct_owned;
  const auto &q = atom_data->atom_type_params.charge;
  double s = 0.0;
  for (size_t j = 0; j < o.type.size(); ++j) {
    const double qj = o.type[j] < q.size() ? q[o.type[j]] : 0.0;
    s += qj * qj;
  }
  return -0.56418958354775628 * alpha * s * k_electrostatic;
}

double Electrostatic_ewald_k::potential(const Vector &r)
{
  double phi = 0.0;
  point_query(this, 1, r, nullptr, &phi);
  return phi;
}

Vector Electrostatic_ewald_k::field(const Vector &r)
{
  Vector e;
  point_query(this, 1, r, &e, nullptr);
  return e;
}

std::vector<double> Electrostatic_ewald_k::potential_k_coef()
{
  cav_ctx *c = ff_ctx(this);
  int nk = 0;
  check(c, cav_num_kvectors(c, &nk), "cav_num_kvectors");
  std::vector<double> sk(2 * (size_t)nk);
  if (nk) check(c, cav_get_sk(c, sk.data()), "cav_get_sk");
  return sk;
}

bool Electrostatic_ewald_slab_correction::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "k_max") {
      const long k = p->integer("k_max");
      if (k < 0) p->fail("'k_max' have to non-negative.");
      kx_max = ky_max = (int)k;
    } else if (t == "kx_max" || t == "ky_max") {
      const long k = p->integer(t);
      if (k < 0) p->fail("'" + t + "' have to non-negative.");
      (t == "kx_max" ? kx_max : ky_max) = (int)k;
    } else if (t == "lz") {
      lz = p->real("lz");
      if (lz < 0) p->fail("'lz' has to be non-negative.");
    } else if (t == "slab_normal_axis") {
      casl::Token a = p->value();
      if (a.kind != casl::Kind::identifier) p->fail("expected 'x', 'y' or 'z'.");
      // an identifier other than x, y, z leaves the axis as it was
      // (electrostatic_ewald_slab_correction.cpp:80-92)
      if (a.text == "x") slab_normal_axis = 0;
      if (a.text == "y") slab_normal_axis = 1;
      if (a.text == "z") slab_normal_axis = 2;
    } else if (t == "k_electrostatic") {
      k_electrostatic = p->real("k_electrostatic");
      if (k_electrostatic < 0) p->fail("k_electrostatic has to be non-negative.");
    } else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      domain = objects->find_domain(p->name("domain"), p);
    else
      return false;
    return true;
  });
}

void Electrostatic_ewald_slab_correction::verify_settings()
{
  if (!atom_data) throw Error("electrostatic_ewald_slab_correction: atom_data = nullptr");
  if (!domain) throw Error("electrostatic_ewald_slab_correction: domain = nullptr");
}

void Electrostatic_ewald_slab_correction::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  check(c, cav_set_elc(c, slab_normal_axis, kx_max, ky_max, lz, k_electrostatic), "cav_set_elc");
}

double Electrostatic_ewald_slab_correction::energy()
{
  cav_ctx *c = ff_ctx(this);
  double e = 0.0;
  check(c, cav_energy_elc(c, &e), "cav_energy_elc");
  return e;
}

double Electrostatic_ewald_slab_correction::potential(const Vector &r)
{
  double phi = 0.0;
  point_query(this, 2, r, nullptr, &phi);
  return phi;
}

Vector Electrostatic_ewald_slab_correction::field(const Vector &r)
{
  Vector e;
  point_query(this, 2, r, &e, nullptr);
  return e;
}

}  // namespace force_field

// -------------------------------------------------------------- constraints --

namespace constraint {

bool Berendsen::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "temperature") {
      temperature = p->real("temperature");
      if (temperature <= 0.0) p->fail("Temperature have to non-negative.");
    } else if (t == "coupling") {
      coupling = p->real("coupling");
      if (coupling <= 0.0) p->fail("coupling have to non-negative.");
    } else if (t == "dt") {
      dt = p->real("dt");
      if (dt <= 0.0) p->fail("dt have to non-negative.");
    } else if (t == "step") {
      step = (int)p->integer("step");
      if (step <= 0) p->fail("step have to non-negative.");
    } else if (t == "lambda_min") {
      lambda_min = p->real("lambda_min");
      if (lambda_min <= 0.0) p->fail("lambda_min have to non-negative.");
    } else if (t == "lambda_max") {
      lambda_max = p->real("lambda_max");
      if (lambda_max <= 0.0) p->fail("lambda_max have to non-negative.");
    } else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else
      return false;
    return true;
  });
}

void Berendsen::verify_settings()
{
  if (!atom_data) throw Error("berendsen: atom_data = nullptr");
  if (dt < 0.0) throw Error("berendsen: dt is not set.");
  if (coupling < 0.0) throw Error("berendsen: coupling is not set.");
}

void Berendsen::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  check(c, cav_set_berendsen(c, temperature, coupling, dt, step, lambda_min, lambda_max), "cav_set_berendsen");
}

bool Nve::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    auto positive = [&](double &v, const char *what) {
      v = p->real(what);
      if (v <= 0.0) p->fail(std::string(what) + " have to non-negative.");
    };
    if (t == "energy_per_dof")
      positive(energy_per_dof, "energy_per_dof");
    else if (t == "energy_tot")
      positive(energy_tot, "energy_tot");
    else if (t == "kb")
      positive(kb, "kb");
    else if (t == "kbt")
      positive(kbt, "kbt");
    else if (t == "temperature")
      positive(temperature, "temperature");
    else if (t == "step") {
      step = (int)p->integer("step");
      if (step <= 0) p->fail("step have to non-negative.");
    } else if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else
      return false;
    return true;
  });
}

void Nve::verify_settings()
{
  if (!atom_data) throw Error("nve: atom_data = nullptr");
}

void Nve::init()
{
  verify_settings();
  cav_ctx *c = atom_data->device();
  check(c, cav_set_nve(c, energy_per_dof, energy_tot, kbt, kb, temperature, step), "cav_set_nve");
}

}  // namespace constraint

// ------------------------------------------------------------------- writer --

namespace writer {

struct Atom_data::Files {
  std::ofstream xyz, energy, temperature, pressure;
};

Atom_data::~Atom_data() { delete files_; }

bool Atom_data::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    auto steps = [&](long &v, bool &on) {
      v = p->integer(t);
      on = v > 0;
    };
    if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "set_domain" || t == "domain")
      (void)objects->find_domain(p->name("domain"), p);
    else if (t == "xyz_step")
      steps(xyz_step, output_xyz);
    else if (t == "energy_step")
      steps(energy_step, output_energy);
    else if (t == "temperature_step")
      steps(temperature_step, output_temperature);
    else if (t == "pressure_step")
      steps(pressure_step, output_pressure);
    else if (t == "file_name_xyz")
      file_name_xyz = p->value().text;
    else if (t == "file_name_energy")
      file_name_energy = p->value().text;
    else if (t == "file_name_temperature")
      file_name_temperature = p->value().text;
    else if (t == "file_name_pressure")
      file_name_pressure = p->value().text;
    else if (t == "xyz_output_id")
      xyz_output_id = p->integer(t) != 0;
    else if (t == "xyz_output_velocity")
      xyz_output_velocity = p->integer(t) != 0;
    else if (t == "xyz_output_acceleration")
      xyz_output_acceleration = p->integer(t) != 0;
    else if (t == "xyz_precision") {
      xyz_precision = (int)p->integer(t);
      if (xyz_precision < 1 || xyz_precision > 17) p->fail("xyz_precision: 1..17");
    } else
      return false;
    return true;
  });
}

void Atom_data::open()
{
  if (!atom_data) throw Error("writer atom_data: atom_data = nullptr");
  files_ = new Files;
  auto path = [&](const std::string &f, const char *ext) { return (dir.empty() ? f : dir + "/" + f) + ext; };
  if (output_xyz) files_->xyz.open(path(file_name_xyz, ".xyz"));
  if (output_energy) files_->energy.open(path(file_name_energy, ".txt"));
  if (output_temperature) files_->temperature.open(path(file_name_temperature, ".txt"));
  if (output_pressure) files_->pressure.open(path(file_name_pressure, ".txt"));
}

bool Atom_data::wants(long i) const
{
  return (output_xyz && i % xyz_step == 0) || (output_energy && i % energy_step == 0) ||
         (output_temperature && i % temperature_step == 0) || (output_pressure && i % pressure_step == 0);
}

// write_serial (writer/atom_data.cpp:678-720) for the outputs above, with the
// owned arrays as the last step left them.
void Atom_data::write(long i, double t)
{
  if (!files_) open();
  if (output_xyz && i % xyz_step == 0) {
    // dump_xyz_serial (writer/atom_data/dump_xyz_serial.cpp:33-71)
    atom_data->sync_from_device();
    const auto &o = atom_data->atom_struct_owned;
    std::ostream &f = files_->xyz;
    f << std::setprecision(xyz_precision);
    f << o.position.size() << "\nAtom\n";
    for (size_t k = 0; k < o.position.size(); ++k) {
      f << o.type[k];
      if (xyz_output_id) f << " " << o.id[k];
      f << " " << o.position[k].x << " " << o.position[k].y << " " << o.position[k].z;
      if (xyz_output_velocity) f << " " << o.velocity[k].x << " " << o.velocity[k].y << " " << o.velocity[k].z;
      if (xyz_output_acceleration)
        f << " " << o.acceleration[k].x << " " << o.acceleration[k].y << " " << o.acceleration[k].z;
      f << "\n";
    }
    f << std::flush;
    f << std::setprecision(6);
  }
  // dump_energy_serial, dump_temperature_serial, dump_pressure_serial
  // (writer/atom_data/dump_energy.cpp:44-139)
  if (output_energy && i % energy_step == 0)
    files_->energy << i << " " << t << " " << atom_data->kinetic_energy() << std::endl;
  if (output_temperature && i % temperature_step == 0)
    files_->temperature << i << " " << t << " " << atom_data->temperature() << std::endl;
  if (output_pressure && i % pressure_step == 0)
    files_->pressure << i << " " << t << " " << atom_data->pressure() << std::endl;
}

}  // namespace writer

// ------------------------------------------------------------ Md_simulator --

bool Md_simulator::read(casl::Parser *p)
{
  return read_commands(p, [&](const std::string &t, bool &done) {
    if (t == "set_atom_data" || t == "atom_data")
      atom_data = objects->find_atom_data(p->name("atom_data"), p);
    else if (t == "add_force_field" || t == "force_field")
      force_field.push_back(objects->find_force_field(p->name("force_field"), p));
    else if (t == "add_neighborlist" || t == "neighborlist")
      neighborlist.push_back(objects->find_neighborlist(p->name("neighborlist"), p));
    else if (t == "add_constraint" || t == "constraint")
      constraint.push_back(objects->find_constraint(p->name("constraint"), p));
    else if (t == "add_writer" || t == "writer")
      writer.push_back(objects->find_writer(p->name("writer"), p));
    else if (t == "integrator_type") {
      casl::Token it = p->value();
      if (it.kind != casl::Kind::identifier || it.text != "velocity_verlet")
        p->fail("integrator_type: only velocity_verlet is on this path");
    } else if (t == "dt")
      dt = p->real("dt");
    else if (t == "initial_step")
      initial_step = p->integer("initial_step");
    else if (t == "final_step")
      final_step = p->integer("final_step");
    else if (t == "final_step_add") {
      final_step += p->integer("final_step_add");
      initial_step = current_step;
    } else if (t == "run")
      run();
    else if (t == "step")
      step();
    else
      return false;
    return true;
  });
}

void Md_simulator::verify_settings()
{
  if (!atom_data) throw Error("md_simulator: atom_data = nullptr");
  if (neighborlist.size() > 1) throw Error("md_simulator: one neighborlist per atom_data on this path");
  int seen = 0;
  for (Force_field *f : force_field) {
    if (f->atom_data != atom_data)
      throw Error(std::string(f->type_name()) + ": force fields must use the simulator's atom_data");
    if (seen & f->mask())
      throw Error(std::string("md_simulator: one ") + f->type_name() + " per atom_data on this path");
    seen |= f->mask();
    if (f->neighborlist && (neighborlist.empty() || f->neighborlist != neighborlist[0]))
      throw Error(std::string(f->type_name()) + ": its neighborlist must be added to the simulator");
  }
  for (Neighborlist *n : neighborlist)
    if (n->atom_data != atom_data) throw Error("neighborlist: must use the simulator's atom_data");
  int ber = 0, nve = 0;
  for (Constraint *c : constraint) {
    if (c->atom_data != atom_data) throw Error(std::string(c->type_name()) + ": must use the simulator's atom_data");
    ber += dynamic_cast<constraint::Berendsen *>(c) != nullptr;
    nve += dynamic_cast<constraint::Nve *>(c) != nullptr;
  }
  if (ber > 1 || nve > 1) throw Error("md_simulator: one berendsen and one nve constraint at most on this path");
  if (ber && nve)
    for (size_t k = 0; k < constraint.size(); ++k)
      if (dynamic_cast<constraint::Nve *>(constraint[k]))
        if (k == 0) throw Error("md_simulator: with both, berendsen must come before nve on this path");
}

bool Md_simulator::dry_run = false;

void Md_simulator::setup()
{
  verify_settings();
  if (dry_run) {
    atom_data->verify_settings();
    for (Neighborlist *n : neighborlist) n->verify_settings();
    for (Force_field *f : force_field) f->verify_settings();
    initialized = true;
    return;
  }
  atom_data->upload();
  for (Neighborlist *n : neighborlist) n->init();
  for (Force_field *f : force_field) f->init();
  for (Constraint *k : constraint) k->init();
  cav_ctx *c = atom_data->device();
  check(c, cav_setup(c), "cav_setup");
  initialized = true;
  atom_data->sync_from_device();
  time = dt * initial_step;  // md_simulator.cpp:397
  for (writer::Atom_data *w : writer) w->write(0, 0.0);  // md_simulator.cpp:545-546
}

void Md_simulator::advance(long nsteps)
{
  if (nsteps <= 0) return;
  if (!initialized) setup();  // the first step sets up (md_simulator.cpp:387-406)
  steps_requested += nsteps;
  if (dry_run) return;
  cav_ctx *c = atom_data->device();
  if (writer.empty()) {
    check(c, cav_set_step_index(c, current_step), "cav_set_step_index");
    check(c, cav_step(c, dt, (int)nsteps), "cav_step");
    current_step += nsteps;
    time += dt * (double)nsteps;
  } else {
    // one step at a time: each writer sees the state at the end of its step
    // (md_simulator.cpp:672-673), time before the step's own advance
    for (long k = 0; k < nsteps; ++k) {
      check(c, cav_set_step_index(c, current_step), "cav_set_step_index");
      check(c, cav_step(c, dt, 1), "cav_step");
      for (writer::Atom_data *w : writer)
        if (w->wants(current_step)) w->write(current_step, time);
      time += dt;
      ++current_step;
    }
  }
  atom_data->sync_from_device();
}

void Md_simulator::step() { advance(1); }

void Md_simulator::run()
{
  verify_settings();
  current_step = initial_step;  // md_simulator.cpp:247-248
  advance(final_step - initial_step);
  if (final_step > initial_step) current_step = final_step;
}

}  // namespace caviar_amd
