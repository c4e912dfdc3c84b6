#include "interpreter.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>

namespace caviar_amd {

Interpreter::Interpreter(int device, int precision) : device_(device), precision_(precision) {}

void Interpreter::run_file(const std::string &path)
{
  casl::Parser p = casl::Parser::from_file(path, &variables);
  run(p);
}

void Interpreter::run(casl::Parser &p)
{
  while (!exit_called && command(p)) {
  }
}

static std::string lower(std::string s)
{
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}

bool Interpreter::command(casl::Parser &p)
{
  casl::Token t;
  do {
    t = p.raw();
  } while (t.kind == casl::Kind::eol);
  if (t.kind == casl::Kind::eof) return false;
  if (t.kind != casl::Kind::identifier) p.fail("Invalid command: '" + t.text + "'");
  const std::string &c = t.text;
  if (c == "exit") {
    exit_called = true;
    return false;
  }
  if (c == "real" || c == "int") {
    define_variable(p, c == "int");
    return true;
  }
  if (c == "domain" || c == "atom_data" || c == "neighborlist" || c == "force_field" || c == "md_simulator" ||
      c == "constraint" || c == "writer") {
    create(p, c);
    return true;
  }
  // an object name: hand the rest of the line to its read()
  bool in_file = true;
  if (auto it = objects.domain.find(c); it != objects.domain.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.atom_data.find(c); it != objects.atom_data.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.neighborlist.find(c); it != objects.neighborlist.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.force_field.find(c); it != objects.force_field.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.md_simulator.find(c); it != objects.md_simulator.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.constraint.find(c); it != objects.constraint.end())
    in_file = it->second->read(&p);
  else if (auto it = objects.writer.find(c); it != objects.writer.end())
    in_file = it->second->read(&p);
  else
    p.fail("Invalid command or object name: " + c +
           " (this build reads the domain / atom_data / neighborlist / lj_cell_list / "
           "electrostatic_ewald_r / electrostatic_ewald_k / md_simulator commands)");
  return in_file;
}

// real NAME [=] EXPR | int NAME [=] EXPR | real NAME (value 0)
void Interpreter::define_variable(casl::Parser &p, bool integer)
{
  casl::Token n = p.raw();
  if (n.kind == casl::Kind::eol) return;
  if (n.kind != casl::Kind::identifier) p.fail("variable NAME expected");
  if (objects.has(n.text)) p.fail("NAME is already used: " + n.text);
  casl::Token t = p.raw();
  double v = 0.0;
  if (t.kind != casl::Kind::eol && t.kind != casl::Kind::eof) {
    if (t.kind != casl::Kind::assign) p.unget(t);
    v = p.expression();
    p.end_of_line();
  }
  if (integer) {
    if (v != std::floor(v)) p.fail("int variable " + n.text + " given a non-integral value");
    variables.ints[n.text] = (long)v;
    variables.reals.erase(n.text);
  } else {
    variables.reals[n.text] = v;
    variables.ints.erase(n.text);
  }
}

void Interpreter::create(casl::Parser &p, const std::string &base)
{
  casl::Token type = p.raw();
  if (type.kind != casl::Kind::identifier) p.fail(base + ": object type expected");
  casl::Token name = p.raw();
  if (name.kind != casl::Kind::identifier) p.fail(base + ": object NAME expected");
  const std::string ty = lower(type.text), nm = name.text;
  if (objects.has(nm) || variables.ints.count(nm) || variables.reals.count(nm))
    p.fail("NAME is already used: " + nm);
  const std::string undefined = "Undefined '" + base + "::" + type.text + "'.";
  if (base == "domain") {
    if (ty != "box") p.fail(undefined);
    auto o = std::make_unique<Domain>();
    o->objects = &objects;
    objects.domain[nm] = std::move(o);
  } else if (base == "atom_data") {
    if (ty != "basic") p.fail(undefined);
    auto o = std::make_unique<Atom_data>();
    o->objects = &objects;
    o->device_index = device_;
    o->precision = precision_;
    objects.atom_data[nm] = std::move(o);
  } else if (base == "neighborlist") {
    // basic (neighborlist/macro/basic.h); cell_list and verlet_list name its two
    // modes the way older scripts do: cells + Verlet list from cells, or brute force
    if (ty != "basic" && ty != "cell_list" && ty != "verlet_list") p.fail(undefined);
    auto o = std::make_unique<Neighborlist>();
    o->objects = &objects;
    if (ty == "cell_list") o->build_cell_list = o->make_verlet_list_from_cell_list = true;
    objects.neighborlist[nm] = std::move(o);
  } else if (base == "force_field") {
    std::unique_ptr<Force_field> o;
    if (ty == "lj_cell_list")
      o = std::make_unique<force_field::Lj_cell_list>();
    else if (ty == "lj")
      o = std::make_unique<force_field::Lj>();
    else if (ty == "electrostatic_short_range")
      o = std::make_unique<force_field::Electrostatic_short_range>();
    else if (ty == "electrostatic_ewald_r")
      o = std::make_unique<force_field::Electrostatic_ewald_r>();
    else if (ty == "electrostatic_ewald_k")
      o = std::make_unique<force_field::Electrostatic_ewald_k>();
    else if (ty == "electrostatic_ewald_slab_correction")
      o = std::make_unique<force_field::Electrostatic_ewald_slab_correction>();
    else
      p.fail(undefined);
    o->objects = &objects;
    objects.force_field[nm] = std::move(o);
  } else if (base == "constraint") {
    std::unique_ptr<Constraint> o;
    if (ty == "berendsen")
      o = std::make_unique<constraint::Berendsen>();
    else if (ty == "nve")
      o = std::make_unique<constraint::Nve>();
    else
      p.fail(undefined);
    o->objects = &objects;
    objects.constraint[nm] = std::move(o);
  } else if (base == "writer") {
    if (ty != "atom_data") p.fail(undefined);
    auto o = std::make_unique<writer::Atom_data>();
    o->objects = &objects;
    objects.writer[nm] = std::move(o);
  } else {
    if (ty != "basic") p.fail(undefined);
    auto o = std::make_unique<Md_simulator>();
    o->objects = &objects;
    objects.md_simulator[nm] = std::move(o);
  }
  objects.order.push_back(nm);
}

}  // namespace caviar_amd
