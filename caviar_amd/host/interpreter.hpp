// CASL interpreter for the objects of the LJ + Ewald path: the command loop of
// the reference's Input::read_command (interpreter/input.cpp:152-200) over the
// object creators (interpreter/object_creator/objects_types.cpp:33-51, type
// names compared case-insensitively) and the `int` / `real` variables
// (object_creator/basic_types.cpp:209-260).
//
// Accepted creations: domain box, atom_data basic, neighborlist basic (also as
// cell_list = cells + Verlet list from cells, verlet_list = brute force),
// force_field lj_cell_list | lj | electrostatic_ewald_r | electrostatic_short_range |
// electrostatic_ewald_k | electrostatic_ewald_slab_correction, constraint berendsen | nve,
// writer atom_data, md_simulator basic.  Anything else is reported as an error naming the
// command, the way the reference reports an unknown command.
#pragma once

#include <string>

#include "objects.hpp"

namespace caviar_amd {

class Interpreter {
 public:
  // device: HIP device index for every Atom_data; precision: CAV_FP64 / CAV_MIXED.
  Interpreter(int device = 0, int precision = CAV_FP64);
  void run_file(const std::string &path);
  void run(casl::Parser &parser);

  Objects objects;
  casl::Variables variables;
  bool exit_called = false;

 private:
  bool command(casl::Parser &p);
  void define_variable(casl::Parser &p, bool integer);
  void create(casl::Parser &p, const std::string &base);
  int device_;
  int precision_;
};

}  // namespace caviar_amd
