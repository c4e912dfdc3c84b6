// caviar_amd_run: run a CASL script of the LJ + Ewald path on the GPU.
//
//   caviar_amd_run [--precision fp64|mixed] [--device N] [--molecules FILE] script.casl [MODE OUT]
//   caviar_amd_run --parse-only script.casl
//
// The script uses the reference's own commands (domain box, atom_data basic,
// neighborlist basic, force_field lj_cell_list / lj / electrostatic_ewald_r /
// electrostatic_short_range / electrostatic_ewald_k, md_simulator basic ... run).  With MODE and OUT the
// state is written to OUT as text:
//   MODE "setup[+er_energy][+ek_energy][+elc_energy][+sk][+field][+thermo]": after the script,
//        Md_simulator::setup (wrap, ghosts, list, one force evaluation) unless
//        a run already set up, then owned positions / velocities /
//        accelerations, ghosts, virial, the Verlet list and the requested
//        energies, S(k) and point fields (at the owned positions shifted by
//        (0.1, -0.2, 0.3)).
//   MODE "run[+thermo]": the same dump after the script's own run ("+thermo":
//        Atom_data::kinetic_energy, temperature and pressure as well).
//   MODE "mrun": the script has no run: the simulator's run is called after
//        the molecule indices are set (--molecules), then the dump.
// --molecules FILE sets atom_struct_owned.molecule_index from FILE (one integer
// per owned atom, in input order) after the script is read, as the
// reference's add_molecule commands would have set them.
// Numbers are printed with 17 significant digits.  --parse-only reads the
// script without touching the GPU (run / step are counted, not executed) and
// prints the settings it collected as JSON.
// Errors print a message and exit with status 1, as the reference does.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "interpreter.hpp"

using namespace caviar_amd;

static void dump_vecs(FILE *f, const char *tag, const std::vector<Vector> &v)
{
  std::fprintf(f, "%s %zu\n", tag, v.size());
  for (const auto &a : v) std::fprintf(f, "%.17g %.17g %.17g\n", a.x, a.y, a.z);
}

static void json_matrix(FILE *f, const std::vector<std::vector<double>> &m)
{
  std::fprintf(f, "[");
  for (size_t i = 0; i < m.size(); ++i) {
    std::fprintf(f, "%s[", i ? ", " : "");
    for (size_t j = 0; j < m[i].size(); ++j) std::fprintf(f, "%s%.17g", j ? ", " : "", m[i][j]);
    std::fprintf(f, "]");
  }
  std::fprintf(f, "]");
}

static void json_vec(FILE *f, const std::vector<double> &v)
{
  std::fprintf(f, "[");
  for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", v[i]);
  std::fprintf(f, "]");
}

static void print_settings(Interpreter &in, Md_simulator *sim)
{
  FILE *f = stdout;
  Atom_data *ad = sim->atom_data;
  std::fprintf(f, "{");
  if (ad && ad->domain) {
    const Domain *d = ad->domain;
    std::fprintf(f, "\"lower\": [%.17g, %.17g, %.17g], \"upper\": [%.17g, %.17g, %.17g], ",
                 d->lower_global.x, d->lower_global.y, d->lower_global.z, d->upper_global.x,
                 d->upper_global.y, d->upper_global.z);
    std::fprintf(f, "\"bc\": [%d, %d, %d], ", d->boundary_condition[0], d->boundary_condition[1],
                 d->boundary_condition[2]);
  }
  if (ad) {
    const auto &o = ad->atom_struct_owned;
    double ps = 0.0, vs = 0.0;
    long ts = 0;
    for (size_t i = 0; i < o.position.size(); ++i) {
      ps += o.position[i].x + 2.0 * o.position[i].y + 3.0 * o.position[i].z;
      vs += o.velocity[i].x + 2.0 * o.velocity[i].y + 3.0 * o.velocity[i].z;
      ts += (long)o.type[i] * (long)(i % 7 + 1);
    }
    std::fprintf(f, "\"n\": %zu, \"ghost_cutoff\": %.17g, \"mass\": ", o.position.size(), ad->ghost_cutoff);
    json_vec(f, ad->atom_type_params.mass);
    std::fprintf(f, ", \"charge\": ");
    json_vec(f, ad->atom_type_params.charge);
    std::fprintf(f, ", \"pos_checksum\": %.17g, \"vel_checksum\": %.17g, \"type_checksum\": %ld, ", ps, vs, ts);
  }
  if (!sim->neighborlist.empty()) {
    const Neighborlist *n = sim->neighborlist[0];
    std::fprintf(f,
                 "\"neighbor\": {\"cutoff\": %.17g, \"cutoff_extra_coef\": %.17g, \"from_cells\": %s, "
                 "\"build_cells\": %s, \"all_atom_test\": %s, \"rebuild_test\": %s}, ",
                 n->cutoff, n->cutoff_extra_coef, n->make_verlet_list_from_cell_list ? "true" : "false",
                 n->build_cell_list ? "true" : "false", n->all_atom_test ? "true" : "false",
                 n->rebuild_test ? "true" : "false");
  }
  for (Force_field *ff : sim->force_field) {
    if (auto *lj = dynamic_cast<force_field::Lj_cell_list *>(ff)) {
      std::fprintf(f, "\"lj\": {\"cutoff\": %.17g, \"epsilon\": ", lj->cutoff);
      json_matrix(f, lj->epsilon);
      std::fprintf(f, ", \"sigma\": ");
      json_matrix(f, lj->sigma);
      std::fprintf(f, "}, ");
    } else if (auto *l = dynamic_cast<force_field::Lj *>(ff)) {
      std::fprintf(f, "\"lj_list\": {\"cutoff\": %.17g, \"epsilon\": ", l->cutoff);
      json_matrix(f, l->epsilon);
      std::fprintf(f, ", \"sigma\": ");
      json_matrix(f, l->sigma);
      std::fprintf(f, ", \"epsilon_atom\": ");
      json_vec(f, l->epsilon_atom);
      std::fprintf(f, ", \"sigma_atom\": ");
      json_vec(f, l->sigma_atom);
      std::fprintf(f, ", \"cutoff_list\": ");
      json_matrix(f, l->cutoff_list);
      std::fprintf(f, ", \"lambda_s\": ");
      json_matrix(f, l->lambda_s);
      std::fprintf(f, ", \"lambda_e\": ");
      json_matrix(f, l->lambda_e);
      std::fprintf(f, ", \"wca\": %s, \"ignore_intra_molecule\": %s, \"jump_tol\": ", l->wca ? "true" : "false",
                   l->ignore_intra_molecule ? "true" : "false");
      if (l->jump_fix)
        std::fprintf(f, "%.17g}, ", l->jump_tol);
      else
        std::fprintf(f, "null}, ");
    } else if (auto *sr = dynamic_cast<force_field::Electrostatic_short_range *>(ff)) {
      std::fprintf(f,
                   "\"short_range\": {\"cutoff\": %.17g, \"k_electrostatic\": %.17g, \"beta\": %.17g, "
                   "\"external_field\": [%.17g, %.17g, %.17g], \"lambda_\": ",
                   sr->cutoff, sr->k_electrostatic, sr->beta, sr->external_field.x, sr->external_field.y,
                   sr->external_field.z);
      if (sr->lambda_is_set)
        json_matrix(f, sr->lambda);
      else
        std::fprintf(f, "null");
      std::fprintf(f, "}, ");
    } else if (auto *er = dynamic_cast<force_field::Electrostatic_ewald_r *>(ff)) {
      std::fprintf(f, "\"ewald_r\": {\"alpha\": %.17g, \"k_electrostatic\": %.17g, \"lambda_\": ", er->alpha,
                   er->k_electrostatic);
      if (er->lambda_is_set)
        json_matrix(f, er->lambda);
      else
        std::fprintf(f, "null");
      std::fprintf(f, "}, ");
    } else if (auto *el = dynamic_cast<force_field::Electrostatic_ewald_slab_correction *>(ff)) {
      std::fprintf(f,
                   "\"elc\": {\"slab_normal_axis\": %d, \"kx_max\": %d, \"ky_max\": %d, \"lz\": %.17g, "
                   "\"k_electrostatic\": %.17g}, ",
                   el->slab_normal_axis, el->kx_max, el->ky_max, el->lz, el->k_electrostatic);
    } else if (auto *ek = dynamic_cast<force_field::Electrostatic_ewald_k *>(ff)) {
      std::fprintf(f,
                   "\"ewald_k\": {\"alpha\": %.17g, \"k_electrostatic\": %.17g, \"kx_max\": %d, "
                   "\"ky_max\": %d, \"kz_max\": %d, \"epsilon_dipole\": ",
                   ek->alpha, ek->k_electrostatic, ek->kx_max, ek->ky_max, ek->kz_max);
      if (ek->dipole)
        std::fprintf(f, "%.17g", ek->epsilon_dipole);
      else
        std::fprintf(f, "null");
      std::fprintf(f, ", \"slab_normal_axis\": ");
      if (ek->slab_geometry)
        std::fprintf(f, "%d", ek->slab_normal_axis);
      else
        std::fprintf(f, "null");
      std::fprintf(f, ", \"slab_lz\": %.17g}, ", ek->slab_lz);
    }
  }
  std::fprintf(f, "\"dt\": %.17g, \"initial_step\": %ld, \"final_step\": %ld, \"objects\": %zu}\n", sim->dt,
               sim->initial_step, sim->final_step, in.objects.order.size());
}

static void dump(Md_simulator *sim, const std::string &mode, const char *out)
{
  Atom_data *ad = sim->atom_data;
  FILE *f = std::fopen(out, "w");
  if (!f) throw Error(std::string("cannot write ") + out);
  ad->sync_from_device();
  const auto &own = ad->atom_struct_owned;
  std::fprintf(f, "N %zu\n", own.position.size());
  dump_vecs(f, "pos", own.position);
  dump_vecs(f, "vel", own.velocity);
  dump_vecs(f, "acc", own.acceleration);
  std::vector<Vector> gpos;
  std::vector<int> gsrc;
  ad->ghosts(gpos, gsrc);
  dump_vecs(f, "ghost_pos", gpos);
  std::fprintf(f, "ghost_id %zu\n", gsrc.size());
  for (int s : gsrc) std::fprintf(f, "%u\n", own.id[s]);
  std::fprintf(f, "virial %.17g\n", ad->virial());
  if (!sim->neighborlist.empty()) {
    const auto nl = sim->neighborlist[0]->neighlist();
    size_t np = 0;
    for (const auto &row : nl) np += row.size();
    std::fprintf(f, "pairs %zu\n", np);
    for (size_t i = 0; i < nl.size(); ++i)
      for (unsigned j : nl[i]) std::fprintf(f, "%zu %u\n", i, j);
  }
  if (mode.find("+thermo") != std::string::npos) {
    std::fprintf(f, "kinetic_energy %.17g\n", ad->kinetic_energy());
    std::fprintf(f, "temperature %.17g\n", ad->temperature());
    std::fprintf(f, "pressure %.17g\n", ad->pressure());
  }
  for (Force_field *ff : sim->force_field) {
    auto *er = dynamic_cast<force_field::Electrostatic_ewald_r *>(ff);
    auto *ek = dynamic_cast<force_field::Electrostatic_ewald_k *>(ff);
    auto *elc = dynamic_cast<force_field::Electrostatic_ewald_slab_correction *>(ff);
    if (elc && mode.find("+elc_energy") != std::string::npos) std::fprintf(f, "energy_elc %.17g\n", elc->energy());
    if (elc && mode.find("+field") != std::string::npos) {
      std::fprintf(f, "field_elc %zu\n", own.position.size());
      for (const auto &p : own.position) {
        const Vector r{p.x + 0.1, p.y - 0.2, p.z + 0.3};
        const Vector e = elc->field(r);
        const double phi = elc->potential(r);
        std::fprintf(f, "%.17g %.17g %.17g %.17g\n", e.x, e.y, e.z, phi);
      }
    }
    if (er && mode.find("+er_energy") != std::string::npos) std::fprintf(f, "energy_r %.17g\n", er->energy());
    if (ek && mode.find("+ek_energy") != std::string::npos) {
      std::fprintf(f, "energy_k %.17g\n", ek->energy());
      std::fprintf(f, "energy_k_self %.17g\n", ek->self_energy());
    }
    if (ek && mode.find("+sk") != std::string::npos) {
      const auto sk = ek->potential_k_coef();
      std::fprintf(f, "sk %zu\n", sk.size() / 2);
      for (size_t k = 0; k < sk.size() / 2; ++k) std::fprintf(f, "%.17g %.17g\n", sk[2 * k], sk[2 * k + 1]);
    }
    if (mode.find("+field") != std::string::npos && (er || ek)) {
      std::fprintf(f, "%s %zu\n", er ? "field_r" : "field_k", own.position.size());
      for (const auto &p : own.position) {
        const Vector r{p.x + 0.1, p.y - 0.2, p.z + 0.3};
        const Vector e = ff->field(r);
        const double phi = ff->potential(r);
        std::fprintf(f, "%.17g %.17g %.17g %.17g\n", e.x, e.y, e.z, phi);
      }
    }
  }
  std::fclose(f);
}

int main(int argc, char **argv)
{
  int precision = CAV_FP64, device = 0;
  bool parse_only = false;
  std::string molecules;
  std::vector<std::string> pos;
  for (int a = 1; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--molecules") && a + 1 < argc) {
      molecules = argv[++a];
      continue;
    }
    if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
      const std::string p = argv[++a];
      if (p == "fp64")
        precision = CAV_FP64;
      else if (p == "mixed")
        precision = CAV_MIXED;
      else {
        std::fprintf(stderr, "unknown precision %s\n", p.c_str());
        return 2;
      }
    } else if (!std::strcmp(argv[a], "--device") && a + 1 < argc)
      device = std::atoi(argv[++a]);
    else if (!std::strcmp(argv[a], "--parse-only"))
      parse_only = true;
    else
      pos.push_back(argv[a]);
  }
  if (pos.size() != 1 && pos.size() != 3) {
    std::fprintf(stderr,
                 "usage: %s [--precision fp64|mixed] [--device N] [--molecules FILE] [--parse-only] script.casl "
                 "[MODE OUT]\n",
                 argv[0]);
    return 2;
  }
  try {
    Interpreter in(device, precision);
    Md_simulator::dry_run = parse_only;
    in.run_file(pos[0]);
    if (in.objects.md_simulator.empty()) throw Error("script must create an md_simulator");
    Md_simulator *sim = in.objects.md_simulator.begin()->second.get();
    if (in.objects.md_simulator.size() > 1) throw Error("one md_simulator per script on this path");
    if (!molecules.empty()) {
      if (!sim->atom_data) throw Error("--molecules: the simulator has no atom_data");
      auto &mol = sim->atom_data->atom_struct_owned.molecule_index;
      FILE *mf = std::fopen(molecules.c_str(), "r");
      if (!mf) throw Error("cannot open " + molecules);
      for (auto &m : mol)
        if (std::fscanf(mf, "%d", &m) != 1) {
          std::fclose(mf);
          throw Error(molecules + ": too few molecule indices");
        }
      std::fclose(mf);
    }
    if (parse_only) {
      print_settings(in, sim);
      return 0;
    }
    if (pos.size() == 3) {
      const std::string mode = pos[1];
      if (mode.rfind("setup", 0) == 0) {
        if (!sim->initialized) sim->setup();
      } else if (mode.rfind("mrun", 0) == 0) {
        sim->run();
      } else if (mode.rfind("run", 0) != 0)
        throw Error("MODE is setup[+...], run[+...] or mrun[+...]");
      dump(sim, mode, pos[2].c_str());
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "caviar_amd_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
