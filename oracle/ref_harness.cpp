// Golden-vector harness for the reference CAVIAR library (built from the
// reference's own sources by oracle/ref.mk into oracle/_ref/).  It runs in the
// development container only; its text output is turned into small fixtures
// under tests/golden/ by tests/golden/make_golden.py.  Nothing on the GPU box
// runs or loads it.
//
// usage:  ref_harness <script.casl> <mode> <out.txt>
//   mode "setup": the script creates every object (no 'run'); the harness then
//          performs Md_simulator::initialize + setup (md_simulator.cpp:303,504),
//          i.e. wrap, ghosts, neighbour build, one force evaluation, and dumps
//          ghosts, the Verlet list, accelerations, virial and energies.
//   mode "run":   the script ends with 'sim run'; the harness dumps the final
//          positions, velocities, accelerations and the last Verlet list.
//   mode "mrun":  as "run", but the script has no 'sim run': the harness sets
//          the molecule indices first (see below), then calls Md_simulator::run.
// An optional fourth argument names a file of molecule indices, one integer
// per owned atom in input order, written into atom_struct_owned.molecule_index
// before setup or run (what the reference's add_molecule commands produce;
// atoms read from an XYZ file otherwise all carry -1).
// "+thermo" dumps Atom_data::kinetic_energy(), temperature() and pressure().
// "+elc_energy" calls Electrostatic_ewald_slab_correction::energy(); "+field"
// also dumps the slab correction's field/potential as field_elc.
// Flags appended to the mode: "+er_energy" also calls Electrostatic_ewald_r::
// energy(), "+ek_energy" calls Electrostatic_ewald_k::energy(), "+sk" dumps
// S(k) (potential_k_coef_cmplx), "+field" dumps field/potential at the owned
// positions shifted by a fixed offset.
#include "caviar/CAVIAR.h"
#include "caviar/interpreter/all.h"
#include "caviar/objects/atom_data.h"
#include "caviar/objects/domain.h"
#include "caviar/objects/force_field.h"
#include "caviar/objects/force_field/electrostatic_ewald_k.h"
#include "caviar/objects/force_field/electrostatic_ewald_r.h"
#include "caviar/objects/force_field/electrostatic_ewald_slab_correction.h"
#include "caviar/objects/md_simulator.h"
#include "caviar/objects/neighborlist.h"

#include <cstdio>
#include <cstring>
#include <string>

using caviar::Vector;

static void dump_vecs(FILE *f, const char *tag, const std::vector<Vector<double>> &v)
{
  std::fprintf(f, "%s %zu\n", tag, v.size());
  for (const auto &a : v)
    std::fprintf(f, "%.17g %.17g %.17g\n", a.x, a.y, a.z);
}

int main(int argc, char **argv)
{
  if (argc < 4)
  {
    std::fprintf(stderr, "usage: %s script.casl mode out.txt\n", argv[0]);
    return 2;
  }
  std::string mode = argv[2];
  char a0[] = "ref_harness", a1[] = "-i";
  char *cargv[] = {a0, a1, argv[1], nullptr};
  caviar::CAVIAR cav(3, cargv);
  cav.execute();

  auto *oc = cav.object_container;
  if (oc->md_simulator.empty() || oc->atom_data.empty())
  {
    std::fprintf(stderr, "script must create atom_data and md_simulator\n");
    return 3;
  }
  auto *sim = oc->md_simulator[0];
  auto *ad = oc->atom_data[0];

  if (argc > 4)
  {
    FILE *mf = std::fopen(argv[4], "r");
    if (!mf)
    {
      std::fprintf(stderr, "cannot open %s\n", argv[4]);
      return 4;
    }
    auto &mol = ad->atom_struct_owned.molecule_index;
    int nmol = 0;
    for (size_t i = 0; i < mol.size(); ++i)
    {
      if (std::fscanf(mf, "%d", &mol[i]) != 1)
      {
        std::fprintf(stderr, "%s: too few molecule indices\n", argv[4]);
        return 4;
      }
      if (mol[i] + 1 > nmol)
        nmol = mol[i] + 1;
    }
    std::fclose(mf);
    // one molecule record per index, with its atoms' ids, as add_molecule
    // leaves them (atom_data.cpp:706-721 looks them up by index)
    ad->molecule_struct_owned.resize(nmol);
    for (size_t i = 0; i < mol.size(); ++i)
      if (mol[i] > -1)
        ad->molecule_struct_owned[mol[i]].atom_list.push_back(ad->atom_struct_owned.id[i]);
  }

  if (mode.rfind("setup", 0) == 0)
  {
    sim->initialize();
    ad->reset_virial();
    sim->setup();
  }
  else if (mode.rfind("mrun", 0) == 0)
  {
    sim->run();
  }

  FILE *f = std::fopen(argv[3], "w");
  const auto &own = ad->atom_struct_owned;
  const auto &gh = ad->atom_struct_ghost;
  std::fprintf(f, "N %zu\n", own.position.size());
  dump_vecs(f, "pos", own.position);
  dump_vecs(f, "vel", own.velocity);
  dump_vecs(f, "acc", own.acceleration);
  dump_vecs(f, "ghost_pos", gh.position);
  std::fprintf(f, "ghost_id %zu\n", gh.id.size());
  for (auto id : gh.id)
    std::fprintf(f, "%u\n", (unsigned)id);
  std::fprintf(f, "virial %.17g\n", ad->virialForce);

  if (!sim->neighborlist.empty())
  {
    auto *nl = sim->neighborlist[0];
    size_t np = 0;
    for (const auto &l : nl->neighlist)
      np += l.size();
    std::fprintf(f, "pairs %zu\n", np);
    for (size_t i = 0; i < nl->neighlist.size(); ++i)
      for (auto j : nl->neighlist[i])
        std::fprintf(f, "%zu %u\n", i, j);
    std::fprintf(f, "no_bins %d %d %d\n", nl->no_bins.x, nl->no_bins.y, nl->no_bins.z);
  }

  if (mode.find("+thermo") != std::string::npos)
  {
    std::fprintf(f, "kinetic_energy %.17g\n", ad->kinetic_energy());
    std::fprintf(f, "temperature %.17g\n", ad->temperature());
    std::fprintf(f, "pressure %.17g\n", ad->pressure());
  }

  for (auto *ff : sim->force_field)
  {
    auto *er = dynamic_cast<caviar::force_field::Electrostatic_ewald_r *>(ff);
    auto *ek = dynamic_cast<caviar::force_field::Electrostatic_ewald_k *>(ff);
    auto *elc = dynamic_cast<caviar::force_field::Electrostatic_ewald_slab_correction *>(ff);
    if (elc && mode.find("+elc_energy") != std::string::npos)
      std::fprintf(f, "energy_elc %.17g\n", elc->energy());
    if (elc && mode.find("+field") != std::string::npos)
    {
      std::fprintf(f, "field_elc %zu\n", own.position.size());
      for (const auto &p : own.position)
      {
        Vector<double> r{p.x + 0.1, p.y - 0.2, p.z + 0.3};
        auto e = elc->field(r);
        double phi = elc->potential(r);
        std::fprintf(f, "%.17g %.17g %.17g %.17g\n", e.x, e.y, e.z, phi);
      }
    }
    if (er && mode.find("+er_energy") != std::string::npos)
      std::fprintf(f, "energy_r %.17g\n", er->energy());
    if (ek && mode.find("+ek_energy") != std::string::npos)
    {
      std::fprintf(f, "energy_k %.17g\n", ek->energy());
      std::fprintf(f, "energy_k_self %.17g\n", ek->self_energy());
    }
    if (ek && mode.find("+sk") != std::string::npos)
    {
      std::fprintf(f, "sk %d\n", ek->n_k_vectors);
      for (int k = 0; k < ek->n_k_vectors; ++k)
        std::fprintf(f, "%.17g %.17g\n", ek->potential_k_coef_cmplx[k].real(),
                     ek->potential_k_coef_cmplx[k].imag());
    }
    if (mode.find("+field") != std::string::npos && (er || ek))
    {
      // point queries at owned positions shifted by (0.1, -0.2, 0.3)
      std::fprintf(f, "%s %zu\n", er ? "field_r" : "field_k", own.position.size());
      for (const auto &p : own.position)
      {
        Vector<double> r{p.x + 0.1, p.y - 0.2, p.z + 0.3};
        auto e = ff->field(r);
        double phi = ff->potential(r);
        std::fprintf(f, "%.17g %.17g %.17g %.17g\n", e.x, e.y, e.z, phi);
      }
    }
  }
  std::fclose(f);
  return 0;
}
