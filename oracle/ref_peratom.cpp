// Per-atom energies and virials of the reference's own FP64 CPU engine, for golden vectors (test infrastructure only: built by
// oracle/ref.mk into oracle/_ref/, run in the build container by tests/golden/make_peratom_golden.py).
//
// Input: the case text of tests/golden/make_golden.py (origin / a / b / c / periodic / atoms N + "tag type x y z" lines /
// charges ...). Output: "eatom tag e" and "vatom tag xx yy zz xy xz yz" per owned atom, the bonded and non-bonded tallies summed
// and every image folded onto its owner (what the LAMMPS adapter's reverse communication does); then the bond table the LAMMPS
// analysis commands read (BondedOptions::bond_table): "bondatom tag total_bo nlp" per owned atom and "bond tag partner_tag bo"
// per bond of an owned atom, partners named by their atom id (an image by its owner's).
// The engine calls are the reference's public interface (include/reaxmetal/bonded.hpp:28, nonbonded.hpp:22 per_atom options;
// system.hpp image expansion; neighbor.hpp far list), used the way tools/energy_tool.cpp:50-121 drives it.
#include <array>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "reaxmetal/bonded.hpp"
#include "reaxmetal/nonbonded.hpp"

using namespace reaxmetal;

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: ref_peratom FFIELD ELEMENTS(comma list) CASE.txt\n");
    return 2;
  }
  try {
    const ForceField ff = read_force_field_file(argv[1], FfieldOptions{});
    std::vector<std::string> elements;
    {
      std::istringstream es(argv[2]);
      std::string e;
      while (std::getline(es, e, ',')) elements.push_back(e);
    }
    ControlParams ctl;
    NeighborCutoffs cut;
    cut.nonb = ff.file_control().nonb_cut;
    cut.bond = ctl.bond_cut;
    cut.hbond = ctl.hbond_cut;
    Box box;
    std::vector<double> x, q;
    std::vector<int> type;
    std::vector<std::int64_t> tag;
    std::ifstream in(argv[3]);
    std::string word;
    while (in >> word) {
      if (word == "origin") {
        in >> box.origin[0] >> box.origin[1] >> box.origin[2];
      } else if (word == "a" || word == "b" || word == "c") {
        auto& v = box.vectors[static_cast<std::size_t>(word[0] - 'a')];
        in >> v[0] >> v[1] >> v[2];
      } else if (word == "periodic") {
        for (std::size_t d = 0; d < 3; ++d) { int p; in >> p; box.periodic[d] = p != 0; }
      } else if (word == "atoms") {
        std::size_t n = 0;
        in >> n;
        for (std::size_t k = 0; k < n; ++k) {
          long long t;
          int ty;
          double r[3];
          in >> t >> ty >> r[0] >> r[1] >> r[2];
          const auto m = ff.match_element(elements.at(static_cast<std::size_t>(ty - 1)));
          if (m.size() != 1) throw std::runtime_error("element not unique in the force field");
          tag.push_back(t);
          type.push_back(m[0]);
          x.insert(x.end(), r, r + 3);
        }
      } else if (word == "charges") {
        double v;
        while (q.size() < tag.size() && (in >> v)) q.push_back(v);
      }
    }
    if (q.size() != tag.size()) throw std::runtime_error("the case needs one charge per atom");
    ExpandOptions eo;
    eo.shell = cut.required_shell();
    const AtomSet a = expand_images(box, x, type, tag, eo);
    a.validate(box);
    const FarList far = build_far_list(a, cut);
    BondedOptions bo;
    bo.per_atom = true;
    bo.bond_table = true;
    const BondedResult rb = compute_bonded_core(ff, ctl, a, far, bo);
    NonbondedOptions no;
    no.per_atom = true;
    const NonbondedResult rn = compute_nonbonded_core(ff, cut, a, far, q, no);
    const std::size_t nl = a.nlocal;
    std::vector<double> e(nl, 0.0);
    std::vector<std::array<double, 6>> v(nl, std::array<double, 6>{});
    for (std::size_t i = 0; i < a.nall(); ++i) {
      const std::size_t o = static_cast<std::size_t>(a.owner[i]);
      e[o] += rb.eatom[i] + rn.eatom[i];
      for (std::size_t c = 0; c < 6; ++c) v[o][c] += rb.vatom[i][c] + rn.vatom[i][c];
    }
    for (std::size_t i = 0; i < nl; ++i) {
      std::printf("eatom %lld %.17g\n", static_cast<long long>(a.tag[i]), e[i]);
      std::printf("vatom %lld %.17g %.17g %.17g %.17g %.17g %.17g\n", static_cast<long long>(a.tag[i]), v[i][0], v[i][1], v[i][2],
                  v[i][3], v[i][4], v[i][5]);
    }
    const BondTable& t = rb.table;
    for (std::size_t i = 0; i < nl; ++i) {
      std::printf("bondatom %lld %.17g %.17g\n", static_cast<long long>(a.tag[i]), t.total_bo[i], t.nlp[i]);
      for (std::int32_t k = t.start[i]; k < t.start[i + 1]; ++k)
        std::printf("bond %lld %lld %.17g\n", static_cast<long long>(a.tag[i]),
                    static_cast<long long>(a.tag[static_cast<std::size_t>(t.nbr[static_cast<std::size_t>(k)])]), t.bo[static_cast<std::size_t>(k)]);
    }
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
