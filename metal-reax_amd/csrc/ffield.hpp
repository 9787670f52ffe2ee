// Run-time ReaxFF force-field reader, control-file reader and pair_style keyword parser of the MI355X engine.
// Host C++ only. The accepted formats and every derived value follow the reference's reader (yeatanir/Metal-Reax
// src/io/ffield_parser.cpp:91-451, src/io/control_parser.cpp:14-54, src/io/pair_settings.cpp:41-91), which itself follows
// pinned LAMMPS reaxff_ffield.cpp / reaxff_control.cpp. Tables are then packed into the flat, type-indexed arrays the
// kernels read (ParamBlob, device_params.hpp); no kernel ever sees an element name or a mass.
#pragma once
#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace rxa {

struct FfieldError : std::runtime_error { using std::runtime_error::runtime_error; };

struct Species {
  std::string name;   // upper-cased, at most 3 characters
  double r_s = 0, valency = 0, mass = 0, r_vdw = 0, epsilon = 0, gamma = 0, r_pi = 0, valency_e = 0, nlp_opt = 0;
  double alpha = 0, gamma_w = 0, valency_boc = 0, p_ovun5 = 0, chi = 0, eta = 0;   // eta stored as 2 x the file value
  int p_hbond = 0;
  double r_pi_pi = 0, p_lp2 = 0, b_o_131 = 0, b_o_132 = 0, b_o_133 = 0, bcut_acks2 = 0;
  double p_ovun2 = 0, p_val3 = 0, valency_val = 0, p_val5 = 0, rcore2 = 0, ecore2 = 0, acore2 = 0;
  double lgcij = 0, lgre = 0;
  bool c2_species = false, heavy = false;
};

struct PairParams {
  double p_bo1 = 0, p_bo2 = 0, p_bo3 = 0, p_bo4 = 0, p_bo5 = 0, p_bo6 = 0;
  double r_s = 0, r_p = 0, r_pp = 0, p_boc3 = 0, p_boc4 = 0, p_boc5 = 0;
  double p_be1 = 0, p_be2 = 0, De_s = 0, De_p = 0, De_pp = 0, p_ovun1 = 0;
  double D = 0, alpha = 0, r_vdW = 0, gamma_w = 0, rcore = 0, ecore = 0, acore = 0, lgcij = 0, lgre = 0;
  double gamma = 0;   // (gamma_i gamma_j)^-1.5
  double v13cor = 0, ovc = 0;
  bool listed = false, triple_stab = false;
};

struct AngleSet { double theta_00 = 0, p_val1 = 0, p_val2 = 0, p_val4 = 0, p_val7 = 0, p_pen1 = 0, p_coa1 = 0; };
struct TorsionSet { double V1 = 0, V2 = 0, V3 = 0, p_tor1 = 0, p_cot1 = 0; };
struct HBondSet { double r0 = -1, p1 = 0, p2 = 0, p3 = 0; };

struct Control {   // control-file values (defaults of the reference / LAMMPS)
  double bond_cut = 5.0, bg_cut = 0.3, thb_cut = 0.001, thb_cutsq = 0.00001, hbond_cut = 7.5;
  int tabulate = 0;
};

struct ForceField {
  std::vector<double> gp;
  int vdw_type = 0;
  std::vector<Species> sp;
  std::vector<PairParams> pair;   // n x n
  std::map<std::array<int, 3>, std::vector<AngleSet>> angles;
  std::map<std::array<int, 4>, TorsionSet> torsions;
  std::map<std::array<int, 3>, HBondSet> hbonds;
  double bo_cut = 0, nonb_low = 0, nonb_cut = 0;
  bool gp37 = false, c2_active = false;
  std::vector<std::string> warnings;
  int ntypes() const { return (int)sp.size(); }
  const PairParams& P(int i, int j) const { return pair[(size_t)i * sp.size() + j]; }
  PairParams& P(int i, int j) { return pair[(size_t)i * sp.size() + j]; }
  int match_element(const std::string& el) const;   // -1 none, -2 ambiguous
  void require_bond_blocks(const std::vector<int>& types) const;   // reference Q-12 strictness
};

ForceField read_ffield(const std::string& text, bool lgvdw);
Control read_control(const std::string& text, std::vector<std::string>* notices);
// text identical to the reference's canonical table dump (src/io/table_dump.cpp:20-84)
std::string table_dump(const ForceField& ff, bool portable);

// `pair_style reaxff/metal <control|NULL> [keyword value]...` (reference src/io/pair_settings.cpp:41-91). The GPU backend is spelled
// `gpu` (the reference's `metal` is accepted as its synonym); `backend cpu64` / `bonded cpu64` select the FP64 device precision of the
// whole step / of the bond-order chain.
struct PairSettings {
  std::string control_file;
  bool checkqeq = true, lgvdw = false, enobonds = true, shellcheck = true, list_blocking = false, selfcheck = false;
  double safezone = 1.2;
  int mincap = 50, minhbonds = 25, tabulate = 0;
  std::string backend = "cpu64", bonded = "gpu";
  std::string precision;   // "fp64" | "fp32" | "mixed": explicit precision keyword of this engine (overrides backend/bonded)
  Control control;
  std::vector<std::string> notices;
  std::string effective_precision() const;
};
PairSettings parse_pair_style(const std::vector<std::string>& args, const std::string& control_text);

}  // namespace rxa
