// Run-time force-field / control / pair_style readers (see ffield.hpp for the contract and the reference lines followed).
#include "ffield.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>

namespace rxa {
namespace {

// Reads the file line by line as LAMMPS' TextFileReader does for ReaxFF (ignore_comments = false): a '!' comment is just more
// words; value lines skip blank lines; a "comment line" consumes exactly one raw line, blank or not.
class TextScanner {
 public:
  explicit TextScanner(const std::string& t) : text_(t) {}
  struct EndOfFile {};

  bool raw_line(std::string& out) {
    if (pos_ >= text_.size()) return false;
    size_t end = text_.find('\n', pos_);
    if (end == std::string::npos) end = text_.size();
    out.assign(text_, pos_, end - pos_);
    pos_ = std::min(text_.size(), end + 1);
    ++lineno_;
    return true;
  }
  bool words_opt(std::vector<std::string>& w) {
    std::string l;
    while (raw_line(l)) {
      split(l, w);
      if (!w.empty()) return true;
    }
    return false;
  }
  std::vector<std::string> words() {
    std::vector<std::string> w;
    if (!words_opt(w)) throw EndOfFile{};
    return w;
  }
  void comment_line() {
    std::string l;
    if (!raw_line(l)) throw EndOfFile{};
  }
  std::string at() const { return "line " + std::to_string(lineno_); }

  static void split(const std::string& l, std::vector<std::string>& w) {
    w.clear();
    std::string cur;
    for (char c : l) {
      if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\f') {
        if (!cur.empty()) { w.push_back(cur); cur.clear(); }
      } else {
        cur.push_back(c);
      }
    }
    if (!cur.empty()) w.push_back(cur);
  }

 private:
  const std::string& text_;
  size_t pos_ = 0;
  int lineno_ = 0;
};

bool parse_real(const std::string& s, double& v) {
  char* e = nullptr;
  v = std::strtod(s.c_str(), &e);
  return e != s.c_str() && (size_t)(e - s.c_str()) == s.size();
}

// consumes the words of one line left to right
class Fields {
 public:
  Fields(std::vector<std::string> w, std::string where) : w_(std::move(w)), where_(std::move(where)) {}
  size_t size() const { return w_.size(); }
  bool left() const { return k_ < w_.size(); }
  void need(size_t n) const {
    if (w_.size() < n) throw FfieldError(where_ + ": expected " + std::to_string(n) + " columns, found " + std::to_string(w_.size()));
  }
  const std::string& word() {
    if (k_ >= w_.size()) throw FfieldError(where_ + ": missing value");
    return w_[k_++];
  }
  double real() {
    const std::string& s = word();
    double v;
    if (!parse_real(s, v)) throw FfieldError(where_ + ": '" + s + "' is not a number");
    if (!std::isfinite(v)) throw FfieldError(where_ + ": '" + s + "' is not finite");
    return v;
  }
  int integer() {
    const std::string& s = word();
    size_t p = 0;
    long long v = 0;
    bool neg = false;
    if (p < s.size() && (s[p] == '+' || s[p] == '-')) neg = s[p++] == '-';
    if (p == s.size()) throw FfieldError(where_ + ": '" + s + "' is not an integer");
    for (; p < s.size(); ++p) {
      if (!std::isdigit((unsigned char)s[p])) throw FfieldError(where_ + ": '" + s + "' is not an integer");
      v = 10 * v + (s[p] - '0');
      if (v > (1LL << 31)) throw FfieldError(where_ + ": integer out of range");
    }
    v = neg ? -v : v;
    if (v > 2147483647LL) throw FfieldError(where_ + ": integer out of range");
    return (int)v;
  }

 private:
  std::vector<std::string> w_;
  std::string where_;
  size_t k_ = 0;
};

std::string symbol3(const std::string& s) {
  std::string r;
  for (char c : s) {
    if (r.size() == 3) break;
    r.push_back((char)std::toupper((unsigned char)c));
  }
  return r;
}

std::string lower(const std::string& s) {
  std::string r;
  for (char c : s) r.push_back((char)std::tolower((unsigned char)c));
  return r;
}

void read_atom(TextScanner& in, Species& s, bool lg, int& vdw_type, std::vector<std::string>& warnings) {
  Fields a(in.words(), in.at());
  if (a.size() < 8 && !lg) throw FfieldError(in.at() + ": this force field file requires lgvdw yes");
  a.need(9);
  s.name = symbol3(a.word());
  s.r_s = a.real(); s.valency = a.real(); s.mass = a.real(); s.r_vdw = a.real(); s.epsilon = a.real(); s.gamma = a.real();
  s.r_pi = a.real(); s.valency_e = a.real();
  s.nlp_opt = 0.5 * (s.valency_e - s.valency);
  Fields b(in.words(), in.at());
  b.need(8);
  s.alpha = b.real(); s.gamma_w = b.real(); s.valency_boc = b.real(); s.p_ovun5 = b.real(); b.word();
  s.chi = b.real(); s.eta = 2.0 * b.real(); s.p_hbond = (int)b.real();
  Fields c(in.words(), in.at());
  c.need(8);
  s.r_pi_pi = c.real(); s.p_lp2 = c.real(); c.word(); s.b_o_131 = c.real(); s.b_o_132 = c.real(); s.b_o_133 = c.real(); s.bcut_acks2 = c.real();
  Fields d(in.words(), in.at());
  d.need(8);
  s.p_ovun2 = d.real(); s.p_val3 = d.real(); d.word(); s.valency_val = d.real(); s.p_val5 = d.real(); s.rcore2 = d.real(); s.ecore2 = d.real(); s.acore2 = d.real();
  if (lg) {   // fifth line: lgcij lgre; it must start with a number (otherwise it is the next element)
    Fields e(in.words(), in.at());
    double probe;
    if (!parse_real(e.word(), probe)) throw FfieldError(in.at() + ": force field is not compatible with lgvdw yes");
    e.need(2);
    s.lgcij = probe;
    s.lgre = e.real();
  }
  // van der Waals variant: decided by the first element; a conflicting later element only warns
  const bool inner = s.rcore2 > 0.01 && s.acore2 > 0.01, shield = s.gamma_w > 0.5;
  const int mine = inner ? (shield ? 3 : 2) : (shield ? 1 : 0);
  if (!mine) throw FfieldError("inconsistent van der Waals parameters: no shielding or inner wall for element " + s.name);
  if (vdw_type && vdw_type != mine) warnings.push_back("van der Waals method of element " + s.name + " differs from earlier elements; keeping the earlier one");
  else vdw_type = mine;
}

}  // namespace

int ForceField::match_element(const std::string& el) const {
  int hit = -1, n = 0;
  for (int i = 0; i < ntypes(); ++i)
    if (lower(sp[i].name) == lower(el)) { hit = i; ++n; }
  return n == 1 ? hit : (n == 0 ? -1 : -2);
}

void ForceField::require_bond_blocks(const std::vector<int>& types) const {
  for (size_t a = 0; a < types.size(); ++a)
    for (size_t b = a; b < types.size(); ++b)
      if (types[a] >= 0 && types[b] >= 0 && !P(types[a], types[b]).listed)
        throw FfieldError("force field has no bond-parameter block for the element pair " + sp[types[a]].name + "-" + sp[types[b]].name +
                          " (LAMMPS would zero-fill it and create a phantom bond of order 1 at every distance up to the bond cutoff)");
}

ForceField read_ffield(const std::string& text, bool lg) {
  ForceField ff;
  TextScanner in(text);
  try {
    {
      std::vector<std::string> w = in.words();
      std::string joined;
      for (auto& x : w) joined += x + " ";
      const bool digits = std::all_of(w[0].begin(), w[0].end(), [](char c) { return c >= '0' && c <= '9'; });
      if (digits && w.size() >= 2 && w[1][0] == '!' && joined.find("general parameters") != std::string::npos)
        throw FfieldError("first line of the ReaxFF potential file must be a comment or empty");
    }
    {
      Fields f(in.words(), in.at());
      const int ng = f.integer();
      if (ng < 38) throw FfieldError("the force field has " + std::to_string(ng) + " general parameters; at least 38 are needed");
      ff.gp.resize(ng);
      for (double& g : ff.gp) { Fields v(in.words(), in.at()); g = v.real(); }
    }
    int nt;
    {
      Fields f(in.words(), in.at());
      nt = f.integer();
      if (nt < 1) throw FfieldError("invalid number of atom types");
      in.comment_line(); in.comment_line(); in.comment_line();
    }
    ff.sp.resize(nt);
    ff.pair.assign((size_t)nt * nt, PairParams{});
    for (int i = 0; i < nt; ++i) read_atom(in, ff.sp[i], lg, ff.vdw_type, ff.warnings);
    for (Species& s : ff.sp) {
      s.c2_species = s.name == "C";
      s.heavy = s.mass > 21.0;
      if (s.mass < 21 && s.valency_val != s.valency_boc) {
        ff.warnings.push_back("Changed valency_val to valency_boc for " + s.name);
        s.valency_val = s.valency_boc;
      }
    }
    {
      Fields f(in.words(), in.at());
      const int nb = f.integer();
      in.comment_line();
      for (int b = 0; b < nb; ++b) {
        Fields l1(in.words(), in.at());
        l1.need(10);
        const int i = l1.integer() - 1, j = l1.integer() - 1;
        if (i < 0 || j < 0) throw FfieldError(in.at() + ": inconsistent force field file");
        const bool use = i < nt && j < nt;
        PairParams t;
        if (use) {
          t.De_s = l1.real(); t.De_p = l1.real(); t.De_pp = l1.real(); t.p_be1 = l1.real(); t.p_bo5 = l1.real(); t.v13cor = l1.real();
          t.p_bo6 = l1.real(); t.p_ovun1 = l1.real();
        }
        Fields l2(in.words(), in.at());
        l2.need(7);
        if (!use) { ff.warnings.push_back(in.at() + ": bond block beyond the number of atom types ignored"); continue; }
        t.p_be2 = l2.real(); t.p_bo3 = l2.real(); t.p_bo4 = l2.real(); l2.word(); t.p_bo1 = l2.real(); t.p_bo2 = l2.real();
        t.ovc = l2.left() ? l2.real() : 0.0;
        for (PairParams* p : {&ff.P(i, j), &ff.P(j, i)}) {
          p->De_s = t.De_s; p->De_p = t.De_p; p->De_pp = t.De_pp; p->p_be1 = t.p_be1; p->p_bo5 = t.p_bo5; p->v13cor = t.v13cor;
          p->p_bo6 = t.p_bo6; p->p_ovun1 = t.p_ovun1; p->p_be2 = t.p_be2; p->p_bo3 = t.p_bo3; p->p_bo4 = t.p_bo4; p->p_bo1 = t.p_bo1;
          p->p_bo2 = t.p_bo2; p->ovc = t.ovc; p->listed = true;
        }
      }
    }
    // mixing rules (after the bond block, before the off-diagonal overrides)
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) {
        const Species &a = ff.sp[i], &b = ff.sp[j];
        for (PairParams* p : {&ff.P(i, j), &ff.P(j, i)}) {
          p->r_s = 0.5 * (b.r_s + a.r_s);
          p->r_p = 0.5 * (b.r_pi + a.r_pi);
          p->r_pp = 0.5 * (b.r_pi_pi + a.r_pi_pi);
          p->p_boc3 = std::sqrt(b.b_o_132 * a.b_o_132);
          p->p_boc4 = std::sqrt(b.b_o_131 * a.b_o_131);
          p->p_boc5 = std::sqrt(b.b_o_133 * a.b_o_133);
          p->D = std::sqrt(b.epsilon * a.epsilon);
          p->alpha = std::sqrt(b.alpha * a.alpha);
          p->r_vdW = 2.0 * std::sqrt(b.r_vdw * a.r_vdw);
          p->gamma_w = std::sqrt(b.gamma_w * a.gamma_w);
          p->gamma = std::pow(b.gamma * a.gamma, -1.5);
          p->rcore = std::sqrt(a.rcore2 * b.rcore2);
          p->ecore = std::sqrt(a.ecore2 * b.ecore2);
          p->acore = std::sqrt(a.acore2 * b.acore2);
          p->lgcij = std::sqrt(a.lgcij * b.lgcij);
          p->lgre = 2.0 * ff.gp[35] * std::sqrt(a.lgre * b.lgre);
        }
      }
    {
      Fields f(in.words(), in.at());
      const int no = f.integer();
      for (int o = 0; o < no; ++o) {
        Fields l(in.words(), in.at());
        l.need(lg ? 9 : 8);
        const int i = l.integer() - 1, j = l.integer() - 1;
        if (i < 0 || j < 0) throw FfieldError(in.at() + ": inconsistent force field file");
        if (i >= nt || j >= nt) continue;
        const double D = l.real(), rv = l.real(), al = l.real(), rs = l.real(), rp = l.real(), rpp = l.real();
        for (PairParams* p : {&ff.P(i, j), &ff.P(j, i)}) {   // only positive values override the mixed ones
          if (D > 0) p->D = D;
          if (rv > 0) p->r_vdW = 2.0 * rv;
          if (al > 0) p->alpha = al;
          if (rs > 0) p->r_s = rs;
          if (rp > 0) p->r_p = rp;
          if (rpp > 0) p->r_pp = rpp;
        }
        if (lg) {
          const double v = l.real();
          if (v >= 0) { ff.P(i, j).lgcij = v; ff.P(j, i).lgcij = v; }
        }
      }
    }
    {
      Fields f(in.words(), in.at());
      const int na = f.integer();
      std::map<std::array<int, 3>, int> used;   // slot counter of the reference: both orientations, twice for j == l
      for (int a = 0; a < na; ++a) {
        Fields l(in.words(), in.at());
        l.need(10);
        const int i = l.integer() - 1, j = l.integer() - 1, k = l.integer() - 1;
        if (i < 0 || j < 0 || k < 0) throw FfieldError(in.at() + ": inconsistent force field file");
        if (i >= nt || j >= nt || k >= nt) continue;
        AngleSet s;
        s.theta_00 = l.real(); s.p_val1 = l.real(); s.p_val2 = l.real(); s.p_coa1 = l.real(); s.p_val7 = l.real(); s.p_pen1 = l.real(); s.p_val4 = l.real();
        const int u1 = ++used[{i, j, k}], u2 = ++used[{k, j, i}];
        if (u1 > 5 || u2 > 5)
          throw FfieldError(in.at() + ": too many valence-angle parameter sets for one triple (the reference array holds 5 slots)");
        ff.angles[{i, j, k}].push_back(s);
        if (i != k) ff.angles[{k, j, i}].push_back(s);
      }
    }
    {
      Fields f(in.words(), in.at());
      const int nq = f.integer();
      std::set<std::array<int, 4>> explicit_entry;
      for (int t = 0; t < nq; ++t) {
        Fields l(in.words(), in.at());
        l.need(9);
        const int i = l.integer() - 1, j = l.integer() - 1, k = l.integer() - 1, m = l.integer() - 1;
        if (i < -1 || j < 0 || k < 0 || m < -1) throw FfieldError(in.at() + ": inconsistent force field file");
        TorsionSet s;
        s.V1 = l.real(); s.V2 = l.real(); s.V3 = l.real(); s.p_tor1 = l.real(); s.p_cot1 = l.real();
        if (i >= 0 && m >= 0) {
          if (i < nt && j < nt && k < nt && m < nt) {
            explicit_entry.insert({i, j, k, m});
            explicit_entry.insert({m, k, j, i});
            ff.torsions[{i, j, k, m}] = s;
            ff.torsions[{m, k, j, i}] = s;
          }
        } else if (j < nt && k < nt) {   // compact 0-X-Y-0: fills every outer pair not set explicitly; later compact lines win
          for (int p = 0; p < nt; ++p)
            for (int o = 0; o < nt; ++o) {
              if (!explicit_entry.count({p, j, k, o})) ff.torsions[{p, j, k, o}] = s;
              if (!explicit_entry.count({o, k, j, p})) ff.torsions[{o, k, j, p}] = s;
            }
        }
      }
    }
    {
      std::vector<std::string> w;
      if (!in.words_opt(w)) {
        ff.warnings.push_back("ReaxFF parameter file has no hydrogen bond parameters");
      } else {
        Fields f(w, in.at());
        const int nh = f.integer();
        for (int h = 0; h < nh; ++h) {
          Fields l(in.words(), in.at());
          l.need(7);
          const int i = l.integer() - 1, j = l.integer() - 1, k = l.integer() - 1;
          if (i < 0 || j < 0 || k < 0) throw FfieldError(in.at() + ": inconsistent force field file");
          if (i >= nt || j >= nt || k >= nt) continue;
          HBondSet s;
          s.r0 = l.real(); s.p1 = l.real(); s.p2 = l.real(); s.p3 = l.real();
          ff.hbonds[{i, j, k}] = s;
        }
      }
    }
  } catch (TextScanner::EndOfFile&) {
    throw FfieldError("unexpected end of the force field file at " + in.at() + " (strict mode)");
  }
  ff.bo_cut = 0.01 * ff.gp[29];
  ff.nonb_low = ff.gp[11];
  ff.nonb_cut = ff.gp[12];
  const double g37 = ff.gp[37];
  if (!(g37 > -2147483649.0 && g37 < 2147483648.0)) throw FfieldError("general parameter 38 is not representable as int");
  ff.gp37 = (double)(int)g37 == 2;
  ff.c2_active = ff.gp[5] > 0.001;
  for (int i = 0; i < ff.ntypes(); ++i)
    for (int j = 0; j < ff.ntypes(); ++j) {
      const Species &a = ff.sp[i], &b = ff.sp[j];
      ff.P(i, j).triple_stab = ff.gp37 || (a.mass == 12.0000 && b.mass == 15.9990) || (b.mass == 12.0000 && a.mass == 15.9990);
    }
  return ff;
}

Control read_control(const std::string& text, std::vector<std::string>* notices) {
  static const std::set<std::string> inactive = {
      "ensemble_type", "nsteps", "dt", "proc_by_dim", "random_vel", "restart_format", "restart_freq", "reposition_atoms",
      "restrict_bonds", "remove_CoM_vel", "debug_level", "reneighbor", "vlist_buffer", "ghost_cutoff", "qeq_freq", "q_err",
      "ilu_refactor", "ilu_droptol", "temp_init", "temp_final", "t_mass", "t_mode", "t_rate", "t_freq", "pressure", "p_mass",
      "pt_mass", "compress", "press_mode", "geo_format", "traj_compress", "traj_method", "molecular_analysis", "ignore",
      "dipole_anal", "freq_dipole_anal", "diffusion_coef", "freq_diffusion_coef", "restrict_type", "traj_title",
      "simulation_name", "energy_update_freq", "atom_info", "atom_velocities", "atom_forces", "bond_info", "angle_info"};
  Control c;
  TextScanner in(text);
  std::vector<std::string> w;
  while (in.words_opt(w)) {
    Fields f(w, in.at());
    const std::string key = f.word();
    if (!f.left()) throw FfieldError(in.at() + ": no value(s) for control parameter: " + key);
    if (inactive.count(key)) { if (notices) notices->push_back("Ignoring inactive control parameter: " + key); continue; }
    if (key == "nbrhood_cutoff") c.bond_cut = f.real();
    else if (key == "bond_graph_cutoff") c.bg_cut = f.real();
    else if (key == "thb_cutoff") c.thb_cut = f.real();
    else if (key == "thb_cutoff_sq") c.thb_cutsq = f.real();
    else if (key == "hbond_cutoff") c.hbond_cut = f.real();
    else if (key == "tabulate_long_range") c.tabulate = f.integer();
    else if (key == "write_freq") { if (f.integer() > 0 && notices) notices->push_back("native trajectory writing is not supported"); }
    else throw FfieldError(in.at() + ": unknown parameter " + key + " in control file");
  }
  return c;
}

std::string table_dump(const ForceField& ff, bool portable) {
  std::string o = "FFIELD-TABLES 1\n";
  char buf[48];
  auto num = [&](double v) { std::snprintf(buf, sizeof buf, " %.17g", v); o += buf; };
  o += "GP " + std::to_string(ff.gp.size()) + " " + std::to_string(ff.vdw_type) + "\n";
  for (size_t i = 0; i < ff.gp.size(); ++i) { o += "G " + std::to_string(i); num(ff.gp[i]); o += "\n"; }
  const int n = ff.ntypes();
  for (int i = 0; i < n; ++i) {
    const Species& s = ff.sp[i];
    o += "S " + std::to_string(i) + " " + s.name;
    const double head[] = {s.r_s, s.valency, s.mass, s.r_vdw, s.epsilon, s.gamma, s.r_pi, s.valency_e, s.nlp_opt, s.alpha, s.gamma_w, s.valency_boc, s.p_ovun5};
    for (double v : head) num(v);
    o += " " + std::to_string(s.p_hbond);
    const double tail[] = {s.chi, s.eta, s.r_pi_pi, s.p_lp2, s.b_o_131, s.b_o_132, s.b_o_133, s.bcut_acks2, s.p_ovun2, s.p_val3, s.valency_val, s.p_val5,
                           s.rcore2, s.ecore2, s.acore2, s.lgcij, s.lgre};
    for (double v : tail) num(v);
    o += "\n";
  }
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const PairParams& t = ff.P(i, j);
      o += "T " + std::to_string(i) + " " + std::to_string(j);
      const double v1[] = {t.p_bo1, t.p_bo2, t.p_bo3, t.p_bo4, t.p_bo5, t.p_bo6, t.r_s, t.r_p, t.r_pp, t.p_boc3, t.p_boc4, t.p_boc5, t.p_be1, t.p_be2,
                           t.De_s, t.De_p, t.De_pp, t.p_ovun1, t.D, t.alpha, t.r_vdW, t.gamma_w, t.rcore, t.ecore, t.acore, t.lgcij, t.lgre};
      for (double v : v1) num(v);
      if (!portable) num(t.gamma);
      num(t.v13cor);
      num(t.ovc);
      o += "\n";
    }
  for (const auto& kv : ff.angles) {
    o += "H3 " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]) + " " + std::to_string(kv.second.size());
    for (const AngleSet& s : kv.second) {
      const double v[] = {s.theta_00, s.p_val1, s.p_val2, s.p_val4, s.p_val7, s.p_pen1, s.p_coa1};
      for (double x : v) num(x);
    }
    o += "\n";
  }
  for (const auto& kv : ff.torsions) {
    o += "H4 " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]) + " " + std::to_string(kv.first[3]);
    const double v[] = {kv.second.V1, kv.second.V2, kv.second.V3, kv.second.p_tor1, kv.second.p_cot1};
    for (double x : v) num(x);
    o += "\n";
  }
  for (const auto& kv : ff.hbonds) {
    const HBondSet& h = kv.second;
    if (h.r0 == -1 && h.p1 == 0 && h.p2 == 0 && h.p3 == 0) continue;
    o += "HB " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]);
    const double v[] = {h.r0, h.p1, h.p2, h.p3};
    for (double x : v) num(x);
    o += "\n";
  }
  return o;
}

namespace {
bool yesno(const std::string& s, const std::string& key) {
  if (s == "yes" || s == "on" || s == "true" || s == "1") return true;
  if (s == "no" || s == "off" || s == "false" || s == "0") return false;
  throw FfieldError("Expected boolean parameter instead of '" + s + "' for pair_style reaxff/metal keyword " + key);
}
}  // namespace

std::string PairSettings::effective_precision() const {
  if (!precision.empty()) return precision;
  if (backend == "cpu64") return "fp64";
  return bonded == "cpu64" ? "mixed" : "fp32";
}

PairSettings parse_pair_style(const std::vector<std::string>& args, const std::string& control_text) {
  if (args.empty()) throw FfieldError("Illegal pair_style command");
  PairSettings p;
  if (args[0] != "NULL") {
    p.control_file = args[0];
    p.control = read_control(control_text, &p.notices);
  }
  for (size_t i = 1; i < args.size(); i += 2) {
    const std::string& key = args[i];
    if (i + 1 >= args.size()) throw FfieldError("Illegal pair_style reaxff/metal command: keyword '" + key + "' needs a value");
    const std::string& v = args[i + 1];
    auto integer = [&](const std::string& s) {
      char* e = nullptr;
      long x = std::strtol(s.c_str(), &e, 10);
      if (s.empty() || *e) throw FfieldError("Expected integer parameter instead of '" + s + "' for pair_style reaxff/metal keyword " + key);
      return (int)x;
    };
    if (key == "checkqeq") p.checkqeq = yesno(v, key);
    else if (key == "enobonds") p.enobonds = yesno(v, key);
    else if (key == "lgvdw") p.lgvdw = yesno(v, key);
    else if (key == "shellcheck") p.shellcheck = yesno(v, key);
    else if (key == "reaxmetal_selfcheck") p.selfcheck = yesno(v, key);
    else if (key == "backend") {
      if (v != "cpu64" && v != "gpu" && v != "metal") throw FfieldError("Illegal pair_style reaxff/metal command: backend must be cpu64 or gpu, got '" + v + "'");
      p.backend = v == "metal" ? "gpu" : v;
    } else if (key == "bonded") {
      if (v != "cpu64" && v != "gpu" && v != "metal") throw FfieldError("Illegal pair_style reaxff/metal command: bonded must be cpu64 or gpu, got '" + v + "'");
      p.bonded = v == "metal" ? "gpu" : v;
    } else if (key == "precision") {
      if (v != "fp64" && v != "fp32" && v != "mixed") throw FfieldError("Illegal pair_style reaxff/metal command: precision must be fp64, fp32 or mixed");
      p.precision = v;
    } else if (key == "safezone") {
      char* e = nullptr;
      p.safezone = std::strtod(v.c_str(), &e);
      if (v.empty() || *e || p.safezone < 0) throw FfieldError("Illegal pair_style reaxff/metal safezone command");
      p.notices.push_back("pair_style reaxff/metal: 'safezone' sizes stock matrices; the engine sizes its own lists");
    } else if (key == "mincap") {
      p.mincap = integer(v);
      if (p.mincap < 0) throw FfieldError("Illegal pair_style reaxff/metal mincap command");
      p.notices.push_back("pair_style reaxff/metal: 'mincap' sizes stock matrices; the engine sizes its own lists");
    } else if (key == "minhbonds") {
      p.minhbonds = integer(v);
      if (p.minhbonds < 0) throw FfieldError("Illegal pair_style reaxff/metal minhbonds command");
      p.notices.push_back("pair_style reaxff/metal: 'minhbonds' sizes stock matrices; the engine sizes its own lists");
    } else if (key == "list/blocking") {
      p.list_blocking = yesno(v, key);
      p.notices.push_back("pair_style reaxff/metal: 'list/blocking' is a Kokkos option and is ignored");
    } else if (key == "tabulate") {
      p.tabulate = integer(v);
      if (p.tabulate < 0) throw FfieldError("Illegal pair_style reaxff/metal tabulate command");
    } else {
      throw FfieldError("Illegal pair_style reaxff/metal command: unknown keyword '" + key + "'");
    }
  }
  if (p.tabulate > 0 || p.control.tabulate > 0)
    p.notices.push_back("pair_style reaxff/metal: 'tabulate' is accepted but the non-bonded terms are evaluated analytically");
  return p;
}

}  // namespace rxa
