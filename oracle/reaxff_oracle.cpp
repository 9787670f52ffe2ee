// ReaxFF oracle: a serial FP64 CPU restatement of the reference's "cpu64" engine (yeatanir/Metal-Reax) and of the stock
// LAMMPS charge equilibration it is paired with. TEST INFRASTRUCTURE ONLY: it is the checker for the HIP engine in
// metal-reax_amd/ and is used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by the product.
//
// What is restated (reference paths are relative to the reference repository root):
//   * ffield / control reader ........ src/io/ffield_parser.cpp:91-451, src/io/control_parser.cpp:14-54, text_reader.hpp
//   * canonical table dump ........... src/io/table_dump.cpp:20-84 (same text format, so both dumps can be compared)
//   * box, image expander ............ src/core/system.cpp:15-59, src/neighbor/neighbor.cpp:18-88
//   * far list, owner-computes rule .. src/neighbor/neighbor.cpp:180-212, 295-310
//   * bonded core (10 terms + grad) .. src/cpu/bonded.cpp:86-800 with the term formulas of include/reaxmetal/terms.hpp:30-260
//   * nonbonded core ................. src/cpu/nonbonded.cpp:18-82, terms.hpp:262-325, terms_host.hpp:17-39
//   * charge equilibration ........... stock LAMMPS fix qeq/reaxff (LAMMPS stable_30Sep2026 @ 8de817dd, fix_qeq_reaxff.cpp:
//                                      compute_H / calculate_H / CG / calculate_Q / init_matvec), whose CG the reference keeps
//                                      line for line in plugin/adapter/fix_qeq_reaxff_metal.cpp:90-123
// The traversal orders are those of the reference (they decide which bond/angle/torsion is seen first and which end of a bond
// counts its energy); the code itself is written for this project. Built with -ffp-contract=off -fno-fast-math, as the
// reference builds its oracle (cmake/ReaxMetalFlags.cmake:16-22).
//
// Threads (orc_set_threads, default 1): the far list, the non-bonded pass, the charge matrix and its product run over OpenMP
// threads for the multi-core CPU baseline. Rows are independent, and every scatter goes into a per-thread buffer that is summed
// in thread order after a static schedule, so results repeat for a given thread count; at one thread every sum runs in the
// serial order, bit for bit the serial oracle (the tests use one thread).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

int g_threads = 1;   // OpenMP threads of the parallel loops (orc_set_threads)
inline int nthreads() { return g_threads; }
inline int thread_id() {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

struct Err : std::runtime_error { using std::runtime_error::runtime_error; };

// ------------------------------------------------------------------------------------------------------------------
// force field tables
// ------------------------------------------------------------------------------------------------------------------
struct Atom1 {   // per element (single body)
  std::string name;
  double r_s = 0, valency = 0, mass = 0, r_vdw = 0, epsilon = 0, gamma = 0, r_pi = 0, valency_e = 0, nlp_opt = 0;
  double alpha = 0, gamma_w = 0, valency_boc = 0, p_ovun5 = 0, chi = 0, eta = 0;
  int p_hbond = 0;
  double r_pi_pi = 0, p_lp2 = 0, b_o_131 = 0, b_o_132 = 0, b_o_133 = 0, bcut_acks2 = 0;
  double p_ovun2 = 0, p_val3 = 0, valency_val = 0, p_val5 = 0, rcore2 = 0, ecore2 = 0, acore2 = 0;
  double lgcij = 0, lgre = 0;
  bool is_c = false, heavy = false;
  double m12 = 0, m16 = 0;   // exact-mass flags of the C-O triple bond rule
};
struct Pair2 {
  double p_bo1 = 0, p_bo2 = 0, p_bo3 = 0, p_bo4 = 0, p_bo5 = 0, p_bo6 = 0;
  double r_s = 0, r_p = 0, r_pp = 0, p_boc3 = 0, p_boc4 = 0, p_boc5 = 0;
  double p_be1 = 0, p_be2 = 0, De_s = 0, De_p = 0, De_pp = 0, p_ovun1 = 0;
  double D = 0, alpha = 0, r_vdW = 0, gamma_w = 0, rcore = 0, ecore = 0, acore = 0, lgcij = 0, lgre = 0;
  double gamma = 0, v13cor = 0, ovc = 0;
  bool listed = false, tbs = false;
};
struct Angle3 { double theta_00 = 0, p_val1 = 0, p_val2 = 0, p_val4 = 0, p_val7 = 0, p_pen1 = 0, p_coa1 = 0; };
struct Tors4 { double V1 = 0, V2 = 0, V3 = 0, p_tor1 = 0, p_cot1 = 0; };
struct Hb3 { double r0 = -1, p1 = 0, p2 = 0, p3 = 0; };

struct FF {
  std::vector<double> gp;
  int vdw_type = 0;
  std::vector<Atom1> at;
  std::vector<Pair2> tb;   // n*n
  std::map<std::array<int, 3>, std::vector<Angle3>> thb;
  std::map<std::array<int, 4>, Tors4> fbp;
  std::map<std::array<int, 3>, Hb3> hbp;
  double bo_cut = 0, nonb_low = 0, nonb_cut = 0;
  bool gp37 = false, c2 = false;
  std::vector<std::string> warnings;
  int n() const { return (int)at.size(); }
  const Pair2& P(int i, int j) const { return tb[(size_t)i * at.size() + j]; }
  Pair2& P(int i, int j) { return tb[(size_t)i * at.size() + j]; }
};

struct Ctl {   // control-file parameters (src/io/control_parser.cpp defaults)
  double bond_cut = 5.0, bg_cut = 0.3, thb_cut = 0.001, thb_cutsq = 0.00001, hbond_cut = 7.5;
  int tabulate = 0;
};

// line reader with LAMMPS TextFileReader semantics (comments are ordinary words; value lines skip blank lines,
// skip_line consumes one raw line)
struct Lines {
  const std::string& s;
  size_t pos = 0;
  int line = 0;
  struct Eof {};
  explicit Lines(const std::string& t) : s(t) {}
  bool raw(std::string& out) {
    if (pos >= s.size()) return false;
    size_t e = s.find('\n', pos);
    if (e == std::string::npos) e = s.size();
    out = s.substr(pos, e - pos);
    pos = e < s.size() ? e + 1 : e;
    ++line;
    return true;
  }
  static std::vector<std::string> words(const std::string& l) {
    std::vector<std::string> w;
    size_t i = 0;
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\f'; };
    while (i < l.size()) {
      while (i < l.size() && ws(l[i])) ++i;
      if (i >= l.size()) break;
      size_t b = i;
      while (i < l.size() && !ws(l[i])) ++i;
      w.push_back(l.substr(b, i - b));
    }
    return w;
  }
  bool next_words(std::vector<std::string>& w) {
    std::string l;
    while (raw(l)) { w = words(l); if (!w.empty()) return true; }
    return false;
  }
  std::vector<std::string> values() { std::vector<std::string> w; if (!next_words(w)) throw Eof{}; return w; }
  void skip() { std::string l; if (!raw(l)) throw Eof{}; }
  std::string where() const { return "line " + std::to_string(line); }
};

struct Cursor {
  const std::vector<std::string>& w;
  size_t k = 0;
  std::string where;
  double d() {
    if (k >= w.size()) throw Err(where + ": missing value");
    const std::string& t = w[k++];
    char* e = nullptr;
    double v = std::strtod(t.c_str(), &e);
    if (e == t.c_str() || (size_t)(e - t.c_str()) != t.size()) throw Err(where + ": bad number '" + t + "'");
    if (!std::isfinite(v)) throw Err(where + ": non-finite number");
    return v;
  }
  int i() {
    if (k >= w.size()) throw Err(where + ": missing value");
    const std::string& t = w[k++];
    size_t p = 0;
    bool neg = false;
    if (p < t.size() && (t[p] == '+' || t[p] == '-')) neg = t[p++] == '-';
    if (p >= t.size()) throw Err(where + ": bad integer");
    long long v = 0;
    for (; p < t.size(); ++p) {
      if (t[p] < '0' || t[p] > '9') throw Err(where + ": bad integer '" + t + "'");
      v = v * 10 + (t[p] - '0');
      if (v > 2147483648LL) throw Err(where + ": integer out of range");
    }
    return (int)(neg ? -v : v);
  }
  std::string s() { if (k >= w.size()) throw Err(where + ": missing value"); return w[k++]; }
  bool more() const { return k < w.size(); }
};

static std::string upper3(const std::string& s) {
  std::string r;
  for (char c : s) { if (r.size() == 3) break; r.push_back((char)std::toupper((unsigned char)c)); }
  return r;
}

// reference: src/io/ffield_parser.cpp:91-451 (itself following LAMMPS reaxff_ffield.cpp)
FF parse_ffield(const std::string& text, bool lg) {
  FF f;
  Lines L(text);
  try {
    {
      auto w = L.values();   // header comment; it must not look like the general-parameter count line
      std::string joined;
      for (auto& x : w) joined += x + " ";
      bool digits = !w[0].empty() && std::all_of(w[0].begin(), w[0].end(), [](char c) { return c >= '0' && c <= '9'; });
      if (digits && w.size() >= 2 && !w[1].empty() && w[1][0] == '!' && joined.find("general parameters") != std::string::npos)
        throw Err("first line of the ffield must be a comment");
    }
    {
      auto w = L.values(); Cursor c{w, 0, L.where()};
      int n = c.i();
      if (n < 38) throw Err("ffield needs at least 38 general parameters");
      f.gp.resize(n);
      for (int i = 0; i < n; ++i) { auto g = L.values(); Cursor cg{g, 0, L.where()}; f.gp[i] = cg.d(); }
    }
    int nt;
    { auto w = L.values(); Cursor c{w, 0, L.where()}; nt = c.i(); if (nt < 1) throw Err("bad number of atom types"); L.skip(); L.skip(); L.skip(); }
    f.at.assign(nt, Atom1{});
    f.tb.assign((size_t)nt * nt, Pair2{});
    for (int i = 0; i < nt; ++i) {
      Atom1& a = f.at[i];
      auto w = L.values();
      if (w.size() < 8 && !lg) throw Err("this force field file requires lgvdw yes");
      if (w.size() < 9) throw Err(L.where() + ": too few columns");
      Cursor c{w, 0, L.where()};
      a.name = upper3(c.s());
      a.r_s = c.d(); a.valency = c.d(); a.mass = c.d(); a.r_vdw = c.d(); a.epsilon = c.d(); a.gamma = c.d(); a.r_pi = c.d(); a.valency_e = c.d();
      a.nlp_opt = 0.5 * (a.valency_e - a.valency);
      w = L.values(); if (w.size() < 8) throw Err(L.where() + ": too few columns");
      Cursor c2{w, 0, L.where()};
      a.alpha = c2.d(); a.gamma_w = c2.d(); a.valency_boc = c2.d(); a.p_ovun5 = c2.d(); c2.s(); a.chi = c2.d(); a.eta = 2.0 * c2.d();
      a.p_hbond = (int)c2.d();
      w = L.values(); if (w.size() < 8) throw Err(L.where() + ": too few columns");
      Cursor c3{w, 0, L.where()};
      a.r_pi_pi = c3.d(); a.p_lp2 = c3.d(); c3.s(); a.b_o_131 = c3.d(); a.b_o_132 = c3.d(); a.b_o_133 = c3.d(); a.bcut_acks2 = c3.d();
      w = L.values(); if (w.size() < 8) throw Err(L.where() + ": too few columns");
      Cursor c4{w, 0, L.where()};
      a.p_ovun2 = c4.d(); a.p_val3 = c4.d(); c4.s(); a.valency_val = c4.d(); a.p_val5 = c4.d(); a.rcore2 = c4.d(); a.ecore2 = c4.d(); a.acore2 = c4.d();
      if (lg) {
        w = L.values();
        char* e = nullptr; std::strtod(w[0].c_str(), &e);
        if (e == w[0].c_str() || (size_t)(e - w[0].c_str()) != w[0].size()) throw Err("ffield is not compatible with lgvdw yes");
        if (w.size() < 2) throw Err(L.where() + ": too few columns");
        Cursor c5{w, 0, L.where()};
        a.lgcij = c5.d(); a.lgre = c5.d();
      }
      // van der Waals variant (first element decides; later conflicts only warn, as the reference)
      const bool inner = a.rcore2 > 0.01 && a.acore2 > 0.01, shield = a.gamma_w > 0.5;
      const int want = inner ? (shield ? 3 : 2) : (shield ? 1 : 0);
      if (want == 0) throw Err("inconsistent van der Waals parameters for element " + a.name);
      if (f.vdw_type != 0 && f.vdw_type != want) f.warnings.push_back("van der Waals method of " + a.name + " differs; keeping the earlier one");
      else f.vdw_type = want;
    }
    for (auto& a : f.at) {
      a.is_c = a.name == "C";
      a.heavy = a.mass > 21.0;
      if (a.mass < 21 && a.valency_val != a.valency_boc) { f.warnings.push_back("Changed valency_val to valency_boc for " + a.name); a.valency_val = a.valency_boc; }
    }
    {
      auto w = L.values(); Cursor c{w, 0, L.where()};
      int nb = c.i();
      L.skip();
      for (int b = 0; b < nb; ++b) {
        w = L.values(); if (w.size() < 10) throw Err(L.where() + ": too few columns");
        Cursor c1{w, 0, L.where()};
        int j = c1.i() - 1, k = c1.i() - 1;
        if (j < 0 || k < 0) throw Err("inconsistent force field file");
        bool in = j < nt && k < nt;
        double v1[8];
        if (in) for (double& v : v1) v = c1.d();
        auto w2 = L.values(); if (w2.size() < 7) throw Err(L.where() + ": too few columns");
        if (!in) continue;
        Cursor c2{w2, 0, L.where()};
        double be2 = c2.d(), bo3 = c2.d(), bo4 = c2.d(); c2.s();
        double bo1 = c2.d(), bo2 = c2.d();
        double ovc = c2.more() ? c2.d() : 0.0;
        for (Pair2* p : {&f.P(j, k), &f.P(k, j)}) {
          p->De_s = v1[0]; p->De_p = v1[1]; p->De_pp = v1[2]; p->p_be1 = v1[3]; p->p_bo5 = v1[4]; p->v13cor = v1[5]; p->p_bo6 = v1[6]; p->p_ovun1 = v1[7];
          p->p_be2 = be2; p->p_bo3 = bo3; p->p_bo4 = bo4; p->p_bo1 = bo1; p->p_bo2 = bo2; p->ovc = ovc; p->listed = true;
        }
      }
    }
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) {
        const Atom1 &a = f.at[i], &b = f.at[j];
        for (Pair2* p : {&f.P(i, j), &f.P(j, i)}) {
          p->r_s = 0.5 * (b.r_s + a.r_s); p->r_p = 0.5 * (b.r_pi + a.r_pi); p->r_pp = 0.5 * (b.r_pi_pi + a.r_pi_pi);
          p->p_boc3 = std::sqrt(b.b_o_132 * a.b_o_132); p->p_boc4 = std::sqrt(b.b_o_131 * a.b_o_131); p->p_boc5 = std::sqrt(b.b_o_133 * a.b_o_133);
          p->D = std::sqrt(b.epsilon * a.epsilon); p->alpha = std::sqrt(b.alpha * a.alpha); p->r_vdW = 2.0 * std::sqrt(b.r_vdw * a.r_vdw);
          p->gamma_w = std::sqrt(b.gamma_w * a.gamma_w); p->gamma = std::pow(b.gamma * a.gamma, -1.5);
          p->rcore = std::sqrt(a.rcore2 * b.rcore2); p->ecore = std::sqrt(a.ecore2 * b.ecore2); p->acore = std::sqrt(a.acore2 * b.acore2);
          p->lgcij = std::sqrt(a.lgcij * b.lgcij); p->lgre = 2.0 * f.gp[35] * std::sqrt(a.lgre * b.lgre);
        }
      }
    {
      auto w = L.values(); Cursor c{w, 0, L.where()};
      int no = c.i();
      for (int o = 0; o < no; ++o) {
        w = L.values(); if (w.size() < (size_t)(8 + (lg ? 1 : 0))) throw Err(L.where() + ": too few columns");
        Cursor c1{w, 0, L.where()};
        int j = c1.i() - 1, k = c1.i() - 1;
        if (j < 0 || k < 0) throw Err("inconsistent force field file");
        if (j >= nt || k >= nt) continue;
        double D = c1.d(), rv = c1.d(), al = c1.d(), rs = c1.d(), rp = c1.d(), rpp = c1.d();
        for (Pair2* p : {&f.P(j, k), &f.P(k, j)}) {
          if (D > 0) p->D = D;
          if (rv > 0) p->r_vdW = 2.0 * rv;
          if (al > 0) p->alpha = al;
          if (rs > 0) p->r_s = rs;
          if (rp > 0) p->r_p = rp;
          if (rpp > 0) p->r_pp = rpp;
        }
        if (lg) { double v = c1.d(); if (v >= 0) { f.P(j, k).lgcij = v; f.P(k, j).lgcij = v; } }
      }
    }
    {
      auto w = L.values(); Cursor c{w, 0, L.where()};
      int n3 = c.i();
      std::map<std::array<int, 3>, int> slots;
      for (int a = 0; a < n3; ++a) {
        w = L.values(); if (w.size() < 10) throw Err(L.where() + ": too few columns");
        Cursor c1{w, 0, L.where()};
        int j = c1.i() - 1, k = c1.i() - 1, l = c1.i() - 1;
        if (j < 0 || k < 0 || l < 0) throw Err("inconsistent force field file");
        if (j >= nt || k >= nt || l >= nt) continue;
        Angle3 s;
        s.theta_00 = c1.d(); s.p_val1 = c1.d(); s.p_val2 = c1.d(); s.p_coa1 = c1.d(); s.p_val7 = c1.d(); s.p_pen1 = c1.d(); s.p_val4 = c1.d();
        slots[{j, k, l}] += 1;
        slots[{l, k, j}] += 1;   // for j == l the same counter twice (two slots per set in the reference)
        if (slots[{j, k, l}] > 5 || slots[{l, k, j}] > 5) throw Err("too many valence-angle parameter sets for one triple");
        f.thb[{j, k, l}].push_back(s);
        if (j != l) f.thb[{l, k, j}].push_back(s);
      }
    }
    {
      auto w = L.values(); Cursor c{w, 0, L.where()};
      int n4 = c.i();
      std::set<std::array<int, 4>> expl;
      for (int a = 0; a < n4; ++a) {
        w = L.values(); if (w.size() < 9) throw Err(L.where() + ": too few columns");
        Cursor c1{w, 0, L.where()};
        int j = c1.i() - 1, k = c1.i() - 1, l = c1.i() - 1, m = c1.i() - 1;
        if (j < -1 || k < 0 || l < 0 || m < -1) throw Err("inconsistent force field file");
        Tors4 t; t.V1 = c1.d(); t.V2 = c1.d(); t.V3 = c1.d(); t.p_tor1 = c1.d(); t.p_cot1 = c1.d();
        if (j >= 0 && m >= 0) {
          if (j < nt && k < nt && l < nt && m < nt) {
            expl.insert({j, k, l, m}); expl.insert({m, l, k, j});
            f.fbp[{j, k, l, m}] = t; f.fbp[{m, l, k, j}] = t;
          }
        } else if (k < nt && l < nt) {
          for (int p = 0; p < nt; ++p)
            for (int o = 0; o < nt; ++o) {
              if (!expl.count({p, k, l, o})) f.fbp[{p, k, l, o}] = t;
              if (!expl.count({o, l, k, p})) f.fbp[{o, l, k, p}] = t;
            }
        }
      }
    }
    {
      std::vector<std::string> w;
      if (!L.next_words(w)) {
        f.warnings.push_back("no hydrogen bond parameters");
      } else {
        Cursor c{w, 0, L.where()};
        int nh = c.i();
        for (int h = 0; h < nh; ++h) {
          auto w1 = L.values(); if (w1.size() < 7) throw Err(L.where() + ": too few columns");
          Cursor c1{w1, 0, L.where()};
          int j = c1.i() - 1, k = c1.i() - 1, l = c1.i() - 1;
          if (j < 0 || k < 0 || l < 0) throw Err("inconsistent force field file");
          if (j >= nt || k >= nt || l >= nt) continue;
          Hb3 hb; hb.r0 = c1.d(); hb.p1 = c1.d(); hb.p2 = c1.d(); hb.p3 = c1.d();
          f.hbp[{j, k, l}] = hb;
        }
      }
    }
  } catch (Lines::Eof&) {
    throw Err("unexpected end of file at " + L.where());
  }
  f.bo_cut = 0.01 * f.gp[29];
  f.nonb_low = f.gp[11];
  f.nonb_cut = f.gp[12];
  f.gp37 = (double)(int)f.gp[37] == 2;
  f.c2 = f.gp[5] > 0.001;
  for (int i = 0; i < f.n(); ++i)
    for (int j = 0; j < f.n(); ++j) {
      const Atom1 &a = f.at[i], &b = f.at[j];
      f.P(i, j).tbs = f.gp37 || (a.mass == 12.0000 && b.mass == 15.9990) || (b.mass == 12.0000 && a.mass == 15.9990);
    }
  return f;
}

Ctl parse_control(const std::string& text) {
  static const std::set<std::string> inactive = {
      "ensemble_type", "nsteps", "dt", "proc_by_dim", "random_vel", "restart_format", "restart_freq", "reposition_atoms",
      "restrict_bonds", "remove_CoM_vel", "debug_level", "reneighbor", "vlist_buffer", "ghost_cutoff", "qeq_freq", "q_err",
      "ilu_refactor", "ilu_droptol", "temp_init", "temp_final", "t_mass", "t_mode", "t_rate", "t_freq", "pressure", "p_mass",
      "pt_mass", "compress", "press_mode", "geo_format", "traj_compress", "traj_method", "molecular_analysis", "ignore",
      "dipole_anal", "freq_dipole_anal", "diffusion_coef", "freq_diffusion_coef", "restrict_type", "traj_title",
      "simulation_name", "energy_update_freq", "atom_info", "atom_velocities", "atom_forces", "bond_info", "angle_info"};
  Ctl c;
  Lines L(text);
  std::vector<std::string> w;
  while (L.next_words(w)) {
    if (w.size() < 2) throw Err("no value for control parameter " + w[0]);
    Cursor v{w, 1, L.where()};
    const std::string& k = w[0];
    if (inactive.count(k)) continue;
    if (k == "nbrhood_cutoff") c.bond_cut = v.d();
    else if (k == "bond_graph_cutoff") c.bg_cut = v.d();
    else if (k == "thb_cutoff") c.thb_cut = v.d();
    else if (k == "thb_cutoff_sq") c.thb_cutsq = v.d();
    else if (k == "hbond_cutoff") c.hbond_cut = v.d();
    else if (k == "tabulate_long_range") c.tabulate = v.i();
    else if (k == "write_freq") v.i();
    else throw Err("unknown parameter " + k + " in control file");
  }
  return c;
}

// same text as src/io/table_dump.cpp (%.17g values, fixed field order)
std::string table_dump(const FF& f, bool portable) {
  std::string o = "FFIELD-TABLES 1\n";
  auto put = [&](double v) { char b[40]; std::snprintf(b, sizeof b, " %.17g", v); o += b; };
  o += "GP " + std::to_string(f.gp.size()) + " " + std::to_string(f.vdw_type) + "\n";
  for (size_t i = 0; i < f.gp.size(); ++i) { o += "G " + std::to_string(i); put(f.gp[i]); o += "\n"; }
  const int n = f.n();
  for (int i = 0; i < n; ++i) {
    const Atom1& s = f.at[i];
    o += "S " + std::to_string(i) + " " + s.name;
    for (double v : {s.r_s, s.valency, s.mass, s.r_vdw, s.epsilon, s.gamma, s.r_pi, s.valency_e, s.nlp_opt, s.alpha, s.gamma_w, s.valency_boc, s.p_ovun5}) put(v);
    o += " " + std::to_string(s.p_hbond);
    for (double v : {s.chi, s.eta, s.r_pi_pi, s.p_lp2, s.b_o_131, s.b_o_132, s.b_o_133, s.bcut_acks2, s.p_ovun2, s.p_val3, s.valency_val, s.p_val5, s.rcore2,
                     s.ecore2, s.acore2, s.lgcij, s.lgre}) put(v);
    o += "\n";
  }
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const Pair2& t = f.P(i, j);
      o += "T " + std::to_string(i) + " " + std::to_string(j);
      for (double v : {t.p_bo1, t.p_bo2, t.p_bo3, t.p_bo4, t.p_bo5, t.p_bo6, t.r_s, t.r_p, t.r_pp, t.p_boc3, t.p_boc4, t.p_boc5, t.p_be1, t.p_be2, t.De_s, t.De_p,
                       t.De_pp, t.p_ovun1, t.D, t.alpha, t.r_vdW, t.gamma_w, t.rcore, t.ecore, t.acore, t.lgcij, t.lgre}) put(v);
      if (!portable) put(t.gamma);
      put(t.v13cor); put(t.ovc);
      o += "\n";
    }
  for (auto& kv : f.thb) {
    o += "H3 " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]) + " " + std::to_string(kv.second.size());
    for (auto& s : kv.second) for (double v : {s.theta_00, s.p_val1, s.p_val2, s.p_val4, s.p_val7, s.p_pen1, s.p_coa1}) put(v);
    o += "\n";
  }
  for (auto& kv : f.fbp) {
    o += "H4 " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]) + " " + std::to_string(kv.first[3]);
    for (double v : {kv.second.V1, kv.second.V2, kv.second.V3, kv.second.p_tor1, kv.second.p_cot1}) put(v);
    o += "\n";
  }
  for (auto& kv : f.hbp) {
    const Hb3& h = kv.second;
    if (h.r0 == -1 && h.p1 == 0 && h.p2 == 0 && h.p3 == 0) continue;
    o += "HB " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " + std::to_string(kv.first[2]);
    for (double v : {h.r0, h.p1, h.p2, h.p3}) put(v);
    o += "\n";
  }
  return o;
}

// ------------------------------------------------------------------------------------------------------------------
// geometry: box, image expander, far list  (reference src/core/system.cpp, src/neighbor/neighbor.cpp)
// ------------------------------------------------------------------------------------------------------------------
using V3 = std::array<double, 3>;
static V3 cross(const V3& a, const V3& b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
static double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

struct Box {
  V3 o{0, 0, 0};
  V3 v[3]{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  bool per[3]{false, false, false};
  double volume() const { return dot(v[0], cross(v[1], v[2])); }
  V3 heights() const {
    double vol = std::fabs(volume());
    auto nrm = [](const V3& a) { return std::sqrt(dot(a, a)); };
    return {vol / nrm(cross(v[1], v[2])), vol / nrm(cross(v[2], v[0])), vol / nrm(cross(v[0], v[1]))};
  }
  V3 frac(const V3& r) const {
    V3 d{r[0] - o[0], r[1] - o[1], r[2] - o[2]};
    double det = volume();
    return {dot(d, cross(v[1], v[2])) / det, dot(d, cross(v[2], v[0])) / det, dot(d, cross(v[0], v[1])) / det};
  }
  V3 shift(int s0, int s1, int s2) const {
    V3 r{0, 0, 0};
    int s[3] = {s0, s1, s2};
    for (int k = 0; k < 3; ++k) for (int c = 0; c < 3; ++c) r[c] += (double)s[k] * v[k][c];
    return r;
  }
};

struct Atoms {   // expanded (ghost-native) set: owned first, then ghosts = owner + lattice shift
  size_t nlocal = 0;
  std::vector<double> x;
  std::vector<int> type;
  std::vector<int64_t> tag;
  std::vector<int> owner;
  size_t nall() const { return type.size(); }
};

Atoms expand(const Box& b, const double* xo, const int* type, const int64_t* tag, size_t n, double shell) {
  if (!(b.volume() > 0)) throw Err("box is degenerate or left-handed");
  Atoms a;
  a.nlocal = n;
  a.x.assign(xo, xo + 3 * n);
  a.type.assign(type, type + n);
  a.tag.assign(tag, tag + n);
  a.owner.resize(n);
  for (size_t i = 0; i < n; ++i) a.owner[i] = (int)i;
  V3 h = b.heights();
  double w[3] = {0, 0, 0};
  int sm[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) if (b.per[d]) { w[d] = shell / h[d]; sm[d] = (int)std::ceil(w[d]) + 1; }
  std::vector<V3> fr(n);
  for (size_t i = 0; i < n; ++i) {
    V3 r{a.x[3 * i], a.x[3 * i + 1], a.x[3 * i + 2]};
    V3 f = b.frac(r);
    double k[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) if (b.per[d]) k[d] = std::floor(f[d]);
    for (int d = 0; d < 3; ++d) if (k[d] != 0.0) for (int c = 0; c < 3; ++c) r[c] -= k[d] * b.v[d][c];
    f = b.frac(r);
    for (int c = 0; c < 3; ++c) a.x[3 * i + c] = r[c];
    fr[i] = f;
  }
  for (int s0 = -sm[0]; s0 <= sm[0]; ++s0)
    for (int s1 = -sm[1]; s1 <= sm[1]; ++s1)
      for (int s2 = -sm[2]; s2 <= sm[2]; ++s2) {
        if (!s0 && !s1 && !s2) continue;
        int s[3] = {s0, s1, s2};
        V3 sh = b.shift(s0, s1, s2);
        for (size_t k = 0; k < n; ++k) {
          bool in = true;
          for (int d = 0; d < 3 && in; ++d) {
            if (!b.per[d]) { if (s[d] != 0) in = false; continue; }
            double g = fr[k][d] + s[d];
            in = g >= -w[d] && g <= 1.0 + w[d];
          }
          if (!in) continue;
          for (int c = 0; c < 3; ++c) a.x.push_back(a.x[3 * k + c] + sh[c]);
          a.type.push_back(a.type[k]);
          a.tag.push_back(a.tag[k]);
          a.owner.push_back((int)k);
        }
      }
  return a;
}

struct Far {   // rows of j > i within the row cutoff (nonb for owned rows, bond for ghost rows), ascending j
  std::vector<size_t> start;
  std::vector<int> nbr;
  std::vector<double> d, dv;
};

Far far_list(const Atoms& a, double nonb, double bond) {
  const size_t N = a.nall();
  Far F;
  F.start.push_back(0);
  if (!N) return F;
  const double cell = std::max(nonb, bond);
  double lo[3], hi[3];
  for (int c = 0; c < 3; ++c) { lo[c] = hi[c] = a.x[c]; }
  for (size_t i = 0; i < N; ++i) for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], a.x[3 * i + c]); hi[c] = std::max(hi[c], a.x[3 * i + c]); }
  int nc[3];
  double cs[3];
  for (int c = 0; c < 3; ++c) {
    double k = std::floor((hi[c] - lo[c]) / cell);
    nc[c] = k < 1 ? 1 : (int)std::min(k, 1000.0);
    cs[c] = std::max(cell, (hi[c] - lo[c]) / nc[c]);
  }
  std::vector<int> cell_of(N);
  std::vector<std::vector<int>> members((size_t)nc[0] * nc[1] * nc[2]);
  for (size_t i = 0; i < N; ++i) {
    int q[3];
    for (int c = 0; c < 3; ++c) { double t = std::floor((a.x[3 * i + c] - lo[c]) / cs[c]); q[c] = t < 0 ? 0 : std::min(nc[c] - 1, (int)t); }
    cell_of[i] = (q[2] * nc[1] + q[1]) * nc[0] + q[0];
    members[cell_of[i]].push_back((int)i);
  }
  struct E { int j; double d, v[3]; };
  auto make_row = [&](size_t i, std::vector<E>& row) {
    const double rc = i < a.nlocal ? nonb : bond, rc2 = rc * rc;
    int c0 = cell_of[i] % nc[0], c1 = (cell_of[i] / nc[0]) % nc[1], c2 = cell_of[i] / (nc[0] * nc[1]);
    row.clear();
    for (int z = std::max(c2 - 1, 0); z <= std::min(c2 + 1, nc[2] - 1); ++z)
      for (int y = std::max(c1 - 1, 0); y <= std::min(c1 + 1, nc[1] - 1); ++y)
        for (int x = std::max(c0 - 1, 0); x <= std::min(c0 + 1, nc[0] - 1); ++x)
          for (int j : members[(z * nc[1] + y) * nc[0] + x]) {
            if ((size_t)j <= i) continue;
            E e;
            for (int c = 0; c < 3; ++c) e.v[c] = a.x[3 * (size_t)j + c] - a.x[3 * i + c];
            double d2 = e.v[0] * e.v[0] + e.v[1] * e.v[1] + e.v[2] * e.v[2];
            if (!(d2 <= rc2)) continue;
            e.j = j; e.d = std::sqrt(d2);
            row.push_back(e);
          }
    std::sort(row.begin(), row.end(), [](const E& p, const E& q) { return p.j < q.j; });
  };
  // rows of a chunk are built independently (in parallel), then appended in row order
  const size_t chunk = 4096;
  std::vector<std::vector<E>> rows(std::min(chunk, N));
  for (size_t i0 = 0; i0 < N; i0 += chunk) {
    const long m = (long)std::min(chunk, N - i0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (long k = 0; k < m; ++k) make_row(i0 + (size_t)k, rows[(size_t)k]);
    for (long k = 0; k < m; ++k) {
      for (auto& e : rows[(size_t)k]) { F.nbr.push_back(e.j); F.d.push_back(e.d); F.dv.insert(F.dv.end(), e.v, e.v + 3); }
      F.start.push_back(F.nbr.size());
    }
  }
  return F;
}

// owner-computes rule of the nonbonded pair loop: 0 none, 1 owned-owned, 2 owned-ghost, 3 self image (SMALL = 1e-4)
static int pair_class(size_t j, size_t nlocal, int64_t ti, int64_t tj, const double* dv) {
  const double SMALL = 0.0001;
  if (j < nlocal) return 1;
  if (ti < tj) return 2;
  if (ti == tj) {
    if (dv[2] > SMALL) return 3;
    if (std::fabs(dv[2]) < SMALL) {
      if (dv[1] > SMALL) return 3;
      if (std::fabs(dv[1]) < SMALL && dv[0] > SMALL) return 3;
    }
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------------------------
// results
// ------------------------------------------------------------------------------------------------------------------
enum Term { E_BOND, E_OV, E_UN, E_LP, E_ANG, E_PEN, E_COA, E_HB, E_TOR, E_CON, E_VDW, E_ELE, E_POL, NTERM };

struct Result {
  double e[NTERM] = {0};
  std::vector<double> grad;          // 3 * nall, dE/dx (ghosts included)
  std::vector<double> total_bo;      // nall
  std::vector<int> c_bonds, c_angles, c_tors, c_hb, c_nb, c_sbo, c_lp;   // per owned atom
  std::vector<double> sbo;   // per owned atom: the SBO value whose branch c_sbo records (its distance to the branch thresholds)
  // bond table of the owned atoms as the reference's BondTable (include/reaxmetal/bonded.hpp:32-37, filled at src/cpu/bonded.cpp:
  // 782-792), partners given as owned atoms (the owner of the image): start[n+1], owner[], bo[], nlp[n]
  std::vector<int> bt_start, bt_owner;
  std::vector<double> bt_bo, bt_nlp;
  long long pairs_oo = 0, pairs_og = 0, pairs_self = 0, nbonds = 0;
};

// ------------------------------------------------------------------------------------------------------------------
// bonded terms (reference src/cpu/bonded.cpp:86-800; formulas of include/reaxmetal/terms.hpp)
// ------------------------------------------------------------------------------------------------------------------
struct Bnd {
  int nbr = 0, sym = 0;
  double d = 0;
  V3 dv{}, dBOp{}, dln_pi{}, dln_pi2{};
  double BO = 0, BO_s = 0, BO_pi = 0, BO_pi2 = 0;
  double C1dbo = 0, C2dbo = 0, C3dbo = 0, C1dbopi = 0, C2dbopi = 0, C3dbopi = 0, C4dbopi = 0, C1dbopi2 = 0, C2dbopi2 = 0, C3dbopi2 = 0, C4dbopi2 = 0;
  double Cdbo = 0, Cdbopi = 0, Cdbopi2 = 0;
};
struct Thb { int thb = 0, pthb = 0; double theta = 0; V3 di{}, dj{}, dk{}; };
struct HbE { int nbr; double scl, d; V3 dv; };

static inline void axpy(V3& a, double s, const V3& b) { a[0] += s * b[0]; a[1] += s * b[1]; a[2] += s * b[2]; }
static inline V3 scl(double s, const V3& b) { return {s * b[0], s * b[1], s * b[2]}; }

// one-sided bond-order correction f1 f4 f5 and its 11 derivative coefficients (terms.hpp:66-132)
static void bo_correct(const Pair2& p, double boc1, double boc2, double vi, double vj, double Dpi, double Dpj, double Dbi, double Dbj, Bnd& b) {
  double BO = b.BO, BOpi = b.BO_pi, BOpi2 = b.BO_pi2;
  if (p.ovc < 0.001 && p.v13cor < 0.001) {
    b.C1dbo = 1; b.C2dbo = b.C3dbo = 0;
    b.C1dbopi = 1; b.C2dbopi = b.C3dbopi = b.C4dbopi = 0;
    b.C1dbopi2 = 1; b.C2dbopi2 = b.C3dbopi2 = b.C4dbopi2 = 0;
  } else {
    double f1 = 1, Cf1ij = 0, Cf1ji = 0;
    if (p.ovc >= 0.001) {
      const double e1i = std::exp(-boc1 * Dpi), e2i = std::exp(-boc2 * Dpi), e1j = std::exp(-boc1 * Dpj), e2j = std::exp(-boc2 * Dpj);
      const double f2 = e1i + e1j, f3 = -1.0 / boc2 * std::log(0.5 * (e2i + e2j));
      f1 = 0.5 * ((vi + f2) / (vi + f2 + f3) + (vj + f2) / (vj + f2 + f3));
      const double ui = vi + (f2 + f3), uj = vj + (f2 + f3), S = e2i + e2j;
      // d f1 / d Delta'_i and d f1 / d Delta'_j (the reference writes the second through its A/B helper sums)
      Cf1ij = 0.5 * (-boc1 * e1i / ui - ((vi + f2) / (ui * ui)) * (-boc1 * e1i + e2i / S) + -boc1 * e1i / uj - ((vj + f2) / (uj * uj)) * (-boc1 * e1i + e2i / S));
      const double A = 0.5 * f3 * (1.0 / (ui * ui) + 1.0 / (uj * uj));
      const double B = -0.5 * ((ui - f3) / (ui * ui) + (uj - f3) / (uj * uj));
      Cf1ji = -A * boc1 * e1j + B * e2j / S;
    }
    double f4 = 1, f5 = 1, Cf45ij = 0, Cf45ji = 0;
    if (p.v13cor >= 0.001) {
      const double ef4 = std::exp(-(p.p_boc4 * (BO * BO) - Dbi) * p.p_boc3 + p.p_boc5);
      const double ef5 = std::exp(-(p.p_boc4 * (BO * BO) - Dbj) * p.p_boc3 + p.p_boc5);
      f4 = 1.0 / (1.0 + ef4); f5 = 1.0 / (1.0 + ef5);
      Cf45ij = -f4 * ef4; Cf45ji = -f5 * ef5;
    }
    const double f45 = f4 * f5;
    const double A0 = f1 * f45;
    const double A1 = -2 * p.p_boc3 * p.p_boc4 * BO * (Cf45ij + Cf45ji);
    const double A2ij = Cf1ij / f1 + p.p_boc3 * Cf45ij, A2ji = Cf1ji / f1 + p.p_boc3 * Cf45ji;
    const double A3ij = A2ij + Cf1ij / f1, A3ji = A2ji + Cf1ji / f1;
    b.BO = BO * A0;
    b.BO_pi = BOpi * A0 * f1;
    b.BO_pi2 = BOpi2 * A0 * f1;
    b.BO_s = b.BO - (b.BO_pi + b.BO_pi2);
    b.C1dbo = A0 + b.BO * A1; b.C2dbo = b.BO * A2ij; b.C3dbo = b.BO * A2ji;
    b.C1dbopi = f1 * f1 * f4 * f5; b.C2dbopi = b.BO_pi * A1; b.C3dbopi = b.BO_pi * A3ij; b.C4dbopi = b.BO_pi * A3ji;
    b.C1dbopi2 = f1 * f1 * f4 * f5; b.C2dbopi2 = b.BO_pi2 * A1; b.C3dbopi2 = b.BO_pi2 * A3ij; b.C4dbopi2 = b.BO_pi2 * A3ji;
  }
  if (b.BO < 1e-10) b.BO = 0;
  if (b.BO_s < 1e-10) b.BO_s = 0;
  if (b.BO_pi < 1e-10) b.BO_pi = 0;
  if (b.BO_pi2 < 1e-10) b.BO_pi2 = 0;
}

struct AQ { double Delta, Delta_e, Delta_boc, Delta_val, vlpex, nlp, Delta_lp, dDelta_lp, Delta_lp_temp; };

static AQ atom_q(double tbo, const Atom1& s, double p_lp1) {   // terms.hpp:141-164
  AQ a;
  a.Delta = tbo - s.valency;
  a.Delta_e = tbo - s.valency_e;
  a.Delta_boc = tbo - s.valency_boc;
  a.Delta_val = tbo - s.valency_val;
  const double ht = (double)(int)(a.Delta_e / 2.0);   // truncation toward zero
  a.vlpex = a.Delta_e - 2.0 * ht;
  const double ex = std::exp(-p_lp1 * ((2.0 + a.vlpex) * (2.0 + a.vlpex)));
  a.nlp = ex - ht;
  a.Delta_lp = s.nlp_opt - a.nlp;
  a.dDelta_lp = 2.0 * p_lp1 * ex * (2.0 + a.vlpex);
  a.Delta_lp_temp = s.heavy ? s.nlp_opt - 0.5 * (s.valency_e - s.valency) : a.Delta_lp;
  return a;
}

static void theta_of(const V3& a, double da, const V3& b, double db, double& th, double& c) {
  c = dot(a, b) / (da * db);
  if (c > 1.0) c = 1.0;
  if (c < -1.0) c = -1.0;
  th = std::acos(c);
}
static void dcos_of(const V3& a, double da, const V3& b, double db, V3& di, V3& dj, V3& dk) {
  const double da2 = da * da, db2 = db * db, inv = 1.0 / (da * db), inv3 = inv * inv * inv, cd = dot(a, b) * inv3;
  for (int t = 0; t < 3; ++t) {
    di[t] = b[t] * inv - cd * db2 * a[t];
    dj[t] = -(b[t] + a[t]) * inv + cd * (db2 * a[t] + da2 * b[t]);
    dk[t] = a[t] * inv - cd * da2 * b[t];
  }
}

void bonded(const FF& ff, const Ctl& ctl, const Atoms& A, const Far& far, bool enobonds, Result& R) {
  const size_t N = A.nall(), n = A.nlocal;
  const auto& gp = ff.gp;
  const double bo_cut = ff.bo_cut, boc1 = gp[0], boc2 = gp[1], p_lp1 = gp[15], p_lp3 = gp[5];
  const double gp3 = gp[3], gp4 = gp[4], gp7 = gp[7], gp10 = gp[10];
  const double ovun3 = gp[32], ovun4 = gp[31], ovun6 = gp[6], ovun7 = gp[8], ovun8 = gp[9];
  std::vector<std::vector<Bnd>> B(N);
  std::vector<V3> dDps(N, V3{0, 0, 0});
  std::vector<double> tbo(N, 0.0), Dp(N, 0.0), Dpb(N, 0.0);
  std::vector<int> mark(N, 0);
  std::vector<std::vector<HbE>> H(N);
  for (size_t i = n; i < N; ++i) mark[i] = 1000;

  // bond list, uncorrected bond orders, H-bond candidate lists
  for (size_t i = 0; i < N; ++i) {
    const int ti = A.type[i];
    if (ti < 0) continue;
    const Atom1& si = ff.at[ti];
    for (size_t p = far.start[i]; p < far.start[i + 1]; ++p) {
      const size_t j = (size_t)far.nbr[p];
      const double d = far.d[p];
      if (!(d <= std::max(ctl.bond_cut, i < n ? ctl.hbond_cut : 0.0))) continue;
      const int tj = A.type[j];
      if (tj < 0) continue;
      const Atom1& sj = ff.at[tj];
      const double* dv = &far.dv[3 * p];
      if (i < n && ctl.hbond_cut > 0 && (si.p_hbond == 1 || si.p_hbond == 2) && d <= ctl.hbond_cut) {
        if (si.p_hbond == 1 && sj.p_hbond == 2) H[i].push_back({(int)j, 1.0, d, {dv[0], dv[1], dv[2]}});
        else if (j < n && si.p_hbond == 2 && sj.p_hbond == 1) H[j].push_back({(int)i, -1.0, d, {dv[0], dv[1], dv[2]}});
      }
      if (!(d <= ctl.bond_cut)) continue;
      const Pair2& t = ff.P(ti, tj);
      double C12 = 0, C34 = 0, C56 = 0, bs = 0, bp = 0, bpp = 0;
      if (si.r_s > 0 && sj.r_s > 0) { C12 = t.p_bo1 * std::pow(d / t.r_s, t.p_bo2); bs = (1.0 + bo_cut) * std::exp(C12); }
      if (si.r_pi > 0 && sj.r_pi > 0) { C34 = t.p_bo3 * std::pow(d / t.r_p, t.p_bo4); bp = std::exp(C34); }
      if (si.r_pi_pi > 0 && sj.r_pi_pi > 0) { C56 = t.p_bo5 * std::pow(d / t.r_pp, t.p_bo6); bpp = std::exp(C56); }
      if (!(bs + bp + bpp >= bo_cut)) continue;
      Bnd bi, bj;
      bi.nbr = (int)j; bj.nbr = (int)i;
      bi.sym = (int)B[j].size(); bj.sym = (int)B[i].size();
      bi.d = bj.d = d;
      bi.dv = {dv[0], dv[1], dv[2]};
      bj.dv = scl(-1.0, bi.dv);
      bi.BO = bj.BO = bs + bp + bpp;
      bi.BO_s = bj.BO_s = bs; bi.BO_pi = bj.BO_pi = bp; bi.BO_pi2 = bj.BO_pi2 = bpp;
      const double rr2 = 1.0 / (d * d);
      const double Cs = t.p_bo2 * C12 * rr2, Cp = t.p_bo4 * C34 * rr2, Cpp = t.p_bo6 * C56 * rr2;
      bi.dln_pi = scl(-bi.BO_pi * Cp, bi.dv);
      bi.dln_pi2 = scl(-bi.BO_pi2 * Cpp, bi.dv);
      bj.dln_pi = scl(-1.0, bi.dln_pi);
      bj.dln_pi2 = scl(-1.0, bi.dln_pi2);
      bi.dBOp = scl(-(bi.BO_s * Cs + bi.BO_pi * Cp + bi.BO_pi2 * Cpp), bi.dv);
      bj.dBOp = scl(-1.0, bi.dBOp);
      axpy(dDps[i], 1.0, bi.dBOp);
      axpy(dDps[j], 1.0, bj.dBOp);
      bi.BO_s -= bo_cut; bi.BO -= bo_cut; bj.BO_s -= bo_cut; bj.BO -= bo_cut;
      tbo[i] += bi.BO; tbo[j] += bj.BO;
      B[i].push_back(bi); B[j].push_back(bj);
      if (mark[j] > mark[i] + 1) mark[j] = mark[i] + 1;
      else if (mark[i] > mark[j] + 1) mark[i] = mark[j] + 1;
    }
  }
  // corrected bond orders
  for (size_t i = 0; i < N; ++i) {
    if (A.type[i] < 0) continue;
    const Atom1& s = ff.at[A.type[i]];
    Dp[i] = tbo[i] - s.valency;
    Dpb[i] = tbo[i] - s.valency_val;
    tbo[i] = 0;
  }
  for (size_t i = 0; i < N; ++i) {
    const int ti = A.type[i];
    if (ti < 0) continue;
    for (Bnd& b : B[i]) {
      const size_t j = (size_t)b.nbr;
      const int tj = A.type[j];
      if (tj < 0) continue;
      if (i < j || mark[j] > 3) {
        bo_correct(ff.P(ti, tj), boc1, boc2, ff.at[ti].valency, ff.at[tj].valency, Dp[i], Dp[j], Dpb[i], Dpb[j], b);
      } else {
        const Bnd& m = B[j][(size_t)b.sym];
        b.BO = m.BO; b.BO_s = m.BO_s; b.BO_pi = m.BO_pi; b.BO_pi2 = m.BO_pi2;
      }
      tbo[i] += b.BO;
    }
  }
  std::vector<AQ> aq(N);
  for (size_t j = 0; j < N; ++j) if (A.type[j] >= 0) aq[j] = atom_q(tbo[j], ff.at[A.type[j]], p_lp1);

  std::vector<double> CdD(N, 0.0);
  R.c_bonds.assign(n, 0); R.c_angles.assign(n, 0); R.c_tors.assign(n, 0); R.c_hb.assign(n, 0); R.c_sbo.assign(n, 0); R.c_lp.assign(n, 0);
  R.sbo.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i) { R.c_lp[i] = (int)(aq[i].Delta_e / 2.0); R.c_bonds[i] = (int)B[i].size(); }
  R.bt_start.assign(n + 1, 0);
  R.bt_owner.clear(); R.bt_bo.clear(); R.bt_nlp.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    R.bt_start[i + 1] = R.bt_start[i] + (int)B[i].size();
    for (const Bnd& b : B[i]) { R.bt_owner.push_back(A.owner[(size_t)b.nbr]); R.bt_bo.push_back(b.BO); }
    R.bt_nlp[i] = aq[i].nlp;
  }
  double e_bond = 0, e_lp = 0, e_ov = 0, e_un = 0;

  // bond energy: owned end with the smaller tag (coordinate tie-break for self images)
  for (size_t i = 0; i < n; ++i) {
    const int ti = A.type[i];
    for (Bnd& b : B[i]) {
      const size_t j = (size_t)b.nbr;
      if (A.tag[i] > A.tag[j]) continue;
      if (A.tag[i] == A.tag[j]) {
        const double *xi = &A.x[3 * i], *xj = &A.x[3 * j];
        if (xj[2] < xi[2]) continue;
        if (xj[2] == xi[2] && xj[1] < xi[1]) continue;
        if (xj[2] == xi[2] && xj[1] == xi[1] && xj[0] < xi[0]) continue;
      }
      const int tj = A.type[j];
      if (ti < 0 || tj < 0) continue;
      const Pair2& t = ff.P(ti, tj);
      const double pw = b.BO_s == 0 ? 0.0 : std::pow(b.BO_s, t.p_be2);
      const double ex = std::exp(t.p_be1 * (1.0 - pw));
      const double CEbo = -t.De_s * ex * (1.0 - t.p_be1 * t.p_be2 * pw);
      e_bond += -t.De_s * b.BO_s * ex - t.De_p * b.BO_pi - t.De_pp * b.BO_pi2;
      b.Cdbo += CEbo;
      b.Cdbopi -= (CEbo + t.De_p);
      b.Cdbopi2 -= (CEbo + t.De_pp);
      if (b.BO >= 1.00 && t.tbs) {   // terminal triple-bond stabilisation (terms.hpp:183-194)
        const double BO = b.BO;
        const double exphu = std::exp(-gp7 * ((BO - 2.50) * (BO - 2.50)));
        const double ea = std::exp(-gp3 * (tbo[i] - BO)), eb = std::exp(-gp3 * (tbo[j] - BO));
        const double eov = std::exp(gp4 * (aq[i].Delta + aq[j].Delta));
        const double hov = 1.0 / (1.0 + 25.0 * eov);
        e_bond += gp10 * exphu * hov * (ea + eb);
        b.Cdbo += gp10 * exphu * hov * (ea + eb) * (gp3 - 2.0 * gp7 * (BO - 2.50));
        CdD[i] += -gp10 * exphu * hov * (gp3 * ea + 25.0 * gp4 * eov * hov * (ea + eb));
        CdD[j] += -gp10 * exphu * hov * (gp3 * eb + 25.0 * gp4 * eov * hov * (ea + eb));
      }
    }
  }
  // lone pair (+ C2 correction), then over/under coordination
  for (size_t i = 0; i < n; ++i) {
    const int ti = A.type[i];
    if (ti < 0) continue;
    const Atom1& s = ff.at[ti];
    if (!B[i].empty() || enobonds) {
      const double Dlp = aq[i].Delta_lp;
      double ev = -75 * Dlp;
      const double e2 = std::exp(ev > 700 ? 700 : ev), inv = 1.0 / (1.0 + e2);
      e_lp += s.p_lp2 * Dlp * inv;
      CdD[i] += (s.p_lp2 * inv + 75 * s.p_lp2 * Dlp * e2 * (inv * inv)) * aq[i].dDelta_lp;
    }
    if (p_lp3 > 0.001 && s.is_c) {
      for (Bnd& b : B[i]) {
        const int tj = A.type[(size_t)b.nbr];
        if (tj < 0 || !ff.at[tj].is_c) continue;
        const double Di = aq[i].Delta;
        const double vov3 = b.BO - Di - 0.040 * std::pow(Di, 4.);
        if (vov3 > 3.) {
          e_lp += p_lp3 * ((vov3 - 3.0) * (vov3 - 3.0));
          b.Cdbo += 2. * p_lp3 * (vov3 - 3.);
          CdD[i] += 2. * p_lp3 * (vov3 - 3.) * (-1. - 0.16 * std::pow(Di, 3.));
        }
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const int ti = A.type[i];
    if (ti < 0) continue;
    const Atom1& s = ff.at[ti];
    const double dfvl = s.heavy ? 0.0 : 1.0;
    double sum1 = 0, sum2 = 0;
    for (const Bnd& b : B[i]) {
      const size_t j = (size_t)b.nbr;
      const int tj = A.type[j];
      if (tj < 0) continue;
      const Pair2& t = ff.P(ti, tj);
      sum1 += t.p_ovun1 * t.De_s * b.BO;
      sum2 += (aq[j].Delta - dfvl * aq[j].Delta_lp_temp) * (b.BO_pi + b.BO_pi2);
    }
    const bool under = !B[i].empty() || enobonds;
    auto cexp = [](double x) { return std::exp(x > 700 ? 700 : x); };
    const double Dlpt = aq[i].Delta_lp_temp, dDlp = aq[i].dDelta_lp;
    const double eo1 = ovun3 * cexp(ovun4 * sum2), ieo1 = 1.0 / (1 + eo1);
    const double Dlc = aq[i].Delta - (dfvl * Dlpt) * ieo1;
    const double eo2 = cexp(s.p_ovun2 * Dlc), ieo2 = 1.0 / (1.0 + eo2);
    const double DlpVi = 1.0 / (Dlc + s.valency + 1e-8);
    const double CEover1 = Dlc * DlpVi * ieo2;
    e_ov += sum1 * CEover1;
    const double CEover2 = sum1 * DlpVi * ieo2 * (1.0 - Dlc * (DlpVi + s.p_ovun2 * eo2 * ieo2));
    const double CEover3 = CEover2 * (1.0 - dfvl * dDlp * ieo1);
    const double CEover4 = CEover2 * (dfvl * Dlpt) * ovun4 * eo1 * (ieo1 * ieo1);
    const double eo2n = cexp(-s.p_ovun2 * Dlc), eo6 = cexp(ovun6 * Dlc), eo8 = ovun7 * cexp(ovun8 * sum2);
    const double ieo2n = 1.0 / (1.0 + eo2n), ieo8 = 1.0 / (1.0 + eo8);
    const double eun = under ? -s.p_ovun5 * (1.0 - eo6) * ieo2n * ieo8 : 0.0;
    const double CEunder1 = ieo2n * (s.p_ovun5 * ovun6 * eo6 * ieo8 + s.p_ovun2 * eun * eo2n);
    const double CEunder2 = -eun * ovun8 * eo8 * ieo8;
    const double CEunder3 = CEunder1 * (1.0 - dfvl * dDlp * ieo1);
    const double CEunder4 = CEunder1 * (dfvl * Dlpt) * ovun4 * eo1 * (ieo1 * ieo1) + CEunder2;
    if (under) e_un += eun;
    CdD[i] += CEover3;
    if (under) CdD[i] += CEunder3;
    for (Bnd& b : B[i]) {
      const size_t j = (size_t)b.nbr;
      const Pair2& t = ff.P(ti, A.type[j]);
      b.Cdbo += CEover1 * t.p_ovun1 * t.De_s;
      const double pis = b.BO_pi + b.BO_pi2, dj = aq[j].Delta - dfvl * aq[j].Delta_lp_temp;
      CdD[j] += CEover4 * (1.0 - dfvl * aq[j].dDelta_lp) * pis;
      b.Cdbopi += CEover4 * dj; b.Cdbopi2 += CEover4 * dj;
      CdD[j] += CEunder4 * (1.0 - dfvl * aq[j].dDelta_lp) * pis;
      b.Cdbopi += CEunder4 * dj; b.Cdbopi2 += CEunder4 * dj;
    }
  }

  // valence angle, penalty, 3-body conjugation; builds the three-body lists the torsion needs
  std::vector<V3> f(N, V3{0, 0, 0});
  double e_ang = 0, e_pen = 0, e_coa = 0, e_tor = 0, e_con = 0, e_hb = 0;
  std::vector<std::vector<std::vector<Thb>>> T(N);
  for (size_t j = 0; j < N; ++j) T[j].resize(B[j].size());
  const double p_val6 = gp[14], p_val8 = gp[33], p_val9 = gp[16], p_val10 = gp[17];
  const double p_pen2 = gp[19], p_pen3 = gp[20], p_pen4 = gp[21], p_coa2 = gp[2], p_coa3 = gp[38], p_coa4 = gp[30];
  for (size_t j = 0; j < N; ++j) {
    const int tj = A.type[j];
    if (tj < 0) continue;
    auto& bj = B[j];
    const double p_val3 = ff.at[tj].p_val3, p_val5 = ff.at[tj].p_val5;
    double SBOp = 0, prod = 1;
    for (const Bnd& b : bj) {
      SBOp += b.BO_pi + b.BO_pi2;
      double q = b.BO * b.BO; q *= q; q *= q;
      prod *= std::exp(-q);
    }
    double vlpadj, dSBO2;
    if (aq[j].vlpex >= 0) { vlpadj = 0; dSBO2 = prod - 1; }
    else { vlpadj = aq[j].nlp; dSBO2 = (prod - 1) * (1 - p_val8 * aq[j].dDelta_lp); }
    const double Dboc = aq[j].Delta_boc;
    const double SBO = SBOp + (1 - prod) * (-Dboc - p_val8 * vlpadj);
    const double dSBO1 = -8 * prod * (Dboc + p_val8 * vlpadj);
    double SBO2, CSBO2;
    if (SBO <= 0) { SBO2 = 0; CSBO2 = 0; }
    else if (SBO <= 1) { SBO2 = std::pow(SBO, p_val9); CSBO2 = p_val9 * std::pow(SBO, p_val9 - 1); }
    else if (SBO < 2) { SBO2 = 2 - std::pow(2 - SBO, p_val9); CSBO2 = p_val9 * std::pow(2 - SBO, p_val9 - 1); }
    else { SBO2 = 2; CSBO2 = 0; }
    const double ev6 = std::exp(p_val6 * Dboc);
    if (j < n) { R.c_sbo[j] = SBO <= 0 ? 0 : (SBO <= 1 ? 1 : (SBO < 2 ? 2 : 3)); R.sbo[j] = SBO; }
    for (size_t pi = 0; pi < bj.size(); ++pi) {
      Bnd& bij = bj[pi];
      const double BOA_ij = bij.BO - ctl.thb_cut;
      if (!(BOA_ij > 0.0 && (j < n || (size_t)bij.nbr < n))) continue;
      const size_t i = (size_t)bij.nbr;
      const int ti = A.type[i];
      auto& li = T[j][pi];
      for (size_t pk = 0; pk < pi; ++pk)
        for (const Thb& e : T[j][pk])
          if (e.thb == (int)i) {
            Thb t; t.thb = bj[pk].nbr; t.pthb = (int)pk; t.theta = e.theta; t.di = e.dk; t.dj = e.dj; t.dk = e.di;
            li.push_back(t);
            break;
          }
      for (size_t pk = pi + 1; pk < bj.size(); ++pk) {
        Bnd& bjk = bj[pk];
        const double BOA_jk = bjk.BO - ctl.thb_cut;
        const size_t k = (size_t)bjk.nbr;
        const int tk = A.type[k];
        Thb t;
        double th, ct;
        theta_of(bij.dv, bij.d, bjk.dv, bjk.d, th, ct);
        dcos_of(bij.dv, bij.d, bjk.dv, bjk.d, t.di, t.dj, t.dk);
        t.thb = (int)k; t.pthb = (int)pk; t.theta = th;
        double st = std::sin(th);
        if (st < 1.0e-5) st = 1.0e-5;
        li.push_back(t);
        if (!(j < n && BOA_jk > 0.0 && bij.BO > ctl.thb_cut && bjk.BO > ctl.thb_cut && bij.BO * bjk.BO > ctl.thb_cutsq)) continue;
        ++R.c_angles[j];
        auto it = ff.thb.find({ti, tj, tk});
        if (it == ff.thb.end()) continue;
        for (const Angle3& P : it->second) {
          if (!(std::fabs(P.p_val1) > 0.001)) continue;
          const double e3ij = std::exp(-p_val3 * std::pow(BOA_ij, P.p_val4)), f7ij = 1.0 - e3ij;
          const double Cf7ij = p_val3 * P.p_val4 * std::pow(BOA_ij, P.p_val4 - 1.0) * e3ij;
          const double e3jk = std::exp(-p_val3 * std::pow(BOA_jk, P.p_val4)), f7jk = 1.0 - e3jk;
          const double Cf7jk = p_val3 * P.p_val4 * std::pow(BOA_jk, P.p_val4 - 1.0) * e3jk;
          const double ev7 = std::exp(-P.p_val7 * Dboc), trm8 = 1.0 + ev6 + ev7;
          const double f8 = p_val5 - ((p_val5 - 1.0) * (2.0 + ev6) / trm8);
          const double Cf8 = ((1.0 - p_val5) / (trm8 * trm8)) * (p_val6 * ev6 * trm8 - (2.0 + ev6) * (p_val6 * ev6 - P.p_val7 * ev7));
          const double th0 = (180.0 - P.theta_00 * (1.0 - std::exp(-p_val10 * (2.0 - SBO2)))) * 3.14159265 / 180.0;
          const double e2t = std::exp(-P.p_val2 * ((th0 - th) * (th0 - th)));
          const double e12t = P.p_val1 >= 0 ? P.p_val1 * (1.0 - e2t) : P.p_val1 * -e2t;
          const double CE1 = Cf7ij * f7jk * f8 * e12t, CE2 = Cf7jk * f7ij * f8 * e12t, CE3 = Cf8 * f7ij * f7jk * e12t;
          const double CE4 = -2.0 * P.p_val1 * P.p_val2 * f7ij * f7jk * f8 * e2t * (th0 - th);
          const double Cth0 = p_val10 * (P.theta_00 * 3.14159265 / 180.0) * std::exp(-p_val10 * (2.0 - SBO2));
          const double CE5 = -CE4 * Cth0 * CSBO2, CE6 = CE5 * dSBO1, CE7 = CE5 * dSBO2, CE8 = -CE4 / st;
          e_ang += f7ij * f7jk * f8 * e12t;
          const double ep2ij = std::exp(-p_pen2 * ((BOA_ij - 2.0) * (BOA_ij - 2.0))), ep2jk = std::exp(-p_pen2 * ((BOA_jk - 2.0) * (BOA_jk - 2.0)));
          const double ep3 = std::exp(-p_pen3 * aq[j].Delta), ep4 = std::exp(p_pen4 * aq[j].Delta), trm = 1.0 + ep3 + ep4;
          const double f9 = (2.0 + ep3) / trm;
          const double Cf9 = (-p_pen3 * ep3 * trm - (2.0 + ep3) * (-p_pen3 * ep3 + p_pen4 * ep4)) / (trm * trm);
          const double epen = P.p_pen1 * f9 * ep2ij * ep2jk;
          e_pen += epen;
          const double CEp1 = epen * Cf9 / f9, tp = -2.0 * p_pen2 * epen, CEp2 = tp * (BOA_ij - 2.0), CEp3 = tp * (BOA_jk - 2.0);
          const double ec2 = std::exp(p_coa2 * aq[j].Delta_val);
          const double dti = tbo[i] - BOA_ij, dtk = tbo[k] - BOA_jk;
          const double ecoa = P.p_coa1 / (1. + ec2) * std::exp(-p_coa3 * (dti * dti)) * std::exp(-p_coa3 * (dtk * dtk)) *
                              std::exp(-p_coa4 * ((BOA_ij - 1.5) * (BOA_ij - 1.5))) * std::exp(-p_coa4 * ((BOA_jk - 1.5) * (BOA_jk - 1.5)));
          e_coa += ecoa;
          const double CEc1 = -2 * p_coa4 * (BOA_ij - 1.5) * ecoa, CEc2 = -2 * p_coa4 * (BOA_jk - 1.5) * ecoa;
          const double CEc3 = -p_coa2 * ec2 * ecoa / (1 + ec2), CEc4 = -2 * p_coa3 * dti * ecoa, CEc5 = -2 * p_coa3 * dtk * ecoa;
          bij.Cdbo += CE1 + CEp2 + (CEc1 - CEc4);
          bjk.Cdbo += CE2 + CEp3 + (CEc2 - CEc5);
          CdD[j] += (CE3 + CE7) + CEp1 + CEc3;
          CdD[i] += CEc4;
          CdD[k] += CEc5;
          for (Bnd& bt : bj) {
            const double c3 = bt.BO * bt.BO * bt.BO;
            bt.Cdbo += CE6 * (c3 * c3 * bt.BO);
            bt.Cdbopi += CE5; bt.Cdbopi2 += CE5;
          }
          axpy(f[i], CE8, t.di); axpy(f[j], CE8, t.dj); axpy(f[k], CE8, t.dk);
        }
      }
    }
  }

  // torsion + 4-body conjugation: owned j, the tag order picks one end of each j-k bond
  const double p_tor2 = gp[23], p_tor3 = gp[24], p_tor4 = gp[25], p_cot2 = gp[27];
  const double MINS = 1e-10;
  auto tan_inv = [&](double s, double c) { return (s >= 0 && s <= MINS) ? c / MINS : ((s <= 0 && s >= -MINS) ? c / -MINS : c / s); };
  for (size_t j = 0; j < n; ++j) {
    const int tj = A.type[j];
    const double Dj = aq[j].Delta_boc;
    for (size_t pk = 0; pk < B[j].size(); ++pk) {
      Bnd& bjk = B[j][pk];
      const size_t k = (size_t)bjk.nbr;
      const double BOA_jk = bjk.BO - ctl.thb_cut;
      if (A.tag[j] > A.tag[k]) continue;
      if (A.tag[j] == A.tag[k]) {
        const double *xj = &A.x[3 * j], *xk = &A.x[3 * k];
        if (xk[2] < xj[2]) continue;
        if (xk[2] == xj[2] && xk[1] < xj[1]) continue;
        if (xk[2] == xj[2] && xk[1] == xj[1] && xk[0] < xj[0]) continue;
      }
      if (!(bjk.BO > ctl.thb_cut && !T[j][pk].empty())) continue;
      const size_t pj = (size_t)bjk.sym;
      if (T[k][pj].empty()) continue;
      const int tk = A.type[k];
      const double Dk = aq[k].Delta_boc, rjk = bjk.d;
      const double et2jk = std::exp(-p_tor2 * BOA_jk), ec2jk = std::exp(-p_cot2 * ((BOA_jk - 1.5) * (BOA_jk - 1.5)));
      const double et3 = std::exp(-p_tor3 * (Dj + Dk)), et4 = std::exp(p_tor4 * (Dj + Dk)), it34 = 1.0 / (1.0 + et3 + et4);
      const double f11 = (2.0 + et3) * it34;
      for (const Thb& pijk : T[j][pk]) {
        Bnd& bij = B[j][(size_t)pijk.pthb];
        if (!(bij.BO > ctl.thb_cut)) continue;
        const size_t i = (size_t)pijk.thb;
        const int ti = A.type[i];
        const double rij = bij.d, BOA_ij = bij.BO - ctl.thb_cut;
        const double sijk = std::sin(pijk.theta), cijk = std::cos(pijk.theta), tijk = tan_inv(sijk, cijk);
        const double et2ij = std::exp(-p_tor2 * BOA_ij), ec2ij = std::exp(-p_cot2 * ((BOA_ij - 1.5) * (BOA_ij - 1.5)));
        for (const Thb& pjkl : T[k][pj]) {
          const size_t l = (size_t)pjkl.thb;
          Bnd& bkl = B[k][(size_t)pjkl.pthb];
          const int tl = A.type[l];
          auto it = ff.fbp.find({ti, tj, tk, tl});
          if (!(i != l && it != ff.fbp.end() && bkl.BO > ctl.thb_cut && bij.BO * bjk.BO * bkl.BO > ctl.thb_cut)) continue;
          const Tors4& P = it->second;
          ++R.c_tors[j];
          const double rkl = bkl.d, BOA_kl = bkl.BO - ctl.thb_cut;
          const double sjkl = std::sin(pjkl.theta), cjkl = std::cos(pjkl.theta), tjkl = tan_inv(sjkl, cjkl);
          const V3 dli{A.x[3 * i] - A.x[3 * l], A.x[3 * i + 1] - A.x[3 * l + 1], A.x[3 * i + 2] - A.x[3 * l + 2]};
          const double rli = std::sqrt(dot(dli, dli));
          // dihedral angle and its derivative (reference Calculate_Omega, bonded.cpp:544-586)
          double s1 = sijk, c1 = cijk, s2 = sjkl, c2 = cjkl;
          if (s1 >= 0 && s1 <= MINS) s1 = MINS; else if (s1 <= 0 && s1 >= -MINS) s1 = -MINS;
          if (s2 >= 0 && s2 <= MINS) s2 = MINS; else if (s2 <= 0 && s2 >= -MINS) s2 = -MINS;
          const V3 &dij = bij.dv, &djk = bjk.dv, &dkl = bkl.dv;
          const double ucos = -dot(dij, djk) * dot(djk, dkl) + (rjk * rjk) * dot(dij, dkl);
          const double usin = -rjk * dot(dij, cross(djk, dkl));
          const double omega = std::atan2(usin, ucos);
          const double htra = rij + c1 * (rkl * c2 - rjk), htrb = rjk - rij * c1 - rkl * c2, htrc = rkl + c2 * (rij * c1 - rjk);
          const double hthd = rij * s1 * (rjk - rkl * c2), hthe = rkl * s2 * (rjk - rij * c1);
          const double hnra = rkl * s1 * s2, hnrc = rij * s1 * s2, hnhd = rij * rkl * c1 * s2, hnhe = rij * rkl * s1 * c2;
          const double tel = (rij * rij) + (rjk * rjk) + (rkl * rkl) - (rli * rli) - 2.0 * (rij * rjk * c1 - rij * rkl * c1 * c2 + rjk * rkl * c2);
          const double poem = 2.0 * rij * rkl * s1 * s2;
          double arg = tel / poem;
          if (arg > 1.0) arg = 1.0;
          if (arg < -1.0) arg = -1.0;
          V3 doi, doj, dok, dol;
          for (int c = 0; c < 3; ++c) {
            doi[c] = (htra - arg * hnra) / rij * dij[c] - dli[c];
            doj[c] = -(htra - arg * hnra) / rij * dij[c] + -htrb / rjk * djk[c];
            dok[c] = -(htrc - arg * hnrc) / rkl * dkl[c] + htrb / rjk * djk[c];
            dol[c] = (htrc - arg * hnrc) / rkl * dkl[c] + dli[c];
          }
          axpy(doi, -(hthd - arg * hnhd) / s1, pijk.dk); doi = scl(2.0 / poem, doi);
          axpy(doj, -(hthd - arg * hnhd) / s1, pijk.dj); axpy(doj, -(hthe - arg * hnhe) / s2, pjkl.di); doj = scl(2.0 / poem, doj);
          axpy(dok, -(hthd - arg * hnhd) / s1, pijk.di); axpy(dok, -(hthe - arg * hnhe) / s2, pjkl.dj); dok = scl(2.0 / poem, dok);
          axpy(dol, -(hthe - arg * hnhe) / s2, pjkl.dk); dol = scl(2.0 / poem, dol);
          const double cw = std::cos(omega), c2w = std::cos(2. * omega), c3w = std::cos(3. * omega);
          const double x = 2.0 - bjk.BO_pi - f11;
          const double et1 = std::exp(P.p_tor1 * (x * x));
          const double et2kl = std::exp(-p_tor2 * BOA_kl), ec2kl = std::exp(-p_cot2 * ((BOA_kl - 1.5) * (BOA_kl - 1.5)));
          const double fn10 = (1.0 - et2ij) * (1.0 - et2jk) * (1.0 - et2kl);
          const double CV = 0.5 * (P.V1 * (1.0 + cw) + P.V2 * et1 * (1.0 - c2w) + P.V3 * (1.0 + c3w));
          e_tor += fn10 * sijk * sjkl * CV;
          const double dfn11 = (-p_tor3 * et3 + (p_tor3 * et3 - p_tor4 * et4) * (2.0 + et3) * it34) * it34;
          const double CT1 = sijk * sjkl * CV;
          const double CT2 = -fn10 * 2.0 * P.p_tor1 * P.V2 * et1 * x * (1.0 - (cw * cw)) * sijk * sjkl;
          const double CT3 = CT2 * dfn11;
          const double CT4 = CT1 * p_tor2 * et2ij * (1.0 - et2jk) * (1.0 - et2kl);
          const double CT5 = CT1 * p_tor2 * (1.0 - et2ij) * et2jk * (1.0 - et2kl);
          const double CT6 = CT1 * p_tor2 * (1.0 - et2ij) * (1.0 - et2jk) * et2kl;
          const double cmn = -fn10 * CV, CT7 = cmn * sjkl * tijk, CT8 = cmn * sijk * tjkl;
          const double CT9 = fn10 * sijk * sjkl * (0.5 * P.V1 - 2.0 * P.V2 * et1 * cw + 1.5 * P.V3 * (c2w + 2.0 * (cw * cw)));
          const double fn12 = ec2ij * ec2jk * ec2kl;
          const double g = 1.0 + ((cw * cw) - 1.0) * sijk * sjkl;
          e_con += P.p_cot1 * fn12 * g;
          const double Cc = -2.0 * fn12 * P.p_cot1 * p_cot2 * g;
          const double CC1 = Cc * (BOA_ij - 1.5e0), CC2 = Cc * (BOA_jk - 1.5e0), CC3 = Cc * (BOA_kl - 1.5e0);
          const double CC4 = -P.p_cot1 * fn12 * ((cw * cw) - 1.0) * sjkl * tijk;
          const double CC5 = -P.p_cot1 * fn12 * ((cw * cw) - 1.0) * sijk * tjkl;
          const double CC6 = 2.0 * P.p_cot1 * fn12 * cw * sijk * sjkl;
          bjk.Cdbopi += CT2;
          CdD[j] += CT3; CdD[k] += CT3;
          bij.Cdbo += CT4 + CC1; bjk.Cdbo += CT5 + CC2; bkl.Cdbo += CT6 + CC3;
          axpy(f[i], CT7 + CC4, pijk.dk); axpy(f[j], CT7 + CC4, pijk.dj); axpy(f[k], CT7 + CC4, pijk.di);
          axpy(f[j], CT8 + CC5, pjkl.di); axpy(f[k], CT8 + CC5, pjkl.dj); axpy(f[l], CT8 + CC5, pjkl.dk);
          axpy(f[i], CT9 + CC6, doi); axpy(f[j], CT9 + CC6, doj); axpy(f[k], CT9 + CC6, dok); axpy(f[l], CT9 + CC6, dol);
        }
      }
    }
  }

  // hydrogen bonds (donor hydrogen j owned)
  if (ctl.hbond_cut > 0) {
    for (size_t j = 0; j < n; ++j) {
      const int tj = A.type[j];
      if (tj < 0 || ff.at[tj].p_hbond != 1) continue;
      std::vector<size_t> hl;
      for (size_t p = 0; p < B[j].size(); ++p) {
        const int ti = A.type[(size_t)B[j][p].nbr];
        if (ti < 0) continue;
        if (ff.at[ti].p_hbond == 2 && B[j][p].BO >= 1e-2) hl.push_back(p);
      }
      for (const HbE& hk : H[j]) {
        const size_t k = (size_t)hk.nbr;
        const int tk = A.type[k];
        if (tk < 0) continue;
        const double rjk = hk.d;
        const V3 djk = scl(hk.scl, hk.dv);
        for (size_t p : hl) {
          Bnd& bij = B[j][p];
          const size_t i = (size_t)bij.nbr;
          if (A.tag[i] == A.tag[k]) continue;
          const int ti = A.type[i];
          if (ti < 0) continue;
          auto it = ff.hbp.find({ti, tj, tk});
          if (it == ff.hbp.end() || it->second.r0 <= 0.0) continue;
          const Hb3& P = it->second;
          ++R.c_hb[j];
          double th, ct;
          theta_of(bij.dv, bij.d, djk, rjk, th, ct);
          V3 dti, dtj, dtk;
          dcos_of(bij.dv, bij.d, djk, rjk, dti, dtj, dtk);
          const double s2 = std::sin(th / 2.0);
          double s4 = s2 * s2; s4 *= s4;
          const double cx = 1.0 - ct;
          const double eh2 = std::exp(-P.p2 * bij.BO), eh3 = std::exp(-P.p3 * (P.r0 / rjk + rjk / P.r0 - 2.0));
          const double ehb = P.p1 * (1.0 - eh2) * eh3 * s4;
          e_hb += ehb;
          const double CH1 = P.p1 * P.p2 * eh2 * eh3 * s4, CH2 = -P.p1 / 2.0 * (1.0 - eh2) * eh3 * cx;
          const double CH3 = -P.p3 * (-P.r0 / (rjk * rjk) + 1.0 / P.r0) * ehb;
          bij.Cdbo += CH1;
          axpy(f[i], CH2, dti); axpy(f[j], CH2, dtj); axpy(f[k], CH2, dtk);
          axpy(f[j], -CH3 / rjk, djk); axpy(f[k], CH3 / rjk, djk);
        }
      }
    }
  }

  // bond-order derivative pass (reference Add_dBond_to_Forces, bonded.cpp:719-766)
  for (size_t i = 0; i < N; ++i)
    for (const Bnd& bij : B[i]) {
      const size_t j = (size_t)bij.nbr;
      if (!(i < j)) continue;
      const Bnd& bji = B[j][(size_t)bij.sym];
      double c = bij.Cdbo + bji.Cdbo;
      const double C1 = bij.C1dbo * c, C2 = bij.C2dbo * c, C3 = bij.C3dbo * c;
      c = bij.Cdbopi + bji.Cdbopi;
      const double P1 = bij.C1dbopi * c, P2 = bij.C2dbopi * c, P3 = bij.C3dbopi * c, P4 = bij.C4dbopi * c;
      c = bij.Cdbopi2 + bji.Cdbopi2;
      const double Q1 = bij.C1dbopi2 * c, Q2 = bij.C2dbopi2 * c, Q3 = bij.C3dbopi2 * c, Q4 = bij.C4dbopi2 * c;
      c = CdD[i] + CdD[j];
      const double D1 = bij.C1dbo * c, D2 = bij.C2dbo * c, D3 = bij.C3dbo * c;
      V3 t = scl(C1 + D1 + P2 + Q2, bij.dBOp);
      axpy(t, C2 + D2 + P3 + Q3, dDps[i]);
      axpy(t, P1, bij.dln_pi);
      axpy(t, Q1, bij.dln_pi2);
      axpy(f[i], 1.0, t);
      t = scl(-(C1 + D1 + P2 + Q2), bij.dBOp);
      axpy(t, C3 + D3 + P4 + Q4, dDps[j]);
      axpy(t, -P1, bij.dln_pi);
      axpy(t, -Q1, bij.dln_pi2);
      axpy(f[j], 1.0, t);
      for (const Bnd& bk : B[i]) axpy(f[(size_t)bk.nbr], -(C2 + D2 + P3 + Q3), bk.dBOp);
      for (const Bnd& bk : B[j]) axpy(f[(size_t)bk.nbr], -(C3 + D3 + P4 + Q4), bk.dBOp);
    }

  R.e[E_BOND] += e_bond; R.e[E_LP] += e_lp; R.e[E_OV] += e_ov; R.e[E_UN] += e_un;
  R.e[E_ANG] += e_ang; R.e[E_PEN] += e_pen; R.e[E_COA] += e_coa; R.e[E_TOR] += e_tor; R.e[E_CON] += e_con; R.e[E_HB] += e_hb;
  R.grad.assign(3 * N, 0.0);
  for (size_t i = 0; i < N; ++i) for (int c = 0; c < 3; ++c) R.grad[3 * i + c] = f[i][c];
  R.total_bo = tbo;
  long long nb = 0;
  for (size_t i = 0; i < N; ++i) nb += (long long)B[i].size();
  R.nbonds = nb / 2;
}

// ------------------------------------------------------------------------------------------------------------------
// taper (reference terms_host.hpp:17-39; LAMMPS reaxff_init_md.cpp / fix_qeq_reaxff.cpp init_taper)
// ------------------------------------------------------------------------------------------------------------------
struct Taper { double c[8]; };
Taper taper(double swa, double swb) {
  const double d1 = swb - swa, d7 = std::pow(d1, 7.0);
  const double a2 = swa * swa, a3 = swa * swa * swa, b2 = swb * swb, b3 = swb * swb * swb;
  Taper t;
  t.c[7] = 20.0 / d7;
  t.c[6] = -70.0 * (swa + swb) / d7;
  t.c[5] = 84.0 * (a2 + 3.0 * swa * swb + b2) / d7;
  t.c[4] = -35.0 * (a3 + 9.0 * a2 * swb + 9.0 * swa * b2 + b3) / d7;
  t.c[3] = 140.0 * (a3 * swb + 3.0 * a2 * b2 + swa * b3) / d7;
  t.c[2] = -210.0 * (a3 * b2 + a2 * b3) / d7;
  t.c[1] = 140.0 * a3 * b3 / d7;
  t.c[0] = (-35.0 * a3 * b2 * b2 + 21.0 * a2 * b3 * b2 - 7.0 * swa * b3 * b3 + b3 * b3 * swb) / d7;
  return t;
}
static void tap_eval(const Taper& T, double r, double& Tap, double& dTap) {
  const double* c = T.c;
  Tap = c[7] * r + c[6];
  Tap = Tap * r + c[5]; Tap = Tap * r + c[4]; Tap = Tap * r + c[3]; Tap = Tap * r + c[2]; Tap = Tap * r + c[1]; Tap = Tap * r + c[0];
  dTap = 7 * c[7] * r + 6 * c[6];
  dTap = dTap * r + 5 * c[5]; dTap = dTap * r + 4 * c[4]; dTap = dTap * r + 3 * c[3]; dTap = dTap * r + 2 * c[2];
  dTap += c[1] / r;
}

// nonbonded (reference src/cpu/nonbonded.cpp:18-82, terms.hpp:282-325)
void nonbonded(const FF& ff, double nonb, const Atoms& A, const Far& far, const double* q_owned, bool lg, Result& R) {
  const size_t N = A.nall(), n = A.nlocal;
  const double p_vdW1 = ff.gp[28];
  const int vt = ff.vdw_type;
  const Taper T = taper(ff.nonb_low, ff.nonb_cut);
  R.c_nb.assign(n, 0);
  double epol = 0;
  // per-thread gradient, energies and pair counts (a static schedule, summed in thread order below)
  const int nth = nthreads();
  std::vector<std::vector<double>> gth((size_t)nth, std::vector<double>(3 * N, 0.0));
  std::vector<double> evdw_t((size_t)nth, 0.0), eele_t((size_t)nth, 0.0);
  std::vector<long long> poo_t((size_t)nth, 0), pog_t((size_t)nth, 0), pself_t((size_t)nth, 0);
#pragma omp parallel for schedule(static) num_threads(nth)
  for (long ii = 0; ii < (long)n; ++ii) {
    const size_t i = (size_t)ii;
    const int tid = thread_id();
    double* g = gth[(size_t)tid].data();
    double& evdw = evdw_t[(size_t)tid];
    double& eele = eele_t[(size_t)tid];
    const int ti = A.type[i];
    if (ti < 0) continue;
    for (size_t p = far.start[i]; p < far.start[i + 1]; ++p) {
      const size_t j = (size_t)far.nbr[p];
      const int tj = A.type[j];
      if (tj < 0) continue;
      const double r = far.d[p];
      if (!(r <= nonb)) continue;
      const double* dv = &far.dv[3 * p];
      const int cls = pair_class(j, n, A.tag[i], A.tag[j], dv);
      if (!cls) continue;
      if (cls == 1) ++poo_t[(size_t)tid]; else if (cls == 2) ++pog_t[(size_t)tid]; else ++pself_t[(size_t)tid];
      ++R.c_nb[i];
      const Pair2& t = ff.P(ti, tj);
      double Tap, dTap;
      tap_eval(T, r, Tap, dTap);
      const double p1i = 1.0 / p_vdW1;
      double CEvd;
      if (vt == 1 || vt == 3) {
        const double pr = std::pow(r, p_vdW1), pg = std::pow(1.0 / t.gamma_w, p_vdW1), fn13 = std::pow(pr + pg, p1i);
        const double e1 = std::exp(t.alpha * (1.0 - fn13 / t.r_vdW)), e2 = std::exp(0.5 * t.alpha * (1.0 - fn13 / t.r_vdW));
        const double ev = t.D * (e1 - 2.0 * e2);
        evdw += Tap * ev;
        const double dfn13 = std::pow(pr + pg, p1i - 1.0) * std::pow(r, p_vdW1 - 2.0);
        CEvd = dTap * ev - Tap * t.D * (t.alpha / t.r_vdW) * (e1 - e2) * dfn13;
      } else {
        const double e1 = std::exp(t.alpha * (1.0 - r / t.r_vdW)), e2 = std::exp(0.5 * t.alpha * (1.0 - r / t.r_vdW));
        const double ev = t.D * (e1 - 2.0 * e2);
        evdw += Tap * ev;
        CEvd = dTap * ev - Tap * t.D * (t.alpha / t.r_vdW) * (e1 - e2) / r;
      }
      if (vt == 2 || vt == 3) {
        const double ec = t.ecore * std::exp(t.acore * (1.0 - (r / t.rcore)));
        evdw += Tap * ec;
        CEvd += dTap * ec + Tap * (-(t.acore / t.rcore) * ec) / r;
        if (lg) {
          const double r5 = std::pow(r, 5.0), r6 = std::pow(r, 6.0), re6 = std::pow(t.lgre, 6.0);
          const double elg = -(t.lgcij / (r6 + re6));
          evdw += Tap * elg;
          CEvd += dTap * elg + Tap * (-6.0 * elg * r5 / (r6 + re6)) / r;
        }
      }
      const double qi = q_owned[i], qj = q_owned[(size_t)A.owner[j]];
      const double d1 = r * r * r + t.gamma, d3 = std::pow(d1, 0.33333333333333);
      eele += 332.06371 * qi * qj * Tap / d3;
      const double CE = CEvd + 332.06371 * qi * qj * (dTap - Tap * r / d1) / d3;
      for (int c = 0; c < 3; ++c) { g[3 * i + c] += -CE * dv[c]; g[3 * j + c] += CE * dv[c]; }
    }
  }
  std::vector<double> g(3 * N, 0.0);
  double evdw = 0, eele = 0;
  for (int t = 0; t < nth; ++t) {
    for (size_t k = 0; k < 3 * N; ++k) g[k] += gth[(size_t)t][k];
    evdw += evdw_t[(size_t)t]; eele += eele_t[(size_t)t];
    R.pairs_oo += poo_t[(size_t)t]; R.pairs_og += pog_t[(size_t)t]; R.pairs_self += pself_t[(size_t)t];
  }
  for (size_t i = 0; i < n; ++i) {
    if (A.type[i] < 0) continue;
    const Atom1& s = ff.at[A.type[i]];
    epol += 23.02 * (s.chi * q_owned[i] + (s.eta / 2.) * (q_owned[i] * q_owned[i]));
  }
  R.e[E_VDW] += evdw; R.e[E_ELE] += eele; R.e[E_POL] += epol;
  if (R.grad.size() != 3 * N) R.grad.assign(3 * N, 0.0);
  for (size_t k = 0; k < 3 * N; ++k) R.grad[k] += g[k];
}

// ------------------------------------------------------------------------------------------------------------------
// charge equilibration: stock fix qeq/reaxff (LAMMPS @ 8de817dd fix_qeq_reaxff.cpp), single process
// ------------------------------------------------------------------------------------------------------------------
struct Qeq {
  // half matrix over owned rows in far-list order, the owner-computes rule of compute_H (EPSILON = SMALL = 1e-4)
  std::vector<size_t> start;
  std::vector<int> col;       // owned column (owner of j)
  std::vector<double> val;
  std::vector<double> eta;    // per owned atom (ffield eta x 2)
};

Qeq qeq_matrix(const FF& ff, const Atoms& A, const Far& far, double swa, double swb, const double* chi_t, const double* eta_t, const double* gam_t) {
  (void)chi_t;
  const size_t n = A.nlocal;
  const Taper T = taper(swa, swb);
  Qeq Q;
  Q.start.push_back(0);
  Q.eta.resize(n);
  const int nt = ff.n();
  std::vector<double> shld((size_t)nt * nt);
  for (int a = 0; a < nt; ++a) for (int b = 0; b < nt; ++b) shld[(size_t)a * nt + b] = std::pow(gam_t[a] * gam_t[b], -1.5);
  struct Ent { int col; double val; };
  auto make_row = [&](size_t i, std::vector<Ent>& row) {
    row.clear();
    for (size_t p = far.start[i]; p < far.start[i + 1]; ++p) {
      const size_t j = (size_t)far.nbr[p];
      const double* dv = &far.dv[3 * p];
      const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      if (!(r2 <= swb * swb)) continue;
      if (!pair_class(j, n, A.tag[i], A.tag[j], dv)) continue;
      const double r = std::sqrt(r2);
      double Tap = T.c[7] * r + T.c[6];
      Tap = Tap * r + T.c[5]; Tap = Tap * r + T.c[4]; Tap = Tap * r + T.c[3]; Tap = Tap * r + T.c[2]; Tap = Tap * r + T.c[1]; Tap = Tap * r + T.c[0];
      const double den = std::pow(r * r * r + shld[(size_t)A.type[i] * nt + A.type[j]], 1.0 / 3.0);
      row.push_back(Ent{A.owner[j], Tap * 14.4 / den});
    }
  };
  // rows of a chunk in parallel, appended in row order
  const size_t chunk = 4096;
  std::vector<std::vector<Ent>> rows(std::min(chunk, std::max<size_t>(n, 1)));
  for (size_t i0 = 0; i0 < n; i0 += chunk) {
    const long m = (long)std::min(chunk, n - i0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (long k = 0; k < m; ++k) make_row(i0 + (size_t)k, rows[(size_t)k]);
    for (long k = 0; k < m; ++k) {
      const size_t i = i0 + (size_t)k;
      Q.eta[i] = eta_t[A.type[i]];
      for (const Ent& e : rows[(size_t)k]) { Q.col.push_back(e.col); Q.val.push_back(e.val); }
      Q.start.push_back(Q.col.size());
    }
  }
  return Q;
}

static void qeq_mv(const Qeq& Q, const double* x, double* y, size_t n) {
  const int nth = nthreads();
  if (nth == 1) {   // the serial order
    for (size_t i = 0; i < n; ++i) y[i] = Q.eta[i] * x[i];
    for (size_t i = 0; i < n; ++i)
      for (size_t p = Q.start[i]; p < Q.start[i + 1]; ++p) {
        const size_t j = (size_t)Q.col[p];
        y[i] += Q.val[p] * x[j];
        y[j] += Q.val[p] * x[i];
      }
    return;
  }
  // per-thread partial products over a static split of the rows, summed in thread order onto the diagonal term
  std::vector<double> buf((size_t)nth * n, 0.0);
#pragma omp parallel for schedule(static) num_threads(nth)
  for (long ii = 0; ii < (long)n; ++ii) {
    const size_t i = (size_t)ii;
    double* yt = buf.data() + (size_t)thread_id() * n;
    for (size_t p = Q.start[i]; p < Q.start[i + 1]; ++p) {
      const size_t j = (size_t)Q.col[p];
      yt[i] += Q.val[p] * x[j];
      yt[j] += Q.val[p] * x[i];
    }
  }
  for (size_t i = 0; i < n; ++i) {
    double v = Q.eta[i] * x[i];
    for (int t = 0; t < nth; ++t) v += buf[(size_t)t * n + i];
    y[i] = v;
  }
}

// the stock CG loop (fix_qeq_reaxff.cpp CG; kept line for line by the reference at fix_qeq_reaxff_metal.cpp:90-123).
// Returns the iteration count as the stock fix does (the loop index at exit); rel = sqrt(sig_new)/|b| at exit.
int qeq_cg(const Qeq& Q, const double* b, double* x, size_t n, int imax, double tol, double* rel) {
  std::vector<double> r(n), d(n), q(n), p(n);
  qeq_mv(Q, x, q.data(), n);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  for (size_t i = 0; i < n; ++i) d[i] = r[i] * (1.0 / Q.eta[i]);
  double bn = 0, sn = 0;
  for (size_t i = 0; i < n; ++i) bn += b[i] * b[i];
  bn = std::sqrt(bn);
  for (size_t i = 0; i < n; ++i) sn += r[i] * d[i];
  int it;
  for (it = 1; it < imax && std::sqrt(sn) / bn > tol; ++it) {
    qeq_mv(Q, d.data(), q.data(), n);
    double tmp = 0;
    for (size_t i = 0; i < n; ++i) tmp += d[i] * q[i];
    const double alpha = sn / tmp;
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
    for (size_t i = 0; i < n; ++i) p[i] = r[i] * (1.0 / Q.eta[i]);
    const double so = sn;
    sn = 0;
    for (size_t i = 0; i < n; ++i) sn += r[i] * p[i];
    const double beta = sn / so;
    for (size_t i = 0; i < n; ++i) d[i] = p[i] + beta * d[i];
  }
  if (rel) *rel = std::sqrt(sn) / bn;
  return it;
}

}  // namespace orc

// ======================================================================================================================
// C interface (ctypes)
// ======================================================================================================================
using namespace orc;

namespace {
thread_local std::string g_err;
struct Handle { FF ff; Ctl ctl; bool lg = false; };
int fail(const std::exception& e) { g_err = e.what(); return -1; }
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ffield + optional control file text; returns an opaque handle or null (orc_last_error)
void* orc_create(const char* ffield, const char* control, int lgvdw) {
  try {
    auto* h = new Handle;
    h->ff = parse_ffield(ffield, lgvdw != 0);
    h->lg = lgvdw != 0;
    if (control && *control) h->ctl = parse_control(control);
    return h;
  } catch (const std::exception& e) { fail(e); return nullptr; }
}
void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

// threads of the parallel loops (1 = the serial oracle); the most this build can use
void orc_set_threads(int t) { g_threads = t < 1 ? 1 : t; }
int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// canonical table dump into buf (returns the full length; truncated if cap is too small)
long orc_table_dump(void* hp, int portable, char* buf, long cap) {
  const std::string d = table_dump(static_cast<Handle*>(hp)->ff, portable != 0);
  if (buf && cap > 0) { long m = std::min<long>(cap - 1, (long)d.size()); std::memcpy(buf, d.data(), (size_t)m); buf[m] = 0; }
  return (long)d.size();
}
int orc_num_types(void* hp) { return static_cast<Handle*>(hp)->ff.n(); }
// element name -> ffield type index (case-insensitive, exactly one match), -1 otherwise
int orc_match(void* hp, const char* el) {
  const FF& f = static_cast<Handle*>(hp)->ff;
  std::string a;
  for (const char* c = el; *c; ++c) a.push_back((char)std::tolower((unsigned char)*c));
  int hit = -1, cnt = 0;
  for (int i = 0; i < f.n(); ++i) {
    std::string b;
    for (char c : f.at[i].name) b.push_back((char)std::tolower((unsigned char)c));
    if (a == b) { hit = i; ++cnt; }
  }
  return cnt == 1 ? hit : -1;
}
// per-type chi, eta (x2 as stored), gamma (what fix qeq/reaxff extracts from the pair style)
void orc_qeq_params(void* hp, double* chi, double* eta, double* gamma) {
  const FF& f = static_cast<Handle*>(hp)->ff;
  for (int i = 0; i < f.n(); ++i) { chi[i] = f.at[i].chi; eta[i] = f.at[i].eta; gamma[i] = f.at[i].gamma; }
}
// cutoffs: out[0]=nonb_cut out[1]=bond_cut out[2]=hbond_cut out[3]=thb_cut out[4]=thb_cutsq out[5]=bo_cut out[6]=nonb_low
void orc_cutoffs(void* hp, double* out) {
  const Handle& h = *static_cast<Handle*>(hp);
  out[0] = h.ff.nonb_cut; out[1] = h.ctl.bond_cut; out[2] = h.ctl.hbond_cut; out[3] = h.ctl.thb_cut; out[4] = h.ctl.thb_cutsq;
  out[5] = h.ff.bo_cut; out[6] = h.ff.nonb_low;
}

// One energy/force evaluation of n owned atoms (periodic images built here, ghost shell `shell` <= 0 -> the required one).
// box: origin[3], a[3], b[3], c[3]; periodic[3]; q: n charges (NULL -> the nonbonded terms are skipped).
// Outputs: energies[13] (bond, ov, un, lp, ang, pen, coa, hb, tor, con, vdW, ele, pol), grad[3n] (dE/dx folded onto owners),
// counts[8] = pairs oo, og, self, bonds, nall, sum angle sets, sum torsions, sum hbonds; census (may be NULL): 7*n ints in the order
// bonds, angle_sets, torsions, hbonds, nonbonded, sbo_region, lp_trunc; total_bo (may be NULL): n; sbo (may be NULL): n, the SBO
// value of each owned atom (the quantity the sbo_region branch tests).
int orc_compute(void* hp, long n, const double* x, const int* type, const long long* tag, const double* box, const int* periodic, const double* q,
                double shell, int enobonds, double* energies, double* grad, long long* counts, int* census, double* total_bo, double* sbo) {
  try {
    const Handle& h = *static_cast<Handle*>(hp);
    Box b;
    for (int c = 0; c < 3; ++c) b.o[c] = box[c];
    for (int k = 0; k < 3; ++k) for (int c = 0; c < 3; ++c) b.v[k][c] = box[3 + 3 * k + c];
    for (int d = 0; d < 3; ++d) b.per[d] = periodic[d] != 0;
    const double nonb = h.ff.nonb_cut, bond = h.ctl.bond_cut, hb = h.ctl.hbond_cut;
    const double need = std::max({nonb, hb, 2.0 * bond});
    std::vector<int64_t> tg(tag, tag + n);
    Atoms A = expand(b, x, type, tg.data(), (size_t)n, shell > 0 ? shell : need);
    Far F = far_list(A, nonb, bond);
    Result R;
    bonded(h.ff, h.ctl, A, F, enobonds != 0, R);
    if (q) nonbonded(h.ff, nonb, A, F, q, h.lg, R);
    for (int t = 0; t < NTERM; ++t) energies[t] = R.e[t];
    for (long i = 0; i < 3 * n; ++i) grad[i] = 0;
    for (size_t i = 0; i < A.nall(); ++i) for (int c = 0; c < 3; ++c) grad[3 * (size_t)A.owner[i] + c] += R.grad[3 * i + c];
    if (counts) {
      long long sa = 0, st = 0, sh = 0;
      for (long i = 0; i < n; ++i) { sa += R.c_angles[i]; st += R.c_tors[i]; sh += R.c_hb[i]; }
      counts[0] = R.pairs_oo; counts[1] = R.pairs_og; counts[2] = R.pairs_self; counts[3] = R.nbonds; counts[4] = (long long)A.nall();
      counts[5] = sa; counts[6] = st; counts[7] = sh;
    }
    if (census) {
      const std::vector<int>* cs[7] = {&R.c_bonds, &R.c_angles, &R.c_tors, &R.c_hb, &R.c_nb, &R.c_sbo, &R.c_lp};
      for (int k = 0; k < 7; ++k) for (long i = 0; i < n; ++i) census[k * n + i] = (cs[k]->size() == (size_t)n) ? (*cs[k])[i] : 0;
    }
    if (total_bo) for (long i = 0; i < n; ++i) total_bo[i] = R.total_bo[i];
    if (sbo) for (long i = 0; i < n; ++i) sbo[i] = R.sbo.size() == (size_t)n ? R.sbo[i] : 0.0;
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

// Charge equilibration of n owned atoms as stock fix qeq/reaxff: H over the images within swb, s and t solved by the stock
// CG from the warm starts s0/t0 (zeros on the first step), q = s - (sum s / sum t) t. chi/eta/gamma per ffield type (eta as stored, x2).
// Outputs q[n], s[n], t[n] (the solutions, i.e. the next history entry), info[4] = iters_s, iters_t, rel_s, rel_t.
int orc_qeq(void* hp, long n, const double* x, const int* type, const long long* tag, const double* box, const int* periodic,
            const double* chi, const double* eta, const double* gamma, double swa, double swb, double tol, int imax,
            const double* s0, const double* t0, double* q, double* s, double* t, double* info) {
  try {
    const Handle& h = *static_cast<Handle*>(hp);
    Box b;
    for (int c = 0; c < 3; ++c) b.o[c] = box[c];
    for (int k = 0; k < 3; ++k) for (int c = 0; c < 3; ++c) b.v[k][c] = box[3 + 3 * k + c];
    for (int d = 0; d < 3; ++d) b.per[d] = periodic[d] != 0;
    const double nonb = h.ff.nonb_cut, bond = h.ctl.bond_cut, hb = h.ctl.hbond_cut;
    const double need = std::max({nonb, hb, 2.0 * bond, swb});
    std::vector<int64_t> tg(tag, tag + n);
    Atoms A = expand(b, x, type, tg.data(), (size_t)n, need);
    Far F = far_list(A, std::max(nonb, swb), bond);
    Qeq Q = qeq_matrix(h.ff, A, F, swa, swb, chi, eta, gamma);
    std::vector<double> bs(n), bt(n);
    for (long i = 0; i < n; ++i) { bs[i] = -chi[type[i]]; bt[i] = -1.0; s[i] = s0 ? s0[i] : 0.0; t[i] = t0 ? t0[i] : 0.0; }
    double rs = 0, rt = 0;
    const int is = qeq_cg(Q, bs.data(), s, (size_t)n, imax, tol, &rs);
    const int it = qeq_cg(Q, bt.data(), t, (size_t)n, imax, tol, &rt);
    double ss = 0, st = 0;
    for (long i = 0; i < n; ++i) { ss += s[i]; st += t[i]; }
    const double u = ss / st;
    for (long i = 0; i < n; ++i) q[i] = s[i] - u * t[i];
    info[0] = is; info[1] = it; info[2] = rs; info[3] = rt;
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

// Bond table of the owned atoms after one bonded evaluation (the reference's BondTable of bonded.cpp:782-792, partners as owned
// atom indices): start[n+1]; owner[], bo[] (capacity cap; the count is returned, -1 on error, and nothing is written past cap);
// total_bo[n], nlp[n] (may be NULL).
long orc_bond_table(void* hp, long n, const double* x, const int* type, const long long* tag, const double* box, const int* periodic,
                    int enobonds, int* start, int* owner, double* bo, long cap, double* total_bo, double* nlp) {
  try {
    const Handle& h = *static_cast<Handle*>(hp);
    Box b;
    for (int c = 0; c < 3; ++c) b.o[c] = box[c];
    for (int k = 0; k < 3; ++k) for (int c = 0; c < 3; ++c) b.v[k][c] = box[3 + 3 * k + c];
    for (int d = 0; d < 3; ++d) b.per[d] = periodic[d] != 0;
    const double nonb = h.ff.nonb_cut, bond = h.ctl.bond_cut, hb = h.ctl.hbond_cut;
    std::vector<int64_t> tg(tag, tag + n);
    Atoms A = expand(b, x, type, tg.data(), (size_t)n, std::max({nonb, hb, 2.0 * bond}));
    Far F = far_list(A, nonb, bond);
    Result R;
    bonded(h.ff, h.ctl, A, F, enobonds != 0, R);
    const long nb = R.bt_start[(size_t)n];
    for (long i = 0; i <= n; ++i) start[i] = R.bt_start[(size_t)i];
    for (long k = 0; k < std::min(nb, cap); ++k) { owner[k] = R.bt_owner[(size_t)k]; bo[k] = R.bt_bo[(size_t)k]; }
    if (total_bo) for (long i = 0; i < n; ++i) total_bo[i] = R.total_bo[(size_t)i];
    if (nlp) for (long i = 0; i < n; ++i) nlp[i] = R.bt_nlp[(size_t)i];
    return nb;
  } catch (const std::exception& e) { fail(e); return -1; }
}

// y = (diag(eta) + H) x with the stock half matrix (checker for the device matvec)
int orc_qeq_matvec(void* hp, long n, const double* x, const int* type, const long long* tag, const double* box, const int* periodic,
                   const double* chi, const double* eta, const double* gamma, double swa, double swb, const double* v, double* y) {
  try {
    const Handle& h = *static_cast<Handle*>(hp);
    Box b;
    for (int c = 0; c < 3; ++c) b.o[c] = box[c];
    for (int k = 0; k < 3; ++k) for (int c = 0; c < 3; ++c) b.v[k][c] = box[3 + 3 * k + c];
    for (int d = 0; d < 3; ++d) b.per[d] = periodic[d] != 0;
    const double need = std::max({h.ff.nonb_cut, h.ctl.hbond_cut, 2.0 * h.ctl.bond_cut, swb});
    std::vector<int64_t> tg(tag, tag + n);
    Atoms A = expand(b, x, type, tg.data(), (size_t)n, need);
    Far F = far_list(A, std::max(h.ff.nonb_cut, swb), h.ctl.bond_cut);
    Qeq Q = qeq_matrix(h.ff, A, F, swa, swb, chi, eta, gamma);
    qeq_mv(Q, v, y, (size_t)n);
    return 0;
  } catch (const std::exception& e) { return fail(e); }
}

}  // extern "C"
