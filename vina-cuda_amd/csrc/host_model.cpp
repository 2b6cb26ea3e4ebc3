// host_model.cpp -- see host_model.h for the reference functions this follows.
#include "host_model.h"

#include <algorithm>
#include <array>
#include <functional>
#include <iomanip>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <unordered_map>

namespace vd {

// ---------------------------------------------------------------------------------------
// Atom kinds (V11/lib/atom_constants.h:117-304)
namespace {
enum { E_H = 0, E_C, E_N, E_O, E_S, E_P, E_F, E_Cl, E_Br, E_I, E_Met, E_NONE };
enum { A_C = 0, A_A, A_N, A_O, A_P, A_S, A_H, A_F, A_I, A_NA, A_OA, A_SA, A_HD, A_Mg, A_Mn, A_Zn, A_Ca, A_Fe, A_Cl, A_Br, A_NONE };
struct Kind { const char* name; double cov; int el; };
const Kind kKinds[A_NONE] = {
    {"C", 0.77, E_C}, {"A", 0.77, E_C}, {"N", 0.75, E_N}, {"O", 0.73, E_O}, {"P", 1.06, E_P},
    {"S", 1.02, E_S}, {"H", 0.37, E_H}, {"F", 0.71, E_F}, {"I", 1.33, E_I}, {"NA", 0.75, E_N},
    {"OA", 0.73, E_O}, {"SA", 1.02, E_S}, {"HD", 0.37, E_H}, {"Mg", 1.30, E_Met}, {"Mn", 1.39, E_Met},
    {"Zn", 1.31, E_Met}, {"Ca", 1.74, E_Met}, {"Fe", 1.25, E_Met}, {"Cl", 0.99, E_Cl}, {"Br", 1.14, E_Br}};
const double kMetalCov = 1.75;
enum { X_C_H = 0, X_C_P, X_N_P, X_N_D, X_N_A, X_N_DA, X_O_P, X_O_D, X_O_A, X_O_DA, X_S_P, X_P_P,
       X_F_H, X_Cl_H, X_Br_H, X_I_H, X_Met_D };

double max_cov() { double m = 0; for (auto& k : kKinds) m = std::max(m, k.cov); return m; }
int ad_from_name(const std::string& s) {
  for (int i = 0; i < A_NONE; ++i) if (s == kKinds[i].name) return i;
  return s == "Se" ? A_S : A_NONE;
}
bool is_non_ad_metal(const std::string& s) {
  static const char* names[] = {"Cu", "Fe", "Na", "K", "Hg", "Co", "U", "Cd", "Ni"};
  for (const char* n : names) if (s == n) return true;
  return false;
}
double sq(double x) { return x * x; }
double d2(const double* a, const double* b) { return sq(a[0] - b[0]) + sq(a[1] - b[1]) + sq(a[2] - b[2]); }

struct TypedAtom { int ad = A_NONE, el = E_NONE, xs = VD_XS_SIZE; double xyz[3]; };
bool heteroatom(const TypedAtom& a) {
  return (a.ad != A_A && a.ad != A_C && a.ad != A_H && a.ad != A_HD && a.ad < A_NONE) || a.xs == X_Met_D;
}
double cov_radius(const TypedAtom& a) { return a.ad < A_NONE ? kKinds[a.ad].cov : kMetalCov; }

// XS type from element and bonded neighbours (V11/lib/model.cpp:771-797)
template <class Bonded>
void assign_xs(std::vector<TypedAtom>& at, const Bonded& bonded) {
  for (size_t i = 0; i < at.size(); ++i) {
    TypedAtom& a = at[i];
    a.el = a.ad < A_NONE ? kKinds[a.ad].el : (a.xs == X_Met_D ? E_Met : E_NONE);
    bool acc = (a.ad == A_OA || a.ad == A_NA), hd = false, het = false;
    for (int j : bonded(i)) { if (at[j].ad == A_HD) hd = true; if (heteroatom(at[j])) het = true; }
    bool don = (a.el == E_Met) || hd;
    switch (a.el) {
      case E_C: a.xs = het ? X_C_P : X_C_H; break;
      case E_N: a.xs = acc && don ? X_N_DA : acc ? X_N_A : don ? X_N_D : X_N_P; break;
      case E_O: a.xs = acc && don ? X_O_DA : acc ? X_O_A : don ? X_O_D : X_O_P; break;
      case E_S: a.xs = X_S_P; break;
      case E_P: a.xs = X_P_P; break;
      case E_F: a.xs = X_F_H; break;
      case E_Cl: a.xs = X_Cl_H; break;
      case E_Br: a.xs = X_Br_H; break;
      case E_I: a.xs = X_I_H; break;
      case E_Met: a.xs = X_Met_D; break;
      default: break;
    }
  }
}

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}
// Fixed-column PDBQT atom record (V11/lib/parse_pdbqt.cpp:101-118)
bool read_atom_record(const std::string& s, TypedAtom& a, unsigned& serial) {
  if (s.size() < 54) throw parse_error("ATOM line too short: " + s);
  try {
    serial = (unsigned)std::stoul(s.substr(6, 5));
    a.xyz[0] = std::stod(s.substr(30, 8)); a.xyz[1] = std::stod(s.substr(38, 8)); a.xyz[2] = std::stod(s.substr(46, 8));
  } catch (...) { throw parse_error("ATOM syntax incorrect: " + s); }
  std::string name = s.size() >= 78 ? trim(s.substr(77)) : std::string();
  a.ad = ad_from_name(name);
  if (is_non_ad_metal(name)) a.xs = X_Met_D;
  if (!(a.ad < A_NONE || a.xs == X_Met_D)) throw parse_error("\"" + name + "\" is not a valid AutoDock type");
  return true;
}
bool tag(const std::string& s, const char* t) { return s.compare(0, std::strlen(t), t) == 0; }
}  // namespace

const double kXsRadius[VD_XS_SIZE] = {1.9, 1.9, 1.8, 1.8, 1.8, 1.8, 1.7, 1.7, 1.7, 1.7, 2.0, 2.1, 1.5, 1.8, 2.0, 2.2, 1.2};
bool xs_is_hydrophobic(int t) { return t == X_C_H || t == X_F_H || t == X_Cl_H || t == X_Br_H || t == X_I_H; }
bool xs_is_acceptor(int t) { return t == X_N_A || t == X_N_DA || t == X_O_A || t == X_O_DA; }
bool xs_is_donor(int t) { return t == X_N_D || t == X_N_DA || t == X_O_D || t == X_O_DA || t == X_Met_D; }

// ---------------------------------------------------------------------------------------
// Scoring function and LUT
double score_terms(int t1, int t2, double r, const double w[5]) {
  const double d = r - (kXsRadius[t1] + kXsRadius[t2]);
  auto step = [](double bad, double good, double x) {  // slope_step, everything.cpp:128-139
    if (bad < good) { if (x <= bad) return 0.0; if (x >= good) return 1.0; }
    else { if (x >= bad) return 0.0; if (x <= good) return 1.0; }
    return (x - bad) / (good - bad);
  };
  double s = w[0] * std::exp(-sq(d / 0.5)) + w[1] * std::exp(-sq((d - 3.0) / 2.0));
  if (d < 0) s += w[2] * d * d;
  if (xs_is_hydrophobic(t1) && xs_is_hydrophobic(t2)) s += w[3] * step(1.5, 0.5, d);
  if ((xs_is_donor(t1) && xs_is_acceptor(t2)) || (xs_is_donor(t2) && xs_is_acceptor(t1))) s += w[4] * step(0.0, -0.7, d);
  return s;
}

void Precalculated::build(const double w[5]) {
  const int n = VD_LUT_BINS;
  smooth_e.assign((size_t)VD_NUM_TYPE_PAIRS * n, 0.0);
  smooth_dor.assign(smooth_e.size(), 0.0);
  fast.assign(smooth_e.size(), 0.0);
  std::vector<double> r(n);
  for (int i = 0; i < n; ++i) r[i] = std::sqrt(i / 32.0);
  for (int t2 = 0; t2 < VD_XS_SIZE; ++t2)
    for (int t1 = 0; t1 <= t2; ++t1) {
      const size_t base = (size_t)type_pair_index(t1, t2) * n;
      double* e = &smooth_e[base];
      for (int i = 0; i < n; ++i) e[i] = score_terms(t1, t2, r[i], w);
      for (int i = 0; i < n; ++i) {
        smooth_dor[base + i] = (i == 0 || i == n - 1) ? 0.0 : (e[i + 1] - e[i - 1]) / ((r[i + 1] - r[i - 1]) * r[i]);
        fast[base + i] = 0.5 * (e[i] + (i + 1 < n ? e[i + 1] : 0.0));
      }
    }
}

void Precalculated::grid_rows(std::vector<float>& out) const {
  const size_t rows = (size_t)VD_XS_SIZE * VD_LUT_BINS;
  out.assign(rows * VD_GRID_ROW + rows, 0.0f);
  for (int t1 = 0; t1 < VD_XS_SIZE; ++t1)
    for (int b = 0; b < VD_LUT_BINS; ++b) {
      const size_t r = (size_t)t1 * VD_LUT_BINS + b;
      for (int t2 = 0; t2 < VD_XS_SIZE; ++t2) {
        const float v = (float)fast[(size_t)type_pair_index(t1, t2) * VD_LUT_BINS + b];
        if (t2 < VD_GRID_ROW) out[r * VD_GRID_ROW + t2] = v;
        else out[rows * VD_GRID_ROW + r] = v;
      }
    }
}

void Precalculated::pair_bins(std::vector<float>& out) const {
  // one extra all-zero row after the 153 type pairs, whose first entry is VD_ZERO_QUAD: empty pair
  // slots and pairs beyond the cutoff read it and contribute nothing.  A pair inside the cutoff
  // has r^2 < 64, so 32 r^2 < 2048 (scaling by 32 is exact) and its bin is at most 2047: the
  // entry after it (bin + 1 <= 2048) lies in the same row.
  out.assign(((size_t)VD_NUM_TYPE_PAIRS + 1) * VD_LUT_BINS * 2, 0.0f);
  for (size_t i = 0; i < (size_t)VD_NUM_TYPE_PAIRS * VD_LUT_BINS; ++i) {
    out[2 * i] = (float)smooth_e[i];
    out[2 * i + 1] = (float)smooth_dor[i];
  }
}

// ---------------------------------------------------------------------------------------
// Receptor (V11/lib/parse_pdbqt.cpp:258-284 + model.cpp:644-797 with every pair FIXED).
// Bond test: r^2 < (1.1 (cov_i + cov_j))^2 and no third atom nearer to both ends.
void parse_receptor_pdbqt(const std::string& text, std::vector<double>& xyz, std::vector<int>& xs) {
  std::vector<TypedAtom> at;
  std::istringstream in(text);
  std::string s;
  int line = 0;
  while (std::getline(in, s)) {
    ++line;
    if (tag(s, "ATOM  ") || tag(s, "HETATM")) { TypedAtom a; unsigned num; read_atom_record(s, a, num); at.push_back(a); }
    else if (s.empty() || tag(s, "TER") || tag(s, "WARNING") || tag(s, "REMARK") || trim(s).empty()) {}
    else if (tag(s, "MODEL")) throw parse_error("multi-MODEL receptor input");
    else throw parse_error("unknown receptor tag at line " + std::to_string(line));
  }
  const int n = (int)at.size();
  const double mc = max_cov();
  // cell list for neighbour candidates within 1.1 * (cov_i + max_cov)
  const double cell = 1.1 * 2 * mc + 0.01;
  std::unordered_map<long long, std::vector<int>> cells;
  auto key = [&](long long x, long long y, long long z) { return (x * 73856093LL) ^ (y * 19349663LL) ^ (z * 83492791LL); };
  std::vector<long long> ci(3 * n);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) ci[3 * i + k] = (long long)std::floor(at[i].xyz[k] / cell);
    cells[key(ci[3 * i], ci[3 * i + 1], ci[3 * i + 2])].push_back(i);
  }
  std::vector<std::vector<int>> bonds(n);
  std::vector<int> rel;
  for (int i = 0; i < n; ++i) {
    const double ri = at[i].ad < A_NONE ? kKinds[at[i].ad].cov : mc;
    const double rcut2 = sq(1.1 * (ri + mc));
    rel.clear();
    for (int dx = -1; dx <= 1; ++dx) for (int dy = -1; dy <= 1; ++dy) for (int dz = -1; dz <= 1; ++dz) {
      auto it = cells.find(key(ci[3 * i] + dx, ci[3 * i + 1] + dy, ci[3 * i + 2] + dz));
      if (it == cells.end()) continue;
      for (int j : it->second) if (j != i && d2(at[i].xyz, at[j].xyz) < rcut2) rel.push_back(j);
    }
    for (int j : rel) {
      if (j <= i) continue;
      const double r2 = d2(at[i].xyz, at[j].xyz);
      if (r2 >= sq(1.1 * (cov_radius(at[i]) + cov_radius(at[j])))) continue;
      bool between = false;
      for (int c : rel)
        if (c != j && d2(at[i].xyz, at[c].xyz) < r2 && d2(at[j].xyz, at[c].xyz) < r2) { between = true; break; }
      if (!between) { bonds[i].push_back(j); bonds[j].push_back(i); }
    }
  }
  assign_xs(at, [&](size_t i) -> const std::vector<int>& { return bonds[i]; });
  xyz.resize(3 * (size_t)n); xs.resize(n);
  for (int i = 0; i < n; ++i) { for (int k = 0; k < 3; ++k) xyz[3 * i + k] = at[i].xyz[k]; xs[i] = at[i].xs; }
}

// ---------------------------------------------------------------------------------------
// Ligand.  The PDBQT ROOT/BRANCH nesting is read into fragments; each fragment becomes a
// rigid node in DFS preorder.  The BRANCH 'to' atom sits in the parent's node (it lies on
// the rotation axis) and is the child's origin (V11/lib/parse_pdbqt.cpp:157-245, 406-447).
namespace {
struct Frag {
  struct Entry { TypedAtom a; unsigned serial; int ctx_line; std::string line; std::vector<std::unique_ptr<Frag>> kids; };
  std::vector<Entry> entries;
  int axis_atom = -1;       // entry index of the BRANCH 'to' atom (-1 for ROOT)
  int from_index = -1, to_index = -1;  // ligand atom indices once flattened
  bool trivial() const {    // parsing_struct::essentially_empty, parse_pdbqt.cpp:230-240
    for (size_t i = 0; i < entries.size(); ++i) {
      if (axis_atom >= 0 && (int)i != axis_atom) return false;
      if (!entries[i].kids.empty()) return false;
    }
    return true;
  }
};

struct LigandReader {
  std::vector<std::string> lines; size_t at = 0;
  std::vector<std::string>* ctx;
  bool next(std::string& s) {
    while (at < lines.size()) {
      s = lines[at++];
      ctx->push_back(s);
      if (s.empty() || tag(s, "REMARK") || tag(s, "WARNING") || trim(s).empty()) continue;
      return true;
    }
    return false;
  }
  void atom_into(Frag& f, const std::string& s, unsigned to) {
    Frag::Entry e; read_atom_record(s, e.a, e.serial); e.line = s; e.ctx_line = (int)ctx->size() - 1;
    if (e.serial == to && to != 0) f.axis_atom = (int)f.entries.size();
    f.entries.push_back(std::move(e));
  }
  void branch(Frag& owner, const std::string& s) {
    unsigned from = 0, to = 0;
    { std::istringstream in(s.substr(6)); if (!(in >> from >> to)) throw parse_error("bad BRANCH line: " + s); }
    Frag::Entry* host = nullptr;
    for (auto& e : owner.entries) if (e.serial == from) host = &e;
    if (!host) throw parse_error("No atom number " + std::to_string(from) + " in this branch");
    host->kids.emplace_back(new Frag());
    Frag& f = *host->kids.back();
    std::string t;
    while (next(t)) {
      if (tag(t, "BRANCH")) branch(f, t);
      else if (tag(t, "ENDBRANCH")) {
        unsigned a = 0, b = 0; std::istringstream in(t.substr(9)); in >> a >> b;
        if (a != from || b != to) throw parse_error("Inconsistent branch numbers");
        if (f.axis_atom < 0) throw parse_error("Atom " + std::to_string(to) + " has not been found in this branch");
        return;
      } else if (tag(t, "ATOM  ") || tag(t, "HETATM")) atom_into(f, t, to);
      else throw parse_error("Unknown or inappropriate tag: " + t);
    }
    throw parse_error("unterminated BRANCH");
  }
};

// Flattening: atom order and node ranges as postprocess_branch produces them.
struct Flattener {
  Ligand& L;
  std::vector<std::vector<unsigned char>> mob;   // 0 fixed, 1 rotor, 2 variable (filled later)
  std::vector<std::tuple<int, int, int>> fixed_to_node;
  std::vector<std::pair<int, int>> rotors;
  explicit Flattener(Ligand& l) : L(l) {}
  int push(Frag::Entry& e, const double* origin) {
    for (auto& k : e.kids) k->from_index = (int)L.atoms.size();
    LigAtom a; a.ad = e.a.ad; a.xs = e.a.xs; a.serial = e.serial; a.line = e.line; a.ctx_line = e.ctx_line;
    for (int k = 0; k < 3; ++k) { a.xyz[k] = e.a.xyz[k]; a.local[k] = e.a.xyz[k] - origin[k]; }
    L.atoms.push_back(a);
    return (int)L.atoms.size() - 1;
  }
  void node(Frag& f, int idx) {
    const double origin[3] = {L.nodes[idx].origin0[0], L.nodes[idx].origin0[1], L.nodes[idx].origin0[2]};
    L.nodes[idx].begin = (int)L.atoms.size();
    for (size_t i = 0; i < f.entries.size(); ++i) {
      if ((int)i != f.axis_atom) push(f.entries[i], origin);
      for (auto& k : f.entries[i].kids)
        if (!k->entries.empty()) { k->to_index = (int)L.atoms.size(); push(k->entries[k->axis_atom], origin); }
    }
    L.nodes[idx].end = (int)L.atoms.size();
    for (int i = L.nodes[idx].begin; i < L.nodes[idx].end; ++i) L.atoms[i].node = idx;
    fixed_to_node.emplace_back(f.from_index, L.nodes[idx].begin, L.nodes[idx].end);
    fixed_to_node.emplace_back(f.to_index, L.nodes[idx].begin, L.nodes[idx].end);
    if (f.from_index >= 0 && f.to_index >= 0) rotors.emplace_back(f.from_index, f.to_index);
    for (auto& e : f.entries)
      for (auto& k : e.kids) {
        if (k->trivial()) continue;
        LigNode c; c.parent = idx; c.depth = L.nodes[idx].depth + 1;
        const double* o = k->entries[k->axis_atom].a.xyz;
        double ax[3], nrm = 0;
        for (int j = 0; j < 3; ++j) { c.origin0[j] = o[j]; ax[j] = o[j] - e.a.xyz[j]; nrm += ax[j] * ax[j]; c.rel_origin[j] = o[j] - origin[j]; }
        nrm = std::sqrt(nrm);
        if (nrm < std::numeric_limits<double>::epsilon()) throw parse_error("zero-length rotatable bond");
        for (int j = 0; j < 3; ++j) c.rel_axis[j] = ax[j] / nrm;
        L.nodes.push_back(c);
        node(*k, (int)L.nodes.size() - 1);
      }
  }
};
}  // namespace

Ligand parse_ligand_pdbqt(const std::string& text) {
  Ligand L;
  LigandReader R; R.ctx = &L.context;
  { std::istringstream in(text); std::string s; while (std::getline(in, s)) { if (!s.empty() && s.back() == '\r') s.pop_back(); R.lines.push_back(s); } }
  Frag root; std::string s; bool torsdof = false;
  if (!R.next(s) || !tag(s, "ROOT")) throw parse_error("expected ROOT");
  while (true) {
    if (!R.next(s)) throw parse_error("unterminated ROOT");
    if (tag(s, "ENDROOT")) break;
    if (tag(s, "ATOM  ") || tag(s, "HETATM")) R.atom_into(root, s, 0);
    else throw parse_error("Unknown or inappropriate tag in ROOT: " + s);
  }
  while (R.next(s)) {
    if (tag(s, "BRANCH")) R.branch(root, s);
    else if (tag(s, "TORSDOF")) {
      if (torsdof) throw parse_error("TORSDOF can occur only once");
      std::istringstream in(s.substr(7)); if (!(in >> L.torsdof)) throw parse_error("bad TORSDOF"); torsdof = true;
    } else if (tag(s, "MODEL")) throw parse_error("Unexpected multi-MODEL input");
    else throw parse_error("Unknown or inappropriate tag: " + s);
  }
  if (root.entries.empty()) throw parse_error("No atoms in the ligand");
  if (!torsdof) throw parse_error("Missing TORSDOF");

  Flattener F(L);
  LigNode r; for (int j = 0; j < 3; ++j) r.origin0[j] = root.entries[0].a.xyz[j];
  L.nodes.push_back(r);
  F.node(root, 0);
  const int n = (int)L.atoms.size();
  if (n > VD_MAX_ATOMS) throw parse_error("ligand has " + std::to_string(n) + " atoms; the limit is " + std::to_string(VD_MAX_ATOMS));
  if ((int)L.nodes.size() > VD_MAX_NODES) throw parse_error("too many rotatable bonds");

  // relative mobility: fixed within a node and between a node and its axis atoms, rotor on the
  // axis, variable otherwise (parse_pdbqt.cpp:368-440)
  std::vector<unsigned char> mob((size_t)n * n, 2);
  auto set = [&](int a, int b, unsigned char v) { mob[(size_t)a * n + b] = mob[(size_t)b * n + a] = v; };
  for (auto& nd : L.nodes) for (int i = nd.begin; i < nd.end; ++i) for (int j = i + 1; j < nd.end; ++j) set(i, j, 0);
  for (auto& t : F.fixed_to_node) { int a = std::get<0>(t); if (a < 0) continue; for (int i = std::get<1>(t); i < std::get<2>(t); ++i) if (i != a) set(a, i, 0); }
  for (auto& rt : F.rotors) set(rt.first, rt.second, 1);
  auto movable = [&](int a, int b) { return a == b ? (unsigned char)0 : mob[(size_t)a * n + b]; };

  // covalent bonds with the bead grouping of model.cpp:562-714 (it fixes the order of each
  // atom's bond list, which the 3-bond neighbourhood search below depends on)
  std::vector<TypedAtom> ta(n);
  for (int i = 0; i < n; ++i) { ta[i].ad = L.atoms[i].ad; ta[i].xs = L.atoms[i].xs; for (int k = 0; k < 3; ++k) ta[i].xyz[k] = L.atoms[i].xyz[k]; }
  std::vector<std::pair<const double*, std::vector<int>>> beads;
  for (int i = 0; i < n; ++i) {
    bool placed = false;
    for (auto& b : beads) if (d2(ta[i].xyz, b.first) < 225.0) { b.second.push_back(i); placed = true; break; }
    if (!placed) beads.push_back({ta[i].xyz, {i}});
  }
  const double mc = max_cov();
  for (int i = 0; i < n; ++i) {
    const double ri = ta[i].ad < A_NONE ? kKinds[ta[i].ad].cov : mc;
    std::vector<int> rel;
    for (auto& b : beads) {
      if (d2(b.first, ta[i].xyz) > sq(15.0 + 1.1 * (ri + mc))) continue;
      for (int j : b.second)
        if (movable(i, j) != 2 && i != j && d2(ta[i].xyz, ta[j].xyz) < sq(1.1 * (ri + mc))) rel.push_back(j);
    }
    for (int j : rel) {
      if (j <= i) continue;
      const double r2 = d2(ta[i].xyz, ta[j].xyz);
      if (r2 >= sq(1.1 * (cov_radius(ta[i]) + cov_radius(ta[j])))) continue;
      bool between = false;
      for (int c : rel) {
        if (c == i || c == j) continue;
        if (movable(i, c) != 2 && movable(j, c) != 2 && d2(ta[i].xyz, ta[c].xyz) < r2 && d2(ta[j].xyz, ta[c].xyz) < r2) { between = true; break; }
      }
      if (!between) { L.atoms[i].bonds.push_back(j); L.atoms[j].bonds.push_back(i); }
    }
  }
  assign_xs(ta, [&](size_t i) -> const std::vector<int>& { return L.atoms[i].bonds; });
  for (int i = 0; i < n; ++i) {
    L.atoms[i].xs = ta[i].xs; L.atoms[i].el = ta[i].el;
    if (ta[i].el != E_H) L.n_heavy++;
  }
  // num_tors of conf_independent_inputs (V11/lib/terms.cpp:54-125): half of each heavy atom's
  // rotatable bonds to heavy atoms that have another heavy neighbour
  auto heavy_neighbours = [&](int a) { int k = 0; for (int b : L.atoms[a].bonds) if (L.atoms[b].el != E_H) ++k; return k; };
  L.num_tors = 0;
  for (int i = 0; i < n; ++i) {
    if (L.atoms[i].el == E_H) continue;
    int rot = 0;
    for (int b : L.atoms[i].bonds)
      if (movable(i, b) == 1 && L.atoms[b].el != E_H && heavy_neighbours(b) > 1) ++rot;
    L.num_tors += 0.5 * rot;
  }
  // intramolecular pairs: variable distance, more than 3 bonds apart, both typed
  // (model.cpp:828-896; the neighbourhood is the depth-limited DFS of model::bonded_to)
  std::vector<int> near;
  std::vector<char> seen(n);
  std::function<void(int, int)> dfs = [&](int a, int depth) {
    if (seen[a]) return;
    seen[a] = 1; near.push_back(a);
    if (depth > 0) for (int b : L.atoms[a].bonds) dfs(b, depth - 1);
  };
  for (int i = 0; i < n; ++i) {
    near.clear(); std::fill(seen.begin(), seen.end(), 0);
    dfs(i, 3);
    for (int j = i + 1; j < n; ++j) {
      if (movable(i, j) != 2 || seen[j]) continue;
      const int t1 = L.atoms[i].xs, t2 = L.atoms[j].xs;
      if (t1 >= VD_XS_SIZE || t2 >= VD_XS_SIZE) continue;
      L.pairs.push_back((uint32_t)i | ((uint32_t)j << 8) | ((uint32_t)type_pair_index(t1, t2) << 16));
    }
  }
  if (L.pairs.size() > VD_MAX_PAIRS) throw parse_error("too many intramolecular pairs");
  // subtree ranges (preorder) and tree depth
  const int N = (int)L.nodes.size();
  for (int k = 0; k < N; ++k) L.nodes[k].subtree_end = k + 1;
  for (int k = N - 1; k >= 1; --k) {
    LigNode& p = L.nodes[L.nodes[k].parent];
    p.subtree_end = std::max(p.subtree_end, L.nodes[k].subtree_end);
  }
  for (auto& nd : L.nodes) L.max_depth = std::max(L.max_depth, nd.depth);
  return L;
}

namespace {

// Chunks from an edge colouring: orient every pair along an Euler tour of the pair graph (a
// dummy vertex joined to the odd-degree atoms closes it), so that each atom is the first atom of
// at most ceil(deg/2) pairs and the second atom of at most ceil(deg/2); then colour the
// bipartite graph (first-atom side, second-atom side) with that many colours by alternating-path
// swaps (Konig's edge-colouring theorem).  A colour class is a valid chunk; classes larger than
// G are cut into pieces of G.
std::vector<std::vector<uint32_t>> chunks_by_colouring(const std::vector<uint32_t>& pairs, int A, int G) {
  const int E = (int)pairs.size(), V = A + 1;
  std::vector<int> ea(E), eb(E), deg(V, 0);
  for (int e = 0; e < E; ++e) { ea[e] = pairs[e] & 255; eb[e] = (pairs[e] >> 8) & 255; ++deg[ea[e]]; ++deg[eb[e]]; }
  // adjacency incl. dummy edges (ids >= E)
  std::vector<int> ux(ea), vy(eb);
  for (int v = 0; v < A; ++v) if (deg[v] & 1) { ux.push_back(A); vy.push_back(v); }
  const int E2 = (int)ux.size();
  std::vector<std::vector<int>> adj(V);
  for (int e = 0; e < E2; ++e) { adj[ux[e]].push_back(e); adj[vy[e]].push_back(e); }
  std::vector<char> used(E2, 0);
  std::vector<size_t> next(V, 0);
  std::vector<int> from(E, -1), to(E, -1);
  for (int start = 0; start < V; ++start) {
    // Hierholzer: walk unused edges; every edge is oriented in the direction it is walked
    std::vector<int> stack{start};
    while (!stack.empty()) {
      const int v = stack.back();
      while (next[v] < adj[v].size() && used[adj[v][next[v]]]) ++next[v];
      if (next[v] == adj[v].size()) { stack.pop_back(); continue; }
      const int e = adj[v][next[v]];
      used[e] = 1;
      const int w = ux[e] == v ? vy[e] : ux[e];
      if (e < E) { from[e] = v; to[e] = w; }
      stack.push_back(w);
    }
  }
  std::vector<int> outd(A, 0), ind(A, 0);
  for (int e = 0; e < E; ++e) { ++outd[from[e]]; ++ind[to[e]]; }
  int D = 1;
  for (int v = 0; v < A; ++v) D = std::max(D, std::max(outd[v], ind[v]));
  std::vector<int> atL((size_t)A * D, -1), atR((size_t)A * D, -1), col(E, -1);
  for (int e = 0; e < E; ++e) {
    const int u = from[e], v = to[e];
    int a = 0; while (atL[(size_t)u * D + a] >= 0) ++a;
    int b = 0; while (atR[(size_t)v * D + b] >= 0) ++b;
    if (atR[(size_t)v * D + a] >= 0) {
      // swap colours a and b on the alternating path that leaves v by its a-coloured edge
      std::vector<int> path;
      int x = v; bool right = true; int c = a;
      while (true) {
        const int f = right ? atR[(size_t)x * D + c] : atL[(size_t)x * D + c];
        if (f < 0) break;
        path.push_back(f);
        x = right ? from[f] : to[f];
        right = !right; c = (c == a) ? b : a;
      }
      for (int f : path) { atL[(size_t)from[f] * D + col[f]] = -1; atR[(size_t)to[f] * D + col[f]] = -1; }
      for (int f : path) {
        col[f] = col[f] == a ? b : a;
        atL[(size_t)from[f] * D + col[f]] = f; atR[(size_t)to[f] * D + col[f]] = f;
      }
    }
    col[e] = a; atL[(size_t)u * D + a] = e; atR[(size_t)v * D + a] = e;
  }
  // Equalise: with K = max(D, ceil(E / G)) colours, move edges from the largest class to the
  // smallest along alternating paths of the two classes that start and end with an edge of the
  // larger one (a bipartite graph has an equitable K-edge-colouring for every K >= D).
  const int K = std::max(D, (E + G - 1) / G);
  {
    std::vector<int> nL((size_t)A * K, -1), nR((size_t)A * K, -1);
    for (int e = 0; e < E; ++e) { nL[(size_t)from[e] * K + col[e]] = e; nR[(size_t)to[e] * K + col[e]] = e; }
    atL.swap(nL); atR.swap(nR);
    std::vector<int> size(K, 0);
    for (int e = 0; e < E; ++e) ++size[col[e]];
    while (true) {
      const int c1 = (int)(std::max_element(size.begin(), size.end()) - size.begin());
      const int c2 = (int)(std::min_element(size.begin(), size.end()) - size.begin());
      if (size[c1] <= G || size[c1] - size[c2] <= 1) break;
      bool moved = false;
      for (int e0 = 0; e0 < E && !moved; ++e0) {
        if (col[e0] != c1) continue;
        // a path end: the left vertex of e0 has no c2 edge (walk from the left side)
        if (atL[(size_t)from[e0] * K + c2] >= 0) continue;
        std::vector<int> path{e0};
        int x = to[e0]; bool right = true; int c = c2;
        while (true) {
          const int f = right ? atR[(size_t)x * K + c] : atL[(size_t)x * K + c];
          if (f < 0) break;
          path.push_back(f);
          x = right ? from[f] : to[f];
          right = !right; c = (c == c1) ? c2 : c1;
        }
        if (col[path.back()] != c1) continue;      // balanced path: swapping gains nothing
        for (int f : path) { atL[(size_t)from[f] * K + col[f]] = -1; atR[(size_t)to[f] * K + col[f]] = -1; }
        for (int f : path) {
          col[f] = col[f] == c1 ? c2 : c1;
          atL[(size_t)from[f] * K + col[f]] = f; atR[(size_t)to[f] * K + col[f]] = f;
        }
        --size[c1]; ++size[c2]; moved = true;
      }
      if (!moved) break;
    }
  }
  std::vector<std::vector<uint32_t>> out;
  for (int c = 0; c < K; ++c) {
    std::vector<uint32_t> cls;
    for (int e = 0; e < E; ++e)
      if (col[e] == c) cls.push_back((uint32_t)from[e] | ((uint32_t)to[e] << 8) | (pairs[e] & 0xFFFF0000u));
    for (size_t k = 0; k < cls.size(); k += G)
      out.emplace_back(cls.begin() + k, cls.begin() + std::min(cls.size(), k + (size_t)G));
  }
  return out;
}

// Greedy first fit over the pairs in list order.
std::vector<std::vector<uint32_t>> chunks_first_fit(const std::vector<uint32_t>& pairs, int G) {
  struct Chunk { uint64_t as[4] = {0, 0, 0, 0}, bs[4] = {0, 0, 0, 0}; int n = 0; };   // one bit per atom (< 256)
  auto has = [](const uint64_t* m, int i) { return (m[i >> 6] >> (i & 63)) & 1u; };
  auto set = [](uint64_t* m, int i) { m[i >> 6] |= 1ull << (i & 63); };
  std::vector<Chunk> chunks;
  std::vector<std::vector<uint32_t>> slots;
  size_t first_open = 0;
  for (uint32_t p : pairs) {
    const int a = p & 255, b = (p >> 8) & 255;
    const uint32_t tp = p >> 16;
    bool placed = false;
    for (size_t c = first_open; c < chunks.size() && !placed; ++c) {
      Chunk& ch = chunks[c];
      if (ch.n >= G) continue;
      int x = -1, y = -1;
      if (!has(ch.as, a) && !has(ch.bs, b)) { x = a; y = b; }
      else if (!has(ch.as, b) && !has(ch.bs, a)) { x = b; y = a; }
      if (x < 0) continue;
      set(ch.as, x); set(ch.bs, y); ++ch.n;
      slots[c].push_back((uint32_t)x | ((uint32_t)y << 8) | (tp << 16));
      placed = true;
    }
    if (!placed) {
      chunks.emplace_back();
      Chunk& ch = chunks.back();
      set(ch.as, a); set(ch.bs, b); ch.n = 1;
      slots.push_back({p});
    }
    while (first_open < chunks.size() && chunks[first_open].n >= G) ++first_open;
  }
  return slots;
}

}  // namespace

const std::vector<uint32_t>& pair_chunks(const Ligand& L, int G) {
  const int which = G == 16 ? 0 : G == 32 ? 1 : 2;
  std::vector<uint32_t>& out = L.chunked[which];
  if (!out.empty() || L.pairs.empty()) return out;
  std::vector<std::vector<uint32_t>> a = chunks_by_colouring(L.pairs, (int)L.atoms.size(), G);
  std::vector<std::vector<uint32_t>> b = chunks_first_fit(L.pairs, G);
  const std::vector<std::vector<uint32_t>>& best = a.size() <= b.size() ? a : b;
  // when every atom has a lane of its own, a pair sits in the slot of its first atom: the lane
  // of atom i then holds every pair whose first atom is i and keeps that side in registers
  const bool rows = (int)L.atoms.size() <= G;
  out.assign(best.size() * (size_t)G, VD_NO_PAIR);
  for (size_t c = 0; c < best.size(); ++c)
    for (size_t k = 0; k < best[c].size(); ++k) out[c * G + (rows ? (best[c][k] & 255) : k)] = best[c][k];
  return out;
}

vd_lig_header pack_ligand(const Ligand& L, std::vector<float>& blob, int G) {
  auto pad16 = [&]() { while (blob.size() % 4) blob.push_back(0.f); };
  auto as_f = [](uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; };
  vd_lig_header h;
  std::memset(&h, 0, sizeof(h));
  h.n_atoms = (int)L.atoms.size(); h.n_nodes = (int)L.nodes.size(); h.n_pairs = (int)L.pairs.size(); h.n_heavy = L.n_heavy;
  h.n_torsions = L.n_torsions();
  int rounds = 0; while ((1 << rounds) < L.max_depth + 1) ++rounds;
  h.jump_rounds = rounds;
  for (int k = 0; k < 3; ++k) h.origin0[k] = (float)L.nodes[0].origin0[k];
  pad16();
  h.off_atoms = (int)(blob.size() / 4);
  uint32_t slot = 0;
  for (const LigAtom& a : L.atoms) {
    const uint32_t heavy = a.el != E_H ? 1u : 0u;
    uint32_t meta = (uint32_t)a.node | ((uint32_t)a.xs << 8) | (heavy << 16) | ((heavy ? slot : 0u) << 17);
    slot += heavy;
    blob.push_back((float)a.local[0]); blob.push_back((float)a.local[1]); blob.push_back((float)a.local[2]); blob.push_back(as_f(meta));
  }
  h.off_nodes = (int)(blob.size() / 4);
  for (const LigNode& nd : L.nodes) {
    blob.push_back((float)nd.rel_origin[0]); blob.push_back((float)nd.rel_origin[1]); blob.push_back((float)nd.rel_origin[2]);
    blob.push_back(as_f((uint32_t)(nd.parent + 1) | ((uint32_t)nd.begin << 8) | ((uint32_t)nd.end << 16)));
  }
  for (const LigNode& nd : L.nodes) {
    blob.push_back((float)nd.rel_axis[0]); blob.push_back((float)nd.rel_axis[1]); blob.push_back((float)nd.rel_axis[2]);
    blob.push_back(as_f((uint32_t)nd.subtree_end));
  }
  // device pair order: conflict-free chunks of G slots (only the order of the energy sum
  // changes), grouped by four chunks with each lane's four words adjacent ([group][lane][4], one
  // 16-byte read per lane and group); the chunk count is padded to a multiple of four
  const std::vector<uint32_t>& pairs = pair_chunks(L, G);
  h.chunk_lanes = G;
  h.n_chunks = (int)(pairs.size() / G);
  h.pair_rows = (int)L.atoms.size() <= G ? 1 : 0;
  // rows: slot l holds pair (l, b); its word carries, in place of the first atom (= l), the slot
  // whose pair has second atom l in the same chunk (G: none), from which lane l pulls that
  // pair's force on atom l.  A slot without a pair names a written coordinate row as partner
  // (atom l, or atom 0 for lanes beyond the ligand) and the all-zero row of the pair table
  // (VD_NUM_TYPE_PAIRS) as type pair, so the device reads every word without decoding it.
  const int A = (int)L.atoms.size();
  auto empty_row = [&](int l, uint32_t src) {
    return src | ((uint32_t)(l < A ? l : 0) << 8) | ((uint32_t)VD_NUM_TYPE_PAIRS << 16);
  };
  std::vector<uint32_t> words(pairs);
  if (h.pair_rows) {
    for (int ch = 0; ch < h.n_chunks; ++ch) {
      std::vector<uint32_t> inv(G, (uint32_t)G);
      for (int l = 0; l < G; ++l) {
        const uint32_t w = pairs[(size_t)ch * G + l];
        if (w != VD_NO_PAIR) inv[(w >> 8) & 255] = (uint32_t)l;
      }
      for (int l = 0; l < G; ++l) {
        const uint32_t w = pairs[(size_t)ch * G + l];
        words[(size_t)ch * G + l] = w == VD_NO_PAIR ? empty_row(l, inv[l]) : (inv[l] | (w & 0xFFFFFF00u));
      }
    }
  }
  const int groups = (h.n_chunks + 3) / 4;
  h.off_pairs = (int)blob.size();
  for (int cg = 0; cg < groups; ++cg)
    for (int l = 0; l < G; ++l)
      for (int u = 0; u < 4; ++u) {
        const int ch = cg * 4 + u;
        const uint32_t pad = h.pair_rows ? empty_row(l, (uint32_t)G) : VD_NO_PAIR;
        blob.push_back(as_f(ch < h.n_chunks ? words[(size_t)ch * G + l] : pad));
      }
  pad16();
  return h;
}

// ---------------------------------------------------------------------------------------
// Synthetic inputs.
void synth_receptor(int n_atoms, double cube, double cavity_radius, uint64_t seed,
                    std::vector<float>& xyz, std::vector<int>& xs) {
  // Poisson-disk dart throwing (min separation 1.2 A) in [-cube/2, cube/2]^3 with a spherical
  // cavity of the given radius emptied at the centre.  XS mix: C_H 35%, C_P 28%, N_P 8%, N_D 6%,
  // N_A 3%, O_A 12%, O_DA 4%, O_P 2%, S_P 2%.
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-cube / 2, cube / 2), U01(0, 1);
  const double sep2 = 1.2 * 1.2, cell = 1.2;
  const int nc = (int)std::ceil(cube / cell) + 1;
  std::vector<std::vector<int>> grid((size_t)nc * nc * nc);
  auto cidx = [&](double v) { int c = (int)std::floor((v + cube / 2) / cell); return std::min(std::max(c, 0), nc - 1); };
  const int types[] = {X_C_H, X_C_P, X_N_P, X_N_D, X_N_A, X_O_A, X_O_DA, X_O_P, X_S_P};
  const double cum[] = {0.35, 0.63, 0.71, 0.77, 0.80, 0.92, 0.96, 0.98, 1.00};
  xyz.clear(); xs.clear();
  long long attempts = 0;
  while ((int)xs.size() < n_atoms && attempts < 200LL * n_atoms) {
    ++attempts;
    double p[3] = {U(rng), U(rng), U(rng)};
    if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] < cavity_radius * cavity_radius) continue;
    int c[3] = {cidx(p[0]), cidx(p[1]), cidx(p[2])};
    bool ok = true;
    for (int dx = -1; dx <= 1 && ok; ++dx) for (int dy = -1; dy <= 1 && ok; ++dy) for (int dz = -1; dz <= 1 && ok; ++dz) {
      int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
      if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) continue;
      for (int j : grid[((size_t)x * nc + y) * nc + z]) {
        double q[3] = {xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]};
        if (d2(p, q) < sep2) { ok = false; break; }
      }
    }
    if (!ok) continue;
    int id = (int)xs.size();
    grid[((size_t)c[0] * nc + c[1]) * nc + c[2]].push_back(id);
    for (int k = 0; k < 3; ++k) xyz.push_back((float)p[k]);
    double u = U01(rng); int t = 0; while (t < 8 && u >= cum[t]) ++t;
    xs.push_back(types[t]);
  }
}

std::string synth_ligand_pdbqt(int n_heavy, int n_rot, uint64_t seed) {
  // Random acyclic tree: 1.5 A bonds at ~109.5 deg, random dihedrals, clash rejection < 2.0 A.
  // Elements: C 70%, N 10% (acceptor NA or donor N-HD), O 15% (OA, donor when an HD is attached),
  // F/Cl 5% on terminal atoms.  n_rot rotatable bonds drawn uniformly among non-terminal bonds.
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U01(0, 1);
  std::normal_distribution<double> N01(0, 1);
  struct A { double p[3]; int parent; std::vector<int> nb; std::string ad; };
  std::vector<A> at;
  auto place = [&](int i, double len, double out[3]) -> bool {
    double dir[3];
    if (at[i].nb.empty()) {
      double n = 0; for (int k = 0; k < 3; ++k) { dir[k] = N01(rng); n += dir[k] * dir[k]; }
      n = std::sqrt(n); for (int k = 0; k < 3; ++k) dir[k] /= n;
    } else {
      const A& j = at[at[i].nb[std::uniform_int_distribution<int>(0, (int)at[i].nb.size() - 1)(rng)]];
      double b[3], n = 0; for (int k = 0; k < 3; ++k) { b[k] = at[i].p[k] - j.p[k]; n += b[k] * b[k]; }
      n = std::sqrt(n); for (int k = 0; k < 3; ++k) b[k] /= n;
      double r[3] = {N01(rng), N01(rng), N01(rng)}, d = r[0] * b[0] + r[1] * b[1] + r[2] * b[2];
      double pp[3], pn = 0; for (int k = 0; k < 3; ++k) { pp[k] = r[k] - d * b[k]; pn += pp[k] * pp[k]; }
      pn = std::sqrt(pn); if (pn < 1e-6) return false;
      const double th = 109.5 * M_PI / 180.0;
      for (int k = 0; k < 3; ++k) dir[k] = std::cos(M_PI - th) * b[k] + std::sin(M_PI - th) * pp[k] / pn;
    }
    for (int k = 0; k < 3; ++k) out[k] = at[i].p[k] + len * dir[k];
    for (size_t q = 0; q < at.size(); ++q) if ((int)q != i && d2(out, at[q].p) < 4.0) return false;
    return true;
  };
  at.push_back(A{{0, 0, 0}, -1, {}, "C"});
  int guard = 0;
  while ((int)at.size() < n_heavy && guard < 100000) {
    ++guard;
    int i = std::uniform_int_distribution<int>(0, (int)at.size() - 1)(rng);
    if (at[i].nb.size() >= 3) continue;
    double p[3];
    if (!place(i, 1.5, p)) continue;
    A a; for (int k = 0; k < 3; ++k) a.p[k] = p[k]; a.parent = i; a.ad = "C";
    at.push_back(a);
    int j = (int)at.size() - 1; at[i].nb.push_back(j); at[j].nb.push_back(i);
  }
  const int nh = (int)at.size();
  // elements
  for (int i = 0; i < nh; ++i) {
    double u = U01(rng);
    bool terminal = at[i].nb.size() == 1;
    if (u < 0.70) at[i].ad = "C";
    else if (u < 0.80) at[i].ad = (U01(rng) < 0.5) ? "NA" : "N";
    else if (u < 0.95) at[i].ad = "OA";
    else at[i].ad = terminal ? (U01(rng) < 0.5 ? "F" : "Cl") : "C";
  }
  // polar hydrogens on donors: every plain N, and about a quarter of the OA atoms
  std::vector<std::pair<int, std::array<double, 3>>> hyd;
  for (int i = 0; i < nh; ++i) {
    bool donor = at[i].ad == "N" || (at[i].ad == "OA" && U01(rng) < 0.27);
    if (!donor || at[i].nb.size() >= 3) continue;
    double p[3];
    bool ok = false;
    for (int t = 0; t < 50 && !ok; ++t) ok = place(i, 1.0, p);
    if (!ok) continue;
    hyd.push_back({i, {p[0], p[1], p[2]}});
  }
  // rotatable bonds among non-terminal heavy-heavy bonds
  std::vector<std::pair<int, int>> cand;
  for (int j = 1; j < nh; ++j) { int i = at[j].parent; if (at[i].nb.size() >= 2 && at[j].nb.size() >= 2) cand.push_back({i, j}); }
  std::shuffle(cand.begin(), cand.end(), rng);
  if ((int)cand.size() > n_rot) cand.resize(n_rot);
  std::vector<std::vector<char>> rot(nh, std::vector<char>(nh, 0));
  for (auto& c : cand) rot[c.first][c.second] = rot[c.second][c.first] = 1;
  // fragments (components without rotatable bonds)
  std::vector<int> frag(nh, -1); int nf = 0;
  for (int s = 0; s < nh; ++s) {
    if (frag[s] >= 0) continue;
    std::vector<int> st{s}; frag[s] = nf;
    while (!st.empty()) { int u = st.back(); st.pop_back(); for (int v : at[u].nb) if (!rot[u][v] && frag[v] < 0) { frag[v] = nf; st.push_back(v); } }
    ++nf;
  }
  // fragment graph and the root minimising the tree height
  std::vector<std::vector<std::pair<int, std::pair<int, int>>>> fg(nf);
  for (auto& c : cand) { fg[frag[c.first]].push_back({frag[c.second], {c.first, c.second}}); fg[frag[c.second]].push_back({frag[c.first], {c.second, c.first}}); }
  int best_root = 0, best_h = 1 << 30;
  for (int r = 0; r < nf; ++r) {
    std::vector<int> dist(nf, -1); std::vector<int> q{r}; dist[r] = 0; int h = 0;
    for (size_t qi = 0; qi < q.size(); ++qi) for (auto& e : fg[q[qi]]) if (dist[e.first] < 0) { dist[e.first] = dist[q[qi]] + 1; h = std::max(h, dist[e.first]); q.push_back(e.first); }
    if (h < best_h) { best_h = h; best_root = r; }
  }
  // serial numbers in output order
  std::ostringstream out;
  out << "REMARK  synthetic ligand seed " << seed << "\n";
  std::vector<int> serial(nh, 0); int next_serial = 1;
  std::vector<std::vector<int>> hyd_of(nh);
  for (size_t k = 0; k < hyd.size(); ++k) hyd_of[hyd[k].first].push_back((int)k);
  auto atom_line = [&](int serial_no, const std::string& ad, const double* p) {
    char buf[128];
    std::string name = ad == "NA" ? "N" : ad == "OA" ? "O" : ad == "HD" ? "H" : ad;
    std::snprintf(buf, sizeof buf, "ATOM  %5d  %-3s UNL     1    %8.3f%8.3f%8.3f  0.00  0.00    %+6.3f %-2s\n",
                  serial_no, name.c_str(), p[0], p[1], p[2], 0.0, ad.c_str());
    return std::string(buf);
  };
  std::function<void(int, int, int)> emit = [&](int f, int entry_atom, int from_frag) {
    // atoms of fragment f, entry atom first, then the rest in index order
    std::vector<int> members;
    if (entry_atom >= 0) members.push_back(entry_atom);
    for (int i = 0; i < nh; ++i) if (frag[i] == f && i != entry_atom) members.push_back(i);
    for (int i : members) {
      serial[i] = next_serial++;
      out << atom_line(serial[i], at[i].ad, at[i].p);
      for (int k : hyd_of[i]) out << atom_line(next_serial++, "HD", hyd[k].second.data());
    }
    for (int i : members)
      for (auto& e : fg[f])
        if (e.second.first == i && e.first != from_frag) {
          int to = e.second.second;
          int to_serial = next_serial;  // assigned when the child emits its entry atom first
          out << "BRANCH " << std::setw(3) << serial[i] << " " << std::setw(3) << to_serial << "\n";
          emit(e.first, to, f);
          out << "ENDBRANCH " << std::setw(3) << serial[i] << " " << std::setw(3) << serial[to] << "\n";
        }
  };
  out << "ROOT\n";
  {
    // root fragment atoms first, branches after ENDROOT
    std::vector<int> members;
    for (int i = 0; i < nh; ++i) if (frag[i] == best_root) members.push_back(i);
    for (int i : members) {
      serial[i] = next_serial++;
      out << atom_line(serial[i], at[i].ad, at[i].p);
      for (int k : hyd_of[i]) out << atom_line(next_serial++, "HD", hyd[k].second.data());
    }
    out << "ENDROOT\n";
    for (int i : members)
      for (auto& e : fg[best_root])
        if (e.second.first == i) {
          int to = e.second.second;
          out << "BRANCH " << std::setw(3) << serial[i] << " " << std::setw(3) << next_serial << "\n";
          emit(e.first, to, best_root);
          out << "ENDBRANCH " << std::setw(3) << serial[i] << " " << std::setw(3) << serial[to] << "\n";
        }
  }
  out << "TORSDOF " << cand.size() << "\n";
  return out.str();
}

std::vector<int> cluster_poses(int n, int heavy, const float* e, const float* hx, double min_rmsd, int max_size) {
  std::vector<int> kept;
  auto rmsd = [&](int a, int b) {
    double acc = 0;
    for (int i = 0; i < 3 * heavy; ++i) acc += sq((double)hx[(size_t)a * 3 * heavy + i] - hx[(size_t)b * 3 * heavy + i]);
    return heavy > 0 ? std::sqrt(acc / heavy) : 0.0;
  };
  for (int t = 0; t < n; ++t) {
    int closest = -1; double best = 0;
    for (size_t i = 0; i < kept.size(); ++i) { double r = rmsd(t, kept[i]); if (closest < 0 || r < best) { best = r; closest = (int)i; } }
    if (closest >= 0 && best < min_rmsd) { if (e[t] < e[kept[closest]]) kept[closest] = t; }
    else if ((int)kept.size() < max_size) kept.push_back(t);
    else if (!kept.empty() && e[t] < e[kept.back()]) kept.back() = t;
    std::stable_sort(kept.begin(), kept.end(), [&](int a, int b) { return e[a] < e[b]; });
  }
  return kept;
}

}  // namespace vd
