/*
 * vina_oracle.cpp -- FP64 CPU oracle of the Vina-CUDA docking hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  A plain restatement of the reference's
 * algorithms in double precision; the product never calls it.  Citations use
 * V11/ = Vina-CUDA_v1.1/ under the reference root.
 *
 * Deliberate deviations from the reference GPU code (the CPU Vina semantics are kept):
 *  - dot products are exact sums over all n = 6+T entries (the reference tree reduction
 *    V11/inc/cuda/quasi_newton.cuh:1510-1551 drops entries when n is not a power of 2);
 *  - mutation picks one of 2+T entities (V11/lib/mutate.cpp:62-68); mutation_quirk=1
 *    restores the reference GPU's 1+T (V11/lib/main_procedure_cu.cpp:591);
 *  - the gyration radius used by orientation mutations is that of the chain's current
 *    conformation about its root origin (V11/lib/model.cpp:1161-1175);
 *  - random numbers come from Philox4x32-10 keyed by (seed, stream) and counted by
 *    (chain, step, purpose, attempt) instead of host mt19937 tables; the Metropolis
 *    draw is independent of the torsion draw.
 */
#include "oracle.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>
#include <sstream>
#include <algorithm>
#include <thread>
#include <mutex>
#include <map>
#include <atomic>
#include <limits>
#include <stdexcept>

namespace {

typedef double fl;
const fl PI = 3.1415926535897931;                       // V11/lib/common.h:239
const fl EPS = std::numeric_limits<fl>::epsilon();      // V11/lib/common.h:274
const fl MAXFL = std::numeric_limits<fl>::max();        // V11/lib/common.h:271

thread_local std::string g_err;
int g_use_inter = 1, g_use_intra = 1;   // term switches for the finite-difference pins
int fail(const std::string& m) { g_err = m; return -1; }

// Worker threads: the hardware threads, capped by OMP_NUM_THREADS when it is set (the CPU share a
// batch host grants the process).
unsigned host_threads() {
  unsigned n = std::max(1u, std::thread::hardware_concurrency());
  if (const char* s = std::getenv("OMP_NUM_THREADS")) {
    const long v = std::strtol(s, nullptr, 10);
    if (v > 0) n = std::min<unsigned>(n, (unsigned)v);
  }
  return n;
}

struct V3 {
  fl v[3];
  V3() { v[0] = v[1] = v[2] = 0; }
  V3(fl a, fl b, fl c) { v[0] = a; v[1] = b; v[2] = c; }
  fl& operator[](int i) { return v[i]; }
  fl operator[](int i) const { return v[i]; }
};
inline V3 operator+(const V3& a, const V3& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator*(fl s, const V3& a) { return V3(s * a[0], s * a[1], s * a[2]); }
inline fl dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& a, const V3& b) {
  return V3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline fl sqr(fl x) { return x * x; }
inline fl dist2(const V3& a, const V3& b) { return sqr(a[0] - b[0]) + sqr(a[1] - b[1]) + sqr(a[2] - b[2]); }

struct Mat { fl m[9]; };  // column-major like V11/lib/matrix.h: m(i,j) = m[i + 3*j]
inline V3 mul(const Mat& M, const V3& v) {
  return V3(M.m[0] * v[0] + M.m[3] * v[1] + M.m[6] * v[2],
            M.m[1] * v[0] + M.m[4] * v[1] + M.m[7] * v[2],
            M.m[2] * v[0] + M.m[5] * v[1] + M.m[8] * v[2]);
}

struct Q { fl a, b, c, d; };  // (w, x, y, z)
inline Q qmul(const Q& p, const Q& q) {  // Hamilton product, boost::math::quaternion operator*
  return Q{p.a * q.a - p.b * q.b - p.c * q.c - p.d * q.d,
           p.a * q.b + p.b * q.a + p.c * q.d - p.d * q.c,
           p.a * q.c - p.b * q.d + p.c * q.a + p.d * q.b,
           p.a * q.d + p.b * q.c - p.c * q.b + p.d * q.a};
}

// V11/lib/common.h:316-340
void normalize_angle(fl& x) {
  if (x > 3 * PI) { fl n = (x - PI) / (2 * PI); x -= 2 * PI * std::ceil(n); normalize_angle(x); }
  else if (x < -3 * PI) { fl n = (-x - PI) / (2 * PI); x += 2 * PI * std::ceil(n); normalize_angle(x); }
  else if (x > PI) x -= 2 * PI;
  else if (x < -PI) x += 2 * PI;
}
fl normalized_angle(fl x) { normalize_angle(x); return x; }

// V11/lib/quaternion.cpp:65-75
Q angle_to_quaternion(const V3& axis, fl angle) {
  normalize_angle(angle);
  fl c = std::cos(angle / 2), s = std::sin(angle / 2);
  return Q{c, s * axis[0], s * axis[1], s * axis[2]};
}
// V11/lib/quaternion.cpp:87-95
Q angle_to_quaternion(const V3& rot) {
  fl angle = std::sqrt(dot(rot, rot));
  if (angle > EPS) return angle_to_quaternion((1 / angle) * rot, angle);
  return Q{1, 0, 0, 0};
}
// V11/lib/quaternion.h:137-147
void quaternion_normalize_approx(Q& q, fl tol = 1e-6) {
  fl s = q.a * q.a + q.b * q.b + q.c * q.c + q.d * q.d;
  if (std::fabs(s - 1) < tol) return;
  fl a = std::sqrt(s);
  q.a /= a; q.b /= a; q.c /= a; q.d /= a;
}
// V11/lib/quaternion.cpp:155-185
Mat quaternion_to_r3(const Q& q) {
  fl a = q.a, b = q.b, c = q.c, d = q.d;
  fl aa = a * a, ab = a * b, ac = a * c, ad = a * d, bb = b * b, bc = b * c, bd = b * d, cc = c * c, cd = c * d, dd = d * d;
  Mat t;
  t.m[0] = aa + bb - cc - dd;   t.m[3] = 2 * (-ad + bc);     t.m[6] = 2 * (ac + bd);
  t.m[1] = 2 * (ad + bc);       t.m[4] = aa - bb + cc - dd;  t.m[7] = 2 * (-ab + cd);
  t.m[2] = 2 * (-ac + bd);      t.m[5] = 2 * (ab + cd);      t.m[8] = aa - bb - cc + dd;
  return t;
}
// V11/lib/quaternion.cpp:235-240
void quaternion_increment(Q& q, const V3& rot) {
  q = qmul(angle_to_quaternion(rot), q);
  quaternion_normalize_approx(q);
}

// ---------------------------------------------------------------------------------------
// Atom typing constants, V11/lib/atom_constants.h
const int XS_SIZE = 17, AD_SIZE = 20;
enum { EL_H = 0, EL_C, EL_N, EL_O, EL_S, EL_P, EL_F, EL_Cl, EL_Br, EL_I, EL_Met, EL_SIZE };
enum { AD_C = 0, AD_A, AD_N, AD_O, AD_P, AD_S, AD_H, AD_F, AD_I, AD_NA, AD_OA, AD_SA, AD_HD,
       AD_Mg, AD_Mn, AD_Zn, AD_Ca, AD_Fe, AD_Cl, AD_Br };
enum { XS_C_H = 0, XS_C_P, XS_N_P, XS_N_D, XS_N_A, XS_N_DA, XS_O_P, XS_O_D, XS_O_A, XS_O_DA,
       XS_S_P, XS_P_P, XS_F_H, XS_Cl_H, XS_Br_H, XS_I_H, XS_Met_D };
const char* AD_NAMES[AD_SIZE] = {"C", "A", "N", "O", "P", "S", "H", "F", "I", "NA", "OA", "SA", "HD",
                                 "Mg", "Mn", "Zn", "Ca", "Fe", "Cl", "Br"};            // :117-138
const fl AD_COV[AD_SIZE] = {0.77, 0.77, 0.75, 0.73, 1.06, 1.02, 0.37, 0.71, 1.33, 0.75, 0.73, 1.02,
                            0.37, 1.30, 1.39, 1.31, 1.74, 1.25, 0.99, 1.14};
const fl METAL_COV = 1.75;                                                               // :141
const int AD_TO_EL[AD_SIZE] = {EL_C, EL_C, EL_N, EL_O, EL_P, EL_S, EL_H, EL_F, EL_I, EL_N, EL_O, EL_S,
                               EL_H, EL_Met, EL_Met, EL_Met, EL_Met, EL_Met, EL_Cl, EL_Br};  // :177-205
const fl XS_RADIUS[XS_SIZE] = {1.9, 1.9, 1.8, 1.8, 1.8, 1.8, 1.7, 1.7, 1.7, 1.7, 2.0, 2.1, 1.5, 1.8, 2.0, 2.2, 1.2};  // :209-227
bool xs_hydrophobic(int t) { return t == XS_C_H || t == XS_F_H || t == XS_Cl_H || t == XS_Br_H || t == XS_I_H; }  // :249
bool xs_acceptor(int t) { return t == XS_N_A || t == XS_N_DA || t == XS_O_A || t == XS_O_DA; }                     // :257
bool xs_donor(int t) { return t == XS_N_D || t == XS_N_DA || t == XS_O_D || t == XS_O_DA || t == XS_Met_D; }        // :264
bool xs_hbond(int a, int b) { return (xs_donor(a) && xs_acceptor(b)) || (xs_donor(b) && xs_acceptor(a)); }         // :276
fl max_covalent_radius() { fl m = 0; for (int i = 0; i < AD_SIZE; ++i) m = std::max(m, AD_COV[i]); return m; }
int string_to_ad(const std::string& s) {
  for (int i = 0; i < AD_SIZE; ++i) if (s == AD_NAMES[i]) return i;
  if (s == "Se") return AD_S;  // atom_equivalence_data
  return AD_SIZE;
}
bool non_ad_metal(const std::string& s) {
  static const char* m[] = {"Cu", "Fe", "Na", "K", "Hg", "Co", "U", "Cd", "Ni"};
  for (auto x : m) if (s == x) return true;
  return false;
}
inline int tri_index(int i, int j) { if (i > j) std::swap(i, j); return i + j * (j + 1) / 2; }  // triangular_matrix_index.h

// ---------------------------------------------------------------------------------------
// Scoring terms, V11/lib/everything.cpp:29-31, 97-216, 381-479
fl slope_step(fl x_bad, fl x_good, fl x) {  // :128-139
  if (x_bad < x_good) { if (x <= x_bad) return 0; if (x >= x_good) return 1; }
  else { if (x >= x_bad) return 0; if (x <= x_good) return 1; }
  return (x - x_bad) / (x_good - x_bad);
}
fl term_eval(int t1, int t2, fl r, const fl* w, fl* parts) {
  fl d = r - (XS_RADIUS[t1] + XS_RADIUS[t2]);
  fl p[5];
  p[0] = std::exp(-sqr((d - 0) / 0.5));                                   // gauss(0, 0.5)
  p[1] = std::exp(-sqr((d - 3) / 2.0));                                   // gauss(3, 2.0)
  p[2] = (d > 0) ? 0 : d * d;                                             // repulsion(0)
  p[3] = (xs_hydrophobic(t1) && xs_hydrophobic(t2)) ? slope_step(1.5, 0.5, d) : 0;  // hydrophobic(0.5, 1.5)
  p[4] = xs_hbond(t1, t2) ? slope_step(0, -0.7, d) : 0;                    // non_dir_h_bond(-0.7, 0)
  fl acc = 0;
  for (int k = 0; k < 5; ++k) { acc += w[k] * p[k]; if (parts) parts[k] = p[k]; }  // weighted_terms.cpp:64-69
  return acc;
}

// Look-up table, V11/lib/precalculate.h:121-140, 255-283, 334-339
struct Lut {
  int n = 0; fl factor = 32, cutoff_sqr = 64;
  std::vector<fl> se, sd, fast;  // [153][n]
  void build(const fl* w) {
    n = int(factor * cutoff_sqr) + 3;
    std::vector<fl> rs(n);
    for (int i = 0; i < n; ++i) rs[i] = std::sqrt(i / factor);
    const int np = XS_SIZE * (XS_SIZE + 1) / 2;
    se.assign(np * n, 0); sd.assign(np * n, 0); fast.assign(np * n, 0);
    for (int t1 = 0; t1 < XS_SIZE; ++t1)
      for (int t2 = t1; t2 < XS_SIZE; ++t2) {
        int tp = tri_index(t1, t2);
        fl* s = &se[tp * n]; fl* d = &sd[tp * n]; fl* f = &fast[tp * n];
        for (int i = 0; i < n; ++i) s[i] = std::min(MAXFL, term_eval(t1, t2, rs[i], w, nullptr));
        for (int i = 0; i < n; ++i) {
          if (i == 0 || i == n - 1) d[i] = 0;
          else d[i] = (s[i + 1] - s[i - 1]) / ((rs[i + 1] - rs[i - 1]) * rs[i]);
          fl f2 = (i + 1 >= n) ? 0 : s[i + 1];
          f[i] = (f2 + s[i]) / 2;
        }
      }
  }
  fl eval_fast(int tp, fl r2) const { return fast[tp * n + int(factor * r2)]; }  // :55-60
  void eval_deriv(int tp, fl r2, fl& e, fl& dor) const {                       // :80-95
    fl rf = factor * r2; int i1 = int(rf); fl rem = rf - i1;
    const fl* s = &se[tp * n]; const fl* d = &sd[tp * n];
    e = s[i1] + rem * (s[i1 + 1] - s[i1]);
    dor = d[i1] + rem * (d[i1 + 1] - d[i1]);
  }
};

// curl, V11/lib/curl.h:31-45
inline void curl(fl& e, V3& deriv, fl v) {
  if (e > 0 && v < 0.1 * MAXFL) { fl t = (v < EPS) ? 0 : v / (v + e); e *= t; deriv = sqr(t) * deriv; }
}

// ---------------------------------------------------------------------------------------
// Covalent bonding + XS typing shared by receptor and ligand models
// (V11/lib/model.cpp:562-714 assign_bonds with its bead acceleration structure, 716-797)
struct AtomT {
  int ad = AD_SIZE, el = EL_SIZE, xs = XS_SIZE;
  V3 coords;           // absolute input coordinates
  V3 local;            // coordinates relative to the owning node's origin
  std::vector<int> bonds;
  unsigned serial = 0;  // PDBQT atom serial number
  bool is_heteroatom() const {
    bool het = ad != AD_A && ad != AD_C && ad != AD_H && ad != AD_HD && ad < AD_SIZE;
    return het || xs == XS_Met_D;
  }
  fl cov() const { return ad < AD_SIZE ? AD_COV[ad] : METAL_COV; }
};

enum { DIST_FIXED = 0, DIST_ROTOR = 1, DIST_VARIABLE = 2 };
struct Mobility {  // symmetric matrix; receptors use all-FIXED
  int n = 0; std::vector<unsigned char> m; bool all_fixed = false;
  int get(int a, int b) const { if (all_fixed) return DIST_FIXED; if (a == b) return DIST_FIXED; return m[a * n + b]; }
  void set(int a, int b, int v) { m[a * n + b] = m[b * n + a] = (unsigned char)v; }
};

void assign_bonds(std::vector<AtomT>& atoms, const Mobility& mob) {
  const fl allow = 1.1, bead_radius = 15;
  const int n = (int)atoms.size();
  // beads (model.cpp:562-585)
  std::vector<std::pair<V3, std::vector<int>>> beads;
  for (int i = 0; i < n; ++i) {
    bool placed = false;
    for (auto& b : beads)
      if (dist2(atoms[i].coords, b.first) < sqr(bead_radius)) { b.second.push_back(i); placed = true; break; }
    if (!placed) beads.push_back({atoms[i].coords, std::vector<int>{i}});
  }
  const fl maxcov = max_covalent_radius();
  for (int i = 0; i < n; ++i) {
    const V3& ci = atoms[i].coords;
    fl ri = atoms[i].ad < AD_SIZE ? AD_COV[atoms[i].ad] : maxcov;
    std::vector<int> relevant;
    fl bead_cut = sqr(bead_radius + allow * (ri + maxcov));
    for (auto& b : beads) {
      if (dist2(b.first, ci) > bead_cut) continue;
      for (int j : b.second) {
        int dt = mob.get(i, j);
        if (dt != DIST_VARIABLE && i != j && dist2(ci, atoms[j].coords) < sqr(allow * (ri + maxcov)))
          relevant.push_back(j);
      }
    }
    for (int j : relevant) {
      if (j <= i) continue;
      fl blen = atoms[i].cov() + atoms[j].cov();
      fl r2 = dist2(ci, atoms[j].coords);
      if (r2 >= sqr(allow * blen)) continue;
      bool between = false;  // atom_exists_between, model.cpp:488-505
      for (int c : relevant) {
        if (c == i || c == j) continue;
        if (mob.get(i, c) != DIST_VARIABLE && mob.get(j, c) != DIST_VARIABLE &&
            dist2(ci, atoms[c].coords) < r2 && dist2(atoms[j].coords, atoms[c].coords) < r2) { between = true; break; }
      }
      if (!between) { atoms[i].bonds.push_back(j); atoms[j].bonds.push_back(i); }
    }
  }
}

void assign_types(std::vector<AtomT>& atoms) {  // model.cpp:771-797
  for (auto& a : atoms) {
    a.el = (a.ad < AD_SIZE) ? AD_TO_EL[a.ad] : EL_SIZE;
    if (a.ad == AD_SIZE && a.xs == XS_Met_D) a.el = EL_Met;
    bool acceptor = (a.ad == AD_OA || a.ad == AD_NA);
    bool hd = false, het = false;
    for (int b : a.bonds) { if (atoms[b].ad == AD_HD) hd = true; if (atoms[b].is_heteroatom()) het = true; }
    bool donor = (a.el == EL_Met || hd);
    switch (a.el) {
      case EL_C: a.xs = het ? XS_C_P : XS_C_H; break;
      case EL_N: a.xs = (acceptor && donor) ? XS_N_DA : (acceptor ? XS_N_A : (donor ? XS_N_D : XS_N_P)); break;
      case EL_O: a.xs = (acceptor && donor) ? XS_O_DA : (acceptor ? XS_O_A : (donor ? XS_O_D : XS_O_P)); break;
      case EL_S: a.xs = XS_S_P; break;
      case EL_P: a.xs = XS_P_P; break;
      case EL_F: a.xs = XS_F_H; break;
      case EL_Cl: a.xs = XS_Cl_H; break;
      case EL_Br: a.xs = XS_Br_H; break;
      case EL_I: a.xs = XS_I_H; break;
      case EL_Met: a.xs = XS_Met_D; break;
      default: break;
    }
  }
}

// PDBQT atom line, V11/lib/parse_pdbqt.cpp:101-118 (fixed columns, 1-based inclusive)
bool parse_atom_line(const std::string& s, AtomT& a, unsigned& number) {
  auto sub = [&](size_t i, size_t j) { if (s.size() < i) return std::string(); return s.substr(i - 1, std::min(j, s.size()) - i + 1); };
  auto trim = [](std::string t) { size_t b = t.find_first_not_of(" \t\r"); if (b == std::string::npos) return std::string(); size_t e = t.find_last_not_of(" \t\r"); return t.substr(b, e - b + 1); };
  if (s.size() < 54) return false;
  number = (unsigned)std::stoul(trim(sub(7, 11)));
  a.coords = V3(std::stod(sub(31, 38)), std::stod(sub(39, 46)), std::stod(sub(47, 54)));
  std::string name = trim(s.size() >= 78 ? s.substr(77) : std::string());  // omit_whitespace(str, 78, 79) reads to end of line
  a.ad = string_to_ad(name);
  if (non_ad_metal(name)) a.xs = XS_Met_D;
  return a.ad < AD_SIZE || a.xs == XS_Met_D;
}
bool starts(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }

// ---------------------------------------------------------------------------------------
// Ligand model: ROOT/BRANCH tree -> rigid nodes in DFS preorder (parse_pdbqt.cpp:286-471,
// 565-627), bonds/types (model.cpp:644-797), intramolecular pairs (model.cpp:828-896).
struct Node { int parent = -1, begin = 0, end = 0; V3 origin0, rel_origin, rel_axis, axis_root; };
struct Pair { int tp, a, b; };
struct Ligand {
  std::vector<AtomT> atoms; std::vector<Node> nodes; std::vector<Pair> pairs;
  std::vector<std::pair<int, int>> rotors;   // atom pairs across rotatable bonds (DISTANCE_ROTOR)
  int torsdof = 0; int heavy = 0;
  int T() const { return (int)nodes.size() - 1; }
};

struct PBranch;  // parsing_struct
struct PAtom { AtomT a; unsigned number; std::vector<PBranch*> ps; };
struct PBranch {
  std::vector<PAtom> atoms; int immobile = -1;      // index into atoms of the BRANCH 'to' atom
  int axis_begin = -1, axis_end = -1;               // ligand atom indices
  ~PBranch() { for (auto& a : atoms) for (auto* p : a.ps) delete p; }
  bool essentially_empty() const {
    for (size_t i = 0; i < atoms.size(); ++i) {
      if (immobile >= 0 && (int)i != immobile) return false;
      if (!atoms[i].ps.empty()) return false;
    }
    return true;
  }
};

struct Parser {
  std::vector<std::string> lines; size_t pos = 0;
  bool next(std::string& s) { if (pos >= lines.size()) return false; s = lines[pos++]; return true; }
};

void parse_branch(Parser& P, PBranch& b, unsigned from, unsigned to);
void parse_branch_aux(Parser& P, const std::string& s, PBranch& p) {  // :332-344
  std::istringstream in(s.substr(6)); unsigned f = 0, t = 0; in >> f >> t;
  for (auto& pa : p.atoms)
    if (pa.number == f) { pa.ps.push_back(new PBranch()); parse_branch(P, *pa.ps.back(), f, t); return; }
  throw std::runtime_error("BRANCH from an atom that is not in this branch");
}
void parse_branch(Parser& P, PBranch& b, unsigned from, unsigned to) {  // :523-563
  std::string s;
  while (P.next(s)) {
    if (s.empty() || starts(s, "WARNING") || starts(s, "REMARK")) continue;
    if (starts(s, "BRANCH")) parse_branch_aux(P, s, b);
    else if (starts(s, "ENDBRANCH")) { if (b.immobile < 0) throw std::runtime_error("branch end atom missing"); return; }
    else if (starts(s, "ATOM  ") || starts(s, "HETATM")) {
      PAtom pa; if (!parse_atom_line(s, pa.a, pa.number)) throw std::runtime_error("bad atom line");
      if (pa.number == to) b.immobile = (int)b.atoms.size();
      b.atoms.push_back(pa);
    } else throw std::runtime_error("unknown tag: " + s);
  }
  throw std::runtime_error("unterminated BRANCH");
}

struct Builder {
  Ligand& L; std::vector<std::vector<int>> mob;  // grows with atoms
  explicit Builder(Ligand& l) : L(l) {}
  int insert(PAtom& pa, const V3& frame_origin) {  // node_t::insert, :183-190
    for (auto* ps : pa.ps) ps->axis_begin = (int)L.atoms.size();
    AtomT a = pa.a; a.local = a.coords - frame_origin; a.serial = pa.number;
    L.atoms.push_back(a);
    return (int)L.atoms.size() - 1;
  }
  void insert_immobiles(PAtom& pa, const V3& frame_origin) {  // :194-214
    for (auto* ps : pa.ps)
      if (!ps->atoms.empty()) { ps->axis_end = (int)L.atoms.size(); insert(ps->atoms[ps->immobile], frame_origin); }
  }
  // postprocess_branch, parse_pdbqt.cpp:406-440.  node index of b already pushed into L.nodes
  void postprocess(PBranch& p, int node_idx, std::vector<std::tuple<int, int, int>>& fixed_ranges,
                   std::vector<std::pair<int, int>>& rotors) {
    const V3 origin = L.nodes[node_idx].origin0;
    L.nodes[node_idx].begin = (int)L.atoms.size();
    for (size_t i = 0; i < p.atoms.size(); ++i) {
      if (!(p.immobile >= 0 && (int)i == p.immobile)) insert(p.atoms[i], origin);
      insert_immobiles(p.atoms[i], origin);
    }
    L.nodes[node_idx].end = (int)L.atoms.size();
    fixed_ranges.push_back(std::make_tuple(p.axis_begin, L.nodes[node_idx].begin, L.nodes[node_idx].end));
    fixed_ranges.push_back(std::make_tuple(p.axis_end, L.nodes[node_idx].begin, L.nodes[node_idx].end));
    if (p.axis_begin >= 0 && p.axis_end >= 0) rotors.push_back({p.axis_begin, p.axis_end});
    for (auto& pa : p.atoms)
      for (auto* ps : pa.ps)
        if (!ps->essentially_empty()) {
          Node c; c.parent = node_idx;
          c.origin0 = ps->atoms[ps->immobile].a.coords;   // segment(origin=immobile atom, axis_root=from atom)
          c.axis_root = pa.a.coords;
          V3 diff = c.origin0 - c.axis_root; fl nrm = std::sqrt(dot(diff, diff));
          if (nrm < EPS) throw std::runtime_error("degenerate rotatable bond");
          c.rel_axis = (1 / nrm) * diff;                   // tree.h:231-245, 269-280 (parent at identity)
          c.rel_origin = c.origin0 - origin;
          L.nodes.push_back(c);
          postprocess(*ps, (int)L.nodes.size() - 1, fixed_ranges, rotors);
        }
  }
};

// model::bonded_to, model.cpp:828-850 (depth-limited DFS with a visited list, in bond order)
void bonded_to(const Ligand& L, int a, int n, std::vector<int>& out) {
  if (std::find(out.begin(), out.end(), a) != out.end()) return;
  out.push_back(a);
  if (n > 0) for (int b : L.atoms[a].bonds) bonded_to(L, b, n - 1, out);
}

Ligand* parse_ligand(const std::string& text) {
  Parser P;
  { std::istringstream in(text); std::string s; while (std::getline(in, s)) { if (!s.empty() && s.back() == '\r') s.pop_back(); P.lines.push_back(s); } }
  PBranch root; std::string s; bool have_torsdof = false; int torsdof = 0;
  // parse_pdbqt_root + parse_pdbqt_aux, :286-366
  while (P.next(s)) {
    if (s.empty() || starts(s, "WARNING") || starts(s, "REMARK")) continue;
    if (starts(s, "ROOT")) break;
    throw std::runtime_error("expected ROOT, got: " + s);
  }
  while (P.next(s)) {
    if (s.empty() || starts(s, "WARNING") || starts(s, "REMARK")) continue;
    if (starts(s, "ENDROOT")) break;
    if (starts(s, "ATOM  ") || starts(s, "HETATM")) {
      PAtom pa; if (!parse_atom_line(s, pa.a, pa.number)) throw std::runtime_error("bad atom line");
      root.atoms.push_back(pa);
    } else throw std::runtime_error("unknown tag in ROOT: " + s);
  }
  while (P.next(s)) {
    if (s.empty() || starts(s, "WARNING") || starts(s, "REMARK")) continue;
    if (starts(s, "BRANCH")) parse_branch_aux(P, s, root);
    else if (starts(s, "TORSDOF")) { std::istringstream in(s.substr(7)); in >> torsdof; have_torsdof = true; }
    else throw std::runtime_error("unknown tag: " + s);
  }
  if (root.atoms.empty()) throw std::runtime_error("no atoms in the ligand");
  if (!have_torsdof) throw std::runtime_error("missing TORSDOF");
  Ligand* L = new Ligand(); L->torsdof = torsdof;
  Builder B(*L);
  Node r; r.origin0 = root.atoms[0].a.coords; L->nodes.push_back(r);  // postprocess_ligand, :442-447
  std::vector<std::tuple<int, int, int>> fixed_ranges; std::vector<std::pair<int, int>> rotors;
  B.postprocess(root, 0, fixed_ranges, rotors);
  const int n = (int)L->atoms.size();
  // mobility matrix (nr_update_matrixes defaults VARIABLE; add_bonds FIXED; set_rotor ROTOR)
  Mobility mob; mob.n = n; mob.m.assign(n * n, DIST_VARIABLE);
  for (auto& nd : L->nodes) for (int i = nd.begin; i < nd.end; ++i) for (int j = i + 1; j < nd.end; ++j) mob.set(i, j, DIST_FIXED);
  for (auto& fr : fixed_ranges) { int a = std::get<0>(fr); if (a < 0) continue; for (int i = std::get<1>(fr); i < std::get<2>(fr); ++i) mob.set(a, i, DIST_FIXED); }
  for (auto& rt : rotors) mob.set(rt.first, rt.second, DIST_ROTOR);
  L->rotors = rotors;
  // add_bonds for axis_end is applied before set_rotor, so the rotor entry wins (order as in :430-432)
  assign_bonds(L->atoms, mob);
  assign_types(L->atoms);
  for (int i = 0; i < n; ++i) {  // initialize_pairs, model.cpp:875-896
    std::vector<int> b3; bonded_to(*L, i, 3, b3);
    for (int j = i + 1; j < n; ++j) {
      if (mob.get(i, j) != DIST_VARIABLE) continue;
      if (std::find(b3.begin(), b3.end(), j) != b3.end()) continue;
      int t1 = L->atoms[i].xs, t2 = L->atoms[j].xs;
      if (t1 < XS_SIZE && t2 < XS_SIZE) L->pairs.push_back(Pair{tri_index(t1, t2), i, j});
    }
  }
  for (int i = 0; i < n; ++i) {
    if (L->atoms[i].el != EL_H) L->heavy++;
    // local coords are relative to the owning node's origin (parse_pdbqt.cpp:184-189)
  }
  return L;
}

// ---------------------------------------------------------------------------------------
// Grid maps: grid_dims / grid::init (V11/main/main.cpp:864-876; V11/lib/grid.cpp:29-43),
// szv_grid cell lists (szv_grid.cpp:28-72, 92-131; brick.h:29-49), populate (cache.cpp:148-208)
struct Grid {
  const Lut* lut = nullptr;
  fl begin[3], end[3]; int n[3];
  fl init[3], range[3], factor[3], factor_inv[3], dim_fl_minus_1[3]; int dim[3];
  std::vector<V3> rxyz; std::vector<int> rxs;
  std::vector<fl> maps;  // [17][dim2][dim1][dim0]
  std::vector<fl> slack; // per map value: bound on the change if an r^2 within slack_eps bins of a bin edge rounds across it
  fl slack_eps = 0;
  // cell lists
  int cdim[3]; std::vector<std::vector<int>> cells;
  void setup(const fl* b, const fl* e, const int* nn) {
    for (int i = 0; i < 3; ++i) {
      begin[i] = b[i]; end[i] = e[i]; n[i] = nn[i];
      dim[i] = n[i] + 1; init[i] = begin[i]; range[i] = end[i] - begin[i];
      dim_fl_minus_1[i] = dim[i] - 1.0; factor[i] = dim_fl_minus_1[i] / range[i]; factor_inv[i] = 1 / factor[i];
    }
    maps.assign((size_t)XS_SIZE * dim[0] * dim[1] * dim[2], 0);
    slack.assign(maps.size(), 0);
  }
  size_t idx(int x, int y, int z) const { return (size_t)x + dim[0] * ((size_t)y + dim[1] * (size_t)z); }
  static fl brick_d2(const V3& lo, const V3& hi, const V3& p) {
    V3 c; for (int i = 0; i < 3; ++i) c[i] = p[i] <= lo[i] ? lo[i] : (p[i] >= hi[i] ? hi[i] : p[i]);
    return dist2(c, p);
  }
  V3 cell_coord(int i, int j, int k) const {
    int id[3] = {i, j, k}; V3 t;
    for (int a = 0; a < 3; ++a) t[a] = init[a] + range[a] * id[a] / cdim[a];
    return t;
  }
  void build_cells() {
    for (int i = 0; i < 3; ++i) { fl nf = (end[i] - begin[i]) / 3; int ni = int(nf); cdim[i] = ni < 1 ? 1 : ni; }
    const fl cut2 = lut->cutoff_sqr;
    std::vector<int> relevant;
    V3 lo(init[0], init[1], init[2]), hi(end[0], end[1], end[2]);
    for (size_t i = 0; i < rxyz.size(); ++i) if (rxs[i] < XS_SIZE && brick_d2(lo, hi, rxyz[i]) < cut2) relevant.push_back((int)i);
    cells.assign((size_t)cdim[0] * cdim[1] * cdim[2], {});
    for (int x = 0; x < cdim[0]; ++x) for (int y = 0; y < cdim[1]; ++y) for (int z = 0; z < cdim[2]; ++z) {
      V3 a = cell_coord(x, y, z), b = cell_coord(x + 1, y + 1, z + 1);
      auto& c = cells[x + cdim[0] * (y + cdim[1] * z)];
      for (int i : relevant) if (brick_d2(a, b, rxyz[i]) < cut2) c.push_back(i);
    }
  }
  static int fl_to_sz(fl x, int mx) { if (x <= 0) return 0; if (x >= mx) return mx; return int(x); }
  const std::vector<int>& possibilities(const V3& p) const {
    int id[3];
    for (int i = 0; i < 3; ++i) id[i] = fl_to_sz((p[i] - init[i]) * cdim[i] / range[i], cdim[i] - 1);
    return cells[id[0] + cdim[0] * (id[1] + cdim[1] * id[2])];
  }
  void populate(bool use_cells) {
    if (use_cells) build_cells();
    std::vector<int> all;
    for (size_t i = 0; i < rxyz.size(); ++i) all.push_back((int)i);
    const size_t plane = (size_t)dim[0] * dim[1] * dim[2];
    unsigned nth = host_threads();
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; ++t) pool.emplace_back([&, t]() {
      fl aff[XS_SIZE], slk[XS_SIZE];
      const int nb = lut->n;
      for (int z = (int)t; z < dim[2]; z += (int)nth)
        for (int y = 0; y < dim[1]; ++y)
          for (int x = 0; x < dim[0]; ++x) {
            for (int k = 0; k < XS_SIZE; ++k) { aff[k] = 0; slk[k] = 0; }
            V3 p(init[0] + factor_inv[0] * x, init[1] + factor_inv[1] * y, init[2] + factor_inv[2] * z);  // grid.h:49-53
            const std::vector<int>& cand = use_cells ? possibilities(p) : all;
            for (int i : cand) {
              int t1 = rxs[i]; if (t1 >= XS_SIZE) continue;
              fl r2 = dist2(rxyz[i], p);
              if (r2 <= lut->cutoff_sqr)
                for (int t2 = 0; t2 < XS_SIZE; ++t2) aff[t2] += lut->eval_fast(tri_index(t1, t2), r2);
              if (slack_eps > 0) {
                const fl b = lut->factor * r2; const int ib = int(b); const fl fr = b - ib;
                const bool in = r2 <= lut->cutoff_sqr;
                for (int t2 = 0; t2 < XS_SIZE; ++t2) {
                  const fl* f = &lut->fast[tri_index(t1, t2) * nb];
                  if (in && fr < slack_eps && ib > 0) slk[t2] += std::fabs(f[ib] - f[ib - 1]);
                  if (in && fr > 1 - slack_eps && ib + 1 < nb) slk[t2] += std::fabs(f[ib + 1] - f[ib]);
                  if (std::fabs(r2 - lut->cutoff_sqr) < slack_eps / lut->factor) slk[t2] += std::fabs(f[std::min(ib, nb - 1)]);
                }
              }
            }
            for (int k = 0; k < XS_SIZE; ++k) { maps[k * plane + idx(x, y, z)] = aff[k]; slack[k * plane + idx(x, y, z)] = slk[k]; }
          }
    });
    for (auto& th : pool) th.join();
  }
  // grid::evaluate_aux, grid.cpp:50-167
  fl evaluate(int t, const V3& loc, fl slope, fl v, V3* deriv) const {
    const size_t plane = (size_t)dim[0] * dim[1] * dim[2];
    const fl* md = &maps[t * plane];
    fl s[3], miss[3] = {0, 0, 0}; int region[3], a[3];
    for (int i = 0; i < 3; ++i) {
      s[i] = (loc[i] - init[i]) * factor[i];
      if (s[i] < 0) { miss[i] = -s[i]; region[i] = -1; a[i] = 0; s[i] = 0; }
      else if (s[i] >= dim_fl_minus_1[i]) { miss[i] = s[i] - dim_fl_minus_1[i]; region[i] = 1; a[i] = dim[i] - 2; s[i] = 1; }
      else { region[i] = 0; a[i] = int(s[i]); s[i] -= a[i]; }
    }
    const fl penalty = slope * (miss[0] * factor_inv[0] + miss[1] * factor_inv[1] + miss[2] * factor_inv[2]);
    int x0 = a[0], y0 = a[1], z0 = a[2], x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    fl f000 = md[idx(x0, y0, z0)], f100 = md[idx(x1, y0, z0)], f010 = md[idx(x0, y1, z0)], f110 = md[idx(x1, y1, z0)];
    fl f001 = md[idx(x0, y0, z1)], f101 = md[idx(x1, y0, z1)], f011 = md[idx(x0, y1, z1)], f111 = md[idx(x1, y1, z1)];
    fl x = s[0], y = s[1], z = s[2], mx = 1 - x, my = 1 - y, mz = 1 - z;
    fl f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
           f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
    if (deriv) {
      fl xg = -f000 * my * mz + f100 * my * mz - f010 * y * mz + f110 * y * mz - f001 * my * z + f101 * my * z - f011 * y * z + f111 * y * z;
      fl yg = -f000 * mx * mz - f100 * x * mz + f010 * mx * mz + f110 * x * mz - f001 * mx * z - f101 * x * z + f011 * mx * z + f111 * x * z;
      fl zg = -f000 * mx * my - f100 * x * my - f010 * mx * y - f110 * x * y + f001 * mx * my + f101 * x * my + f011 * mx * y + f111 * x * y;
      V3 gr(xg, yg, zg);
      curl(f, gr, v);
      for (int i = 0; i < 3; ++i) (*deriv)[i] = factor[i] * (region[i] == 0 ? gr[i] : 0) + slope * region[i];
      return f + penalty;
    }
    V3 dummy; curl(f, dummy, v);
    return f + penalty;
  }
};

// ---------------------------------------------------------------------------------------
// Kinematics (tree.h:205-209, 282-292; conf.h) and energy/gradient (model.cpp:1241-1259,
// 1324-1334; cache.cpp:73-88; tree.h:169-178, 341-348, 373-380, 416-422)
struct Conf { V3 pos; Q q; std::vector<fl> tors; fl e = 0; };
Conf conf_from(const Ligand& L, const fl* c) {
  Conf x; x.pos = V3(c[0], c[1], c[2]); x.q = Q{c[3], c[4], c[5], c[6]};
  x.tors.assign(c + 7, c + 7 + L.T()); return x;
}
void conf_to(const Conf& x, fl* c) {
  for (int i = 0; i < 3; ++i) c[i] = x.pos[i];
  c[3] = x.q.a; c[4] = x.q.b; c[5] = x.q.c; c[6] = x.q.d;
  for (size_t i = 0; i < x.tors.size(); ++i) c[7 + i] = x.tors[i];
}
struct Frames { std::vector<V3> origin, axis; std::vector<Q> q; std::vector<Mat> R; };

void set_conf(const Ligand& L, const Conf& c, std::vector<V3>& coords, Frames& F) {
  const int N = (int)L.nodes.size();
  F.origin.resize(N); F.axis.resize(N); F.q.resize(N); F.R.resize(N);
  coords.resize(L.atoms.size());
  F.origin[0] = c.pos; F.q[0] = c.q; F.R[0] = quaternion_to_r3(c.q); F.axis[0] = V3();
  for (int i = L.nodes[0].begin; i < L.nodes[0].end; ++i) coords[i] = F.origin[0] + mul(F.R[0], L.atoms[i].local);
  for (int k = 1; k < N; ++k) {  // preorder: parent done first
    const Node& nd = L.nodes[k]; int p = nd.parent;
    F.origin[k] = F.origin[p] + mul(F.R[p], nd.rel_origin);
    F.axis[k] = mul(F.R[p], nd.rel_axis);
    Q q = qmul(angle_to_quaternion(F.axis[k], c.tors[k - 1]), F.q[p]);
    quaternion_normalize_approx(q);
    F.q[k] = q; F.R[k] = quaternion_to_r3(q);
    for (int i = nd.begin; i < nd.end; ++i) coords[i] = F.origin[k] + mul(F.R[k], L.atoms[i].local);
  }
}

fl noncache_terms(const Grid& g, const Ligand& L, const std::vector<V3>& coords, fl v, fl slope,
                  std::vector<V3>* minus_forces);

struct EvalCtx {
  const Grid* grid; const Ligand* L; fl v_intra, v_inter, slope;
  long long evals = 0;
  bool noncache = false;  // receptor atoms directly (non_cache) instead of the grid maps
  std::vector<V3> coords, forces; Frames F;
  fl e_inter = 0;
  fl operator()(const Conf& c, std::vector<fl>& g) {
    ++evals;
    set_conf(*L, c, coords, F);
    const int n = (int)L->atoms.size();
    forces.assign(n, V3());
    fl e = 0;
    if (noncache && g_use_inter) e = noncache_terms(*grid, *L, coords, v_inter, slope, &forces);
    for (int i = 0; i < n && g_use_inter && !noncache; ++i) {  // cache::eval_deriv
      int t = L->atoms[i].xs; if (t >= XS_SIZE) continue;
      V3 d; e += grid->evaluate(t, coords[i], slope, v_inter, &d); forces[i] = d;
    }
    e_inter = e;
    const fl cut2 = grid->lut->cutoff_sqr;
    for (const Pair& p : L->pairs) {  // eval_interacting_pairs_deriv
      if (!g_use_intra) break;
      V3 r = coords[p.b] - coords[p.a]; fl r2 = dot(r, r);
      if (r2 < cut2) {
        fl pe, dor; grid->lut->eval_deriv(p.tp, r2, pe, dor);
        V3 force = dor * r; curl(pe, force, v_intra);
        e += pe; forces[p.a] = forces[p.a] - force; forces[p.b] = forces[p.b] + force;
      }
    }
    // derivative: reverse preorder folds children into parents
    const int N = (int)L->nodes.size();
    std::vector<V3> Fs(N), Ts(N);
    for (int k = 0; k < N; ++k)
      for (int i = L->nodes[k].begin; i < L->nodes[k].end; ++i) {
        Fs[k] = Fs[k] + forces[i]; Ts[k] = Ts[k] + cross(coords[i] - F.origin[k], forces[i]);
      }
    g.assign(6 + N - 1, 0);
    for (int k = N - 1; k >= 1; --k) {
      g[6 + k - 1] = dot(Ts[k], F.axis[k]);
      int p = L->nodes[k].parent;
      Fs[p] = Fs[p] + Fs[k];
      Ts[p] = Ts[p] + cross(F.origin[k] - F.origin[p], Fs[k]) + Ts[k];
    }
    for (int i = 0; i < 3; ++i) { g[i] = Fs[0][i]; g[3 + i] = Ts[0][i]; }
    return e;
  }
};

// conformation increment: rigid_conf::increment (conf.h:118-124), torsions_increment (:60-66),
// same as output_type_cl_increment (V11/inc/cuda/mutate_conf.cuh:8-29)
void increment(Conf& x, const std::vector<fl>& d, fl factor) {
  for (int i = 0; i < 3; ++i) x.pos[i] += factor * d[i];
  quaternion_increment(x.q, V3(factor * d[3], factor * d[4], factor * d[5]));
  for (size_t k = 0; k < x.tors.size(); ++k) { x.tors[k] += normalized_angle(factor * d[6 + k]); normalize_angle(x.tors[k]); }
}

fl vdot(const std::vector<fl>& a, const std::vector<fl>& b) { fl s = 0; for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i]; return s; }

// ---------------------------------------------------------------------------------------
// Optimisers, written over an abstract point type so that the same code runs on conformations
// and on closed-form test functions.
//   RILC-BFGS + Lewis-Overton line search: V11/inc/cuda/quasi_newton.cuh:1710-1806, 2172-2389
//   dense BFGS + backtracking line search: V11/lib/bfgs.h:30-140 (= quasi_newton.cuh:1859-1916, 2391-2499)
template <class P, class F, class Inc>
int line_search_lo(F& f, Inc& inc, fl& stp, P& x, fl& fx, std::vector<fl>& g, const std::vector<fl>& d,
                   const P& xp, const std::vector<fl>& gp) {
  int count = 0; bool brackt = false; fl mu = 0, nu = 1e20;
  fl dginit = vdot(gp, d);
  if (0.0 < dginit) return -1;
  fl finit = fx, dgtest = 1e-4 * dginit, dstest = 0.1 * dginit;
  while (true) {
    x = xp; inc(x, d, stp);
    fx = f(x, g); ++count;
    if (fx > finit + stp * dgtest) { nu = stp; brackt = true; }
    else { if (vdot(g, d) < dstest) mu = stp; else return count; }
    if (10 <= count) return (fx > finit) ? -1 : count;
    stp = brackt ? 0.5 * (mu + nu) : stp * 2.0;
  }
}

struct NoHooks {   // observers of a RILC-BFGS run (QuickVina-W records points through them)
  template <class P> void iter(const P&, fl, const std::vector<fl>&) {}   // an iteration ended, search continues
  template <class P> void stop(const P&, fl, const std::vector<fl>&) {}   // line search failed / step budget spent
  template <class P> void done(const P&, fl, const std::vector<fl>&) {}   // the loop ended by one of those two
};

template <class P, class F, class Inc, class H = NoHooks>
fl rilc_bfgs(F& f, Inc& inc, P& x, std::vector<fl>& g, int max_steps, H* hooks = nullptr) {
  fl fx = f(x, g);
  const int n = (int)g.size();
  fl fx_orig = fx; std::vector<fl> g_orig = g; P x_orig = x;
  std::vector<fl> d(n), s(n), y(n);
  for (int i = 0; i < n; ++i) d[i] = -g[i];
  if (!(std::sqrt(vdot(g, g)) >= 1e-5)) return fx;
  fl step = 1.0 / std::sqrt(vdot(d, d));
  int k = 1;
  while (true) {
    P xp = x; std::vector<fl> gp = g; fl fxp = fx;
    int ls = line_search_lo(f, inc, step, x, fx, g, d, xp, gp);
    if (ls < 0) { x = xp; g = gp; fx = fxp; if (hooks) hooks->stop(x, fx, g); break; }
    if (!(std::sqrt(vdot(g, g)) >= 1e-5)) return fx;
    if (max_steps != 0 && max_steps <= k) { if (hooks) hooks->stop(x, fx, g); break; }
    ++k;
    for (int i = 0; i < n; ++i) { s[i] = step * d[i]; y[i] = g[i] - gp[i]; }
    fl ys = vdot(y, s), yy = vdot(y, y);
    for (int i = 0; i < n; ++i) d[i] = -g[i];
    fl cau = vdot(s, s) * std::sqrt(vdot(gp, gp)) * 1.0e-6;
    if (ys > cau) {
      fl alpha = vdot(s, d) / ys;
      for (int i = 0; i < n; ++i) d[i] = d[i] + alpha * y[i];   // sign as in quasi_newton.cuh:2342-2350
      for (int i = 0; i < n; ++i) d[i] = d[i] * (ys / yy);
      fl beta = vdot(y, d) / ys;
      for (int i = 0; i < n; ++i) d[i] = d[i] + (alpha - beta) * s[i];
    }
    step = 1.0;
    if (hooks) hooks->iter(x, fx, g);
  }
  if (hooks) hooks->done(x, fx, g);
  if (!(fx <= fx_orig)) { fx = fx_orig; x = x_orig; g = g_orig; }
  return fx;
}

template <class P, class F, class Inc>
fl dense_bfgs(F& f, Inc& inc, P& x, std::vector<fl>& g, int max_steps) {
  fl f0 = f(x, g), f_orig = f0; std::vector<fl> g_orig = g; P x_orig = x;
  const int n = (int)g.size();
  std::vector<fl> h(n * (n + 1) / 2, 0);  // packed triangular, index i + j(j+1)/2 for i <= j
  auto H = [&](int i, int j) -> fl& { if (i > j) std::swap(i, j); return h[i + j * (j + 1) / 2]; };
  for (int i = 0; i < n; ++i) H(i, i) = 1;
  std::vector<fl> g_new(n), p(n), y(n), mhy(n);
  P x_new = x;
  for (int step = 0; step < max_steps; ++step) {
    for (int i = 0; i < n; ++i) { fl sm = 0; for (int j = 0; j < n; ++j) sm += H(i, j) * g[j]; p[i] = -sm; }
    // line_search, bfgs.h:66-82
    fl f1 = 0, alpha = 1; const fl pg = vdot(p, g);
    for (int trial = 0; trial < 10; ++trial) {
      x_new = x; inc(x_new, p, alpha); f1 = f(x_new, g_new);
      if (f1 - f0 < 1e-4 * alpha * pg) break;
      alpha *= 0.5;
    }
    for (int i = 0; i < n; ++i) y[i] = g_new[i] - g[i];
    f0 = f1; x = x_new;
    if (!(std::sqrt(vdot(g, g)) >= 1e-5)) break;
    g = g_new;
    if (step == 0) { fl yy = vdot(y, y); if (std::fabs(yy) > EPS) { fl dg = alpha * vdot(y, p) / yy; for (int i = 0; i < n; ++i) H(i, i) = dg; } }
    // bfgs_update, bfgs.h:49-63
    fl yp = vdot(y, p);
    if (alpha * yp < EPS) continue;
    for (int i = 0; i < n; ++i) { fl sm = 0; for (int j = 0; j < n; ++j) sm += H(i, j) * y[j]; mhy[i] = -sm; }
    fl yhy = -vdot(y, mhy), r = 1 / (alpha * yp);
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j)
        H(i, j) += alpha * r * (mhy[i] * p[j] + mhy[j] * p[i]) + alpha * alpha * (r * r * yhy + r) * p[i] * p[j];
  }
  if (!(f0 <= f_orig)) { f0 = f_orig; x = x_orig; g = g_orig; }
  return f0;
}

struct ConfInc { void operator()(Conf& x, const std::vector<fl>& d, fl a) const { increment(x, d, a); } };
struct VecInc { void operator()(std::vector<fl>& x, const std::vector<fl>& d, fl a) const { for (size_t i = 0; i < x.size(); ++i) x[i] += a * d[i]; } };

fl local_search(EvalCtx& ctx, Conf& x, int max_steps, int mode);

// QuickVina2 visited history and its gate (QV2/inc/cuda/visited.cuh:7-226, used at
// QV2/inc/cuda/quasi_newton.cuh:1545-1550), with the intended semantics: check() answers
// "interesting" unless some dimension shows the new point on the same monotone slope as the
// neighbour (the reference returns false at its end as well, so its gate never lets a local
// search through once the history is full), torsion dimensions use the torsion derivative (the
// reference reads the position derivative there), and entries are stored by value.
struct VEntry { std::vector<fl> x; fl f; uint32_t zero = 0, pos = 0; };
struct Visited {
  int n = 0, count = 0, p = 0; bool full = false;
  int rounds = 10;   // capacity = rounds x n: 10 for QuickVina2, 5 for QuickVina-W's individual buffer
  std::vector<VEntry> list;
  static std::vector<fl> flat(const Conf& c) {
    std::vector<fl> v{c.pos[0], c.pos[1], c.pos[2], c.q.a, c.q.b, c.q.c, c.q.d};
    v.insert(v.end(), c.tors.begin(), c.tors.end());
    return v;
  }
  // coordinate compared for change dimension i: position, the first three quaternion
  // components (as the reference compares them), torsions
  static fl coord(const std::vector<fl>& x, int i) { return i < 3 ? x[i] : i < 6 ? x[i] : x[i + 1]; }
  void add(const Conf& c, fl f, const std::vector<fl>& g) {
    VEntry e; e.x = flat(c); e.f = f;
    for (size_t i = 0; i < g.size(); ++i) { if (g[i] == 0) e.zero |= 1u << i; else if (g[i] > 0) e.pos |= 1u << i; }
    if (count == 0) { n = (int)g.size(); list.assign(rounds * n, VEntry()); }
    if (!full) { list[count++] = e; if (count >= rounds * n) { full = true; p = 0; } }
    else { list[p] = e; p = (p + 1) % (rounds * n); }
  }
  bool check(const std::vector<fl>& x, fl f, const std::vector<fl>& g, const VEntry& nb) const {
    const bool y_bigger = (f - nb.f) > 0;
    for (int i = 0; i < (int)g.size(); ++i) {
      if ((nb.zero >> i & 1u) || g[i] == 0) continue;
      const bool now_pos = g[i] > 0;
      if (now_pos != (bool)(nb.pos >> i & 1u)) continue;
      const bool x_bigger = (coord(x, i) - coord(nb.x, i)) > 0;
      if (now_pos ? (x_bigger != y_bigger) : (x_bigger == y_bigger)) continue;
      return false;
    }
    return true;
  }
  bool interesting(const Conf& c, fl f, const std::vector<fl>& g) const {
    if (!full) return true;
    return nearest_check(list, flat(c), f, g, n) < 0;
  }
  // The up-to-max_checks nearest of entries (squared distance over the 7+T conformation numbers,
  // lowest index among equals) in order of distance: -1 as soon as one does not explain the
  // point (interesting), else the number examined.
  int nearest_check(const std::vector<VEntry>& entries, const std::vector<fl>& x, fl f, const std::vector<fl>& g,
                    int max_checks) const {
    const int len = (int)entries.size();
    std::vector<fl> dist(len);
    for (int k = 0; k < len; ++k) { fl d = 0; for (size_t j = 0; j < x.size(); ++j) d += sqr(entries[k].x[j] - x[j]); dist[k] = d; }
    std::vector<char> picked(len, 0);
    int it = 0;
    for (; it < std::min(max_checks, len); ++it) {
      int best = -1;
      for (int k = 0; k < len; ++k) if (!picked[k] && (best < 0 || dist[k] < dist[best])) best = k;
      picked[best] = 1;
      if (check(x, f, g, entries[best])) return -1;
    }
    return it;
  }
  // QuickVina-W's individual check (individual_interesting_cl): interesting until the buffer is
  // full, then the up-to-(4n - excluded) nearest entries decide.
  int interesting_qvw(const Conf& c, fl f, const std::vector<fl>& g, int excluded) const {
    if (!full) return -1;
    return nearest_check(list, flat(c), f, g, 4 * n - excluded);
  }
};

// QuickVina-W's global buffer (QVW/inc/cuda/visited.cuh: add_to_global_buffer,
// getPointsWithinCutoff_cl, global_interesting_cl; get_n and set_Binary_Id_cl for the cell),
// intended semantics: one buffer per ligand shared by all its chains; the box is cut into 2^L
// cells per axis, L = (get_n(chains x search depth) - 3) / 3 levels (get_n: the smallest multiple
// of 3, i, with chains x depth < 2^(i+1)), capped at 4; each cell keeps a ring of the last 8
// local-search end points added in it.  The G-check takes the cell holding the point's position
// (the reference takes the cell of the last point added and never adds, as its only call
// passes global_flag = true) and examines up to n = 6+T of its entries, nearest first.
struct GlobalVisited {
  int per_axis = 1; fl lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  std::vector<std::vector<VEntry>> cells; std::vector<long long> counter;
  std::mutex m;
  static int levels_for(long long chains, long long depth) {
    int i = 0;
    while (!(chains * depth < (2LL << i))) i += 3;
    return std::max(0, std::min(4, (i - 3) / 3));
  }
  void init(const fl* b, const fl* e, int levels) {
    per_axis = 1 << levels;
    for (int k = 0; k < 3; ++k) { lo[k] = b[k]; hi[k] = e[k]; }
    cells.assign((size_t)per_axis * per_axis * per_axis, {});
    counter.assign(cells.size(), 0);
  }
  int cell_of(const Conf& c) const {
    int id[3];
    for (int k = 0; k < 3; ++k) {
      const fl t = (c.pos[k] - lo[k]) / (hi[k] - lo[k]) * per_axis;
      id[k] = t >= 0 ? std::min(per_axis - 1, (int)t) : 0;
    }
    return (id[2] * per_axis + id[1]) * per_axis + id[0];
  }
  void add(const Conf& c, fl f, const std::vector<fl>& g) {
    VEntry e; e.x = Visited::flat(c); e.f = f;
    for (size_t i = 0; i < g.size(); ++i) { if (g[i] == 0) e.zero |= 1u << i; else if (g[i] > 0) e.pos |= 1u << i; }
    std::lock_guard<std::mutex> lk(m);
    const int cl = cell_of(c);
    if (cells[cl].size() < 8) cells[cl].push_back(e); else cells[cl][counter[cl] % 8] = e;
    ++counter[cl];
  }
  int interesting(const Visited& rule, const Conf& c, fl f, const std::vector<fl>& g) {
    std::vector<VEntry> here;
    { std::lock_guard<std::mutex> lk(m); here = cells[cell_of(c)]; }
    return rule.nearest_check(here, Visited::flat(c), f, g, (int)g.size());
  }
};

// QuickVina-W's local search (QVW/inc/cuda/quasi_newton.cuh:2180-2420): the G-check, then the
// individual check with the examined count excluded; a point neither finds interesting is not
// searched.  A searched start point, every iteration's end point and the point where the search
// stops go to the individual buffer; the search's last point goes to the global buffer.
struct QvwHooks {
  Visited* ind; GlobalVisited* glob;
  template <class P> void iter(const P& x, fl f, const std::vector<fl>& g) { ind->add(x, f, g); }
  template <class P> void stop(const P& x, fl f, const std::vector<fl>& g) { ind->add(x, f, g); }
  template <class P> void done(const P& x, fl f, const std::vector<fl>& g) { glob->add(x, f, g); }
};
bool qvw_gate(Visited& ind, GlobalVisited& glob, const Conf& x, fl f, const std::vector<fl>& g) {
  const int ret = glob.interesting(ind, x, f, g);
  return !(ret >= 0 && ind.interesting_qvw(x, f, g, ret) >= 0);
}

fl local_search_qvw(EvalCtx& ctx, Conf& x, int max_steps, Visited& ind, GlobalVisited& glob) {
  std::vector<fl> g;
  const fl f = ctx(x, g);
  if (!qvw_gate(ind, glob, x, f, g)) { x.e = f; return f; }
  ind.add(x, f, g);
  --ctx.evals;             // the search below starts by evaluating the same point: count it once
  ConfInc inc;
  QvwHooks h{&ind, &glob};
  const fl e = rilc_bfgs(ctx, inc, x, g, max_steps, &h);
  x.e = e;
  return e;
}

fl local_search(EvalCtx& ctx, Conf& x, int max_steps, int mode, Visited* visited) {
  if (visited) {
    std::vector<fl> g;
    const fl f = ctx(x, g);
    if (!visited->interesting(x, f, g)) { x.e = f; return f; }
    visited->add(x, f, g);
    --ctx.evals;           // the search below starts by evaluating the same point: count it once
  }
  return local_search(ctx, x, max_steps, mode);
}

fl local_search(EvalCtx& ctx, Conf& x, int max_steps, int mode) {
  std::vector<fl> g;
  ConfInc inc;
  fl e = mode == 1 ? rilc_bfgs(ctx, inc, x, g, max_steps) : dense_bfgs(ctx, inc, x, g, max_steps);
  x.e = e;
  return e;
}

// ---------------------------------------------------------------------------------------
// After the search (V11/main/main.cpp:420-470): each kept pose is refined against the receptor
// atoms themselves, and the final score is adjusted by the ligand's torsion count.

// non_cache::eval and ::eval_deriv (V11/lib/non_cache.cpp:38-128): per typed movable atom, the
// coordinates clamped to the box, slope x the distance outside it, and every typed receptor
// atom within the cutoff through the table -- fast values for the energy alone, interpolated
// values and dor for the derivative -- capped by curl(v).  minus_forces receives dE/dx.
fl noncache_terms(const Grid& g, const Ligand& L, const std::vector<V3>& coords, fl v, fl slope,
                  std::vector<V3>* minus_forces) {
  const Lut& lut = *g.lut;
  fl e = 0;
  const int n = (int)L.atoms.size();
  if (minus_forces) minus_forces->assign(n, V3());
  for (int i = 0; i < n; ++i) {
    const int t1 = L.atoms[i].xs;
    if (t1 >= XS_SIZE) continue;
    V3 adj = coords[i], oob;
    fl pen = 0;
    for (int j = 0; j < 3; ++j) {
      if (g.n[j] <= 0) continue;
      if (coords[i][j] < g.begin[j]) { adj[j] = g.begin[j]; oob[j] = -1; pen += std::fabs(coords[i][j] - g.begin[j]); }
      else if (coords[i][j] > g.end[j]) { adj[j] = g.end[j]; oob[j] = 1; pen += std::fabs(coords[i][j] - g.end[j]); }
    }
    pen *= slope;
    fl this_e = 0; V3 deriv;
    for (size_t b = 0; b < g.rxyz.size(); ++b) {
      const int t2 = g.rxs[b];
      if (t2 >= XS_SIZE) continue;
      const V3 r = adj - g.rxyz[b];
      const fl r2 = dot(r, r);
      if (r2 >= lut.cutoff_sqr) continue;
      const int tp = tri_index(t1, t2);
      if (minus_forces) { fl pe, dor; lut.eval_deriv(tp, r2, pe, dor); this_e += pe; deriv = deriv + dor * r; }
      else this_e += lut.eval_fast(tp, r2);
    }
    if (minus_forces) { curl(this_e, deriv, v); (*minus_forces)[i] = deriv + slope * oob; }
    else if (this_e > 0 && v < 0.1 * MAXFL) this_e *= (v < EPS) ? 0 : v / (v + this_e);
    e += this_e + pen;
  }
  return e;
}

// eval_interacting_pairs (V11/lib/model.cpp:1201-1214): fast table values, curl on the energy.
fl intra_fast(const Lut& lut, const Ligand& L, const std::vector<V3>& coords, fl v) {
  fl e = 0;
  for (const Pair& p : L.pairs) {
    const fl r2 = dist2(coords[p.a], coords[p.b]);
    if (r2 >= lut.cutoff_sqr) continue;
    fl t = lut.eval_fast(p.tp, r2);
    if (t > 0 && v < 0.1 * MAXFL) t *= (v < EPS) ? 0 : v / (v + t);
    e += t;
  }
  return e;
}

// conf_independent_inputs::num_tors (V11/lib/terms.cpp:54-125): half the rotatable bonds seen
// from each heavy atom, counting a bond only when the other atom is heavy and has more than one
// heavy neighbour (terminal CH3, OH, NH2 are not counted).
fl num_tors(const Ligand& L) {
  auto rotatable = [&](int a, int b) {
    for (auto& r : L.rotors) if ((r.first == a && r.second == b) || (r.first == b && r.second == a)) return true;
    return false;
  };
  auto heavy_nb = [&](int a) { int k = 0; for (int b : L.atoms[a].bonds) if (L.atoms[b].el != EL_H) ++k; return k; };
  fl nt = 0;
  for (int i = 0; i < (int)L.atoms.size(); ++i) {
    if (L.atoms[i].el == EL_H) continue;
    int ar = 0;
    for (int b : L.atoms[i].bonds)
      if (rotatable(i, b) && L.atoms[b].el != EL_H && heavy_nb(b) > 1) ++ar;
    nt += 0.5 * ar;
  }
  return nt;
}

// weighted_terms::conf_independent with Vina's only such term, num_tors_div
// (V11/lib/everything.cpp:286-303, weighted_terms.cpp:83-89): the weight stored for it is
// 5 w_rot / 0.1 - 1 (main.cpp:649-654), so the divisor is 1 + w_rot * num_tors; smooth_div.
fl conf_independent(const Ligand& L, fl e, fl w_rot) {
  const fl stored = 5 * w_rot / 0.1 - 1, w = 0.1 * (stored + 1);
  const fl y = 1 + w * num_tors(L) / 5.0;
  if (std::fabs(e) < EPS) return 0;
  if (std::fabs(y) < EPS) return e * y > 0 ? MAXFL : -MAXFL;
  return e / y;
}

// model::rmsd_lower_bound / rmsd_upper_bound (V11/lib/model.cpp:1453-1521), heavy atoms.
fl rmsd_lb_asym(const Ligand& L, const std::vector<V3>& x, const std::vector<V3>& y) {
  fl sum = 0; int cnt = 0;
  for (size_t i = 0; i < L.atoms.size(); ++i) {
    if (L.atoms[i].el == EL_H) continue;
    fl r2 = MAXFL;
    for (size_t j = 0; j < L.atoms.size(); ++j)
      if (L.atoms[j].el == L.atoms[i].el && L.atoms[j].el != EL_H) r2 = std::min(r2, dist2(x[i], y[j]));
    sum += r2; ++cnt;
  }
  return cnt ? std::sqrt(sum / cnt) : 0;
}
fl rmsd_ub(const Ligand& L, const std::vector<V3>& x, const std::vector<V3>& y) {
  fl sum = 0; int cnt = 0;
  for (size_t i = 0; i < L.atoms.size(); ++i) if (L.atoms[i].el != EL_H) { sum += dist2(x[i], y[i]); ++cnt; }
  return cnt ? std::sqrt(sum / cnt) : 0;
}

bool within_box(const Grid& g, const Ligand& L, const std::vector<V3>& coords, fl margin = 0.0001) {  // non_cache.h:43, .cpp:58-66
  for (size_t i = 0; i < L.atoms.size(); ++i) {
    if (L.atoms[i].el == EL_H) continue;
    for (int j = 0; j < 3; ++j)
      if (g.n[j] > 0 && (coords[i][j] < g.begin[j] - margin || coords[i][j] > g.end[j] + margin)) return false;
  }
  return true;
}

// refine_structure (V11/main/main.cpp:139-157): dense BFGS (bfgs.h) on the receptor atoms
// directly, inter and intra capped at v, with the out-of-box slope raised 100 -> 1e10 until the
// heavy atoms lie within the box; a pose that never gets inside scores MAXFL.
fl refine(const Grid& g, const Ligand& L, Conf& c, int max_steps, fl v, std::vector<V3>& coords) {
  EvalCtx ctx{&g, &L, v, v, 0};
  ctx.noncache = true;
  ConfInc inc;
  fl e = 0;
  Frames F;
  for (int p = 0; p < 5; ++p) {
    ctx.slope = 100 * std::pow(10.0, 2.0 * p);
    std::vector<fl> gr;
    e = dense_bfgs(ctx, inc, c, gr, max_steps);
    set_conf(L, c, coords, F);
    if (within_box(g, L, coords)) break;
  }
  if (!within_box(g, L, coords)) e = MAXFL;
  return e;
}

// main.cpp:425-470 for one ligand: refine every pose, sort, rescore against the best pose's
// intramolecular energy (model::eval_adjusted, model.cpp:1428-1431), sort again,
// remove_redundant (main.cpp:172-178) and keep up to num_modes poses within energy_range of
// the best, with rmsd lower/upper bounds against the best (model.cpp:1453-1521).
struct Mode { Conf c; fl e = 0, lb = 0, ub = 0, refined = 0; std::vector<V3> coords; };
std::vector<Mode> postprocess(const Grid& g, const Ligand& L, std::vector<Conf> poses, int max_steps,
                              fl w_rot, int num_modes, fl energy_range, fl min_rmsd, fl slope) {
  const fl v = 1000;
  std::vector<Mode> out(poses.size());
  for (size_t i = 0; i < poses.size(); ++i) {
    out[i].c = poses[i];
    out[i].e = out[i].refined = refine(g, L, out[i].c, max_steps, v, out[i].coords);
  }
  auto by_e = [](const Mode& a, const Mode& b) { return a.e < b.e; };
  std::stable_sort(out.begin(), out.end(), by_e);
  if (out.empty()) return out;
  Frames F; std::vector<V3> coords;
  set_conf(L, out[0].c, coords, F);
  const fl best_intra = intra_fast(*g.lut, L, coords, v);
  for (Mode& m : out) {
    if (!(m.e < 0.1 * MAXFL)) continue;
    set_conf(L, m.c, coords, F);
    const fl e = noncache_terms(g, L, coords, v, slope, nullptr) + intra_fast(*g.lut, L, coords, v);
    m.e = conf_independent(L, e - best_intra, w_rot);
  }
  std::stable_sort(out.begin(), out.end(), by_e);
  // remove_redundant: add_to_output_container with max_size = in.size(), on heavy-atom coordinates
  auto heavy_rmsd = [&](const Mode& a, const Mode& b) {
    fl acc = 0; int cnt = 0;
    for (size_t i = 0; i < L.atoms.size(); ++i) if (L.atoms[i].el != EL_H) { acc += dist2(a.coords[i], b.coords[i]); ++cnt; }
    return cnt ? std::sqrt(acc / cnt) : 0.0;
  };
  std::vector<Mode> kept;
  for (const Mode& t : out) {
    int closest = -1; fl best = MAXFL;
    for (size_t i = 0; i < kept.size(); ++i) { fl r = heavy_rmsd(t, kept[i]); if (i == 0 || r < best) { best = r; closest = (int)i; } }
    if (closest >= 0 && best < min_rmsd) { if (t.e < kept[closest].e) kept[closest] = t; }
    else if (kept.size() < out.size()) kept.push_back(t);
    else if (!kept.empty() && t.e < kept.back().e) kept.back() = t;
    std::stable_sort(kept.begin(), kept.end(), by_e);
  }
  std::vector<Mode> modes;
  for (const Mode& m : kept) {
    if ((int)modes.size() >= num_modes || !(m.e < 0.1 * MAXFL) || m.e > kept[0].e + energy_range) break;
    Mode r = m;
    r.lb = std::max(rmsd_lb_asym(L, m.coords, kept[0].coords), rmsd_lb_asym(L, kept[0].coords, m.coords));
    r.ub = rmsd_ub(L, m.coords, kept[0].coords);
    modes.push_back(r);
  }
  return modes;
}

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the random stream layout shared with the GPU
inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
  uint32_t c[4] = {c_in[0], c_in[1], c_in[2], c_in[3]}, k[2] = {k_in[0], k_in[1]};
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0], n1 = lo1, n2 = hi0 ^ c[3] ^ k[1], n3 = lo0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u;
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}
inline fl u01(uint32_t x) { return (x >> 8) * (1.0 / 16777216.0); }
const uint32_t INIT_STEP = 0xFFFFFFFFu;

Conf initial_conf(const Ligand& L, const orc_mc_params& p, int chain) {
  // conf::randomize (conf.h:127-130, 214-217; quaternion.cpp:206-219; random.cpp:125-130)
  uint32_t key[2] = {p.seed, p.stream}, w[4];
  Conf x;
  uint32_t c0[4] = {(uint32_t)chain, INIT_STEP, 0, 0}; philox(c0, key, w);
  for (int i = 0; i < 3; ++i) x.pos[i] = p.box_lo[i] + (p.box_hi[i] - p.box_lo[i]) * u01(w[i]);
  for (uint32_t a = 0;; ++a) {
    uint32_t c1[4] = {(uint32_t)chain, INIT_STEP, 1, a}; philox(c1, key, w);
    fl r0 = std::sqrt(-2 * std::log(u01(w[0]) + 1.0 / 16777216.0)), r1 = std::sqrt(-2 * std::log(u01(w[2]) + 1.0 / 16777216.0));
    fl n0 = r0 * std::cos(2 * PI * u01(w[1])), n1 = r0 * std::sin(2 * PI * u01(w[1]));
    fl n2 = r1 * std::cos(2 * PI * u01(w[3])), n3 = r1 * std::sin(2 * PI * u01(w[3]));
    fl nrm = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2 + n3 * n3);
    if (nrm > 1e-6) { x.q = Q{n0 / nrm, n1 / nrm, n2 / nrm, n3 / nrm}; break; }
  }
  x.tors.resize(L.T());
  for (int j = 0; j < L.T(); ++j) {
    uint32_t c2[4] = {(uint32_t)chain, INIT_STEP, 2 + (uint32_t)(j / 4), 0}; philox(c2, key, w);
    x.tors[j] = -PI + 2 * PI * u01(w[j % 4]);
  }
  return x;
}

fl gyration_radius(const Ligand& L, const std::vector<V3>& coords, const V3& origin) {  // model.cpp:1161-1175
  fl acc = 0; int cnt = 0;
  for (size_t i = 0; i < L.atoms.size(); ++i) if (L.atoms[i].el != EL_H) { acc += dist2(coords[i], origin); ++cnt; }
  return cnt > 0 ? std::sqrt(acc / cnt) : 0;
}

// One Monte Carlo chain with the reference GPU schedule (V11/inc/cuda/kernel2.cu:290-426):
// mutate -> local search -> Metropolis(T) -> on accept, if better than best: second local
// search, keep best.  Mutation per V11/lib/mutate.cpp:62-127.
// One MC chain's state (monte_carlo.cpp:93-127 / kernel2.cu:219-426 as a step function).
struct McState {
  Conf tmp, best; fl rg = 0, best_e = 0; std::vector<V3> best_heavy; Visited vis;
};
// What one MC step decided (orc_mc_trace): the entity mutated, the Metropolis draw, the
// conformation before and after the mutation, the candidate's energy after its local search, the
// current energy it was compared with, and the decision.
struct StepTrace {
  int which = -1, accept = 0; fl u = 0, e_prev = 0, e_cand = 0, rg = 0;
  Conf before, mutated;
};
void mc_init(EvalCtx& ctx, const orc_mc_params& p, int chain, McState& st, const double* starts = nullptr) {
  const Ligand& L = *ctx.L;
  st.tmp = initial_conf(L, p, chain);
  if (starts) {   // the caller's starting conformation (kernel_monte's ric, V11/inc/cuda/kernel2.cuh:27)
    st.tmp = conf_from(L, starts + (size_t)chain * (7 + L.T()));
    const fl n = std::sqrt(st.tmp.q.a * st.tmp.q.a + st.tmp.q.b * st.tmp.q.b + st.tmp.q.c * st.tmp.q.c + st.tmp.q.d * st.tmp.q.d);
    st.tmp.q = Q{st.tmp.q.a / n, st.tmp.q.b / n, st.tmp.q.c / n, st.tmp.q.d / n};
  }
  st.tmp.e = 0;
  Frames F; std::vector<V3> coords;
  set_conf(L, st.tmp, coords, F);
  st.rg = gyration_radius(L, coords, st.tmp.pos);
  st.best_e = std::numeric_limits<fl>::infinity();
  if (p.policy == 2) st.vis.rounds = 5;
}
void mc_step(EvalCtx& ctx, const orc_mc_params& p, int chain, int step, McState& st, GlobalVisited* glob,
             StepTrace* tr = nullptr) {
  const Ligand& L = *ctx.L;
  uint32_t key[2] = {p.seed, p.stream}, w[4];
  const int T = L.T();
  const int n_ent = p.mutation_quirk ? (1 + T) : (2 + T);
  Visited* vp = p.policy == 1 ? &st.vis : nullptr;
  auto search = [&](Conf& c) {
    if (p.policy == 2) local_search_qvw(ctx, c, p.bfgs_steps, st.vis, *glob);
    else local_search(ctx, c, p.bfgs_steps, p.mode, vp);
  };
  Conf cand = st.tmp;
  uint32_t c0[4] = {(uint32_t)chain, (uint32_t)step, 0, 0}; philox(c0, key, w);
  int which = std::min(int(u01(w[0]) * n_ent), n_ent - 1);
  fl u_metro = u01(w[1]); fl tor = -PI + 2 * PI * u01(w[2]);
  if (which <= 1) {
    V3 sph;
    for (uint32_t a = 0;; ++a) {
      uint32_t c1[4] = {(uint32_t)chain, (uint32_t)step, 1, a}; philox(c1, key, w);
      sph = V3(2 * u01(w[0]) - 1, 2 * u01(w[1]) - 1, 2 * u01(w[2]) - 1);
      if (dot(sph, sph) < 1) break;
    }
    if (which == 0) cand.pos = cand.pos + p.amplitude * sph;
    else if (st.rg > EPS) quaternion_increment(cand.q, (p.amplitude / st.rg) * sph);
  } else {
    cand.tors[which - 2] = tor;
  }
  if (tr) { tr->which = which; tr->u = u_metro; tr->before = st.tmp; tr->mutated = cand; tr->rg = st.rg; tr->e_prev = st.tmp.e; }
  search(cand);
  bool accept = (step == 0) || (cand.e < st.tmp.e) || (u_metro < std::exp((st.tmp.e - cand.e) / p.temperature));  // monte_carlo.cpp:35-39
  if (tr) { tr->e_cand = cand.e; tr->accept = accept; }
  if (accept) {
    st.tmp = cand;
    Frames F; std::vector<V3> coords;
    if (st.tmp.e < st.best_e) {
      search(st.tmp);
      if (st.tmp.e < st.best_e) {
        st.best = st.tmp; st.best_e = st.tmp.e;
        set_conf(L, st.best, coords, F);
        st.best_heavy.clear();
        for (size_t i = 0; i < L.atoms.size(); ++i) if (L.atoms[i].el != EL_H) st.best_heavy.push_back(coords[i]);
      }
    }
    set_conf(L, st.tmp, coords, F);
    st.rg = gyration_radius(L, coords, st.tmp.pos);
  }
}

}  // namespace

// =========================================================================================
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
void orc_set_terms(int use_inter, int use_intra) { g_use_inter = use_inter; g_use_intra = use_intra; }

double orc_term_eval(int t1, int t2, double r, const double w[5]) { return term_eval(t1, t2, r, w, nullptr); }

void* orc_lut_create(const double w[5]) { Lut* l = new Lut(); l->build(w); return l; }
void orc_lut_destroy(void* l) { delete (Lut*)l; }
int orc_lut_size(void) { return int(32 * 64) + 3; }
int orc_lut_get(void* lp, int tp, double* se, double* sd, double* fast) {
  Lut* l = (Lut*)lp; if (tp < 0 || tp >= 153) return fail("tp out of range");
  for (int i = 0; i < l->n; ++i) { se[i] = l->se[tp * l->n + i]; sd[i] = l->sd[tp * l->n + i]; fast[i] = l->fast[tp * l->n + i]; }
  return 0;
}
double orc_lut_eval_fast(void* lp, int tp, double r2) { return ((Lut*)lp)->eval_fast(tp, r2); }
int orc_lut_eval_deriv(void* lp, int tp, double r2, double* e, double* dor) { ((Lut*)lp)->eval_deriv(tp, r2, *e, *dor); return 0; }

int orc_receptor_parse(const char* text, int max_atoms, double* xyz, int* xs, int* ad) {
  try {
    std::vector<AtomT> atoms;
    std::istringstream in(text); std::string s;
    while (std::getline(in, s)) {
      if (!s.empty() && s.back() == '\r') s.pop_back();
      if (starts(s, "ATOM  ") || starts(s, "HETATM")) {
        AtomT a; unsigned num; if (!parse_atom_line(s, a, num)) return fail("bad receptor atom line: " + s);
        atoms.push_back(a);
      }
    }
    Mobility mob; mob.all_fixed = true;
    assign_bonds(atoms, mob);
    assign_types(atoms);
    if ((int)atoms.size() > max_atoms) return fail("too many atoms");
    for (size_t i = 0; i < atoms.size(); ++i) {
      for (int k = 0; k < 3; ++k) xyz[3 * i + k] = atoms[i].coords[k];
      xs[i] = atoms[i].xs; ad[i] = atoms[i].ad;
    }
    return (int)atoms.size();
  } catch (std::exception& e) { return fail(e.what()); }
}

void* orc_grid_create(void* lut, const double begin[3], const double end[3], const int n[3]) {
  Grid* g = new Grid(); g->lut = (const Lut*)lut; g->setup(begin, end, n); return g;
}
void orc_grid_destroy(void* g) { delete (Grid*)g; }
int orc_grid_set_receptor(void* gp, int n, const double* xyz, const int* xs) {
  Grid* g = (Grid*)gp; g->rxyz.resize(n); g->rxs.assign(xs, xs + n);
  for (int i = 0; i < n; ++i) g->rxyz[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  return 0;
}
int orc_grid_populate(void* gp, int use_cells) { ((Grid*)gp)->populate(use_cells != 0); return 0; }
int orc_grid_populate_slack(void* gp, int use_cells, double eps, double* slack_out) {
  Grid* g = (Grid*)gp; g->slack_eps = eps; g->populate(use_cells != 0); g->slack_eps = 0;
  std::memcpy(slack_out, g->slack.data(), g->slack.size() * sizeof(double));
  return 0;
}
int orc_grid_get_maps(void* gp, double* out) { Grid* g = (Grid*)gp; std::memcpy(out, g->maps.data(), g->maps.size() * sizeof(double)); return 0; }
int orc_grid_set_maps(void* gp, const double* in) { Grid* g = (Grid*)gp; std::memcpy(g->maps.data(), in, g->maps.size() * sizeof(double)); return 0; }
int orc_grid_cell_candidates(void* gp, const double p[3], int* out, int max_out) {
  Grid* g = (Grid*)gp; if (g->cells.empty()) g->build_cells();
  const std::vector<int>& c = g->possibilities(V3(p[0], p[1], p[2]));
  for (size_t i = 0; i < c.size() && (int)i < max_out; ++i) out[i] = c[i];
  return (int)c.size();
}
double orc_grid_eval(void* gp, int t, const double xyz[3], double slope, double v, double deriv[3]) {
  V3 d; double e = ((Grid*)gp)->evaluate(t, V3(xyz[0], xyz[1], xyz[2]), slope, v, &d);
  for (int i = 0; i < 3; ++i) deriv[i] = d[i];
  return e;
}

void* orc_ligand_parse(const char* text) {
  try { return parse_ligand(text); } catch (std::exception& e) { fail(e.what()); return nullptr; }
}
void orc_ligand_destroy(void* l) { delete (Ligand*)l; }
int orc_ligand_sizes(void* lp, int s[5]) {
  Ligand* L = (Ligand*)lp; s[0] = (int)L->atoms.size(); s[1] = (int)L->nodes.size(); s[2] = (int)L->pairs.size();
  s[3] = L->heavy; s[4] = L->torsdof; return 0;
}
int orc_ligand_atoms(void* lp, int* xs, int* el, int* ad, int* node, double* local, double* coords0) {
  Ligand* L = (Ligand*)lp;
  for (size_t k = 0; k < L->nodes.size(); ++k) for (int i = L->nodes[k].begin; i < L->nodes[k].end; ++i) node[i] = (int)k;
  for (size_t i = 0; i < L->atoms.size(); ++i) {
    xs[i] = L->atoms[i].xs; el[i] = L->atoms[i].el; ad[i] = L->atoms[i].ad;
    for (int k = 0; k < 3; ++k) { local[3 * i + k] = L->atoms[i].local[k]; coords0[3 * i + k] = L->atoms[i].coords[k]; }
  }
  return 0;
}
int orc_ligand_nodes(void* lp, int* parent, int* begin, int* end, double* ro, double* ra, double* o0) {
  Ligand* L = (Ligand*)lp;
  for (size_t k = 0; k < L->nodes.size(); ++k) {
    const Node& n = L->nodes[k]; parent[k] = n.parent; begin[k] = n.begin; end[k] = n.end;
    for (int j = 0; j < 3; ++j) { ro[3 * k + j] = n.rel_origin[j]; ra[3 * k + j] = n.rel_axis[j]; o0[3 * k + j] = n.origin0[j]; }
  }
  return 0;
}
int orc_ligand_pairs(void* lp, int* tp, int* a, int* b) {
  Ligand* L = (Ligand*)lp;
  for (size_t i = 0; i < L->pairs.size(); ++i) { tp[i] = L->pairs[i].tp; a[i] = L->pairs[i].a; b[i] = L->pairs[i].b; }
  return 0;
}
int orc_ligand_bonds(void* lp, int* bi, int* bj, int max_out) {
  Ligand* L = (Ligand*)lp; int c = 0;
  for (size_t i = 0; i < L->atoms.size(); ++i)
    for (int j : L->atoms[i].bonds) if ((int)i < j) { if (c < max_out) { bi[c] = (int)i; bj[c] = j; } ++c; }
  return c;
}

int orc_set_conf(void* lp, const double* conf, double* coords_out) {
  Ligand* L = (Ligand*)lp; Conf c = conf_from(*L, conf); std::vector<V3> coords; Frames F;
  set_conf(*L, c, coords, F);
  for (size_t i = 0; i < coords.size(); ++i) for (int k = 0; k < 3; ++k) coords_out[3 * i + k] = coords[i][k];
  return 0;
}
double orc_eval_deriv(void* gp, void* lp, const double* conf, double v_intra, double v_inter, double slope,
                      double* grad, double* coords_out, double* e_inter_out) {
  Ligand* L = (Ligand*)lp;
  EvalCtx ctx{(Grid*)gp, L, v_intra, v_inter, slope};
  std::vector<fl> g; Conf c = conf_from(*L, conf);
  double e = ctx(c, g);
  for (size_t i = 0; i < g.size(); ++i) grad[i] = g[i];
  if (coords_out) for (size_t i = 0; i < ctx.coords.size(); ++i) for (int k = 0; k < 3; ++k) coords_out[3 * i + k] = ctx.coords[i][k];
  if (e_inter_out) *e_inter_out = ctx.e_inter;
  return e;
}
int orc_conf_increment(void* lp, double* conf, const double* d, double factor) {
  Ligand* L = (Ligand*)lp; Conf c = conf_from(*L, conf);
  std::vector<fl> dv(d, d + 6 + L->T()); increment(c, dv, factor); conf_to(c, conf); return 0;
}
double orc_local_search(void* gp, void* lp, double* conf, int max_steps, int mode, double v_intra, double v_inter,
                        double slope, int* n_evals) {
  Ligand* L = (Ligand*)lp;
  EvalCtx ctx{(Grid*)gp, L, v_intra, v_inter, slope};
  Conf c = conf_from(*L, conf);
  double e = local_search(ctx, c, max_steps, mode);
  conf_to(c, conf);
  if (n_evals) *n_evals = (int)ctx.evals;
  return e;
}

int orc_visited_feed(int T, int m, const double* x, const double* f, const double* g, int mode, int* interesting) {
  try {
    if (T < 0 || m < 0 || (mode != 1 && mode != 2)) throw std::runtime_error("bad arguments");
    Visited v;
    const int n = 6 + T;
    for (int k = 0; k < m; ++k) {
      Conf c;
      const double* xk = x + (size_t)k * (7 + T);
      c.pos = V3(xk[0], xk[1], xk[2]); c.q = Q{xk[3], xk[4], xk[5], xk[6]};
      c.tors.assign(xk + 7, xk + 7 + T);
      std::vector<fl> gk(g + (size_t)k * n, g + (size_t)(k + 1) * n);
      const bool it = v.interesting(c, f[k], gk);
      interesting[k] = it ? 1 : 0;
      if (mode == 2 || it) v.add(c, f[k], gk);
    }
    return 0;
  } catch (std::exception& e) { return fail(e.what()); }
}

int orc_qvw_feed(int T, int m, const double* x, const double* f, const double* g, const int* chain, const int* op,
                 const double box_lo[3], const double box_hi[3], int levels, int* out) {
  try {
    if (T < 0 || m < 0 || levels < 0 || levels > 4) throw std::runtime_error("bad arguments");
    GlobalVisited glob;
    glob.init(box_lo, box_hi, levels);
    std::map<int, Visited> ind;
    const int n = 6 + T;
    for (int k = 0; k < m; ++k) {
      Conf c;
      const double* xk = x + (size_t)k * (7 + T);
      c.pos = V3(xk[0], xk[1], xk[2]); c.q = Q{xk[3], xk[4], xk[5], xk[6]};
      c.tors.assign(xk + 7, xk + 7 + T);
      std::vector<fl> gk(g + (size_t)k * n, g + (size_t)(k + 1) * n);
      Visited& v = ind[chain[k]];
      v.rounds = 5;
      out[k] = 0;
      if (op[k] == 0) out[k] = qvw_gate(v, glob, c, f[k], gk) ? 1 : 0;
      else if (op[k] == 1) v.add(c, f[k], gk);
      else if (op[k] == 2) { glob.add(c, f[k], gk); out[k] = glob.cell_of(c); }
      else if (op[k] == 3) out[k] = glob.interesting(v, c, f[k], gk);
      else throw std::runtime_error("bad op");
    }
    return 0;
  } catch (std::exception& e) { return fail(e.what()); }
}
int orc_qvw_levels(long long chains, long long depth) { return GlobalVisited::levels_for(chains, depth); }

double orc_opt_test(int kind, int n, double* x, const double* coef, int max_steps, int mode, int* n_evals) {
  return orc_opt_trace(kind, n, x, coef, max_steps, mode, 0, nullptr, nullptr, n_evals);
}

double orc_opt_trace(int kind, int n, double* x, const double* coef, int max_steps, int mode, int max_pts,
                     double* pts, double* fs, int* n_evals) {
  int evals = 0;
  auto f = [&](const std::vector<fl>& p, std::vector<fl>& g) -> fl {
    if (evals < max_pts) for (int i = 0; i < n; ++i) pts[(size_t)evals * n + i] = p[i];
    ++evals; g.assign(n, 0); fl v = 0;
    if (kind == 0) { for (int i = 0; i < n; ++i) { v += 0.5 * coef[i] * p[i] * p[i]; g[i] = coef[i] * p[i]; } }
    else {
      for (int i = 0; i + 1 < n; ++i) {
        fl a = p[i + 1] - p[i] * p[i], b = 1 - p[i];
        v += 100 * a * a + b * b;
        g[i] += -400 * a * p[i] - 2 * b; g[i + 1] += 200 * a;
      }
    }
    if (evals <= max_pts) fs[evals - 1] = v;
    return v;
  };
  std::vector<fl> p(x, x + n), g;
  VecInc inc;
  fl v = mode == 1 ? rilc_bfgs(f, inc, p, g, max_steps) : dense_bfgs(f, inc, p, g, max_steps);
  for (int i = 0; i < n; ++i) x[i] = p[i];
  if (n_evals) *n_evals = evals;
  return v;
}

int orc_initial_conf(void* lp, const orc_mc_params* p, int chain, double* conf) {
  Ligand* L = (Ligand*)lp; conf_to(initial_conf(*L, *p, chain), conf); return 0;
}

int orc_mc_run(void* gp, void* lp, const orc_mc_params* p, double* e, double* conf, double* heavy_coords, long long* evals) {
  return orc_mc_run_from(gp, lp, p, nullptr, e, conf, heavy_coords, evals);
}

int orc_mc_run_from(void* gp, void* lp, const orc_mc_params* p, const double* starts, double* e, double* conf,
                    double* heavy_coords, long long* evals) {
  Ligand* L = (Ligand*)lp; Grid* G = (Grid*)gp;
  const int T = L->T(), H = L->heavy;
  std::atomic<int> next(0); std::atomic<long long> total(0);
  int nth = p->threads > 0 ? p->threads : (int)host_threads();
  std::vector<McState> st(p->chains);
  auto output = [&](int c) {
    e[c] = st[c].best.e;
    conf_to(st[c].best, conf + (size_t)c * (7 + T));
    const std::vector<V3>& heavy = st[c].best_heavy;
    for (int i = 0; i < H; ++i) for (int k = 0; k < 3; ++k) heavy_coords[((size_t)c * H + i) * 3 + k] = i < (int)heavy.size() ? heavy[i][k] : 0;
  };
  if (p->policy != 2) {
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t) pool.emplace_back([&]() {   // parallel_mc.cpp:65-77: one task per chain
      EvalCtx ctx{G, L, p->v_intra, p->v_inter, p->slope};
      for (int c = next++; c < p->chains; c = next++) {
        mc_init(ctx, *p, c, st[c], starts);
        for (int s = 0; s < p->search_depth; ++s) mc_step(ctx, *p, c, s, st[c], nullptr);
        output(c);
      }
      total += ctx.evals;
    });
    for (auto& th : pool) th.join();
  } else {
    // QuickVina-W: the chains share the global buffer, so they advance together, one MC step of
    // every chain before the next step (as the device's resident chains do); threads take the
    // chains of a step in any order
    GlobalVisited glob;
    glob.init(p->box_lo, p->box_hi, GlobalVisited::levels_for(p->chains, p->search_depth));
    for (int s = -1; s < p->search_depth; ++s) {
      next = 0;
      std::vector<std::thread> pool;
      for (int t = 0; t < nth; ++t) pool.emplace_back([&]() {
        EvalCtx ctx{G, L, p->v_intra, p->v_inter, p->slope};
        for (int c = next++; c < p->chains; c = next++) {
          if (s < 0) mc_init(ctx, *p, c, st[c], starts);
          else mc_step(ctx, *p, c, s, st[c], &glob);
        }
        total += ctx.evals;
      });
      for (auto& th : pool) th.join();
    }
    for (int c = 0; c < p->chains; ++c) output(c);
  }
  if (evals) *evals = total;
  return 0;
}

int orc_mc_trace(void* gp, void* lp, const orc_mc_params* p, int chain, int* which, double* u, double* e_prev,
                 double* e_cand, int* accept, double* rg, double* before, double* mutated) {
  Ligand* L = (Ligand*)lp; Grid* G = (Grid*)gp;
  if (p->policy == 2) return fail("orc_mc_trace: the QuickVina-W policy shares a buffer across chains");
  const int nc = 7 + L->T();
  EvalCtx ctx{G, L, p->v_intra, p->v_inter, p->slope};
  McState st;
  mc_init(ctx, *p, chain, st);
  for (int s = 0; s < p->search_depth; ++s) {
    StepTrace tr;
    mc_step(ctx, *p, chain, s, st, nullptr, &tr);
    which[s] = tr.which; u[s] = tr.u; e_prev[s] = tr.e_prev; e_cand[s] = tr.e_cand; accept[s] = tr.accept; rg[s] = tr.rg;
    conf_to(tr.before, before + (size_t)s * nc);
    conf_to(tr.mutated, mutated + (size_t)s * nc);
  }
  return 0;
}

int orc_ligand_serials(void* lp, int* serial) {
  Ligand* L = (Ligand*)lp;
  for (size_t i = 0; i < L->atoms.size(); ++i) serial[i] = (int)L->atoms[i].serial;
  return 0;
}
double orc_num_tors(void* lp) { return num_tors(*(Ligand*)lp); }
double orc_conf_independent(void* lp, double e, double w_rot) { return conf_independent(*(Ligand*)lp, e, w_rot); }
static std::vector<V3> to_v3(const double* xyz, size_t n) {
  std::vector<V3> c(n); for (size_t i = 0; i < n; ++i) c[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]); return c;
}
double orc_noncache_eval(void* gp, void* lp, const double* xyz, double v, double slope, double* minus_forces) {
  Grid* g = (Grid*)gp; Ligand* L = (Ligand*)lp;
  std::vector<V3> f;
  fl e = noncache_terms(*g, *L, to_v3(xyz, L->atoms.size()), v, slope, minus_forces ? &f : nullptr);
  if (minus_forces) for (size_t i = 0; i < f.size(); ++i) for (int k = 0; k < 3; ++k) minus_forces[3 * i + k] = f[i][k];
  return e;
}
double orc_intra_fast(void* gp, void* lp, const double* xyz, double v) {
  Ligand* L = (Ligand*)lp; return intra_fast(*((Grid*)gp)->lut, *L, to_v3(xyz, L->atoms.size()), v);
}
int orc_within(void* gp, void* lp, const double* xyz) {
  Ligand* L = (Ligand*)lp; return within_box(*(Grid*)gp, *L, to_v3(xyz, L->atoms.size())) ? 1 : 0;
}
int orc_rmsd_bounds(void* lp, const double* x, const double* y, double* lb, double* ub) {
  Ligand* L = (Ligand*)lp; const size_t n = L->atoms.size();
  std::vector<V3> a = to_v3(x, n), b = to_v3(y, n);
  *lb = std::max(rmsd_lb_asym(*L, a, b), rmsd_lb_asym(*L, b, a));
  *ub = rmsd_ub(*L, a, b);
  return 0;
}
double orc_noncache_eval_deriv(void* gp, void* lp, const double* conf, double v, double slope, double* grad) {
  Ligand* L = (Ligand*)lp;
  EvalCtx ctx{(Grid*)gp, L, v, v, slope};
  ctx.noncache = true;
  std::vector<fl> g; Conf c = conf_from(*L, conf);
  double e = ctx(c, g);
  for (size_t i = 0; i < g.size(); ++i) grad[i] = g[i];
  return e;
}
double orc_refine(void* gp, void* lp, double* conf, int max_steps, double v, double* coords_out) {
  Ligand* L = (Ligand*)lp; Conf c = conf_from(*L, conf); std::vector<V3> coords;
  fl e = refine(*(Grid*)gp, *L, c, max_steps, v, coords);
  conf_to(c, conf);
  if (coords_out) for (size_t i = 0; i < coords.size(); ++i) for (int k = 0; k < 3; ++k) coords_out[3 * i + k] = coords[i][k];
  return e;
}
int orc_postprocess(void* gp, void* lp, int n, const double* confs, int max_steps, double w_rot, int num_modes,
                    double energy_range, double min_rmsd, double slope, double* conf_out, double* e_out,
                    double* lb_out, double* ub_out, double* refined_out) {
  Ligand* L = (Ligand*)lp; const int cs = 7 + L->T();
  std::vector<Conf> poses;
  for (int i = 0; i < n; ++i) poses.push_back(conf_from(*L, confs + (size_t)i * cs));
  std::vector<Mode> m = postprocess(*(Grid*)gp, *L, poses, max_steps, w_rot, num_modes, energy_range, min_rmsd, slope);
  for (size_t i = 0; i < m.size(); ++i) {
    conf_to(m[i].c, conf_out + i * cs);
    e_out[i] = m[i].e; lb_out[i] = m[i].lb; ub_out[i] = m[i].ub;
    if (refined_out) refined_out[i] = m[i].refined;
  }
  return (int)m.size();
}

int orc_cluster(int n, int heavy, const double* e, const double* hc, double min_rmsd, int max_size, int* kept) {
  std::vector<int> out;  // indices, kept sorted by energy (output_container::sort)
  auto rmsd = [&](int a, int b) {
    fl acc = 0;
    for (int i = 0; i < heavy; ++i) for (int k = 0; k < 3; ++k) acc += sqr(hc[((size_t)a * heavy + i) * 3 + k] - hc[((size_t)b * heavy + i) * 3 + k]);
    return heavy > 0 ? std::sqrt(acc / heavy) : 0.0;
  };
  for (int t = 0; t < n; ++t) {
    int closest = -1; fl best = MAXFL;
    for (size_t i = 0; i < out.size(); ++i) { fl r = rmsd(t, out[i]); if (i == 0 || r < best) { best = r; closest = (int)i; } }
    if (closest >= 0 && best < min_rmsd) { if (e[t] < e[out[closest]]) out[closest] = t; }
    else if ((int)out.size() < max_size) out.push_back(t);
    else if (!out.empty() && e[t] < e[out.back()]) out.back() = t;
    std::stable_sort(out.begin(), out.end(), [&](int a, int b) { return e[a] < e[b]; });
  }
  for (size_t i = 0; i < out.size(); ++i) kept[i] = out[i];
  return (int)out.size();
}

}  // extern "C"
