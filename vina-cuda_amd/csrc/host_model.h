// host_model.h -- host-side model building for libvdock.
//
// Builds what the HIP kernels consume: the FP32 look-up tables (from the FP64 scoring
// function), typed receptor atoms, and packed ligand records (rigid-node tree in DFS
// preorder, atoms with node-local coordinates, intramolecular pairs).  Follows the
// reference's host code:
//   V11/lib/precalculate.h:121-140, 255-339  (LUT)
//   V11/lib/everything.cpp:97-216, 381-479   (terms)
//   V11/lib/parse_pdbqt.cpp:101-118, 286-652 (PDBQT tree)
//   V11/lib/model.cpp:562-896                (bonds, XS typing, pairs)
//   V11/lib/main_procedure_cu.cpp:551-763    (flattening for the device)
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include <stdexcept>
#include "vd_common.h"

namespace vd {

struct parse_error : std::runtime_error {
  explicit parse_error(const std::string& m) : std::runtime_error(m) {}
};

// XS typing helpers (V11/lib/atom_constants.h:209-282)
extern const double kXsRadius[VD_XS_SIZE];
bool xs_is_hydrophobic(int t);
bool xs_is_acceptor(int t);
bool xs_is_donor(int t);
inline int type_pair_index(int a, int b) { if (a > b) { int t = a; a = b; b = t; } return a + b * (b + 1) / 2; }

// Weighted sum of the five Vina terms for XS types t1, t2 at distance r.
double score_terms(int t1, int t2, double r, const double w[5]);

// FP64 tables of precalculate (smooth.first, smooth.second, fast) for all 153 pairs.
struct Precalculated {
  std::vector<double> smooth_e, smooth_dor, fast;   // [153][VD_LUT_BINS]
  void build(const double w[5]);
  // FP32 device layouts:
  //   grid rows  [t1][bin][VD_GRID_ROW]: fast value for t2 = 0..15, then the column [t1][bin] of t2 = 16
  //   pair bins  [tp][bin] = (e_i, dor_i); the 16 bytes at bin i are the interpolation quad
  //              (e_i, dor_i, e_{i+1}, dor_{i+1})
  void grid_rows(std::vector<float>& out) const;
  void pair_bins(std::vector<float>& out) const;
};

struct LigAtom {
  int ad = 20, el = 11, xs = VD_XS_SIZE;
  double xyz[3] = {0, 0, 0};     // absolute input coordinates
  double local[3] = {0, 0, 0};   // relative to the owning node's origin
  int node = 0;
  unsigned serial = 0;
  int ctx_line = -1;             // index of its line in Ligand::context
  std::string line;              // the PDBQT line, for writing poses back out
  std::vector<int> bonds;
};
struct LigNode {
  int parent = -1, begin = 0, end = 0, subtree_end = 0, depth = 0;
  double origin0[3] = {0, 0, 0}, rel_origin[3] = {0, 0, 0}, rel_axis[3] = {0, 0, 0};
};
struct Ligand {
  std::vector<LigAtom> atoms;
  std::vector<LigNode> nodes;      // preorder, node 0 = ROOT
  std::vector<uint32_t> pairs;     // a | b << 8 | tp << 16
  int torsdof = 0, n_heavy = 0, max_depth = 0;
  std::vector<std::string> context; // every input line in file order (for output)
  double num_tors = 0;             // torsion count of the final score's divisor (terms.cpp:97-125)
  int n_torsions() const { return (int)nodes.size() - 1; }
  // conflict-free pair chunks per chain-group width (16, 32, 64), built on first use
  mutable std::vector<uint32_t> chunked[3];
};

// Pairs in chunks of G slots such that within a chunk every atom is the first atom of at most
// one pair and the second atom of at most one pair (each pair may be written either way round;
// padding slots hold VD_NO_PAIR).  Greedy first fit; the chunk count is at least
// max(ceil(P / G), ceil(max pair degree / 2)).
const std::vector<uint32_t>& pair_chunks(const Ligand& L, int G);

Ligand parse_ligand_pdbqt(const std::string& text);
// Typed receptor atoms (rigid receptor only): xyz[3n], xs[n] (17 = no XS type)
void parse_receptor_pdbqt(const std::string& text, std::vector<double>& xyz, std::vector<int>& xs);

// Appends a ligand's packed record to a device blob (16-byte words), with its pairs chunked for
// chain groups of G lanes; returns the header.
vd_lig_header pack_ligand(const Ligand& L, std::vector<float>& blob, int G);

// Synthetic inputs (recipes of the benchmark configurations).
void synth_receptor(int n_atoms, double cube, double cavity_radius, uint64_t seed,
                    std::vector<float>& xyz, std::vector<int>& xs);
std::string synth_ligand_pdbqt(int n_heavy, int n_rot, uint64_t seed);

// add_to_output_container over heavy-atom coordinates (V11/lib/coords.cpp:24-62).
// Feeds results in the given order; returns indices kept, sorted by energy.
std::vector<int> cluster_poses(int n, int heavy, const float* e, const float* heavy_xyz,
                               double min_rmsd, int max_size);

// After the search (post.cpp).  conf_independent: the final score's torsion divisor
// (num_tors_div, V11/lib/everything.cpp:296-303).  rmsd_bounds: rmsd lower/upper bound between
// two all-atom poses (model.cpp:1453-1521).  select_modes: sorting, eval_adjusted,
// remove_redundant and the num_modes / energy_range cut of V11/main/main.cpp:425-470 over n
// refined poses (coords[n][A][3]).  write_models: the PDBQT output of n modes.
struct Mode { int pose = 0; double e = 0, lb = 0, ub = 0; };
double conf_independent(double e, double num_tors, double w_rot);
void rmsd_bounds(const Ligand& L, const float* x, const float* y, double& lb, double& ub);
std::vector<Mode> select_modes(const Ligand& L, int n, const float* e_refined, const float* e_inter,
                               const float* e_intra, const float* coords, double w_rot, int num_modes,
                               double energy_range, double min_rmsd);
std::string write_models(const Ligand& L, int n, const double* e, const double* lb, const double* ub,
                         const float* coords);

}  // namespace vd
