// post.cpp -- host side of the step after the search: final scores, mode selection, rmsd
// bounds and the PDBQT writer (V11/main/main.cpp:89-102, 158-178, 425-501).  The refinement and
// the energies it feeds in come from the device (vd_refine_kernel).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "host_model.h"

namespace vd {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kMax = std::numeric_limits<double>::max();
bool not_max(double x) { return x < 0.1 * kMax; }   // V11/lib/common.h not_max
bool not_max_f(float x) { return x < 0.1f * std::numeric_limits<float>::max(); }
double d2(const float* a, const float* b) {
  const double x = (double)a[0] - b[0], y = (double)a[1] - b[1], z = (double)a[2] - b[2];
  return x * x + y * y + z * z;
}
}  // namespace

// weighted_terms::conf_independent with Vina's one such term, num_tors_div (V11/lib/
// everything.cpp:296-303): the weight stored for it is 5 w_rot / 0.1 - 1 (main.cpp:862) and the
// term reads w = 0.1 (stored + 1), so the divisor is 1 + w num_tors / 5 = 1 + w_rot num_tors;
// smooth_div (V11/lib/everything.cpp) guards the quotient.
double conf_independent(double e, double num_tors, double w_rot) {
  const double stored = 5 * w_rot / 0.1 - 1;
  const double w = 0.1 * (stored + 1);
  const double y = 1 + w * num_tors / 5.0;
  if (std::fabs(e) < kEps) return 0;
  if (std::fabs(y) < kEps) return e * y > 0 ? kMax : -kMax;
  return e / y;
}

// model::rmsd_lower_bound / rmsd_upper_bound (V11/lib/model.cpp:1453-1521) over all atoms
// x[A][3], y[A][3] in the ligand's order: heavy atoms only; the lower bound matches each atom to
// the nearest atom of the same element in the other pose, both ways, and takes the larger.
void rmsd_bounds(const Ligand& L, const float* x, const float* y, double& lb, double& ub) {
  const int n = (int)L.atoms.size();
  auto asym = [&](const float* a, const float* b) {
    double sum = 0; int cnt = 0;
    for (int i = 0; i < n; ++i) {
      if (L.atoms[i].el == 0) continue;
      double r2 = kMax;
      for (int j = 0; j < n; ++j)
        if (L.atoms[j].el == L.atoms[i].el) r2 = std::min(r2, d2(a + 3 * i, b + 3 * j));
      sum += r2; ++cnt;
    }
    return cnt ? std::sqrt(sum / cnt) : 0.0;
  };
  lb = std::max(asym(x, y), asym(y, x));
  double sum = 0; int cnt = 0;
  for (int i = 0; i < n; ++i)
    if (L.atoms[i].el != 0) { sum += d2(x + 3 * i, y + 3 * i); ++cnt; }
  ub = cnt ? std::sqrt(sum / cnt) : 0.0;
}

// V11/main/main.cpp:425-470 once every pose is refined: sort by refined energy, score each pose
// that reached the box with eval_adjusted against the intramolecular energy of the best
// (model.cpp:1428-1431), sort again, remove_redundant (add_to_output_container over heavy-atom
// coordinates with max size = pose count, main.cpp:172-178 and coords.cpp:47-62), then keep up
// to num_modes poses within energy_range of the best, each with rmsd bounds to the best.
std::vector<Mode> select_modes(const Ligand& L, int n, const float* e_refined, const float* e_inter,
                               const float* e_intra, const float* coords, double w_rot, int num_modes,
                               double energy_range, double min_rmsd) {
  const int A = (int)L.atoms.size();
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::vector<double> e(n);
  for (int i = 0; i < n; ++i) e[i] = not_max_f(e_refined[i]) ? (double)e_refined[i] : kMax;
  auto by_e = [&](int a, int b) { return e[a] < e[b]; };
  std::stable_sort(order.begin(), order.end(), by_e);
  if (n == 0) return {};
  const double best_intra = e_intra[order[0]];
  for (int i = 0; i < n; ++i)
    if (not_max(e[i])) e[i] = conf_independent((double)e_inter[i] + (double)e_intra[i] - best_intra, L.num_tors, w_rot);
  std::stable_sort(order.begin(), order.end(), by_e);
  auto heavy_rmsd = [&](int a, int b) {
    double s = 0; int cnt = 0;
    for (int i = 0; i < A; ++i)
      if (L.atoms[i].el != 0) { s += d2(coords + ((size_t)a * A + i) * 3, coords + ((size_t)b * A + i) * 3); ++cnt; }
    return cnt ? std::sqrt(s / cnt) : 0.0;
  };
  std::vector<int> kept;
  for (int t : order) {
    int closest = -1; double best = kMax;
    for (size_t i = 0; i < kept.size(); ++i) {
      const double r = heavy_rmsd(t, kept[i]);
      if (i == 0 || r < best) { best = r; closest = (int)i; }
    }
    if (closest >= 0 && best < min_rmsd) { if (e[t] < e[kept[closest]]) kept[closest] = t; }
    else if ((int)kept.size() < n) kept.push_back(t);
    else if (!kept.empty() && e[t] < e[kept.back()]) kept.back() = t;
    std::stable_sort(kept.begin(), kept.end(), by_e);
  }
  std::vector<Mode> modes;
  for (int k : kept) {
    if ((int)modes.size() >= num_modes || !not_max(e[k]) || e[k] > e[kept[0]] + energy_range) break;
    Mode m;
    m.pose = k; m.e = e[k];
    rmsd_bounds(L, coords + (size_t)k * A * 3, coords + (size_t)kept[0] * A * 3, m.lb, m.ub);
    modes.push_back(m);
  }
  return modes;
}

// write_all_output / model::write_model (V11/main/main.cpp:89-102, model.h:201-205): per mode
// "MODEL i", the REMARK VINA RESULT line (vina_remark, main.cpp:158-170), every input line with
// the atom lines' coordinates replaced in columns 31-54 (coords_to_pdbqt_string and
// string_write_coord, V11/lib/model.cpp:1063-1094), "ENDMDL".
std::string write_models(const Ligand& L, int n, const double* e, const double* lb, const double* ub,
                         const float* coords) {
  const int A = (int)L.atoms.size();
  std::vector<int> line_atom(L.context.size(), -1);
  for (int i = 0; i < A; ++i)
    if (L.atoms[i].ctx_line >= 0 && L.atoms[i].ctx_line < (int)line_atom.size()) line_atom[L.atoms[i].ctx_line] = i;
  std::string out;
  char buf[128];
  for (int m = 0; m < n; ++m) {
    std::snprintf(buf, sizeof buf, "MODEL %d\n", m + 1);
    out += buf;
    std::snprintf(buf, sizeof buf, "REMARK VINA RESULT: %9.1f  %9.3f  %9.3f\n", e[m], lb[m], ub[m]);
    out += buf;
    for (size_t li = 0; li < L.context.size(); ++li) {
      std::string s = L.context[li];
      const int a = line_atom[li];
      if (a >= 0) {
        if (s.size() <= 30 + 16 + 8) throw parse_error("atom line too short to hold coordinates: " + s);
        for (int k = 0; k < 3; ++k) {
          const int len = std::snprintf(buf, sizeof buf, "%8.3f", (double)coords[((size_t)m * A + a) * 3 + k]);
          if (len != 8) throw parse_error("coordinate does not fit the 8-column PDBQT field");
          s.replace(30 + 8 * k, 8, buf, 8);
        }
      }
      out += s;
      out += '\n';
    }
    out += "ENDMDL\n";
  }
  return out;
}

}  // namespace vd
