// vdock_dock -- a C++ consumer of include/vdock.h: the reference's virtual-screening run
// (--ligand_directory, V11/main/main.cpp:898-926) with main_procedure_cl
// (V11/lib/main_procedure_cu.cpp:131-1006) and the per-ligand block of main_procedure after it
// (V11/main/main.cpp:423-501) on libvdock, as a maintainer of the reference would call it.
//
//   vdock_dock --receptor R.pdbqt --ligand_directory DIR --output_directory OUT
//              --center_x X --center_y Y --center_z Z --size_x SX --size_y SY --size_z SZ
//              --thread N --seed S [--num_modes 9] [--energy_range 3] [--rilc_bfgs 1]
//
// Each ligand's chains are first reduced to its pose list the way cl_to_vina and
// add_to_output_container do (V11/lib/main_procedure_cu.cpp:84-124, 921-926): the kept chains'
// conformations, energies and heavy-atom coordinates in an output container per ligand.  The
// post-processing call then refines those poses on the device, and the modes are written as
// <ligand>_out.pdbqt in the output directory.
//
// Exit status: 0 done, 2 bad options, 3 no usable GPU, 4 any other library error.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "vdock.h"

namespace {

// One pose of the output container (output_type, V11/lib/conf.h:447-458): the conformation
// (position, orientation quaternion w,x,y,z, torsions), its energy and heavy-atom coordinates.
struct OutputType {
  std::vector<double> conf;
  double e = 0;
  std::vector<std::array<double, 3>> coords;
};
using OutputContainer = std::vector<OutputType>;

struct Failure : std::runtime_error {
  int status;
  Failure(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void check(vd_ctx* ctx, int rc, const char* what) {
  if (rc >= 0) return;
  const std::string msg = std::string(what) + ": " + (ctx ? vd_last_error(ctx) : "no context");
  throw Failure(rc == VD_ERR_NO_DEVICE ? 3 : 4, msg);
}

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Failure(2, "cannot read " + path);
  std::stringstream s;
  s << f.rdbuf();
  return s.str();
}

// cl_to_vina + add_to_output_container: the chains in order, kept when no kept pose lies within
// min_rmsd (heavy atoms) or, when one does, in its place if better (vd_cluster), at most max_size.
OutputContainer pose_list(vd_ctx* ctx, int k, int id, double min_rmsd, int max_size) {
  vd_ligand_info inf;
  check(ctx, vd_ligand_info_get(ctx, id, &inf), "vd_ligand_info_get");
  const int chains = vd_batch_results(ctx, k, 0, nullptr, nullptr, nullptr);
  check(ctx, chains, "vd_batch_results");
  const int nc = 7 + inf.n_torsions;
  std::vector<float> e(chains), conf((size_t)chains * nc), heavy((size_t)chains * inf.n_heavy * 3);
  check(ctx, vd_batch_results(ctx, k, chains, e.data(), conf.data(), heavy.data()), "vd_batch_results");
  std::vector<int> kept(max_size);
  const int n = vd_cluster(chains, inf.n_heavy, e.data(), heavy.data(), min_rmsd, max_size, kept.data());
  check(ctx, n, "vd_cluster");
  OutputContainer out;
  for (int i = 0; i < n; ++i) {
    const int c = kept[i];
    OutputType o;
    o.conf.assign(conf.begin() + (size_t)c * nc, conf.begin() + (size_t)(c + 1) * nc);
    o.e = e[c];
    for (int a = 0; a < inf.n_heavy; ++a) {
      const float* x = &heavy[((size_t)c * inf.n_heavy + a) * 3];
      o.coords.push_back({x[0], x[1], x[2]});
    }
    out.push_back(std::move(o));
  }
  return out;
}

int run(const std::map<std::string, std::string>& opt) {
  auto need = [&](const char* k) {
    auto it = opt.find(k);
    if (it == opt.end()) throw Failure(2, std::string("missing --") + k);
    return it->second;
  };
  auto num = [&](const char* k) { return std::strtod(need(k).c_str(), nullptr); };
  auto get = [&](const char* k, const char* dflt) { auto it = opt.find(k); return it == opt.end() ? std::string(dflt) : it->second; };
  const std::string rec = need("receptor"), dir = need("ligand_directory"), out_dir = need("output_directory");
  const double center[3] = {num("center_x"), num("center_y"), num("center_z")};
  const double size[3] = {num("size_x"), num("size_y"), num("size_z")};
  const int thread = std::atoi(need("thread").c_str());
  const long seed = std::atol(need("seed").c_str());
  const int num_modes = std::atoi(get("num_modes", "9").c_str());
  const double energy_range = std::strtod(get("energy_range", "3").c_str(), nullptr);
  const int rilc = std::atoi(get("rilc_bfgs", "1").c_str());
  if (thread < 1000) throw Failure(2, "thread must be at least 1000");
  if (num_modes < 1) throw Failure(2, "num_modes must be 1 or greater");

  // the directory's .pdbqt entries in sorted order (main.cpp:898-926)
  std::vector<std::string> names;
  if (DIR* d = opendir(dir.c_str())) {
    while (dirent* ent = readdir(d)) {
      const std::string n = ent->d_name;
      if (n.size() > 6 && n.compare(n.size() - 6, 6, ".pdbqt") == 0) names.push_back(n);
    }
    closedir(d);
  } else {
    throw Failure(2, "cannot open the ligand directory " + dir);
  }
  std::sort(names.begin(), names.end());
  mkdir(out_dir.c_str(), 0755);

  vd_ctx* ctx = nullptr;
  const int rc = vd_create(0, &ctx);
  if (rc == VD_ERR_NO_DEVICE) throw Failure(3, "no usable GPU");
  check(nullptr, rc, "vd_create");
  struct Guard { vd_ctx* c; ~Guard() { vd_destroy(c); } } guard{ctx};
  const double w[6] = {-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 0.05846};
  check(ctx, vd_set_weights(ctx, w), "vd_set_weights");
  int n_rec = 0;
  check(ctx, vd_set_receptor_pdbqt(ctx, read_file(rec).c_str(), &n_rec), "vd_set_receptor_pdbqt");
  vd_box box;
  check(ctx, vd_box_from_center(center, size, 0.375, &box), "vd_box_from_center");
  check(ctx, vd_grid_populate(ctx, &box), "vd_grid_populate");

  std::vector<int> ids;
  std::vector<std::string> kept_names;
  for (const std::string& n : names) {
    int id = -1;
    const int r = vd_ligand_add_pdbqt(ctx, read_file(dir + "/" + n).c_str(), &id);
    if (r == VD_ERR_PARSE) { std::printf("Parse error in file \"%s\": %s\n", n.c_str(), vd_last_error(ctx)); continue; }
    check(ctx, r, "vd_ligand_add_pdbqt");
    ids.push_back(id);
    kept_names.push_back(n);
  }
  if (ids.empty()) throw Failure(2, "No valid ligands in the input directory");

  vd_dock_params prm;
  vd_dock_params_default(&prm);
  prm.chains = thread;
  prm.seed = (uint32_t)seed;
  prm.local_search = rilc == 1 ? VD_LS_RILC_BFGS : VD_LS_BFGS;
  vd_stats st;
  check(ctx, vd_dock_batch(ctx, (int)ids.size(), ids.data(), &prm, &st), "vd_dock_batch");
  std::printf("docked %zu ligands x %d chains: %.2f ms, %llu evaluations\n", ids.size(), thread, st.kernel_ms,
              (unsigned long long)st.evals);

  // main_procedure_cl's output: one container per ligand (num_saved_mins 20, min_rmsd 1)
  std::vector<OutputContainer> outs;
  for (size_t k = 0; k < ids.size(); ++k) {
    outs.push_back(pose_list(ctx, (int)k, ids[k], 1.0, 20));
    std::printf("pose list %s %zu %.6f\n", kept_names[k].c_str(), outs.back().size(), outs.back().empty() ? 0.0 : outs.back()[0].e);
  }

  // the step after the search, every ligand in one launch
  vd_post_params post;
  vd_post_params_default(&post);
  post.num_modes = num_modes;
  post.energy_range = (float)energy_range;
  post.local_search = prm.local_search;
  vd_stats pst;
  check(ctx, vd_postprocess_docked(ctx, 20, &post, &pst), "vd_postprocess_docked");
  for (size_t k = 0; k < ids.size(); ++k) {
    vd_ligand_info inf;
    check(ctx, vd_ligand_info_get(ctx, ids[k], &inf), "vd_ligand_info_get");
    const int n = vd_post_results(ctx, (int)k, 0, nullptr, nullptr, nullptr);
    check(ctx, n, "vd_post_results");
    std::vector<vd_mode> modes(std::max(n, 1));
    std::vector<float> conf((size_t)std::max(n, 1) * (7 + inf.n_torsions)), xyz((size_t)std::max(n, 1) * inf.n_atoms * 3);
    check(ctx, vd_post_results(ctx, (int)k, n, modes.data(), conf.data(), xyz.data()), "vd_post_results");
    std::printf("Refining ligand %s results...\nmode |   affinity | dist from best mode\n"
                "     | (kcal/mol) | rmsd l.b.| rmsd u.b.\n-----+------------+----------+----------\n",
                kept_names[k].c_str());
    for (int i = 0; i < n; ++i)
      std::printf("%4d    %9.1f  %9.3f  %9.3f\n", i + 1, modes[i].e, modes[i].rmsd_lb, modes[i].rmsd_ub);
    const int len = vd_write_pdbqt(ctx, ids[k], n, modes.data(), xyz.data(), nullptr, 0);
    check(ctx, len, "vd_write_pdbqt");
    std::string text((size_t)len + 1, '\0');
    check(ctx, vd_write_pdbqt(ctx, ids[k], n, modes.data(), xyz.data(), &text[0], len + 1), "vd_write_pdbqt");
    text.resize(len);
    const std::string& nm = kept_names[k];
    std::ofstream(out_dir + "/" + nm.substr(0, nm.size() - 6) + "_out.pdbqt", std::ios::binary) << text;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, std::string> opt;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a.rfind("--", 0) != 0 || i + 1 >= argc) {
      std::fprintf(stderr, "usage: %s --receptor R --ligand_directory D --output_directory O --center_x X ... --thread N --seed S\n", argv[0]);
      return 2;
    }
    opt[a.substr(2)] = argv[++i];
  }
  try {
    return run(opt);
  } catch (const Failure& f) {
    std::fprintf(stderr, "vdock_dock: %s\n", f.what());
    return f.status;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "vdock_dock: %s\n", e.what());
    return 4;
  }
}
