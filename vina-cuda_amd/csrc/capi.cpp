// capi.cpp -- the C ABI of include/vdock.h on top of host_model and the HIP kernels.
#include "../../include/vdock.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <tuple>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "host_model.h"
#include "kernels.h"

namespace {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};
struct ArgError : std::runtime_error {
  int code;
  ArgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define HIP_CHECK(x)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (p) HIP_CHECK(hipFree(p));
    p = nullptr; n = 0;
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T)));
    n = std::max<size_t>(count, 1);
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Page-locked host staging for large device-to-host copies (hipHostMalloc), grown on demand and
// kept: copies into it run at the link's rate and leave the calling thread free until the wait.
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (p) HIP_CHECK(hipHostFree(p));
    p = nullptr; n = 0;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault));
    n = std::max<size_t>(count, 1);
  }
  void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
};

struct BatchState {   // results of the last batch, kept on the device
  std::vector<int> ids;
  std::vector<longlong4> off;      // per slot: e, conf, heavy offsets
  longlong4 total{};               // element counts of e, conf, heavy
  std::vector<int> T, H, chains;
  DevBuf<float> e, conf, heavy;
};

struct Packed {
  std::vector<float> blob;
  std::vector<vd_lig_header> headers;
  int Amax = 0, Nmax = 0, Tmax = 0;
};

// Device buffers, stream and events of one MC launch, and the host side of a launch in flight
// (the sources of its asynchronous copies stay alive until it is collected).
struct McSlot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf<float4> blob;
  DevBuf<vd_lig_header> headers;
  DevBuf<uint32_t> streams;
  DevBuf<int2> items;
  DevBuf<longlong4> offs;
  DevBuf<int> queue;
  DevBuf<unsigned long long> evals;
  DevBuf<float> hist;                        // QuickVina2 / QuickVina-W per-chain histories
  DevBuf<float> glob;                        // QuickVina-W global buffers
  DevBuf<long long> glob_off;
  DevBuf<int> glob_axis, glob_T;
  Packed pk;
  std::vector<vd_lig_header> h_headers;
  std::vector<longlong4> h_offs;
  std::vector<uint32_t> h_streams;
  std::vector<int2> h_items;
  std::vector<long long> h_goff;
  std::vector<int> h_gaxis, h_gT;
  std::vector<unsigned long long> h_evals;
  McArgs a{};
  int G = 0, W = 0, blocks = 0;
  bool persistent = false, timed = false, busy = false;
  vd_launch rec{};
  void release() {
    blob.release(); headers.release(); streams.release(); items.release(); offs.release(); queue.release();
    evals.release(); hist.release(); glob.release(); glob_off.release(); glob_axis.release(); glob_T.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    ev0 = ev1 = nullptr; stream = nullptr;   // the streams are the process's (device_stream)
  }
};
constexpr int kMcSlots = 4;   // at most four hardware queues per process on these devices

// A persistent batch captured as a hipGraph (vd_set_graphs): every stage of one vd_dock_batch
// call -- the uploads of the ligand records and launch tables from pinned staging, the clears of
// the work-queue counter and the evaluation counts (fill kernels, so the graph holds no memset
// node), the QuickVina-W buffer clear, the MC launch of every class (two branches when the classes
// overlap) and the read-back of the evaluation counts -- replayed when the same ligands are docked
// again with the same parameters.  The seed may differ between replays: it is a field of the MC
// kernels' argument, updated in the instantiated graph's kernel nodes.
struct GraphClass {
  int slot = 0;                       // launch slot whose device buffers the graph uses
  McArgs a{};                         // the kernel argument as captured; P.seed follows the replays
  hipGraphNode_t node = nullptr;      // its kernel node (null: not found, seed changes re-capture)
  hipKernelNodeParams kp{};
  void* kargs[1] = {nullptr};
  vd_launch rec{};
  std::vector<vd_lig_header> headers;
  size_t evals_at = 0;                // first evaluation count of its ligands in GraphEntry::evals
  std::vector<const void*> ptrs;      // the slot's device buffers at capture time
};
struct GraphEntry {
  std::vector<int> ids;
  vd_dock_params p{};                 // seed zeroed
  std::vector<std::pair<int, int>> shapes;   // (lanes per chain, waves per SIMD) per ligand
  bool overlap = false;
  std::vector<const void*> outs;     // the batch's result buffers at capture time
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t seed = 0;                  // the seed the instantiated graph holds
  std::vector<GraphClass> cls;
  PinnedBuf<unsigned char> staging;   // every host-to-device copy's source
  PinnedBuf<unsigned long long> evals;
  GraphEntry() = default;
  GraphEntry(const GraphEntry&) = delete;
  GraphEntry& operator=(const GraphEntry&) = delete;
  ~GraphEntry() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    staging.release(); evals.release();
  }
};
constexpr size_t kMaxGraphs = 4;      // captured batches kept per context, least recently used dropped

struct PostLigand {   // modes of one ligand after refinement and selection
  int id = 0;
  std::vector<vd_mode> modes;
  std::vector<float> conf, coords;   // per mode: 7+T floats, 3A floats
};
struct PostState { std::vector<PostLigand> ligs; };

}  // namespace

struct vd_ctx {
  int device = 0;
  bool host_only = false;      // device == -1: host model only, no HIP calls
  hipStream_t stream = nullptr;
  std::string err;
  double weights[6] = {-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 0.05846};
  vd::Precalculated pre;
  DevBuf<float> lut_rows, lutp;
  std::vector<double> rec_xyz;   // receptor coordinates as parsed (FP64)
  std::vector<int> rec_xs;
  // maps
  bool maps_ready = false;
  vd_box box{};
  vd_grid_desc gd{};
  DevBuf<float> maps;
  DevBuf<float4> corners;   // the maps as the searches read them: eight corners per cell
  double corners_ms = 0.0;
  DevBuf<float4> grid_atoms;
  DevBuf<int2> grid_cells;
  double grid_ms = 0, eval_ms = 0;
  // ligands
  std::vector<vd::Ligand> ligs;
  // MC launches: slot 0 runs on `stream`, slots 1.. on their own streams (created on first use)
  McSlot mc[kMcSlots];
  bool overlap_classes = true;               // batch classes run at the same time
  hipEvent_t ev_base = nullptr;
  bool graphs_on = false;                    // persistent batches run as captured graphs (vd_set_graphs)
  hipStream_t cap[2] = {nullptr, nullptr};   // this context's capture-only streams (nothing runs on them)
  std::vector<std::unique_ptr<GraphEntry>> graphs;   // least recently used first
  DevBuf<float4> blob;                       // ligand record of vd_eval
  DevBuf<vd_lig_header> headers;
  DevBuf<float> io_a, io_b, io_c;
  BatchState batch;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int n_cu = 0;
  int lds_max = 64 * 1024;                   // LDS one workgroup may take (queried from the device)
  unsigned char* scratch = nullptr;          // this device's scratch array (vd_dock_noalloc)
  DevBuf<float> init;                        // caller's starting conformations (vd_dock, vd_local_search)
  std::map<int, std::pair<int, int>> tuned;  // size bucket -> (lanes per chain, waves per SIMD)
  bool autotune_on = false;                   // timing-based choice of the pair above (vd_set_autotune)
  DevBuf<float> tune_e, tune_conf, tune_heavy;
  std::vector<vd_launch> launches;           // MC launches of the last docking call
  // after the search: receptor cells of the direct evaluation (built on first use per box)
  bool nc_ready = false;
  vdr::NcD nck{};
  DevBuf<double4> nc_atoms;
  DevBuf<int> nc_start;
  DevBuf<double> lutd;                       // FP64 pair table quads of the refinement
  DevBuf<double4> ref_geo;
  DevBuf<uint32_t> ref_pairs;
  DevBuf<vdr::RefHeader> ref_hdr;
  DevBuf<double> ref_in, ref_out, ref_xyz;
  DevBuf<double4> ref_res;
  DevBuf<int> ref_lig, ref_evals;
  DevBuf<long long> ref_coff, ref_xoff;
  PostState post;
  PinnedBuf<float> pin_e, pin_conf, pin_heavy;   // per-chain results of a batch, for pose lists
  PinnedBuf<double> pin_cout, pin_xout;          // refinement outputs
  PinnedBuf<double4> pin_res;
};

namespace {

// The device address of each device's scratch array, looked up once per device (with that device
// current) and kept for the process: every context on a device shares its array.
std::mutex g_scratch_mutex;
std::map<int, unsigned char*> g_scratch;
unsigned char* scratch_for_current_device(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end()) return it->second;
  void* p = nullptr;
  HIP_CHECK(vdl::scratch_address(&p));
  if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
  g_scratch[device] = static_cast<unsigned char*>(p);
  return g_scratch[device];
}

// The library's streams, one set of kMcSlots per device for the whole process, created on first
// use and kept until the process ends.  A process's streams are served by at most four hardware
// queues; two streams that share one run their kernels one after the other, so the two launches of
// an overlapped batch would serialise on a context created after others.  Every context on a
// device therefore uses the same streams: slot 0 carries all of a context's work, slots 1.. the
// concurrent launches of a batch.
std::mutex g_stream_mutex;
std::map<int, std::vector<hipStream_t>> g_streams;
hipStream_t device_stream(int device, int k) {
  std::lock_guard<std::mutex> lock(g_stream_mutex);
  std::vector<hipStream_t>& v = g_streams[device];
  if (v.empty()) v.assign(kMcSlots, nullptr);
  if (!v[k]) HIP_CHECK(hipStreamCreate(&v[k]));
  return v[k];
}

int guard(vd_ctx* ctx, const std::function<void()>& f) {
  try {
    f();
    return VD_OK;
  } catch (ArgError& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (vd::parse_error& e) {
    if (ctx) ctx->err = e.what();
    return VD_ERR_PARSE;
  } catch (HipError& e) {
    if (ctx) ctx->err = e.what();
    return VD_ERR_HIP;
  } catch (std::exception& e) {
    if (ctx) ctx->err = e.what();
    return VD_ERR_ARG;
  }
}

void upload_lut(vd_ctx* c) {
  c->pre.build(c->weights);
  if (c->host_only) return;
  std::vector<float> rows, bins;
  c->pre.grid_rows(rows);
  c->pre.pair_bins(bins);
  c->lut_rows.ensure(rows.size());
  c->lutp.ensure(bins.size());
  HIP_CHECK(hipMemcpy(c->lut_rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->lutp.p, bins.data(), bins.size() * 4, hipMemcpyHostToDevice));
  // FP64 quads (e_i, dor_i, e_i+1, dor_i+1) for the refinement after the search
  std::vector<double> qd((size_t)VD_NUM_TYPE_PAIRS * VD_LUT_BINS * 4);
  for (int tp = 0; tp < VD_NUM_TYPE_PAIRS; ++tp)
    for (int b = 0; b < VD_LUT_BINS; ++b) {
      const size_t o = (size_t)tp * VD_LUT_BINS, b1 = std::min(b + 1, VD_LUT_BINS - 1);
      double* q = &qd[(o + b) * 4];
      q[0] = c->pre.smooth_e[o + b]; q[1] = c->pre.smooth_dor[o + b]; q[2] = c->pre.smooth_e[o + b1]; q[3] = c->pre.smooth_dor[o + b1];
    }
  c->lutd.ensure(qd.size());
  HIP_CHECK(hipMemcpy(c->lutd.p, qd.data(), qd.size() * 8, hipMemcpyHostToDevice));
}

vd_grid_desc make_desc(const vd_box& b) {
  // grid::init (V11/lib/grid.cpp:29-43), computed in double and stored in float
  vd_grid_desc g{};
  for (int i = 0; i < 3; ++i) {
    const int dim = b.n[i] + 1;
    const double range = b.end[i] - b.begin[i];
    const double dfm1 = dim - 1.0, factor = dfm1 / range;
    g.init[i] = (float)b.begin[i];
    g.factor[i] = (float)factor;
    g.factor_inv[i] = (float)(1.0 / factor);
    g.dim_fl_minus_1[i] = (float)dfm1;
    g.dim[i] = dim;
  }
  g.plane = (long long)g.dim[0] * g.dim[1] * g.dim[2];
  return g;
}

void check_box(const vd_box* b) {
  if (!b) throw ArgError(VD_ERR_ARG, "box is NULL");
  for (int i = 0; i < 3; ++i) {
    if (b->n[i] < 1 || !(b->end[i] > b->begin[i])) throw ArgError(VD_ERR_ARG, "empty search box");
    if (b->n[i] > 4096) throw ArgError(VD_ERR_CAPACITY, "search box too large");
  }
}

void need_device(vd_ctx* c) {
  if (c->host_only) throw ArgError(VD_ERR_NO_DEVICE, "host-only context (created with device -1)");
  HIP_CHECK(hipSetDevice(c->device));
}

const vd::Ligand& lig_at(vd_ctx* c, int id) {
  if (id < 0 || id >= (int)c->ligs.size()) throw ArgError(VD_ERR_ARG, "unknown ligand id " + std::to_string(id));
  return c->ligs[id];
}

void check_params(const vd_dock_params* p) {
  if (!p) throw ArgError(VD_ERR_ARG, "params is NULL");
  if (p->chains < 1) throw ArgError(VD_ERR_ARG, "chains must be >= 1");
  if (p->search_depth < 0 || p->bfgs_steps < 0) throw ArgError(VD_ERR_ARG, "negative search_depth or bfgs_steps");
  if (p->local_search != VD_LS_BFGS && p->local_search != VD_LS_RILC_BFGS) throw ArgError(VD_ERR_ARG, "bad local_search");
  if (p->policy != VD_MC_VINA && p->policy != VD_MC_QVINA2 && p->policy != VD_MC_QVINAW) throw ArgError(VD_ERR_ARG, "bad policy");
  if (p->policy != VD_MC_VINA && p->local_search != VD_LS_RILC_BFGS)
    throw ArgError(VD_ERR_ARG, "the QuickVina2 and QuickVina-W policies run with RILC-BFGS only");
  if (p->chains_per_wave != 0 && p->chains_per_wave != 1 && p->chains_per_wave != 2 && p->chains_per_wave != 4)
    throw ArgError(VD_ERR_ARG, "chains_per_wave must be 0, 1, 2 or 4");
  if (p->waves_per_simd != 0 && (p->waves_per_simd < 2 || p->waves_per_simd > 4))
    throw ArgError(VD_ERR_ARG, "waves_per_simd must be 0, 2, 3 or 4");
}

// QuickVina-W global buffer levels (QVW/inc/cuda/visited.cuh get_n): the smallest multiple of 3,
// i, with chains x depth < 2^(i+1); (i - 3) / 3 halvings per axis, at least 0, at most 4.
int qvw_levels(long long chains, long long depth) {
  int i = 0;
  while (!(chains * depth < (2LL << i))) i += 3;
  return std::max(0, std::min(4, (i - 3) / 3));
}

int heur_depth(int A, int T, int rilc) {
  double s = (0.24 * A + 0.29 * T - 3.41) * (rilc ? 2.0 : 2.4);
  if (s < 1) s = 1;
  return (int)s;
}
int heur_bfgs(int A, int rilc) { return (unsigned)((rilc ? 30 : 25) + A) / 3; }

// Group width G (lanes per chain) that a ligand allows: torsion lanes 1..T and the six rigid
// lanes G-6..G-1 must not overlap, and every node needs a lane.
int min_group(const vd::Ligand& L) {
  const int T = L.n_torsions();
  if (T + 7 <= 16) return 16;
  if (T + 7 <= 32) return 32;
  if (T + 7 <= 64) return 64;
  return 0;
}
// Autotuner buckets: atoms in [0,24), [24,48), [48,80), [80,..) x torsions in the ranges that
// admit 4, 2 or 1 chains per wavefront.
int size_bucket(const vd::Ligand& L) {
  const int a = (int)L.atoms.size(), t = L.n_torsions();
  const int ab = a < 24 ? 0 : a < 48 ? 1 : a < 80 ? 2 : 3;
  const int tb = t + 7 <= 16 ? 0 : t + 7 <= 32 ? 1 : 2;
  return ab * 4 + tb;
}
// (lanes per chain, waves per SIMD) per size bucket when timing-based tuning is off: what the
// autotuner picks on MI355X (gfx950, 256 CUs) in batched screening (1,024 ligands of 10-60 heavy
// atoms, 256 chains each): 16 lanes for ligands under 80 atoms with at most 9 torsions, 64 lanes
// for the rest, 3 waves per SIMD.  Two widths keep any batch at two classes, which run at the
// same time.  A fixed choice makes a run depend only on its inputs and seed: a chain's FP32 sums
// take another order under another width or register budget, and its trajectory drifts from
// there, so choices taken from timings that flip between processes would give another result
// for the same seed.
const std::pair<int, int> kDefaultPick[16] = {
    {16, 3}, {64, 3}, {64, 3}, {64, 3},    // fewer than 24 atoms; torsions <= 9, <= 25, more
    {16, 3}, {64, 3}, {64, 3}, {64, 3},    // 24 to 47
    {16, 3}, {64, 3}, {64, 3}, {64, 3},    // 48 to 79
    {64, 3}, {64, 3}, {64, 3}, {64, 3}};   // 80 and more
std::pair<int, int> bucket_pick(vd_ctx* c, const vd::Ligand& L) {
  const int b = size_bucket(L);
  auto it = c->tuned.find(b);
  return it != c->tuned.end() ? it->second : kDefaultPick[b];
}
int pick_group(vd_ctx* c, const vd::Ligand& L, int cpw) {
  const int gmin = min_group(L);
  if (gmin == 0) throw ArgError(VD_ERR_CAPACITY, "ligand has too many torsions");
  int G = gmin;
  if (cpw == 1) G = 64;
  else if (cpw == 2) G = 32;
  else if (cpw == 4) G = 16;
  else G = bucket_pick(c, L).first;
  if (G < gmin) G = gmin;
  return G;
}

int pick_waves(vd_ctx* c, const vd::Ligand& L, int G) {
  const std::pair<int, int> pick = bucket_pick(c, L);
  return pick.first == G ? pick.second : 2;
}

vd_mc_consts make_consts(vd_ctx* c, const vd_dock_params* p, int depth, int bfgs) {
  vd_mc_consts m{};
  for (int i = 0; i < 3; ++i) { m.box_lo[i] = (float)c->box.begin[i]; m.box_hi[i] = (float)c->box.end[i]; }
  m.v_intra = p->v_intra; m.v_inter = p->v_inter; m.slope = p->slope;
  m.temperature = p->temperature; m.amplitude = p->amplitude;
  m.search_depth = depth; m.bfgs_steps = bfgs; m.local_search = p->local_search;
  m.mutation_quirk = p->mutation_quirk; m.policy = p->policy; m.seed = p->seed; m.chains = p->chains;
  return m;
}

vdk::EvalK make_evalk(vd_ctx* c, float v_intra, float v_inter, float slope) {
  vdk::EvalK k;
  k.corners = c->corners.p; k.gd = c->gd; k.lutp = reinterpret_cast<const float2*>(c->lutp.p);
  k.cd0 = c->gd.dim[0] - 1; k.cd01 = (c->gd.dim[0] - 1) * (c->gd.dim[1] - 1);
  k.cplane = (long long)k.cd01 * (c->gd.dim[2] - 1);
  k.v_intra = v_intra; k.v_inter = v_inter; k.slope = slope;
  return k;
}

// Ligand records for chain groups of G lanes.
Packed pack(vd_ctx* c, int G, const std::vector<int>& ids, const std::vector<int>* depth = nullptr,
            const std::vector<int>* bfgs = nullptr) {
  Packed pk;
  for (size_t k = 0; k < ids.size(); ++k) {
    const vd::Ligand& L = lig_at(c, ids[k]);
    pk.headers.push_back(vd::pack_ligand(L, pk.blob, G));
    if (depth) pk.headers.back().search_depth = (*depth)[k];
    if (bfgs) pk.headers.back().bfgs_steps = (*bfgs)[k];
    pk.Amax = std::max(pk.Amax, (int)L.atoms.size());
    pk.Nmax = std::max(pk.Nmax, (int)L.nodes.size());
    pk.Tmax = std::max(pk.Tmax, L.n_torsions());
  }
  return pk;
}

double elapsed_ms(vd_ctx* c) {
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  return ms;
}

// The corner records of the context's maps, timed on the context's stream.
void make_corners(vd_ctx* c) {
  c->corners.ensure((size_t)2 * VD_XS_SIZE * (c->gd.dim[0] - 1) * (c->gd.dim[1] - 1) * (c->gd.dim[2] - 1));
  HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  HIP_CHECK(vdl::launch_corners(c->maps.p, c->corners.p, c->gd, c->stream));
  HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->corners_ms = elapsed_ms(c);
}

// Algorithmic bytes of one evaluation in the MC kernel: the eight map corners of every typed
// atom (8 x 4 B) and one interpolation quad of the pair table per intramolecular pair (16 B).
double eval_bytes(const vd_lig_header& h) { return 32.0 * h.n_heavy + 16.0 * h.n_pairs; }

// Kernel mode of a launch: 0 dense BFGS, 1 RILC-BFGS, 2 with the QuickVina2 history, 3 QuickVina-W.
int mc_mode(const vd_mc_consts& P) {
  return P.local_search == 0 ? 0 : (P.policy == VD_MC_QVINA2 ? 2 : P.policy == VD_MC_QVINAW ? 3 : 1);
}

// A tighter register budget only pays when the LDS of a launch lets the extra waves in: the
// loosest budget (waves per SIMD) that reaches the residency W reaches.
int loosest_budget(int G, int W, int mode, int lds) {
  if (mode == 0) return 2;
  const int occ = vdl::max_resident_mc_blocks(G, W, mode, lds);
  for (int w = 2; w < W; ++w)
    if (vdl::max_resident_mc_blocks(G, w, mode, lds) >= occ) return w;
  return W;
}

// One MC launch in parts, so that the launches of different lanes-per-chain classes of a batch
// can run at the same time on their own streams: mc_prepare does the host work and sizes the
// slot's device buffers (allocation may synchronise the device, so every launch of a batch is
// prepared before any is enqueued); mc_enqueue_copies and mc_enqueue_kernel issue the copies and
// the kernel on the slot's stream; mc_collect waits for it and reads its evaluation counts.
void mc_prepare(vd_ctx* c, McSlot& s, int G, int W, Packed&& pk, std::vector<uint32_t>&& streams,
                std::vector<int2>&& items, std::vector<longlong4>&& off, const vd_mc_consts& P, float* d_e,
                float* d_conf, float* d_heavy, bool persistent, bool timed, double share = 1.0) {
  // a launch left in flight by a call that failed part-way still reads this slot's host copies
  if (s.busy) { (void)hipStreamSynchronize(s.stream); s.busy = false; }
  s.pk = std::move(pk); s.h_streams = std::move(streams); s.h_items = std::move(items);
  s.h_headers = s.pk.headers; s.h_offs = std::move(off);
  s.h_goff.clear(); s.h_gaxis.clear(); s.h_gT.clear();
  s.timed = timed;
  long long glob_floats = 0;
  if (P.policy == VD_MC_QVINAW) {
    // one global buffer per ligand slot, 2^levels cells per axis (levels from chains x depth)
    for (size_t k = 0; k < s.h_headers.size(); ++k) {
      const int axis = 1 << qvw_levels(P.chains, s.h_headers[k].search_depth);
      s.h_headers[k].reserved = axis;
      s.h_offs[k].w = glob_floats;
      s.h_goff.push_back(glob_floats); s.h_gaxis.push_back(axis); s.h_gT.push_back(s.h_headers[k].n_torsions);
      glob_floats += (vdk::GlobalHist::floats(axis, s.h_headers[k].n_torsions) + 63) / 64 * 64;
    }
    s.glob.ensure(glob_floats);
    s.glob_off.ensure(s.h_goff.size()); s.glob_axis.ensure(s.h_gaxis.size()); s.glob_T.ensure(s.h_gT.size());
  }
  s.blob.ensure(s.pk.blob.size() / 4 + 1);
  s.headers.ensure(s.h_headers.size());
  s.streams.ensure(s.h_streams.size());
  s.items.ensure(s.h_items.size());
  s.offs.ensure(s.h_offs.size());
  s.queue.ensure(1);
  s.evals.ensure(s.h_headers.size());
  McArgs& a = s.a;
  a = McArgs{};
  a.blob = s.blob.p; a.headers = s.headers.p; a.streams = s.streams.p; a.items = s.items.p;
  a.n_items = (int)s.h_items.size();
  a.K = make_evalk(c, P.v_intra, P.v_inter, P.slope);
  a.P = P;
  a.out_e = d_e; a.out_conf = d_conf; a.out_heavy = d_heavy; a.out_off = s.offs.p;
  a.slot_evals = s.evals.p;
  a.Amax = s.pk.Amax; a.Nmax = s.pk.Nmax;
  const int lds = vdl::mc_lds_bytes(G, s.pk.Amax, s.pk.Nmax, P.local_search == 0);
  if (lds > c->lds_max)
    throw ArgError(VD_ERR_CAPACITY, "ligand too large for the LDS layout (" + std::to_string(lds) + " bytes of " +
                                        std::to_string(c->lds_max) + " per workgroup)");
  const int mode = mc_mode(P);
  W = loosest_budget(G, W, mode, lds);
  int blocks = (int)s.h_items.size();
  if (persistent) {
    a.queue = s.queue.p;
    const int per_cu = std::max(1, (int)std::ceil(share * vdl::max_resident_mc_blocks(G, W, mode, lds)));
    blocks = std::min<long long>((long long)per_cu * c->n_cu, (long long)s.h_items.size());
    blocks = std::max(blocks, 1);
  }
  if (P.policy == VD_MC_QVINA2 || P.policy == VD_MC_QVINAW) {
    // one visited-history region per chain slot of every launched block (blocks keep their
    // slot for all the items they take from the queue)
    a.hist_stride = vdk::History::floats_per_chain(s.pk.Tmax, P.policy == VD_MC_QVINAW ? 5 : 10);
    s.hist.ensure((size_t)blocks * (64 / G) * a.hist_stride);
    a.hist = s.hist.p;
  }
  if (P.policy == VD_MC_QVINAW) a.glob = s.glob.p;
  s.G = G; s.W = W; s.blocks = blocks; s.persistent = persistent;
  s.rec = vd_launch{};
  s.rec.lanes_per_chain = G; s.rec.waves_per_simd = W; s.rec.blocks = blocks; s.rec.items = (int)s.h_items.size();
  s.rec.ligands = (int)s.h_headers.size(); s.rec.persistent = persistent ? 1 : 0;
  s.rec.mode = mode;
  s.h_evals.assign(s.h_headers.size(), 0);
}

void mc_enqueue_copies(McSlot& s) {
  hipStream_t st = s.stream;
  if (!s.h_goff.empty()) {
    HIP_CHECK(hipMemcpyAsync(s.glob_off.p, s.h_goff.data(), s.h_goff.size() * 8, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(s.glob_axis.p, s.h_gaxis.data(), s.h_gaxis.size() * 4, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(s.glob_T.p, s.h_gT.data(), s.h_gT.size() * 4, hipMemcpyHostToDevice, st));
    HIP_CHECK(vdl::glob_init(s.glob.p, s.glob_off.p, s.glob_axis.p, s.glob_T.p, (int)s.h_goff.size(), st));
  }
  HIP_CHECK(hipMemcpyAsync(s.blob.p, s.pk.blob.data(), s.pk.blob.size() * 4, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(s.headers.p, s.h_headers.data(), s.h_headers.size() * sizeof(vd_lig_header), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(s.streams.p, s.h_streams.data(), s.h_streams.size() * 4, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(s.items.p, s.h_items.data(), s.h_items.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(s.offs.p, s.h_offs.data(), s.h_offs.size() * sizeof(longlong4), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemsetAsync(s.evals.p, 0, s.h_headers.size() * sizeof(unsigned long long), st));
  if (s.persistent) HIP_CHECK(vdl::reset_counter(s.queue.p, st));
}

// The kernel.  Copies may run as kernels that need a
// CU slot, so every slot's copies are enqueued before any slot's kernel: a persistent grid
// already holding the CUs would otherwise hold back the next slot's copies until it ends.
void mc_enqueue_kernel(McSlot& s) {
  hipStream_t st = s.stream;
  if (s.timed) HIP_CHECK(hipEventRecord(s.ev0, st));
  HIP_CHECK(vdl::launch_mc(s.G, s.W, s.a, s.blocks, st));
  if (s.timed) HIP_CHECK(hipEventRecord(s.ev1, st));
  s.busy = true;
}

// base: an event recorded before every launch of the call (start_ms is measured from it), or null.
void mc_collect(vd_ctx* c, McSlot& s, vd_stats* st, hipEvent_t base, bool add_ms = true) {
  if (!s.busy) return;
  s.busy = false;
  // read back after the wait: a copy into pageable host memory blocks the calling thread until
  // it is done, which would hold back the other slots' kernels
  HIP_CHECK(hipStreamSynchronize(s.stream));
  HIP_CHECK(hipMemcpyAsync(s.h_evals.data(), s.evals.p, s.h_evals.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s.stream));
  HIP_CHECK(hipStreamSynchronize(s.stream));
  vd_launch& rec = s.rec;
  if (s.timed) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    rec.kernel_ms = ms;
    if (base) { HIP_CHECK(hipEventElapsedTime(&ms, base, s.ev0)); rec.start_ms = ms; }
  }
  for (size_t k = 0; k < s.h_evals.size(); ++k) {
    rec.evals += s.h_evals[k];
    rec.eval_bytes += (double)s.h_evals[k] * eval_bytes(s.h_headers[k]);
  }
  if (s.timed) c->launches.push_back(rec);
  if (st) {
    st->evals += rec.evals;
    st->blocks += s.blocks;
    st->chains_per_wave = 64 / s.G;
    if (add_ms) st->kernel_ms += rec.kernel_ms;
  }
}

// Runs one MC launch over the given ligand slots (all with the same G) and work items, on the
// context's stream, and waits for it.
void run_mc(vd_ctx* c, int G, int W, const Packed& pk, const std::vector<uint32_t>& streams, const std::vector<int2>& items,
            const std::vector<longlong4>& off, const vd_mc_consts& P, float* d_e, float* d_conf, float* d_heavy,
            bool persistent, vd_stats* st, bool timed, const float* d_init = nullptr, bool ls_only = false) {
  McSlot& s = c->mc[0];
  mc_prepare(c, s, G, W, Packed(pk), std::vector<uint32_t>(streams), std::vector<int2>(items),
             std::vector<longlong4>(off), P, d_e, d_conf, d_heavy, persistent, timed);
  s.a.init_conf = d_init;
  s.a.ls_only = ls_only ? 1 : 0;
  mc_enqueue_copies(s);
  mc_enqueue_kernel(s);
  mc_collect(c, s, st, nullptr);
}

struct LigPlan { int id, G, W, depth, bfgs, T, H, chains; };
// single_launch: the ligand gets a launch of its own (vd_dock, the plain per-ligand loop).  Its
// only parallelism is its chains, so unless they fill two waves per SIMD at the tuned width, it
// runs one chain per wavefront, the most wavefronts such a launch can have.
LigPlan plan_for(vd_ctx* c, int id, const vd_dock_params* p, bool single_launch) {
  const vd::Ligand& L = lig_at(c, id);
  LigPlan pl;
  pl.id = id; pl.T = L.n_torsions(); pl.H = L.n_heavy; pl.chains = p->chains;
  pl.G = pick_group(c, L, p->chains_per_wave);
  if (single_launch && p->chains_per_wave == 0 && (long long)p->chains * pl.G / 64 < 8LL * std::max(c->n_cu, 1)) pl.G = 64;
  pl.W = p->waves_per_simd ? p->waves_per_simd : pick_waves(c, L, pl.G);
  if (p->local_search != VD_LS_RILC_BFGS) pl.W = 2;
  const int rilc = p->local_search == VD_LS_RILC_BFGS;
  pl.depth = p->search_depth > 0 ? p->search_depth : heur_depth((int)L.atoms.size(), pl.T, rilc);
  pl.bfgs = p->bfgs_steps > 0 ? p->bfgs_steps : heur_bfgs((int)L.atoms.size(), rilc);
  pl.bfgs = std::max(pl.bfgs, 1);
  return pl;
}

// Autotuner: for every size bucket of the batch not tuned yet, a representative ligand docks the
// same fixed work -- 64 chains per CU (enough to fill every wavefront slot at four chains per
// wavefront and four wavefronts per SIMD), at most six MC steps -- under each admissible
// (lanes per chain, waves per SIMD) pair, in a persistent launch whose LDS layout is sized for the
// whole batch as the batched launch's is.  The pairs within 10 % of the fastest run again and
// keep their shorter time; among those within 2 % of the fastest, the first in a fixed preference
// order wins (wider groups first, then 3, 4, 2 waves per SIMD), so that timing noise does not flip
// the choice between runs.  Every chain follows the same trajectory under every pair up to
// rounding (the random stream does not depend on the layout), so the work is the same.
void autotune(vd_ctx* c, const std::vector<int>& ids, const vd_dock_params* p) {
  std::map<int, int> rep;
  for (int id : ids) { const int b = size_bucket(lig_at(c, id)); if (!c->tuned.count(b) && !rep.count(b)) rep[b] = id; }
  if (rep.empty()) return;
  int Amax = 0, Nmax = 0;
  for (int id : ids) {
    const vd::Ligand& L = lig_at(c, id);
    Amax = std::max(Amax, (int)L.atoms.size()); Nmax = std::max(Nmax, (int)L.nodes.size());
  }
  const int rilc = p->local_search == VD_LS_RILC_BFGS;
  const int chains = std::max(1, c->n_cu) * 64;
  for (auto& kv : rep) {
    const vd::Ligand& L = lig_at(c, kv.second);
    const int gmin = min_group(L);
    const int bfgs = std::max(1, p->bfgs_steps > 0 ? p->bfgs_steps : heur_bfgs((int)L.atoms.size(), rilc));
    const int depth = std::min(6, p->search_depth > 0 ? p->search_depth
                                                      : heur_depth((int)L.atoms.size(), L.n_torsions(), rilc));
    c->tune_e.ensure(chains); c->tune_conf.ensure((size_t)chains * (7 + L.n_torsions()));
    c->tune_heavy.ensure((size_t)chains * L.n_heavy * 3);
    struct Trial { int G, W, W_run; double ms; };
    std::vector<Trial> trials;   // preference order
    for (int G = 64; G >= gmin; G /= 2)
      for (int W : {3, 4, 2})
        if (rilc || W == 2) trials.push_back({G, W, W, 1e300});
    auto run = [&](Trial& t) {
      vd_dock_params q = *p;
      q.chains = chains; q.chains_per_wave = 64 / t.G;
      const int cpw = 64 / t.G;
      std::vector<int> dv{depth}, bv{bfgs};
      Packed pk = pack(c, t.G, {kv.second}, &dv, &bv);
      pk.Amax = std::max(pk.Amax, Amax); pk.Nmax = std::max(pk.Nmax, Nmax);
      std::vector<int2> items;
      for (int c0 = 0; c0 < chains; c0 += cpw) items.push_back(make_int2(0, c0));
      std::vector<longlong4> off{make_longlong4(0, 0, 0, 0)};
      vd_stats st{};
      run_mc(c, t.G, t.W, pk, {(uint32_t)kv.second}, items, off, make_consts(c, &q, depth, bfgs), c->tune_e.p,
             c->tune_conf.p, c->tune_heavy.p, true, &st, true);
      t.ms = std::min(t.ms, st.kernel_ms);
      t.W_run = c->launches.back().waves_per_simd;   // the budget the launch actually took
    };
    for (Trial& t : trials) run(t);
    double best = 1e300;
    for (const Trial& t : trials) best = std::min(best, t.ms);
    for (Trial& t : trials) if (t.ms <= 1.10 * best) run(t);
    best = 1e300;
    for (const Trial& t : trials) best = std::min(best, t.ms);
    for (const Trial& t : trials)
      if (t.ms <= 1.02 * best) { c->tuned[kv.first] = std::make_pair(t.G, t.W_run); break; }
  }
}

// Prepares the batch result buffers for the ligands in `ids`.
void prepare_batch(vd_ctx* c, const std::vector<LigPlan>& plans) {
  BatchState& b = c->batch;
  b.ids.clear(); b.off.clear(); b.T.clear(); b.H.clear(); b.chains.clear();
  long long oe = 0, oc = 0, oh = 0;
  for (const LigPlan& pl : plans) {
    b.ids.push_back(pl.id); b.T.push_back(pl.T); b.H.push_back(pl.H); b.chains.push_back(pl.chains);
    b.off.push_back(make_longlong4(oe, oc, oh, 0));
    oe += pl.chains; oc += (long long)pl.chains * (7 + pl.T); oh += (long long)pl.chains * pl.H * 3;
  }
  b.e.ensure(oe); b.conf.ensure(oc); b.heavy.ensure(oh);
  b.total = make_longlong4(oe, oc, oh, 0);
}

// f(0..n-1) on up to 16 host threads (independent items); rethrows the first failure.
void parallel_for(int n, const std::function<void(int)>& f) {
  const int threads = std::min({n, 16, (int)std::max(1u, std::thread::hardware_concurrency())});
  if (threads <= 1) { for (int i = 0; i < n; ++i) f(i); return; }
  std::atomic<int> next{0};
  std::exception_ptr failure;
  std::mutex m;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (int i; (i = next++) < n;) {
        try { f(i); } catch (...) { std::lock_guard<std::mutex> g(m); if (!failure) failure = std::current_exception(); }
      }
    });
  for (std::thread& t : pool) t.join();
  if (failure) std::rethrow_exception(failure);
}

// The device pair layout of every (ligand, lanes per chain) of a batch, built on host threads
// before the packing: each layout is an edge colouring of the ligand's pair graph, about half a
// millisecond for a 40-atom ligand, cached on the ligand.  Each distinct pair once, so that no
// two threads fill one ligand's cache.
void prepare_pair_chunks(vd_ctx* c, const std::vector<LigPlan>& plans) {
  std::vector<std::pair<int, int>> todo;
  for (const LigPlan& pl : plans) todo.push_back(std::make_pair(pl.id, pl.G));
  std::sort(todo.begin(), todo.end());
  todo.erase(std::unique(todo.begin(), todo.end()), todo.end());
  parallel_for((int)todo.size(), [&](int k) { (void)vd::pair_chunks(lig_at(c, todo[k].first), todo[k].second); });
}

}  // namespace

namespace {

// Receptor cells of the direct evaluation after the search: the box is cut into cells of about
// 2 A (n per axis exactly, so the cells tile it), and each cell lists every typed receptor atom
// within the cutoff of its brick grown by 1e-3 A (a point the device's float arithmetic puts in
// the cell is covered even when it lies a rounding error outside), as FP64 coordinates and XS
// type.  Coarser cells when the lists would pass 2^26 entries (very large boxes).
void build_nc(vd_ctx* c) {
  if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no search box: call vd_grid_populate first");
  if (c->rec_xs.empty()) throw ArgError(VD_ERR_ARG, "no receptor atoms: call vd_set_receptor first");
  const vd_box& b = c->box;
  const double grow = std::sqrt((double)VD_CUTOFF_SQR) + 1e-3;
  std::vector<int> typed;
  for (size_t i = 0; i < c->rec_xs.size(); ++i) {
    if (c->rec_xs[i] >= VD_XS_SIZE) continue;
    bool near = true;
    for (int k = 0; k < 3; ++k) {
      const double x = c->rec_xyz[3 * i + k];
      if (x < b.begin[k] - grow || x > b.end[k] + grow) near = false;
    }
    if (near) typed.push_back((int)i);
  }
  for (double target = 2.0;; target *= 1.5) {
    int n[3]; double span[3], cs[3];
    for (int k = 0; k < 3; ++k) {
      span[k] = b.end[k] - b.begin[k];
      n[k] = std::max(1, (int)std::ceil(span[k] / target));
      cs[k] = span[k] / n[k];
    }
    // bins of the same size over the box grown by the cutoff speed up the brick tests
    int nb[3], reach[3];
    for (int k = 0; k < 3; ++k) {
      nb[k] = (int)std::ceil((span[k] + 2 * grow) / cs[k]) + 1;
      reach[k] = (int)std::ceil(2 * grow / cs[k]) + 1;
    }
    std::vector<std::vector<int>> bins((size_t)nb[0] * nb[1] * nb[2]);
    for (int i : typed) {
      int bi[3];
      for (int k = 0; k < 3; ++k)
        bi[k] = std::min(nb[k] - 1, std::max(0, (int)std::floor((c->rec_xyz[3 * i + k] - (b.begin[k] - grow)) / cs[k])));
      bins[((size_t)bi[2] * nb[1] + bi[1]) * nb[0] + bi[0]].push_back(i);
    }
    const long long ncell = (long long)n[0] * n[1] * n[2];
    std::vector<int> start(ncell + 1, 0);
    std::vector<double4> list;
    bool too_big = false;
    for (int z = 0; z < n[2] && !too_big; ++z)
      for (int y = 0; y < n[1] && !too_big; ++y)
        for (int x = 0; x < n[0]; ++x) {
          const int id[3] = {x, y, z};
          double lo[3], hi[3];
          for (int k = 0; k < 3; ++k) { lo[k] = b.begin[k] + cs[k] * id[k]; hi[k] = b.begin[k] + cs[k] * (id[k] + 1); }
          for (int bz = z; bz <= std::min(nb[2] - 1, z + reach[2]); ++bz)
            for (int by = y; by <= std::min(nb[1] - 1, y + reach[1]); ++by)
              for (int bx = x; bx <= std::min(nb[0] - 1, x + reach[0]); ++bx)
                for (int i : bins[((size_t)bz * nb[1] + by) * nb[0] + bx]) {
                  double dd = 0;
                  for (int k = 0; k < 3; ++k) {
                    const double v = c->rec_xyz[3 * i + k];
                    const double d = v < lo[k] ? lo[k] - v : (v > hi[k] ? v - hi[k] : 0.0);
                    dd += d * d;
                  }
                  if (dd >= grow * grow) continue;
                  list.push_back(make_double4(c->rec_xyz[3 * i], c->rec_xyz[3 * i + 1], c->rec_xyz[3 * i + 2], (double)c->rec_xs[i]));
                }
          const long long cell = ((long long)z * n[1] + y) * n[0] + x;
          start[cell + 1] = (int)list.size();
          if (list.size() > (size_t)(1 << 26)) { too_big = true; break; }
        }
    if (too_big) continue;
    c->nc_atoms.ensure(list.size());
    c->nc_start.ensure(start.size());
    if (!list.empty())
      HIP_CHECK(hipMemcpy(c->nc_atoms.p, list.data(), list.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->nc_start.p, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
    vdr::NcD& k = c->nck;
    k.atoms = c->nc_atoms.p; k.start = c->nc_start.p;
    for (int j = 0; j < 3; ++j) {
      k.lo[j] = b.begin[j]; k.hi[j] = b.end[j]; k.inv_cell[j] = n[j] / span[j]; k.n[j] = n[j];
    }
    c->nc_ready = true;
    return;
  }
}

void check_post(const vd_post_params* p) {
  if (!p) throw ArgError(VD_ERR_ARG, "post params is NULL");
  if (p->max_steps < 0) throw ArgError(VD_ERR_ARG, "negative max_steps");
  if (p->local_search != VD_LS_BFGS && p->local_search != VD_LS_RILC_BFGS) throw ArgError(VD_ERR_ARG, "bad local_search");
  if (p->num_modes < 1) throw ArgError(VD_ERR_ARG, "num_modes must be >= 1");
  if (!(p->min_rmsd >= 0) || !(p->energy_range >= 0) || !(p->v >= 0) || !(p->slope >= 0))
    throw ArgError(VD_ERR_ARG, "negative or NaN post-processing parameter");
}

// FP64 ligand record of the refinement (vdr::RefHeader): atom words (local xyz, node | xs << 8 |
// heavy << 16), node origin words (rel_origin, (parent + 1) | begin << 8 | end << 16), node axis
// words (rel_axis, subtree end | depth << 8 | (last child + 1) << 16 | (previous sibling + 1) << 24);
// pair words in the conflict-free chunks of 64.
vdr::RefHeader pack_ref(const vd::Ligand& L, std::vector<double4>& geo, std::vector<uint32_t>& pw, int max_steps) {
  vdr::RefHeader h{};
  h.A = (int)L.atoms.size(); h.N = (int)L.nodes.size(); h.T = L.n_torsions(); h.heavy = L.n_heavy;
  h.max_steps = max_steps;
  h.off_atoms = (long long)geo.size();
  for (const vd::LigAtom& a : L.atoms) {
    const uint32_t meta = (uint32_t)a.node | ((uint32_t)a.xs << 8) | ((a.el != 0 ? 1u : 0u) << 16);
    geo.push_back(make_double4(a.local[0], a.local[1], a.local[2], (double)meta));
  }
  for (const vd::LigNode& nd : L.nodes)
    geo.push_back(make_double4(nd.rel_origin[0], nd.rel_origin[1], nd.rel_origin[2],
                               (double)((uint32_t)(nd.parent + 1) | ((uint32_t)nd.begin << 8) | ((uint32_t)nd.end << 16))));
  // tree links of the level-by-level kinematics and adjoint fold (nodes are in preorder, so a
  // parent precedes its children)
  const int N = (int)L.nodes.size();
  std::vector<int> last_child(N, -1), prev_sibling(N, -1);
  for (int k = 1; k < N; ++k) {
    const int par = L.nodes[k].parent;
    prev_sibling[k] = last_child[par];
    last_child[par] = k;
  }
  for (int k = 0; k < N; ++k) {
    const vd::LigNode& nd = L.nodes[k];
    const uint32_t links = (uint32_t)nd.subtree_end | ((uint32_t)nd.depth << 8) | ((uint32_t)(last_child[k] + 1) << 16) |
                           ((uint32_t)(prev_sibling[k] + 1) << 24);
    geo.push_back(make_double4(nd.rel_axis[0], nd.rel_axis[1], nd.rel_axis[2], (double)links));
  }
  const std::vector<uint32_t>& ch = vd::pair_chunks(L, 64);
  h.C = (int)(ch.size() / 64);
  h.off_pairs = (long long)pw.size();
  pw.insert(pw.end(), ch.begin(), ch.end());
  return h;
}

// One refinement launch over poses of several ligands: ligand ids[k] owns poses[k] poses, their
// conformations concatenated in conf.  Fills per pose the refined conformation, all-atom
// coordinates and vd_refined.
struct RefineOut {
  std::vector<float> conf, coords;
  std::vector<vd_refined> res;
  std::vector<long long> coff, xoff;   // per pose offsets into conf / coords
  unsigned long long evals = 0;
  double ms = 0;
};
RefineOut run_refine(vd_ctx* c, const std::vector<int>& ids, const std::vector<int>& poses, const float* conf,
                     const vd_post_params* p) {
  check_post(p);
  need_device(c);
  if (!c->nc_ready) build_nc(c);
  std::vector<double4> geo;
  std::vector<uint32_t> pw;
  std::vector<vdr::RefHeader> hdr;
  int Amax = 1, Nmax = 1, nmax = 1;
  for (int id : ids) {
    const vd::Ligand& L = lig_at(c, id);
    if (L.n_torsions() + 7 > 64) throw ArgError(VD_ERR_CAPACITY, "ligand has too many torsions for the refinement");
    const int steps = p->max_steps > 0 ? p->max_steps : heur_bfgs((int)L.atoms.size(), p->local_search);
    hdr.push_back(pack_ref(L, geo, pw, steps));
    Amax = std::max(Amax, (int)L.atoms.size()); Nmax = std::max(Nmax, (int)L.nodes.size());
    nmax = std::max(nmax, 6 + L.n_torsions());
  }
  RefineOut r;
  std::vector<int> pose_lig;
  long long nc = 0, nx = 0;
  for (size_t k = 0; k < ids.size(); ++k) {
    const vd::Ligand& L = lig_at(c, ids[k]);
    if (poses[k] < 0) throw ArgError(VD_ERR_ARG, "negative pose count");
    for (int i = 0; i < poses[k]; ++i) {
      pose_lig.push_back((int)k);
      r.coff.push_back(nc); r.xoff.push_back(nx);
      nc += 7 + L.n_torsions(); nx += 3 * (long long)L.atoms.size();
    }
  }
  const int n = (int)pose_lig.size();
  r.conf.assign(nc, 0.0f); r.coords.assign(nx, 0.0f); r.res.assign(n, vd_refined{});
  if (n == 0) return r;
  std::vector<double> cin(nc);
  for (long long i = 0; i < nc; ++i) {
    if (!std::isfinite(conf[i])) throw ArgError(VD_ERR_ARG, "non-finite conformation value");
    cin[i] = conf[i];
  }
  const int lds = vdl::refine_lds_bytes(Amax, Nmax, nmax);
  if (lds > c->lds_max)
    throw ArgError(VD_ERR_CAPACITY, "ligand too large for the refinement's LDS layout (" + std::to_string(lds) +
                                        " bytes of " + std::to_string(c->lds_max) + " per workgroup)");
  if (pw.empty()) pw.push_back(VD_NO_PAIR);
  c->ref_geo.ensure(geo.size()); c->ref_pairs.ensure(pw.size()); c->ref_hdr.ensure(hdr.size());
  c->ref_lig.ensure(n); c->ref_evals.ensure(n); c->ref_res.ensure(n);
  c->ref_coff.ensure(n); c->ref_xoff.ensure(n);
  c->ref_in.ensure(nc); c->ref_out.ensure(nc); c->ref_xyz.ensure(nx);
  HIP_CHECK(hipMemcpyAsync(c->ref_geo.p, geo.data(), geo.size() * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_pairs.p, pw.data(), pw.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_hdr.p, hdr.data(), hdr.size() * sizeof(vdr::RefHeader), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_lig.p, pose_lig.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_coff.p, r.coff.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_xoff.p, r.xoff.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->ref_in.p, cin.data(), nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
  RefineArgs a{};
  a.geo = c->ref_geo.p; a.pairs = c->ref_pairs.p; a.hdr = c->ref_hdr.p; a.pose_lig = c->ref_lig.p;
  a.conf_off = c->ref_coff.p; a.coords_off = c->ref_xoff.p;
  a.conf_in = c->ref_in.p; a.conf_out = c->ref_out.p; a.coords_out = c->ref_xyz.p;
  a.res = c->ref_res.p; a.evals = c->ref_evals.p;
  a.K.lut = reinterpret_cast<const double4*>(c->lutd.p);
  a.K.nc = c->nck;
  a.K.v = p->v;
  a.final_slope = p->slope;
  a.Amax = Amax; a.Nmax = Nmax; a.nmax = nmax;
  HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  HIP_CHECK(vdl::launch_refine(a, n, c->stream));
  HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  std::vector<int> ev(n);
  c->pin_cout.ensure(nc); c->pin_xout.ensure(nx); c->pin_res.ensure(n);
  const double* cout_ = c->pin_cout.p;
  const double* xout = c->pin_xout.p;
  const double4* res = c->pin_res.p;
  HIP_CHECK(hipMemcpyAsync(c->pin_cout.p, c->ref_out.p, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->pin_xout.p, c->ref_xyz.p, nx * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->pin_res.p, c->ref_res.p, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(ev.data(), c->ref_evals.p, n * sizeof(int), hipMemcpyDeviceToHost));
  r.ms = elapsed_ms(c);
  for (long long i = 0; i < nc; ++i) r.conf[i] = (float)cout_[i];
  for (long long i = 0; i < nx; ++i) r.coords[i] = (float)xout[i];
  const double fmax = std::numeric_limits<float>::max();
  for (int i = 0; i < n; ++i) {
    r.res[i].e = res[i].x < fmax ? (float)res[i].x : std::numeric_limits<float>::max();
    r.res[i].e_inter = (float)std::min(res[i].y, fmax); r.res[i].e_intra = (float)std::min(res[i].z, fmax);
    r.res[i].inside = res[i].w != 0.0;
    r.evals += (unsigned long long)ev[i];
  }
  return r;
}

std::vector<vd_mode> select_modes(vd_ctx* c, const vd::Ligand& L, int n, const vd_refined* r, const float* coords,
                                  const vd_post_params* p) {
  std::vector<float> e(n), ei(n), ea(n);
  for (int i = 0; i < n; ++i) { e[i] = r[i].e; ei[i] = r[i].e_inter; ea[i] = r[i].e_intra; }
  std::vector<vd::Mode> m = vd::select_modes(L, n, e.data(), ei.data(), ea.data(), coords, c->weights[5], p->num_modes,
                                             p->energy_range, p->min_rmsd);
  std::vector<vd_mode> out(m.size());
  for (size_t i = 0; i < m.size(); ++i) {
    out[i].e = (float)m[i].e; out[i].rmsd_lb = (float)m[i].lb; out[i].rmsd_ub = (float)m[i].ub; out[i].pose = m[i].pose;
  }
  return out;
}

// Refine + select for many ligands; the results replace c->post.
void postprocess(vd_ctx* c, const std::vector<int>& ids, const std::vector<int>& poses, const float* conf,
                 const vd_post_params* p, vd_stats* st) {
  RefineOut r = run_refine(c, ids, poses, conf, p);
  PostState ps;
  ps.ligs.resize(ids.size());
  std::vector<size_t> first(ids.size() + 1, 0);
  for (size_t k = 0; k < ids.size(); ++k) first[k + 1] = first[k] + poses[k];
  parallel_for((int)ids.size(), [&](int k) {
    const vd::Ligand& L = c->ligs[ids[k]];
    const int A = (int)L.atoms.size(), T = L.n_torsions(), n = poses[k];
    PostLigand& pl = ps.ligs[k];
    pl.id = ids[k];
    const float* xyz = n > 0 ? r.coords.data() + r.xoff[first[k]] : nullptr;
    pl.modes = select_modes(c, L, n, r.res.data() + first[k], xyz, p);
    for (const vd_mode& m : pl.modes) {
      const size_t q = first[k] + m.pose;
      pl.conf.insert(pl.conf.end(), r.conf.begin() + r.coff[q], r.conf.begin() + r.coff[q] + 7 + T);
      pl.coords.insert(pl.coords.end(), r.coords.begin() + r.xoff[q], r.coords.begin() + r.xoff[q] + 3 * A);
    }
  });
  c->post = std::move(ps);
  if (st) { st->kernel_ms = r.ms; st->evals = r.evals; st->chains_per_wave = 1; st->blocks = (int)r.res.size(); }
}

}  // namespace

// =============================================================================================
extern "C" {

const char* vd_version(void) { return "vdock 0.1 (gfx950)"; }

int vd_create(int device, vd_ctx** out) {
  if (!out) return VD_ERR_ARG;
  *out = nullptr;
  if (device == -1) {
    vd_ctx* c = new vd_ctx();
    c->host_only = true;
    int rc = guard(c, [&]() { upload_lut(c); });
    if (rc != VD_OK) { delete c; return rc; }
    *out = c;
    return VD_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return VD_ERR_NO_DEVICE;
  vd_ctx* c = new vd_ctx();
  int rc = guard(c, [&]() {
    c->device = device;
    HIP_CHECK(hipSetDevice(device));
    c->stream = device_stream(device, 0);
    HIP_CHECK(hipEventCreate(&c->ev0));
    HIP_CHECK(hipEventCreate(&c->ev1));
    HIP_CHECK(hipEventCreate(&c->ev_base));
    for (int k = 0; k < kMcSlots; ++k) {
      HIP_CHECK(hipEventCreate(&c->mc[k].ev0));
      HIP_CHECK(hipEventCreate(&c->mc[k].ev1));
    }
    c->mc[0].stream = c->stream;
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    c->lds_max = (int)prop.sharedMemPerBlock;
    c->scratch = scratch_for_current_device(device);
    upload_lut(c);
  });
  if (rc != VD_OK) { delete c; return rc; }
  *out = c;
  return VD_OK;
}

void vd_destroy(vd_ctx* c) {
  if (!c) return;
  if (c->host_only) { delete c; return; }
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->lut_rows.release(); c->lutp.release(); c->maps.release(); c->corners.release(); c->grid_atoms.release();
  c->grid_cells.release();
  c->graphs.clear();
  for (hipStream_t& cs : c->cap)
    if (cs) { (void)hipStreamDestroy(cs); cs = nullptr; }
  for (int k = 0; k < kMcSlots; ++k) {
    if (c->mc[k].stream) (void)hipStreamSynchronize(c->mc[k].stream);
    c->mc[k].release();
  }
  if (c->ev_base) (void)hipEventDestroy(c->ev_base);
  c->blob.release(); c->headers.release(); c->io_a.release(); c->io_b.release(); c->io_c.release();
  c->batch.e.release(); c->batch.conf.release(); c->batch.heavy.release();
  c->tune_e.release(); c->tune_conf.release(); c->tune_heavy.release();
  c->nc_atoms.release(); c->nc_start.release(); c->ref_in.release(); c->ref_out.release(); c->ref_xyz.release();
  c->ref_res.release(); c->ref_lig.release(); c->ref_evals.release(); c->ref_coff.release(); c->ref_xoff.release();
  c->lutd.release(); c->ref_geo.release(); c->ref_pairs.release(); c->ref_hdr.release();
  c->pin_e.release(); c->pin_conf.release(); c->pin_heavy.release();
  c->pin_cout.release(); c->pin_xout.release(); c->pin_res.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  delete c;
}

const char* vd_last_error(const vd_ctx* c) { return c ? c->err.c_str() : "no context"; }

int vd_set_weights(vd_ctx* c, const double w[6]) {
  if (!c || !w) return VD_ERR_ARG;
  return guard(c, [&]() {
    if (!c->host_only) HIP_CHECK(hipSetDevice(c->device));
    std::memcpy(c->weights, w, sizeof(c->weights));
    upload_lut(c);
    c->maps_ready = false;
  });
}

int vd_get_lut(vd_ctx* c, int tp, double* se, double* sd, double* fast) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    if (tp < 0 || tp >= VD_NUM_TYPE_PAIRS) throw ArgError(VD_ERR_ARG, "type pair out of range");
    const size_t b = (size_t)tp * VD_LUT_BINS;
    for (int i = 0; i < VD_LUT_BINS; ++i) {
      if (se) se[i] = c->pre.smooth_e[b + i];
      if (sd) sd[i] = c->pre.smooth_dor[b + i];
      if (fast) fast[i] = c->pre.fast[b + i];
    }
  });
}

int vd_set_receptor(vd_ctx* c, int n, const float* xyz, const int* xs) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    if (n < 0 || (n > 0 && (!xyz || !xs))) throw ArgError(VD_ERR_ARG, "bad receptor arrays");
    for (int i = 0; i < 3 * n; ++i) if (!std::isfinite(xyz[i])) throw ArgError(VD_ERR_ARG, "non-finite receptor coordinate");
    for (int i = 0; i < n; ++i) if (xs[i] < 0 || xs[i] > VD_XS_SIZE) throw ArgError(VD_ERR_ARG, "bad receptor XS type");
    c->rec_xyz.assign(xyz, xyz + 3 * (size_t)n);
    c->rec_xs.assign(xs, xs + n);
    c->maps_ready = false;
    c->nc_ready = false;
  });
}

int vd_set_receptor_pdbqt(vd_ctx* c, const char* text, int* n_atoms) {
  if (!c || !text) return VD_ERR_ARG;
  return guard(c, [&]() {
    std::vector<double> xyz; std::vector<int> xs;
    vd::parse_receptor_pdbqt(text, xyz, xs);
    c->rec_xyz.swap(xyz); c->rec_xs.swap(xs);
    c->maps_ready = false;
    c->nc_ready = false;
    if (n_atoms) *n_atoms = (int)c->rec_xs.size();
  });
}

int vd_get_receptor(vd_ctx* c, int max, float* xyz, int* xs) {
  if (!c) return VD_ERR_ARG;
  const int n = (int)c->rec_xs.size();
  for (int i = 0; i < std::min(n, max); ++i) {
    if (xyz) for (int k = 0; k < 3; ++k) xyz[3 * i + k] = (float)c->rec_xyz[3 * i + k];
    if (xs) xs[i] = c->rec_xs[i];
  }
  return n;
}

int vd_box_from_center(const double center[3], const double size[3], double granularity, vd_box* box) {
  if (!center || !size || !box || !(granularity > 0)) return VD_ERR_ARG;
  for (int i = 0; i < 3; ++i) {
    if (!(size[i] > 0)) return VD_ERR_ARG;
    box->n[i] = (int)std::ceil(size[i] / granularity);
    const double span = granularity * box->n[i];
    box->begin[i] = center[i] - span / 2;
    box->end[i] = box->begin[i] + span;
  }
  return VD_OK;
}

int vd_grid_populate(vd_ctx* c, const vd_box* box) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    check_box(box);
    need_device(c);
    c->box = *box;
    c->gd = make_desc(*box);
    // receptor cell list over the box grown by the cutoff (typed atoms only)
    const double cell = 4.0, grow = 8.0 + 1e-3;
    GridCells gcl{};
    double org[3];
    for (int i = 0; i < 3; ++i) {
      org[i] = box->begin[i] - grow;
      gcl.origin[i] = (float)org[i];
      gcl.n[i] = (int)std::floor((box->end[i] + grow - org[i]) / cell) + 1;
    }
    gcl.cell = (float)cell;
    const long long ncell = (long long)gcl.n[0] * gcl.n[1] * gcl.n[2];
    std::vector<int> count(ncell + 1, 0);
    std::vector<long long> cidx(c->rec_xs.size(), -1);
    for (size_t i = 0; i < c->rec_xs.size(); ++i) {
      if (c->rec_xs[i] >= VD_XS_SIZE) continue;
      int ci[3]; bool ok = true;
      for (int k = 0; k < 3; ++k) {
        const double v = (c->rec_xyz[3 * i + k] - org[k]) / cell;
        if (!(v >= 0) || v >= gcl.n[k]) { ok = false; break; }
        ci[k] = (int)v;
      }
      if (!ok) continue;
      cidx[i] = ((long long)ci[2] * gcl.n[1] + ci[1]) * gcl.n[0] + ci[0];
      count[cidx[i] + 1]++;
    }
    for (long long k = 0; k < ncell; ++k) count[k + 1] += count[k];
    std::vector<int2> ranges(ncell);
    for (long long k = 0; k < ncell; ++k) ranges[k] = make_int2(count[k], count[k + 1]);
    std::vector<float4> sorted(std::max(count[ncell], 1));
    std::vector<int> fill(count.begin(), count.end() - 1);
    for (size_t i = 0; i < c->rec_xs.size(); ++i) {
      if (cidx[i] < 0) continue;
      float4 a;
      a.x = (float)(c->rec_xyz[3 * i] - org[0]); a.y = (float)(c->rec_xyz[3 * i + 1] - org[1]); a.z = (float)(c->rec_xyz[3 * i + 2] - org[2]);
      uint32_t t = (uint32_t)c->rec_xs[i]; std::memcpy(&a.w, &t, 4);
      sorted[fill[cidx[i]]++] = a;
    }
    c->grid_atoms.ensure(sorted.size());
    c->grid_cells.ensure(ranges.size());
    HIP_CHECK(hipMemcpyAsync(c->grid_atoms.p, sorted.data(), sorted.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->grid_cells.p, ranges.data(), ranges.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    c->maps.ensure((size_t)VD_XS_SIZE * c->gd.plane);
    HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    HIP_CHECK(vdl::launch_grid(c->grid_atoms.p, c->grid_cells.p, gcl, reinterpret_cast<const float4*>(c->lut_rows.p), c->maps.p,
                               c->gd, c->stream));
    HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->grid_ms = elapsed_ms(c);
    make_corners(c);
    c->maps_ready = true;
    c->nc_ready = false;
  });
}

double vd_grid_last_ms(const vd_ctx* c) { return c ? c->grid_ms : 0.0; }

int vd_grid_download(vd_ctx* c, float* maps) {
  if (!c || !maps) return VD_ERR_ARG;
  return guard(c, [&]() {
    need_device(c);
    if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no maps: call vd_grid_populate or vd_grid_upload first");
    HIP_CHECK(hipMemcpy(maps, c->maps.p, (size_t)VD_XS_SIZE * c->gd.plane * 4, hipMemcpyDeviceToHost));
  });
}

int vd_grid_upload(vd_ctx* c, const vd_box* box, const float* maps) {
  if (!c || !maps) return VD_ERR_ARG;
  return guard(c, [&]() {
    check_box(box);
    need_device(c);
    c->box = *box;
    c->gd = make_desc(*box);
    c->maps.ensure((size_t)VD_XS_SIZE * c->gd.plane);
    HIP_CHECK(hipMemcpy(c->maps.p, maps, (size_t)VD_XS_SIZE * c->gd.plane * 4, hipMemcpyHostToDevice));
    make_corners(c);
    c->maps_ready = true;
    c->nc_ready = false;
  });
}

int vd_ligand_add_pdbqt(vd_ctx* c, const char* text, int* id) {
  if (!c || !text) return VD_ERR_ARG;
  return guard(c, [&]() {
    vd::Ligand L = vd::parse_ligand_pdbqt(text);
    if (min_group(L) == 0) throw ArgError(VD_ERR_CAPACITY, "ligand has more than 57 torsions");
    c->ligs.push_back(std::move(L));
    if (id) *id = (int)c->ligs.size() - 1;
  });
}

int vd_ligand_info_get(vd_ctx* c, int id, vd_ligand_info* info) {
  if (!c || !info) return VD_ERR_ARG;
  return guard(c, [&]() {
    const vd::Ligand& L = lig_at(c, id);
    info->n_atoms = (int)L.atoms.size(); info->n_torsions = L.n_torsions(); info->n_pairs = (int)L.pairs.size();
    info->n_heavy = L.n_heavy; info->torsdof = L.torsdof; info->tree_depth = L.max_depth;
    info->num_tors = (float)L.num_tors;
  });
}

int vd_ligand_export(vd_ctx* c, int id, int* xs, int* el, int* node, float* local, int* parent, int* begin, int* end,
                     float* rel_origin, float* rel_axis, int* pairs) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    const vd::Ligand& L = lig_at(c, id);
    for (size_t i = 0; i < L.atoms.size(); ++i) {
      if (xs) xs[i] = L.atoms[i].xs;
      if (el) el[i] = L.atoms[i].el;
      if (node) node[i] = L.atoms[i].node;
      if (local) for (int k = 0; k < 3; ++k) local[3 * i + k] = (float)L.atoms[i].local[k];
    }
    for (size_t k = 0; k < L.nodes.size(); ++k) {
      if (parent) parent[k] = L.nodes[k].parent;
      if (begin) begin[k] = L.nodes[k].begin;
      if (end) end[k] = L.nodes[k].end;
      for (int j = 0; j < 3; ++j) {
        if (rel_origin) rel_origin[3 * k + j] = (float)L.nodes[k].rel_origin[j];
        if (rel_axis) rel_axis[3 * k + j] = (float)L.nodes[k].rel_axis[j];
      }
    }
    if (pairs)
      for (size_t p = 0; p < L.pairs.size(); ++p) {
        pairs[3 * p] = (int)(L.pairs[p] >> 16); pairs[3 * p + 1] = (int)(L.pairs[p] & 255); pairs[3 * p + 2] = (int)((L.pairs[p] >> 8) & 255);
      }
  });
}

int vd_ligand_count(const vd_ctx* c) { return c ? (int)c->ligs.size() : 0; }

int vd_eval(vd_ctx* c, int id, long long n, const float* conf, float* e, float* grad, float v_intra, float v_inter,
            float slope, int cpw, int device_ptrs) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    need_device(c);
    if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no maps");
    if (n < 0 || (n > 0 && (!conf || !e || !grad))) throw ArgError(VD_ERR_ARG, "bad eval arrays");
    if (cpw != 0 && cpw != 1 && cpw != 2 && cpw != 4) throw ArgError(VD_ERR_ARG, "chains_per_wave must be 0, 1, 2 or 4");
    const vd::Ligand& L = lig_at(c, id);
    if (n == 0) return;
    const int T = L.n_torsions();
    int G = pick_group(c, L, cpw);
    Packed pk = pack(c, G, {id});
    c->blob.ensure(pk.blob.size() / 4 + 1);
    c->headers.ensure(1);
    HIP_CHECK(hipMemcpyAsync(c->blob.p, pk.blob.data(), pk.blob.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->headers.p, pk.headers.data(), sizeof(vd_lig_header), hipMemcpyHostToDevice, c->stream));
    const float* d_conf = conf; float* d_e = e; float* d_g = grad;
    if (!device_ptrs) {
      c->io_a.ensure((size_t)n * (7 + T)); c->io_b.ensure(n); c->io_c.ensure((size_t)n * (6 + T));
      HIP_CHECK(hipMemcpyAsync(c->io_a.p, conf, (size_t)n * (7 + T) * 4, hipMemcpyHostToDevice, c->stream));
      d_conf = c->io_a.p; d_e = c->io_b.p; d_g = c->io_c.p;
    }
    EvalArgs a{};
    a.blob = c->blob.p; a.header = c->headers.p; a.host_header = pk.headers[0];
    a.K = make_evalk(c, v_intra, v_inter, slope);
    a.conf = d_conf; a.n_conf = n; a.e = d_e; a.grad = d_g;
    const long long per_block = 64 / G;
    const int blocks = (int)std::min<long long>((n + per_block - 1) / per_block, (long long)c->n_cu * 16);
    HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    HIP_CHECK(vdl::launch_eval(G, a, blocks, c->stream));
    HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    if (!device_ptrs) {
      HIP_CHECK(hipMemcpyAsync(e, d_e, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipMemcpyAsync(grad, d_g, (size_t)n * (6 + T) * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->eval_ms = elapsed_ms(c);
  });
}

double vd_eval_last_ms(const vd_ctx* c) { return c ? c->eval_ms : 0.0; }

int vd_dock_params_default(vd_dock_params* p) {
  if (!p) return VD_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->chains = 8192; p->search_depth = 0; p->bfgs_steps = 0;
  p->local_search = VD_LS_RILC_BFGS; p->policy = VD_MC_VINA; p->mutation_quirk = 0; p->chains_per_wave = 0;
  p->seed = 42; p->temperature = 1.2f; p->amplitude = 2.0f;
  p->v_intra = 10.0f; p->v_inter = 10.0f; p->slope = 1e6f;
  return VD_OK;
}

int vd_dock_heuristics(int n_atoms, int n_torsions, int rilc, int* depth, int* bfgs) {
  if (n_atoms < 0 || n_torsions < 0) return VD_ERR_ARG;
  if (depth) *depth = heur_depth(n_atoms, n_torsions, rilc);
  if (bfgs) *bfgs = heur_bfgs(n_atoms, rilc);
  return VD_OK;
}

}  // extern "C"

namespace {

// Uploads n starting conformations [n][7+T] (finite; quaternions are normalised on the device).
const float* upload_starts(vd_ctx* c, const float* conf, long long n, int T) {
  const size_t m = (size_t)n * (7 + T);
  for (size_t i = 0; i < m; ++i)
    if (!std::isfinite(conf[i])) throw ArgError(VD_ERR_ARG, "non-finite starting conformation value");
  for (long long k = 0; k < n; ++k) {
    const float* q = conf + (size_t)k * (7 + T) + 3;
    if (!(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] > 1e-12f))
      throw ArgError(VD_ERR_ARG, "zero orientation quaternion in a starting conformation");
  }
  c->init.ensure(m);
  HIP_CHECK(hipMemcpyAsync(c->init.p, conf, m * 4, hipMemcpyHostToDevice, c->stream));
  return c->init.p;
}

// One plain launch of ligand id's chains (vd_dock, vd_local_search): results in the batch buffers.
void dock_one(vd_ctx* c, int id, const vd_dock_params* p, const float* init_conf, bool ls_only, vd_stats* st,
              LigPlan& pl) {
  check_params(p);
  need_device(c);
  if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no maps");
  if (st) std::memset(st, 0, sizeof(*st));
  if (c->autotune_on && (p->chains_per_wave == 0 || p->waves_per_simd == 0)) autotune(c, {id}, p);
  c->launches.clear();
  pl = plan_for(c, id, p, true);
  if (ls_only) pl.depth = 0;
  prepare_batch(c, {pl});
  const float* d_init = init_conf ? upload_starts(c, init_conf, p->chains, pl.T) : nullptr;
  std::vector<int> dv{pl.depth}, bv{pl.bfgs};
  Packed pk = pack(c, pl.G, {id}, &dv, &bv);
  const int cpw = 64 / pl.G;
  std::vector<int2> items;
  for (int c0 = 0; c0 < p->chains; c0 += cpw) items.push_back(make_int2(0, c0));
  run_mc(c, pl.G, pl.W, pk, {(uint32_t)id}, items, c->batch.off, make_consts(c, p, pl.depth, pl.bfgs), c->batch.e.p,
         c->batch.conf.p, c->batch.heavy.p, false, st, true, d_init, ls_only);
}

}  // namespace

extern "C" {

int vd_dock(vd_ctx* c, int id, const vd_dock_params* p, const float* init_conf, float* e, float* conf, float* heavy,
            vd_stats* st) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() {
    LigPlan pl;
    dock_one(c, id, p, init_conf, false, st, pl);
    if (e) HIP_CHECK(hipMemcpy(e, c->batch.e.p, (size_t)p->chains * 4, hipMemcpyDeviceToHost));
    if (conf) HIP_CHECK(hipMemcpy(conf, c->batch.conf.p, (size_t)p->chains * (7 + pl.T) * 4, hipMemcpyDeviceToHost));
    if (heavy) HIP_CHECK(hipMemcpy(heavy, c->batch.heavy.p, (size_t)p->chains * pl.H * 12, hipMemcpyDeviceToHost));
  });
}

int vd_local_search(vd_ctx* c, int id, int n, const float* conf_in, const vd_dock_params* p, float* conf_out, float* e,
                    vd_stats* st) {
  if (!c || n < 1 || !conf_in || !p) return VD_ERR_ARG;
  return guard(c, [&]() {
    vd_dock_params q = *p;
    q.chains = n;
    LigPlan pl;
    dock_one(c, id, &q, conf_in, true, st, pl);
    if (e) HIP_CHECK(hipMemcpy(e, c->batch.e.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (conf_out) HIP_CHECK(hipMemcpy(conf_out, c->batch.conf.p, (size_t)n * (7 + pl.T) * 4, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

namespace {

// The no-allocation docking call: the plain launch of vd_dock with every buffer beyond the
// caller's inputs and outputs carved out of this device's scratch array.
void dock_noalloc(vd_ctx* c, int id, const vd_dock_params* p, const float* d_init, float* d_e, float* d_conf,
                  float* d_heavy, vd_stats* st) {
  check_params(p);
  need_device(c);
  if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no maps");
  if (!d_e || !d_conf || !d_heavy) throw ArgError(VD_ERR_ARG, "output pointers are NULL");
  if (!c->scratch) throw ArgError(VD_ERR_ARG, "no scratch array on this device");
  if (st) std::memset(st, 0, sizeof(*st));
  LigPlan pl = plan_for(c, id, p, true);
  std::vector<int> dv{pl.depth}, bv{pl.bfgs};
  Packed pk = pack(c, pl.G, {id}, &dv, &bv);
  const int cpw = 64 / pl.G;
  const int blocks = (p->chains + cpw - 1) / cpw;
  vd_mc_consts P = make_consts(c, p, pl.depth, pl.bfgs);
  const int mode = mc_mode(P);
  const int lds = vdl::mc_lds_bytes(pl.G, pk.Amax, pk.Nmax, mode == 0);
  if (lds > c->lds_max) throw ArgError(VD_ERR_CAPACITY, "ligand too large for the LDS layout");
  // scratch layout, 256-byte aligned regions
  size_t used = 0;
  auto take = [&](size_t bytes) { const size_t o = used; used += (bytes + 255) / 256 * 256; return o; };
  const size_t o_hdr = take(sizeof(vd_lig_header)), o_stream = take(4), o_off = take(sizeof(longlong4));
  const size_t o_evals = take(8), o_blob = take(pk.blob.size() * 4);
  long long hist_stride = 0, glob_floats = 0;
  int axis = 0;
  size_t o_hist = 0, o_glob = 0, o_goff = 0, o_gaxis = 0, o_gT = 0;
  if (mode == 2 || mode == 3) {
    hist_stride = vdk::History::floats_per_chain(pl.T, mode == 3 ? 5 : 10);
    o_hist = take((size_t)blocks * cpw * hist_stride * 4);
  }
  if (mode == 3) {
    axis = 1 << qvw_levels(p->chains, pl.depth);
    pk.headers[0].reserved = axis;
    glob_floats = vdk::GlobalHist::floats(axis, pl.T);
    o_glob = take((size_t)glob_floats * 4);
    o_goff = take(8); o_gaxis = take(4); o_gT = take(4);
  }
  if (used > VD_NOALLOC_SCRATCH_BYTES)
    throw ArgError(VD_ERR_CAPACITY, "vd_dock_noalloc needs " + std::to_string(used) + " bytes of scratch; the array holds " +
                                        std::to_string((size_t)VD_NOALLOC_SCRATCH_BYTES) +
                                        " (fewer chains, or vd_dock, which allocates what it needs)");
  unsigned char* base = c->scratch;
  hipStream_t s = c->stream;
  const uint32_t stream_id = (uint32_t)id;
  const longlong4 off = make_longlong4(0, 0, 0, 0);
  HIP_CHECK(hipMemcpyAsync(base + o_hdr, pk.headers.data(), sizeof(vd_lig_header), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(base + o_stream, &stream_id, 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(base + o_off, &off, sizeof(longlong4), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(base + o_blob, pk.blob.data(), pk.blob.size() * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(vdl::zero_u64(reinterpret_cast<unsigned long long*>(base + o_evals), 1, s));
  const long long goff = 0;
  const int T = pl.T;
  if (mode == 3) {
    HIP_CHECK(hipMemcpyAsync(base + o_goff, &goff, 8, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(base + o_gaxis, &axis, 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(base + o_gT, &T, 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(vdl::glob_init(reinterpret_cast<float*>(base + o_glob), reinterpret_cast<const long long*>(base + o_goff),
                             reinterpret_cast<const int*>(base + o_gaxis), reinterpret_cast<const int*>(base + o_gT), 1, s));
  }
  McArgs a{};
  a.blob = reinterpret_cast<const float4*>(base + o_blob);
  a.headers = reinterpret_cast<const vd_lig_header*>(base + o_hdr);
  a.streams = reinterpret_cast<const uint32_t*>(base + o_stream);
  a.items = nullptr;
  a.n_items = blocks;
  a.init_conf = d_init;
  a.ls_only = 0;
  a.K = make_evalk(c, P.v_intra, P.v_inter, P.slope);
  a.P = P;
  a.out_e = d_e; a.out_conf = d_conf; a.out_heavy = d_heavy;
  a.out_off = reinterpret_cast<const longlong4*>(base + o_off);
  a.slot_evals = reinterpret_cast<unsigned long long*>(base + o_evals);
  a.Amax = pk.Amax; a.Nmax = pk.Nmax;
  if (mode == 2 || mode == 3) { a.hist = reinterpret_cast<float*>(base + o_hist); a.hist_stride = hist_stride; }
  if (mode == 3) a.glob = reinterpret_cast<float*>(base + o_glob);
  const int W = loosest_budget(pl.G, pl.W, mode, lds);
  McSlot& slot = c->mc[0];
  if (slot.busy) { HIP_CHECK(hipStreamSynchronize(slot.stream)); slot.busy = false; }
  HIP_CHECK(hipEventRecord(slot.ev0, s));
  HIP_CHECK(vdl::launch_mc(pl.G, W, a, blocks, s));
  HIP_CHECK(hipEventRecord(slot.ev1, s));
  unsigned long long evals = 0;
  HIP_CHECK(hipMemcpyAsync(&evals, base + o_evals, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, slot.ev0, slot.ev1));
  vd_launch rec{};
  rec.lanes_per_chain = pl.G; rec.mode = mode; rec.waves_per_simd = W; rec.blocks = blocks; rec.items = blocks;
  rec.ligands = 1; rec.kernel_ms = ms; rec.evals = evals; rec.eval_bytes = (double)evals * eval_bytes(pk.headers[0]);
  c->launches.assign(1, rec);
  if (st) { st->kernel_ms = ms; st->evals = evals; st->chains_per_wave = cpw; st->blocks = blocks; }
}

}  // namespace

extern "C" {

int vd_dock_noalloc(vd_ctx* c, int id, const vd_dock_params* p, const float* d_init_conf, float* d_e, float* d_conf,
                    float* d_heavy, vd_stats* st) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() { dock_noalloc(c, id, p, d_init_conf, d_e, d_conf, d_heavy, st); });
}

}  // extern "C"

namespace {

// The launch classes of a persistent batch: ligand indices per (lanes per chain, waves per SIMD).
using Classes = std::map<std::pair<int, int>, std::vector<int>>;
Classes batch_classes(const std::vector<LigPlan>& plans) {
  Classes classes;
  for (size_t k = 0; k < plans.size(); ++k) classes[std::make_pair(plans[k].G, plans[k].W)].push_back((int)k);
  return classes;
}

// Overlap applies to batches of exactly two classes.  Measured on the screening batch (a 16-lane
// class whose waves mostly wait, a 64-lane class that keeps the vector units busy): 3 % shorter
// steps.  With more classes it lost: on a 4,096-ligand batch of five classes a fourth stream
// shared a hardware queue with the third (four queues per process), and capped classes that
// outlast the uncapped one keep half their CU slots empty to the end.
bool overlap_of(const vd_ctx* c, const Classes& classes) { return c->overlap_classes && classes.size() == 2; }

// Host work of one class of a persistent batch -- its ligand records, launch tables and work items
// in decreasing estimated cost -- prepared into launch slot s.  share: the fraction of its resident
// blocks per CU the class launches.  With overlap, every class but the last (the widest groups)
// takes half, which leaves room for the last class's waves beside it; the last class launches its
// full grid, whose blocks that do not fit at first take over the CU slots the others give up when
// their queues run dry.  Measured on the screening batch (16- and 64-lane classes): 0.5 to 0.75 of
// the residency are equally good, below 0.4 the first class runs too long.
void prepare_class(vd_ctx* c, McSlot& s, int G, int W, const std::vector<int>& ks, const std::vector<LigPlan>& plans,
                   const vd_dock_params* p, double share) {
  std::vector<int> lig_ids, dv, bv;
  for (int k : ks) { lig_ids.push_back(plans[k].id); dv.push_back(plans[k].depth); bv.push_back(plans[k].bfgs); }
  Packed pk = pack(c, G, lig_ids, &dv, &bv);
  std::vector<uint32_t> streams;
  std::vector<longlong4> off;
  for (int k : ks) { streams.push_back((uint32_t)plans[k].id); off.push_back(c->batch.off[k]); }
  // cost estimate per chain: steps x iterations x (atoms + pairs)
  std::vector<std::pair<double, int>> order;
  for (size_t s = 0; s < ks.size(); ++s)
    order.push_back({-(double)dv[s] * bv[s] * (pk.headers[s].n_atoms + 0.5 * pk.headers[s].n_chunks * G + 16), (int)s});
  std::stable_sort(order.begin(), order.end());
  const int cpw = 64 / G;
  std::vector<int2> items;
  for (auto& o : order)
    for (int c0 = 0; c0 < p->chains; c0 += cpw) items.push_back(make_int2(o.second, c0));
  if (!s.stream) s.stream = device_stream(c->device, (int)(&s - c->mc));
  mc_prepare(c, s, G, W, std::move(pk), std::move(streams), std::move(items), std::move(off), make_consts(c, p, 0, 0),
             c->batch.e.p, c->batch.conf.p, c->batch.heavy.p, true, true, share);
}

// Every device buffer of a launch slot that a captured graph reads or writes.  DevBuf::ensure
// frees and reallocates only when it grows, so a replay is valid while these are unchanged.
std::vector<const void*> slot_ptrs(const McSlot& s) {
  return {s.blob.p, s.headers.p, s.streams.p, s.items.p, s.offs.p, s.queue.p, s.evals.p,
          s.hist.p, s.glob.p, s.glob_off.p, s.glob_axis.p, s.glob_T.p};
}

// Captures the batch's stages (classes already prepared into slots 0..n-1) into a graph.
std::unique_ptr<GraphEntry> capture_batch(vd_ctx* c, const Classes& classes, bool overlap) {
  auto g = std::make_unique<GraphEntry>();
  const int n = (int)classes.size();
  // pinned staging: every host-to-device copy's source, 256-byte aligned pieces
  size_t bytes = 0, n_evals = 0;
  auto piece = [&](size_t b) { const size_t at = bytes; bytes += (b + 255) / 256 * 256; return at; };
  struct At { size_t goff, gaxis, gT, blob, headers, streams, items, offs; };
  std::vector<At> at(n);
  for (int k = 0; k < n; ++k) {
    const McSlot& s = c->mc[k];
    at[k] = At{piece(s.h_goff.size() * 8), piece(s.h_gaxis.size() * 4), piece(s.h_gT.size() * 4),
               piece(s.pk.blob.size() * 4), piece(s.h_headers.size() * sizeof(vd_lig_header)),
               piece(s.h_streams.size() * 4), piece(s.h_items.size() * sizeof(int2)),
               piece(s.h_offs.size() * sizeof(longlong4))};
    n_evals += s.h_headers.size();
  }
  g->staging.ensure(bytes);
  g->evals.ensure(n_evals);
  unsigned char* base = g->staging.p;
  size_t evals_at = 0;
  for (int k = 0; k < n; ++k) {
    const McSlot& s = c->mc[k];
    auto put = [&](size_t off, const void* src, size_t b) { if (b) std::memcpy(base + off, src, b); };
    put(at[k].goff, s.h_goff.data(), s.h_goff.size() * 8);
    put(at[k].gaxis, s.h_gaxis.data(), s.h_gaxis.size() * 4);
    put(at[k].gT, s.h_gT.data(), s.h_gT.size() * 4);
    put(at[k].blob, s.pk.blob.data(), s.pk.blob.size() * 4);
    put(at[k].headers, s.h_headers.data(), s.h_headers.size() * sizeof(vd_lig_header));
    put(at[k].streams, s.h_streams.data(), s.h_streams.size() * 4);
    put(at[k].items, s.h_items.data(), s.h_items.size() * sizeof(int2));
    put(at[k].offs, s.h_offs.data(), s.h_offs.size() * sizeof(longlong4));
    GraphClass gc;
    gc.slot = k; gc.a = s.a; gc.rec = s.rec; gc.headers = s.h_headers; gc.evals_at = evals_at;
    gc.ptrs = slot_ptrs(s);
    evals_at += s.h_headers.size();
    g->cls.push_back(std::move(gc));
  }
  // captured on streams of this context's own: the launch streams are shared by every context of
  // the process, and work another context issued to a stream while it captures would be captured
  for (int k = 0; k < 2; ++k)
    if (!c->cap[k]) HIP_CHECK(hipStreamCreate(&c->cap[k]));
  hipStream_t main = c->cap[0], side = c->cap[1];
  HIP_CHECK(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
  try {
    if (overlap) {   // two classes: the second on a branch of its own
      HIP_CHECK(hipEventRecord(c->ev_base, main));
      HIP_CHECK(hipStreamWaitEvent(side, c->ev_base, 0));
    }
    auto h2d = [&](hipStream_t st, void* dst, size_t off, size_t b) {
      if (b) HIP_CHECK(hipMemcpyAsync(dst, base + off, b, hipMemcpyHostToDevice, st));
    };
    // every class's copies before any kernel, as in the eager path
    for (int k = 0; k < n; ++k) {
      McSlot& s = c->mc[k];
      hipStream_t st = overlap && k == 1 ? side : main;
      if (!s.h_goff.empty()) {
        h2d(st, s.glob_off.p, at[k].goff, s.h_goff.size() * 8);
        h2d(st, s.glob_axis.p, at[k].gaxis, s.h_gaxis.size() * 4);
        h2d(st, s.glob_T.p, at[k].gT, s.h_gT.size() * 4);
        HIP_CHECK(vdl::glob_init(s.glob.p, s.glob_off.p, s.glob_axis.p, s.glob_T.p, (int)s.h_goff.size(), st));
      }
      h2d(st, s.blob.p, at[k].blob, s.pk.blob.size() * 4);
      h2d(st, s.headers.p, at[k].headers, s.h_headers.size() * sizeof(vd_lig_header));
      h2d(st, s.streams.p, at[k].streams, s.h_streams.size() * 4);
      h2d(st, s.items.p, at[k].items, s.h_items.size() * sizeof(int2));
      h2d(st, s.offs.p, at[k].offs, s.h_offs.size() * sizeof(longlong4));
      HIP_CHECK(vdl::zero_u64(s.evals.p, (int)s.h_headers.size(), st));
      HIP_CHECK(vdl::reset_counter(s.queue.p, st));
    }
    for (int k = 0; k < n; ++k) {
      McSlot& s = c->mc[k];
      HIP_CHECK(vdl::launch_mc(s.G, s.W, s.a, s.blocks, overlap && k == 1 ? side : main));
    }
    if (overlap) {
      HIP_CHECK(hipEventRecord(c->mc[1].ev1, side));
      HIP_CHECK(hipStreamWaitEvent(main, c->mc[1].ev1, 0));
    }
    for (int k = 0; k < n; ++k) {
      const McSlot& s = c->mc[k];
      HIP_CHECK(hipMemcpyAsync(g->evals.p + g->cls[k].evals_at, s.evals.p, s.h_headers.size() * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, main));
    }
  } catch (...) {
    hipGraph_t partial = nullptr;
    (void)hipStreamEndCapture(main, &partial);
    if (partial) (void)hipGraphDestroy(partial);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(main, &g->graph));
  HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  // the MC kernel node of each class, for seed updates: the launched function, and the captured
  // argument's ligand records (classes with the same instantiation differ there)
  size_t nn = 0;
  HIP_CHECK(hipGraphGetNodes(g->graph, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn) HIP_CHECK(hipGraphGetNodes(g->graph, nodes.data(), &nn));
  for (hipGraphNode_t node : nodes) {
    hipGraphNodeType type;
    HIP_CHECK(hipGraphNodeGetType(node, &type));
    if (type != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams kp{};
    HIP_CHECK(hipGraphKernelNodeGetParams(node, &kp));
    for (GraphClass& gc : g->cls) {
      if (gc.node || kp.func != vdl::mc_kernel_fn(gc.rec.lanes_per_chain, gc.rec.waves_per_simd, gc.a)) continue;
      if (!kp.kernelParams || !kp.kernelParams[0] || static_cast<const McArgs*>(kp.kernelParams[0])->blob != gc.a.blob)
        continue;
      gc.node = node; gc.kp = kp;
      break;
    }
  }
  g->seed = g->cls.empty() ? 0 : g->cls[0].a.P.seed;
  return g;
}

// Runs a persistent batch as a graph: replays the captured graph of the same ligands and
// parameters when one is kept and its buffers are still those it captured, or captures one.
// Returns false when the batch has more classes than launch slots (then it runs eagerly).
bool graph_dock(vd_ctx* c, const std::vector<int>& ids, const std::vector<LigPlan>& plans, const Classes& classes,
                const vd_dock_params* p, vd_stats* st) {
  if (classes.size() > (size_t)kMcSlots) return false;
  const bool overlap = overlap_of(c, classes);
  vd_dock_params key = *p;
  key.seed = 0;
  std::vector<std::pair<int, int>> shapes;
  for (const LigPlan& pl : plans) shapes.push_back(std::make_pair(pl.G, pl.W));
  const std::vector<const void*> outs{c->batch.e.p, c->batch.conf.p, c->batch.heavy.p};
  std::unique_ptr<GraphEntry> g;
  for (size_t k = 0; k < c->graphs.size(); ++k) {
    GraphEntry& e = *c->graphs[k];
    if (e.ids != ids || std::memcmp(&e.p, &key, sizeof key) != 0 || e.shapes != shapes || e.overlap != overlap) continue;
    bool valid = e.outs == outs;
    for (const GraphClass& gc : e.cls) valid = valid && gc.ptrs == slot_ptrs(c->mc[gc.slot]);
    bool nodes = true;   // a kernel node to update per class
    for (const GraphClass& gc : e.cls) nodes = nodes && gc.node;
    if (valid && (e.seed == p->seed || nodes)) g = std::move(c->graphs[k]);
    c->graphs.erase(c->graphs.begin() + k);
    break;
  }
  if (!g) {
    int slot = 0;
    for (auto& cl : classes) {
      const bool last = std::next(classes.find(cl.first)) == classes.end();
      prepare_class(c, c->mc[slot++], cl.first.first, cl.first.second, cl.second, plans, p, (overlap && !last) ? 0.5 : 1.0);
    }
    g = capture_batch(c, classes, overlap);
    g->ids = ids; g->p = key; g->shapes = shapes; g->overlap = overlap; g->outs = outs;
  }
  // a launch that a failed eager call left in flight still uses its slot's buffers, which the
  // graph rewrites from another stream: wait for it first
  for (McSlot& s : c->mc)
    if (s.busy) { HIP_CHECK(hipStreamSynchronize(s.stream)); s.busy = false; }
  if (g->seed != p->seed) {
    for (GraphClass& gc : g->cls) {
      gc.a.P.seed = p->seed;
      gc.kargs[0] = &gc.a;
      gc.kp.kernelParams = gc.kargs;
      gc.kp.extra = nullptr;
      HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec, gc.node, &gc.kp));
    }
    g->seed = p->seed;
  }
  HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  HIP_CHECK(hipGraphLaunch(g->exec, c->stream));
  HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  for (const GraphClass& gc : g->cls) {
    // launches of one graph are timed together: each record carries the graph's span
    vd_launch rec = gc.rec;
    rec.kernel_ms = ms; rec.start_ms = 0; rec.evals = 0; rec.eval_bytes = 0;
    for (size_t k = 0; k < gc.headers.size(); ++k) {
      const unsigned long long v = g->evals.p[gc.evals_at + k];
      rec.evals += v;
      rec.eval_bytes += (double)v * eval_bytes(gc.headers[k]);
    }
    c->launches.push_back(rec);
    if (st) { st->evals += rec.evals; st->blocks += rec.blocks; st->chains_per_wave = 64 / rec.lanes_per_chain; }
  }
  if (st) st->kernel_ms += ms;
  c->graphs.push_back(std::move(g));
  if (c->graphs.size() > kMaxGraphs) c->graphs.erase(c->graphs.begin());
  return true;
}

}  // namespace

extern "C" {

static int dock_many(vd_ctx* c, int n, const int* ids, const vd_dock_params* p, vd_stats* st, bool persistent) {
  if (!c || n < 0 || (n > 0 && !ids)) return VD_ERR_ARG;
  return guard(c, [&]() {
    check_params(p);
    need_device(c);
    if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no maps");
    if (st) std::memset(st, 0, sizeof(*st));
    if (c->autotune_on && (p->chains_per_wave == 0 || p->waves_per_simd == 0))
      autotune(c, std::vector<int>(ids, ids + n), p);
    c->launches.clear();
    std::vector<LigPlan> plans;
    for (int k = 0; k < n; ++k) plans.push_back(plan_for(c, ids[k], p, !persistent));
    prepare_pair_chunks(c, plans);
    prepare_batch(c, plans);
    if (n == 0) return;
    if (!persistent) {
      // the plain per-ligand loop: one launch per ligand
      for (int k = 0; k < n; ++k) {
        const LigPlan& pl = plans[k];
        std::vector<int> dv{pl.depth}, bv{pl.bfgs};
        Packed pk = pack(c, pl.G, {pl.id}, &dv, &bv);
        const int cpw = 64 / pl.G;
        std::vector<int2> items;
        for (int c0 = 0; c0 < p->chains; c0 += cpw) items.push_back(make_int2(0, c0));
        std::vector<longlong4> off{c->batch.off[k]};
        run_mc(c, pl.G, pl.W, pk, {(uint32_t)pl.id}, items, off, make_consts(c, p, pl.depth, pl.bfgs), c->batch.e.p,
               c->batch.conf.p, c->batch.heavy.p, false, st, true);
      }
      return;
    }
    // one persistent launch per (group width, register budget) class; items ordered by decreasing
    // estimated cost.  With class overlap the launches are all prepared first (buffer growth may
    // synchronise the device), then enqueued on one stream each, then collected.
    const Classes classes = batch_classes(plans);
    const bool overlap = overlap_of(c, classes);
    if (c->graphs_on && graph_dock(c, std::vector<int>(ids, ids + n), plans, classes, p, st)) return;
    int used = 0;
    auto flush = [&]() {
      if (used == 0) return;
      HIP_CHECK(hipEventRecord(c->ev_base, c->stream));
      for (int k = 1; k < used; ++k) HIP_CHECK(hipStreamWaitEvent(c->mc[k].stream, c->ev_base, 0));
      for (int k = 0; k < used; ++k) mc_enqueue_copies(c->mc[k]);
      for (int k = 0; k < used; ++k) mc_enqueue_kernel(c->mc[k]);
      // launches that ran at the same time count once: device time from first start to last end
      double t0 = 1e300, t1 = -1e300;
      for (int k = 0; k < used; ++k) {
        mc_collect(c, c->mc[k], st, c->ev_base, false);
        t0 = std::min(t0, c->mc[k].rec.start_ms);
        t1 = std::max(t1, c->mc[k].rec.start_ms + c->mc[k].rec.kernel_ms);
      }
      if (st) st->kernel_ms += t1 - t0;
      used = 0;
    };
    for (auto& cl : classes) {
      if (used == (overlap ? kMcSlots : 1)) flush();
      const bool last = std::next(classes.find(cl.first)) == classes.end();
      prepare_class(c, c->mc[used], cl.first.first, cl.first.second, cl.second, plans, p, (overlap && !last) ? 0.5 : 1.0);
      ++used;
    }
    flush();
  });
}

int vd_set_autotune(vd_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return VD_ERR_ARG;
  c->autotune_on = on == 1;
  if (!c->autotune_on) c->tuned.clear();
  return VD_OK;
}

int vd_set_class_overlap(vd_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return VD_ERR_ARG;
  c->overlap_classes = on == 1;
  return VD_OK;
}

int vd_set_graphs(vd_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return VD_ERR_ARG;
  return guard(c, [&]() {
    c->graphs_on = on == 1;
    if (!c->graphs_on && !c->host_only) {
      HIP_CHECK(hipStreamSynchronize(c->stream));
      c->graphs.clear();
    }
  });
}

int vd_dock_batch(vd_ctx* c, int n, const int* ids, const vd_dock_params* p, vd_stats* st) {
  return dock_many(c, n, ids, p, st, true);
}
int vd_dock_batch_plain(vd_ctx* c, int n, const int* ids, const vd_dock_params* p, vd_stats* st) {
  return dock_many(c, n, ids, p, st, false);
}

int vd_ligand_pair_chunks(vd_ctx* c, int id, int lanes, int max, uint32_t* words) {
  if (!c || max < 0 || (max > 0 && !words)) return VD_ERR_ARG;
  int n = 0;
  const int rc = guard(c, [&]() {
    if (lanes != 16 && lanes != 32 && lanes != 64) throw ArgError(VD_ERR_ARG, "lanes must be 16, 32 or 64");
    const std::vector<uint32_t>& w = vd::pair_chunks(lig_at(c, id), lanes);
    n = (int)w.size();
    for (int k = 0; k < n && k < max; ++k) words[k] = w[k];
  });
  return rc == VD_OK ? n : rc;
}

int vd_last_launches(vd_ctx* c, int max, vd_launch* out) {
  if (!c || max < 0 || (max > 0 && !out)) return VD_ERR_ARG;
  const int n = (int)c->launches.size();
  for (int k = 0; k < n && k < max; ++k) out[k] = c->launches[k];
  return n;
}

int vd_batch_results(vd_ctx* c, int k, int max_chains, float* e, float* conf, float* heavy) {
  if (!c || max_chains < 0) return VD_ERR_ARG;
  int count = 0;
  const int rc = guard(c, [&]() {
    need_device(c);
    BatchState& b = c->batch;
    if (k < 0 || k >= (int)b.ids.size()) throw ArgError(VD_ERR_ARG, "batch index out of range");
    const longlong4 o = b.off[k];
    count = b.chains[k];
    const size_t ch = (size_t)std::min(count, max_chains);
    if (ch == 0) return;
    if (e) HIP_CHECK(hipMemcpy(e, b.e.p + o.x, ch * 4, hipMemcpyDeviceToHost));
    if (conf) HIP_CHECK(hipMemcpy(conf, b.conf.p + o.y, ch * (7 + b.T[k]) * 4, hipMemcpyDeviceToHost));
    if (heavy) HIP_CHECK(hipMemcpy(heavy, b.heavy.p + o.z, ch * b.H[k] * 12, hipMemcpyDeviceToHost));
  });
  return rc == VD_OK ? count : rc;
}

int vd_batch_best(vd_ctx* c, int max, float* best_e) {
  if (!c || max < 0 || (max > 0 && !best_e)) return VD_ERR_ARG;
  int n = 0;
  const int rc = guard(c, [&]() {
    BatchState& b = c->batch;
    n = (int)b.ids.size();
    const int m = std::min(n, max);
    if (m == 0) return;
    need_device(c);
    size_t tot = 0;
    for (int k = 0; k < m; ++k) tot = std::max<size_t>(tot, (size_t)(b.off[k].x + b.chains[k]));
    std::vector<float> e(tot);
    if (tot) HIP_CHECK(hipMemcpy(e.data(), b.e.p, tot * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) {
      float v = INFINITY;
      for (int i = 0; i < b.chains[k]; ++i) v = std::min(v, e[b.off[k].x + i]);
      best_e[k] = v;
    }
  });
  return rc == VD_OK ? n : rc;
}

int vd_cluster(int n, int n_heavy, const float* e, const float* heavy, double min_rmsd, int max_size, int* kept) {
  if (n < 0 || n_heavy < 0 || (n > 0 && (!e || !heavy || !kept)) || max_size < 1) return VD_ERR_ARG;
  std::vector<int> k = vd::cluster_poses(n, n_heavy, e, heavy, min_rmsd, max_size);
  for (size_t i = 0; i < k.size(); ++i) kept[i] = k[i];
  return (int)k.size();
}

int vd_post_params_default(vd_post_params* p) {
  if (!p) return VD_ERR_ARG;
  p->max_steps = 0; p->local_search = VD_LS_RILC_BFGS; p->num_modes = 9; p->energy_range = 3.0f;
  p->min_rmsd = 1.0f; p->v = 1000.0f; p->slope = 1e6f;
  return VD_OK;
}

int vd_refine(vd_ctx* c, int id, int n, const float* conf_in, const vd_post_params* p, float* conf_out, float* coords_out,
              vd_refined* out) {
  if (!c || n < 0 || (n > 0 && !conf_in)) return VD_ERR_ARG;
  return guard(c, [&]() {
    RefineOut r = run_refine(c, {id}, {n}, conf_in, p);
    if (conf_out) std::memcpy(conf_out, r.conf.data(), r.conf.size() * sizeof(float));
    if (coords_out) std::memcpy(coords_out, r.coords.data(), r.coords.size() * sizeof(float));
    if (out) std::memcpy(out, r.res.data(), r.res.size() * sizeof(vd_refined));
  });
}

int vd_select_modes(vd_ctx* c, int id, int n, const vd_refined* r, const float* coords, const vd_post_params* p,
                    vd_mode* modes) {
  if (!c || n < 0 || (n > 0 && (!r || !coords)) || !modes) return VD_ERR_ARG;
  int count = 0;
  const int rc = guard(c, [&]() {
    check_post(p);
    std::vector<vd_mode> m = select_modes(c, lig_at(c, id), n, r, coords, p);
    std::copy(m.begin(), m.end(), modes);
    count = (int)m.size();
  });
  return rc == VD_OK ? count : rc;
}

int vd_postprocess_batch(vd_ctx* c, int n, const int* ids, const int* poses, const float* conf, const vd_post_params* p,
                         vd_stats* st) {
  if (!c || n < 0 || (n > 0 && (!ids || !poses))) return VD_ERR_ARG;
  return guard(c, [&]() {
    std::vector<int> vi(ids, ids + n), vp(poses, poses + n);
    long long tot = 0;
    for (int k = 0; k < n; ++k) tot += vp[k];
    if (tot > 0 && !conf) throw ArgError(VD_ERR_ARG, "conf is NULL");
    postprocess(c, vi, vp, conf, p, st);
  });
}

int vd_postprocess_docked(vd_ctx* c, int max_poses, const vd_post_params* p, vd_stats* st) {
  if (!c || max_poses < 1) return VD_ERR_ARG;
  return guard(c, [&]() {
    need_device(c);
    BatchState& b = c->batch;
    const int nl = (int)b.ids.size();
    c->pin_e.ensure(b.total.x); c->pin_conf.ensure(b.total.y); c->pin_heavy.ensure(b.total.z);
    const float* e = c->pin_e.p;
    const float* cf = c->pin_conf.p;
    const float* hv = c->pin_heavy.p;
    if (b.total.x) HIP_CHECK(hipMemcpyAsync(c->pin_e.p, b.e.p, b.total.x * 4, hipMemcpyDeviceToHost, c->stream));
    if (b.total.y) HIP_CHECK(hipMemcpyAsync(c->pin_conf.p, b.conf.p, b.total.y * 4, hipMemcpyDeviceToHost, c->stream));
    if (b.total.z) HIP_CHECK(hipMemcpyAsync(c->pin_heavy.p, b.heavy.p, b.total.z * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<std::vector<int>> kept(nl);
    parallel_for(nl, [&](int k) {
      const longlong4 o = b.off[k];
      kept[k] = vd::cluster_poses(b.chains[k], b.H[k], e + o.x, hv + o.z, 1.0, max_poses);
    });
    std::vector<int> ids(b.ids), poses(nl);
    std::vector<float> conf;
    for (int k = 0; k < nl; ++k) {
      const int T = b.T[k];
      poses[k] = (int)kept[k].size();
      const float* base = cf + b.off[k].y;
      for (int i : kept[k]) conf.insert(conf.end(), base + (size_t)i * (7 + T), base + (size_t)(i + 1) * (7 + T));
    }
    postprocess(c, ids, poses, conf.data(), p, st);
  });
}

int vd_post_results(vd_ctx* c, int k, int max_modes, vd_mode* modes, float* conf, float* coords) {
  if (!c || max_modes < 0) return VD_ERR_ARG;
  int count = 0;
  const int rc = guard(c, [&]() {
    if (k < 0 || k >= (int)c->post.ligs.size()) throw ArgError(VD_ERR_ARG, "post-processing index out of range");
    const PostLigand& pl = c->post.ligs[k];
    count = (int)pl.modes.size();
    const size_t m = (size_t)std::min(count, max_modes);
    if (m == 0) return;
    const size_t per_conf = pl.conf.size() / pl.modes.size(), per_xyz = pl.coords.size() / pl.modes.size();
    if (modes) std::copy(pl.modes.begin(), pl.modes.begin() + m, modes);
    if (conf) std::copy(pl.conf.begin(), pl.conf.begin() + m * per_conf, conf);
    if (coords) std::copy(pl.coords.begin(), pl.coords.begin() + m * per_xyz, coords);
  });
  return rc == VD_OK ? count : rc;
}

int vd_write_pdbqt(vd_ctx* c, int id, int n, const vd_mode* modes, const float* coords, char* out, int cap) {
  if (!c || n < 0 || (n > 0 && (!modes || !coords)) || cap < 0 || (cap > 0 && !out)) return VD_ERR_ARG;
  int len = 0;
  const int rc = guard(c, [&]() {
    std::vector<double> e(n), lb(n), ub(n);
    for (int i = 0; i < n; ++i) { e[i] = modes[i].e; lb[i] = modes[i].rmsd_lb; ub[i] = modes[i].rmsd_ub; }
    const std::string s = vd::write_models(lig_at(c, id), n, e.data(), lb.data(), ub.data(), coords);
    if (s.size() > (size_t)0x7FFFFFFF) throw ArgError(VD_ERR_CAPACITY, "output too large");
    len = (int)s.size();
    if (cap > 0) {
      const size_t m = std::min<size_t>(s.size(), (size_t)cap - 1);
      std::memcpy(out, s.data(), m);
      out[m] = '\0';
    }
  });
  return rc == VD_OK ? len : rc;
}

int vd_device_alloc(vd_ctx* c, size_t bytes, void** ptr) {
  if (!c || !ptr) return VD_ERR_ARG;
  return guard(c, [&]() { need_device(c); *ptr = nullptr; HIP_CHECK(hipMalloc(ptr, std::max<size_t>(bytes, 1))); });
}

int vd_device_free(vd_ctx* c, void* ptr) {
  if (!c) return VD_ERR_ARG;
  return guard(c, [&]() { need_device(c); if (ptr) HIP_CHECK(hipFree(ptr)); });
}

int vd_copy(vd_ctx* c, void* dst, const void* src, size_t bytes, int direction) {
  if (!c || (bytes && (!dst || !src))) return VD_ERR_ARG;
  return guard(c, [&]() {
    need_device(c);
    const hipMemcpyKind k = direction == 0 ? hipMemcpyHostToDevice : direction == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int vd_random_confs(vd_ctx* c, int id, long long n, uint32_t seed, float* d_conf) {
  if (!c || n < 0 || (n > 0 && !d_conf)) return VD_ERR_ARG;
  return guard(c, [&]() {
    need_device(c);
    if (!c->maps_ready) throw ArgError(VD_ERR_ARG, "no box: populate or upload maps first");
    const vd::Ligand& L = lig_at(c, id);
    float lo[3], hi[3];
    for (int i = 0; i < 3; ++i) { lo[i] = (float)c->box.begin[i]; hi[i] = (float)c->box.end[i]; }
    if (n) HIP_CHECK(vdl::random_confs(d_conf, n, L.n_torsions(), seed, lo, hi, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int vd_synth_receptor(int n_atoms, double cube, double cavity_radius, uint64_t seed, float* xyz, int* xs) {
  if (n_atoms < 0 || !xyz || !xs || !(cube > 0)) return VD_ERR_ARG;
  std::vector<float> x; std::vector<int> t;
  vd::synth_receptor(n_atoms, cube, cavity_radius, seed, x, t);
  std::memcpy(xyz, x.data(), x.size() * 4);
  std::memcpy(xs, t.data(), t.size() * 4);
  return (int)t.size();
}

int vd_synth_ligand(int n_heavy, int n_rot, uint64_t seed, char* out, int cap) {
  if (n_heavy < 1 || n_rot < 0) return VD_ERR_ARG;
  std::string s = vd::synth_ligand_pdbqt(n_heavy, n_rot, seed);
  if (out && (int)s.size() < cap) { std::memcpy(out, s.c_str(), s.size() + 1); }
  return (int)s.size() + 1;
}

}  // extern "C"
