// kernels.hip -- HIP kernels of libvdock for gfx950 and their launch wrappers.
//
//  vd_grid_kernel     17-map precompute (replaces kernel1/kernel_grid, V11/inc/cuda/kernel1.cu:93-249;
//                     CPU semantics cache::populate, V11/lib/cache.cpp:148-208)
//  vd_eval_kernel     energy + adjoint gradient of given conformations (m_eval_deriv_update,
//                     V11/inc/cuda/quasi_newton.cuh:1294-1339; model::eval_deriv, model.cpp:1324-1334)
//  vd_mc_kernel       Monte Carlo + local search chains (kernel2_update, V11/inc/cuda/kernel2.cu:219-564),
//                     either one work item per workgroup (plain per-ligand launch) or a persistent
//                     grid pulling (ligand, chain block) items from a device queue.
//  vd_refine_kernel   refinement of the kept poses against the receptor atoms and the parts of
//                     their final score (refine_structure + eval_adjusted, V11/main/main.cpp:139-157,
//                     425-440; the reference runs these on the CPU).
#include <hip/hip_runtime.h>
#include "dock_device.hpp"
#include "refine_device.hpp"
#include "kernels.h"

using namespace vdk;

namespace {

// ------------------------------------------------------------------------------------------
// Grid maps.  A workgroup is one wavefront and owns a compact tile of points (x fastest), LP
// lanes per point: lane l works on point l / LP and on the maps t2 = (l % LP) * 16/LP ..
// +16/LP-1, the first lane of a point also on t2 = 16.  The receptor atoms that can reach the tile
// come from a host-built cell list (cells of VD_GCELL A over the box grown by the 8 A cutoff, one
// contiguous atom range per x-row of cells).  Each range is read 64 atoms at a time, one per lane
// in a single coalesced load; a ballot of the lanes whose atom lies within the cutoff of the tile
// box lists the atoms worth a pass, and the loop takes them in index order (GU per pass, all their
// table reads in flight before any is added), reading each atom from its lane with readlane.  For
// every atom within the cutoff of a point, one 16-float row of the map LUT (rows indexed
// [t1][bin]) and one value of the column after the rows give the contribution to all 17 maps;
// the LP lanes of the point read the row's 16-byte chunks side by side, so one load instruction
// fetches whole rows of 64 / LP points, one cache line each.  Every point adds its atoms in the
// order of the cell list, whatever the tile shape or LP, so neither changes a map value.
constexpr int GU = 2;   // receptor atoms per pass of the grid kernel's loop

__device__ __forceinline__ float lane_f(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

template <int LP> struct GridTile {
  static constexpr int NP = 64 / LP;                 // points per wavefront
  static constexpr int TX = 4, TY = NP >= 32 ? 4 : 2, TZ = NP / (TX * TY);
  static constexpr int K = VD_GRID_ROW / LP;         // row values per lane
  static constexpr int CH = K / 4;                   // 16-byte chunks per lane
};

template <int LP>
__global__ __launch_bounds__(64) void vd_grid_kernel(const float4* __restrict__ atoms, const int2* __restrict__ cells,
                                                     GridCells gc, const float4* __restrict__ lut_rows,
                                                     const float* __restrict__ lut_last, float* __restrict__ maps,
                                                     vd_grid_desc gd, int tiles_x, int tiles_y) {
  using T = GridTile<LP>;
  constexpr int TX = T::TX, TY = T::TY, TZ = T::TZ, K = T::K, CH = T::CH;
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int pt = lane / LP, q = lane % LP;
  const int tx0 = (tile % tiles_x) * TX, ty0 = ((tile / tiles_x) % tiles_y) * TY, tz0 = (tile / (tiles_x * tiles_y)) * TZ;
  const int lx = pt % TX, ly = (pt / TX) % TY, lz = pt / (TX * TY);
  const int x = tx0 + lx, y = ty0 + ly, z = tz0 + lz;
  const bool inside = x < gd.dim[0] && y < gd.dim[1] && z < gd.dim[2];
  // probe coordinates as grid::index_to_argument (V11/lib/grid.h:49-53), relative to the cell origin
  const float px = gd.init[0] + gd.factor_inv[0] * x - gc.origin[0];
  const float py = gd.init[1] + gd.factor_inv[1] * y - gc.origin[1];
  const float pz = gd.init[2] + gd.factor_inv[2] * z - gc.origin[2];
  // tile bounds (relative), grown by the cutoff, -> cell range
  const float lo_x = gd.init[0] + gd.factor_inv[0] * tx0 - gc.origin[0], hi_x = lo_x + gd.factor_inv[0] * (TX - 1);
  const float lo_y = gd.init[1] + gd.factor_inv[1] * ty0 - gc.origin[1], hi_y = lo_y + gd.factor_inv[1] * (TY - 1);
  const float lo_z = gd.init[2] + gd.factor_inv[2] * tz0 - gc.origin[2], hi_z = lo_z + gd.factor_inv[2] * (TZ - 1);
  const float cut = 8.0f, inv = 1.0f / gc.cell;
  const int cx0 = max(0, (int)floorf((lo_x - cut) * inv)), cx1 = min(gc.n[0] - 1, (int)floorf((hi_x + cut) * inv));
  const int cy0 = max(0, (int)floorf((lo_y - cut) * inv)), cy1 = min(gc.n[1] - 1, (int)floorf((hi_y + cut) * inv));
  const int cz0 = max(0, (int)floorf((lo_z - cut) * inv)), cz1 = min(gc.n[2] - 1, (int)floorf((hi_z + cut) * inv));
  const bool first = q == 0;   // this lane also sums t2 = 16
  float acc[K], acc16 = 0.0f;
#pragma unroll
  for (int t = 0; t < K; ++t) acc[t] = 0.0f;
  for (int cz = cz0; cz <= cz1; ++cz)
    for (int cy = cy0; cy <= cy1; ++cy) {
      const int row = (cz * gc.n[1] + cy) * gc.n[0];
      const int begin = cells[row + cx0].x, end = cells[row + cx1].y;   // cells of one x-row are contiguous
      for (int base = begin; base < end; base += 64) {
        // this lane's atom of the chunk, and whether it can reach the tile at all
        const int k = base + lane;
        const float4 mine = k < end ? atoms[k] : make_float4(0, 0, 0, 0);
        const float ddx = fmaxf(fmaxf(lo_x - mine.x, mine.x - hi_x), 0.0f);
        const float ddy = fmaxf(fmaxf(lo_y - mine.y, mine.y - hi_y), 0.0f);
        const float ddz = fmaxf(fmaxf(lo_z - mine.z, mine.z - hi_z), 0.0f);
        uint64_t todo = __ballot(k < end && ddx * ddx + ddy * ddy + ddz * ddz <= VD_CUTOFF_SQR);
        while (todo) {
          // the next GU atoms of the list, in index order (fewer when fewer are left)
          int j[GU];
          bool has[GU];
#pragma unroll
          for (int u = 0; u < GU; ++u) {
            has[u] = todo != 0;
            j[u] = has[u] ? __builtin_ctzll(todo) : 0;
            todo &= todo - 1;
          }
          float4 rv[GU][CH];
          float last[GU];
          bool use[GU];
#pragma unroll
          for (int u = 0; u < GU; ++u) {
            const float ax = lane_f(mine.x, j[u]), ay = lane_f(mine.y, j[u]), az = lane_f(mine.z, j[u]);
            const int t1 = __builtin_amdgcn_readlane(__float_as_int(mine.w), j[u]);
            const float dx = ax - px, dy = ay - py, dz = az - pz;
            const float r2 = dx * dx + dy * dy + dz * dz;
            use[u] = has[u] && inside && r2 <= VD_CUTOFF_SQR;
            last[u] = 0.0f;
            if (use[u]) {
              const size_t r = (size_t)t1 * VD_LUT_BINS + (int)(VD_LUT_FACTOR * r2);
              const float4* rw = lut_rows + r * (VD_GRID_ROW / 4) + q * CH;
#pragma unroll
              for (int c = 0; c < CH; ++c) rv[u][c] = rw[c];
              if (first) last[u] = lut_last[r];
            }
          }
#pragma unroll
          for (int u = 0; u < GU; ++u) {
            if (!use[u]) continue;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              acc[4 * c] += rv[u][c].x; acc[4 * c + 1] += rv[u][c].y;
              acc[4 * c + 2] += rv[u][c].z; acc[4 * c + 3] += rv[u][c].w;
            }
            acc16 += last[u];
          }
        }
      }
    }
  if (inside) {
    const size_t o = (size_t)x + (size_t)gd.dim[0] * ((size_t)y + (size_t)gd.dim[1] * z);
#pragma unroll
    for (int t = 0; t < K; ++t) maps[(size_t)(q * K + t) * gd.plane + o] = acc[t];
    if (first) maps[(size_t)VD_GRID_ROW * gd.plane + o] = acc16;
  }
}

// Corner records of the maps (the layout the searches read, described at EvalK): thread per
// (type, cell), x fastest; reads eight values of the canonical maps, writes 32 bytes.
__global__ __launch_bounds__(256) void vd_corner_kernel(const float* __restrict__ maps, float4* __restrict__ corners,
                                                        vd_grid_desc gd, long long total) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int c0 = gd.dim[0] - 1, c1 = gd.dim[1] - 1, c2 = gd.dim[2] - 1;
  const long long per_type = (long long)c0 * c1 * c2;
  const int t = (int)(id / per_type);
  long long r = id - (long long)t * per_type;
  const int i = (int)(r % c0); r /= c0;
  const int j = (int)(r % c1);
  const int k = (int)(r / c1);
  const long long d0 = gd.dim[0], d01 = (long long)gd.dim[0] * gd.dim[1];
  const float* b = maps + (long long)t * gd.plane + i + d0 * j + d01 * k;
  corners[2 * id] = make_float4(b[0], b[1], b[d0], b[d0 + 1]);
  corners[2 * id + 1] = make_float4(b[d01], b[d01 + 1], b[d01 + d0], b[d01 + d0 + 1]);
}

// ------------------------------------------------------------------------------------------
// LDS layout of one wave: ligand (atoms 4A, node_o 4N, node_a 4N floats), the conformation
// stash, then per chain group the node area 12N, xyz 4A, frc 4A, the pair-force exchange row
// 4(G+1), and for dense BFGS H G*G + 2G.  The node area holds the frame rows from the kinematics
// until the atoms are placed, then the per-node force and moment sums of the adjoint (8N): an
// evaluation reads the frames only in its intermolecular pass, which ends (with a fence) before
// the sums are cleared.  The pair words stay in global memory (read once per evaluation,
// coalesced, L2-resident).
__host__ __device__ inline int lig_floats(int A, int N) { return 4 * A + 8 * N + kStashSlots * kStashSlotFloats; }
__host__ __device__ inline int group_floats(int A, int N, int G, int dense) {
  return 12 * N + 8 * A + 4 * (G + 1) + (dense ? G * G + 2 * G : 0);
}

template <int G>
__device__ __forceinline__ void stage_ligand(const float4* __restrict__ blob, const vd_lig_header& h, float* lds, int lane) {
  float4* dst = reinterpret_cast<float4*>(lds);
  for (int i = lane; i < h.n_atoms; i += 64) dst[i] = blob[h.off_atoms + i];
  for (int i = lane; i < 2 * h.n_nodes; i += 64) dst[h.n_atoms + i] = blob[h.off_nodes + i];
}

template <int G>
__device__ __forceinline__ void make_views(const float4* __restrict__ blob, const vd_lig_header& h, float* lds, int Amax,
                                           int Nmax, int dense, int group, LigView& L, Scratch& S, float*& H) {
  float4* base = reinterpret_cast<float4*>(lds);
  L.atom = base; L.node_o = base + h.n_atoms; L.node_a = base + h.n_atoms + h.n_nodes;
  L.pairs = reinterpret_cast<const uint4*>(reinterpret_cast<const uint32_t*>(blob) + h.off_pairs);
  L.A = h.n_atoms; L.N = h.n_nodes; L.C = h.n_chunks; L.T = h.n_torsions; L.H = h.n_heavy; L.rounds = h.jump_rounds;
  L.rows = h.pair_rows;
  S.stash = lds + lig_floats(Amax, Nmax) - kStashSlots * kStashSlotFloats;
  float* g = lds + lig_floats(Amax, Nmax) + group * group_floats(Amax, Nmax, G, dense);
  S.frame = reinterpret_cast<float4*>(g);
  S.nacc = S.frame;
  S.xyz = reinterpret_cast<float4*>(g + 12 * Nmax);
  S.frc = S.xyz + Amax;
  S.fx = S.frc + Amax;
  H = reinterpret_cast<float*>(S.fx + G + 1);
}

// Writes a chain's best energy, conformation and heavy-atom coordinates (call after
// set_coords of the best conformation).
template <int G>
__device__ void write_chain(const LigView& L, const Scratch& S, const Conf& best, float best_e, uint32_t chain, int l,
                            float* out_e, float* out_conf, float* out_heavy) {
  if (l == 0) {
    out_e[chain] = best_e;
    float* c = out_conf + (size_t)chain * (7 + L.T);
    c[0] = best.px; c[1] = best.py; c[2] = best.pz; c[3] = best.q.w; c[4] = best.q.x; c[5] = best.q.y; c[6] = best.q.z;
  }
  if (l >= 1 && l <= L.T) out_conf[(size_t)chain * (7 + L.T) + 6 + l] = best.tor;
  for (int i = l; i < L.A; i += G) {
    const uint32_t meta = fbits(L.atom[i].w);
    if ((meta >> 16) & 1u) {
      const int slot = (meta >> 17) & 255;
      float* o = out_heavy + ((size_t)chain * L.H + slot) * 3;
      const float4 y = S.xyz[i];
      o[0] = best.px + y.x; o[1] = best.py + y.y; o[2] = best.pz + y.z;
    }
  }
}

// Initial conformation of a chain: the caller's (the role of kernel_monte's ric,
// V11/inc/cuda/kernel2.cuh:27; the reference host draws them at main_procedure_cu.cpp:584-635), or
// conf::randomize (V11/lib/conf.h:127-130, 214-217; random_orientation, V11/lib/quaternion.cpp:206-219)
// from the shared random stream.  `given` points at the chain's 7+T numbers or is null.
template <int G>
__device__ Conf initial_conf(const vd_mc_consts& P, uint32_t stream, uint32_t chain, int l, int T,
                             const float* given = nullptr) {
  Conf c;
  if (given) {
    c.px = given[0]; c.py = given[1]; c.pz = given[2];
    c.q = qnormalize(Quat{given[3], given[4], given[5], given[6]});
    c.tor = (l >= 1 && l <= T) ? given[6 + l] : 0.0f;
    return c;
  }
  const U4 w = philox(chain, VD_INIT_STEP, 0, 0, P.seed, stream);
  c.px = P.box_lo[0] + (P.box_hi[0] - P.box_lo[0]) * u01(w.x);
  c.py = P.box_lo[1] + (P.box_hi[1] - P.box_lo[1]) * u01(w.y);
  c.pz = P.box_lo[2] + (P.box_hi[2] - P.box_lo[2]) * u01(w.z);
  c.q = Quat{1, 0, 0, 0};
  bool done = false;
  for (uint32_t a = 0; a < 8 && !done; ++a) {
    const U4 v = philox(chain, VD_INIT_STEP, 1, a, P.seed, stream);
    const float r0 = sqrtf(-2.0f * logf(u01(v.x) + 1.0f / 16777216.0f)), r1 = sqrtf(-2.0f * logf(u01(v.z) + 1.0f / 16777216.0f));
    float s0, c0, s1, c1;
    sincosf(2.0f * kPi * u01(v.y), &s0, &c0); sincosf(2.0f * kPi * u01(v.w), &s1, &c1);
    const float n0 = r0 * c0, n1 = r0 * s0, n2 = r1 * c1, n3 = r1 * s1;
    const float nrm = sqrtf(n0 * n0 + n1 * n1 + n2 * n2 + n3 * n3);
    if (nrm > 1e-6f) { c.q = Quat{n0 / nrm, n1 / nrm, n2 / nrm, n3 / nrm}; done = true; }
  }
  c.tor = 0.0f;
  if (l >= 1 && l <= T) {
    const int j = l - 1;
    const U4 v = philox(chain, VD_INIT_STEP, 2 + (uint32_t)(j / 4), 0, P.seed, stream);
    const uint32_t word = (j % 4 == 0) ? v.x : (j % 4 == 1) ? v.y : (j % 4 == 2) ? v.z : v.w;
    c.tor = -kPi + 2.0f * kPi * u01(word);
  }
  return c;
}

// One MC chain per group with the reference GPU schedule (V11/inc/cuda/kernel2.cu:219-426):
// mutate -> local search -> Metropolis(T) -> on accept, if better than the best so far: second
// local search, keep the best.  Mutation as V11/lib/mutate.cpp:62-127 (see initial_conf for the
// random stream); the orientation step uses the gyration radius of the current conformation.
// ls_only: one local search from the starting conformation, no MC steps (vd_local_search).
// The local search is RILC-BFGS with the Lewis-Overton line search (quasi_newton.cuh:1710-1806,
// 2172-2389) or, with POL = -1, dense BFGS with its backtracking line search (--rilc_bfgs 0,
// kernel2.cu:427-563 with bfgs_optimized, quasi_newton.cuh:2391-2499; H is the group's matrix,
// see dense_minus_hv).  It runs as a flat state machine: every pass evaluates one conformation
// per chain group and each group then advances its own state (start of a local search,
// line-search trial, end of a BFGS iteration, end of a local search, Metropolis step, next
// mutation) as far as it can without another evaluation.  Chains of a wavefront thus never wait
// for each other's line searches or local searches; they only wait at the end of the chain.  The
// gyration radius of each evaluated conformation comes from the evaluation itself and travels
// with the conformation.
// POL: -1 dense BFGS; 0 Vina; 1 QuickVina2 gate at the start of every local search (see
// History); 2 QuickVina-W gate (qvw_interesting) with its records: the searched start point, each
// BFGS iteration's end point and the point where the search stops go to the chain's own buffer,
// the search's last point to the ligand's global buffer (QVW/inc/cuda/quasi_newton.cuh:2236-2410).
template <int G, int POL>
__device__ void mc_chain_flat(const LigView& L, const Scratch& S, const EvalK& K, const vd_mc_consts& P, int depth,
                              int max_steps, uint32_t stream, uint32_t chain, bool valid, int l, float* out_e,
                              float* out_conf, float* out_heavy, int& evals, History* hist, const GlobalHist* gh,
                              const float* given, bool ls_only, float* H) {
  const uint32_t k0 = P.seed, k1 = stream;
  const int n_ent = P.mutation_quirk ? (1 + L.T) : (2 + L.T);
  enum { kMutate = 0, kStart = 1, kTrial = 2, kDone = 3 };
  // MC state; the current conformation lives in the stash (kSlotTmp)
  {
    const Conf c0 = initial_conf<G>(P, stream, chain, l, L.T, given);
    stash_put<G>(S.stash, kSlotTmp, c0);
    stash_put<G>(S.stash, kSlotBest, c0);
    set_coords<G>(L, S, c0, l);
  }
  float rg = gyration_radius<G>(L, S, l);
  float tmp_e = 0.0f, best_e = INFINITY, u_metro = 0.0f;
  // ls_kind: 1 the search of a mutated candidate, 2 the second search of an accepted one,
  // 0 a lone search from the starting conformation (ls_only)
  int step = 0, phase = (valid && depth > 0) ? kMutate : kDone, ls_kind = 1;
  if (ls_only) {
    phase = valid ? kStart : kDone;
    ls_kind = 0;
  }
  // local-search state
  Conf x = stash_get<G>(S.stash, kSlotTmp);
  float x_rg = rg, g = 0.0f, d = 0.0f, gp = 0.0f;
  float fx = 0.0f, fxp = 0.0f, fx_orig = 0.0f, orig_rg = 0.0f, prev_rg = 0.0f;
  float stp = 0.0f, finit = 0.0f, dgtest = 0.0f, dstest = 0.0f, mu = 0.0f, nu = 0.0f;
  bool brackt = false;
  int count = 0, k = 0;
  float dn_pg = 0.0f;   // dense BFGS: p.g of the iteration's direction
  int dn_step = 0;      // dense BFGS: iterations done in this search
  while (__any(phase != kDone)) {
    // (1) MC step start: draws and mutation (mutate.cpp:62-127); the candidate is evaluated next
    const bool mut = phase == kMutate;
    if (__any(mut)) {
      const U4 w = philox(chain, (uint32_t)step, 0, 0, k0, k1);
      const int which = min((int)(u01(w.x) * (float)n_ent), n_ent - 1);
      const float tor = -kPi + 2.0f * kPi * u01(w.z);
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
      bool need = mut && which <= 1;
      for (uint32_t a = 0; a < 64 && __any(need); ++a) {   // P(64 rejections) ~ 1e-20
        const U4 v = philox(chain, (uint32_t)step, 1, a, k0, k1);
        const float vx = 2.0f * u01(v.x) - 1.0f, vy = 2.0f * u01(v.y) - 1.0f, vz = 2.0f * u01(v.z) - 1.0f;
        if (need && vx * vx + vy * vy + vz * vz < 1.0f) { sx = vx; sy = vy; sz = vz; need = false; }
      }
      if (mut) {
        Conf cand = stash_get<G>(S.stash, kSlotTmp);
        if (which == 0) { cand.px += P.amplitude * sx; cand.py += P.amplitude * sy; cand.pz += P.amplitude * sz; }
        else if (which == 1) { if (rg > kEpsFl) { const float f = P.amplitude / rg; cand.q = qincrement(cand.q, f * sx, f * sy, f * sz); } }
        else if (l == which - 1) cand.tor = tor;
        x = cand; u_metro = u01(w.y); ls_kind = 1; phase = kStart;
      }
    }
    // (2) one evaluation per group: the search start point, or the line-search trial point
    const bool trial = phase == kTrial, start = phase == kStart;
    const Conf xt = increment<G>(stash_get<G>(S.stash, kSlotPrev), d, stp, l, L.T);
    const Conf xe = csel(trial, xt, x);
    float ge, rge;
    const float fe = eval_deriv<G>(L, S, K, xe, l, ge, &rge);
    // (3) consume it
    bool ls_done = false, iter_end = false, begin = false;
    int result = -1;
    if (start || trial) { ++evals; x = xe; fx = fe; g = ge; x_rg = rge; }
    if (trial) ++count;
    if (start) { fx_orig = fe; orig_rg = rge; stash_put<G>(S.stash, kSlotOrig, x); }
    bool search = start;
    if constexpr (POL == 1) {
      if (__any(start)) {
        const bool interesting = hist_interesting<G>(*hist, x, fx, g, l, L.T, start);
        hist_add<G>(*hist, x, fx, g, l, L.T, start && interesting);
        search = start && interesting;
      }
    }
    if constexpr (POL == 2) {
      if (__any(start)) {
        const bool interesting = qvw_interesting<G>(*hist, *gh, x, fx, g, l, L.T, start);
        hist_add<G>(*hist, x, fx, g, l, L.T, start && interesting);
        search = start && interesting;
      }
    }
    bool rec_own = false, rec_glob = false;   // QuickVina-W records of this pass
    if constexpr (POL < 0) {   // dense BFGS (quasi_newton.cuh:2391-2499); x, fx, g hold the point evaluated
      // a search starts from the identity matrix; no step limit -> the start point is the result
      dense_identity<G>(H, l, L.T, start);
      if (start) { dn_step = 0; if (max_steps > 0) begin = true; else ls_done = true; }
      // backtracking: accept when f1 - f0 < 1e-4 alpha p.g, else halve alpha; after ten trials the
      // last point is taken (alpha halved once more)
      if (trial) {
        const bool ok = fx - fxp < 1e-4f * stp * dn_pg;
        if (!ok) stp *= 0.5f;
        if (ok || count >= 10) iter_end = true;
      }
      if (__any(iter_end)) {   // end of an iteration: the step is taken; the update
        const float y = g - gp;
        const float gnorm = sqrtf(gsum<G>(gp * gp));
        bool upd = false;
        if (iter_end) { if (!(gnorm >= 1e-5f)) ls_done = true; else upd = true; }
        const float yy = gsum<G>(y * y), yp = gsum<G>(y * d);
        dense_set_diag<G>(H, stp * yp / yy, l, L.T, iter_end && dn_step == 0 && upd && fabsf(yy) > kEpsFl);   // first step: scaled diagonal
        const float mhy = dense_minus_hv<G>(H, H + G * G, y, l, L.T);
        const float yhy = -gsum<G>(y * mhy);
        const float r = 1.0f / (stp * yp);
        dense_update<G>(H, H + G * G, mhy, d, stp, r, yhy, iter_end && upd && !(stp * yp < kEpsFl), l, L.T);
        if (iter_end && !ls_done) { ++dn_step; if (dn_step < max_steps) begin = true; else ls_done = true; }
      }
      if (__any(begin)) {      // an iteration starts at x: p = -H g
        if (begin) { stash_put<G>(S.stash, kSlotPrev, x); gp = g; fxp = fx; prev_rg = x_rg; }
        const float pn = dense_minus_hv<G>(H, H + G * G, g, l, L.T);
        const float pg = gsum<G>(pn * g);
        if (begin) { d = pn; dn_pg = pg; stp = 1.0f; count = 0; phase = kTrial; }
      }
    }
    if constexpr (POL >= 0) {
    if (__any(start)) {
      const float gg = gsum<G>(g * g);
      if (start) {
        d = -g;
        if (search && sqrtf(gg) >= 1e-5f) { stp = 1.0f / sqrtf(gg); k = 1; begin = true; }
        else ls_done = true;
      }
    }
    if (__any(trial)) {   // Lewis-Overton line search (quasi_newton.cuh:2172-2389)
      const float gtd = gsum<G>(g * d);
      if (trial) {
        bool done = false;
        if (fx > finit + stp * dgtest) { nu = stp; brackt = true; }
        else if (gtd < dstest) { mu = stp; }
        else { result = count; done = true; }
        if (!done && count >= 10) { result = (fx > finit) ? -1 : count; done = true; }
        if (!done) stp = brackt ? 0.5f * (mu + nu) : 2.0f * stp;
        iter_end = done;
      }
    }
    if (__any(iter_end)) {   // end of a BFGS iteration, direction update (quasi_newton.cuh:2342-2350)
      const float gnorm = sqrtf(gsum<G>(g * g));
      bool update = false;
      if (iter_end) {
        if (result < 0) {
          x = stash_get<G>(S.stash, kSlotPrev); g = gp; fx = fxp; x_rg = prev_rg; ls_done = true;
          rec_own = rec_glob = true;
        }
        else if (!(gnorm >= 1e-5f)) ls_done = true;
        else if ((max_steps != 0 && max_steps <= k) || k >= 1000) { ls_done = true; rec_own = rec_glob = true; }   // 1000: hard bound
        else { ++k; update = true; rec_own = true; }
      }
      const float sv = stp * d, y = g - gp;
      const float ys = gsum<G>(y * sv), yy = gsum<G>(y * y), ss = gsum<G>(sv * sv), gpgp = gsum<G>(gp * gp);
      const float nd = -g;
      const float alpha = gsum<G>(sv * nd) / ys;
      const float nd2 = (nd + alpha * y) * (ys / yy);
      const float beta = gsum<G>(y * nd2) / ys;
      const float nd3 = nd2 + (alpha - beta) * sv;
      if (update) { d = (ys > ss * sqrtf(gpgp) * 1.0e-6f) ? nd3 : nd; stp = 1.0f; begin = true; }
    }
    if (__any(begin)) {      // a BFGS iteration starts: its line search begins at x along d
      if (begin) { stash_put<G>(S.stash, kSlotPrev, x); gp = g; fxp = fx; prev_rg = x_rg; }
      const float dginit = gsum<G>(g * d);
      if (begin) {
        if (0.0f < dginit) { ls_done = true; rec_own = rec_glob = true; }   // no descent: the iteration fails, x stays
        else {
          finit = fx; dgtest = 1e-4f * dginit; dstest = 0.1f * dginit;
          mu = 0.0f; nu = 1.0e20f; brackt = false; count = 0; phase = kTrial;
        }
      }
    }
    }   // RILC-BFGS
    if constexpr (POL == 2) {
      if (__any(rec_own)) hist_add<G>(*hist, x, fx, g, l, L.T, rec_own);
      if (__any(rec_glob)) glob_add<G>(*gh, x, fx, g, l, L.T, rec_glob);
    }
    if (ls_done) {           // end of a local search: keep the better of its start and its result
      if (!(fx <= fx_orig)) { fx = fx_orig; x = stash_get<G>(S.stash, kSlotOrig); x_rg = orig_rg; }
      bool step_done = true;
      if (ls_kind == 0) {    // a lone local search: its result is the chain's output
        stash_put<G>(S.stash, kSlotBest, x); best_e = fx; phase = kDone; step_done = false;
      } else if (ls_kind == 1) {    // Metropolis (kernel2.cu:364-392)
        const bool accept = step == 0 || fx < tmp_e || u_metro < expf((tmp_e - fx) / P.temperature);
        if (accept) { stash_put<G>(S.stash, kSlotTmp, x); tmp_e = fx; rg = x_rg; }
        if (accept && tmp_e < best_e) { ls_kind = 2; phase = kStart; step_done = false; }   // second search of x
      } else {
        stash_put<G>(S.stash, kSlotTmp, x); tmp_e = fx; rg = x_rg;
        if (tmp_e < best_e) { stash_put<G>(S.stash, kSlotBest, x); best_e = tmp_e; }
      }
      if (step_done) { ++step; phase = step < depth ? kMutate : kDone; }
    }
  }
  const Conf best = stash_get<G>(S.stash, kSlotBest);
  set_coords<G>(L, S, best, l);
  if (valid) write_chain<G>(L, S, best, best_e, chain, l, out_e, out_conf, out_heavy);
}

// W = minimum waves per SIMD the register allocation must allow (2 = unconstrained); a tuning knob:
// more resident waves hide more latency, at the price of register spills.
template <int G, int MODE, int W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, 8))) void vd_mc_kernel(McArgs a) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  constexpr int CPW = 64 / G;
  const int lane = threadIdx.x, group = lane / G, l = lane % G;
  int cur = -1;
  for (int it = 0;; ++it) {
    int item;
    if (a.queue) {
      int v = 0;
      if (lane == 0) v = atomicAdd(a.queue, 1);
      item = __shfl(v, 0);
    } else {
      item = (it == 0) ? (int)blockIdx.x : a.n_items;
    }
    if (item >= a.n_items) break;
    // (ligand slot, first chain of the block); without an item list, one ligand's chain blocks in order
    const int2 wi = a.items ? a.items[item] : make_int2(0, item * CPW);
    const vd_lig_header h = a.headers[wi.x];
    if (wi.x != cur) {
      __syncthreads();
      stage_ligand<G>(a.blob, h, lds, lane);
      __syncthreads();
      cur = wi.x;
    }
    LigView L; Scratch S; float* H;
    make_views<G>(a.blob, h, lds, a.Amax, a.Nmax, MODE == 0, group, L, S, H);
    const uint32_t chain = (uint32_t)(wi.y + group);
    const bool valid = (int)chain < a.P.chains;
    const float* given = (a.init_conf && valid) ? a.init_conf + a.out_off[wi.x].y + (size_t)chain * (7 + h.n_torsions)
                                                : nullptr;
    int evals = 0;
    History hist;
    if constexpr (MODE == 2) hist.init(a.hist + ((long long)blockIdx.x * CPW + group) * a.hist_stride, h.n_torsions);
    if constexpr (MODE == 3) hist.init(a.hist + ((long long)blockIdx.x * CPW + group) * a.hist_stride, h.n_torsions, 5);
    GlobalHist gh{};
    if constexpr (MODE == 3) {
      gh.base = a.glob + a.out_off[wi.x].w; gh.per_axis = h.reserved; gh.T = h.n_torsions;
      for (int k = 0; k < 3; ++k) { gh.lo[k] = a.P.box_lo[k]; gh.hi[k] = a.P.box_hi[k]; }
    }
    mc_chain_flat<G, MODE - 1>(L, S, a.K, a.P, h.search_depth, h.bfgs_steps, a.streams[wi.x], chain, valid, l,
                               a.out_e + a.out_off[wi.x].x, a.out_conf + a.out_off[wi.x].y,
                               a.out_heavy + a.out_off[wi.x].z, evals, &hist, &gh, given, a.ls_only != 0, H);
    if (a.slot_evals) {
      int sum = 0;
      for (int gi = 0; gi < CPW; ++gi) sum += __shfl(evals, gi * G);
      if (lane == 0) atomicAdd(a.slot_evals + wi.x, (unsigned long long)sum);
    }
  }
}

template <int G>
__global__ __launch_bounds__(64) void vd_eval_kernel(EvalArgs a) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  constexpr int CPW = 64 / G;
  const int lane = threadIdx.x, group = lane / G, l = lane % G;
  const vd_lig_header h = *a.header;
  stage_ligand<G>(a.blob, h, lds, lane);
  __syncthreads();
  LigView L; Scratch S; float* H;
  make_views<G>(a.blob, h, lds, h.n_atoms, h.n_nodes, 0, group, L, S, H);
  const int T = h.n_torsions, nc = 7 + T, ng = 6 + T;
  for (long long base = (long long)blockIdx.x * CPW; base < a.n_conf; base += (long long)gridDim.x * CPW) {
    const long long c = base + group;
    const bool valid = c < a.n_conf;
    const float* x = a.conf + (valid ? c : 0) * nc;
    Conf cf;
    cf.px = x[0]; cf.py = x[1]; cf.pz = x[2]; cf.q = Quat{x[3], x[4], x[5], x[6]};
    cf.tor = (l >= 1 && l <= T) ? x[6 + l] : 0.0f;
    float g;
    const float e = eval_deriv<G>(L, S, a.K, cf, l, g);
    if (valid) {
      if (l == 0) a.e[c] = e;
      float* go = a.grad + c * ng;
      if (l >= 1 && l <= T) go[5 + l] = g;
      if (l >= G - 6) go[l - (G - 6)] = g;
    }
  }
}

// ------------------------------------------------------------------------------------------
// After the search: refine_structure (V11/main/main.cpp:139-157) of one pose per workgroup (one
// wavefront), in FP64: dense BFGS on the receptor atoms themselves with the out-of-box slope
// raised 1e2 -> 1e10 until the heavy atoms lie within the box (a pose that never gets inside
// reports max_fl), then the parts of the final score at the refined pose (fast tables, slope
// a.final_slope).
__global__ __launch_bounds__(64) void vd_refine_kernel(RefineArgs a) {
  extern __shared__ double lds_d[];
  const int pose = blockIdx.x, l = threadIdx.x;
  const vdr::RefHeader h = a.hdr[a.pose_lig[pose]];
  vdr::Pose P;
  vdr::carve(P, lds_d, a.Amax, a.Nmax);
  P.node_o = a.geo + h.off_atoms + h.A;
  P.node_a = a.geo + h.off_atoms + h.A + h.N;
  P.pairs = a.pairs + h.off_pairs;
  P.A = h.A; P.N = h.N; P.T = h.T; P.C = h.C; P.n = 6 + h.T;
  for (int i = l; i < h.A; i += 64) const_cast<double4*>(P.atom)[i] = a.geo[h.off_atoms + i];
  vdr::load_links(P, l);
  const double* ci = a.conf_in + a.conf_off[pose];
  vdr::Conf x;
  x.pos = vdr::D3{ci[0], ci[1], ci[2]};
  x.q = vdr::DQ{ci[3], ci[4], ci[5], ci[6]};
  x.tor = (l >= 6 && l < 6 + h.T) ? ci[7 + (l - 6)] : 0.0;
  __syncthreads();
  double e = 0, slope = 100;
  bool inside = false;
  int evals = 0;
  for (int p = 0; p < 5; ++p, slope *= 100) {
    e = vdr::bfgs(P, a.K, slope, x, h.max_steps, l, evals);
    vdr::set_coords(P, x, l);
    inside = vdr::within(P, a.K, l);
    if (inside) break;
  }
  if (!inside) e = vdr::kMaxFl;
  const double e_inter = vdr::lane_sum<64>(vdr::nc_inter<true>(P, a.K, a.final_slope, l));
  const double e_intra = vdr::lane_sum<64>(vdr::pair_terms<true>(P, a.K, l));
  double* co = a.conf_out + a.conf_off[pose];
  if (l == 0) {
    co[0] = x.pos.x; co[1] = x.pos.y; co[2] = x.pos.z; co[3] = x.q.a; co[4] = x.q.b; co[5] = x.q.c; co[6] = x.q.d;
    a.res[pose] = make_double4(e, e_inter, e_intra, inside ? 1.0 : 0.0);
    a.evals[pose] = evals;
  }
  if (l >= 6 && l < 6 + h.T) co[7 + (l - 6)] = x.tor;
  double* xyz = a.coords_out + a.coords_off[pose];
  for (int i = l; i < h.A; i += 64) {
    const double4 y = P.xyz[i];
    xyz[3 * i] = y.x; xyz[3 * i + 1] = y.y; xyz[3 * i + 2] = y.z;
  }
}

}  // namespace

// Device memory of the no-allocation docking entry point (vd_dock_noalloc): one array per device,
// held until the process ends.  Host code never passes the array's name to a kernel or a copy: it
// takes the device address from scratch_address once per device and uses that.
static __device__ __align__(256) unsigned char scratch[VD_NOALLOC_SCRATCH_BYTES];

namespace vdl {

hipError_t scratch_address(void** p) { return hipGetSymbolAddress(p, HIP_SYMBOL(scratch)); }

int refine_lds_bytes(int Amax, int Nmax, int nmax) { return 8 * vdr::pose_lds_doubles(Amax, Nmax, nmax); }

hipError_t launch_refine(const RefineArgs& a, int poses, hipStream_t s) {
  hipLaunchKernelGGL(vd_refine_kernel, dim3(poses), dim3(64), refine_lds_bytes(a.Amax, a.Nmax, a.nmax), s, a);
  return hipGetLastError();
}

int mc_lds_bytes(int G, int Amax, int Nmax, int dense) {
  return 4 * (lig_floats(Amax, Nmax) + (64 / G) * group_floats(Amax, Nmax, G, dense));
}

// Lanes per grid point of the grid kernel.
constexpr int kGridLanesPerPoint = 2;

hipError_t launch_grid(const float4* atoms, const int2* cells, const GridCells& gc, const float4* lut_rows, float* maps,
                       const vd_grid_desc& gd, hipStream_t s) {
  using T = GridTile<kGridLanesPerPoint>;
  const int tx = (gd.dim[0] + T::TX - 1) / T::TX, ty = (gd.dim[1] + T::TY - 1) / T::TY, tz = (gd.dim[2] + T::TZ - 1) / T::TZ;
  const float* lut_last = reinterpret_cast<const float*>(lut_rows) + (size_t)VD_XS_SIZE * VD_LUT_BINS * VD_GRID_ROW;
  hipLaunchKernelGGL(vd_grid_kernel<kGridLanesPerPoint>, dim3(tx * ty * tz), dim3(64), 0, s, atoms, cells, gc, lut_rows,
                     lut_last, maps, gd, tx, ty);
  return hipGetLastError();
}

hipError_t launch_corners(const float* maps, float4* corners, const vd_grid_desc& gd, hipStream_t s) {
  const long long total = (long long)VD_XS_SIZE * (gd.dim[0] - 1) * (gd.dim[1] - 1) * (gd.dim[2] - 1);
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(vd_corner_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, maps, corners, gd, total);
  return hipGetLastError();
}

hipError_t launch_eval(int G, const EvalArgs& a, int blocks, hipStream_t s) {
  const int A = a.host_header.n_atoms, N = a.host_header.n_nodes;
  const int lds = mc_lds_bytes(G, A, N, 0);
  switch (G) {
    case 64: hipLaunchKernelGGL(vd_eval_kernel<64>, dim3(blocks), dim3(64), lds, s, a); break;
    case 32: hipLaunchKernelGGL(vd_eval_kernel<32>, dim3(blocks), dim3(64), lds, s, a); break;
    case 16: hipLaunchKernelGGL(vd_eval_kernel<16>, dim3(blocks), dim3(64), lds, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MODE, int W>
static hipError_t launch_mc_mode(int G, const McArgs& a, int blocks, int lds, hipStream_t s) {
  switch (G) {
    case 64: hipLaunchKernelGGL((vd_mc_kernel<64, MODE, W>), dim3(blocks), dim3(64), lds, s, a); break;
    case 32: hipLaunchKernelGGL((vd_mc_kernel<32, MODE, W>), dim3(blocks), dim3(64), lds, s, a); break;
    case 16: hipLaunchKernelGGL((vd_mc_kernel<16, MODE, W>), dim3(blocks), dim3(64), lds, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mc(int G, int W, const McArgs& a, int blocks, hipStream_t s) {
  const int lds = mc_lds_bytes(G, a.Amax, a.Nmax, a.P.local_search == 0);
  if (a.P.local_search == 0) return launch_mc_mode<0, 2>(G, a, blocks, lds, s);
  if (a.P.policy == 1) {
    if (!a.hist) return hipErrorInvalidValue;
    switch (W) {
      case 3: return launch_mc_mode<2, 3>(G, a, blocks, lds, s);
      case 4: return launch_mc_mode<2, 4>(G, a, blocks, lds, s);
      default: return launch_mc_mode<2, 2>(G, a, blocks, lds, s);
    }
  }
  if (a.P.policy == 2) {
    if (!a.hist || !a.glob) return hipErrorInvalidValue;
    switch (W) {
      case 3: return launch_mc_mode<3, 3>(G, a, blocks, lds, s);
      case 4: return launch_mc_mode<3, 4>(G, a, blocks, lds, s);
      default: return launch_mc_mode<3, 2>(G, a, blocks, lds, s);
    }
  }
  switch (W) {
    case 3: return launch_mc_mode<1, 3>(G, a, blocks, lds, s);
    case 4: return launch_mc_mode<1, 4>(G, a, blocks, lds, s);
    default: return launch_mc_mode<1, 2>(G, a, blocks, lds, s);
  }
}

template <int MODE, int W>
static const void* mc_fn_mode(int G) {
  switch (G) {
    case 64: return reinterpret_cast<const void*>(&vd_mc_kernel<64, MODE, W>);
    case 32: return reinterpret_cast<const void*>(&vd_mc_kernel<32, MODE, W>);
    case 16: return reinterpret_cast<const void*>(&vd_mc_kernel<16, MODE, W>);
    default: return nullptr;
  }
}

// The instantiation launch_mc launches for (G, W, a), the same dispatch.
const void* mc_kernel_fn(int G, int W, const McArgs& a) {
  if (a.P.local_search == 0) return mc_fn_mode<0, 2>(G);
  const int mode = a.P.policy == 1 ? 2 : a.P.policy == 2 ? 3 : 1;
  const int w = (W == 3 || W == 4) ? W : 2;
  switch (mode * 8 + w) {
    case 1 * 8 + 2: return mc_fn_mode<1, 2>(G);
    case 1 * 8 + 3: return mc_fn_mode<1, 3>(G);
    case 1 * 8 + 4: return mc_fn_mode<1, 4>(G);
    case 2 * 8 + 2: return mc_fn_mode<2, 2>(G);
    case 2 * 8 + 3: return mc_fn_mode<2, 3>(G);
    case 2 * 8 + 4: return mc_fn_mode<2, 4>(G);
    case 3 * 8 + 2: return mc_fn_mode<3, 2>(G);
    case 3 * 8 + 3: return mc_fn_mode<3, 3>(G);
    default: return mc_fn_mode<3, 4>(G);
  }
}

__global__ void vd_fill_int(int* p, int v) { *p = v; }
__global__ void vd_zero_u64_kernel(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}
hipError_t zero_u64(unsigned long long* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(vd_zero_u64_kernel, dim3(1), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// Clears QuickVina-W global buffers: counters 0, conformation numbers 1e30 (never near),
// energies 0, zero-derivative masks all ones (an entry nobody wrote never explains a point).
__global__ void vd_glob_init_kernel(float* base, const long long* off, const int* per_axis, const int* T, int n) {
  for (int k = blockIdx.y; k < n; k += gridDim.y) {
  const long long cells = (long long)per_axis[k] * per_axis[k] * per_axis[k];
  const int t = T[k];
  float* b = base + off[k];
  const long long per_cell = 8LL * (12 + t);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < cells + cells * per_cell;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < cells) { reinterpret_cast<int*>(b)[i] = 0; continue; }
    const long long j = (i - cells) % per_cell;      // position inside the cell's record
    float v = 0.0f;
    if (j < 8LL * (7 + t)) v = 1e30f;
    else if (j >= 8LL * (8 + t) && j < 8LL * (8 + t) + 16) v = __uint_as_float(0xFFFFFFFFu);   // Z masks
    b[i] = v;
  }
  }
}
hipError_t glob_init(float* base, const long long* off, const int* per_axis, const int* T, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(vd_glob_init_kernel, dim3(64, std::min(n, 1024)), dim3(256), 0, s, base, off, per_axis, T, n);
  return hipGetLastError();
}

// Random conformations of the adjoint benchmark: translation uniform in the box, orientation a
// normalised 4-Gaussian (uniform on rotations), torsions uniform in [-pi, pi).
__global__ void vd_random_confs_kernel(float* conf, long long n, int T, uint32_t seed, float3 lo, float3 hi) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float* x = conf + c * (7 + T);
  const uint32_t cl = (uint32_t)c, ch = (uint32_t)(c >> 32);
  U4 w = philox(cl, ch, 0xFFFFFFFEu, 0, seed, 0x5EEDu);
  x[0] = lo.x + (hi.x - lo.x) * u01(w.x); x[1] = lo.y + (hi.y - lo.y) * u01(w.y); x[2] = lo.z + (hi.z - lo.z) * u01(w.z);
  w = philox(cl, ch, 0xFFFFFFFEu, 1, seed, 0x5EEDu);
  const float r0 = sqrtf(-2.0f * logf(u01(w.x) + 1.0f / 16777216.0f)), r1 = sqrtf(-2.0f * logf(u01(w.z) + 1.0f / 16777216.0f));
  float s0, c0, s1, c1;
  sincosf(2.0f * kPi * u01(w.y), &s0, &c0); sincosf(2.0f * kPi * u01(w.w), &s1, &c1);
  float q0 = r0 * c0, q1 = r0 * s0, q2 = r1 * c1, q3 = r1 * s1;
  const float inv = rsqrtf(fmaxf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3, 1e-30f));
  x[3] = q0 * inv; x[4] = q1 * inv; x[5] = q2 * inv; x[6] = q3 * inv;
  for (int j = 0; j < T; j += 4) {
    w = philox(cl, ch, 0xFFFFFFFEu, 2 + (uint32_t)(j / 4), seed, 0x5EEDu);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    for (int k = 0; k < 4 && j + k < T; ++k) x[7 + j + k] = -kPi + 2.0f * kPi * u01(ws[k]);
  }
}
hipError_t random_confs(float* conf, long long n, int T, uint32_t seed, const float lo[3], const float hi[3], hipStream_t s) {
  const int threads = 256;
  const long long blocks = (n + threads - 1) / threads;
  hipLaunchKernelGGL(vd_random_confs_kernel, dim3((unsigned)blocks), dim3(threads), 0, s, conf, n, T, seed,
                     make_float3(lo[0], lo[1], lo[2]), make_float3(hi[0], hi[1], hi[2]));
  return hipGetLastError();
}
hipError_t reset_counter(int* p, hipStream_t s) {
  hipLaunchKernelGGL(vd_fill_int, dim3(1), dim3(1), 0, s, p, 0);
  return hipGetLastError();
}

template <int MODE, int W>
static int occupancy(int G, int lds_bytes) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (G) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, vd_mc_kernel<64, MODE, W>, 64, lds_bytes); break;
    case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, vd_mc_kernel<32, MODE, W>, 64, lds_bytes); break;
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, vd_mc_kernel<16, MODE, W>, 64, lds_bytes); break;
    default: break;
  }
  return e == hipSuccess ? n : 0;
}

// mode: 0 dense BFGS, 1 RILC-BFGS, 2 RILC-BFGS with the QuickVina2 history, 3 with QuickVina-W's
int max_resident_mc_blocks(int G, int W, int mode, int lds_bytes) {
  if (mode == 0) return occupancy<0, 2>(G, lds_bytes);
  if (mode == 2) return W == 3 ? occupancy<2, 3>(G, lds_bytes) : W == 4 ? occupancy<2, 4>(G, lds_bytes) : occupancy<2, 2>(G, lds_bytes);
  if (mode == 3) return W == 3 ? occupancy<3, 3>(G, lds_bytes) : W == 4 ? occupancy<3, 4>(G, lds_bytes) : occupancy<3, 2>(G, lds_bytes);
  return W == 3 ? occupancy<1, 3>(G, lds_bytes) : W == 4 ? occupancy<1, 4>(G, lds_bytes) : occupancy<1, 2>(G, lds_bytes);
}

}  // namespace vdl
