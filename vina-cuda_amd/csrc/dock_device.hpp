// dock_device.hpp -- device code of the docking hot path for gfx950 (wave64).
//
// One chain is owned by a group of G lanes (G = 64, 32 or 16; 64/G chains per wavefront).
// A workgroup is exactly one wavefront, so __syncthreads() is a wave-local LDS fence.
// Every collective (group shuffles, evaluations, fences) is executed in wave-uniform control
// flow; per-chain decisions only predicate register updates.  This replaces the reference's
// 32-thread blocks with shared-memory reductions (V11/inc/cuda/quasi_newton.cuh:352-428,
// 905-975, 1102-1180, 1510-1551).
//
// Lane layout of a chain (G lanes, lane l):
//   kinematics / adjoint: lane k < N owns rigid node k (preorder, node 0 = ROOT)
//   atoms:                atom i is handled by lane i % G
//   pairs:                pair p by lane p % G
//   change vector (6+T):  torsion k-1 in lane k (1 <= k <= T); position x,y,z and
//                         orientation x,y,z in lanes G-6 .. G-1; other lanes hold 0
//   conformation:         position and orientation quaternion replicated in all lanes of the
//                         group, torsion k-1 in lane k.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vd_common.h"

namespace vdk {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kEpsFl = 2.220446049250313e-16f;   // epsilon_fl as the reference passes it to the device

// ----------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): the random stream shared with the host oracle.
struct U4 { uint32_t x, y, z, w; };
__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ----------------------------------------------------------------------------------------
// Group collectives (G divides 64).  Sums: a row-of-16 all-reduce with DPP (quad_perm xor 1,
// xor 2, row_half_mirror, row_mirror), then xor-16/xor-32 lane shuffles across rows.  Every step
// adds the same two operands in both partner lanes, so all lanes of a group hold bit-identical
// sums (group decisions taken from them agree in every lane).
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int G> __device__ __forceinline__ float gsum(float v) {
  v += dpp<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);   // row_half_mirror
  v += dpp<0x140>(v);   // row_mirror
  if constexpr (G >= 32) v += __shfl_xor(v, 16);
  if constexpr (G >= 64) v += __shfl_xor(v, 32);
  return v;
}
template <int G> __device__ __forceinline__ float gget(float v, int src) { return __shfl(v, src, G); }
template <int G> __device__ __forceinline__ int ggeti(int v, int src) { return __shfl(v, src, G); }
// broadcast from a fixed lane of the group (a wave-uniform readlane when the group is the wave)
template <int G, int SRC> __device__ __forceinline__ float gbcast(float v) {
  if constexpr (G == 64) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), SRC));
  else return __shfl(v, SRC, G);
}

// Segmented inclusive scan of six sums over the lanes of a group, segments = runs of equal
// `seg` (-1 = no segment): DPP row shifts by 1, 2, 4, 8 inside each row of 16 lanes, then the
// last lane of a row feeds the next row (row_bcast:15) and lane 31 feeds rows 2 and 3
// (row_bcast:31), each added only where the segment continues.  Lanes whose source lies outside
// the row read 0 and segment -1.
template <int CTRL, int ROWS> __device__ __forceinline__ float dppf(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS> __device__ __forceinline__ int dppi(int v) {
  return __builtin_amdgcn_update_dpp(-1, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS> __device__ __forceinline__ void seg_step(float (&v)[6], int seg) {
  const int s = dppi<CTRL, ROWS>(seg);
  float t[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) t[c] = dppf<CTRL, ROWS>(v[c]);
  if (s == seg && seg >= 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] += t[c];
  }
}
template <int G> __device__ __forceinline__ void seg_scan(float (&v)[6], int seg) {
  seg_step<0x111, 0xF>(v, seg);   // row_shr:1
  seg_step<0x112, 0xF>(v, seg);   // row_shr:2
  seg_step<0x114, 0xF>(v, seg);   // row_shr:4
  seg_step<0x118, 0xF>(v, seg);   // row_shr:8
  if constexpr (G >= 32) seg_step<0x142, 0xA>(v, seg);   // row_bcast:15 into rows 1 and 3
  if constexpr (G >= 64) seg_step<0x143, 0xC>(v, seg);   // row_bcast:31 into rows 2 and 3
}

// ----------------------------------------------------------------------------------------
// Rotations.  Quaternions are (w, x, y, z); products are Hamilton products.
struct Quat { float w, x, y, z; };
__device__ __forceinline__ Quat qmul(const Quat& a, const Quat& b) {
  return Quat{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
              a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
__device__ __forceinline__ Quat qnormalize(Quat q) {
  const float s = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  const float inv = rsqrtf(s);
  return Quat{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
// rotation matrix, rows r0 r1 r2 (V11/lib/quaternion.cpp:155-185)
__device__ __forceinline__ void qmat(const Quat& q, float m[9]) {
  const float aa = q.w * q.w, ab = q.w * q.x, ac = q.w * q.y, ad = q.w * q.z, bb = q.x * q.x, bc = q.x * q.y,
              bd = q.x * q.z, cc = q.y * q.y, cd = q.y * q.z, dd = q.z * q.z;
  m[0] = aa + bb - cc - dd; m[1] = 2 * (-ad + bc);    m[2] = 2 * (ac + bd);
  m[3] = 2 * (ad + bc);     m[4] = aa - bb + cc - dd; m[5] = 2 * (-ab + cd);
  m[6] = 2 * (-ac + bd);    m[7] = 2 * (ab + cd);     m[8] = aa - bb - cc + dd;
}
__device__ __forceinline__ float3 qrot(const Quat& q, float3 v) {
  float m[9]; qmat(q, m);
  return make_float3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                     m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
// V11/lib/common.h:316-340
__device__ __forceinline__ float normalize_angle(float x) {
  if (!(x >= -3 * kPi && x <= 3 * kPi)) {
    if (x > 3 * kPi) x -= 2 * kPi * ceilf((x - kPi) / (2 * kPi));
    else if (x < -3 * kPi) x += 2 * kPi * ceilf((-x - kPi) / (2 * kPi));
  }
  if (x > kPi) x -= 2 * kPi;
  else if (x < -kPi) x += 2 * kPi;
  return x;
}
// V11/lib/quaternion.cpp:65-75
__device__ __forceinline__ Quat axis_angle(float3 a, float angle) {
  angle = normalize_angle(angle);
  float s, c; __sincosf(0.5f * angle, &s, &c);
  return Quat{c, s * a.x, s * a.y, s * a.z};
}
// quaternion_increment (V11/lib/quaternion.cpp:87-95, 235-240)
__device__ __forceinline__ Quat qincrement(Quat q, float rx, float ry, float rz) {
  const float ang = sqrtf(rx * rx + ry * ry + rz * rz);
  if (ang > kEpsFl) {
    const float inv = 1.0f / ang;
    q = qmul(axis_angle(make_float3(rx * inv, ry * inv, rz * inv), ang), q);
  }
  return qnormalize(q);
}

// ----------------------------------------------------------------------------------------
struct LigView {               // one ligand, staged in LDS for the wave
  const float4* atom;          // A: local xyz + meta (node | xs << 8 | heavy << 16)
  const float4* node_o;        // N: rel_origin xyz + bits (parent+1 | atom begin << 8 | atom end << 16)
  const float4* node_a;        // N: rel_axis xyz + subtree_end
  const uint4* pairs;          // C chunks of G pair words in groups of four, [group][lane][4], global
  int A, N, C, T, H, rounds;
  int rows;                    // 1: slot l of every chunk holds a pair whose first atom is l
};
struct Scratch {               // per chain, LDS
  float4* xyz;                 // A, coordinates relative to the root origin (w unused)
  float4* frc;                 // A, dE/dx (w unused)
  float4* fx;                  // G + 1: pair-force exchange row (entry G stays zero)
  float4* frame;               // 3N rows (R row, origin component)
  float4* nacc;                // 2N: per-node force and moment sums
  float* stash;                // per wave: conformations parked out of registers, [slot][field][lane]
};
// Maps as the searches read them (corner records): per type, one 32-byte record per grid cell (cell (i,j,k) spans
// points i..i+1, j..j+1, k..k+1), [type][k][j][i], holding the cell's eight corner values as
// {f000, f100, f010, f110} {f001, f101, f011, f111}.  One lookup is then two 16-byte loads from
// one cache line instead of four row gathers.
struct EvalK {
  const float4* __restrict__ corners; vd_grid_desc gd;
  int cd0, cd01;                 // cells per x-row, per z-plane
  long long cplane;              // cells per type
  const float2* __restrict__ lutp;  // pair table, (e, dor) per bin: [tp][bin]
  float v_intra, v_inter, slope;
  // the interpolation quad (e_i, dor_i, e_{i+1}, dor_{i+1}) of entry i: one 16-byte load at 8-byte
  // alignment (global memory takes unaligned vector loads), half the table's footprint in the
  // caches against a table that stores each quad whole
  __device__ __forceinline__ float4 quad(uint32_t i) const {
    return *reinterpret_cast<const float4*>(lutp + i);
  }
};
struct Conf {                  // register view of one conformation
  float px, py, pz; Quat q; float tor;
};

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

// A product rounded on its own, never fused with a following add into an FMA.
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}

// Trilinear map value and derivative at p (grid::evaluate_aux, V11/lib/grid.cpp:50-167), from the
// corner records of one type (cd0, cd01: cells per x-row, per z-plane).
// prel = position relative to the map origin.  NaN/inf inputs stay in bounds.
__device__ __forceinline__ float map_eval(const float4* __restrict__ m, int cd0, int cd01, const vd_grid_desc& gd,
                                          float3 prel, float slope, float v, float3& deriv) {
  float s[3] = {prel.x * gd.factor[0], prel.y * gd.factor[1], prel.z * gd.factor[2]};
  float miss[3]; int region[3], a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (s[i] >= 0.0f && s[i] < gd.dim_fl_minus_1[i]) {
      region[i] = 0; a[i] = (int)s[i]; s[i] -= (float)a[i]; miss[i] = 0.0f;
      a[i] = min(a[i], gd.dim[i] - 2);
    } else if (s[i] >= gd.dim_fl_minus_1[i]) {
      miss[i] = s[i] - gd.dim_fl_minus_1[i]; region[i] = 1; a[i] = gd.dim[i] - 2; s[i] = 1.0f;
    } else {
      miss[i] = -s[i]; region[i] = -1; a[i] = 0; s[i] = 0.0f;
    }
  }
  const float penalty = slope * (miss[0] * gd.factor_inv[0] + miss[1] * gd.factor_inv[1] + miss[2] * gd.factor_inv[2]);
  const float4* b = m + 2 * ((long long)cd01 * a[2] + cd0 * a[1] + a[0]);
  const float4 lo = b[0], hi = b[1];
  const float f000 = lo.x, f100 = lo.y, f010 = lo.z, f110 = lo.w;
  const float f001 = hi.x, f101 = hi.y, f011 = hi.z, f111 = hi.w;
  const float x = s[0], y = s[1], z = s[2], mx = 1 - x, my = 1 - y, mz = 1 - z;
  float f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
            f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
  float gx = (f100 - f000) * my * mz + (f110 - f010) * y * mz + (f101 - f001) * my * z + (f111 - f011) * y * z;
  float gy = (f010 - f000) * mx * mz + (f110 - f100) * x * mz + (f011 - f001) * mx * z + (f111 - f101) * x * z;
  float gz = (f001 - f000) * mx * my + (f101 - f100) * x * my + (f011 - f010) * mx * y + (f111 - f110) * x * y;
  {  // curl (V11/lib/curl.h:31-38); hardware reciprocal for the quotient; below zero the factor is
     // 1, which leaves f and its gradient exactly as they are
    const float t = f > 0.0f ? ((v < kEpsFl) ? 0.0f : v * __builtin_amdgcn_rcpf(v + f)) : 1.0f;
    f = mul_rn(f, t); const float t2 = t * t; gx *= t2; gy *= t2; gz *= t2;
  }
  deriv.x = gd.factor[0] * (region[0] == 0 ? gx : 0.0f) + slope * region[0];
  deriv.y = gd.factor[1] * (region[1] == 0 ? gy : 0.0f) + slope * region[1];
  deriv.z = gd.factor[2] * (region[2] == 0 ? gz : 0.0f) + slope * region[2];
  return f + penalty;
}

// Forward kinematics (tree.h:205-209, 282-292).  Node transforms compose along the tree as
// q_k = q_parent (x) axis_angle(rel_axis_k, theta_k), o_k = o_parent + R_parent rel_origin_k,
// which equals the reference's lab-frame update because axis_angle(R a, t) R = R axis_angle(a, t).
// The per-path products are formed by pointer jumping in ceil(log2(depth+1)) rounds.
// Writes frame rows to LDS; returns the node's lab-frame axis in `ax` and origin in `org`.
template <int G>
__device__ __forceinline__ void kinematics(const LigView& L, const Scratch& S, const Conf& c, int l,
                                           float3& ax, float3& org) {
  Quat q{1, 0, 0, 0}; float3 t = make_float3(0, 0, 0); int anc = -1;
  float3 rel_axis = make_float3(0, 0, 0);
  if (l == 0) { q = c.q; }
  else if (l < L.N) {
    const float4 no = L.node_o[l], na = L.node_a[l];
    rel_axis = make_float3(na.x, na.y, na.z);
    q = axis_angle(rel_axis, c.tor);
    t = make_float3(no.x, no.y, no.z);
    anc = (int)(fbits(no.w) & 255u) - 1;
  }
  for (int r = 0; r < L.rounds; ++r) {
    const int src = anc >= 0 ? anc : l;
    const Quat qa{gget<G>(q.w, src), gget<G>(q.x, src), gget<G>(q.y, src), gget<G>(q.z, src)};
    const float3 ta = make_float3(gget<G>(t.x, src), gget<G>(t.y, src), gget<G>(t.z, src));
    const int aa = ggeti<G>(anc, src);
    if (anc >= 0) {
      const float3 rt = qrot(qa, t);
      t = make_float3(ta.x + rt.x, ta.y + rt.y, ta.z + rt.z);
      q = qmul(qa, q);
      anc = aa;
    }
  }
  q = qnormalize(q);
  float m[9]; qmat(q, m);
  org = t;
  ax = make_float3(m[0] * rel_axis.x + m[1] * rel_axis.y + m[2] * rel_axis.z,
                   m[3] * rel_axis.x + m[4] * rel_axis.y + m[5] * rel_axis.z,
                   m[6] * rel_axis.x + m[7] * rel_axis.y + m[8] * rel_axis.z);
  if (l < L.N) {
    S.frame[3 * l + 0] = make_float4(m[0], m[1], m[2], t.x);
    S.frame[3 * l + 1] = make_float4(m[3], m[4], m[5], t.y);
    S.frame[3 * l + 2] = make_float4(m[6], m[7], m[8], t.z);
  }
}

__device__ __forceinline__ float3 atom_position(const LigView& L, const Scratch& S, int i, uint32_t& meta) {
  const float4 a = L.atom[i];
  meta = fbits(a.w);
  const int nd = meta & 255;
  const float4 r0 = S.frame[3 * nd], r1 = S.frame[3 * nd + 1], r2 = S.frame[3 * nd + 2];
  return make_float3(r0.x * a.x + r0.y * a.y + r0.z * a.z + r0.w, r1.x * a.x + r1.y * a.y + r1.z * a.z + r1.w,
                     r2.x * a.x + r2.y * a.y + r2.z * a.z + r2.w);
}

// Coordinates only (model::set).  Leaves xyz (relative to the root origin) in LDS.
template <int G>
__device__ void set_coords(const LigView& L, const Scratch& S, const Conf& c, int l) {
  float3 ax, org;
  kinematics<G>(L, S, c, l, ax, org);
  __syncthreads();
  for (int i = l; i < L.A; i += G) {
    uint32_t meta; const float3 y = atom_position(L, S, i, meta);
    S.xyz[i] = make_float4(y.x, y.y, y.z, 0.0f);
  }
  __syncthreads();
}

// Radius of gyration of heavy atoms about the root origin (model::gyration_radius,
// V11/lib/model.cpp:1161-1175); call after set_coords.
template <int G>
__device__ float gyration_radius(const LigView& L, const Scratch& S, int l) {
  float acc = 0.0f, cnt = 0.0f;
  for (int i = l; i < L.A; i += G) {
    const uint32_t meta = fbits(L.atom[i].w);
    if ((meta >> 16) & 1u) { const float4 y = S.xyz[i]; acc += y.x * y.x + y.y * y.y + y.z * y.z; cnt += 1.0f; }
  }
  acc = gsum<G>(acc); cnt = gsum<G>(cnt);
  return cnt > 0.0f ? sqrtf(acc / cnt) : 0.0f;
}

// Curl constants of the pair terms: the cap factor is vs / (vd + e), which is v / (v + e), or 0
// when v < epsilon (V11/lib/curl.h:4-13); computed once per evaluation.
struct CurlK {
  float vs, vd;
  __device__ explicit CurlK(float v) : vs(v < kEpsFl ? 0.0f : v), vd(v < kEpsFl ? 1.0f : v) {}
};

// One pair term: value and derivative/r from the pair table quad, with the curl cap
// (V11/lib/curl.h:4-13) above zero.  The cap's quotient uses the hardware reciprocal (about one
// ulp, far inside the parity tolerance).
__device__ __forceinline__ void pair_term(const float4 q, float rem, const CurlK& ck, float& e, float& dor) {
  e = q.x + rem * (q.z - q.x);
  dor = q.y + rem * (q.w - q.y);
  // the factor, not the results, is selected: multiplying by 1 leaves e and dor exactly as they are
  const float t = e > 0.0f ? ck.vs * __builtin_amdgcn_rcpf(ck.vd + e) : 1.0f;
  e = mul_rn(e, t);
  dor *= t * t;
}

// Pair terms when every atom has its own lane (L.rows): slot l of each chunk holds pair (l, b).
// Lane l keeps the force on atom l in registers: its own pairs give the first-atom side, and the
// second-atom side comes from the one slot per chunk whose pair has second atom l, through a
// per-group LDS exchange row (every lane writes its pair force to its entry, then reads the
// source slot's entry; entry G stays zero for lanes without a source).  The words need no
// decoding: a slot without a pair names a written coordinate row as its partner and the all-zero
// row of the pair table as its type pair, so it reads zero quads and contributes exact zeros,
// and a lane without a source names entry G.  Four chunks per pass (one 16-byte read of the
// lane's pair words, prefetched a pass ahead); the four partner reads, then the four table reads,
// are issued before the first is used.
template <int G>
__device__ float pair_terms_rows(const LigView& L, const Scratch& S, const EvalK& K, int l) {
  constexpr int U = 4;
  const int groups = (L.C + U - 1) / U;
  const uint4* pw = L.pairs + l;
  const float4 own = l < L.A ? S.xyz[l] : make_float4(0, 0, 0, 0);
  const CurlK ck(K.v_intra);
  if (l == 0) S.fx[G] = make_float4(0, 0, 0, 0);
  float e_intra = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;
  uint4 w_next = groups > 0 ? pw[0] : make_uint4(0, 0, 0, 0);
  for (int cg = 0; cg < groups; ++cg) {
    const uint4 w4 = w_next;
    if (cg + 1 < groups) w_next = pw[(cg + 1) * G];
    const uint32_t w[U] = {w4.x, w4.y, w4.z, w4.w};
    float4 pb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) pb[u] = S.xyz[(w[u] >> 8) & 255u];
    float rx[U], ry[U], rz[U], rem[U];
    uint32_t idx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      rx[u] = pb[u].x - own.x; ry[u] = pb[u].y - own.y; rz[u] = pb[u].z - own.z;
      const float r2 = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
      const float rf = VD_LUT_FACTOR * r2;
      rem[u] = __builtin_amdgcn_fractf(rf);   // rf - trunc(rf), exact for rf >= 0
      idx[u] = r2 < VD_CUTOFF_SQR ? (w[u] >> 16) * (uint32_t)VD_LUT_BINS + (uint32_t)rf : (uint32_t)VD_ZERO_QUAD;
    }
    float4 q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = K.quad(idx[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float e, dor;
      pair_term(q[u], rem[u], ck, e, dor);
      e_intra += e;
      const float fx = dor * rx[u], fy = dor * ry[u], fz = dor * rz[u];
      ax -= fx; ay -= fy; az -= fz;
      S.fx[l] = make_float4(fx, fy, fz, 0.0f);
      const float4 in = S.fx[w[u] & 255u];
      ax += in.x; ay += in.y; az += in.z;
    }
  }
  if (l < L.A) {
    float4 f = S.frc[l];
    f.x += ax; f.y += ay; f.z += az;
    S.frc[l] = f;
  }
  return e_intra;
}

// Pair terms in the general case: lane l takes slot l of each chunk; within a chunk the first
// atoms are distinct and so are the second atoms, so the force scatter is a plain read-modify-
// write of LDS (first atoms, then second atoms), in chunk order: no atomics, and the same
// summation order in every run.
template <int G>
__device__ float pair_terms_scatter(const LigView& L, const Scratch& S, const EvalK& K, int l) {
  constexpr int U = 4;
  const int groups = (L.C + U - 1) / U;
  const uint4* pw = L.pairs + l;
  const CurlK ck(K.v_intra);
  float e_intra = 0.0f;
  uint4 w_next = groups > 0 ? pw[0] : make_uint4(0, 0, 0, 0);
  for (int cg = 0; cg < groups; ++cg) {
    const uint4 w4 = w_next;
    if (cg + 1 < groups) w_next = pw[(cg + 1) * G];
    const uint32_t w[U] = {w4.x, w4.y, w4.z, w4.w};
    bool ok[U];
    int a[U], b[U];
    uint32_t idx[U];
    float rx[U], ry[U], rz[U], rem[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = w[u] != VD_NO_PAIR;
      a[u] = ok[u] ? (int)(w[u] & 255) : 0;
      b[u] = ok[u] ? (int)((w[u] >> 8) & 255) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float4 pa = S.xyz[a[u]], pb = S.xyz[b[u]];
      rx[u] = pb.x - pa.x; ry[u] = pb.y - pa.y; rz[u] = pb.z - pa.z;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float r2 = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
      ok[u] = ok[u] && r2 < VD_CUTOFF_SQR;
      const float rf = VD_LUT_FACTOR * r2;
      rem[u] = __builtin_amdgcn_fractf(rf);   // rf - trunc(rf), exact for rf >= 0
      idx[u] = ok[u] ? (w[u] >> 16) * (uint32_t)VD_LUT_BINS + (uint32_t)rf : (uint32_t)VD_ZERO_QUAD;
    }
    float4 q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = K.quad(idx[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float e, dor;
      pair_term(q[u], rem[u], ck, e, dor);
      if (!ok[u]) continue;
      e_intra += e;
      const float fx = dor * rx[u], fy = dor * ry[u], fz = dor * rz[u];
      float4 fa = S.frc[a[u]];
      fa.x -= fx; fa.y -= fy; fa.z -= fz;
      S.frc[a[u]] = fa;
      float4 fb = S.frc[b[u]];
      fb.x += fx; fb.y += fy; fb.z += fz;
      S.frc[b[u]] = fb;
    }
  }
  return e_intra;
}

// Energy and gradient of one conformation per group (model::eval_deriv,
// V11/lib/model.cpp:1324-1334; device form quasi_newton.cuh:1294-1339).
// Returns e (group-uniform); g = this lane's gradient component.  rg (when given) receives the
// gyration radius of the evaluated conformation, as gyration_radius computes it.
template <int G>
__device__ float eval_deriv(const LigView& L, const Scratch& S, const EvalK& K, const Conf& c, int l, float& g,
                            float* rg = nullptr) {
  float3 ax, org;
  kinematics<G>(L, S, c, l, ax, org);
  __syncthreads();
  // intermolecular term: trilinear maps (cache::eval_deriv, V11/lib/cache.cpp:73-88)
  const float3 prel = make_float3(c.px - K.gd.init[0], c.py - K.gd.init[1], c.pz - K.gd.init[2]);
  float e_inter = 0.0f, racc = 0.0f;
  // two atoms per lane in flight: both sets of map reads are issued before either is consumed
  for (int i0 = l; i0 < L.A; i0 += 2 * G) {
    float3 d[2], y[2];
    float ev[2];
    bool heavy[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + u * G;
      d[u] = make_float3(0, 0, 0); ev[u] = 0.0f; y[u] = make_float3(0, 0, 0); heavy[u] = false;
      if (i < L.A) {
        uint32_t meta; y[u] = atom_position(L, S, i, meta);
        heavy[u] = (meta >> 16) & 1u;
        const int xs = (meta >> 8) & 255;
        if (xs < VD_XS_SIZE)
          ev[u] = map_eval(K.corners + 2 * xs * K.cplane, K.cd0, K.cd01, K.gd,
                           make_float3(prel.x + y[u].x, prel.y + y[u].y, prel.z + y[u].z), K.slope, K.v_inter, d[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = i0 + u * G;
      if (i < L.A) {
        S.xyz[i] = make_float4(y[u].x, y[u].y, y[u].z, 0.0f);
        S.frc[i] = make_float4(d[u].x, d[u].y, d[u].z, 0.0f);
        e_inter += ev[u];
        if (heavy[u]) racc += y[u].x * y[u].x + y[u].y * y[u].y + y[u].z * y[u].z;
      }
    }
  }
  if (rg) { const float acc = gsum<G>(racc); *rg = L.H > 0 ? sqrtf(acc / (float)L.H) : 0.0f; }
  __syncthreads();
  // intramolecular pairs (eval_interacting_pairs_deriv, V11/lib/model.cpp:1241-1259)
  const float e_intra = L.rows ? pair_terms_rows<G>(L, S, K, l) : pair_terms_scatter<G>(L, S, K, l);
  __syncthreads();
  // adjoint: per-node force and moment (about the root origin) sums.  A node's atoms are one
  // contiguous range, so the sums are a segmented inclusive scan over the atom lanes (segment =
  // node; untyped atoms carry 0); the last lane of each segment adds its sums to the node's
  // record.  Atoms beyond G lanes take further passes, each adding in turn.
  if (l < L.N) { S.nacc[2 * l] = make_float4(0, 0, 0, 0); S.nacc[2 * l + 1] = make_float4(0, 0, 0, 0); }
  __syncthreads();
  for (int i0 = 0; i0 < L.A; i0 += G) {
    const int i = i0 + l;
    float v[6] = {0, 0, 0, 0, 0, 0};
    int node = -1;
    if (i < L.A) {
      const float4 fr = S.frc[i], p = S.xyz[i];
      node = (int)(fbits(L.atom[i].w) & 255u);
      v[0] = fr.x; v[1] = fr.y; v[2] = fr.z;
      v[3] = p.y * fr.z - p.z * fr.y; v[4] = p.z * fr.x - p.x * fr.z; v[5] = p.x * fr.y - p.y * fr.x;
    }
    seg_scan<G>(v, node);
    const int next = __shfl_down(node, 1, G);
    if (node >= 0 && (l == G - 1 || next != node)) {
      float4 a = S.nacc[2 * node], b = S.nacc[2 * node + 1];
      a.x += v[0]; a.y += v[1]; a.z += v[2]; b.x += v[3]; b.y += v[4]; b.z += v[5];
      S.nacc[2 * node] = a; S.nacc[2 * node + 1] = b;
    }
  }
  __syncthreads();
  // subtree sums: the subtree of node k is the preorder range [k, subtree_end_k); the torque
  // about o_k is M - o_k x F (tree.h:341-348, 373-380; torsion derivative = torque . axis)
  float F[3] = {0, 0, 0}, M[3] = {0, 0, 0};
  if (l < L.N) {
    const int end = (int)fbits(L.node_a[l].w);
    for (int j = l; j < end; ++j) {
      const float4 f = S.nacc[2 * j], m = S.nacc[2 * j + 1];
      F[0] += f.x; F[1] += f.y; F[2] += f.z; M[0] += m.x; M[1] += m.y; M[2] += m.z;
    }
  }
  const float tx = M[0] - (org.y * F[2] - org.z * F[1]);
  const float ty = M[1] - (org.z * F[0] - org.x * F[2]);
  const float tz = M[2] - (org.x * F[1] - org.y * F[0]);
  const float gt = tx * ax.x + ty * ax.y + tz * ax.z;
  const float r0 = gbcast<G, 0>(F[0]), r1 = gbcast<G, 0>(F[1]), r2 = gbcast<G, 0>(F[2]);
  const float r3 = gbcast<G, 0>(tx), r4 = gbcast<G, 0>(ty), r5 = gbcast<G, 0>(tz);
  const int c6 = l - (G - 6);
  g = (l >= 1 && l <= L.T) ? gt
      : c6 == 0 ? r0 : c6 == 1 ? r1 : c6 == 2 ? r2 : c6 == 3 ? r3 : c6 == 4 ? r4 : c6 == 5 ? r5 : 0.0f;
  return gsum<G>(e_inter + e_intra);
}

// x <- x (+) factor*d (output_type_cl_increment, V11/inc/cuda/mutate_conf.cuh:8-29;
// conf.h:60-66, 118-124).  d is the lane's change component.
template <int G>
__device__ __forceinline__ Conf increment(const Conf& x, float d, float factor, int l, int T) {
  Conf y = x;
  const float d0 = gbcast<G, G - 6>(d), d1 = gbcast<G, G - 5>(d), d2 = gbcast<G, G - 4>(d);
  const float d3 = gbcast<G, G - 3>(d), d4 = gbcast<G, G - 2>(d), d5 = gbcast<G, G - 1>(d);
  y.px += factor * d0; y.py += factor * d1; y.pz += factor * d2;
  y.q = qincrement(y.q, factor * d3, factor * d4, factor * d5);
  if (l >= 1 && l <= T) y.tor = normalize_angle(y.tor + normalize_angle(factor * d));
  return y;
}

__device__ __forceinline__ Conf csel(bool p, const Conf& a, const Conf& b) { return p ? a : b; }

// Conformations that are only read now and then live in LDS instead of registers (the MC kernel
// is register-bound).  A slot holds, per group, the position and orientation once (8 floats,
// written by the group's lane 0) and, per lane, its torsion: [slot][group][8] then [slot][64].
// LDS operations of one wavefront complete in order; the wavefront fences keep the compiler
// from moving a read of a slot above the write that fills it.
constexpr int kStashSlots = 4;
constexpr int kStashSlotFloats = 4 * 8 + 64;
enum { kSlotOrig = 0, kSlotPrev = 1, kSlotTmp = 2, kSlotBest = 3 };
template <int G>
__device__ __forceinline__ void stash_put(float* st, int slot, const Conf& c) {
  const int lane = threadIdx.x & 63;
  float* p = st + slot * kStashSlotFloats;
  if (lane % G == 0) {
    float4* u = reinterpret_cast<float4*>(p + (lane / G) * 8);
    u[0] = make_float4(c.px, c.py, c.pz, c.q.w);
    u[1] = make_float4(c.q.x, c.q.y, c.q.z, 0.0f);
  }
  p[32 + lane] = c.tor;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}
template <int G>
__device__ __forceinline__ Conf stash_get(const float* st, int slot) {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int lane = threadIdx.x & 63;
  const float* p = st + slot * kStashSlotFloats;
  const float4* u = reinterpret_cast<const float4*>(p + (lane / G) * 8);
  const float4 a = u[0], b = u[1];
  Conf c;
  c.px = a.x; c.py = a.y; c.pz = a.z; c.q = Quat{a.w, b.x, b.y, b.z}; c.tor = p[32 + lane];
  return c;
}

// ----------------------------------------------------------------------------------------
// QuickVina2 visited history of one chain (QV2/inc/cuda/visited.cuh:7-226) and the gate that
// skips a local search whose start point is explained by its neighbours
// (QV2/inc/cuda/quasi_newton.cuh:1545-1550).  Semantics as intended by the reference: a
// neighbour explains the point when some change dimension shows both on one monotone slope
// (same derivative sign, coordinate and energy differences ordered as that sign says); the point
// is interesting unless each of its n nearest neighbours (n = 6+T, squared distance over the
// 7+T conformation numbers) explains it; torsion dimensions use the torsion derivative.
// Storage is a ring of cap = 10 n entries in global memory, one region per resident chain:
//   X[c][cap]  conformation number c: c = 0 orientation z, 1..T torsions, T+1..T+6 position
//              x,y,z and orientation w,x,y (the lane-layout order: lane 0, lanes 1..T,
//              lanes G-6..G-1)
//   F[cap]     energy;   Z[cap], Pm[cap]  lane masks of zero / positive derivatives.
// Entries are compared on the coordinate of their dimension lane: lanes G-6..G-4 position,
// G-3..G-1 orientation w,x,y (as the reference compares the first components of its
// flattened vector), lane k torsion k-1.
constexpr int kHistMaxRounds = 10;                // cap = 10 (6+T) <= 10 G entries -> 10 per lane
struct History {
  float* X; float* F; unsigned long long* Z; unsigned long long* Pm;
  int cap, count, p;
  // rounds x (6+T) entries: 10 for QuickVina2, 5 for QuickVina-W's individual buffer
  __host__ __device__ static long long floats_per_chain(int T, int rounds = 10) {
    const long long cap = (long long)rounds * (6 + T); return cap * (12 + T);
  }
  __device__ void init(float* base, int T, int rounds = 10) {
    cap = rounds * (6 + T); count = 0; p = 0;
    X = base; F = base + (long long)cap * (7 + T);
    Z = reinterpret_cast<unsigned long long*>(F + cap); Pm = Z + cap;
  }
};
// Entries to examine: the first `avail` of a store laid out as History (row stride cap).
struct HistView {
  const float* X; const float* F; const unsigned long long* Z; const unsigned long long* Pm;
  int cap, avail;
};
template <int G> __device__ __forceinline__ int hist_comp(int l, int T) {
  return l <= T ? l : (l >= G - 6 ? T + 1 + (l - (G - 6)) : -1);
}
template <int G> __device__ __forceinline__ int hist_lane(int c, int T) { return c <= T ? c : (G - 6) + (c - T - 1); }
template <int G> __device__ __forceinline__ float conf_coord(const Conf& x, int l, int T) {
  const int c6 = l - (G - 6);
  return l == 0 ? x.q.z : l <= T ? x.tor
       : c6 == 0 ? x.px : c6 == 1 ? x.py : c6 == 2 ? x.pz : c6 == 3 ? x.q.w : c6 == 4 ? x.q.x : c6 == 5 ? x.q.y : 0.0f;
}
template <int G> __device__ __forceinline__ unsigned long long group_ballot(bool v) {
  const unsigned long long b = __ballot(v);
  if constexpr (G == 64) return b;
  else return (b >> ((threadIdx.x & 63) / G * G)) & ((1ull << G) - 1ull);
}
template <int G> __device__ __forceinline__ bool is_dim_lane(int l, int T) { return (l >= 1 && l <= T) || l >= G - 6; }

// The gate functions below are compiled out of line (noinline).  Inlined into the search loop,
// their temporaries raise the register pressure of the whole loop, and at three waves per SIMD the
// kernels then spill inside it on every evaluation; called once per local search, the calls cost
// little.

// Adds (x, f, g) when on; every lane of the wave calls it.
template <int G>
__device__ __attribute__((noinline)) void hist_add(History& h, const Conf& x, float f, float g, int l, int T, bool on) {
  const bool dim = is_dim_lane<G>(l, T);
  const unsigned long long z = group_ballot<G>(dim && g == 0.0f), ps = group_ballot<G>(dim && g > 0.0f);
  const int slot = h.count < h.cap ? h.count : h.p;
  if (on) {
    const int c = hist_comp<G>(l, T);
    if (c >= 0) h.X[(long long)c * h.cap + slot] = conf_coord<G>(x, l, T);
    if (l == 0) { h.F[slot] = f; h.Z[slot] = z; h.Pm[slot] = ps; }
    if (h.count < h.cap) ++h.count; else h.p = (h.p + 1 == h.cap) ? 0 : h.p + 1;
  }
  __syncthreads();                                 // entries visible to the whole wave
}

// Examines, nearest first (lowest index among equals), up to max_checks of the view's entries:
// true as soon as one does not explain (x, f, g), else false; checked = entries examined.
// Groups with need == false examine nothing.  Every lane of the wave calls it.
template <int G>
__device__ __attribute__((noinline)) bool hist_check(const HistView& h, const Conf& x, float f, float g, int l, int T, bool need,
                           int max_checks, int& checked) {
  checked = 0;
  if (!__any(need)) return false;
  const float coord = conf_coord<G>(x, l, T);
  float d[kHistMaxRounds];
#pragma unroll
  for (int r = 0; r < kHistMaxRounds; ++r) d[r] = 0.0f;
  for (int c = 0; c < 7 + T; ++c) {
    const float xc = gget<G>(coord, hist_lane<G>(c, T));
    const float* row = h.X + (long long)c * h.cap;
#pragma unroll
    for (int r = 0; r < kHistMaxRounds; ++r) {
      const int e = l + G * r;
      if (need && e < h.avail) { const float t = row[e] - xc; d[r] += t * t; }
    }
  }
  const bool dim = is_dim_lane<G>(l, T);
  const int cl = hist_comp<G>(l, T);
  uint32_t picked = 0;
  bool found = false;
  for (int it = 0; __any(need && !found && it < max_checks); ++it) {
    // next nearest entry: smallest distance, lowest index among equals
    float bd = INFINITY; int be = h.cap;
#pragma unroll
    for (int r = 0; r < kHistMaxRounds; ++r) {
      const int e = l + G * r;
      if (e < h.avail && !((picked >> r) & 1u) && d[r] < bd) { bd = d[r]; be = e; }
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const float od = __shfl_xor(bd, off, G); const int oe = __shfl_xor(be, off, G);
      if (od < bd || (od == bd && oe < be)) { bd = od; be = oe; }
    }
    const bool pick = need && !found && it < max_checks && be < h.avail;
    if (pick) checked = it + 1;
    if (pick && be % G == l) picked |= 1u << (be / G);
    bool explains = false;
    if (pick && dim && !(g == 0.0f)) {
      const unsigned long long nz = h.Z[be], np = h.Pm[be];
      if (!((nz >> l) & 1ull)) {
        const bool now_pos = g > 0.0f;
        if (now_pos == (bool)((np >> l) & 1ull)) {
          const bool x_bigger = (coord - h.X[(long long)cl * h.cap + be]) > 0.0f;
          const bool y_bigger = (f - h.F[be]) > 0.0f;
          explains = now_pos ? (x_bigger == y_bigger) : (x_bigger != y_bigger);
        }
      }
    }
    const bool any_explains = group_ballot<G>(explains) != 0ull;
    if (pick && !any_explains) found = true;
  }
  return need && found;
}

// QuickVina2's gate: true when (x, f, g) is worth a local search -- always until the history is
// full, then unless each of the 6+T nearest entries explains it.  Every lane of the wave calls it.
template <int G>
__device__ __attribute__((noinline)) bool hist_interesting(const History& h, const Conf& x, float f, float g, int l, int T, bool on) {
  const bool need = on && h.count == h.cap;
  int checked;
  const bool found = hist_check<G>(HistView{h.X, h.F, h.Z, h.Pm, h.cap, h.cap}, x, f, g, l, T, need, 6 + T, checked);
  return !need || found;
}

// QuickVina-W's global buffer of one ligand: per_axis^3 cells over the box (x fastest), each a
// ring of 8 entries laid out as History with cap 8, after an int counter per cell (adds so far).
// Shared by every chain of the ligand through global memory: a slot is claimed with an atomic
// add, then written; a reader may meet an entry while it is being written (the search heuristic
// tolerates it; unwritten entries carry all-zero-derivative masks and never explain a point).
struct GlobalHist {
  float* base; int per_axis, T;
  float lo[3], hi[3];
  __device__ int cells() const { return per_axis * per_axis * per_axis; }
  __device__ int cell_of(float px, float py, float pz) const {
    const float p[3] = {px, py, pz};
    int id[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float t = (p[k] - lo[k]) / (hi[k] - lo[k]) * (float)per_axis;   // NaN lands in cell 0
      id[k] = t >= 0.0f ? (int)fminf(t, (float)(per_axis - 1)) : 0;
    }
    return (id[2] * per_axis + id[1]) * per_axis + id[0];
  }
  __device__ int* counter(int cell) const { return reinterpret_cast<int*>(base) + cell; }
  __device__ HistView view(int cell, int avail) const {
    float* c = base + cells() + (long long)cell * 8 * (12 + T);
    const float* F = c + 8 * (7 + T);
    const unsigned long long* Z = reinterpret_cast<const unsigned long long*>(F + 8);
    return HistView{c, F, Z, Z + 8, 8, avail};
  }
  __host__ __device__ static long long floats(int per_axis, int T) {
    const long long n = (long long)per_axis * per_axis * per_axis;
    return n + n * 8 * (12 + T);
  }
};

// Adds (x, f, g) to the global buffer cell of x's position when on.  Every lane calls it.
template <int G>
__device__ __attribute__((noinline)) void glob_add(const GlobalHist& gh, const Conf& x, float f, float g, int l, int T, bool on) {
  const bool dim = is_dim_lane<G>(l, T);
  const unsigned long long z = group_ballot<G>(dim && g == 0.0f), ps = group_ballot<G>(dim && g > 0.0f);
  const int cell = gh.cell_of(x.px, x.py, x.pz);
  int slot = 0;
  if (on && l == 0) slot = atomicAdd(gh.counter(cell), 1) & 7;
  slot = ggeti<G>(slot, 0);
  if (on) {
    const HistView v = gh.view(cell, 8);
    float* X = const_cast<float*>(v.X);
    const int c = hist_comp<G>(l, T);
    if (c >= 0) X[c * 8 + slot] = conf_coord<G>(x, l, T);
    if (l == 0) {
      const_cast<float*>(v.F)[slot] = f;
      const_cast<unsigned long long*>(v.Z)[slot] = z;
      const_cast<unsigned long long*>(v.Pm)[slot] = ps;
    }
  }
}

// QuickVina-W's gate (QVW/inc/cuda/quasi_newton.cuh:2236-2245, intended semantics): the G-check
// examines up to 6+T nearest entries of the cell holding x; if none says interesting, the
// individual buffer decides -- interesting until full, then by its 4(6+T) - (examined) nearest.
template <int G>
__device__ __attribute__((noinline)) bool qvw_interesting(const History& h, const GlobalHist& gh, const Conf& x, float f, float g, int l, int T,
                                bool on) {
  const int cell = gh.cell_of(x.px, x.py, x.pz);
  const int avail = on ? min(__atomic_load_n(gh.counter(cell), __ATOMIC_RELAXED), 8) : 0;
  int checked = 0;
  const bool glob = hist_check<G>(gh.view(cell, avail), x, f, g, l, T, on && avail > 0, 6 + T, checked);
  const bool full = h.count == h.cap;
  int c2;
  const bool ind = hist_check<G>(HistView{h.X, h.F, h.Z, h.Pm, h.cap, h.cap}, x, f, g, l, T, on && !glob && full,
                                 4 * (6 + T) - checked, c2);
  return glob || !full || ind;
}

// Dense BFGS with backtracking line search (V11/lib/bfgs.h:30-140; device form
// quasi_newton.cuh:1859-1916, 2391-2499).  H (n x n, n = 6+T <= G) lives in LDS, row i in the
// lane that owns change component i.  The loops over a row stay rolled: unrolled G-long, they
// kept G LDS values live at once and spilled up to 1.3 KB per lane.
// Dense BFGS (bfgs.h:30-140 as the reference GPU code runs it, quasi_newton.cuh:2391-2499): the
// matrix of each chain group, over the n = 6 + T dimensions of the change vector (torsion k - 1 in
// lane k, position and orientation in lanes G-6 .. G-1; see the lane layout above).  Dimension c
// of lane l is dense_dim(l); the lanes of the other G - n lanes hold zeros in every vector, so
// their rows and columns of a G x G matrix would stay zero and only ever meet zeros: the matrix
// keeps the n x n block, in the group's G x G area.  Lane l owns row dense_dim(l), kept
// column-major (entry (c, c') at H[c' n + c]) so that the lanes' accesses at one c' are
// consecutive words: no LDS bank conflicts (row-major, the lanes would hit one bank).  Sums run
// over the dimensions in lane order.  vbuf (2 G floats after the G x G area) broadcasts vectors.
// Every lane of the wavefront calls these (they fence); `on` selects the groups whose matrix is
// written.
template <int G> __device__ __forceinline__ int dense_dim(int l, int T) {
  return (l >= 1 && l <= T) ? l - 1 : (l >= G - 6 ? T + (l - (G - 6)) : -1);
}
template <int G> __device__ __forceinline__ int dense_lane(int c, int T) { return c < T ? c + 1 : (G - 6) + (c - T); }
template <int G>
__device__ __forceinline__ void dense_identity(float* H, int l, int T, bool on) {
  const int n = 6 + T, c = dense_dim<G>(l, T);
  if (on && c >= 0) {
#pragma unroll 4
    for (int j = 0; j < n; ++j) H[j * n + c] = (j == c) ? 1.0f : 0.0f;
  }
}
// -(H v) for this lane's dimension (-0 for the other lanes)
template <int G>
__device__ __forceinline__ float dense_minus_hv(const float* H, float* vbuf, float v, int l, int T) {
  const int n = 6 + T, c = dense_dim<G>(l, T);
  __syncthreads(); vbuf[l] = v; __syncthreads();
  float acc = 0.0f;
  if (c >= 0) {
#pragma unroll 4
    for (int j = 0; j < n; ++j) acc += H[j * n + c] * vbuf[dense_lane<G>(j, T)];
  }
  __syncthreads();
  return -acc;
}
// H_cj += alpha r (mhy_c p_j + mhy_j p_c) + alpha^2 (r^2 yhy + r) p_c p_j  (mhy = -H y, r = 1/(alpha y.p))
template <int G>
__device__ __forceinline__ void dense_update(float* H, float* vbuf, float mhy, float p, float alpha, float r, float yhy,
                                             bool on, int l, int T) {
  const int n = 6 + T, c = dense_dim<G>(l, T);
  float* pb = vbuf + G;
  __syncthreads();
  vbuf[l] = mhy; pb[l] = p;
  __syncthreads();
  if (on && c >= 0)
#pragma unroll 4
    for (int j = 0; j < n; ++j) {
      const int lj = dense_lane<G>(j, T);
      H[j * n + c] += alpha * r * (mhy * pb[lj] + vbuf[lj] * p) + alpha * alpha * (r * r * yhy + r) * p * pb[lj];
    }
  __syncthreads();
}
// the diagonal entry of this lane's dimension
template <int G>
__device__ __forceinline__ void dense_set_diag(float* H, float v, int l, int T, bool on) {
  const int n = 6 + T, c = dense_dim<G>(l, T);
  if (on && c >= 0) H[c * n + c] = v;
}

}  // namespace vdk
