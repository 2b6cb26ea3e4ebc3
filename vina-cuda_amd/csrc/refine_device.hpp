// refine_device.hpp -- FP64 device code of the step after the search: refine_structure against
// the receptor atoms (non_cache) and the parts of the final score, one pose per wavefront.
//
// The reference runs this step on the CPU in FP64 (V11/main/main.cpp:139-157, 425-440).  Near a
// minimum its BFGS line search decides on energy changes of 1e-7 kcal/mol, finer than FP32 atom
// positions resolve, so this path computes in FP64 throughout (the MI355X vector ALUs issue FP64
// FMAs at the rate of unpacked FP32) and keeps the reference's operation order where that is
// cheap: each node's kinematics and each parent's fold of its children take the same operations
// in the same order as the reference's preorder walks (the nodes of one tree depth at once, one
// lane per node), BFGS dot products and matrix rows are summed in index order, and nothing is
// contracted into FMAs, so the optimiser's decisions match an FP64 CPU run except where a
// summation order differs.
//
// Lane layout (lane l of the wavefront):
//   change vector (n = 6 + T <= 64):  component j in lane j -- 0-2 position, 3-5 orientation,
//                                     6 + k torsion k (the reference's change order)
//   conformation:  position and orientation replicated in every lane, torsion k in lane 6 + k
//   atoms:         atom i in lane i % 64; direct intermolecular term: one atom per row of 16
//   pairs:         slot l of each conflict-free chunk of 64 pair words
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vd_common.h"

// Every FP64 operation of this file is rounded on its own, as in the CPU code it restates: no
// multiply and add are fused into an FMA, in the small helpers as in the kernel bodies (hipcc
// fuses by default, and which products it fuses changes with the surrounding code).  The
// compiler's default is restored at the end of the file.
#pragma clang fp contract(off)

namespace vdr {

constexpr double kEps = 2.220446049250313e-16;      // epsilon_fl
constexpr double kMaxFl = 1.7976931348623157e308;   // max_fl
constexpr double kPi = 3.141592653589793;

// Ligand record of the FP64 path: atoms (local x, y, z relative to the node origin, meta =
// node | xs << 8 | heavy << 16 as a double), nodes (rel_origin + parent, rel_axis + atom range and
// subtree end), pairs in chunks of 64 words (a | b << 8 | tp << 16, VD_NO_PAIR for empty slots,
// each atom at most once as a and once as b per chunk).
struct RefHeader {
  int A, N, T, C;          // atoms, nodes, torsions, pair chunks
  long long off_atoms;     // double4 index of A atom words, then N origin words, N axis words
  long long off_pairs;     // uint32 index of C * 64 pair words
  int max_steps, heavy;
};
// Receptor atoms of the direct evaluation: cells n[0] x n[1] x n[2] over the box [lo, hi]
// (x fastest); cell c lists, from start[c], every typed receptor atom within the cutoff of its
// brick as (x, y, z, XS type).
struct NcD {
  const double4* __restrict__ atoms;
  const int* __restrict__ start;
  double lo[3], hi[3], inv_cell[3];
  int n[3];
};
struct Consts {
  const double4* __restrict__ lut;   // FP64 pair table quads (e_i, dor_i, e_i+1, dor_i+1), [tp][bin]
  NcD nc;
  double v;                          // curl cap of inter and intra (authentic_v)
};

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 scale(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
struct DQ { double a, b, c, d; };   // w, x, y, z
// Hamilton product (boost::math::quaternion operator*)
__device__ __forceinline__ DQ qmul(DQ p, DQ q) {
  return {p.a * q.a - p.b * q.b - p.c * q.c - p.d * q.d, p.a * q.b + p.b * q.a + p.c * q.d - p.d * q.c,
          p.a * q.c - p.b * q.d + p.c * q.a + p.d * q.b, p.a * q.d + p.b * q.c - p.c * q.b + p.d * q.a};
}
// normalize_angle (V11/lib/common.h:316-340)
__device__ inline double normalize_angle(double x) {
  for (int guard = 0; guard < 64; ++guard) {
    if (x > 3 * kPi) { const double n = (x - kPi) / (2 * kPi); x -= 2 * kPi * ceil(n); }
    else if (x < -3 * kPi) { const double n = (-x - kPi) / (2 * kPi); x += 2 * kPi * ceil(n); }
    else break;
  }
  if (x > kPi) x -= 2 * kPi;
  else if (x < -kPi) x += 2 * kPi;
  return x;
}
// angle_to_quaternion (V11/lib/quaternion.cpp:65-95)
__device__ inline DQ axis_angle(D3 axis, double angle) {
  angle = normalize_angle(angle);
  const double c = cos(angle / 2), s = sin(angle / 2);
  return {c, s * axis.x, s * axis.y, s * axis.z};
}
// quaternion_normalize_approx (V11/lib/quaternion.h:137-147)
__device__ inline DQ normalize_approx(DQ q) {
  const double s = q.a * q.a + q.b * q.b + q.c * q.c + q.d * q.d;
  if (fabs(s - 1) < 1e-6) return q;
  const double a = sqrt(s);
  return {q.a / a, q.b / a, q.c / a, q.d / a};
}
// quaternion_to_r3 (V11/lib/quaternion.cpp:155-185), row-major
__device__ inline void qmat(DQ q, double r[9]) {
  const double aa = q.a * q.a, ab = q.a * q.b, ac = q.a * q.c, ad = q.a * q.d, bb = q.b * q.b, bc = q.b * q.c,
               bd = q.b * q.d, cc = q.c * q.c, cd = q.c * q.d, dd = q.d * q.d;
  r[0] = aa + bb - cc - dd; r[1] = 2 * (-ad + bc);    r[2] = 2 * (ac + bd);
  r[3] = 2 * (ad + bc);     r[4] = aa - bb + cc - dd; r[5] = 2 * (-ab + cd);
  r[6] = 2 * (-ac + bd);    r[7] = 2 * (ab + cd);     r[8] = aa - bb - cc + dd;
}
__device__ __forceinline__ D3 mat_vec(const double* r, D3 v) {
  return {r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z, r[6] * v.x + r[7] * v.y + r[8] * v.z};
}
__device__ __forceinline__ uint32_t meta_of(double w) { return (uint32_t)w; }

// The value of v in lane j (j wave-uniform), in every lane.
__device__ __forceinline__ double lane_value(double v, int j) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, j);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Butterfly sum over `width` lanes (a power of two up to 64); every lane ends with the same sum.
template <int WIDTH>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
  for (int m = 1; m < WIDTH; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

// LDS of one pose (doubles): atoms 4A, frames 20N (R row-major 9, origin 3, lab axis 3, q 4, pad),
// xyz 4A, frc 4A, node sums 8N, tree links N (one int per node: previous sibling + 1), torsion
// half-angle cosines 64 and sines 64, gradient 64, H n x n.
struct Pose {
  const double4* atom;      // LDS copy of the atom words
  const double4* node_o;    // global: rel_origin xyz + (parent + 1) | begin << 8 | end << 16
  const double4* node_a;    // global: rel_axis xyz + subtree end
  const uint32_t* pairs;    // global
  int A, N, T, C, n;
  double* frame; double4* xyz; double4* frc; double4* nsum; int* sib;
  double* tcos; double* tsin; double* grad; double* H;
};
__host__ __device__ inline int pose_lds_doubles(int A, int N, int n) { return 12 * A + 29 * N + 3 * 64 + n * n; }

__device__ inline void carve(Pose& P, double* lds, int Amax, int Nmax) {
  double* p = lds;
  P.atom = reinterpret_cast<const double4*>(p); p += 4 * Amax;
  P.frame = p; p += 20 * Nmax;
  P.xyz = reinterpret_cast<double4*>(p); p += 4 * Amax;
  P.frc = reinterpret_cast<double4*>(p); p += 4 * Amax;
  P.nsum = reinterpret_cast<double4*>(p); p += 8 * Nmax;
  P.sib = reinterpret_cast<int*>(p); p += Nmax;
  P.tcos = p; p += 64;
  P.tsin = p; p += 64;
  P.grad = p; p += 64;
  P.H = p;
}

// Conformation: position and orientation in every lane, torsion k in lane 6 + k.
struct Conf { D3 pos; DQ q; double tor; };

// Tree links of node k: depth, last child and previous sibling (-1: none).
struct Links { int depth, last_child, prev_sibling; };
__device__ __forceinline__ Links links_of(const Pose& P, int k) {
  const uint32_t w = meta_of(P.node_a[k].w);
  return Links{(int)((w >> 8) & 255u), (int)((w >> 16) & 255u) - 1, (int)(w >> 24) - 1};
}
// Copies every node's previous sibling into LDS (the adjoint fold walks siblings there).  Call
// once per pose, before the first evaluation.
__device__ inline void load_links(const Pose& P, int l) {
  if (l < P.N) P.sib[l] = links_of(P, l).prev_sibling;
  __syncthreads();
}
// Depth of the deepest node (wave-uniform).
__device__ __forceinline__ int max_depth(const Pose& P, int l) {
  const int d = l < P.N ? links_of(P, l).depth : 0;
  int D = 0;
  while (__any(d > D)) ++D;
  return D;
}

// model::set for one pose (tree.h:205-209, 282-292 as the oracle's set_conf restates it):
// q_k = angle_to_quaternion(R_parent rel_axis, theta_k) * q_parent (normalize_approx),
// o_k = o_parent + R_parent rel_origin, node by node from the root: lane k computes node k once
// every node of the depth above is done (the operations of the reference's preorder walk, each
// node's from its parent's frame); then every lane places its atoms, x_i = o_node + R_node
// local_i.  The half-angle cosine and sine of angle_to_quaternion are computed beforehand by each
// torsion's own lane.  Leaves frames and absolute coordinates in LDS.
__device__ inline void set_coords(const Pose& P, const Conf& c, int l) {
  if (l >= 6 && l < 6 + P.T) {
    const double a = normalize_angle(c.tor);
    P.tcos[l - 6] = cos(a / 2);
    P.tsin[l - 6] = sin(a / 2);
  }
  if (l == 0) {
    double* f = P.frame;
    qmat(c.q, f);
    f[9] = c.pos.x; f[10] = c.pos.y; f[11] = c.pos.z; f[12] = 0; f[13] = 0; f[14] = 0;
    f[15] = c.q.a; f[16] = c.q.b; f[17] = c.q.c; f[18] = c.q.d;
  }
  const int depth = l < P.N ? links_of(P, l).depth : 0;
  for (int d = 1; __any(d <= depth); ++d) {
    __syncthreads();
    if (l >= 1 && l < P.N && depth == d) {
      const int k = l;
      const double4 no = P.node_o[k], na = P.node_a[k];
      const int par = (int)(meta_of(no.w) & 255u) - 1;
      const double* fp = P.frame + 20 * par;
      double* fk = P.frame + 20 * k;
      const D3 axis = mat_vec(fp, D3{na.x, na.y, na.z});
      const D3 org = add(D3{fp[9], fp[10], fp[11]}, mat_vec(fp, D3{no.x, no.y, no.z}));
      const double hc = P.tcos[k - 1], hs = P.tsin[k - 1];
      const DQ q = normalize_approx(qmul(DQ{hc, hs * axis.x, hs * axis.y, hs * axis.z}, DQ{fp[15], fp[16], fp[17], fp[18]}));
      qmat(q, fk);
      fk[9] = org.x; fk[10] = org.y; fk[11] = org.z; fk[12] = axis.x; fk[13] = axis.y; fk[14] = axis.z;
      fk[15] = q.a; fk[16] = q.b; fk[17] = q.c; fk[18] = q.d;
    }
  }
  __syncthreads();
  for (int i = l; i < P.A; i += 64) {
    const double4 a = P.atom[i];
    const double* f = P.frame + 20 * (int)(meta_of(a.w) & 255u);
    const D3 y = add(D3{f[9], f[10], f[11]}, mat_vec(f, D3{a.x, a.y, a.z}));
    P.xyz[i] = make_double4(y.x, y.y, y.z, 0.0);
  }
  __syncthreads();
}

// non_cache::eval_deriv (FAST = false) / non_cache::eval (FAST = true), V11/lib/non_cache.cpp:
// 38-128: each typed atom is clamped to the box (slope x the distance outside joins the energy,
// +-slope the force); the receptor atoms listed for the cell of the clamped point are scored
// with the table (interpolated value and dor, or the bin's fast value); the atom's sum is capped
// by curl.  One atom per row of 16 lanes.  FAST = false writes dE/dx of every atom to frc.
// Returns this lane's share of the energy.
template <bool FAST>
__device__ inline double nc_inter(const Pose& P, const Consts& K, double slope, int l) {
  const int row = l >> 4, rl = l & 15;
  double e_lane = 0;
  for (int i0 = 0; i0 < P.A; i0 += 4) {
    const int i = i0 + row;
    const int t1 = i < P.A ? (int)((meta_of(P.atom[i].w) >> 8) & 255u) : VD_XS_SIZE;
    double e = 0, dx = 0, dy = 0, dz = 0, pen = 0;
    double ob[3] = {0, 0, 0};
    if (t1 < VD_XS_SIZE) {
      const double4 y = P.xyz[i];
      double p[3] = {y.x, y.y, y.z};
      int ci[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (p[k] < K.nc.lo[k]) { pen += fabs(p[k] - K.nc.lo[k]); p[k] = K.nc.lo[k]; ob[k] = -1; }
        else if (p[k] > K.nc.hi[k]) { pen += fabs(p[k] - K.nc.hi[k]); p[k] = K.nc.hi[k]; ob[k] = 1; }
        const double cf = (p[k] - K.nc.lo[k]) * K.nc.inv_cell[k];   // NaN lands in cell 0, overflow in the last
        ci[k] = cf >= 0 ? (int)fmin(cf, (double)(K.nc.n[k] - 1)) : 0;
      }
      const int cell = (ci[2] * K.nc.n[1] + ci[1]) * K.nc.n[0] + ci[0];
      const int end = K.nc.start[cell + 1];
      // four candidates per pass: their atom words, then their table quads, are all loaded
      // before the first is used; each lane still adds its candidates in list order
      constexpr int U = 4;
      for (int k0 = K.nc.start[cell] + rl; k0 < end; k0 += 16 * U) {
        double4 a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = K.nc.atoms[k0 + 16 * u < end ? k0 + 16 * u : k0];
        double rx[U], ry[U], rz[U], rem[U];
        int idx[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          rx[u] = p[0] - a[u].x; ry[u] = p[1] - a[u].y; rz[u] = p[2] - a[u].z;
          const double r2 = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
          ok[u] = k0 + 16 * u < end && r2 < VD_CUTOFF_SQR;
          const int t2 = (int)a[u].w;
          const int tp = t1 <= t2 ? t1 + t2 * (t2 + 1) / 2 : t2 + t1 * (t1 + 1) / 2;
          const double rf = 32.0 * r2;
          const int b = ok[u] ? (int)rf : 0;
          rem[u] = rf - b;
          idx[u] = ok[u] ? tp * VD_LUT_BINS + b : 0;
        }
        double4 q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = K.lut[idx[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!ok[u]) continue;
          if constexpr (FAST) {
            e += (q[u].z + q[u].x) / 2;
          } else {
            e += q[u].x + rem[u] * (q[u].z - q[u].x);
            const double dor = q[u].y + rem[u] * (q[u].w - q[u].y);
            dx += dor * rx[u]; dy += dor * ry[u]; dz += dor * rz[u];
          }
        }
      }
    }
    e = lane_sum<16>(e);
    if constexpr (!FAST) { dx = lane_sum<16>(dx); dy = lane_sum<16>(dy); dz = lane_sum<16>(dz); }
    if (e > 0 && K.v < 0.1 * kMaxFl) {   // curl (V11/lib/curl.h:4-13, 31-38)
      const double t = K.v < kEps ? 0 : K.v / (K.v + e);
      e *= t;
      const double t2 = t * t;
      dx *= t2; dy *= t2; dz *= t2;
    }
    if (rl == 0 && i < P.A) {
      if constexpr (!FAST) P.frc[i] = make_double4(dx + slope * ob[0], dy + slope * ob[1], dz + slope * ob[2], 0.0);
      e_lane += e + pen * slope;
    }
  }
  return e_lane;
}

// Intramolecular pairs (eval_interacting_pairs_deriv, V11/lib/model.cpp:1241-1259, FAST = false;
// eval_interacting_pairs, model.cpp:1201-1214, FAST = true), each term capped by curl.  Within a
// chunk the first atoms are distinct and so are the second atoms, so the force scatter is a
// plain read-modify-write of LDS in chunk order.  Returns this lane's share of the energy.
template <bool FAST>
__device__ inline double pair_terms(const Pose& P, const Consts& K, int l) {
  double e_lane = 0;
  uint32_t w_next = P.C > 0 ? P.pairs[l] : VD_NO_PAIR;
  for (int ch = 0; ch < P.C; ++ch) {
    const uint32_t w = w_next;
    if (ch + 1 < P.C) w_next = P.pairs[(ch + 1) * 64 + l];
    const bool ok = w != VD_NO_PAIR;
    const int a = ok ? (int)(w & 255u) : 0, b = ok ? (int)((w >> 8) & 255u) : 0;
    const double4 pa = P.xyz[a], pb = P.xyz[b];
    const double rx = pb.x - pa.x, ry = pb.y - pa.y, rz = pb.z - pa.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    double fx = 0, fy = 0, fz = 0;
    if (ok && r2 < VD_CUTOFF_SQR) {
      const double rf = 32.0 * r2;
      const int bin = (int)rf;
      const double4 q = K.lut[(int)(w >> 16) * VD_LUT_BINS + bin];
      if constexpr (FAST) {
        double t = (q.z + q.x) / 2;
        if (t > 0 && K.v < 0.1 * kMaxFl) t *= K.v < kEps ? 0 : K.v / (K.v + t);
        e_lane += t;
      } else {
        const double rem = rf - bin;
        double e = q.x + rem * (q.z - q.x);
        double dor = q.y + rem * (q.w - q.y);
        fx = dor * rx; fy = dor * ry; fz = dor * rz;
        if (e > 0 && K.v < 0.1 * kMaxFl) {
          const double t = K.v < kEps ? 0 : K.v / (K.v + e);
          e *= t;
          const double t2 = t * t;
          fx *= t2; fy *= t2; fz *= t2;
        }
        e_lane += e;
      }
    }
    if constexpr (!FAST) {
      if (ok) { double4 f = P.frc[a]; f.x -= fx; f.y -= fy; f.z -= fz; P.frc[a] = f; }
      if (ok) { double4 f = P.frc[b]; f.x += fx; f.y += fy; f.z += fz; P.frc[b] = f; }
    }
  }
  return e_lane;
}

// model::eval_deriv over non_cache with the intramolecular pairs (model.cpp:1324-1334) at the
// given slope.  Returns e (wave-uniform); g = this lane's gradient component (0 beyond n).
// The adjoint follows the oracle's order: per-node force and torque sums over the node's atoms
// in index order, then every parent folds in its children in decreasing index order, which is
// the order the reference's reverse-preorder walk adds them in; parents one depth at a time from
// the deepest, so a child is complete before it is folded.
__device__ inline double eval_deriv(const Pose& P, const Consts& K, double slope, const Conf& c, int l, double& g) {
  set_coords(P, c, l);
  double e = nc_inter<false>(P, K, slope, l);
  __syncthreads();
  e += pair_terms<false>(P, K, l);
  __syncthreads();
  if (l < P.N) {
    const double4 no = P.node_o[l];
    const uint32_t bits = meta_of(no.w);
    const int begin = (int)((bits >> 8) & 255u), end = (int)((bits >> 16) & 255u);
    const double* f = P.frame + 20 * l;
    const D3 o{f[9], f[10], f[11]};
    D3 F{0, 0, 0}, T{0, 0, 0};
    for (int i = begin; i < end; ++i) {
      const double4 fr = P.frc[i], x = P.xyz[i];
      const D3 fi{fr.x, fr.y, fr.z};
      F = add(F, fi);
      T = add(T, cross(sub(D3{x.x, x.y, x.z}, o), fi));
    }
    P.nsum[2 * l] = make_double4(F.x, F.y, F.z, 0.0);
    P.nsum[2 * l + 1] = make_double4(T.x, T.y, T.z, 0.0);
  }
  const Links lk = l < P.N ? links_of(P, l) : Links{-1, -1, -1};
  for (int d = max_depth(P, l) - 1; d >= 0; --d) {
    __syncthreads();
    if (l < P.N && lk.depth == d && lk.last_child >= 0) {
      const double* fp = P.frame + 20 * l;
      double4 Fp = P.nsum[2 * l], Tp = P.nsum[2 * l + 1];
      for (int k = lk.last_child; k >= 0; k = P.sib[k]) {
        const double* f = P.frame + 20 * k;
        const double4 Fk = P.nsum[2 * k], Tk = P.nsum[2 * k + 1];
        P.grad[6 + k - 1] = dot(D3{Tk.x, Tk.y, Tk.z}, D3{f[12], f[13], f[14]});
        const D3 arm = sub(D3{f[9], f[10], f[11]}, D3{fp[9], fp[10], fp[11]});
        const D3 c = cross(arm, D3{Fk.x, Fk.y, Fk.z});
        Fp.x += Fk.x; Fp.y += Fk.y; Fp.z += Fk.z;
        Tp.x = (Tp.x + c.x) + Tk.x; Tp.y = (Tp.y + c.y) + Tk.y; Tp.z = (Tp.z + c.z) + Tk.z;
      }
      P.nsum[2 * l] = Fp; P.nsum[2 * l + 1] = Tp;
    }
  }
  __syncthreads();
  if (l == 0) {
    const double4 F0 = P.nsum[0], T0 = P.nsum[1];
    P.grad[0] = F0.x; P.grad[1] = F0.y; P.grad[2] = F0.z; P.grad[3] = T0.x; P.grad[4] = T0.y; P.grad[5] = T0.z;
  }
  __syncthreads();
  g = l < P.n ? P.grad[l] : 0.0;
  return lane_sum<64>(e);
}

// x (+) factor * d (conf.h:60-66, 118-124; quaternion_increment, quaternion.cpp:235-240).
__device__ inline Conf increment(const Conf& x, double d, double factor, int l, int T) {
  Conf y = x;
  const double d0 = __shfl(d, 0), d1 = __shfl(d, 1), d2 = __shfl(d, 2);
  const double d3 = __shfl(d, 3), d4 = __shfl(d, 4), d5 = __shfl(d, 5);
  y.pos.x += factor * d0; y.pos.y += factor * d1; y.pos.z += factor * d2;
  const D3 rot{factor * d3, factor * d4, factor * d5};
  const double angle = sqrt(dot(rot, rot));
  DQ q = x.q;
  if (angle > kEps) q = qmul(axis_angle(scale(1 / angle, rot), angle), q);
  y.q = normalize_approx(q);
  if (l >= 6 && l < 6 + T) y.tor = normalize_angle(x.tor + normalize_angle(factor * d));
  return y;
}

// Dot product of two lane vectors summed in index order (every lane gets the same value): each
// lane's product, then the products of lanes 0 .. n-1 added in turn.
__device__ inline double ordered_dot(const Pose& P, double a, double b, int l) {
  const double ab = a * b;
  double s = 0;
  for (int j = 0; j < P.n; ++j) s += lane_value(ab, j);
  return s;
}
// -(H v) for this lane's row, summed in index order.  The rows are read after a fence, so every
// update of H before the call is seen.
__device__ inline double minus_hv(const Pose& P, double v, int l) {
  __syncthreads();
  double s = 0;
  for (int j = 0; j < P.n; ++j) {
    const double vj = lane_value(v, j);
    if (l < P.n) s += P.H[l * P.n + j] * vj;
  }
  return -s;
}

// bfgs (V11/lib/bfgs.h:30-140): H starts as the identity, p = -H g, backtracking line search of
// at most 10 halvings from alpha = 1 with the Armijo constant 1e-4, the first step rescales the
// diagonal, the rank-two update skips steps with alpha y.p < epsilon; a result above the start
// reverts to the start.  No early stop beyond the gradient-norm test.  Returns f; x is updated.
__device__ inline double bfgs(const Pose& P, const Consts& K, double slope, Conf& x, int max_steps, int l, int& evals) {
  const int n = P.n;
  if (l < n)
    for (int j = 0; j < n; ++j) P.H[l * n + j] = (j == l) ? 1.0 : 0.0;
  double g;
  double f0 = eval_deriv(P, K, slope, x, l, g);
  ++evals;
  const double f_orig = f0;
  const Conf x_orig = x;
  for (int step = 0; step < max_steps; ++step) {
    const double p = minus_hv(P, g, l);
    const double pg = ordered_dot(P, p, g, l);
    double f1 = 0, alpha = 1, g_new = 0;
    Conf x_new = x;
    for (int trial = 0; trial < 10; ++trial) {
      x_new = increment(x, p, alpha, l, P.T);
      f1 = eval_deriv(P, K, slope, x_new, l, g_new);
      ++evals;
      if (f1 - f0 < 1e-4 * alpha * pg) break;
      alpha *= 0.5;
    }
    const double y = g_new - g;
    f0 = f1;
    x = x_new;
    if (!(sqrt(ordered_dot(P, g, g, l)) >= 1e-5)) break;
    g = g_new;
    if (step == 0) {
      const double yy = ordered_dot(P, y, y, l);
      if (fabs(yy) > kEps) {
        const double dg = alpha * ordered_dot(P, y, p, l) / yy;
        __syncthreads();
        if (l < n) P.H[l * n + l] = dg;
        __syncthreads();
      }
    }
    const double yp = ordered_dot(P, y, p, l);
    if (alpha * yp < kEps) continue;
    const double mhy = minus_hv(P, y, l);
    const double yhy = -ordered_dot(P, y, mhy, l);
    const double r = 1 / (alpha * yp);
    // lane l updates H[l][j] for j >= l and mirrors it, as the reference's packed triangle
    __syncthreads();
    for (int j = 0; j < n; ++j) {
      const double pj = lane_value(p, j), mhyj = lane_value(mhy, j);
      if (l < n && j >= l) {
        const double h = P.H[l * n + j] + (alpha * r * (mhy * pj + mhyj * p) + alpha * alpha * (r * r * yhy + r) * p * pj);
        P.H[l * n + j] = h;
        P.H[j * n + l] = h;
      }
    }
    __syncthreads();
  }
  if (!(f0 <= f_orig)) { f0 = f_orig; x = x_orig; }
  return f0;
}

// non_cache::within (V11/lib/non_cache.cpp:58-66, margin 1e-4 from non_cache.h:43): every heavy
// atom inside the box.  Call after set_coords.
__device__ inline bool within(const Pose& P, const Consts& K, int l) {
  bool out = false;
  for (int i = l; i < P.A; i += 64) {
    if (!((meta_of(P.atom[i].w) >> 16) & 1u)) continue;
    const double4 y = P.xyz[i];
    const double p[3] = {y.x, y.y, y.z};
    for (int k = 0; k < 3; ++k)
      if (p[k] < K.nc.lo[k] - 1e-4 || p[k] > K.nc.hi[k] + 1e-4) out = true;
  }
  return !__any(out);
}

}  // namespace vdr

#pragma clang fp contract(fast)
