// vd_common.h -- definitions shared by the host side and the HIP kernels of libvdock.
//
// Scoring constants restate V11/lib/atom_constants.h (XS types, radii, flags) and
// V11/lib/precalculate.h:255-283 (LUT geometry: factor 32, cutoff 8 A, 2051 bins).
#pragma once
#include <stdint.h>

#define VD_XS_SIZE 17
#define VD_NUM_TYPE_PAIRS 153        // 17*18/2 triangular pairs, index i + j(j+1)/2 for i <= j
#define VD_LUT_BINS 2051             // int(32 * 64) + 3
#define VD_LUT_FACTOR 32.0f
#define VD_CUTOFF_SQR 64.0f
#define VD_ZERO_QUAD (VD_NUM_TYPE_PAIRS * VD_LUT_BINS)   // first entry of the all-zero row after the pair table
#define VD_GRID_ROW 16               // per (t1, bin) row of the map LUT: the values of t2 = 0..15 (64 B,
                                     // four aligned 16-byte loads); the t2 = 16 values follow all rows
                                     // as a [t1][bin] column
#define VD_MAX_ATOMS 255             // per ligand: atom indices are 8-bit fields of the pair words and node
                                     // atom ranges (the reference caps at 128, MAX_NUM_OF_ATOMS, V11/inc/kernel2.h:35)
#define VD_MAX_NODES 64              // rigid nodes (torsions + 1)
#define VD_MAX_PAIRS (VD_MAX_ATOMS * (VD_MAX_ATOMS - 1) / 2)   // every pair of atoms (reference: 2048)

// Device ligand record: all ligands of a job live in one blob of 16-byte words.
// Atom word:  x,y,z = coordinates relative to the owning node's origin; w = meta bits
//             (node | xs << 8 | heavy << 16 | heavy slot << 17).
// Node words: o = (rel_origin xyz, parent); a = (rel_axis xyz, subtree_end) -- preorder, the
//             subtree of node k is the node range [k, subtree_end).
// Pair word:  a | b << 8 | tp << 16.  Pairs are stored in chunks of `chunk_lanes` slots (one per
//             lane of a chain group); within a chunk every atom is the `a` of at most one pair
//             and the `b` of at most one pair, so the force scatter of a chunk needs no atomics.
//             Unused slots hold VD_NO_PAIR.  When the ligand has at most chunk_lanes atoms
//             (pair_rows = 1), the pair in slot l has first atom l and its word is
//             src | b << 8 | tp << 16 instead: src = the slot of the same chunk whose pair has
//             second atom l (chunk_lanes: none); a slot without a pair has b = l (0 for lanes
//             beyond the ligand) and tp = VD_NUM_TYPE_PAIRS, the all-zero row of the pair table.
//             The area is laid out in groups of four chunks, [group][lane][4].
#define VD_NO_PAIR 0xFFFFFFFFu
struct vd_lig_header {
  int n_atoms, n_nodes, n_pairs, n_heavy;
  int off_atoms;   // float4 index into the blob
  int off_nodes;   // float4 index: n_nodes origin words then n_nodes axis words
  int off_pairs;   // uint32 index into the blob (viewed as uint32)
  int jump_rounds; // ceil(log2(tree depth + 1)) pointer-jumping rounds
  float origin0[3];
  int n_torsions;
  int search_depth;   // MC steps for this ligand (filled per launch)
  int bfgs_steps;     // local-search iterations for this ligand (filled per launch)
  int n_chunks;       // pair chunks; the pair area holds n_chunks * chunk_lanes words
  int chunk_lanes;    // lanes per chain group the chunks were built for
  int pair_rows;      // 1: slot l of every chunk holds a pair whose first atom is l
  int reserved;
};

// Uniform-grid description of the 17 maps, plain FP32 layout map[t][z][y][x].
struct vd_grid_desc {
  float init[3];
  float factor[3];
  float factor_inv[3];
  float dim_fl_minus_1[3];
  int dim[3];
  long long plane;   // dim0*dim1*dim2
};

// Docking constants passed to the MC kernels.
struct vd_mc_consts {
  float box_lo[3], box_hi[3];
  float v_intra, v_inter, slope;
  float temperature, amplitude;
  int search_depth, bfgs_steps, local_search;   // local_search: 1 RILC-BFGS, 0 dense BFGS
  int mutation_quirk;                           // 1: one of 1+T entities (reference GPU host tables)
  int policy;                                   // 0 Vina schedule, 1 QuickVina2 visited-history gate
  uint32_t seed;
  int chains;                                   // chains per ligand
};

#define VD_INIT_STEP 0xFFFFFFFFu

// Device memory of the no-allocation docking entry point (vd_dock_noalloc): one array per device
// for the ligand record, launch bookkeeping and search histories of one call.
#define VD_NOALLOC_SCRATCH_BYTES (4u << 20)
