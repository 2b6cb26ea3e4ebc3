// kernels.h -- launch interface between the C ABI layer and the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "vd_common.h"
#include "dock_device.hpp"
#include "refine_device.hpp"

// Receptor cell list for the grid kernel: atoms sorted by cell (x fastest), one (begin, end)
// range per cell; cells of `cell` A cover the box grown by the cutoff, starting at `origin`.
struct GridCells {
  float origin[3];
  float cell;
  int n[3];
};

struct McArgs {
  const float4* blob;
  const vd_lig_header* headers;     // per ligand slot
  const uint32_t* streams;          // random stream id per ligand slot
  const int2* items;                // (ligand slot, first chain); nullptr: item k = (0, k x chains per wave)
  int n_items;
  const float* init_conf;           // nullptr: starting conformations drawn from the random stream; else
                                    // per chain [7+T] at the slot's conformation offset (out_off.y)
  int ls_only;                      // 1: one local search from each starting conformation, no MC steps
  int* queue;                       // nullptr: item = blockIdx.x
  vdk::EvalK K;
  vd_mc_consts P;
  float* out_e; float* out_conf; float* out_heavy;
  const longlong4* out_off;         // per ligand slot: (e, conf, heavy) element offsets
  unsigned long long* slot_evals;    // per ligand slot: evaluations performed (zeroed by the caller)
  int Amax, Nmax;
  float* hist;                      // QuickVina2 / QuickVina-W per-chain histories (nullptr: Vina policy),
  long long hist_stride;            // floats per chain region; region = block * chains/wave + group
  float* glob;                      // QuickVina-W global buffers: ligand slot k at glob + out_off[k].w,
                                    // header.reserved cells per axis
};

struct EvalArgs {
  const float4* blob;
  const vd_lig_header* header;      // device copy
  vd_lig_header host_header;
  vdk::EvalK K;
  const float* conf; long long n_conf;
  float* e; float* grad;
};

// Refinement launch (vd_refine_kernel, FP64): one pose per workgroup.
struct RefineArgs {
  const double4* geo;               // ligand atom and node words (refine_device.hpp RefHeader)
  const uint32_t* pairs;            // pair chunks of 64 words
  const vdr::RefHeader* hdr;        // per ligand slot; max_steps = the step budget per stage
  const int* pose_lig;              // ligand slot per pose
  const long long* conf_off;        // per pose: offset of its conformation (7 + T doubles)
  const long long* coords_off;      // per pose: offset of its coordinates (3 A doubles)
  const double* conf_in;
  double* conf_out; double* coords_out;
  double4* res;                     // per pose: (refined e, inter, intra, inside the box)
  int* evals;                       // per pose: evaluations
  vdr::Consts K;
  double final_slope;
  int Amax, Nmax, nmax;
};

namespace vdl {
int refine_lds_bytes(int Amax, int Nmax, int nmax);
hipError_t launch_refine(const RefineArgs& a, int poses, hipStream_t s);
int mc_lds_bytes(int G, int Amax, int Nmax, int dense);
int max_resident_mc_blocks(int G, int W, int mode, int lds_bytes);
hipError_t launch_grid(const float4* atoms, const int2* cells, const GridCells& gc, const float4* lut_rows, float* maps,
                       const vd_grid_desc& gd, hipStream_t s);
// Corner records of the maps (VD_XS_SIZE * (dim0-1)(dim1-1)(dim2-1) * 2 float4) from the
// canonical maps, both on the device.
hipError_t launch_corners(const float* maps, float4* corners, const vd_grid_desc& gd, hipStream_t s);
hipError_t launch_eval(int G, const EvalArgs& a, int blocks, hipStream_t s);
hipError_t launch_mc(int G, int W, const McArgs& a, int blocks, hipStream_t s);
// The kernel function launch_mc launches for these arguments (to find its node in a captured graph).
const void* mc_kernel_fn(int G, int W, const McArgs& a);
hipError_t reset_counter(int* p, hipStream_t s);
hipError_t glob_init(float* base, const long long* off, const int* per_axis, const int* T, int n, hipStream_t s);
hipError_t random_confs(float* conf, long long n, int T, uint32_t seed, const float lo[3], const float hi[3], hipStream_t s);
hipError_t zero_u64(unsigned long long* p, int n, hipStream_t s);
// The device address of this device's scratch array (the memory of the no-allocation entry
// point); call with the device current.
hipError_t scratch_address(void** p);
}  // namespace vdl
