/*
 * vdock.h -- C ABI of libvdock, the MI355X (gfx950) docking hot path.
 *
 * Plain C types and pointers only.  Every call returns VD_OK (0) or a negative vd_status and
 * never throws; the message of the last failure is available from vd_last_error().  A context
 * works on one device and must be used from one host thread at a time.  The contexts of a
 * process on one device share the library's HIP streams (four per device, made on first use and
 * kept until the process ends), so contexts used from several threads at once run their device
 * work in turn.
 * Device buffers are owned by the library; host buffers belong to the caller and are copied.
 *
 * The reference path this replaces (paths under the reference root, V11 = Vina-CUDA_v1.1):
 *   kernel_grid   V11/inc/cuda/kernel1.cuh:30-41   -> vd_grid_populate
 *   kernel_monte  V11/inc/cuda/kernel2.cuh:25-38, defined at kernel2.cu:567   -> vd_dock / vd_dock_batch
 *   main_procedure_cl V11/lib/main_procedure_cu.h:2-4 (host driver)  -> the calls below in sequence
 *   cache::populate   V11/lib/cache.h:73            -> vd_grid_populate
 *   model::eval_deriv V11/lib/model.cpp:1324-1334   -> vd_eval
 *   refine_structure + rescoring + output, V11/main/main.cpp:139-178, 423-501
 *                     -> vd_refine, vd_postprocess_batch / vd_postprocess_docked, vd_write_pdbqt
 */
#ifndef VDOCK_H
#define VDOCK_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct vd_ctx vd_ctx;

enum {
  VD_OK = 0,
  VD_ERR_ARG = -1,        /* invalid argument or call order */
  VD_ERR_CAPACITY = -2,   /* input beyond a supported size (atoms, torsions, pairs, box) */
  VD_ERR_HIP = -3,        /* HIP runtime failure; see vd_last_error */
  VD_ERR_PARSE = -4,      /* malformed PDBQT input */
  VD_ERR_NO_DEVICE = -5   /* no usable GPU */
};

/* Library version string. */
const char* vd_version(void);

/* Context on HIP device `device` (replaces cudaSetDevice + stream setup,
 * V11/lib/main_procedure_cu.cpp:219-228).  Default scoring weights are installed. */
int vd_create(int device, vd_ctx** out);
void vd_destroy(vd_ctx* ctx);
const char* vd_last_error(const vd_ctx* ctx);

/* Scoring weights w[0..5] = gauss1, gauss2, repulsion, hydrophobic, hydrogen bond, rot
 * (V11/main/main.cpp:649-654, 856-862) -> FP64 look-up tables (precalculate,
 * V11/lib/precalculate.h:255-283) uploaded as FP32 (replaces the pre_cl upload,
 * V11/lib/main_procedure_cu.cpp:300-356). */
int vd_set_weights(vd_ctx* ctx, const double w[6]);
/* FP64 tables of type pair tp (0..152): smooth value, smooth derivative/r, fast (2051 each). */
int vd_get_lut(vd_ctx* ctx, int tp, double* smooth_e, double* smooth_dor, double* fast);

/* Receptor atoms: coordinates xyz[3n] and XS types (17 = untyped, ignored)
 * (replaces the pa_cl upload, V11/lib/main_procedure_cu.cpp:246-272). */
int vd_set_receptor(vd_ctx* ctx, int n, const float* xyz, const int* xs);
/* Same from a rigid receptor PDBQT text (parse_receptor_pdbqt, V11/lib/parse_pdbqt.cpp:643-652). */
int vd_set_receptor_pdbqt(vd_ctx* ctx, const char* text, int* n_atoms);
/* Copies up to max receptor atoms; returns the count (>= 0) or an error. */
int vd_get_receptor(vd_ctx* ctx, int max, float* xyz, int* xs);

/* Search box as Vina builds it: n = ceil(size/granularity) intervals per axis, centred
 * (V11/main/main.cpp:864-876).  Maps have n+1 points per axis. */
typedef struct {
  double begin[3];
  double end[3];
  int n[3];
} vd_box;
int vd_box_from_center(const double center[3], const double size[3], double granularity, vd_box* box);

/* All 17 XS maps over the box on the GPU (replaces kernel_grid, V11/inc/cuda/kernel1.cuh:30-41,
 * and cache::populate, V11/lib/cache.cpp:148-208).  Maps stay resident on the device in the
 * plain FP32 layout map[t][z][y][x], beside the copy the searches read: per map and grid cell,
 * the cell's eight corner values (32 bytes), built from the plain maps by the same call.  The
 * device then holds 17*4*(n0+1)*(n1+1)*(n2+1) + 17*32*n0*n1*n2 bytes of maps (314 MB for an
 * 80^3-step box); VD_ERR_HIP if that does not fit. */
int vd_grid_populate(vd_ctx* ctx, const vd_box* box);
/* Kernel time of the last vd_grid_populate in milliseconds (the map precompute, without the
 * corner copy). */
double vd_grid_last_ms(const vd_ctx* ctx);
/* Host copies of the resident maps: 17*(n0+1)*(n1+1)*(n2+1) floats. */
int vd_grid_download(vd_ctx* ctx, float* maps);
/* Maps from the host in the plain layout (as vd_grid_download returns them); builds the corner
 * copy as vd_grid_populate does. */
int vd_grid_upload(vd_ctx* ctx, const vd_box* box, const float* maps);

/* Ligands (parse_ligand_pdbqt, V11/lib/parse_pdbqt.cpp:618-627, flattened for the device as
 * V11/lib/main_procedure_cu.cpp:551-763 does).  Returns the ligand id. */
int vd_ligand_add_pdbqt(vd_ctx* ctx, const char* text, int* id);
typedef struct {
  int n_atoms;      /* movable atoms, hydrogens included */
  int n_torsions;   /* rigid nodes - 1 = degrees of freedom beyond 6 */
  int n_pairs;      /* intramolecular interacting pairs */
  int n_heavy;      /* non-hydrogen atoms (the pose coordinates reported per chain) */
  int torsdof;      /* TORSDOF record */
  int tree_depth;   /* depth of the rigid-node tree */
  float num_tors;   /* torsion count of the final score's divisor (conf_independent_inputs,
                       V11/lib/terms.cpp:97-125) */
} vd_ligand_info;
int vd_ligand_info_get(vd_ctx* ctx, int id, vd_ligand_info* info);
/* Flattened model: per atom xs, element class (0 = H), node, node-local xyz; per node parent,
 * atom range [begin, end), rel_origin, rel_axis; pairs as (type pair, a, b) triples. */
int vd_ligand_export(vd_ctx* ctx, int id, int* xs, int* el, int* node, float* local, int* parent,
                     int* begin, int* end, float* rel_origin, float* rel_axis, int* pairs);
int vd_ligand_count(const vd_ctx* ctx);
/* Device layout of the intramolecular pairs for chain groups of `lanes` lanes (16, 32, 64):
 * chunks of `lanes` pair words (a | b << 8 | tp << 16, 0xFFFFFFFF = empty slot), each atom at most
 * once as a and at most once as b per chunk.  Copies up to max words; returns the word count
 * (chunks x lanes).  Layout of this implementation, no counterpart in the reference (which
 * scatters pair forces without synchronisation, V11/inc/cuda/quasi_newton.cuh:941-944). */
int vd_ligand_pair_chunks(vd_ctx* ctx, int id, int lanes, int max, uint32_t* words);

/* Energy and gradient of n conformations of ligand `id` (model::eval_deriv,
 * V11/lib/model.cpp:1324-1334; device form V11/inc/cuda/quasi_newton.cuh:1294-1339).
 * conf[n][7+T] = position, quaternion (w,x,y,z), torsions; grad[n][6+T] = force, torque,
 * torsion derivatives; e[n].  Curl caps v_intra / v_inter and the out-of-box slope as the
 * reference's hunt_cap (10) and slope (1e6).  device_ptrs = 1: the three pointers are device
 * pointers (no copies).  chains_per_wave: 1, 2 or 4 (0 = pick). */
int vd_eval(vd_ctx* ctx, int id, long long n, const float* conf, float* e, float* grad,
            float v_intra, float v_inter, float slope, int chains_per_wave, int device_ptrs);
/* Kernel time of the last vd_eval in milliseconds. */
double vd_eval_last_ms(const vd_ctx* ctx);

/* Docking parameters (monte_carlo, V11/lib/monte_carlo.h:60-70; kernel_monte arguments). */
enum { VD_LS_BFGS = 0, VD_LS_RILC_BFGS = 1 };
enum { VD_MC_VINA = 0, VD_MC_QVINA2 = 1, VD_MC_QVINAW = 2 };
typedef struct {
  int chains;            /* MC chains per ligand (reference: --thread) */
  int search_depth;      /* MC steps per chain; 0 = heuristic of V11/main/main.cpp:354-380 */
  int bfgs_steps;        /* local-search iterations; 0 = heuristic of V11/main/main.cpp:383-391 */
  int local_search;      /* VD_LS_RILC_BFGS (reference --rilc_bfgs 1) or VD_LS_BFGS */
  int policy;            /* VD_MC_VINA, VD_MC_QVINA2 or VD_MC_QVINAW (the last two with RILC-BFGS) */
  int mutation_quirk;    /* 1: mutate one of 1+T entities like the reference GPU host tables */
  int chains_per_wave;   /* 1, 2, 4; 0 = per ligand size bucket (vd_set_autotune) */
  int waves_per_simd;    /* register budget of the MC kernel: 2, 3 or 4 resident waves per SIMD; 0 = per size bucket */
  uint32_t seed;         /* chain c of ligand id draws from the Philox stream keyed (seed, id) at
                            counter c: ids count from 0 in the order of vd_add_ligand, as the
                            reference's ligands count in the order of its batch */
  float temperature;     /* 1.2 */
  float amplitude;       /* mutation amplitude 2 */
  float v_intra, v_inter;/* curl caps during search: hunt_cap (10, 10) */
  float slope;           /* out-of-box penalty slope 1e6 */
} vd_dock_params;
int vd_dock_params_default(vd_dock_params* p);
/* search_depth and bfgs_steps heuristics of V11/main/main.cpp:354-391 for a ligand. */
int vd_dock_heuristics(int n_atoms, int n_torsions, int rilc, int* search_depth, int* bfgs_steps);

typedef struct {
  double kernel_ms;            /* device time of the docking kernel(s); launches that run at the
                                  same time count once (first start to last end) */
  unsigned long long evals;    /* energy/gradient evaluations performed */
  int chains_per_wave;         /* used by the last launch (plain mode) */
  int blocks;                  /* workgroups launched */
} vd_stats;

/* One MC kernel launch of the last vd_dock / vd_dock_batch / vd_dock_batch_plain call (the
 * autotuner's trial launches excluded), in launch order.  The kernel is
 * vd_mc_kernel<lanes_per_chain, mode, waves_per_simd>; mode 0 dense BFGS, 1 RILC-BFGS,
 * 2 RILC-BFGS with the QuickVina2 history.  eval_bytes = algorithmic bytes of its evaluations:
 * per evaluation 32 B per heavy atom (eight map corners) and 16 B per intramolecular pair (one
 * interpolation quad of the pair table).  No counterpart in the reference (which prints only
 * per-ligand kernel times, V11/lib/main_procedure_cu.cpp:975-994); it serves the roofline measurement. */
typedef struct {
  int lanes_per_chain, mode, waves_per_simd;
  int blocks, items, ligands, persistent;
  double kernel_ms;
  unsigned long long evals;
  double eval_bytes;
  double start_ms;   /* start of the kernel after the first enqueue of the call (launches of one
                        batch overlap when their classes run at the same time) */
} vd_launch;
/* Copies up to max records; returns the number of launches (>= 0) or an error. */
int vd_last_launches(vd_ctx* ctx, int max, vd_launch* out);
/* on = 1 (default): when a vd_dock_batch call has two lanes-per-chain classes, their launches run
 * at the same time, each on its own stream: the narrower class takes half of its resident blocks
 * per CU and the wider class fills the rest, so that a class whose waves mostly wait shares every
 * CU with one that keeps the vector units busy, and the wider class's waiting workgroups take
 * over the slots the other gives up when its queue runs dry.  Batches of one class, or of three
 * or more, run one class after the other.  on = 0: always one after the other.  No counterpart
 * in the reference (one kernel_monte launch per ligand). */
int vd_set_class_overlap(vd_ctx* ctx, int on);
/* on = 1: vd_dock_batch runs as a captured hipGraph.  The first call for a set of ligands and
 * parameters captures its stages -- uploads of the ligand records and launch tables (from pinned
 * staging the graph owns), the work-queue and evaluation-count clears, the QuickVina-W buffer
 * clear, the MC launch of every class (concurrent branches when the classes overlap) and the
 * read-back of the evaluation counts -- and later calls with the same ligand ids (same order) and
 * the same parameters replay it without packing or re-uploading anything on the host; the seed may
 * differ between calls (the graph's kernel nodes are updated).  A replay gives the results of the
 * eager path bit for bit.  Up to four captured batches are kept per context (least recently used
 * dropped); a batch whose buffers another call has grown since is captured again, and batches of
 * more than four classes run eagerly.  on = 0 (default) drops the captured graphs.  vd_last_launches
 * then gives each launch the whole graph's span.  No counterpart in the reference (one
 * kernel_monte launch per ligand, main_procedure_cu.cpp:527-956). */
int vd_set_graphs(vd_ctx* ctx, int on);
/* on = 0 (default): with chains_per_wave or waves_per_simd left at 0, each ligand's
 * (lanes per chain, waves per SIMD) pair comes from a fixed table per size bucket (atoms x
 * torsions), so that under the Vina and QuickVina2 policies the same inputs and seed give the same
 * results in every process (QuickVina-W's shared buffer is filled in the order chains happen to
 * run, so its results depend on scheduling).  on = 1:
 * the pair is timed per bucket on first use (trial launches of a representative ligand) and the
 * fastest kept for the context; timings that flip between processes then change a run's FP32
 * rounding and with it its trajectories.  Turning it off forgets the timed pairs.  No counterpart
 * in the reference (one fixed launch shape). */
int vd_set_autotune(vd_ctx* ctx, int on);

/* Plain per-ligand launch (replaces kernel_monte, V11/inc/cuda/kernel2.cuh:25-38, defined at kernel2.cu:567).
 * init_conf: each chain's starting conformation [chains][7+T] (position, quaternion w,x,y,z,
 * torsions; host pointer), the role of kernel_monte's ric (drawn by the reference host at
 * V11/lib/main_procedure_cu.cpp:584-635), or NULL to draw them from the chain's random stream.
 * Per chain results: e[chains], conf[chains][7+T], heavy[chains][n_heavy][3] (host pointers,
 * any may be NULL).  The box is the one of the resident maps. */
int vd_dock(vd_ctx* ctx, int id, const vd_dock_params* p, const float* init_conf, float* e, float* conf,
            float* heavy, vd_stats* st);

/* One local search from each of n given conformations conf_in[n][7+T] of ligand id, with no
 * Monte Carlo steps: the search the MC kernel runs after every mutation (p->local_search:
 * RILC-BFGS with the Lewis-Overton line search, V11/inc/cuda/quasi_newton.cuh:1710-1806,
 * 2172-2389, or dense BFGS, quasi_newton.cuh:2391-2499 / V11/lib/bfgs.h:30-140), p->bfgs_steps
 * iterations (0 = heuristic), p's curl caps and slope; the better of the start and the search's
 * end point, as the reference keeps it.  Outputs conf_out[n][7+T] and e[n] (host pointers, either
 * may be NULL).  Overwrites the results of the last batch. */
int vd_local_search(vd_ctx* ctx, int id, int n, const float* conf_in, const vd_dock_params* p, float* conf_out,
                    float* e, vd_stats* st);

/* vd_dock for callers that cannot allocate (code under HIP graph capture, tight serving loops):
 * no counterpart in the reference.  All pointers are device pointers the caller owns: d_init_conf
 * [chains][7+T] or NULL (drawn), d_e [chains], d_conf [chains][7+T], d_heavy [chains][n_heavy][3].
 * Every other device buffer of the call (the ligand record, launch bookkeeping, the QuickVina2 /
 * QuickVina-W histories) lives in a 4 MiB array per device (VD_NOALLOC_SCRATCH_BYTES in the
 * library), which the device holds until the process ends: the call allocates no device memory,
 * and returns VD_ERR_CAPACITY when a ligand and chain count need more than the array.  The
 * launch shape is the fixed per-bucket table (no timing trials).  All calls on one device share
 * its array: calls that overlap on one device, from any context or thread, race. */
int vd_dock_noalloc(vd_ctx* ctx, int id, const vd_dock_params* p, const float* d_init_conf, float* d_e,
                    float* d_conf, float* d_heavy, vd_stats* st);

/* Many ligands of uneven size in one persistent launch fed by a device work queue.
 * Results stay on the device; fetch them per ligand with vd_batch_results. */
int vd_dock_batch(vd_ctx* ctx, int n, const int* ids, const vd_dock_params* p, vd_stats* st);
/* Same work as vd_dock_batch issued as one plain launch per ligand (the per-ligand baseline). */
int vd_dock_batch_plain(vd_ctx* ctx, int n, const int* ids, const vd_dock_params* p, vd_stats* st);
/* Results of ligand k (0..n-1) of the last batch (the per-chain output of kernel_monte that
 * cl_to_vina reads, V11/lib/main_procedure_cu.cpp:84-124): the first min(max_chains, chains) chains
 * into e[], conf[][7+T], heavy[][n_heavy][3] (host pointers, any may be NULL).  Returns the number
 * of chains the ligand was docked with (>= 0; max_chains = 0 asks for it) or an error. */
int vd_batch_results(vd_ctx* ctx, int k, int max_chains, float* e, float* conf, float* heavy);
/* Best energy per ligand of the last batch: copies at most max values into best_e (NULL when
 * max = 0) and returns the batch's ligand count (>= 0) or an error. */
int vd_batch_best(vd_ctx* ctx, int max, float* best_e);

/* Output pose list: add_to_output_container over chains in order (min_rmsd, max_size;
 * V11/lib/coords.cpp:47-62).  Writes kept chain indices sorted by energy; returns the count. */
int vd_cluster(int n, int n_heavy, const float* e, const float* heavy, double min_rmsd, int max_size, int* kept);

/* ---- After the search: refinement, final scores and output.  The reference runs these on the
 * CPU after kernel_monte (V11/main/main.cpp:139-178, 423-501); here the refinement runs on the
 * device, one pose per wavefront, many ligands per launch. ---- */
typedef struct {
  int max_steps;        /* BFGS steps per slope stage; 0 = the reference's bfgs_steps heuristic
                           (V11/main/main.cpp:383-391): (30 + n_atoms) / 3 after a RILC-BFGS search,
                           (25 + n_atoms) / 3 otherwise */
  int local_search;     /* the search that produced the poses (selects that heuristic) */
  int num_modes;        /* modes kept per ligand (--num_modes, 9) */
  float energy_range;   /* kcal/mol above the best mode (--energy_range, 3) */
  float min_rmsd;       /* remove_redundant distance (1 A, main.cpp:466) */
  float v;              /* curl cap of the refinement and the final score (authentic_v, 1000) */
  float slope;          /* out-of-box slope of the final score (the non_cache slope, 1e6) */
} vd_post_params;
int vd_post_params_default(vd_post_params* p);

/* refine_structure (V11/main/main.cpp:139-157) of n poses of ligand id: dense BFGS (V11/lib/bfgs.h)
 * on the receptor atoms themselves (non_cache::eval_deriv, V11/lib/non_cache.cpp:82-128) plus the
 * intramolecular pairs, both capped at v, the out-of-box slope raised 1e2, 1e4, .. 1e10 until every
 * heavy atom lies within the box (non_cache::within, margin 1e-4).  conf_in/conf_out [n][7+T] (may
 * alias), coords_out [n][n_atoms][3] all atoms, the ligand's internal order (NULL allowed).  The
 * box is the one of the resident maps. */
typedef struct {
  float e;          /* refined energy (output_type::e); FLT_MAX when the pose never got inside the box */
  float e_inter;    /* non_cache::eval at the refined pose: fast table, cap v, slope p->slope
                       (V11/lib/non_cache.cpp:38-56) */
  float e_intra;    /* model::eval_intramolecular at the refined pose, cap v (V11/lib/model.cpp:1375-1417) */
  int inside;       /* 1 when the heavy atoms are within the box */
} vd_refined;
int vd_refine(vd_ctx* ctx, int id, int n, const float* conf_in, const vd_post_params* p,
              float* conf_out, float* coords_out, vd_refined* out);

/* One mode of the output (output_type after main.cpp:425-470). */
typedef struct {
  float e;          /* final score: conf_independent(e_inter + e_intra - e_intra of the best mode),
                       num_tors_div (V11/lib/model.cpp:1428-1431, everything.cpp:296-303) */
  float rmsd_lb;    /* rmsd lower bound to the best mode (model.cpp:1453-1508) */
  float rmsd_ub;    /* rmsd upper bound to the best mode (model.cpp:1510-1521) */
  int pose;         /* index of the pose it came from */
} vd_mode;
/* The selection of V11/main/main.cpp:425-470 over n refined poses of ligand id (host only, no
 * device work): sort by refined energy, eval_adjusted against the best pose's intramolecular
 * energy, sort, remove_redundant (min_rmsd), keep up to num_modes poses within energy_range of
 * the best; coords [n][n_atoms][3] as vd_refine returns them.  Writes at most min(n,
 * p->num_modes) entries into modes; returns the number of modes. */
int vd_select_modes(vd_ctx* ctx, int id, int n, const vd_refined* r, const float* coords,
                    const vd_post_params* p, vd_mode* modes);

/* Both steps for many ligands, every pose in one launch: ligand ids[k] has poses[k] poses, their
 * conformations concatenated in conf (sum poses[k] x (7+T_k) floats).  vd_dock_batch's output
 * feeds it through vd_postprocess_docked.  Results per ligand: vd_post_results. */
int vd_postprocess_batch(vd_ctx* ctx, int n, const int* ids, const int* poses, const float* conf,
                         const vd_post_params* p, vd_stats* st);
/* The same for every ligand of the last vd_dock_batch / vd_dock_batch_plain: each ligand's chains
 * are first reduced to its pose list as the reference's host does (add_to_output_container with
 * min_rmsd 1 and max_poses, V11/lib/main_procedure_cu.cpp:908-926; max_poses 20 = num_saved_mins). */
int vd_postprocess_docked(vd_ctx* ctx, int max_poses, const vd_post_params* p, vd_stats* st);
/* Modes of ligand k of the last post-processing call: the first min(max_modes, count) modes into
 * modes[], their conformations conf[][7+T] and all-atom coordinates coords[][n_atoms][3] (NULL
 * allowed).  Returns the number of modes the ligand has (>= 0; max_modes = 0 asks for it) or an
 * error. */
int vd_post_results(vd_ctx* ctx, int k, int max_modes, vd_mode* modes, float* conf, float* coords);

/* PDBQT text of n modes of ligand id (write_all_output, V11/main/main.cpp:89-102; vina_remark,
 * main.cpp:158-170; model::write_context, V11/lib/model.cpp:1063-1110): per mode "MODEL i", the
 * "REMARK VINA RESULT:" line with score, rmsd lb, rmsd ub, then the ligand's input lines with
 * each atom line's coordinates (columns 31-54, %8.3f) replaced, then "ENDMDL".
 * coords [n][n_atoms][3].  Writes at most cap bytes including the terminating NUL and returns the
 * full length without it (so cap = 0 asks for the size), or an error. */
int vd_write_pdbqt(vd_ctx* ctx, int id, int n, const vd_mode* modes, const float* coords, char* out, int cap);

/* Device memory on the context's GPU (for callers that keep inputs resident in HBM).
 * vd_copy direction: 0 host->device, 1 device->host, 2 device->device; synchronous. */
int vd_device_alloc(vd_ctx* ctx, size_t bytes, void** ptr);
int vd_device_free(vd_ctx* ctx, void* ptr);
int vd_copy(vd_ctx* ctx, void* dst, const void* src, size_t bytes, int direction);
/* n random conformations of ligand id written to device memory d_conf[n][7+T]: translation
 * uniform in the current box, uniform random orientation, torsions uniform in [-pi, pi). */
int vd_random_confs(vd_ctx* ctx, int id, long long n, uint32_t seed, float* d_conf);

/* Synthetic inputs of the benchmark recipes: Poisson-disk receptor with a central cavity,
 * random acyclic ligand with n_rot rotatable bonds (PDBQT text; returns its length, or the
 * needed size when cap is too small). */
int vd_synth_receptor(int n_atoms, double cube, double cavity_radius, uint64_t seed, float* xyz, int* xs);
int vd_synth_ligand(int n_heavy, int n_rot, uint64_t seed, char* out, int cap);

#ifdef __cplusplus
}
#endif
#endif
