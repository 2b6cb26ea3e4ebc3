/*
 * oracle.h -- C interface of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates, in plain double-precision C++, the
 * reference algorithms of the docking hot path (AutoDock Vina 1.1.2 as forked by
 * Vina-CUDA v1.1).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker; the product library (libvdock) never links
 * against it.  Every function cites the reference file:line it follows (paths relative
 * to the reference root, V11 = Vina-CUDA_v1.1/).
 *
 * All handles are opaque; every call returns 0 on success or a negative code and
 * leaves a message retrievable through orc_last_error().
 */
#ifndef VDOCK_ORACLE_H
#define VDOCK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* orc_last_error(void);

/* ---- Philox4x32-10 counter-based generator (the random stream shared with the GPU) ---- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* switch the intermolecular / intramolecular terms of orc_eval_deriv (both on by default) */
void orc_set_terms(int use_inter, int use_intra);

/* ---- scoring terms and look-up table (V11/lib/everything.cpp, precalculate.h) ---- */
double orc_term_eval(int t1, int t2, double r, const double w[5]);
/* LUT handle: weights w[0..4] = gauss1, gauss2, repulsion, hydrophobic, hbond */
void* orc_lut_create(const double w[5]);
void  orc_lut_destroy(void* lut);
int   orc_lut_size(void);                 /* bins per pair: 2051 */
/* copies smooth.first, smooth.second and fast for type pair tp (153 pairs) */
int   orc_lut_get(void* lut, int tp, double* smooth_e, double* smooth_dor, double* fast);
double orc_lut_eval_fast(void* lut, int tp, double r2);
int   orc_lut_eval_deriv(void* lut, int tp, double r2, double* e, double* dor);

/* ---- receptor parse + typing (V11/lib/parse_pdbqt.cpp:258-284, model.cpp:644-797) ---- */
/* returns n atoms kept (all ATOM/HETATM lines); xyz[3n], xs[n] (17 = untyped), ad[n] */
int orc_receptor_parse(const char* text, int max_atoms, double* xyz, int* xs, int* ad);

/* ---- grid maps (V11/lib/cache.cpp:148-208, grid.cpp:29-167, szv_grid.cpp) ---- */
/* grid handle from box corners and interval counts n[3] (points per axis = n+1) */
void* orc_grid_create(void* lut, const double begin[3], const double end[3], const int n[3]);
void  orc_grid_destroy(void* grid);
int   orc_grid_set_receptor(void* grid, int n, const double* xyz, const int* xs);
/* use_cells=1: szv_grid cell lists (the reference CPU path); 0: brute force over atoms */
int   orc_grid_populate(void* grid, int use_cells);
/* populate, plus per map value the sum of |fast[i+-1]-fast[i]| over atoms whose 32 r^2 lies
 * within eps bins of a bin edge (the change an FP32 rounding of r^2 can cause) */
int   orc_grid_populate_slack(void* grid, int use_cells, double eps, double* slack_out);
/* maps out[17][nz+1][ny+1][nx+1], x fastest (array3d.h:78) */
int   orc_grid_get_maps(void* grid, double* out);
int   orc_grid_set_maps(void* grid, const double* maps);
/* average number of cell-list candidates per cell, and the candidate list of the cell
 * that holds point p (for culling identity checks) */
int   orc_grid_cell_candidates(void* grid, const double p[3], int* out, int max_out);
double orc_grid_eval(void* grid, int t, const double xyz[3], double slope, double v, double deriv[3]);

/* ---- ligand model (parse_pdbqt.cpp:286-652, model.cpp:644-896) ---- */
void* orc_ligand_parse(const char* text);
void  orc_ligand_destroy(void* lig);
/* sizes[0]=atoms, [1]=nodes (torsions+1), [2]=pairs, [3]=heavy atoms, [4]=TORSDOF */
int   orc_ligand_sizes(void* lig, int sizes[5]);
/* per atom: xs, el, ad, node, local xyz, input xyz */
int   orc_ligand_atoms(void* lig, int* xs, int* el, int* ad, int* node, double* local, double* coords0);
/* per node: parent, atom begin, atom end, rel_origin, rel_axis, origin0 */
int   orc_ligand_nodes(void* lig, int* parent, int* begin, int* end, double* rel_origin, double* rel_axis, double* origin0);
/* pairs: tp, a, b */
int   orc_ligand_pairs(void* lig, int* tp, int* a, int* b);
/* covalent bonds (i<j) as assigned by model::assign_bonds; returns count */
int   orc_ligand_bonds(void* lig, int* i, int* j, int max_out);

/* ---- kinematics + energy/gradient (tree.h, model.cpp:1148-1334) ----
 * conf layout: pos[3], quat[4] (w,x,y,z), torsions[T]; gradient: force[3], torque[3], torsions[T] */
int    orc_set_conf(void* lig, const double* conf, double* coords_out);
double orc_eval_deriv(void* grid, void* lig, const double* conf, double v_intra, double v_inter,
                      double slope, double* grad, double* coords_out, double* e_inter_out);
/* conformation increment x <- x (+) factor*d (conf.h:118-124, 60-66; quaternion.cpp:235-240) */
int    orc_conf_increment(void* lig, double* conf, const double* d, double factor);

/* ---- local search (quasi_newton.cuh:1710-1806, 2172-2389; bfgs.h:95-140) ---- */
/* mode 1 = RILC-BFGS, 0 = dense BFGS; conf updated in place, returns energy, *n_evals */
double orc_local_search(void* grid, void* lig, double* conf, int max_steps, int mode,
                        double v_intra, double v_inter, double slope, int* n_evals);
/* the same optimisers on closed-form test functions over R^n (kind 0 = quadratic with
 * diagonal curvature coef[0..n), 1 = Rosenbrock); x updated in place, returns f */
double orc_opt_test(int kind, int n, double* x, const double* coef, int max_steps, int mode, int* n_evals);
/* the same, recording the first max_pts evaluated points pts[max_pts][n] and their values fs[max_pts] */
double orc_opt_trace(int kind, int n, double* x, const double* coef, int max_steps, int mode, int max_pts,
                     double* pts, double* fs, int* n_evals);

/* ---- QuickVina2 visited history (QV2/inc/cuda/visited.cuh:7-226, intended semantics) ----
 * Feeds m points of a ligand with T torsions into one fresh history in order: x[m][7+T]
 * conformations, f[m] energies, g[m][6+T] derivatives.  interesting[k] = the gate's answer for
 * point k; mode 1 adds a point when it is interesting (the gate), mode 2 adds every point. */
/* QuickVina-W state (per-chain individual buffers, one global buffer of 2^levels cells per axis
 * over the box) driven by a sequence of operations: op 0 gate query (out = 1 interesting),
 * 1 add to chain's individual buffer, 2 add to the global buffer (out = cell), 3 G-check alone
 * (out = -1 interesting, else entries examined). */
int orc_qvw_feed(int T, int m, const double* x, const double* f, const double* g, const int* chain, const int* op,
                 const double box_lo[3], const double box_hi[3], int levels, int* out);
int orc_qvw_levels(long long chains, long long depth);
int orc_visited_feed(int T, int m, const double* x, const double* f, const double* g, int mode, int* interesting);

/* ---- Monte Carlo (kernel2.cu:219-426 schedule; mutate.cpp:62-127; monte_carlo.cpp:35-39) ---- */
typedef struct {
  int    chains, search_depth, bfgs_steps, mode;   /* mode: 1 RILC-BFGS, 0 dense BFGS */
  int    mutation_quirk;                           /* 1: mutate one of T+1 entities as the GPU reference does */
  uint32_t seed, stream;
  double box_lo[3], box_hi[3];
  double v_intra, v_inter, slope, temperature, amplitude;
  int    threads;
  int    policy;                                   /* 0 Vina, 1 QuickVina2 visited-history gate, 2 QuickVina-W */
} orc_mc_params;
/* per chain: e[c], conf[c*(7+T)], heavy coords[c*3*H]; returns total evals through *evals */
int orc_mc_run(void* grid, void* lig, const orc_mc_params* p, double* e, double* conf,
               double* heavy_coords, long long* evals);
/* the same from given starting conformations starts[chains][7+T] (quaternions normalised) */
int orc_mc_run_from(void* grid, void* lig, const orc_mc_params* p, const double* starts, double* e, double* conf,
                    double* heavy_coords, long long* evals);
/* One chain of orc_mc_run (policy 0 or 1), run alone, recording per MC step s < search_depth: the
 * entity mutated which[s] (0 position, 1 orientation, 2+k torsion k), the Metropolis draw u[s], the
 * current energy before the step e_prev[s], the candidate's energy after its local search
 * e_cand[s], the decision accept[s], the gyration radius the orientation step used rg[s], and the
 * conformations before[s][7+T] (current) and mutated[s][7+T] (candidate before its search). */
int orc_mc_trace(void* grid, void* lig, const orc_mc_params* p, int chain, int* which, double* u, double* e_prev,
                 double* e_cand, int* accept, double* rg, double* before, double* mutated);
/* initial conformation of chain c from the shared random stream */
int orc_initial_conf(void* lig, const orc_mc_params* p, int chain, double* conf);

/* ---- result clustering (coords.cpp:24-62) ----
 * feeds n results in order into an output container (min_rmsd, max_size), returns kept
 * count and their indices sorted by energy */
/* after the search (V11/main/main.cpp:139-157, 423-470): non_cache terms on the receptor atoms,
 * intramolecular energy with fast table values, num_tors, conf_independent, rmsd bounds,
 * refine_structure and the whole per-ligand post-processing. Coordinates are all atoms, xyz. */
int    orc_ligand_serials(void* lig, int* serial);
double orc_num_tors(void* lig);
double orc_conf_independent(void* lig, double e, double w_rot);
double orc_noncache_eval(void* grid, void* lig, const double* xyz, double v, double slope, double* minus_forces);
double orc_intra_fast(void* grid, void* lig, const double* xyz, double v);
int    orc_within(void* grid, void* lig, const double* xyz);
int    orc_rmsd_bounds(void* lig, const double* x, const double* y, double* lb, double* ub);
double orc_noncache_eval_deriv(void* grid, void* lig, const double* conf, double v, double slope, double* grad);
double orc_refine(void* grid, void* lig, double* conf, int max_steps, double v, double* coords_out);
int    orc_postprocess(void* grid, void* lig, int n, const double* confs, int max_steps, double w_rot, int num_modes,
                       double energy_range, double min_rmsd, double slope, double* conf_out, double* e_out,
                       double* lb_out, double* ub_out, double* refined_out);

int orc_cluster(int n, int heavy, const double* e, const double* heavy_coords, double min_rmsd,
                int max_size, int* kept_idx);

#ifdef __cplusplus
}
#endif
#endif
