/*
 * emagddm_fetidp.h -- C ABI of the MI355X FETI-DP solver for lowest-order
 * Nedelec (edge) curl-curl problems  curl(nu curl A) + kappa A = J  on tets.
 *
 * Plain pointers and sizes only.  Unless a function says "device pointer",
 * every array argument is host memory; the handle owns all device memory.
 * Calls on one handle must be serialized (as EMLinearSolverBase: "set_matrix()
 * and solve() need external synchronization", include/numeric/
 * em_linear_solver.h:76-90).
 *
 * Status codes map 1:1 onto numeric::SolverStatus
 * (include/numeric/em_linear_solver.h:26-32).
 */
#ifndef EMAGDDM_FETIDP_H
#define EMAGDDM_FETIDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fetidp_status {
  FETIDP_SUCCESS = 0,          /* SolverStatus::SUCCESS */
  FETIDP_NUMERICAL_ERROR = 1,  /* SolverStatus::NUMERICAL_ERROR (non-SPD block, breakdown) */
  FETIDP_DIVERGED = 2,         /* SolverStatus::DIVERGED */
  FETIDP_MAX_ITER_REACHED = 3, /* SolverStatus::MAX_ITER_REACHED */
  FETIDP_INVALID_INPUT = 4     /* SolverStatus::INVALID_INPUT */
};

enum fetidp_precision {
  FETIDP_FP64 = 0,        /* FP64 factors and operators on MFMA */
  FETIDP_FP32_REFINE = 1  /* FP32 factors/operators on MFMA inside FP64 refinement
                             (flexible CG on the assembled system) */
};

/* Solver configuration.  Defaults (fetidp_config_default) mirror
 * IterativeSolverConfig{tolerance = 1e-8, max_iterations = 1000}
 * (include/numeric/iterative_solvers.h:59-75). */
typedef struct fetidp_config {
  int32_t precision;        /* enum fetidp_precision */
  int32_t max_iterations;   /* PCPG iteration cap per solve */
  double tolerance;         /* target ||f - K u|| / ||f|| on the assembled free-dof system */
  double dual_tolerance;    /* first PCPG stop ||r||/||d||; tightened x0.1 until tolerance holds */
  int32_t max_refinements;  /* FP32 mode: outer FP64 flexible-CG steps */
  int32_t verbose;          /* bit 0 (value 1): report an FP32-to-FP64 factor fallback and
                               its reason on stderr; bit 1 (value 2): record
                               per-kernel-class event timings */
  double inner_tolerance;   /* FP32 mode: PCPG stop ||r||/||d|| of each refinement step */
  int32_t graphs;           /* 1: replay each PCPG iteration as one captured HIP graph
                               (ignored while event profiling is on) */
  int32_t precond_precision; /* Dirichlet blocks: 0 auto (FP64 when max nu / (min kappa
                               h_min^2) > 1e6, else FP32), 1 FP32, 2 FP64.  With deluxe
                               scaling the blocks are FP64 and this selects the face
                               operators D_F = A_F^-1 B_F: FP32 (0, 1; formed in FP64) or
                               FP64 (2) */
  int32_t scaling;          /* interface scaling of the preconditioner: 0 auto (deluxe when
                               kappa/nu varies by more than 1e4 between materials or when
                               neighbouring subdomains' bounding boxes overlap by more
                               than one mean edge length, i.e. not a box partition; else
                               stiffness), 1 stiffness rho_t/(rho_s+rho_t), 2 deluxe
                               (S_s,FF + S_t,FF)^-1 S_t,FF per face (FP64) */
  int32_t side_stream;      /* where the diagonal-tile factorizations of the column steps
                               run: 2 at the front of the column-step launch, dispatched
                               first at raised wave priority (default); 1 on a second
                               stream beside it; 0 as a launch of their own, one stream */
  int32_t tile_waves;       /* throughput tile kernels: 4 = capped at 128 registers for four
                               waves per SIMD (default), 0 = the compiler's own allocation */
  int32_t coarse_solve;     /* coarse problem S~ (order n_c): 0 auto (grouped when n_c > 12288,
                               else dense), 1 dense: the factorization forms -S~^-1 and each
                               iteration applies it as one symmetric product (n_c^3 flops at
                               setup, n_c^2/2 words per iteration), 2 grouped: S~ = L L^T
                               (envelope) and each iteration sweeps forward and backward over
                               groups of tile columns with the explicit inverses of the
                               groups' diagonal blocks.  Read at fetidp_setup; changing it
                               afterwards rebuilds the coarse plan and drops the factor. */
  int32_t coarse_group;     /* grouped coarse solve: tile columns (64 rows) per group, 1..64;
                               0 = 32 */
  int32_t reserved;
} fetidp_config;

typedef struct fetidp_stats {
  int32_t status;
  int32_t iterations;       /* total PCPG iterations (all refinement steps) */
  int32_t refinements;      /* outer refinement steps taken */
  int32_t fell_back;        /* FP32_REFINE only, 1: the answer comes from an FP64 factor
                               (the accuracy estimate nu/(kappa h^2) 2^-24 exceeded 5e-3,
                               the FP32 factor was not positive definite, or the outer
                               refinement stalled) */
  double residual_norm;     /* final ||f - K u||_2 / ||f||_2 (FP64, assembled system) */
  double dual_residual;     /* final PCPG ||r|| / ||d|| */
  double initial_dual_residual; /* ||d - F lambda0|| / ||d|| */
  double factor_ms;         /* numeric setup: integration, assembly, factorization, coarse */
  double solve_ms;          /* RHS condensation, PCPG, recovery, residual check */
  double symbolic_ms;       /* host topology + plans + upload (fetidp_setup) */
  double guess_ms;          /* learned lambda0: features + network (0 without a model) */
  int32_t learned_guess;    /* 1: PCPG started from the attached model's lambda0 */
  int32_t precond_fp64;     /* 1: the Dirichlet blocks were kept in FP64 */
  int32_t deluxe;           /* 1: deluxe interface scaling was used */
  int32_t reserved;
} fetidp_stats;

typedef struct fetidp_solver* fetidp_handle;

void fetidp_config_default(fetidp_config* cfg);

/* Replaces EMSolverFactory::create_solver for the new FETIDP solver type
 * (include/numeric/em_solver_factory.h:213-279). NULL on failure. */
fetidp_handle fetidp_create(const fetidp_config* cfg);

/* Change the configuration of an existing handle (as IterativeSolverConfig is
 * set on a solver, include/numeric/iterative_solvers.h:59-75).  A change of
 * precision, precond_precision or scaling drops the current factor (call
 * fetidp_factor again). */
int32_t fetidp_set_config(fetidp_handle h, const fetidp_config* cfg);
int32_t fetidp_get_config(fetidp_handle h, fetidp_config* cfg);

/* Symbolic setup.  Mesh, materials, partition and PEC edges; replaces the
 * EMMeshData -> GlobalEdgeIDGenerator -> EMDOFManager chain
 * (src/fe_em/global_edge_id_generator.cpp:30-110, src/fe_em/em_dof_manager.cpp:29-65)
 * and FETIDPContext::initialize (src/fetidp/fetidp_context.cpp:19-35).
 *   xyz[3*n_nodes], tet2node[4*n_tets] (Nedelec TET4 local node order),
 *   tet_material[n_tets] in [0,n_mat), nu[n_mat] = 1/mu, kappa[n_mat] (sigma role),
 *   tet_subdomain[n_tets] >= 0, pec_edge_nodes[2*n_pec]: DIRICHLET VECTOR_EDGE_ONLY
 *   edge node pairs (EMBoundaryMarker, include/fe_em/em_mesh_data.hpp:166-176). */
int32_t fetidp_setup(fetidp_handle h, int64_t n_nodes, const double* xyz, int64_t n_tets,
                     const int32_t* tet2node, const int32_t* tet_material, int32_t n_mat,
                     const double* nu, const double* kappa, const int32_t* tet_subdomain,
                     int64_t n_pec, const int32_t* pec_edge_nodes);

/* Replace material values (same n_mat) without redoing the symbolic setup. */
int32_t fetidp_set_materials(fetidp_handle h, int32_t n_mat, const double* nu, const double* kappa);

/* Source J sampled at the 4-point rule nodes of every tet, jq[12*n_tets]
 * (point q has barycentric weight (5+3sqrt5)/20 on local node q, (5-sqrt5)/20
 * on the others); F_e(i) = sum_q |T|/4 W_i(x_q).J_q.  Replaces the zero
 * MagneticVector3DIntegrator::computeSourceVector
 * (src/numeric/magnetic_vector_3d_integrator.cpp:343-364).  After
 * fetidp_factor a new source is assembled at once and reuses the factor
 * (factor once, solve many). */
int32_t fetidp_set_rhs_tet_J(fetidp_handle h, const double* jq);

/* Assembled free-dof right-hand side f[n_free] (replaces set_rhs by J). */
int32_t fetidp_set_rhs_free(fetidp_handle h, const double* f);

/* Numeric setup on the GPU: element integration, subdomain assembly, batched
 * partial Cholesky (MFMA), local dual operators, coarse Schur complement and
 * its factorization.  Analogue of set_matrix (include/numeric/em_linear_solver.h:110-130). */
int32_t fetidp_factor(fetidp_handle h);

/* PCPG solve; lambda0[n_lambda] may be NULL (start from 0).  x_free[n_free]
 * (may be NULL) receives the solution in EMDOFManager free-dof order.
 * Analogue of EMLinearSolverBase::solve -> SolverResult
 * (include/numeric/em_linear_solver.h:53-61,132-160). */
int32_t fetidp_solve(fetidp_handle h, const double* lambda0, double* x_free, fetidp_stats* stats);

/* Sizes: out[0..9] = n_edges, n_free, n_lambda, n_primal, n_sub, n_tets,
 * pool elements, operator elements (W+P), coarse padded order, max local order. */
int32_t fetidp_get_sizes(fetidp_handle h, int64_t* out);

/* Preconditioner choices the current configuration resolves to after setup (host only,
 * no GPU work): *deluxe = 1 when deluxe scaling will be used, *precond_fp64 = 1 when the
 * Dirichlet blocks will be stored in FP64, *box_overlap = largest overlap of neighbouring
 * subdomains' bounding boxes in mean edge lengths (the partition-shape input of the
 * automatic scaling rule).  Any pointer may be NULL.  No reference counterpart (the
 * reference has no domain decomposition preconditioner). */
int32_t fetidp_get_choices(fetidp_handle h, int32_t* deluxe, int32_t* precond_fp64,
                           double* box_overlap);

/* Topology readback: tet_free[6*n_tets] (free dof per local edge, -1 PEC, in
 * ElementGeometry order), edge_nodes[2*n_edges], lam_free[n_lambda],
 * primal_free[n_primal].  Any pointer may be NULL. */
int32_t fetidp_get_topology(fetidp_handle h, int32_t* tet_free, int32_t* edge_nodes,
                            int32_t* lam_free, int32_t* primal_free);

/* Global free-dof CSR of K + M and the assembled RHS (after fetidp_factor /
 * a set_rhs call): row_ptr[n_free+1], col[nnz], val[nnz], f[n_free]. */
int64_t fetidp_get_global_nnz(fetidp_handle h);
int32_t fetidp_get_global_system(fetidp_handle h, int32_t* row_ptr, int32_t* col, double* val,
                                 double* f);

/* Dual-space operators after fetidp_factor and a RHS: q = F p, z = M^-1 r,
 * d (dual right-hand side), last lambda.  Vectors have n_lambda entries. */
int32_t fetidp_apply_dual_operator(fetidp_handle h, const double* p, double* q);
int32_t fetidp_apply_preconditioner(fetidp_handle h, const double* r, double* z);
int32_t fetidp_get_dual_rhs(fetidp_handle h, double* d);
int32_t fetidp_get_lambda(fetidp_handle h, double* lam);
/* Per-iteration PCPG dual residual history of the last solve (up to cap). */
int32_t fetidp_get_history(fetidp_handle h, double* hist, int32_t cap);
/* Global relative residual after each recovery / refinement step of the last solve. */
int32_t fetidp_get_refinement_history(fetidp_handle h, double* hist, int32_t cap);
/* Coarse Schur complement S~ (n_primal^2, row-major) after fetidp_factor. */
int32_t fetidp_get_coarse_matrix(fetidp_handle h, double* S);

/* Event profile when cfg.verbose & 2: out[2k], out[2k+1] = milliseconds and
 * launch count of kernel class k (0 integrate, 1 pool assembly, 2..4 local tile
 * modes 0..2, 5 operator extraction, 6..8 coarse tile modes 0..2, 9 dual GEMV
 * of F, 10 rest of F, 11 preconditioner GEMV, 12 forward solve, 13 interior
 * solve, 14-15 spare), then the executed flops of local modes 0..2 and coarse modes 0..2
 * (2*64^3 per tile product: padding, the full diagonal-tile share and the TRSM as a
 * product with the stored inverse included), then the useful flops of the same six
 * classes (real rows only, triangular TRSM, lower half of the diagonal share).  Returns
 * the number of values written. */
int32_t fetidp_get_profile(fetidp_handle h, double* out, int32_t cap);
void fetidp_reset_profile(fetidp_handle h);

/* Learned initial multiplier (include/emagddm_mpnn.h).  With a model attached
 * and lambda0 == NULL, fetidp_solve starts PCPG from lambda0 = s_z * y, where
 * y is the network's output on the interface graph for the features of
 * fetidp_get_lambda_features; model == NULL detaches.  The handle does not own
 * the model, which must outlive its use.  No reference counterpart: the
 * reference always starts from zero (src/numeric/iterative_solvers.cpp:424-503). */
struct emag_mpnn;
int32_t fetidp_set_lambda0_model(fetidp_handle h, struct emag_mpnn* model);

/* Interface graph of the multipliers: nodes = dual edges (lambda order), i ~ j
 * when their mesh edges share a node and their unit tangents are not
 * orthogonal; coef = t_i . t_j.  Returns nnz (< 0 on error); arrays
 * row_ptr[n_lambda+1], col[nnz], coef[nnz] are filled when non-NULL. */
int64_t fetidp_get_dual_graph(fetidp_handle h, int32_t* row_ptr, int32_t* col, float* coef);

/* Network input features (n_lambda x 16, FP32) of the current factor and RHS,
 * computed on the GPU from d, z = M^-1 d and the stiffness scaling; scales[0..1]
 * = rms(d), rms(z) (the output scale s_z).  After fetidp_factor. */
int32_t fetidp_get_lambda_features(fetidp_handle h, float* features, double* scales);

/* Autotuner: times `repetitions` factor + solve runs of the current problem for every
 * combination of the launch choices that keep the arithmetic identical (side_stream,
 * graphs, tile_waves), checks that all return bitwise the same solution, and keeps
 * the fastest in the handle's configuration.  When the coarse solve is grouped it then
 * times, with that launch choice, the group widths 8, 16, 24, 32, 48 and 64 tile columns
 * that fit the coarse matrix (a tile shape: the width changes the order of the coarse
 * sums and so the PCPG trajectory, hence a width is kept only if its solve succeeds with
 * the residual within the tolerance) and keeps the fastest (then refactors).  timings[5k..5k+4] = {side_stream, graphs, tile_waves, coarse
 * group width (0: dense coarse inverse), ms per factor + solve} for variant k (up to cap
 * values).  Returns the number of variants, or -status on error.  Needs a RHS.
 * No reference counterpart. */
int32_t fetidp_autotune(fetidp_handle h, int32_t repetitions, double* timings, int32_t cap);

const char* fetidp_last_error(fetidp_handle h);
void fetidp_destroy(fetidp_handle h);

/* Stand-alone batched element kernel (device pointers, HIP stream or NULL):
 * K_e, M_e packed upper triangles (21 per tet, structure-of-arrays: entry k of
 * tet t at [k*n_tets + t]) and F_e (6 per tet, same layout; jq may be NULL).
 * Replaces MagneticVector3DIntegrator::computeAllMatrices
 * (src/numeric/magnetic_vector_3d_integrator.cpp:104-135) with D1-D4 corrected. */
int32_t emag_nedelec_integrate(int64_t n_tets, const double* d_xyz, const int32_t* d_tet2node,
                               const int32_t* d_tet_material, const double* d_nu,
                               const double* d_kappa, const double* d_jq, double* d_Ke,
                               double* d_Me, double* d_Fe, void* stream);

/* Self-test of the batched tile factorization on one dense SPD matrix (host
 * arrays, row-major m x m, m and p_cols multiples of 64): eliminates the first
 * p_cols columns with the same MFMA kernels and schedule shape the solver uses
 * and returns the lower triangle: L in the eliminated columns, the Schur
 * complement A22 - L21 L21^T in the trailing block.  No reference counterpart. */
int32_t emag_tile_cholesky_test(int32_t precision, int64_t m, int64_t p_cols, const double* A,
                                double* out);

/* Known-answer probe of the 16x16x4 MFMA accumulator lane layout the tile
 * kernels rely on: out[512] = [2 products][64 lanes][4 registers] holding the
 * row index p, then the column index q, of each accumulator element.  No
 * reference counterpart (device self-test). */
int32_t emag_mfma_layout_probe(int32_t precision, double* out);

#ifdef __cplusplus
}
#endif

#endif /* EMAGDDM_FETIDP_H */
