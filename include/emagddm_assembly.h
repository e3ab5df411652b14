/*
 * emagddm_assembly.h -- C ABI of the GPU global Nedelec assembly and the FP64 residual
 * operator, without any FETI-DP factorization.
 *
 * Replaces EMAssembly::assemble and its getters (include/numeric/em_assembly.hpp:236-326,
 * src/numeric/em_assembly.cpp:144-411) for TET4 edge elements: the real Nedelec
 * integrator instead of DummyEMIntegrator (:39-125) and the edge-orientation sign in the
 * scatter (:654-718); and the free-dof numbering of EMDOFManager::build
 * (src/fe_em/em_dof_manager.cpp:29-65) with O(N) radix sorts instead of the quadratic
 * renumbering (:250-281).  The CSR has int32 row_ptr / col with sorted, unique columns
 * per row (the reference's duplicates are summed, as its solvers sum them:
 * src/numeric/sparse_converter.cpp:60).
 *
 * Device pointers are HIP device memory; `stream` is a hipStream_t or NULL (the null
 * stream).  Status codes as enum fetidp_status (include/emagddm_fetidp.h).
 */
#ifndef EMAGDDM_ASSEMBLY_H
#define EMAGDDM_ASSEMBLY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emag_assembly* emag_assembly_handle;

/* Symbolic part (host): global edge ids in first-appearance order, orientation signs,
 * PEC edges -> -1, free numbering, the CSR pattern and the contribution plans, uploaded
 * to the device that is current at the first emag_assemble (creating the handle touches
 * no GPU).  tet2node[4 n_tets] (Nedelec TET4 local node order), pec_edge_nodes[2 n_pec]
 * (DIRICHLET VECTOR_EDGE_ONLY edge node pairs).  The first emag_assemble also allocates
 * the element workspace (48 doubles per tet).  NULL on failure. */
emag_assembly_handle emag_assembly_create(int64_t n_nodes, int64_t n_tets, const int32_t* tet2node,
                                          int64_t n_pec, const int32_t* pec_edge_nodes);

/* n_free and the CSR nonzero count. */
int32_t emag_assembly_sizes(emag_assembly_handle h, int64_t* n_free, int64_t* nnz);

/* Host copies: row_ptr[n_free + 1], col[nnz]; tet_free[6 n_tets] (free dof per local
 * edge in ElementGeometry order, -1 PEC: the Local2Global map) and tet_sign[6 n_tets]
 * (+-1 orientation of the reference Whitney function against the global edge).  Any
 * pointer may be NULL. */
int32_t emag_assembly_pattern(emag_assembly_handle h, int32_t* row_ptr, int32_t* col,
                              int32_t* tet_free, int8_t* tet_sign);

/* Numeric part on the GPU, on `stream` (asynchronous: results are ready when the stream
 * is).  Inputs: xyz[3 n_nodes], tet_material[n_tets], nu[n_mat] (= 1/mu), kappa[n_mat],
 * jq[12 n_tets] (J at the 4-point rule nodes, may be NULL).  Outputs (each may be NULL,
 * CSR values in the pattern's order): K[nnz] curl-curl stiffness (getStiffnessMatrix),
 * M[nnz] kappa-weighted mass (getMassMatrix), A[nnz] = K + M (the system the FETI-DP
 * solver factors), f[n_free] source vector (getSourceVector; needs jq). */
int32_t emag_assemble(emag_assembly_handle h, const double* d_xyz, const int32_t* d_tet_material,
                      const double* d_nu, const double* d_kappa, const double* d_jq, double* d_K,
                      double* d_M, double* d_A, double* d_f, void* stream);

const char* emag_assembly_last_error(emag_assembly_handle h);
void emag_assembly_destroy(emag_assembly_handle h);

/* Tree-cotree gauge for kappa = 0 (magnetostatics), after EMDOFManager::applyTreeGauge
 * (src/fe_em/em_dof_manager.cpp:448-495) and TreeGauge (include/fe_em/tree_gauge.hpp:
 * 131-293): a spanning tree of the mesh graph over the free edges, grown breadth-first from
 * all PEC nodes at once (the PEC boundary is one contracted root, so every other node gets
 * exactly one tree edge; a component without PEC starts from its lowest node).  The tree
 * edges are returned as node pairs, tree_edge_nodes[2 * n_tree] (may be NULL to count):
 * given to emag_assembly_create or fetidp_setup together with the PEC pairs they are set
 * to zero like the PEC edges, and the curl-curl system on the remaining cotree edges is
 * SPD for kappa = 0 on a simply connected domain.  Returns n_tree (at most n_nodes), or
 * -1 on bad input.  Host only. */
int64_t emag_tree_gauge(int64_t n_nodes, int64_t n_tets, const int32_t* tet2node, int64_t n_pec,
                        const int32_t* pec_edge_nodes, int32_t* tree_edge_nodes);

/* FP64 residual operator: r = f - A x over a CSR (device arrays, row_ptr[n + 1]), and
 * *rel_norm = ||r||_2 / ||f||_2 (0 when f = 0) on the host after the stream finishes.
 * d_r may be NULL when only the norm is wanted.  The library allocates the workspace (the
 * residual vector when d_r is NULL, and the 512 partial sums of the two norms) on each
 * call.  The reference computes the same norm on the host
 * (src/numeric/iterative_solvers.cpp:445-474). */
int32_t emag_csr_residual(int64_t n, const int32_t* d_row_ptr, const int32_t* d_col,
                          const double* d_val, const double* d_x, const double* d_f, double* d_r,
                          double* rel_norm, void* stream);

/* emag_csr_residual with no allocation: no reference counterpart.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the workspace is a static device array of the
 * library (8 MiB + 4 KiB on every GPU the library runs on, held until the process ends).
 * d_r == NULL holds the residual in that array, so then n may be at most 1,048,576 rows
 * (larger n: INVALID_INPUT, nothing launched); with d_r given, any n.  Same launches and
 * bitwise the same result as emag_csr_residual.  All calls on one GPU share the array:
 * calls that overlap on one GPU (other host threads, other streams) race. */
int32_t emag_csr_residual_noalloc(int64_t n, const int32_t* d_row_ptr, const int32_t* d_col,
                                  const double* d_val, const double* d_x, const double* d_f,
                                  double* d_r, double* rel_norm, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EMAGDDM_ASSEMBLY_H */
