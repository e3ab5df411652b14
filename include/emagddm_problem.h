/*
 * emagddm_problem.h -- C ABI of the on-disk problem format: one curl-curl FETI-DP
 * problem (tet mesh, materials, tet -> subdomain partition, PEC edges and the source
 * sampled at the 4-point rule) in a versioned little-endian binary file.
 *
 * The reference has the reader interface but no reader: fe_em::MeshReader::read
 * (include/fe_em/mesh_reader.hpp:32-56) is pure virtual with no implementation.  These
 * entry points are what a MeshReader for this format calls
 * (include/reference/numeric/fetidp_solver.h: fe_em::FETIDPProblemReader).
 *
 * File layout, version 1 (every field little-endian, every array 8-byte aligned):
 *   char     magic[8]      "EMAGFDP" followed by a zero byte
 *   uint32   version       1
 *   uint32   flags         bit 0: jq present
 *   int64    n_nodes, n_tets, n_pec
 *   int32    n_mat, n_sub
 *   double   xyz[3 n_nodes]
 *   int32    tet2node[4 n_tets]       0-based node indices, Nedelec TET4 local order
 *   int32    tet_material[n_tets]     in [0, n_mat)
 *   int32    tet_subdomain[n_tets]    in [0, n_sub)
 *   double   nu[n_mat], kappa[n_mat]  nu = 1/mu, kappa in the sigma role
 *   int32    pec_edge_nodes[2 n_pec]  DIRICHLET VECTOR_EDGE_ONLY edge node pairs
 *   double   jq[12 n_tets]            only when flags & 1 (J at the 4 quadrature points)
 *   uint64   checksum                 of every byte after the magic up to here
 * int32 arrays of odd length are followed by 4 zero bytes.
 */
#ifndef EMAGDDM_PROBLEM_H
#define EMAGDDM_PROBLEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMAG_PROBLEM_VERSION 1

/* A problem in host memory.  emag_problem_read allocates every array (malloc); free
 * them with emag_problem_free.  For emag_problem_write the caller owns the arrays and
 * jq may be NULL. */
typedef struct emag_problem {
  int64_t n_nodes, n_tets, n_pec;
  int32_t n_mat, n_sub;
  double* xyz;
  int32_t* tet2node;
  int32_t* tet_material;
  int32_t* tet_subdomain;
  double* nu;
  double* kappa;
  int32_t* pec_edge_nodes;
  double* jq;
} emag_problem;

/* Status codes as fetidp_status: 0 success, 4 invalid input (bad file, failed
 * validation, I/O error).  Validation on read: magic, version, sizes against the file
 * length, checksum, node / material / subdomain indices in range, every subdomain in
 * [0, n_sub) non-empty. */
int32_t emag_problem_write(const char* path, const emag_problem* p);
int32_t emag_problem_read(const char* path, emag_problem* p);
void emag_problem_free(emag_problem* p);
/* Message of the last failed read/write on this thread ("" after a success). */
const char* emag_problem_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* EMAGDDM_PROBLEM_H */
