// Adjoint (vertex-gradient) kernels of the quadrature-mode and approximate-mode fills.
#pragma once

#include "common.hpp"
#include "mesh_dev.hpp"

namespace frico {

// Extra per-mesh data of the adjoint:
//   tri_v[3T]       vertex indices of each triangle
//   qbeta[3q]       barycentric weights of regular-rule point p on the triangle's vertices
//   qsub_beta[9q]   the same for the 3q sub-triangle points (centroid expanded to 1/3 each)
struct AdjMesh {
    const int* tri_v;
    const double* qbeta;
    const double* qsub_beta;
};

// gvert[6*nv] (complex x,y,z per vertex) = sum_ij W_j I_i dZ(j,i)/dx at fixed edge lengths.
// aw, bw: 3T complex workspaces.
int launch_adjoint_contract(const DevMesh& m, const AdjMesh& am, double k, double inv_ksq, const int2* near_pairs,
                            int64_t nnear, const c128* I, const c128* W, c128* aw, c128* bw, double* gvert, int64_t nv,
                            hipStream_t st);
// The same contraction for the approximate-mode Z (compute_matrix_approx, src/emw_solver.cpp:144-220):
// cent_rule (T double4, written here) and third ({1/3, 1/3, 1/3} on the device) are workspaces of
// its one-point rule.
int launch_adjoint_contract_approx(const DevMesh& m, const AdjMesh& am, double k, double inv_ksq, const c128* I,
                                   const c128* W, c128* aw, c128* bw, double4* cent_rule, const double* third,
                                   double* gvert, int64_t nv, hipStream_t st);

} // namespace frico
