// Device-side views of the mesh, RWG table and incident waves.
#pragma once

#include "common.hpp"

namespace frico {

// Per-triangle and per-RWG arrays in HBM.  All geometry is FP64.
//   cen[3T]      centroid (p0+p1+p2)/3                       src/geom_mesh.cpp:64-69
//   area[T]      0.5 |e0 x e1|                               src/geom_mesh.cpp:71-78
//   sqrtA[T]     sqrt(area)  (approximate-mode self term)    src/emw_solver.cpp:175-176
//   alpha[9T]    centroid - free vertex of local edge m, m=0..2; the free vertex of local edge m
//                is vertex {2,0,1}[m]                        src/rwg_basis.cpp:47-50
//   qreg[T*q]    regular Dunavant rule, {x,y,z} relative to the centroid, w = weight * Kahan area
//   qsub[T*3q]   sub-triangle rule (p0,p1,c),(p1,p2,c),(p2,p0,c), same convention
//   tri_rwg[3T]  RWG on local edge m of triangle t, or -1; tri_sgn[3T] = -1 (t is T-), +1 (T+)
//   rwg_t[2N]    (T-, T+);  rwg_le[2N] local edge index in T-, T+;  rwg_len[N] hypot(dx,dy)
//   rwg_Am/Ap[N] areas of T-/T+
struct DevMesh {
    int T, N, q;
    const double* cen;
    const double* area;
    const double* sqrtA;
    const double* alpha;
    const double4* qreg;
    const double4* qsub;
    const int* tri_rwg;
    const double* tri_sgn;
    const int* rwg_t;
    const int* rwg_le;
    const double* rwg_len;
    const double* rwg_Am;
    const double* rwg_Ap;
};

// One incident plane wave (src/sources.h:44-65): unit vector toward the source and the
// complex amplitude vector E0 = Eth theta_hat + Eph phi_hat.
struct PlaneWaveDev {
    double dir[3];
    double e0r[3];
    double e0i[3];
};

int launch_fill_quad(const DevMesh& m, double k, double inv_ksq, int prec, void* Cbuf, void* Z, int64_t ldz,
                     const int2* near_pairs, int64_t nnear, hipStream_t st, KernelTimer* kt);
int launch_fill_approx(const DevMesh& m, double k, double inv_ksq, int prec, void* Z, int64_t ldz, hipStream_t st);
int launch_rhs_planewave(const DevMesh& m, double k, double omega_mu0, const PlaneWaveDev* waves, int nang, c128* B,
                         int64_t ldb, hipStream_t st);
int launch_scatter(const DevMesh& m, const c128* I, int64_t ldi, int nrhs, c128* TS, hipStream_t st);
// E and H at npts points from the triangle currents TS of ncols current vectors; point p uses
// column cols[p] (cols null: column 0), NaN fields for a column outside [0, ncols).
int launch_fields(const DevMesh& m, double k, double omega_mu0, const c128* TS, int ncols, const double* pts,
                  const int* cols, int npts, double h_sign, c128* E, c128* H, hipStream_t st);

} // namespace frico
