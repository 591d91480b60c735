/*
 * frico_gpu.h -- C ABI of the MI355X (gfx950) EFIE Method-of-Moments library.
 *
 * Plain C: pointers, sizes and status codes; no exceptions cross it and no C++ or HIP types
 * appear in the signatures.  Complex numbers are interleaved (re, im) doubles, the layout of
 * std::complex<double>; matrices are column-major with an explicit leading dimension, like the
 * reference's Eigen matrices (src/eigen.h:36-40).  Device memory is owned by the context, except
 * the caller's device buffers of frico_fields_noalloc.
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference
 * source root).  Status convention: 0 = OK; -k = argument k invalid (LAPACK style); > 0 =
 * singular pivot at that 1-based column (zgetrf's info); FRICO_ERR_* for the rest.  The text
 * of the last error of the calling thread is returned by frico_last_error().
 */
#ifndef FRICO_GPU_H
#define FRICO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRICO_OK 0
#define FRICO_ERR_HIP (-1000)    /* a HIP runtime call failed */
#define FRICO_ERR_NOMEM (-1001)  /* device allocation failed */
#define FRICO_ERR_STATE (-1002)  /* call order (e.g. solve before factor, no mesh) */
#define FRICO_ERR_MESH (-1003)   /* connectivity error: lonely edge or edge in > 2 triangles */
#define FRICO_ERR_NOCONV (-1004) /* iterative refinement did not reach the tolerance */
#define FRICO_ERR_SYNC (-1005)   /* a device-wide barrier exceeded its wait limit */
#define FRICO_ERR_LIMIT (-1006)  /* problem exceeds a documented size limit */

/* Matrix fill modes: compute_matrix (src/emw_solver.cpp:45-142) and compute_matrix_approx
 * (src/emw_solver.cpp:144-220). */
#define FRICO_FILL_QUAD 0
#define FRICO_FILL_APPROX 1

/* Precisions.  FRICO_C128: complex128 everywhere.  FRICO_C64: complex64 storage. */
#define FRICO_C128 0
#define FRICO_C64 1

/* Factorisation / solve forms (frico_factor):
 *   FRICO_SOLVE_C128      complex128 LU of Z, in place (zgetrf semantics)
 *   FRICO_SOLVE_C64_IR    complex64 LU of a copy of Z, FP64 residual b - Z x against the kept Z,
 *                         iterative refinement to the requested tolerance
 *   FRICO_SOLVE_C64_SYM_IR  for a symmetric Z (solve_symmetric, src/lapack_wrappers.cpp:135-177):
 *                         complex64 L D L^T of the lower triangle of a copy of Z without pivoting,
 *                         refined as FRICO_SOLVE_C64_IR.  A zero pivot, or a refinement that stops
 *                         converging, falls back to FRICO_SOLVE_C64_IR by itself
 *                         (frico_factored_form tells which factorisation the answer came from).
 *   FRICO_SOLVE_C64_NP_IR complex64 LU of a copy of Z without pivoting, refined as
 *                         FRICO_SOLVE_C64_IR; no grid-wide pivot search per column, so the
 *                         factorisation is bound by its trailing products.  Falls back to
 *                         FRICO_SOLVE_C64_IR like the symmetric form.  For matrices whose
 *                         unpivoted factors stay bounded, as EFIE matrices' do. */
#define FRICO_SOLVE_C128 0
#define FRICO_SOLVE_C64_IR 1
#define FRICO_SOLVE_C64_SYM_IR 2
#define FRICO_SOLVE_C64_NP_IR 3

typedef struct frico_ctx frico_ctx;
typedef struct frico_msh frico_msh;

/* One RWG basis function (src/rwg_basis.h:41-62), in matrix order. */
typedef struct frico_rwg {
    int64_t itminus, itplus;     /* T- < T+ (src/input_gmsh.cpp:243-262) */
    int32_t loc_eminus, loc_eplus;
    int64_t edge_index;          /* position in the sorted unique edge list */
    int32_t has_interface, interface_tag;
    double length;               /* hypot(dx, dy) of the edge (src/geom_mesh.cpp:46-52) */
    double Aminus, Aplus;        /* 0.5 |e0 x e1| (src/geom_mesh.cpp:71-78) */
    double pminus[3], pplus[3];  /* free vertices, evmap {2,0,1} (src/rwg_basis.cpp:47-50) */
} frico_rwg;

typedef struct frico_solve_info {
    int32_t iterations;          /* refinement sweeps performed (0 for FRICO_SOLVE_C128) */
    int32_t converged;           /* 1 if every column met the tolerance */
    double max_rel_residual;     /* max over columns of ||b - Z x|| / ||b|| (if computed) */
} frico_solve_info;

typedef struct frico_profile {
    double fill_ms, rhs_ms, factor_ms, solve_ms, refine_ms; /* per phase, HIP events */
    double gemm_ms, gemm_flops;  /* trailing-update GEMM launches of the last factorisations */
    int64_t gemm_launches;
    double gemm_bytes;           /* their algorithmic bytes: C read and written once, operands once */
    double tripair_ms, tripair_pairs; /* triangle-pair kernel of the quadrature fill */
    int64_t tripair_launches;
} frico_profile;

/* Text of the last error on this thread ("" if none). */
const char* frico_last_error(void);

/* Context: one device, one HIP stream (default priority), owned device buffers. */
int frico_ctx_create(int device, frico_ctx** out);
int frico_ctx_destroy(frico_ctx* ctx);
int frico_ctx_synchronize(frico_ctx* ctx);
/* Enables HIP-event timing of phases and of every GEMM / fill launch (off by default). */
int frico_ctx_set_profiling(frico_ctx* ctx, int enable);
int frico_ctx_get_profile(frico_ctx* ctx, frico_profile* out);
/* Block width of the LU: 32, 64 or 128, or 0 (the default) for 128 from N = 6144 up and 64 below.
 * Panels of 128 columns are factored as two 64-column panel launches and give K = 128 trailing
 * products. */
int frico_ctx_set_block(frico_ctx* ctx, int nb);

/*
 * Mesh ingestion -> connectivity -> RWG function space, and the per-triangle quadrature data of
 * Dunavant degree `degree` (0..20).  Replaces the connectivity part of load_from_gmsh
 * (src/input_gmsh.cpp:120-140, 143-199, 216-322) and make_function_space
 * (src/rwg_basis.cpp:27-68); vertices must already be compacted.  iface_edges are the edges of
 * 1D entities (pairs of vertex indices) with their entity tags.
 */
int frico_mesh_load(frico_ctx* ctx, const double* vertices, int64_t nv, const int64_t* triangles,
                    const int32_t* tri_tags, int64_t nt, const int64_t* iface_edges,
                    const int32_t* iface_tags, int64_t nif, int degree);
int64_t frico_num_unknowns(const frico_ctx* ctx);
int64_t frico_num_triangles(const frico_ctx* ctx);
int frico_rwg_table(const frico_ctx* ctx, frico_rwg* out);

/*
 * Delta-gap edge orientation signs: reorient_deltagap_edges (src/emw_solver.cpp:222-350).
 * signs[N] receives +-1 for the gap functions when the reference would reorient (then
 * *applied = 1), otherwise all ones and *applied = 0.
 */
int frico_reorient_deltagap(frico_ctx* ctx, const int32_t* ifaces, int nif, int force,
                            double* signs, int* applied);

/* Impedance matrix Z (N x N) of the loaded mesh at `freq`, kept on the device:
 * compute_matrix / compute_matrix_approx (src/emw_solver.h:85-86).  prec = FRICO_C128 or
 * FRICO_C64 (complex64 bulk, FP64 near/self entries). */
int frico_fill(frico_ctx* ctx, double freq, int mode, int prec);
/* Z = (Z + Z^T)/2 (src/emw_solver.h:118-120), complex128 matrix only. */
int frico_symmetrize(frico_ctx* ctx);
/* Copies Z (complex128 after a C128 fill; complex64 values widened after a C64 fill). */
int frico_matrix_get(frico_ctx* ctx, double* Z, int64_t ldz);
/* Loads an arbitrary complex128 N x N matrix as Z (N from the mesh, or n if no mesh). */
int frico_matrix_set(frico_ctx* ctx, const double* Z, int64_t n, int64_t ldz);

/* Right-hand sides, one column per excitation, kept on the device.
 * update_rhs(delta_gap) (src/emw_solver.cpp:352-378): column 0 = sum over the gap functions of
 * -j l/(omega mu0) * V * sign; signs may be NULL.  update_rhs(plane_wave)
 * (src/emw_solver.cpp:380-416, src/sources.h:34-75): column a for incidence (theta[a], phi[a])
 * with complex amplitudes eth[2a..2a+1], eph[2a..2a+1]. */
int frico_rhs_deltagap(frico_ctx* ctx, double freq, const int32_t* ifaces, int nif, double v_re,
                       double v_im, const double* signs);
int frico_rhs_planewave(frico_ctx* ctx, double freq, int64_t nang, const double* theta,
                        const double* phi, const double* eth, const double* eph);
int frico_rhs_set(frico_ctx* ctx, const double* B, int64_t ldb, int64_t nrhs);
int frico_rhs_get(frico_ctx* ctx, double* B, int64_t ldb);

/* LU of Z: zgetrf (src/lapack_wrappers.cpp:69) in the given form.  Returns > 0 for a zero
 * pivot (the 1-based column), like zgetrf's info. */
int frico_factor(frico_ctx* ctx, int form);
/* Form of the current factorisation (FRICO_SOLVE_*), or -1 if there is none. */
int frico_factored_form(const frico_ctx* ctx);
/* X = Z^{-1} B for every right-hand side column: zgetrs (src/lapack_wrappers.cpp:83,124),
 * followed by refinement for FRICO_SOLVE_C64_IR until max_c ||b_c - Z x_c||/||b_c|| <= tol or
 * maxit sweeps.  info may be NULL. */
int frico_solve(frico_ctx* ctx, double tol, int maxit, frico_solve_info* info);
/* Relative residual ||b_c - Z x_c|| / ||b_c|| per column against a freshly supplied Z (the
 * current Z must be unfactored, e.g. after a new frico_fill). */
int frico_residual(frico_ctx* ctx, double* rel /* nrhs */);
int frico_solution_get(frico_ctx* ctx, double* X, int64_t ldx);
int frico_solution_set(frico_ctx* ctx, const double* X, int64_t ldx, int64_t nrhs);
int64_t frico_num_rhs(const frico_ctx* ctx);

/* RWG -> triangle scatter of solution column `col` (src/emw_solver.h:137-164):
 * tri_AJ (T x 3 complex, row-major [t][xyz]) and tri_AdivJ (T complex). */
int frico_triangle_currents(frico_ctx* ctx, int64_t col, double* tri_AJ, double* tri_AdivJ);

/* E and H at points from the solution columns (eval_fields, src/emw_postpro_common.cpp:31-66).
 * Point p uses solution column cols[p] (cols may be NULL: column 0).  E, H: npts x 3 complex,
 * row-major.  h_sign = +1 keeps the reference's exp(+jkR) in H (:60-61), -1 uses exp(-jkR). */
int frico_fields(frico_ctx* ctx, double freq, const double* pts, int64_t npts, const int64_t* cols,
                 double h_sign, double* E, double* H);

/*
 * frico_fields_noalloc: frico_fields on device buffers, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Every pointer is DEVICE memory of the
 * context's device: currents = ncols current vectors (N x ncols complex128, leading dimension ldi,
 * e.g. the caller's own solve), pts = npts x 3 doubles, cols = npts column indices (NULL: column 0
 * for every point), E and H = npts x 3 complex128 outputs, row-major [point][xyz] (either may be
 * NULL).  A point whose column is outside [0, ncols) gets NaN fields.  The triangle currents go to
 * a fixed per-device scratch of FRICO_FIELDS_NOALLOC_MAX_BYTES: 64 * T * ncols bytes must fit in it
 * (T = triangles), otherwise FRICO_ERR_LIMIT.  The work is only enqueued on the context's stream
 * (frico_ctx_synchronize waits for it).  Calls that overlap on one GPU (contexts on the same
 * device whose streams run concurrently) race on the shared scratch.
 */
#define FRICO_FIELDS_NOALLOC_MAX_BYTES (64LL * 1024 * 1024)
int frico_fields_noalloc(frico_ctx* ctx, double freq, const double* currents, int64_t ldi, int64_t ncols,
                         const double* pts, int64_t npts, const int32_t* cols, double h_sign, double* E,
                         double* H);

/*
 * gmsh .msh ingestion (format 4.1 ASCII): load_from_gmsh (src/input_gmsh.cpp:36-406) without the
 * gmsh library.  The arrays are what the reference reads back through gmsh's API, in its order:
 * nodes by entity (dim, tag) and file order within an entity; triangles of the surfaces by tag,
 * then file order; line elements of the curves likewise; vertices no triangle uses dropped (the
 * others keep their order).  Pass them to frico_mesh_load.  Physical groups are listed by name
 * (the reference's std::map), each with its dimension, tag and entity tags.
 * Returns FRICO_ERR_MESH with a message for a file it cannot read.
 */
int frico_msh_read(const char* path, frico_msh** out);
int frico_msh_free(frico_msh* msh);
int frico_msh_sizes(const frico_msh* msh, int64_t* nv, int64_t* nt, int64_t* nb, int32_t* ngroups);
int frico_msh_arrays(const frico_msh* msh, double* vertices, int64_t* triangles, int32_t* tags,
                     int64_t* bedges, int32_t* btags);
int frico_msh_group(const frico_msh* msh, int32_t i, char* name, int64_t name_size, int32_t* dim,
                    int32_t* tag, int32_t* entities, int32_t max_entities, int32_t* nentities);

/*
 * Near-field sampling planes (the -P argument, src/main.cpp:164-181).
 * frico_sampling_plane_parse: parse_sampling_plane (src/sampling_planes.cpp:30-106):
 *   "xy|yz|xz:width:height:h" or "nx:ny:nz:width:height:h".  axis[3] is the reference's `normal`:
 *   the axis about which its mesh turns the rectangle by 90 degrees ("xy" -> z, "yz" -> y,
 *   "xz" -> x), normalised in the long form.  Returns -1 for an invalid specification.
 * frico_sampling_plane_mesh: the plane's mesh (make_sampling_plane_mesh, :108-156): the rectangle
 *   [-w/2, w/2] x [-h/2, h/2] in z = 0 turned by pi/2 about axis, as a structured mesh of
 *   ceil(width/h) x ceil(height/h) cells in right triangles (the reference meshes it with gmsh).
 *   vertices: nv x 3 row-major, triangles: nt x 3; pass NULL for both to get nv and nt only.
 * eval_fields(mesh) (src/emw_postpro_common.cpp:77-90) is frico_fields at the mesh's vertices.
 */
int frico_sampling_plane_parse(const char* spec, double* axis, double* width, double* height,
                               double* h);
int frico_sampling_plane_mesh(const double* axis, double width, double height, double h,
                              double* vertices, int64_t max_vertices, int64_t* triangles,
                              int64_t max_triangles, int64_t* nv, int64_t* nt);

/*
 * Text outputs, byte for byte as the reference's std::println writes them ("{}" of a double is
 * the shortest round-trip std::to_chars form):
 * frico_write_port_sweep: port_sweep.txt, write_sweep_data (src/emw_ana_antenna.cpp:150-169);
 *   zin[2*i], zin[2*i+1] = Re, Im of Zin at step i.
 * frico_write_radiation_diagram: polar_<n>.txt, write_radiation_diagrams (:171-203); gain = 3 x 360
 *   (planes xy, yz, xz), as frico_antenna_sweep's gain_out for one frequency.
 * frico_write_bistatic_rcs: bistatic_rcs_<n>.txt (src/emw_ana_bistatic_rcs.cpp:97-117); table =
 *   359 x 4, as frico_bistatic_sweep's table_out for one frequency.
 * frico_format_double: the "{}" form of one double (for callers writing their own lines).
 */
int frico_write_port_sweep(const char* path, double Z0, int64_t nsteps, const double* freqs,
                           const double* zin, const double* swr);
int frico_write_radiation_diagram(const char* path, double freq, const double* gain);
int frico_write_bistatic_rcs(const char* path, const double* table);
int frico_format_double(double v, char* buf, int64_t size);

/*
 * HDF5 outputs (no HDF5 library needed; the file structures of the reference's own output,
 * share/hdf5/antenna_analysis/biquad.h5: superblock 0, version-1 object headers, symbol-table groups,
 * contiguous datasets, std::complex<double> as the compound {r, i}):
 * frico_write_antenna_h5: write_hdf5 (src/emw_ana_antenna.cpp:207-234): /frico/antenna_analysis/
 *   rdiag_dist, sweep_steps (uint64), Z0 and per step <n>/frequency, gain_XY, gain_YZ, gain_XZ (360
 *   each), port/voltage, current, power, impedance, gamma (complex) and port/swr.  port = nsteps x 10
 *   and gain = nsteps x 1080 as frico_antenna_sweep's port_out and gain_out; (v_re, v_im) the gap
 *   voltage.
 * frico_write_linear_system_h5: the -d dump (src/emw_ana_antenna.cpp:260-268): /frico/Z (n x n,
 *   entry [i][j] = Z(i, j) of the column-major Z), /frico/V and /frico/I (n x nrhs; I may be NULL,
 *   as in the bistatic sweep's dump, src/emw_solver.cpp:494-500).
 * frico_dump_linear_system: the same for a context: its complex128 matrix as it stands (the filled
 *   Z, or after frico_factor(FRICO_SOLVE_C128) the LU factors, as the reference's dump holds
 *   zgetrf's in-place factors), its right-hand sides and, if solved, its solution.
 */
int frico_write_antenna_h5(const char* path, double rdiag_dist, double Z0, int64_t nsteps,
                           const double* freqs, double v_re, double v_im, const double* port,
                           const double* gain);
int frico_write_linear_system_h5(const char* path, int64_t n, const double* Z, int64_t ldz,
                                 int64_t nrhs, const double* V, int64_t ldv, const double* I,
                                 int64_t ldi);
int frico_dump_linear_system(frico_ctx* ctx, const char* path);

/*
 * Frequency points of a sweep as init_sweep builds them (src/emw_solver.cpp:460-471): start, then
 * freq += step while freq <= end (accumulated, so rounding can drop the end point).  *count
 * receives the number of points; freqs (max_freqs entries, may be NULL) receives them.  Returns -5
 * if freqs is too short, -3 for a non-positive step.  The sweeps below take such lists.
 */
int frico_sweep_frequencies(double start, double end, double step, double* freqs, int64_t max_freqs,
                            int64_t* count);

/*
 * Whole analyses, mirroring the reference entry points for one frequency each.
 *
 * frico_antenna_sweep: do_sweep(simulation&, antenna_analysis&) (src/emw_ana_antenna.cpp:236-295)
 *   per frequency: fill (mode), optional symmetrisation, delta-gap RHS, solve, port values
 *   (src/emw_ana_antenna.cpp:76-105) and gain on the xy/yz/xz cuts (:33-65).
 *   port_out[10*f ..]: I(re,im) P(re,im) Zin(re,im) Gamma(re,im) SWR, residual
 *   gain_out[1080*f ..]: 3 x 360 gains, planes xy, yz, xz.
 * frico_bistatic_sweep: do_sweep(simulation&, const bistatic_rcs_analysis&)
 *   (src/emw_ana_bistatic_rcs.cpp:120-151); table_out[359*4*f ..] = VV_theta HH_theta VV_phi HH_phi.
 * frico_monostatic_sweep: nang plane waves at (theta[a], phi[a]) solved as one block; the
 *   back-scattered sigma_VV, sigma_HH at (R, theta[a], phi[a]) (formula of
 *   src/emw_ana_bistatic_rcs.cpp:62-82); sigma_out[2*nang*f ..].
 * `form` selects FRICO_SOLVE_C128 or FRICO_SOLVE_C64_IR (then tol/maxit apply).  As in run_context
 *   (src/emw_solver.h:124-128), approximate-mode fills (and antenna sweeps with force_symmetry) are
 *   solved as symmetric systems: FRICO_SOLVE_C64_IR becomes FRICO_SOLVE_C64_SYM_IR there.
 */
int frico_antenna_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode,
                        const int32_t* ifaces, int nif, double v_re, double v_im, double R,
                        double Z0, int force_symmetry, int force_reorient, int form, double tol,
                        int maxit, double* port_out, double* gain_out);
int frico_bistatic_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode, double theta_r,
                         double phi_r, double eth, double eph, double R, int form, double tol,
                         int maxit, double* table_out);
int frico_monostatic_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode, int64_t nang,
                           const double* theta, const double* phi, double eth, double eph,
                           double R, int form, double tol, int maxit, double* sigma_out,
                           frico_solve_info* info_out);

/*
 * Adjoint pair (no counterpart in the reference; it differentiates the antenna analysis of
 * src/emw_ana_antenna.cpp:33-105 through compute_matrix, src/emw_solver.cpp:45-142).
 * frico_adjoint_set_fill: the fill of the forward entry points below, FRICO_FILL_QUAD (the default,
 *   compute_matrix) or FRICO_FILL_APPROX (compute_matrix_approx, src/emw_solver.cpp:144-220, the
 *   reference's -A); the backward pass differentiates the fill its forward pass used.
 * frico_adjoint_forward: fill at `freq` (quadrature mode: Dunavant degree 0 to 7 of the loaded
 *   mesh), delta-gap excitation on ifaces, complex128 LU and solve; out[0..2] = Re Zin, Im Zin and
 *   the gain G = 4 pi |obs|^2 |S| / Re(V conj(I_port)/2) at the point obs[3].
 * frico_adjoint_backward: grad[3*v + c] = g_zin_re dRe(Zin)/dx + g_zin_im dIm(Zin)/dx + g_gain dG/dx
 *   for every vertex coordinate, from two solves with Z^T on the forward factors and the
 *   device contraction of dZ/dx with the forward and adjoint currents.
 */
int frico_adjoint_set_fill(frico_ctx* ctx, int mode);
int frico_adjoint_forward(frico_ctx* ctx, double freq, const int32_t* ifaces, int nif, double v_re,
                          double v_im, const double* obs, double* out);
int frico_adjoint_backward(frico_ctx* ctx, double g_zin_re, double g_zin_im, double g_gain,
                           double* grad);

/*
 * RCS adjoint pair (no counterpart in the reference; it differentiates bistatic_RCS_unit_Ei,
 * src/emw_ana_bistatic_rcs.cpp:62-82, of the currents of update_rhs(plane_wave),
 * src/emw_solver.cpp:380-416, through compute_matrix, src/emw_solver.cpp:45-142).
 * frico_adjoint_rcs_forward: fill at `freq` in the mode of frico_adjoint_set_fill (quadrature mode:
 *   Dunavant degree 0 to 7; the plane-wave moments keep the loaded degree's rule), one plane
 *   wave from (theta_i, phi_i) with complex amplitudes eth[2], eph[2], complex128 LU and solve;
 *   out[0] = sigma_VV, out[1] = sigma_HH at sph2rect(R, theta_s, phi_s) (monostatic: theta_s =
 *   theta_i, phi_s = phi_i).
 * frico_adjoint_rcs_backward: grad[3*v + c] = g_vv dsigma_VV/dx + g_hh dsigma_HH/dx for every vertex
 *   coordinate, from one solve with Z^T on the forward factors, the device contraction of dZ/dx
 *   with the forward and adjoint currents, and the vertex derivatives of the plane-wave moments and
 *   of the field sums.
 */
int frico_adjoint_rcs_forward(frico_ctx* ctx, double freq, double theta_i, double phi_i, const double* eth,
                              const double* eph, double theta_s, double phi_s, double R, double* out);
int frico_adjoint_rcs_backward(frico_ctx* ctx, double g_vv, double g_hh, double* grad);

/*
 * Batched uneven antenna workloads (the reference runs one mesh and one frequency at a time:
 * do_sweep, src/emw_ana_antenna.cpp:236-295, over init_sweep's frequencies,
 * src/emw_solver.cpp:460-471).  A batch owns `workers` contexts on one device; meshes are added
 * once and then swept together.
 * frico_batch_add_mesh: mesh arrays as for frico_mesh_load plus the feed interface tags; returns
 *   the mesh index (>= 0) or an error.
 * frico_batch_antenna: for every mesh and every freqs[f], quadrature-mode fill (fill_prec
 *   FRICO_C128 or FRICO_C64), delta-gap excitation of v on the feed (oriented once per mesh,
 *   src/emw_ana_antenna.cpp:238), factor + solve in `form`, and the input impedance
 *   zin[2*(m*nfreq + f) ..] = V / I_port (compute_port_values, src/emw_ana_antenna.cpp:76-105);
 *   info[m*nfreq + f] is the solve status.  Meshes run largest first on up to `workers` concurrent
 *   contexts (fewer when ceil(Nmax / R) panel workgroups per context would exceed the CUs).
 *   The unpivoted form with a complex64 fill (256 <= N < 4096) runs each mesh's frequencies in
 *   lockstep: one factorisation launch sequence and one refinement sequence for all of them;
 *   systems that stall are retried in lockstep with the row order of one frequency's pivoted LU.
 *   FRICO_BATCH_GRAPHS replays a HIP graph per mesh, recorded from the kernel launches, for the
 *   complex64 LU (FRICO_SOLVE_C64_IR or FRICO_SOLVE_C64_NP_IR) and each refinement sweep; the
 *   unpivoted form falls back to the pivoted LU on a zero pivot or a stalled refinement, and the
 *   mesh's later frequencies then take the pivoted LU directly (the worker count is sized for
 *   that fall-back's panel).  The
 *   symmetric form is rejected (-5): the batch's quadrature-mode Z is not symmetric, and
 *   run_context solves symmetrically only approximate-mode or symmetrised matrices
 *   (src/emw_solver.h:124-128).
 *   FRICO_BATCH_AUTOTUNE picks the LU block width per size class (N / 256) by
 *   timing the candidates on the first system of that class.
 * frico_batch_block_sizes: nb[m] = block width used for mesh m by the last frico_batch_antenna;
 *   returns the number of meshes whose frequencies after the first ran as graph replays.
 * frico_batch_factored_forms: forms[m*nfreq + f] = the factorisation (FRICO_SOLVE_*) the last
 *   frico_batch_antenna solved unit (m, f) with, after any fall-back (frico_factored_form of the
 *   unit's context); returns the number of units.
 */
#define FRICO_BATCH_GRAPHS 1
#define FRICO_BATCH_AUTOTUNE 2
typedef struct frico_batch frico_batch;
int frico_batch_create(int device, int workers, frico_batch** out);
int frico_batch_destroy(frico_batch* batch);
int frico_batch_add_mesh(frico_batch* batch, const double* vertices, int64_t nv,
                         const int64_t* triangles, const int32_t* tri_tags, int64_t nt,
                         const int64_t* iface_edges, const int32_t* iface_tags, int64_t nif,
                         int degree, const int32_t* feed, int nfeed);
int frico_batch_num_meshes(const frico_batch* batch);
/* Concurrent contexts used by the last frico_batch_antenna: min(workers, GPU_MAX_HW_QUEUES
 * (default 4), CUs / ceil(Nmax / R)). */
int frico_batch_workers_used(const frico_batch* batch);
int frico_batch_antenna(frico_batch* batch, const double* freqs, int nfreq, int fill_prec, int form,
                        double tol, int maxit, int flags, double v_re, double v_im, double* zin,
                        frico_solve_info* info);
int frico_batch_block_sizes(const frico_batch* batch, int32_t* nb);
int frico_batch_factored_forms(const frico_batch* batch, int32_t* forms);

#ifdef __cplusplus
}
#endif

#endif /* FRICO_GPU_H */
