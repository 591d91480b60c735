// A compiled C++ caller of the C ABI (include/frico_gpu.h): the run_context binding a FRICO
// maintainer would add (src/emw_solver.h:92-167), for an antenna analysis with a delta-gap source.
//
//   run_context <mesh.msh> <feed physical group> <freq Hz> <degree> <approx 0|1> <out file>
//
// Reads a gmsh 4.1 ASCII mesh through frico_msh_read (load_from_gmsh, src/input_gmsh.cpp:36-406),
// loads it, and per run_context: fills Z (compute_matrix or compute_matrix_approx), writes the
// delta-gap right-hand side (update_rhs, src/emw_solver.cpp:222-378: reorientation, then the
// moments), factors and solves (solve_general / solve_symmetric by run_context's rule,
// src/emw_solver.h:124-128), and reads the currents back.  Then, as compute_port_values and
// compute_gain do (src/emw_ana_antenna.cpp:33-105), it forms the input impedance from the port
// current and the gain on the xy, yz and xz circles of radius 10 m from frico_fields.  The same
// analysis through frico_antenna_sweep is written beside it.  Output: text lines
//   zin <re> <im>          (this binding)
//   zin_sweep <re> <im>    (frico_antenna_sweep)
//   gain <1080 values>     (this binding), gain_sweep <1080 values>
#include "frico_gpu.h"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using cd = std::complex<double>;

namespace {

constexpr double PI = 3.14159265358979323846;

struct Mesh {
    std::vector<double> V;
    std::vector<int64_t> T, B;
    std::vector<int32_t> tags, btags;
    std::vector<int32_t> feed;
};

int fail(const char* what, int rc)
{
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, frico_last_error());
    return 1;
}

bool read_mesh(const char* path, const char* feed_group, Mesh& m)
{
    frico_msh* h = nullptr;
    if (frico_msh_read(path, &h) != FRICO_OK)
        return false;
    int64_t nv = 0, nt = 0, nb = 0;
    int32_t ng = 0;
    frico_msh_sizes(h, &nv, &nt, &nb, &ng);
    m.V.resize(3 * nv);
    m.T.resize(3 * nt);
    m.tags.resize(nt);
    m.B.resize(2 * nb);
    m.btags.resize(nb);
    frico_msh_arrays(h, m.V.data(), m.T.data(), m.tags.data(), m.B.data(), m.btags.data());
    // the source's interfaces are the entity tags of its physical group (src/main.cpp:183-199)
    for (int32_t g = 0; g < ng; g++) {
        char name[256];
        int32_t dim = 0, tag = 0, ne = 0;
        frico_msh_group(h, g, name, sizeof name, &dim, &tag, nullptr, 0, &ne);
        if (std::strcmp(name, feed_group) == 0) {
            m.feed.resize(ne);
            frico_msh_group(h, g, nullptr, 0, nullptr, nullptr, m.feed.data(), ne, &ne);
        }
    }
    frico_msh_free(h);
    return !m.feed.empty();
}

// update_rhs(sim, ctx, delta_gap): the reference reorients the gap edges once
// (reorient_deltagap_edges) and accumulates -j l / (omega mu0) V sign into V.
int upload_rhs(frico_ctx* ctx, double freq, const Mesh& m, cd voltage, std::vector<double>& signs)
{
    const int64_t N = frico_num_unknowns(ctx);
    signs.assign((size_t)N, 1.0);
    int applied = 0;
    int rc = frico_reorient_deltagap(ctx, m.feed.data(), (int)m.feed.size(), 0, signs.data(), &applied);
    if (rc)
        return rc;
    return frico_rhs_deltagap(ctx, freq, m.feed.data(), (int)m.feed.size(), voltage.real(), voltage.imag(),
                              applied ? signs.data() : nullptr);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s mesh.msh feed_group freq degree approx out\n", argv[0]);
        return 2;
    }
    const char* path = argv[1];
    const char* feed_group = argv[2];
    const double freq = std::atof(argv[3]);
    const int degree = std::atoi(argv[4]);
    const bool approx = std::atoi(argv[5]) != 0;
    Mesh m;
    if (!read_mesh(path, feed_group, m))
        return fail("frico_msh_read / feed group", -1);

    frico_ctx* ctx = nullptr;
    int rc = frico_ctx_create(0, &ctx);
    if (rc)
        return fail("frico_ctx_create", rc);
    rc = frico_mesh_load(ctx, m.V.data(), (int64_t)m.V.size() / 3, m.T.data(), m.tags.data(), (int64_t)m.tags.size(),
                         m.B.data(), m.btags.data(), (int64_t)m.btags.size(), degree);
    if (rc)
        return fail("frico_mesh_load", rc);
    const int64_t N = frico_num_unknowns(ctx);

    // ---- run_context ------------------------------------------------------------------------
    const cd voltage = 1.0;
    const int mode = approx ? FRICO_FILL_APPROX : FRICO_FILL_QUAD;
    if ((rc = frico_fill(ctx, freq, mode, FRICO_C128)))
        return fail("frico_fill", rc);
    std::vector<double> signs;
    if ((rc = upload_rhs(ctx, freq, m, voltage, signs)))
        return fail("upload_rhs", rc);
    // solve_symmetric where run_context calls it (approximate mode), solve_general elsewhere
    const int form = approx ? FRICO_SOLVE_C64_SYM_IR : FRICO_SOLVE_C64_IR;
    if ((rc = frico_factor(ctx, form)) != 0)
        return fail("frico_factor", rc);
    frico_solve_info info{};
    if ((rc = frico_solve(ctx, 1e-12, 10, &info)))
        return fail("frico_solve", rc);
    std::vector<cd> I((size_t)N);
    if ((rc = frico_solution_get(ctx, reinterpret_cast<double*>(I.data()), N)))
        return fail("frico_solution_get", rc);

    // ---- compute_port_values: I_port = sum l I sign over the gap functions --------------------
    std::vector<frico_rwg> bf((size_t)N);
    frico_rwg_table(ctx, bf.data());
    cd ip = 0.0, power = 0.0;
    for (int64_t n = 0; n < N; n++)
        if (bf[n].has_interface)
            for (int32_t f : m.feed)
                if (bf[n].interface_tag == f) {
                    const cd In = bf[n].length * I[n] * signs[n];
                    ip += In;
                    power += 0.5 * voltage * std::conj(In);
                }
    const cd zin = voltage / ip;
    const double pwr = std::real(power);

    // ---- compute_gain: |S| of E x conj(H) / 2 on three circles of radius 10 m ---------------------
    const double R = 10.0;
    std::vector<double> pts(3 * 1080);
    for (int plane = 0; plane < 3; plane++)
        for (int d = 0; d < 360; d++) {
            const double a = d * PI / 180.0, c = R * std::cos(a), s = R * std::sin(a);
            double* p = &pts[3 * (plane * 360 + d)];
            if (plane == 0) { p[0] = c; p[1] = s; p[2] = 0; }       // xy
            else if (plane == 1) { p[0] = 0; p[1] = c; p[2] = s; }  // yz
            else { p[0] = -c; p[1] = 0; p[2] = s; }                 // xz
        }
    std::vector<cd> E(3 * 1080), H(3 * 1080);
    if ((rc = frico_fields(ctx, freq, pts.data(), 1080, nullptr, 1.0, reinterpret_cast<double*>(E.data()),
                           reinterpret_cast<double*>(H.data()))))
        return fail("frico_fields", rc);
    std::vector<double> gain(1080);
    for (int q = 0; q < 1080; q++) {
        const cd* e = &E[3 * q];
        const cd* h = &H[3 * q];
        const cd S[3] = {0.5 * (e[1] * std::conj(h[2]) - e[2] * std::conj(h[1])),
                         0.5 * (e[2] * std::conj(h[0]) - e[0] * std::conj(h[2])),
                         0.5 * (e[0] * std::conj(h[1]) - e[1] * std::conj(h[0]))};
        const double absS = std::sqrt(std::norm(S[0]) + std::norm(S[1]) + std::norm(S[2]));
        const double* p = &pts[3 * q];
        const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); // R = norm(P) (:58)
        gain[q] = 4.0 * PI * r * r * absS / pwr;
    }

    // ---- the same analysis through frico_antenna_sweep -------------------------------------------
    double port[10], gsweep[1080];
    if ((rc = frico_antenna_sweep(ctx, &freq, 1, mode, m.feed.data(), (int)m.feed.size(), voltage.real(), voltage.imag(),
                                  R, 50.0, 0, 0, FRICO_SOLVE_C64_IR, 1e-12, 10, port, gsweep)))
        return fail("frico_antenna_sweep", rc);

    FILE* out = std::fopen(argv[6], "w");
    if (!out)
        return fail("fopen", -1);
    std::fprintf(out, "N %lld sweeps %d residual %.17g\n", (long long)N, info.iterations, info.max_rel_residual);
    std::fprintf(out, "zin %.17g %.17g\n", zin.real(), zin.imag());
    std::fprintf(out, "zin_sweep %.17g %.17g\n", port[4], port[5]);
    std::fprintf(out, "gain");
    for (double g : gain)
        std::fprintf(out, " %.17g", g);
    std::fprintf(out, "\ngain_sweep");
    for (double g : gsweep)
        std::fprintf(out, " %.17g", g);
    std::fprintf(out, "\n");
    std::fclose(out);
    frico_ctx_destroy(ctx);
    std::printf("run_context: N=%lld Zin=%.10f%+.10fj (%d sweeps)\n", (long long)N, zin.real(), zin.imag(),
                info.iterations);
    return 0;
}
