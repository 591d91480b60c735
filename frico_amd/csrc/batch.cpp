// Batched uneven antenna workloads: many meshes, each at many frequencies, on one GPU.
//
// The reference runs one mesh at a time and its frequencies one after another
// (do_sweep, src/emw_ana_antenna.cpp:236-295; init_sweep, src/emw_solver.cpp:460-471).  Here the
// unit of work is one mesh at all of its frequencies:
//   - units are ordered by decreasing N (longest processing time first) and taken by W worker
//     threads, each with its own context (stream, workspace), so W systems are in flight at once;
//   - per unit the worker loads the mesh, orients the delta gap once, and picks the LU block width
//     from a table filled by timing the candidates on the first system of each size class;
//   - per frequency: impedance fill and delta-gap RHS are enqueued eagerly; the complex64
//     conversion, the LU (pivoted, or without pivoting for FRICO_SOLVE_C64_NP_IR, which returns
//     to the pivoted LU eagerly on a zero pivot or a stalled refinement) and the first refinement
//     sweep are one HIP graph, built on the second frequency from the recorded kernel launches and
//     replayed for the rest, and every further sweep is a second graph.  The graphs are assembled node by node on the worker's thread
//     (no stream capture), so the other workers' launches and allocations cannot disturb them.
// Co-residency: a panel launch needs ceil(N / R) workgroups resident at once (one per CU), so the
// worker count is capped at CUs / ceil(Nmax / R); it is also capped at the process's hardware
// queues (GPU_MAX_HW_QUEUES, 4 by default), beyond which streams serialise behind each other.
#include "ctx.hpp"
#include "solver.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <complex>
#include <map>
#include <mutex>
#include <string>
#include <thread>

using namespace frico;

struct frico_batch {
    struct Mesh {
        std::vector<double> V;
        std::vector<int64_t> T;
        std::vector<int32_t> tags;
        std::vector<int64_t> iface;
        std::vector<int32_t> itags;
        std::vector<int32_t> feed;
        int degree = 5;
        int64_t N = 0;
    };
    int device = 0;
    int cus = 0;
    std::vector<frico_ctx*> ctx;
    std::vector<Mesh> meshes;
    std::vector<int32_t> mesh_nb; // block width used for each mesh by the last run
    std::vector<int32_t> unit_form; // factorisation each (mesh, frequency) was solved with (FRICO_SOLVE_*)
    std::atomic<int> graph_units{0}; // units of the last run that replayed graphs
    int workers_used = 0;            // concurrent contexts of the last run
    std::mutex mu;
    std::map<int64_t, int> tuned; // size class (N / 256) -> block width
};

namespace {

using cd = std::complex<double>;

// Internal edges (shared by exactly two triangles) = RWG unknowns, for ordering only.
int64_t count_internal_edges(const int64_t* tri, int64_t nt)
{
    std::vector<std::pair<int64_t, int64_t>> e;
    e.reserve((size_t)(3 * nt));
    for (int64_t t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            const int64_t a = tri[3 * t + k], b = tri[3 * t + (k + 1) % 3];
            e.emplace_back(std::min(a, b), std::max(a, b));
        }
    std::sort(e.begin(), e.end());
    int64_t n = 0;
    for (size_t i = 0; i < e.size();) {
        size_t j = i;
        while (j < e.size() && e[j] == e[i])
            j++;
        n += (j - i) == 2;
        i = j;
    }
    return n;
}

struct RunOpts {
    const double* freqs;
    int nfreq;
    int fill_prec;
    int form;
    double tol;
    int maxit;
    bool graphs;
    bool autotune;
    cd voltage;
};

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    bool exact = true; // residual form the graph was recorded with
    void reset()
    {
        if (x)
            (void)hipGraphExecDestroy(x);
        if (g)
            (void)hipGraphDestroy(g);
        x = nullptr;
        g = nullptr;
    }
    ~Graph() { reset(); }
};

// Builds an executable graph from the kernels fn launches on the context's stream (recorded, not
// run: see GraphRecorder).  fn must enqueue kernels only.
template <typename F> int record(frico_ctx* c, Graph& gr, F fn)
{
    GraphRecorder rec;
    rec.st = c->st;
    FRICO_HIP_CHECK(hipGraphCreate(&rec.graph, 0));
    graph_recorder() = &rec;
    const int rc = fn();
    graph_recorder() = nullptr;
    if (rc || rec.err != hipSuccess) {
        (void)hipGraphDestroy(rec.graph);
        return rc ? rc : hip_fail(rec.err, "hipGraphAddKernelNode", __FILE__, __LINE__);
    }
    gr.g = rec.graph;
    FRICO_HIP_CHECK(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
    return 0;
}

// Eager factor + solve (used for the first frequency, for tuning and when graphs are off).
int solve_eager(frico_ctx* c, const RunOpts& o, frico_solve_info* info)
{
    int rc = frico_factor(c, o.form);
    if (rc)
        return rc;
    return frico_solve(c, o.tol, o.maxit, info);
}

// Complex64 LU + refinement by graph replay; g1 = conversion + LU + first sweep, g2 = one sweep.
int solve_graphs(frico_ctx* c, const RunOpts& o, Graph& g1, Graph& g2, frico_solve_info* info)
{
    LuWork wk;
    int rc = factor_prepare(c, o.form, wk, false);
    if (rc)
        return rc;
    // the state factor_finish leaves, set before recording: the recorded sweeps choose their path
    // by it (the factors' form, the right-hand sides carried through the factorisation)
    c->factored = o.form;
    c->aug_valid = c->aug_cols > 0;
    c->aug_version = c->rhs_version;
    c->wv_ok = wk.wv != nullptr;
    if ((rc = refine_prepare(c)))
        return rc;
    std::vector<double> bn;
    if ((rc = col_norms(c, c->B.p, c->nrhs, bn)))
        return rc;
    // the residual form follows the tolerance (as in frico_solve) and is part of the recording
    c->resid_exact = !(o.tol >= 1e-10);
    if (g1.x && g1.exact != c->resid_exact) {
        g1.reset();
        g2.reset();
    }
    g1.exact = g2.exact = c->resid_exact;
    if (!g1.x && ((rc = record(c, g1, [&] {
                      int r = factor_enqueue(c, o.form, wk);
                      return r ? r : refine_enqueue(c, true);
                  })) ||
                  (rc = record(c, g2, [&] { return refine_enqueue(c, false); }))))
        return rc;
    FRICO_HIP_CHECK(hipGraphLaunch(g1.x, c->st));
    if ((rc = factor_finish(c, o.form, wk)))
        return rc; // > 0: a zero pivot (the unpivoted form); the caller takes the pivoted LU
    frico_solve_info inf{0, 0, -1.0};
    for (int it = 1; it <= o.maxit; it++) {
        if (it > 1)
            FRICO_HIP_CHECK(hipGraphLaunch(g2.x, c->st));
        double worst = 0;
        if ((rc = read_refine_norms(c, bn, &worst)))
            return rc;
        inf.iterations = it;
        inf.max_rel_residual = worst;
        if (worst <= o.tol) {
            inf.converged = 1;
            break;
        }
    }
    c->have_x = true;
    *info = inf;
    if (!inf.converged)
        return set_error("refinement stopped at relative residual %.3e > %.3e after %d sweeps", inf.max_rel_residual,
                         o.tol, inf.iterations),
               FRICO_ERR_NOCONV;
    return 0;
}

// Block width for this size class: the table entry, or (autotune) the faster of the candidates
// on the current system, timed as factor + solve.  The complex128 form factors Z in place, so the
// system at `freq` is filled again before every pass and once more at the end.
int pick_block(frico_batch* b, frico_ctx* c, const RunOpts& o, double freq, int* nb_out)
{
    const bool refill = o.form == FRICO_SOLVE_C128;
    const int64_t cls = c->N / 256;
    {
        std::lock_guard<std::mutex> lock(b->mu);
        auto it = b->tuned.find(cls);
        if (it != b->tuned.end() || !o.autotune) {
            *nb_out = it != b->tuned.end() ? it->second : 64;
            return 0;
        }
    }
    // the unpivoted form's wider steps (nb = 128: paired trailing products) pay from a few
    // thousand unknowns; the pivoted panel's per-column cost favours narrower blocks
    const bool np = o.form == FRICO_SOLVE_C64_NP_IR && c->N >= 1536;
    const int cand[2] = {np ? 128 : 64, np ? 64 : 32};
    double best = 1e300;
    int best_nb = 64;
    int rc = 0;
    for (int nb : cand) {
        rc = frico_ctx_set_block(c, nb);
        if (rc)
            return rc;
        frico_solve_info inf{};
        // one untimed pass (workspace allocation, first launches), then the timed one
        for (int pass = 0; pass < 2; pass++) {
            if (refill && (rc = frico_fill(c, freq, FRICO_FILL_QUAD, o.fill_prec)))
                return rc;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = solve_eager(c, o, &inf)))
                return rc;
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (pass == 1 && t < best) {
                best = t;
                best_nb = nb;
            }
        }
    }
    if (refill && (rc = frico_fill(c, freq, FRICO_FILL_QUAD, o.fill_prec)))
        return rc;
    std::lock_guard<std::mutex> lock(b->mu);
    b->tuned.emplace(cls, best_nb);
    *nb_out = b->tuned[cls];
    return 0;
}

// compute_port_values (src/emw_ana_antenna.cpp:76-105): Zin = V / I_port, I_port = sum l I sign over the feed
cd port_zin(const frico_batch::Mesh& m, const std::vector<frico_rwg>& bf, const std::vector<double>& signs, bool applied,
            const cd* I, int64_t N, cd voltage)
{
    cd ip = 0.0;
    for (int64_t n = 0; n < N; n++) {
        if (!bf[n].has_interface)
            continue;
        for (int32_t itf : m.feed)
            if (bf[n].interface_tag == itf)
                ip += bf[n].length * I[n] * (applied ? signs[n] : 1.0);
    }
    return voltage / ip;
}

// only: the frequencies to solve (null: all)
int run_unit(frico_batch* b, frico_ctx* c, int mi, const RunOpts& o, double* zin, frico_solve_info* infos,
             const std::vector<char>* only = nullptr)
{
    const frico_batch::Mesh& m = b->meshes[mi];
    int rc = frico_mesh_load(c, m.V.data(), (int64_t)(m.V.size() / 3), m.T.data(), m.tags.data(),
                             (int64_t)m.tags.size(), m.iface.empty() ? nullptr : m.iface.data(),
                             m.itags.empty() ? nullptr : m.itags.data(), (int64_t)m.itags.size(), m.degree);
    if (rc)
        return rc;
    const int64_t N = c->N;
    // reorient_deltagap_edges once per sweep (src/emw_ana_antenna.cpp:238)
    std::vector<double> signs((size_t)N);
    int applied = 0;
    if ((rc = frico_reorient_deltagap(c, m.feed.data(), (int)m.feed.size(), 0, signs.data(), &applied)))
        return rc;
    std::vector<frico_rwg> bf((size_t)N);
    if ((rc = frico_rwg_table(c, bf.data())))
        return rc;
    std::vector<cd> I((size_t)N);
    Graph g1, g2;
    int nb = 64;
    // the form of this unit's solves: once the unpivoted factors of a frequency have fallen back to
    // the pivoted LU (a zero pivot, or a refinement that stalls: factors that grow without
    // interchanges), the mesh's other frequencies go to the pivoted LU at once rather than each
    // paying for the unpivoted factorisation and sweeps first
    RunOpts ou = o;
    bool first = true;
    for (int f = 0; f < o.nfreq; f++) {
        if (only && !(*only)[f])
            continue;
        const double freq = o.freqs[f];
        if ((rc = frico_fill(c, freq, FRICO_FILL_QUAD, o.fill_prec)) ||
            (rc = frico_rhs_deltagap(c, freq, m.feed.data(), (int)m.feed.size(), o.voltage.real(), o.voltage.imag(),
                                     applied ? signs.data() : nullptr)))
            return rc;
        frico_solve_info& inf = infos[(size_t)mi * o.nfreq + f];
        if (first) {
            first = false;
            if ((rc = pick_block(b, c, ou, freq, &nb)) || (rc = frico_ctx_set_block(c, nb)))
                return rc;
            b->mesh_nb[mi] = nb;
            rc = solve_eager(c, ou, &inf);
        } else if (ou.graphs && (ou.form == FRICO_SOLVE_C64_IR || ou.form == FRICO_SOLVE_C64_NP_IR)) {
            rc = solve_graphs(c, ou, g1, g2, &inf);
            if (ou.form == FRICO_SOLVE_C64_NP_IR && (rc > 0 || rc == FRICO_ERR_NOCONV)) {
                // a zero pivot or a stalled refinement of the unpivoted factors: the pivoted LU, as
                // frico_factor and frico_solve do
                RunOpts op = ou;
                op.form = FRICO_SOLVE_C64_IR;
                rc = solve_eager(c, op, &inf);
            }
        } else {
            rc = solve_eager(c, ou, &inf);
        }
        if (rc)
            return rc;
        if (ou.form == FRICO_SOLVE_C64_NP_IR && c->factored == FRICO_SOLVE_C64_IR) {
            ou.form = FRICO_SOLVE_C64_IR;
            g1.reset(); // recorded for the unpivoted form: recorded again for the pivoted one
            g2.reset();
        }
        b->unit_form[(size_t)mi * o.nfreq + f] = c->factored;
        if (f == o.nfreq - 1 && g1.x)
            b->graph_units++;
        if ((rc = frico_solution_get(c, reinterpret_cast<double*>(I.data()), N)))
            return rc;
        const cd z = port_zin(m, bf, signs, applied != 0, I.data(), N, o.voltage);
        zin[2 * ((size_t)mi * o.nfreq + f)] = z.real();
        zin[2 * ((size_t)mi * o.nfreq + f) + 1] = z.imag();
    }
    return 0;
}

// Buffers of a worker's lockstep units (system z of a unit at z times the per-system size).
struct LockBufs {
    DevBuf<c64> Z, LU, D;
    DevBuf<c128> B, X, R;
    DevBuf<double> norms;
    DevBuf<unsigned long long> words;
    DevBuf<int> info, perm;
    DevBuf<unsigned char> split;
};

// FRICO_BATCH_LOCKSTEP=0 keeps one system per launch for the unpivoted form.
bool lockstep_on()
{
    static const bool on = [] {
        const char* e = std::getenv("FRICO_BATCH_LOCKSTEP");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

bool lockstep_fits(const frico_ctx* c, const RunOpts& o, int64_t N)
{
    return lockstep_on() && o.form == FRICO_SOLVE_C64_NP_IR && o.fill_prec == FRICO_C64 && o.nfreq >= 2 &&
           o.nfreq <= 65535 && N >= 256 && N + 1 <= 4096 && c->gemm_f16;
}

// Row gathers for a static row order of the unpivoted factors: dst(i, j) = src(perm[i], j).
__global__ __launch_bounds__(256) void gather_rows_c64(const c64* __restrict__ src, int64_t lds, c64* __restrict__ dst,
                                                       int64_t ldd, int64_t rows, const int* __restrict__ perm)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i < rows)
        dst[i + j * ldd] = src[perm[i] + j * lds];
}
// dst(i, j) = complex64(src(perm[i], j)), or without perm the conversion alone; column j of dst
// at dst + j sd (systems as columns)
__global__ __launch_bounds__(256) void gather_c128_c64(const c128* __restrict__ src, int64_t lds, c64* __restrict__ dst,
                                                       int64_t sd, int64_t rows, const int* __restrict__ perm)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i < rows) {
        const c128 v = src[(perm ? perm[i] : i) + j * lds];
        dst[i + j * sd] = c64{(float)v.re, (float)v.im};
    }
}

struct UnitMesh {
    const frico_batch::Mesh* m;
    std::vector<double> signs;
    int applied = 0;
    std::vector<frico_rwg> bf;
    int64_t N = 0;
};

// One lockstep pass over the frequencies sel of a loaded mesh: the systems are filled one after
// another into the unit's buffers, factored together (lu_nopiv_batched, the right-hand side carried
// as column N) and refined together (chained solves and matrix-vector residuals of all systems in
// one launch each) until every system has met the tolerance or stalled.  perm (device, or null): a
// static row order, the factors being those of P Z (the refinement's residual is of Z itself, and
// each sweep's right-hand side is taken through P).  The frequencies that do not converge are
// returned in failed.
int lockstep_pass(frico_batch* b, frico_ctx* c, int mi, const UnitMesh& um, const RunOpts& o, double* zin,
                  frico_solve_info* infos, LockBufs& lb, const std::vector<int>& sel, const int* perm,
                  std::vector<int>& failed)
{
    const frico_batch::Mesh& m = *um.m;
    const int64_t N = um.N;
    const int nz = (int)sel.size();
    const int64_t sa = N * (N + 1), sz = N * N;
    const int64_t acap = (N + 127) / 128 * 128, bcap = (N + 1 + 127) / 128 * 128;
    const size_t pa = (size_t)acap * 4 * HSPLIT_K, pb = (size_t)bcap * 4 * HSPLIT_K; // halves a system
    const size_t abytes = (pa * 2 + (size_t)acap * sizeof(int)) * nz, bbytes = (pb * 2 + (size_t)bcap * sizeof(int)) * nz;
    const size_t nw = lu_chain_words(N, 1);
    int rc;
    if ((rc = lb.Z.ensure((size_t)(sz * nz))) || (rc = lb.LU.ensure((size_t)(sa * nz))) ||
        (rc = lb.D.ensure((size_t)(N * nz))) || (rc = lb.B.ensure((size_t)(N * nz))) ||
        (rc = lb.X.ensure((size_t)(N * nz))) || (rc = lb.R.ensure((size_t)(N * nz))) ||
        (rc = lb.norms.ensure((size_t)nz)) || (rc = lb.words.ensure(nw * nz)) || (rc = lb.info.ensure((size_t)nz)) ||
        (rc = lb.split.ensure(4 * (abytes + bbytes))))
        return rc;
    BatchLu w;
    {
        unsigned char* p = lb.split.p;
        for (int i = 0; i < 4; i++) {
            w.ha[i] = reinterpret_cast<_Float16*>(p);
            w.hea[i] = reinterpret_cast<int*>(p + pa * 2 * nz);
            p += abytes;
            w.hb[i] = reinterpret_cast<_Float16*>(p);
            w.heb[i] = reinterpret_cast<int*>(p + pb * 2 * nz);
            p += bbytes;
        }
        w.hacap = acap;
        w.hbcap = bcap;
        w.spa = (int64_t)pa;
        w.sea = acap;
        w.spb = (int64_t)pb;
        w.seb = bcap;
        w.info = lb.info.p;
    }
    hipStream_t st = c->st;
    const unsigned gr = (unsigned)((N + 255) / 256);
    for (int z = 0; z < nz; z++) {
        const double freq = o.freqs[sel[z]];
        if ((rc = frico_fill(c, freq, FRICO_FILL_QUAD, FRICO_C64)) ||
            (rc = frico_rhs_deltagap(c, freq, m.feed.data(), (int)m.feed.size(), o.voltage.real(), o.voltage.imag(),
                                     um.applied ? um.signs.data() : nullptr)))
            return rc;
        if ((rc = launch_copy_bytes(c->Z32.p, lb.Z.p + z * sz, (size_t)sz * sizeof(c64), st)) ||
            (rc = launch_copy_bytes(c->B.p, lb.B.p + z * N, (size_t)N * sizeof(c128), st)))
            return rc;
        if (perm)
            launch_kernel(&gather_rows_c64, dim3(gr, (unsigned)N), dim3(256), 0, st, (const c64*)c->Z32.p, N,
                          lb.LU.p + z * sa, N, N, perm);
        else if ((rc = launch_copy_bytes(c->Z32.p, lb.LU.p + z * sa, (size_t)sz * sizeof(c64), st)))
            return rc;
        launch_kernel(&gather_c128_c64, dim3(gr, 1), dim3(256), 0, st, (const c128*)c->B.p, N, lb.LU.p + z * sa + sz, N,
                      N, perm);
        FRICO_HIP_CHECK(hipGetLastError());
    }
    if ((rc = lu_nopiv_batched(w, lb.LU.p, sa, N, N, 1, nz, st)))
        return rc;
    std::vector<int> zp((size_t)nz);
    std::vector<double> bn((size_t)nz), rn((size_t)nz), prev((size_t)nz, HUGE_VAL);
    if ((rc = launch_col_norm2(lb.B.p, N, N, nz, lb.norms.p, st)))
        return rc;
    FRICO_HIP_CHECK(hipMemcpyAsync(rn.data(), lb.norms.p, (size_t)nz * sizeof(double), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipMemcpyAsync(zp.data(), lb.info.p, (size_t)nz * sizeof(int), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipStreamSynchronize(st));
    for (int z = 0; z < nz; z++)
        bn[z] = std::sqrt(rn[z]);
    // 0: refining, 1: converged, 2: not converging
    std::vector<int> state((size_t)nz, 0);
    for (int z = 0; z < nz; z++)
        if (zp[z] > 0)
            state[z] = 2;
    const int64_t N16 = N * (int64_t)sizeof(c128);
    for (int it = 1; it <= o.maxit; it++) {
        if (it == 1) {
            // the carried column (L^{-1} P B): the first sweep is the backward half
            for (int z = 0; z < nz; z++)
                if ((rc = launch_copy_bytes(lb.LU.p + z * sa + sz, lb.D.p + z * N, (size_t)N * sizeof(c64), st)))
                    return rc;
            if ((rc = lu_chain_solve_batched(lb.LU.p, sa, N, N, lb.D.p, N, N, 1, nz, lb.words.p, false, st)) ||
                (rc = launch_zero_bytes(lb.X.p, (size_t)(N16 * nz), st)))
                return rc;
        } else {
            launch_kernel(&gather_c128_c64, dim3(gr, (unsigned)nz), dim3(256), 0, st, (const c128*)lb.R.p, N, lb.D.p, N, N,
                          perm);
            FRICO_HIP_CHECK(hipGetLastError());
            if ((rc = lu_chain_solve_batched(lb.LU.p, sa, N, N, lb.D.p, N, N, 1, nz, lb.words.p, true, st)))
                return rc;
        }
        if ((rc = launch_add_correction(lb.X.p, N, lb.D.p, N, N, nz, st)) ||
            (rc = launch_copy_bytes(lb.B.p, lb.R.p, (size_t)(N16 * nz), st)) ||
            (rc = gemv_sub_mixed_batched(N, 1, N, lb.Z.p, N, sz, lb.X.p, N, N, lb.R.p, N, N, nz, st)) ||
            (rc = launch_col_norm2(lb.R.p, N, N, nz, lb.norms.p, st)))
            return rc;
        FRICO_HIP_CHECK(hipMemcpyAsync(rn.data(), lb.norms.p, (size_t)nz * sizeof(double), hipMemcpyDeviceToHost, st));
        FRICO_HIP_CHECK(hipStreamSynchronize(st));
        bool active = false;
        for (int z = 0; z < nz; z++) {
            if (state[z] != 0)
                continue;
            const double r = std::sqrt(rn[z]), rel = bn[z] > 0 ? r / bn[z] : r;
            frico_solve_info& inf = infos[(size_t)mi * o.nfreq + sel[z]];
            inf.iterations = it;
            inf.max_rel_residual = std::isfinite(rel) ? rel : HUGE_VAL;
            if (rel <= o.tol) {
                state[z] = 1;
                inf.converged = 1;
            } else if (!(rel < 0.5 * prev[z])) {
                state[z] = 2; // stalled: factors that grew without interchanges
            } else {
                prev[z] = rel;
                active = true;
            }
        }
        if (!active)
            break;
    }
    std::vector<cd> X((size_t)(N * nz));
    FRICO_HIP_CHECK(hipMemcpyAsync(X.data(), lb.X.p, (size_t)(N16 * nz), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipStreamSynchronize(st));
    failed.clear();
    for (int z = 0; z < nz; z++) {
        const int f = sel[z];
        if (state[z] == 1) {
            const cd zz = port_zin(m, um.bf, um.signs, um.applied != 0, X.data() + z * N, N, o.voltage);
            zin[2 * ((size_t)mi * o.nfreq + f)] = zz.real();
            zin[2 * ((size_t)mi * o.nfreq + f) + 1] = zz.imag();
            b->unit_form[(size_t)mi * o.nfreq + f] = FRICO_SOLVE_C64_NP_IR;
        } else {
            infos[(size_t)mi * o.nfreq + f] = frico_solve_info{0, 0, -1.0};
            failed.push_back(f);
        }
    }
    return 0;
}

// One mesh at all of its frequencies in lockstep (lockstep_pass).  Frequencies whose unpivoted
// factors fail (a zero pivot, or a refinement that stalls: factors that grow without interchanges)
// are solved once more in lockstep with the row order of the pivoted LU of the first of them (the
// partial-pivoting order of one frequency keeps the factors of its neighbours bounded as well); what
// still fails goes to the pivoted LU one system at a time (run_unit).
int run_unit_lockstep(frico_batch* b, frico_ctx* c, int mi, const RunOpts& o, double* zin, frico_solve_info* infos,
                      LockBufs& lb)
{
    const frico_batch::Mesh& m = b->meshes[mi];
    int rc = frico_mesh_load(c, m.V.data(), (int64_t)(m.V.size() / 3), m.T.data(), m.tags.data(),
                             (int64_t)m.tags.size(), m.iface.empty() ? nullptr : m.iface.data(),
                             m.itags.empty() ? nullptr : m.itags.data(), (int64_t)m.itags.size(), m.degree);
    if (rc)
        return rc;
    UnitMesh um;
    um.m = &m;
    um.N = c->N;
    const int64_t N = um.N;
    if (!lockstep_fits(c, o, N))
        return run_unit(b, c, mi, o, zin, infos);
    um.signs.assign((size_t)N, 0.0);
    if ((rc = frico_reorient_deltagap(c, m.feed.data(), (int)m.feed.size(), 0, um.signs.data(), &um.applied)))
        return rc;
    um.bf.resize((size_t)N);
    if ((rc = frico_rwg_table(c, um.bf.data())))
        return rc;
    b->mesh_nb[mi] = 128;
    std::vector<int> all((size_t)o.nfreq), failed;
    for (int f = 0; f < o.nfreq; f++)
        all[f] = f;
    if ((rc = lockstep_pass(b, c, mi, um, o, zin, infos, lb, all, nullptr, failed)))
        return rc;
    if (failed.empty())
        return 0;
    // the pivoted LU of the first failed frequency, and its row order for the others
    const int f0 = failed[0];
    std::vector<char> one((size_t)o.nfreq, 0);
    one[f0] = 1;
    RunOpts op = o;
    op.form = FRICO_SOLVE_C64_IR;
    op.graphs = false;
    if ((rc = run_unit(b, c, mi, op, zin, infos, &one)))
        return rc;
    std::vector<int> rest(failed.begin() + 1, failed.end()), still;
    if (!rest.empty() && c->factored == FRICO_SOLVE_C64_IR) {
        std::vector<int> ipiv((size_t)N), perm((size_t)N);
        FRICO_HIP_CHECK(hipMemcpy(ipiv.data(), c->ipiv.p, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; i++)
            perm[i] = (int)i;
        for (int64_t i = 0; i < N; i++) {
            const int p = ipiv[i];
            if (p < 0 || p >= N)
                return set_error("batch: pivot %d of row %lld outside 0..%lld", p, (long long)i, (long long)N - 1), -1;
            std::swap(perm[i], perm[p]);
        }
        if ((rc = lb.perm.upload(perm.data(), (size_t)N, c->st)) ||
            (rc = lockstep_pass(b, c, mi, um, o, zin, infos, lb, rest, lb.perm.p, still)))
            return rc;
        rest.swap(still);
    }
    if (rest.empty())
        return 0;
    std::vector<char> mask((size_t)o.nfreq, 0);
    for (int f : rest)
        mask[f] = 1;
    return run_unit(b, c, mi, op, zin, infos, &mask);
}

} // namespace

extern "C" {

int frico_batch_create(int device, int workers, frico_batch** out)
{
    if (!out)
        return set_error("null output pointer"), -3;
    if (workers < 1)
        return set_error("workers must be >= 1"), -2;
    auto* b = new frico_batch;
    b->device = device;
    for (int w = 0; w < workers; w++) {
        frico_ctx* c = nullptr;
        const int rc = frico_ctx_create(device, &c);
        if (rc) {
            for (auto* x : b->ctx)
                frico_ctx_destroy(x);
            delete b;
            return rc;
        }
        c->lookahead = false;
        b->ctx.push_back(c);
    }
    b->cus = b->ctx[0]->max_groups;
    *out = b;
    return 0;
}

int frico_batch_destroy(frico_batch* b)
{
    if (!b)
        return 0;
    for (auto* c : b->ctx)
        frico_ctx_destroy(c);
    delete b;
    return 0;
}

int frico_batch_add_mesh(frico_batch* b, const double* vertices, int64_t nv, const int64_t* triangles,
                         const int32_t* tri_tags, int64_t nt, const int64_t* iface_edges, const int32_t* iface_tags,
                         int64_t nif, int degree, const int32_t* feed, int nfeed)
{
    if (!b)
        return set_error("null batch"), -1;
    if (!vertices || nv < 3 || !triangles || !tri_tags || nt < 1)
        return set_error("empty mesh"), -2;
    if (!feed || nfeed < 1)
        return set_error("no feed interface"), -11;
    frico_batch::Mesh m;
    m.V.assign(vertices, vertices + 3 * nv);
    m.T.assign(triangles, triangles + 3 * nt);
    m.tags.assign(tri_tags, tri_tags + nt);
    if (nif > 0 && iface_edges && iface_tags) {
        m.iface.assign(iface_edges, iface_edges + 2 * nif);
        m.itags.assign(iface_tags, iface_tags + nif);
    }
    m.feed.assign(feed, feed + nfeed);
    m.degree = degree;
    m.N = count_internal_edges(triangles, nt);
    b->meshes.push_back(std::move(m));
    return (int)b->meshes.size() - 1;
}

int frico_batch_num_meshes(const frico_batch* b) { return b ? (int)b->meshes.size() : -1; }

int frico_batch_workers_used(const frico_batch* b) { return b ? b->workers_used : -1; }

int frico_batch_antenna(frico_batch* b, const double* freqs, int nfreq, int fill_prec, int form, double tol, int maxit,
                        int flags, double v_re, double v_im, double* zin, frico_solve_info* info)
{
    if (!b)
        return set_error("null batch"), -1;
    if (!freqs || nfreq < 1)
        return set_error("no frequencies"), -2;
    if (fill_prec != FRICO_C128 && fill_prec != FRICO_C64)
        return set_error("unknown fill precision %d", fill_prec), -4;
    if (form != FRICO_SOLVE_C128 && form != FRICO_SOLVE_C64_IR && form != FRICO_SOLVE_C64_SYM_IR &&
        form != FRICO_SOLVE_C64_NP_IR)
        return set_error("unknown factorisation form %d", form), -5;
    if (form == FRICO_SOLVE_C128 && fill_prec != FRICO_C128)
        return set_error("the complex128 form needs a complex128 fill"), -4;
    // batches fill in quadrature mode, whose Z is not symmetric (the sub-triangle rule on the row
    // basis, src/emw_solver.cpp:67-78); run_context takes solve_symmetric only for approximate-mode
    // or symmetrised matrices (src/emw_solver.h:124-128)
    if (form == FRICO_SOLVE_C64_SYM_IR)
        return set_error("the symmetric form needs a symmetric Z; batched quadrature-mode fills are not"), -5;
    if (maxit < 1)
        return set_error("maxit must be >= 1"), -7;
    if (!zin || !info)
        return set_error("null output"), -11;
    const int nm = (int)b->meshes.size();
    if (nm == 0)
        return 0;
    RunOpts o{freqs, nfreq, fill_prec, form, tol, maxit, (flags & FRICO_BATCH_GRAPHS) != 0,
              (flags & FRICO_BATCH_AUTOTUNE) != 0, cd(v_re, v_im)};
    // longest processing time first
    std::vector<int> order((size_t)nm);
    for (int i = 0; i < nm; i++)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->meshes[a].N > b->meshes[c].N; });
    // co-residency of the panel launches of all workers
    int64_t nmax = b->meshes[order[0]].N;
    const int R = form == FRICO_SOLVE_C128 ? panel_rows<double>(64) : panel_rows<float>(64);
    const int64_t gmax = std::max<int64_t>(1, (nmax + R - 1) / R);
    // streams beyond the process's hardware queues share queues and serialise behind each other
    int hwq = 4; // HIP's default
    if (const char* q = std::getenv("GPU_MAX_HW_QUEUES"))
        hwq = std::max(1, std::atoi(q));
    const int W = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::min<int64_t>((int64_t)b->ctx.size(), hwq), b->cus / gmax));
    b->workers_used = W;
    b->mesh_nb.assign((size_t)nm, 0);
    b->unit_form.assign((size_t)nm * nfreq, -1);
    b->graph_units = 0;
    for (int i = 0; i < nm * nfreq; i++)
        info[i] = frico_solve_info{0, 0, -1.0};

    std::atomic<int> next{0};
    std::atomic<int> failed{0};
    std::mutex emu;
    int first_rc = 0;
    std::string first_msg;
    auto worker = [&](int w) {
        frico_ctx* c = b->ctx[w];
        LockBufs lb;
        if (hipSetDevice(b->device) != hipSuccess) {
            failed = 1;
            return;
        }
        for (;;) {
            if (failed.load())
                return;
            const int k = next.fetch_add(1);
            if (k >= nm)
                return;
            const int rc = lockstep_on() && o.form == FRICO_SOLVE_C64_NP_IR
                               ? run_unit_lockstep(b, c, order[k], o, zin, info, lb)
                               : run_unit(b, c, order[k], o, zin, info);
            if (rc) {
                std::lock_guard<std::mutex> lock(emu);
                if (!first_rc) {
                    first_rc = rc;
                    first_msg = std::string("mesh ") + std::to_string(order[k]) + ": " + frico_last_error();
                }
                failed = 1;
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 0; w < W; w++)
        th.emplace_back(worker, w);
    for (auto& t : th)
        t.join();
    if (first_rc)
        return set_error("%s", first_msg.c_str()), first_rc;
    if (failed)
        return set_error("a batch worker could not select the device"), FRICO_ERR_HIP;
    return 0;
}

int frico_batch_factored_forms(const frico_batch* b, int32_t* forms)
{
    if (!b || !forms)
        return -1;
    for (size_t i = 0; i < b->unit_form.size(); i++)
        forms[i] = b->unit_form[i];
    return (int)b->unit_form.size();
}

int frico_batch_block_sizes(const frico_batch* b, int32_t* nb)
{
    if (!b || !nb)
        return -1;
    for (size_t i = 0; i < b->mesh_nb.size(); i++)
        nb[i] = b->mesh_nb[i];
    return b->graph_units.load();
}

} // extern "C"
