// C ABI of the MI355X EFIE library: context, mesh ingestion, fill, RHS, factor/solve, fields.
#include "ctx.hpp"
#include "solver.hpp"

#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <list>
#include <mutex>
#include <tuple>

namespace {
thread_local std::string g_err;
}

namespace frico {

void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

GraphRecorder*& graph_recorder()
{
    thread_local GraphRecorder* r = nullptr;
    return r;
}

hipError_t allow_lds(const void* func, size_t bytes)
{
    static std::mutex mu;
    static std::vector<std::tuple<int, const void*, size_t>> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& d : done)
        if (std::get<0>(d) == dev && std::get<1>(d) == func && std::get<2>(d) >= bytes)
            return hipSuccess;
    e = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess)
        done.emplace_back(dev, func, bytes);
    return e;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return e == hipErrorOutOfMemory ? FRICO_ERR_NOMEM : FRICO_ERR_HIP;
}

} // namespace frico

using namespace frico;

namespace {

int check_ctx(frico_ctx* c)
{
    if (!c) {
        set_error("null context");
        return -1;
    }
    g_err.clear();
    return 0;
}

int set_dev(frico_ctx* c)
{
    FRICO_HIP_CHECK(hipSetDevice(c->device));
    return 0;
}

void phase_begin(frico_ctx* c, KernelTimer& t)
{
    if (c->profiling)
        t.begin(c->st);
}
void phase_end(frico_ctx* c, KernelTimer& t)
{
    if (c->profiling)
        t.end(c->st, 0.0);
}

int ensure_lu_work(frico_ctx* c)
{
    const int64_t npan = (c->N + 31) / 32 + 3; // panels at the smallest block width, two sub-panel lists
    const int gmax = std::min(c->max_groups, LU_MAX_GROUPS);
    int rc;
    if ((rc = c->llw.ensure(lu_exchange_words(gmax))))
        return rc;
    if ((rc = c->flags.ensure(2)))
        return rc;
    if ((rc = c->dinv.ensure((size_t)((c->N + 63) / 64) * 2 * 4096 * sizeof(c128))))
        return rc;
    if ((rc = c->perm.ensure((size_t)npan * lu_perm_stride())))
        return rc;
    const size_t nev = (size_t)(2 * npan + 1);
    while (c->lu_events.size() < nev) {
        hipEvent_t e;
        FRICO_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->lu_events.push_back(e);
    }
    return 0;
}

LuWork lu_work(frico_ctx* c)
{
    LuWork w;
    w.ll = c->llw.p;
    w.nll = c->llw.n;
    w.flags = c->flags.p;
    w.dinv = c->dinv.p;
    w.perm = c->perm.p;
    w.nperm = (int64_t)(c->perm.n / lu_perm_stride());
    if (c->lookahead) {
        w.side = c->side;
        w.events = c->lu_events.data();
        w.nevents = (int)c->lu_events.size();
    }
    w.max_groups = std::min(c->max_groups, LU_MAX_GROUPS);
    return w;
}

double kwave(double freq, double* inv_ksq, double* omega_out)
{
    const double omega = 2.0 * PI * freq;
    if (inv_ksq)
        *inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    if (omega_out)
        *omega_out = omega;
    return omega * std::sqrt(MU0 * EPS0);
}

} // namespace

extern "C" {

const char* frico_last_error(void) { return g_err.c_str(); }

int frico_ctx_create(int device, frico_ctx** out)
{
    g_err.clear();
    if (!out)
        return set_error("null output pointer"), -2;
    int ndev = 0;
    FRICO_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return set_error("device %d not present (%d devices)", device, ndev), -1;
    auto* c = new frico_ctx;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "stream create", __FILE__, __LINE__);
    }
    if (const char* la = std::getenv("FRICO_LU_LOOKAHEAD"))
        c->lookahead = std::atoi(la) != 0;
    if (const char* fs = std::getenv("FRICO_LU_FUSED_SOLVE"))
        c->fused_solve = std::atoi(fs) != 0;
    if (const char* au = std::getenv("FRICO_LU_AUGMENT"))
        c->augment = std::atoi(au) != 0;
    if (const char* sr = std::getenv("FRICO_STREAM_RESIDUAL"))
        c->stream_residual = std::atoi(sr) != 0;
    if (const char* rp = std::getenv("FRICO_RESID_PAIR"))
        c->resid_pair = std::atoi(rp) != 0;
    if (const char* sf = std::getenv("FRICO_SOLVE_F16"))
        c->solve_f16 = std::atoi(sf) != 0;
    if (const char* gf = std::getenv("FRICO_GEMM_F16"))
        c->gemm_f16 = std::atoi(gf) != 0;
    if (const char* ds = std::getenv("FRICO_LU_DIRECT_SOLVES"))
        c->direct_solves = std::atoi(ds) != 0;
    if (const char* fs = std::getenv("FRICO_LU_FUSED_STEP"))
        c->fused_step = std::atoi(fs) != 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->max_groups = prop.multiProcessorCount; // one panel workgroup per CU (LDS-limited)
    }
    // the device address of frico_fields_noalloc's scratch, resolved once per device
    if (int rc = noalloc_init(device)) {
        (void)hipStreamDestroy(c->st);
        (void)hipStreamDestroy(c->side);
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

int frico_ctx_destroy(frico_ctx* c)
{
    if (!c)
        return 0;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    (void)hipStreamSynchronize(c->side);
    for (auto e : c->lu_events)
        (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->side);
    (void)hipStreamDestroy(c->st);
    delete c;
    return 0;
}

int frico_ctx_synchronize(frico_ctx* c)
{
    if (check_ctx(c))
        return -1;
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    return 0;
}

int frico_ctx_set_profiling(frico_ctx* c, int enable)
{
    if (check_ctx(c))
        return -1;
    c->profiling = enable != 0;
    for (KernelTimer* t : {&c->t_fill, &c->t_rhs, &c->t_factor, &c->t_solve, &c->t_refine, &c->t_gemm, &c->t_tripair}) {
        t->reset();
        t->enabled = c->profiling;
    }
    return 0;
}

int frico_ctx_get_profile(frico_ctx* c, frico_profile* out)
{
    if (check_ctx(c))
        return -1;
    if (!out)
        return set_error("null output"), -2;
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    for (KernelTimer* t : {&c->t_fill, &c->t_rhs, &c->t_factor, &c->t_solve, &c->t_refine, &c->t_gemm, &c->t_tripair})
        t->collect();
    out->fill_ms = c->t_fill.total_ms;
    out->rhs_ms = c->t_rhs.total_ms;
    out->factor_ms = c->t_factor.total_ms;
    out->solve_ms = c->t_solve.total_ms;
    out->refine_ms = c->t_refine.total_ms;
    out->gemm_ms = c->t_gemm.total_ms;
    out->gemm_flops = c->t_gemm.total_work;
    out->gemm_launches = c->t_gemm.launches;
    out->gemm_bytes = c->t_gemm.total_bytes;
    out->tripair_ms = c->t_tripair.total_ms;
    out->tripair_pairs = c->t_tripair.total_work;
    out->tripair_launches = c->t_tripair.launches;
    return 0;
}

int frico_ctx_set_block(frico_ctx* c, int nb)
{
    if (check_ctx(c))
        return -1;
    if (nb != 0 && nb != 32 && nb != 64 && nb != 128)
        return set_error("block size must be 0 (automatic), 32, 64 or 128"), -2;
    c->nb = nb;
    return 0;
}

// Block width of the next factorisation: the set width, or (0) 128 from N = 6144 up and 64 below.
// At N = 12160 the 128-wide steps (two 64-column panel launches, K = 128 trailing products) took
// the complex64 factorisation from 85.1 to 84.3 ms and the complex128 one from 144.9 to 135.6.
static int block_width(const frico_ctx* c) { return c->nb ? c->nb : (c->N >= 6144 ? 128 : 64); }

int frico_mesh_load(frico_ctx* c, const double* vertices, int64_t nv, const int64_t* triangles, const int32_t* tri_tags,
                    int64_t nt, const int64_t* iface_edges, const int32_t* iface_tags, int64_t nif, int degree)
{
    if (check_ctx(c))
        return -1;
    if (!vertices || nv < 3)
        return set_error("need at least 3 vertices"), -2;
    if (!triangles || nt < 1)
        return set_error("need at least one triangle"), -4;
    if (nif < 0 || (nif > 0 && !iface_edges))
        return set_error("bad interface edge list"), -7;
    if (degree < 0 || degree > FRICO_DUNAVANT_MAX_DEGREE)
        return set_error("Dunavant quadrature: degree %d out of range 0..%d", degree, FRICO_DUNAVANT_MAX_DEGREE), -10;
    if (nt > (int64_t)1 << 30)
        return set_error("too many triangles"), FRICO_ERR_LIMIT;
    int rc = set_dev(c);
    if (rc)
        return rc;
    HostMesh hm;
    hm.nv = nv;
    hm.nt = nt;
    hm.V.resize(nv);
    for (int64_t i = 0; i < nv; i++)
        hm.V[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
    hm.T.resize(nt);
    hm.tag.resize(nt);
    std::vector<std::pair<int64_t, int64_t>> all;
    all.reserve(3 * nt);
    auto mk = [](int64_t a, int64_t b) { return std::make_pair(std::min(a, b), std::max(a, b)); };
    for (int64_t t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++) {
            const int64_t v = triangles[3 * t + k];
            if (v < 0 || v >= nv)
                return set_error("triangle %lld references vertex %lld", (long long)t, (long long)v), -3;
            hm.T[t][k] = v;
        }
        hm.tag[t] = tri_tags ? tri_tags[t] : 0;
        all.push_back(mk(hm.T[t][0], hm.T[t][1]));
        all.push_back(mk(hm.T[t][1], hm.T[t][2]));
        all.push_back(mk(hm.T[t][2], hm.T[t][0]));
    }
    // sorted unique edge list (src/input_gmsh.cpp:136-140)
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    hm.edges = std::move(all);
    auto offset = [&](std::pair<int64_t, int64_t> e) -> int64_t {
        auto it = std::lower_bound(hm.edges.begin(), hm.edges.end(), e);
        return (it != hm.edges.end() && *it == e) ? (int64_t)(it - hm.edges.begin()) : -1;
    };
    std::vector<Neighbours> nb(hm.edges.size());
    std::vector<std::pair<int64_t, int>> ifl;
    for (int64_t b = 0; b < nif; b++) {
        const int64_t o = offset(mk(iface_edges[2 * b], iface_edges[2 * b + 1]));
        if (o < 0)
            return set_error("lonely edge: 1D entity %d edge (%lld, %lld) is not on any triangle",
                             iface_tags ? iface_tags[b] : 0, (long long)iface_edges[2 * b], (long long)iface_edges[2 * b + 1]),
                   FRICO_ERR_MESH;
        ifl.push_back({o, iface_tags ? iface_tags[b] : 0});
    }
    // T- = first (lower) triangle, T+ = second; a third one is an error (src/input_gmsh.cpp:216-262)
    for (int64_t t = 0; t < nt; t++) {
        const auto& tr = hm.T[t];
        const std::pair<int64_t, int64_t> es[3] = {mk(tr[0], tr[1]), mk(tr[1], tr[2]), mk(tr[2], tr[0])};
        for (int ie = 0; ie < 3; ie++) {
            Neighbours& en = nb[offset(es[ie])];
            if (en.tp >= 0)
                return set_error("edge shared by more than two triangles (tags %d, %d and %d): no RWG basis",
                                 hm.tag[en.tm], hm.tag[en.tp], hm.tag[t]),
                       FRICO_ERR_MESH;
            if (en.tm >= 0) {
                en.tp = t;
                en.lp = ie;
            } else {
                en.tm = t;
                en.lm = ie;
            }
        }
    }
    for (auto& [o, tg] : ifl)
        if (nb[o].tp >= 0) {
            nb[o].has_iface = true;
            nb[o].iface = tg;
        }
    auto tri_area = [&](int64_t t) {
        const auto& tr = hm.T[t];
        return 0.5 * norm(cross(hm.V[tr[1]] - hm.V[tr[0]], hm.V[tr[2]] - hm.V[tr[1]]));
    };
    const int evmap[3] = {2, 0, 1};
    for (size_t ie = 0; ie < nb.size(); ie++) {
        const Neighbours& en = nb[ie];
        if (en.tp < 0)
            continue;
        frico_rwg b{};
        b.itminus = en.tm;
        b.itplus = en.tp;
        b.loc_eminus = en.lm;
        b.loc_eplus = en.lp;
        b.edge_index = (int64_t)ie;
        b.has_interface = en.has_iface;
        b.interface_tag = en.iface;
        const vec3 d = hm.V[hm.edges[ie].second] - hm.V[hm.edges[ie].first];
        b.length = std::hypot(d.x, d.y);
        b.Aminus = tri_area(en.tm);
        b.Aplus = tri_area(en.tp);
        const vec3 pm = hm.V[hm.T[en.tm][evmap[en.lm]]], pp = hm.V[hm.T[en.tp][evmap[en.lp]]];
        b.pminus[0] = pm.x;
        b.pminus[1] = pm.y;
        b.pminus[2] = pm.z;
        b.pplus[0] = pp.x;
        b.pplus[1] = pp.y;
        b.pplus[2] = pp.z;
        hm.bf.push_back(b);
    }
    const int64_t N = (int64_t)hm.bf.size();
    if (N < 1)
        return set_error("mesh has no internal edge: no unknowns"), FRICO_ERR_MESH;
    if (N > 65535)
        return set_error("N = %lld exceeds the 65535-unknown launch limit", (long long)N), FRICO_ERR_LIMIT;

    // per-triangle data
    hm.degree = degree;
    hm.q = frico_dunavant_npoints[degree == 0 ? 0 : degree - 1];
    hm.cen.resize(3 * nt);
    hm.area.resize(nt);
    hm.sqrtA.resize(nt);
    hm.alpha.resize(9 * nt);
    hm.tri_rwg.assign(3 * nt, -1);
    hm.tri_sgn.assign(3 * nt, 0.0);
    hm.qreg.clear();
    hm.qsub.clear();
    hm.qreg.reserve(nt * hm.q);
    hm.qsub.reserve(nt * 3 * hm.q);
    std::vector<QP> qp;
    for (int64_t t = 0; t < nt; t++) {
        const auto& tr = hm.T[t];
        const vec3 p0 = hm.V[tr[0]], p1 = hm.V[tr[1]], p2 = hm.V[tr[2]];
        const vec3 c = (p0 + p1 + p2) / 3.0; // barycenter (src/geom_mesh.cpp:64-69)
        hm.cen[3 * t] = c.x;
        hm.cen[3 * t + 1] = c.y;
        hm.cen[3 * t + 2] = c.z;
        hm.area[t] = tri_area(t);
        hm.sqrtA[t] = std::sqrt(hm.area[t]);
        for (int m = 0; m < 3; m++) {
            const vec3 a = c - hm.V[tr[evmap[m]]];
            hm.alpha[9 * t + 3 * m] = a.x;
            hm.alpha[9 * t + 3 * m + 1] = a.y;
            hm.alpha[9 * t + 3 * m + 2] = a.z;
        }
        qp.clear();
        dunavant(degree, p0, p1, p2, qp);
        for (auto& x : qp)
            hm.qreg.push_back(double4{x.p.x - c.x, x.p.y - c.y, x.p.z - c.z, x.w});
        qp.clear();
        // sub-triangle rule (src/quadratures.cpp:2685-2702)
        dunavant(degree, p0, p1, c, qp);
        dunavant(degree, p1, p2, c, qp);
        dunavant(degree, p2, p0, c, qp);
        for (auto& x : qp)
            hm.qsub.push_back(double4{x.p.x - c.x, x.p.y - c.y, x.p.z - c.z, x.w});
    }
    hm.rwg_t.resize(2 * N);
    hm.rwg_le.resize(2 * N);
    hm.rwg_len.resize(N);
    hm.rwg_Am.resize(N);
    hm.rwg_Ap.resize(N);
    for (int64_t n = 0; n < N; n++) {
        const frico_rwg& b = hm.bf[n];
        hm.rwg_t[2 * n] = (int)b.itminus;
        hm.rwg_t[2 * n + 1] = (int)b.itplus;
        hm.rwg_le[2 * n] = b.loc_eminus;
        hm.rwg_le[2 * n + 1] = b.loc_eplus;
        hm.rwg_len[n] = b.length;
        hm.rwg_Am[n] = b.Aminus;
        hm.rwg_Ap[n] = b.Aplus;
        hm.tri_rwg[3 * b.itminus + b.loc_eminus] = (int)n;
        hm.tri_sgn[3 * b.itminus + b.loc_eminus] = -1.0;
        hm.tri_rwg[3 * b.itplus + b.loc_eplus] = (int)n;
        hm.tri_sgn[3 * b.itplus + b.loc_eplus] = 1.0;
    }
    // pairs (i, j) whose row function j has a triangle of the column function i
    hm.near.clear();
    for (int64_t i = 0; i < N; i++) {
        int js[6], nj = 0;
        for (int s = 0; s < 2; s++) {
            const int t = hm.rwg_t[2 * i + s];
            for (int e = 0; e < 3; e++) {
                const int j = hm.tri_rwg[3 * t + e];
                if (j < 0)
                    continue;
                bool dup = false;
                for (int x = 0; x < nj; x++)
                    dup |= js[x] == j;
                if (!dup)
                    js[nj++] = j;
            }
        }
        std::sort(js, js + nj);
        for (int x = 0; x < nj; x++)
            hm.near.push_back(int2{(int)i, js[x]});
    }

    // adjoint data: triangle vertices and barycentric weights of every rule point
    hm.tri_v.resize(3 * nt);
    for (int64_t t = 0; t < nt; t++)
        for (int kk = 0; kk < 3; kk++)
            hm.tri_v[3 * t + kk] = (int)hm.T[t][kk];
    {
        std::vector<std::array<double, 3>> bc;
        dunavant_bary(degree, bc);
        hm.qbeta.clear();
        for (auto& b : bc)
            hm.qbeta.insert(hm.qbeta.end(), b.begin(), b.end());
        hm.qsub_beta.clear();
        const int corners[3][2] = {{0, 1}, {1, 2}, {2, 0}};
        for (int s = 0; s < 3; s++)
            for (auto& b : bc)
                for (int kk = 0; kk < 3; kk++)
                    hm.qsub_beta.push_back((kk == corners[s][0] ? b[0] : 0.0) + (kk == corners[s][1] ? b[1] : 0.0) +
                                           b[2] / 3.0);
    }

    // upload
    hipStream_t st = c->st;
    if ((rc = c->d_cen.upload(hm.cen.data(), hm.cen.size(), st)) || (rc = c->d_area.upload(hm.area.data(), hm.area.size(), st)) ||
        (rc = c->d_sqrtA.upload(hm.sqrtA.data(), hm.sqrtA.size(), st)) ||
        (rc = c->d_alpha.upload(hm.alpha.data(), hm.alpha.size(), st)) ||
        (rc = c->d_tri_sgn.upload(hm.tri_sgn.data(), hm.tri_sgn.size(), st)) ||
        (rc = c->d_rwg_len.upload(hm.rwg_len.data(), hm.rwg_len.size(), st)) ||
        (rc = c->d_rwg_Am.upload(hm.rwg_Am.data(), hm.rwg_Am.size(), st)) ||
        (rc = c->d_rwg_Ap.upload(hm.rwg_Ap.data(), hm.rwg_Ap.size(), st)) ||
        (rc = c->d_qreg.upload(hm.qreg.data(), hm.qreg.size(), st)) ||
        (rc = c->d_qsub.upload(hm.qsub.data(), hm.qsub.size(), st)) ||
        (rc = c->d_tri_rwg.upload(hm.tri_rwg.data(), hm.tri_rwg.size(), st)) ||
        (rc = c->d_rwg_t.upload(hm.rwg_t.data(), hm.rwg_t.size(), st)) ||
        (rc = c->d_rwg_le.upload(hm.rwg_le.data(), hm.rwg_le.size(), st)) ||
        (rc = c->d_near.upload(hm.near.data(), hm.near.size(), st)) ||
        (rc = c->d_tri_v.upload(hm.tri_v.data(), hm.tri_v.size(), st)) ||
        (rc = c->d_qbeta.upload(hm.qbeta.data(), hm.qbeta.size(), st)) ||
        (rc = c->d_qsub_beta.upload(hm.qsub_beta.data(), hm.qsub_beta.size(), st)))
        return rc;
    FRICO_HIP_CHECK(hipStreamSynchronize(st));
    c->hm = std::move(hm);
    c->N = N;
    c->have_mesh = true;
    DevMesh& m = c->dm;
    m.T = (int)nt;
    m.N = (int)N;
    m.q = c->hm.q;
    m.cen = c->d_cen.p;
    m.area = c->d_area.p;
    m.sqrtA = c->d_sqrtA.p;
    m.alpha = c->d_alpha.p;
    m.qreg = c->d_qreg.p;
    m.qsub = c->d_qsub.p;
    m.tri_rwg = c->d_tri_rwg.p;
    m.tri_sgn = c->d_tri_sgn.p;
    m.rwg_t = c->d_rwg_t.p;
    m.rwg_le = c->d_rwg_le.p;
    m.rwg_len = c->d_rwg_len.p;
    m.rwg_Am = c->d_rwg_Am.p;
    m.rwg_Ap = c->d_rwg_Ap.p;
    c->am.tri_v = c->d_tri_v.p;
    c->am.qbeta = c->d_qbeta.p;
    c->am.qsub_beta = c->d_qsub_beta.p;
    c->zprec = -1;
    c->factored = -1;
    c->nrhs = 0;
    c->have_x = false;
    c->adj = frico_ctx::AdjointState{};
    return 0;
}

int64_t frico_num_unknowns(const frico_ctx* c) { return c ? c->N : -1; }
int64_t frico_num_triangles(const frico_ctx* c) { return c && c->have_mesh ? c->hm.nt : -1; }
int64_t frico_num_rhs(const frico_ctx* c) { return c ? c->nrhs : -1; }

int frico_rwg_table(const frico_ctx* c, frico_rwg* out)
{
    if (!c || !c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!out)
        return set_error("null output"), -2;
    std::memcpy(out, c->hm.bf.data(), c->hm.bf.size() * sizeof(frico_rwg));
    return 0;
}

int frico_reorient_deltagap(frico_ctx* c, const int32_t* ifaces, int nif, int force, double* signs, int* applied)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!signs || !applied)
        return set_error("null output"), -5;
    const HostMesh& hm = c->hm;
    for (int64_t n = 0; n < c->N; n++)
        signs[n] = 1.0;
    *applied = 0;
    // gap functions and whether T+ does not come after T- in surface-tag order
    bool fix = false;
    std::list<std::pair<std::pair<int64_t, int64_t>, int64_t>> edges;
    for (int64_t n = 0; n < c->N; n++) {
        const frico_rwg& b = hm.bf[n];
        if (!b.has_interface)
            continue;
        for (int f = 0; f < nif; f++)
            if (b.interface_tag == ifaces[f]) {
                if (hm.tag[b.itplus] <= hm.tag[b.itminus])
                    fix = true;
                edges.push_back({hm.edges[b.edge_index], n});
            }
    }
    if (!(fix || force))
        return 0;
    if (edges.empty())
        return set_error("no edges identified for the delta-gap"), -2;
    // chain the gap edges end to end (src/emw_solver.cpp:271-315)
    struct Seg {
        int64_t a, b, n;
    };
    std::list<Seg> chain;
    chain.push_back({edges.front().first.first, edges.front().first.second, edges.front().second});
    edges.pop_front();
    while (!edges.empty()) {
        const size_t before = edges.size();
        for (auto it = edges.begin(); it != edges.end();) {
            const Seg f = chain.front(), bk = chain.back();
            const int64_t e0 = it->first.first, e1 = it->first.second, n = it->second;
            if (e0 == f.a)
                chain.push_front({e1, e0, n});
            else if (e1 == f.a)
                chain.push_front({e0, e1, n});
            else if (e0 == bk.b)
                chain.push_back({e0, e1, n});
            else if (e1 == bk.b)
                chain.push_back({e1, e0, n});
            else {
                ++it;
                continue;
            }
            it = edges.erase(it);
        }
        if (edges.size() == before)
            return set_error("delta-gap edges are not topologically connected"), -2;
    }
    auto tri_bar = [&](int64_t t) {
        const auto& tr = hm.T[t];
        return (hm.V[tr[0]] + hm.V[tr[1]] + hm.V[tr[2]]) / 3.0;
    };
    auto edge_bar = [&](int64_t e) { return (hm.V[hm.edges[e].first] + hm.V[hm.edges[e].second]) / 2.0; };
    const frico_rwg& b0 = hm.bf[chain.front().n];
    vec3 dir = tri_bar(b0.itplus) - edge_bar(b0.edge_index);
    double sign = 1;
    for (const Seg& s : chain) {
        const frico_rwg& b = hm.bf[s.n];
        const vec3 cur = tri_bar(b.itplus) - edge_bar(b.edge_index);
        if (dot(dir, cur) < 0)
            sign *= -1;
        signs[s.n] = sign;
        dir = cur;
    }
    *applied = 1;
    return 0;
}

int frico_fill(frico_ctx* c, double freq, int mode, int prec)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!(freq > 0))
        return set_error("frequency must be positive"), -2;
    if (mode != FRICO_FILL_QUAD && mode != FRICO_FILL_APPROX)
        return set_error("unknown fill mode %d", mode), -3;
    if (prec != FRICO_C128 && prec != FRICO_C64)
        return set_error("unknown precision %d", prec), -4;
    int rc = set_dev(c);
    if (rc)
        return rc;
    const int64_t N = c->N, T = c->hm.nt;
    void* Z = nullptr;
    if (prec == FRICO_C128) {
        if ((rc = c->Z64.ensure((size_t)(N * N))))
            return rc;
        Z = c->Z64.p;
    } else {
        if ((rc = c->Z32.ensure((size_t)(N * N))))
            return rc;
        Z = c->Z32.p;
    }
    if (mode == FRICO_FILL_QUAD) {
        const size_t csz = (size_t)9 * T * T * (prec == FRICO_C128 ? sizeof(c128) : sizeof(c64));
        if ((rc = c->Cbuf.ensure(csz)))
            return rc;
    }
    double inv_ksq;
    const double k = kwave(freq, &inv_ksq, nullptr);
    phase_begin(c, c->t_fill);
    if (mode == FRICO_FILL_QUAD)
        rc = launch_fill_quad(c->dm, k, inv_ksq, prec, c->Cbuf.p, Z, N, c->d_near.p, (int64_t)c->hm.near.size(), c->st,
                              c->profiling ? &c->t_tripair : nullptr);
    else
        rc = launch_fill_approx(c->dm, k, inv_ksq, prec, Z, N, c->st);
    phase_end(c, c->t_fill);
    if (rc)
        return rc;
    c->zprec = prec;
    c->z64_factored = false;
    c->factored = -1;
    return 0;
}

int frico_symmetrize(frico_ctx* c)
{
    if (check_ctx(c))
        return -1;
    if (c->zprec != FRICO_C128 || c->z64_factored)
        return set_error("symmetrize needs an unfactored complex128 Z"), FRICO_ERR_STATE;
    return launch_symmetrize(c->Z64.p, c->N, c->N, c->st);
}

int frico_matrix_get(frico_ctx* c, double* Zh, int64_t ldz)
{
    if (check_ctx(c))
        return -1;
    if (c->zprec < 0)
        return set_error("no matrix"), FRICO_ERR_STATE;
    if (!Zh || ldz < c->N)
        return set_error("bad output / leading dimension"), -3;
    const int64_t N = c->N;
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    if (c->zprec == FRICO_C128) {
        FRICO_HIP_CHECK(hipMemcpy2D(Zh, ldz * sizeof(c128), c->Z64.p, N * sizeof(c128), N * sizeof(c128), N,
                                    hipMemcpyDeviceToHost));
    } else {
        std::vector<c64> tmp((size_t)(N * N));
        FRICO_HIP_CHECK(hipMemcpy(tmp.data(), c->Z32.p, tmp.size() * sizeof(c64), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; i++)
            for (int64_t j = 0; j < N; j++) {
                Zh[2 * (j + i * ldz)] = tmp[j + i * N].re;
                Zh[2 * (j + i * ldz) + 1] = tmp[j + i * N].im;
            }
    }
    return 0;
}

int frico_matrix_set(frico_ctx* c, const double* Zh, int64_t n, int64_t ldz)
{
    if (check_ctx(c))
        return -1;
    if (c->have_mesh && n != c->N)
        return set_error("n must equal the number of unknowns of the loaded mesh"), -3;
    if (!Zh || n < 1 || ldz < n)
        return set_error("bad matrix arguments"), -2;
    int rc = set_dev(c);
    if (rc)
        return rc;
    if (!c->have_mesh)
        c->N = n;
    if ((rc = c->Z64.ensure((size_t)(n * n))))
        return rc;
    FRICO_HIP_CHECK(hipMemcpy2DAsync(c->Z64.p, n * sizeof(c128), Zh, ldz * sizeof(c128), n * sizeof(c128), n,
                                     hipMemcpyHostToDevice, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    c->zprec = FRICO_C128;
    c->z64_factored = false;
    c->factored = -1;
    return 0;
}

static int rhs_alloc(frico_ctx* c, int64_t nrhs)
{
    int rc;
    const size_t sz = (size_t)(c->N * nrhs);
    if ((rc = c->B.ensure(sz)) || (rc = c->X.ensure(sz)))
        return rc;
    c->nrhs = nrhs;
    c->have_x = false;
    c->rhs_version++;
    return 0;
}

int frico_rhs_set(frico_ctx* c, const double* Bh, int64_t ldb, int64_t nrhs)
{
    if (check_ctx(c))
        return -1;
    if (c->N < 1)
        return set_error("no system size"), FRICO_ERR_STATE;
    if (!Bh || ldb < c->N || nrhs < 1)
        return set_error("bad rhs arguments"), -2;
    int rc = set_dev(c);
    if (rc || (rc = rhs_alloc(c, nrhs)))
        return rc;
    FRICO_HIP_CHECK(hipMemcpy2DAsync(c->B.p, c->N * sizeof(c128), Bh, ldb * sizeof(c128), c->N * sizeof(c128), nrhs,
                                     hipMemcpyHostToDevice, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    return 0;
}

int frico_rhs_get(frico_ctx* c, double* Bh, int64_t ldb)
{
    if (check_ctx(c))
        return -1;
    if (c->nrhs < 1)
        return set_error("no rhs"), FRICO_ERR_STATE;
    if (!Bh || ldb < c->N)
        return set_error("bad output"), -2;
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    FRICO_HIP_CHECK(hipMemcpy2D(Bh, ldb * sizeof(c128), c->B.p, c->N * sizeof(c128), c->N * sizeof(c128), c->nrhs,
                                hipMemcpyDeviceToHost));
    return 0;
}

int frico_rhs_deltagap(frico_ctx* c, double freq, const int32_t* ifaces, int nif, double v_re, double v_im,
                       const double* signs)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!(freq > 0))
        return set_error("frequency must be positive"), -2;
    if (nif < 0 || (nif > 0 && !ifaces))
        return set_error("bad interface list"), -3;
    int rc = set_dev(c);
    if (rc || (rc = rhs_alloc(c, 1)))
        return rc;
    const double omega = 2.0 * PI * freq;
    std::vector<c128> V((size_t)c->N, c128{0, 0});
    for (int64_t n = 0; n < c->N; n++) {
        const frico_rwg& b = c->hm.bf[n];
        if (!b.has_interface)
            continue;
        for (int f = 0; f < nif; f++)
            if (b.interface_tag == ifaces[f]) {
                // v = (0, -l/(omega mu0)); V += v * voltage * sign
                const double vi = -b.length / (omega * MU0);
                const double s = signs ? signs[n] : 1.0;
                V[n].re += (-vi * v_im) * s;
                V[n].im += (vi * v_re) * s;
            }
    }
    phase_begin(c, c->t_rhs);
    FRICO_HIP_CHECK(hipMemcpyAsync(c->B.p, V.data(), V.size() * sizeof(c128), hipMemcpyHostToDevice, c->st));
    phase_end(c, c->t_rhs);
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    return 0;
}

int frico_rhs_planewave(frico_ctx* c, double freq, int64_t nang, const double* theta, const double* phi,
                        const double* eth, const double* eph)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!(freq > 0))
        return set_error("frequency must be positive"), -2;
    if (nang < 1 || nang > 65535 || !theta || !phi || !eth || !eph)
        return set_error("bad incidence arrays"), -3;
    int rc = set_dev(c);
    if (rc || (rc = rhs_alloc(c, nang)))
        return rc;
    std::vector<PlaneWaveDev> w((size_t)nang);
    for (int64_t a = 0; a < nang; a++) {
        const double th = theta[a], ph = phi[a];
        const double dir[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
        const double tht[3] = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), -std::sin(th)};
        const double pht[3] = {-std::sin(ph), std::cos(ph), 0.0};
        for (int d = 0; d < 3; d++) {
            w[a].dir[d] = dir[d];
            w[a].e0r[d] = eth[2 * a] * tht[d] + eph[2 * a] * pht[d];
            w[a].e0i[d] = eth[2 * a + 1] * tht[d] + eph[2 * a + 1] * pht[d];
        }
    }
    if ((rc = c->waves.upload(w.data(), w.size(), c->st)))
        return rc;
    double omega;
    const double k = kwave(freq, nullptr, &omega);
    phase_begin(c, c->t_rhs);
    rc = launch_rhs_planewave(c->dm, k, omega * MU0, c->waves.p, (int)nang, c->B.p, c->N, c->st);
    phase_end(c, c->t_rhs);
    return rc;
}

static int read_lu_flags(frico_ctx* c)
{
    int h[2] = {0, 0};
    FRICO_HIP_CHECK(hipMemcpyAsync(h, c->flags.p, sizeof h, hipMemcpyDeviceToHost, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    if (h[0])
        return set_error("LU panel barrier exceeded its wait limit"), FRICO_ERR_SYNC;
    if (h[1] > 0)
        set_error("singular matrix: zero pivot at column %d", h[1]);
    return h[1];
}

} // extern "C"

namespace frico {

int factor_prepare(frico_ctx* c, int form, LuWork& wk, bool lookahead)
{
    if (c->zprec < 0)
        return set_error("no matrix to factor"), FRICO_ERR_STATE;
    if (form != FRICO_SOLVE_C128 && form != FRICO_SOLVE_C64_IR && form != FRICO_SOLVE_C64_SYM_IR &&
        form != FRICO_SOLVE_C64_NP_IR)
        return set_error("unknown factorisation form %d", form), -2;
    if (form == FRICO_SOLVE_C128 && (c->zprec != FRICO_C128 || c->z64_factored))
        return set_error("complex128 factorisation needs an unfactored complex128 Z"), FRICO_ERR_STATE;
    if (form != FRICO_SOLVE_C128 && c->zprec == FRICO_C128 && c->z64_factored)
        return set_error("Z was overwritten by a complex128 factorisation"), FRICO_ERR_STATE;
    int rc = set_dev(c);
    if (rc || (rc = ensure_lu_work(c)) || (rc = c->ipiv.ensure((size_t)c->N)))
        return rc;
    const int64_t N = c->N;
    // the complex64 factorisation carries the current right-hand sides along (forward elimination
    // inside the trailing products); the first refinement sweep then only substitutes backward
    c->aug_cols = (form != FRICO_SOLVE_C128 && c->augment && c->nrhs > 0) ? c->nrhs : 0;
    c->aug_valid = false;
    if (form != FRICO_SOLVE_C128 && (rc = c->LU32.ensure((size_t)(N * (N + c->aug_cols)))))
        return rc;
    c->factored = -1; // the factors are about to be overwritten
    c->nb_factored = block_width(c);
    if (form == FRICO_SOLVE_C64_NP_IR && (rc = c->tbuf.ensure((size_t)(LU_NB_MAX * N))))
        return rc;
    wk = lu_work(c);
    if (form == FRICO_SOLVE_C64_NP_IR) {
        wk.tbuf = c->tbuf.p;
        wk.ntbuf = (int64_t)c->tbuf.n;
        wk.direct_solves = c->direct_solves;
        wk.fused_step = c->fused_step;
        if (c->gemm_f16) {
            // L21 of four steps and U12 of two products a stream, padded to 128-row tiles
            const int64_t acap = (N + 127) / 128 * 128, bcap = (N + c->aug_cols + 127) / 128 * 128;
            const size_t abytes = split_f16_bytes(acap), bbytes = split_f16_bytes(bcap);
            if ((rc = c->hsplit.ensure(4 * (abytes + bbytes))))
                return rc;
            unsigned char* p = c->hsplit.p;
            wk.nsplit = 4;
            for (int i = 0; i < 4; i++) {
                wk.ha[i] = reinterpret_cast<_Float16*>(p);
                wk.hea[i] = reinterpret_cast<int*>(p + (size_t)acap * 4 * HSPLIT_K * 2);
                p += abytes;
                wk.hb[i] = reinterpret_cast<_Float16*>(p);
                wk.heb[i] = reinterpret_cast<int*>(p + (size_t)bcap * 4 * HSPLIT_K * 2);
                p += bbytes;
            }
            wk.hacap = acap;
            wk.hbcap = bcap;
        }
    }
    if (!lookahead) {
        wk.side = nullptr;
        wk.events = nullptr;
        wk.nevents = 0;
    }
    // W products for the fused solves: one more N x N matrix of the factor's type, taken only
    // when it fits with room to spare
    c->wv_ok = false;
    const bool blocked = form == FRICO_SOLVE_C64_NP_IR && lu_blocked_solves();
    if (c->fused_solve && c->nb_factored % 64 == 0 && N > 64 && !blocked) {
        const size_t bytes = (size_t)(N * N) * (form == FRICO_SOLVE_C128 ? sizeof(c128) : sizeof(c64));
        size_t fr = 0, tot = 0;
        if (c->wv.n >= bytes || (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > bytes + c->wv.n + ((size_t)2 << 30))) {
            if ((rc = c->wv.ensure(bytes)))
                return rc;
            wk.wv = c->wv.p;
        }
    }
    c->wv_ok = wk.wv != nullptr;
    return 0;
}

int factor_enqueue(frico_ctx* c, int form, LuWork& wk)
{
    const int64_t N = c->N;
    int rc;
    c->sexp_ok = false;
    phase_begin(c, c->t_factor);
    if (form == FRICO_SOLVE_C128) {
        rc = lu_factor<double>(wk, c->Z64.p, N, N, c->ipiv.p, c->nb_factored, c->st, c->profiling ? &c->t_gemm : nullptr);
    } else {
        if (c->zprec == FRICO_C128)
            rc = launch_convert_c128_c64(c->Z64.p, N, c->LU32.p, N, N, N, c->st);
        else
            rc = launch_copy_bytes(c->Z32.p, c->LU32.p, (size_t)(N * N) * sizeof(c64), c->st);
        if (!rc && c->aug_cols > 0)
            rc = launch_convert_c128_c64(c->B.p, N, c->LU32.p + N * N, N, N, c->aug_cols, c->st);
        KernelTimer* kt = c->profiling ? &c->t_gemm : nullptr;
        if (!rc && form == FRICO_SOLVE_C64_SYM_IR)
            rc = ldlt_factor<float>(wk, c->LU32.p, N, N, c->ipiv.p, c->nb_factored, c->st, kt, c->aug_cols);
        else if (!rc && form == FRICO_SOLVE_C64_NP_IR)
            rc = lu_nopiv_factor<float>(wk, c->LU32.p, N, N, c->ipiv.p, c->nb_factored, c->st, kt, c->aug_cols);
        else if (!rc)
            rc = lu_factor<float>(wk, c->LU32.p, N, N, c->ipiv.p, c->nb_factored, c->st, kt, c->aug_cols);
        // factors without interchanges are solved blockwise: their row exponents per block of the
        // solves' columns, for the solves' split products
        if (!rc && c->solve_f16 && lu_blocked_solves() &&
            (form == FRICO_SOLVE_C64_NP_IR || form == FRICO_SOLVE_C64_SYM_IR)) {
            const int W = lu_solve_block_rows();
            const int64_t nblk = (N + W - 1) / W;
            if (!(rc = c->sexp.ensure((size_t)(nblk * N))) &&
                !(rc = launch_rowexp_blocks(c->LU32.p, N, N, W, c->sexp.p, N, c->st)))
                c->sexp_ok = true;
        }
    }
    phase_end(c, c->t_factor);
    return rc;
}

int factor_finish(frico_ctx* c, int form, const LuWork& wk)
{
    if (form == FRICO_SOLVE_C128)
        c->z64_factored = true;
    c->factored = form;
    c->aug_valid = c->aug_cols > 0;
    c->aug_version = c->rhs_version;
    c->wv_ok = wk.wv != nullptr;
    return read_lu_flags(c);
}

int col_norms(frico_ctx* c, const c128* X, int64_t cols, std::vector<double>& out)
{
    int rc;
    if ((rc = c->norms.ensure((size_t)cols)))
        return rc;
    if ((rc = launch_col_norm2(X, c->N, c->N, cols, c->norms.p, c->st)))
        return rc;
    out.resize((size_t)cols);
    FRICO_HIP_CHECK(hipMemcpyAsync(out.data(), c->norms.p, (size_t)cols * sizeof(double), hipMemcpyDeviceToHost, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    for (auto& v : out)
        v = std::sqrt(v);
    return 0;
}

} // namespace frico

extern "C" {

int frico_factor(frico_ctx* c, int form)
{
    if (check_ctx(c))
        return -1;
    LuWork wk;
    int rc = factor_prepare(c, form, wk, true);
    if (rc || (rc = factor_enqueue(c, form, wk)))
        return rc;
    rc = factor_finish(c, form, wk);
    // a zero pivot of an unpivoted factorisation: the pivoted LU instead
    if (rc > 0 && (form == FRICO_SOLVE_C64_SYM_IR || form == FRICO_SOLVE_C64_NP_IR))
        return frico_factor(c, FRICO_SOLVE_C64_IR);
    return rc;
}

int frico_factored_form(const frico_ctx* c)
{
    return c ? c->factored : -1;
}

// R = B - Z X with the kept (unfactored) Z.
static int residual(frico_ctx* c, KernelTimer* kt)
{
    const int64_t N = c->N, nr = c->nrhs;
    int rc;
    if ((rc = c->R.ensure((size_t)(N * nr))))
        return rc;
    if ((rc = launch_copy_bytes(c->B.p, c->R.p, (size_t)(N * nr) * sizeof(c128), c->st)))
        return rc;
    if (c->zprec == FRICO_C128) {
        if (c->z64_factored)
            return set_error("residual needs the unfactored Z"), FRICO_ERR_STATE;
        return gemm_sub<double>(N, nr, N, c->Z64.p, N, c->X.p, N, c->R.p, N, c->st, kt, c->resid_exact);
    }
    return gemm_sub_mixed(N, nr, N, c->Z32.p, N, c->X.p, N, c->R.p, N, c->st, kt);
}

} // extern "C"

namespace frico {

int refine_prepare(frico_ctx* c)
{
    const int64_t N = c->N, nr = c->nrhs;
    int rc;
    if ((rc = c->R.ensure((size_t)(N * nr))) || (rc = c->D32.ensure((size_t)(N * nr))) ||
        (rc = c->norms.ensure((size_t)nr)))
        return rc;
    if (c->stream_residual && c->resid_pair && (rc = c->R2.ensure((size_t)(N * nr))))
        return rc;
    // (also before a factorisation recorded into a graph, which then sets sexp_ok)
    if (c->solve_f16 && (rc = c->bexp.ensure((size_t)(2 * nr))))
        return rc;
    if (c->wv_ok && (rc = c->ysolve.ensure((size_t)(N * nr) * sizeof(c64))))
        return rc;
    if (nr <= 4 && (rc = c->chain.ensure(lu_chain_words(N, nr))))
        return rc;
    return 0;
}

int refine_enqueue(frico_ctx* c, bool first, bool overlap)
{
    const int64_t N = c->N, nr = c->nrhs;
    int rc;
    KernelTimer& tt = first ? c->t_solve : c->t_refine;
    const bool carried = first && c->aug_valid && c->aug_version == c->rhs_version && c->aug_cols == nr;
    // X += D and R = B - Z X taken block by block beside the backward substitution: needs the FP64
    // Z, the second stream, its events and the carried or fused solve
    ResidStream rs;
    const bool stream = overlap && c->stream_residual && c->side && c->zprec == FRICO_C128 && !c->z64_factored &&
                        (carried || c->wv_ok || (c->factored == FRICO_SOLVE_C64_NP_IR && lu_blocked_solves())) &&
                        (int64_t)c->lu_events.size() >= (N + 127) / 128 + 1;
    if (stream) {
        rs.side = c->side;
        rs.ev = c->lu_events.data();
        rs.nev = (int)c->lu_events.size();
        rs.Z = c->Z64.p;
        rs.ldz = N;
        rs.S = c->X.p;
        rs.lds = N;
        rs.R = c->R.p;
        rs.ldr = N;
        rs.exact = c->resid_exact;
        rs.R2 = c->resid_pair && c->R2.p ? c->R2.p : nullptr;
        if (c->sexp_ok && c->bexp.p && (c->factored == FRICO_SOLVE_C64_NP_IR || c->factored == FRICO_SOLVE_C64_SYM_IR)) {
            rs.aexp = c->sexp.p;
            rs.ldae = N;
            rs.bexp = c->bexp.p;
            rs.bexp2 = c->bexp.p + nr;
        }
    }
    phase_begin(c, tt);
    if (first && (rc = launch_zero_bytes(c->X.p, (size_t)(N * nr) * sizeof(c128), c->st)))
        return rc;
    if (first && !carried && (rc = launch_copy_bytes(c->B.p, c->R.p, (size_t)(N * nr) * sizeof(c128), c->st)))
        return rc;
    // the correction's right-hand side: the carried L^{-1} P B, or the current residual
    if (carried)
        rc = launch_copy_bytes(c->LU32.p + N * N, c->D32.p, (size_t)(N * nr) * sizeof(c64), c->st);
    else
        rc = launch_convert_c128_c64(c->R.p, N, c->D32.p, N, N, nr, c->st);
    if (rc)
        return rc;
    // the residual of this sweep starts from B (and its second accumulator from zero)
    if (stream && (rc = launch_copy_bytes(c->B.p, c->R.p, (size_t)(N * nr) * sizeof(c128), c->st)))
        return rc;
    if (stream && rs.R2 && (rc = launch_zero_bytes(rs.R2, (size_t)(N * nr) * sizeof(c128), c->st)))
        return rc;
    ChainSolve cs;
    cs.words = c->chain.p;
    cs.nwords = c->chain.n;
    if (carried)
        rc = lu_solve_upper<float>(c->LU32.p, N, N, c->D32.p, N, nr, c->st, stream ? &rs : nullptr, &cs);
    else
        // the factors without pivoting have no interchanges to apply
        rc = lu_solve<float>(c->LU32.p, N, N, c->factored == FRICO_SOLVE_C64_NP_IR ? nullptr : c->perm.p,
                             c->nb_factored, c->D32.p, N, nr, c->st, c->dinv.p, c->wv_ok ? c->wv.p : nullptr,
                             c->ysolve.p, stream ? &rs : nullptr, &cs);
    if (rc)
        return rc;
    if (!stream && (rc = launch_add_correction(c->X.p, N, c->D32.p, N, N, nr, c->st)))
        return rc;
    phase_end(c, tt);
    if (!stream) {
        phase_begin(c, c->t_refine);
        if ((rc = residual(c, nullptr)))
            return rc;
        phase_end(c, c->t_refine);
    }
    return launch_col_norm2(c->R.p, N, N, nr, c->norms.p, c->st);
}

int read_refine_norms(frico_ctx* c, const std::vector<double>& bn, double* worst)
{
    const int64_t nr = c->nrhs;
    std::vector<double> rn((size_t)nr);
    FRICO_HIP_CHECK(hipMemcpyAsync(rn.data(), c->norms.p, (size_t)nr * sizeof(double), hipMemcpyDeviceToHost, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    double w = 0;
    for (int64_t j = 0; j < nr; j++) {
        const double r = std::sqrt(rn[j]);
        const double rel = bn[j] > 0 ? r / bn[j] : r;
        // a NaN or infinite residual (factors broken down, a non-finite Z or B) is the worst
        // possible one: std::max would drop a NaN and report the column as converged
        w = std::isfinite(rel) ? std::max(w, rel) : HUGE_VAL;
    }
    *worst = w;
    return 0;
}

} // namespace frico

extern "C" {

int frico_solve(frico_ctx* c, double tol, int maxit, frico_solve_info* info)
{
    if (check_ctx(c))
        return -1;
    if (c->factored < 0)
        return set_error("solve before factor"), FRICO_ERR_STATE;
    if (c->nrhs < 1)
        return set_error("no right-hand side"), FRICO_ERR_STATE;
    int rc = set_dev(c);
    if (rc)
        return rc;
    const int64_t N = c->N, nr = c->nrhs;
    frico_solve_info inf{0, 1, -1.0};
    if (c->factored == FRICO_SOLVE_C128) {
        phase_begin(c, c->t_solve);
        if ((rc = launch_copy_bytes(c->B.p, c->X.p, (size_t)(N * nr) * sizeof(c128), c->st)))
            return rc;
        if (c->wv_ok && (rc = c->ysolve.ensure((size_t)(N * nr) * sizeof(c128))))
            return rc;
        rc = lu_solve<double>(c->Z64.p, N, N, c->perm.p, c->nb_factored, c->X.p, N, nr, c->st, c->dinv.p,
                              c->wv_ok ? c->wv.p : nullptr, c->ysolve.p);
        phase_end(c, c->t_solve);
        if (rc)
            return rc;
        FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
        c->have_x = true;
        if (info)
            *info = inf;
        return 0;
    }
    if (maxit < 1)
        return set_error("maxit must be >= 1"), -3;
    c->resid_exact = !(tol >= 1e-10);
    if ((rc = refine_prepare(c)))
        return rc;
    std::vector<double> bn;
    if ((rc = col_norms(c, c->B.p, nr, bn)))
        return rc;
    inf.converged = 0;
    int sweeps = 0;
    for (;;) {
        // unpivoted factors are left for the pivoted LU when a sweep fails to halve the residual
        // before the tolerance is met, or when maxit sweeps do not meet it
        const bool sym = c->factored == FRICO_SOLVE_C64_SYM_IR || c->factored == FRICO_SOLVE_C64_NP_IR;
        double prev = HUGE_VAL;
        for (int it = 1; it <= maxit; it++) {
            double worst = 0;
            if ((rc = refine_enqueue(c, it == 1, true)) || (rc = read_refine_norms(c, bn, &worst)))
                return rc;
            inf.iterations = ++sweeps;
            inf.max_rel_residual = worst;
            if (worst <= tol) {
                inf.converged = 1;
                break;
            }
            if (sym && !(worst < 0.5 * prev))
                break;
            prev = worst;
        }
        if (inf.converged || !sym)
            break;
        // the pivoted LU's fused solves need buffers the unpivoted factors' blocked solves did not
        if ((rc = frico_factor(c, FRICO_SOLVE_C64_IR)) || (rc = refine_prepare(c)))
            return rc;
    }
    c->have_x = true;
    if (info)
        *info = inf;
    if (!inf.converged)
        return set_error("refinement stopped at relative residual %.3e > %.3e after %d sweeps", inf.max_rel_residual, tol,
                         inf.iterations),
               FRICO_ERR_NOCONV;
    return 0;
}

int frico_residual(frico_ctx* c, double* rel)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_x)
        return set_error("no solution"), FRICO_ERR_STATE;
    if (!rel)
        return set_error("null output"), -2;
    int rc = set_dev(c);
    c->resid_exact = true; // a check of the answer: always the exact form
    if (rc || (rc = residual(c, nullptr)))
        return rc;
    std::vector<double> bn, rn;
    if ((rc = col_norms(c, c->B.p, c->nrhs, bn)) || (rc = col_norms(c, c->R.p, c->nrhs, rn)))
        return rc;
    for (int64_t j = 0; j < c->nrhs; j++)
        rel[j] = bn[j] > 0 ? rn[j] / bn[j] : rn[j];
    return 0;
}

int frico_solution_get(frico_ctx* c, double* Xh, int64_t ldx)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_x)
        return set_error("no solution"), FRICO_ERR_STATE;
    if (!Xh || ldx < c->N)
        return set_error("bad output"), -2;
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    FRICO_HIP_CHECK(hipMemcpy2D(Xh, ldx * sizeof(c128), c->X.p, c->N * sizeof(c128), c->N * sizeof(c128), c->nrhs,
                                hipMemcpyDeviceToHost));
    return 0;
}

int frico_solution_set(frico_ctx* c, const double* Xh, int64_t ldx, int64_t nrhs)
{
    if (check_ctx(c))
        return -1;
    if (!Xh || ldx < c->N || nrhs < 1 || (c->nrhs > 0 && nrhs != c->nrhs))
        return set_error("bad solution arguments"), -2;
    int rc = set_dev(c);
    if (rc)
        return rc;
    if (c->nrhs < 1 && (rc = rhs_alloc(c, nrhs)))
        return rc;
    FRICO_HIP_CHECK(hipMemcpy2DAsync(c->X.p, c->N * sizeof(c128), Xh, ldx * sizeof(c128), c->N * sizeof(c128), nrhs,
                                     hipMemcpyHostToDevice, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    c->have_x = true;
    return 0;
}

static int scatter_all(frico_ctx* c)
{
    int rc;
    const int64_t T = c->hm.nt;
    if (c->nrhs > 65535)
        return set_error("too many columns"), FRICO_ERR_LIMIT;
    if ((rc = c->TS.ensure((size_t)(4 * T * c->nrhs))))
        return rc;
    return launch_scatter(c->dm, c->X.p, c->N, (int)c->nrhs, c->TS.p, c->st);
}

int frico_triangle_currents(frico_ctx* c, int64_t col, double* AJ, double* AdivJ)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh || !c->have_x)
        return set_error("no solution"), FRICO_ERR_STATE;
    if (col < 0 || col >= c->nrhs)
        return set_error("column out of range"), -2;
    int rc = set_dev(c);
    if (rc || (rc = scatter_all(c)))
        return rc;
    const int64_t T = c->hm.nt;
    std::vector<c128> h((size_t)(4 * T));
    FRICO_HIP_CHECK(hipMemcpyAsync(h.data(), c->TS.p + 4 * T * col, h.size() * sizeof(c128), hipMemcpyDeviceToHost, c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    for (int64_t t = 0; t < T; t++) {
        for (int d = 0; d < 3; d++) {
            if (AJ) {
                AJ[2 * (3 * t + d)] = h[4 * t + d].re;
                AJ[2 * (3 * t + d) + 1] = h[4 * t + d].im;
            }
        }
        if (AdivJ) {
            AdivJ[2 * t] = h[4 * t + 3].re;
            AdivJ[2 * t + 1] = h[4 * t + 3].im;
        }
    }
    return 0;
}

int frico_fields(frico_ctx* c, double freq, const double* pts, int64_t npts, const int64_t* cols, double h_sign,
                 double* E, double* H)
{
    if (check_ctx(c))
        return -1;
    if (!c->have_mesh || !c->have_x)
        return set_error("no solution"), FRICO_ERR_STATE;
    if (!(freq > 0))
        return set_error("frequency must be positive"), -2;
    if (!pts || npts < 1 || npts > (1 << 30))
        return set_error("bad points"), -3;
    if (!E && !H)
        return set_error("null outputs"), -7;
    int rc = set_dev(c);
    if (rc || (rc = scatter_all(c)))
        return rc;
    std::vector<int> cl;
    if (cols) {
        cl.resize((size_t)npts);
        for (int64_t p = 0; p < npts; p++) {
            if (cols[p] < 0 || cols[p] >= c->nrhs)
                return set_error("point %lld uses column %lld out of range", (long long)p, (long long)cols[p]), -5;
            cl[p] = (int)cols[p];
        }
        if ((rc = c->pcols.upload(cl.data(), cl.size(), c->st)))
            return rc;
    }
    if ((rc = c->pts.upload(pts, (size_t)(3 * npts), c->st)) || (rc = c->EH.ensure((size_t)(6 * npts))))
        return rc;
    double omega;
    const double k = kwave(freq, nullptr, &omega);
    if ((rc = launch_fields(c->dm, k, omega * MU0, c->TS.p, (int)c->nrhs, c->pts.p, cols ? c->pcols.p : nullptr,
                            (int)npts, h_sign >= 0 ? 1.0 : -1.0, c->EH.p, c->EH.p + 3 * npts, c->st)))
        return rc;
    // E and H are npts x 3 complex each: the staging buffer holds E, then H
    if (E)
        FRICO_HIP_CHECK(hipMemcpyAsync(E, c->EH.p, (size_t)(3 * npts) * sizeof(c128), hipMemcpyDeviceToHost, c->st));
    if (H)
        FRICO_HIP_CHECK(hipMemcpyAsync(H, c->EH.p + 3 * npts, (size_t)(3 * npts) * sizeof(c128), hipMemcpyDeviceToHost,
                                       c->st));
    FRICO_HIP_CHECK(hipStreamSynchronize(c->st));
    return 0;
}

} // extern "C"
