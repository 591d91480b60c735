/*
 * cpu_ref — FP64 CPU restatement of frico's EFIE Method-of-Moments hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker or the CPU baseline; the product library never links or
 * calls it.
 *
 * Every function restates the reference's algorithm from its source read as text, keeping
 * its loop structure, accumulation order and arithmetic expression order, with the
 * citation (path:line, relative to the reference root) next to it.  The Dunavant table
 * is the published data (D. A. Dunavant 1985) at the 15 digits the reference uses.
 *
 * Parity pinning (see tests/test_oracle_*.py): quadrature exactness and weight sums,
 * RWG invariants (T- < T+, divergence, normal continuity), numpy second restatement of
 * the fill on tiny meshes, Z symmetry properties, LU backward error and scipy zgetrf,
 * Hertzian-dipole closed form for the fields, loose physics pins (dipole Zin, gain,
 * plate physical-optics RCS) and the reference-held dipole impedance table.
 *
 * Build: oracle/Makefile (g++ -O2 -fopenmp -shared).  ABI: plain extern "C".
 */
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <list>
#include <stdexcept>
#include <vector>

#include "../frico_amd/csrc/dunavant_rules.inc"

namespace {

using cd = std::complex<double>;

/* constants: src/constants.h:23-24 */
constexpr double MU0 = 1.256637061435917e-06;
constexpr double EPS0 = 8.8541878188e-12;

/* vec3 arithmetic: src/geom_point.h:30-138 */
struct vec3 {
    double x, y, z;
};
inline vec3 operator+(const vec3& a, const vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(const vec3& a, const vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator*(const vec3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator/(const vec3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(const vec3& a, const vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(const vec3& a, const vec3& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const vec3& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

struct cvec3 {
    cd x, y, z;
};

struct qpoint {
    vec3 p;
    double w;
};

struct edge_t {
    int64_t iv0, iv1; /* iv0 < iv1: src/geom_mesh.h:35-57 */
    bool operator<(const edge_t& o) const { return iv0 < o.iv0 || (iv0 == o.iv0 && iv1 < o.iv1); }
    bool operator==(const edge_t& o) const { return iv0 == o.iv0 && iv1 == o.iv1; }
};
inline edge_t mk_edge(int64_t a, int64_t b) { return {std::min(a, b), std::max(a, b)}; }

struct neighbours {
    int64_t itminus = -1;
    int loc_eminus = -1;
    int64_t itplus = -1;
    int loc_eplus = -1;
    bool has_iface = false;
    int iface = 0;
};

struct rwg {
    double Aminus, Aplus;
    int64_t itminus, itplus;
    vec3 pminus, pplus;
    double length;
    int64_t edge_index, matrix_index;
    bool has_iface;
    int iface;
    int loc_eminus, loc_eplus;
};

} // namespace

struct orc_mesh {
    std::vector<vec3> V;
    std::vector<std::array<int64_t, 3>> T;
    std::vector<int> tag;
    std::vector<edge_t> edges;
    std::vector<neighbours> nb;
    std::vector<rwg> bf;
    /* per-triangle quadrature caches (same values the reference regenerates per pair) */
    int cached_degree = -1;
    std::vector<std::vector<qpoint>> q_reg, q_sub;
};

namespace {

/* measure(triangle) = 0.5 |e0 x e1|: src/geom_mesh.cpp:71-78 */
double tri_measure(const orc_mesh& m, int64_t t)
{
    const auto& tr = m.T[t];
    vec3 e0 = m.V[tr[1]] - m.V[tr[0]];
    vec3 e1 = m.V[tr[2]] - m.V[tr[1]];
    return 0.5 * norm(cross(e0, e1));
}

/* barycenter(triangle): src/geom_mesh.cpp:64-69 */
vec3 tri_bar(const orc_mesh& m, int64_t t)
{
    const auto& tr = m.T[t];
    return (m.V[tr[0]] + m.V[tr[1]] + m.V[tr[2]]) / 3.0;
}

/* measure(edge) = hypot(dx, dy) -- z ignored, as the reference does: src/geom_mesh.cpp:46-52 */
double edge_measure(const orc_mesh& m, const edge_t& e)
{
    vec3 d = m.V[e.iv1] - m.V[e.iv0];
    return std::hypot(d.x, d.y);
}

vec3 edge_bar(const orc_mesh& m, const edge_t& e) { return (m.V[e.iv0] + m.V[e.iv1]) / 2.0; }

/* area_triangle_kahan: src/quadratures.cpp:41-57 */
double area_kahan(const vec3& p0, const vec3& p1, const vec3& p2)
{
    const double l10 = norm(p1 - p0);
    const double l20 = norm(p2 - p0);
    const double l21 = norm(p2 - p1);
    std::array<double, 3> len{l10, l20, l21};
    std::sort(len.begin(), len.end());
    const double a = len[2], b = len[1], c = len[0];
    return 0.25 * std::sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)));
}

inline vec3 bary(const vec3& p0, const vec3& p1, const vec3& p2, double l0, double l1, double l2)
{
    return p0 * l0 + p1 * l1 + p2 * l2;
}

/* dunavant(degree, p0, p1, p2): src/quadratures.cpp:321-372 (orbit expansion order :348-368) */
std::vector<qpoint> dunavant(int degree, const vec3& p0, const vec3& p1, const vec3& p2)
{
    if (degree > FRICO_DUNAVANT_MAX_DEGREE || degree < 0)
        throw std::invalid_argument("Dunavant quadrature: degree out of range");
    int rule = (degree == 0) ? 0 : degree - 1;
    double area = area_kahan(p0, p1, p2);
    std::vector<qpoint> ret;
    ret.reserve(frico_dunavant_npoints[rule]);
    for (int o = frico_dunavant_orbit_begin[rule]; o < frico_dunavant_orbit_begin[rule + 1]; o++) {
        const auto& d = frico_dunavant_orbits[o];
        double l0 = d.a, l1 = d.b, l2 = d.c, w = d.w * area;
        switch (d.orbit) {
        case 1:
            ret.push_back({bary(p0, p1, p2, l0, l1, l2), w});
            break;
        case 3:
            ret.push_back({bary(p0, p1, p2, l0, l1, l2), w});
            ret.push_back({bary(p0, p1, p2, l1, l2, l0), w});
            ret.push_back({bary(p0, p1, p2, l1, l0, l2), w});
            break;
        case 6:
            ret.push_back({bary(p0, p1, p2, l0, l1, l2), w});
            ret.push_back({bary(p0, p1, p2, l0, l2, l1), w});
            ret.push_back({bary(p0, p1, p2, l1, l0, l2), w});
            ret.push_back({bary(p0, p1, p2, l1, l2, l0), w});
            ret.push_back({bary(p0, p1, p2, l2, l0, l1), w});
            ret.push_back({bary(p0, p1, p2, l2, l1, l0), w});
            break;
        }
    }
    return ret;
}

/* integrate: src/quadratures.cpp:2678-2683 */
std::vector<qpoint> integrate(const orc_mesh& m, int64_t t, int degree)
{
    const auto& tr = m.T[t];
    return dunavant(degree, m.V[tr[0]], m.V[tr[1]], m.V[tr[2]]);
}

/* integrate_subtri: src/quadratures.cpp:2685-2702 */
std::vector<qpoint> integrate_subtri(const orc_mesh& m, int64_t t, int degree)
{
    const auto& tr = m.T[t];
    vec3 bar = tri_bar(m, t);
    auto ret = dunavant(degree, m.V[tr[0]], m.V[tr[1]], bar);
    auto q = dunavant(degree, m.V[tr[1]], m.V[tr[2]], bar);
    ret.insert(ret.end(), q.begin(), q.end());
    q = dunavant(degree, m.V[tr[2]], m.V[tr[0]], bar);
    ret.insert(ret.end(), q.begin(), q.end());
    return ret;
}

void cache_rules(orc_mesh& m, int degree)
{
    if (m.cached_degree == degree)
        return;
    m.q_reg.resize(m.T.size());
    m.q_sub.resize(m.T.size());
    for (size_t t = 0; t < m.T.size(); t++) {
        m.q_reg[t] = integrate(m, (int64_t)t, degree);
        m.q_sub[t] = integrate_subtri(m, (int64_t)t, degree);
    }
    m.cached_degree = degree;
}

inline double kwave(double freq)
{
    double omega = 2.0 * M_PI * freq;
    return omega * std::sqrt(MU0 * EPS0);
}

} // namespace

extern "C" {

/*
 * Build mesh connectivity and the RWG function space.
 *   edges: every triangle pushes (v0,v1),(v1,v2),(v2,v0); sort + unique   src/input_gmsh.cpp:120-140
 *   boundary (1D entity) edges must exist among them, else lonely_edge    src/input_gmsh.cpp:143-199
 *   T-/T+ by first/second visit, T- < T+; >2 triangles -> error           src/input_gmsh.cpp:216-262
 *   interface tag of internal edges that are 1D entity edges              src/input_gmsh.cpp:264-269
 *   RWG per internal edge in edge order, free vertex evmap {2,0,1}         src/rwg_basis.cpp:27-68
 * Vertices must already be compacted (every vertex used), as load_from_gmsh does
 * (src/input_gmsh.cpp:335-394).
 * Returns 0, -1 (lonely edge), -2 (edge shared by more than two triangles), -3 (bad index).
 */
int orc_mesh_build(const double* verts, int64_t nv, const int64_t* tris, const int32_t* tags,
                   int64_t nt, const int64_t* bedges, const int32_t* btags, int64_t nbe,
                   orc_mesh** out)
{
    auto* m = new orc_mesh;
    m->V.resize(nv);
    for (int64_t i = 0; i < nv; i++)
        m->V[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    m->T.resize(nt);
    m->tag.resize(nt);
    for (int64_t t = 0; t < nt; t++) {
        for (int k = 0; k < 3; k++) {
            int64_t v = tris[3 * t + k];
            if (v < 0 || v >= nv) {
                delete m;
                return -3;
            }
            m->T[t][k] = v;
        }
        m->tag[t] = tags ? tags[t] : 0;
        const auto& tr = m->T[t];
        m->edges.push_back(mk_edge(tr[0], tr[1]));
        m->edges.push_back(mk_edge(tr[1], tr[2]));
        m->edges.push_back(mk_edge(tr[2], tr[0]));
    }
    std::sort(m->edges.begin(), m->edges.end());
    m->edges.erase(std::unique(m->edges.begin(), m->edges.end()), m->edges.end());

    auto offset = [&](const edge_t& e) -> int64_t {
        auto it = std::lower_bound(m->edges.begin(), m->edges.end(), e);
        if (it != m->edges.end() && *it == e)
            return it - m->edges.begin();
        return -1;
    };

    std::vector<std::pair<int64_t, int>> bptr;
    for (int64_t b = 0; b < nbe; b++) {
        int64_t o = offset(mk_edge(bedges[2 * b], bedges[2 * b + 1]));
        if (o < 0) {
            delete m;
            return -1;
        }
        bptr.push_back({o, btags ? btags[b] : 0});
    }

    m->nb.assign(m->edges.size(), neighbours{});
    std::vector<int> flags(m->edges.size(), 0);
    for (int64_t t = 0; t < nt; t++) {
        const auto& tr = m->T[t];
        std::array<edge_t, 3> es{mk_edge(tr[0], tr[1]), mk_edge(tr[1], tr[2]), mk_edge(tr[2], tr[0])};
        for (int ie = 0; ie < 3; ie++) {
            int64_t o = offset(es[ie]);
            auto& en = m->nb[o];
            if (en.itplus >= 0) {
                delete m;
                return -2;
            }
            if (flags[o]) {
                if (en.itminus < t) {
                    en.itplus = t;
                    en.loc_eplus = ie;
                } else {
                    en.itplus = en.itminus;
                    en.itminus = t;
                    en.loc_eplus = en.loc_eminus;
                    en.loc_eminus = ie;
                }
            } else {
                en.itminus = t;
                en.loc_eminus = ie;
                flags[o] = 1;
            }
        }
    }
    for (auto& [o, tg] : bptr) {
        if (m->nb[o].itplus >= 0) {
            m->nb[o].has_iface = true;
            m->nb[o].iface = tg;
        }
    }

    const int evmap[3] = {2, 0, 1};
    int64_t midx = 0;
    for (size_t ie = 0; ie < m->nb.size(); ie++) {
        const auto& en = m->nb[ie];
        if (en.itplus < 0)
            continue;
        const auto& Tm = m->T[en.itminus];
        const auto& Tp = m->T[en.itplus];
        rwg b;
        b.Aminus = tri_measure(*m, en.itminus);
        b.Aplus = tri_measure(*m, en.itplus);
        b.itminus = en.itminus;
        b.itplus = en.itplus;
        b.pminus = m->V[Tm[evmap[en.loc_eminus]]];
        b.pplus = m->V[Tp[evmap[en.loc_eplus]]];
        b.length = edge_measure(*m, m->edges[ie]);
        b.edge_index = (int64_t)ie;
        b.matrix_index = midx++;
        b.has_iface = en.has_iface;
        b.iface = en.iface;
        b.loc_eminus = en.loc_eminus;
        b.loc_eplus = en.loc_eplus;
        m->bf.push_back(b);
    }
    *out = m;
    return 0;
}

void orc_mesh_free(orc_mesh* m) { delete m; }
int64_t orc_mesh_num_rwg(const orc_mesh* m) { return (int64_t)m->bf.size(); }
int64_t orc_mesh_num_edges(const orc_mesh* m) { return (int64_t)m->edges.size(); }

/* Export the RWG table (one row per basis function, in matrix order). */
void orc_mesh_rwg(const orc_mesh* m, int64_t* itm, int64_t* itp, int32_t* lem, int32_t* lep,
                  int64_t* edge, int32_t* has_iface, int32_t* iface, double* len, double* Am,
                  double* Ap, double* pm, double* pp)
{
    for (size_t i = 0; i < m->bf.size(); i++) {
        const auto& b = m->bf[i];
        itm[i] = b.itminus;
        itp[i] = b.itplus;
        lem[i] = b.loc_eminus;
        lep[i] = b.loc_eplus;
        edge[i] = b.edge_index;
        has_iface[i] = b.has_iface;
        iface[i] = b.iface;
        len[i] = b.length;
        Am[i] = b.Aminus;
        Ap[i] = b.Aplus;
        pm[3 * i] = b.pminus.x;
        pm[3 * i + 1] = b.pminus.y;
        pm[3 * i + 2] = b.pminus.z;
        pp[3 * i] = b.pplus.x;
        pp[3 * i + 1] = b.pplus.y;
        pp[3 * i + 2] = b.pplus.z;
    }
}

void orc_mesh_edges(const orc_mesh* m, int64_t* e2)
{
    for (size_t i = 0; i < m->edges.size(); i++) {
        e2[2 * i] = m->edges[i].iv0;
        e2[2 * i + 1] = m->edges[i].iv1;
    }
}

/* Quadrature rule on three points (for tests): returns the number of points written. */
int orc_dunavant(int degree, const double* p0, const double* p1, const double* p2, double* pts,
                 double* w)
{
    try {
        auto q = dunavant(degree, {p0[0], p0[1], p0[2]}, {p1[0], p1[1], p1[2]}, {p2[0], p2[1], p2[2]});
        for (size_t i = 0; i < q.size(); i++) {
            pts[3 * i] = q[i].p.x;
            pts[3 * i + 1] = q[i].p.y;
            pts[3 * i + 2] = q[i].p.z;
            w[i] = q[i].w;
        }
        return (int)q.size();
    } catch (...) {
        return -1;
    }
}

int orc_subtri(const orc_mesh* m, int64_t t, int degree, double* pts, double* w)
{
    try {
        auto q = integrate_subtri(*m, t, degree);
        for (size_t i = 0; i < q.size(); i++) {
            pts[3 * i] = q[i].p.x;
            pts[3 * i + 1] = q[i].p.y;
            pts[3 * i + 2] = q[i].p.z;
            w[i] = q[i].w;
        }
        return (int)q.size();
    } catch (...) {
        return -1;
    }
}

/*
 * Quadrature-mode impedance matrix: src/emw_solver.cpp:45-142.
 * Z is N x N complex128 column-major; entry (j, i) = Z[j + i*N] (src/emw_solver.cpp:139).
 * The reference regenerates both rules per (i, j); the rules depend only on the triangle and
 * the split flag, so they are computed once per triangle here (identical values).
 */
int orc_fill_quad(orc_mesh* m, double freq, int degree, double* Zout)
{
    if (degree < 0 || degree > FRICO_DUNAVANT_MAX_DEGREE)
        return -2;
    cache_rules(*m, degree);
    const int64_t N = (int64_t)m->bf.size();
    cd* Z = reinterpret_cast<cd*>(Zout);
    double omega = 2.0 * M_PI * freq;
    double k = omega * std::sqrt(MU0 * EPS0);
    double inv_ksq = 1. / (omega * omega * MU0 * EPS0);

#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ii = 0; ii < N; ii++) {
        const rwg& ibf = m->bf[ii];
        const auto& iqps_minus = m->q_reg[ibf.itminus];
        const auto& iqps_plus = m->q_reg[ibf.itplus];
        for (int64_t jj = 0; jj < N; jj++) {
            const rwg& jbf = m->bf[jj];
            bool split_minus = (jbf.itminus == ibf.itminus) || (jbf.itminus == ibf.itplus);
            bool split_plus = (jbf.itplus == ibf.itminus) || (jbf.itplus == ibf.itplus);
            const auto& jqps_minus = split_minus ? m->q_sub[jbf.itminus] : m->q_reg[jbf.itminus];
            const auto& jqps_plus = split_plus ? m->q_sub[jbf.itplus] : m->q_reg[jbf.itplus];

            cd entry = 0.0;
            double prodl = 0.25 * ibf.length * jbf.length * M_1_PI;

            auto block = [&](const std::vector<qpoint>& iq, const std::vector<qpoint>& jq,
                             const vec3& pi, const vec3& pj, double invAA, double sgn) {
                for (const auto& iqp : iq) {
                    for (const auto& jqp : jq) {
                        double v = 0.25 * dot(iqp.p - pi, jqp.p - pj);
                        double s = inv_ksq;
                        double w = iqp.w * jqp.w * invAA;
                        double Rij = norm(iqp.p - jqp.p);
                        cd g = std::exp(cd{0, -k * Rij}) / Rij;
                        if (sgn > 0)
                            entry += prodl * w * (v - s) * g;
                        else
                            entry -= prodl * w * (v - s) * g;
                    }
                }
            };
            block(iqps_minus, jqps_minus, ibf.pminus, jbf.pminus, 1. / (ibf.Aminus * jbf.Aminus), +1);
            block(iqps_minus, jqps_plus, ibf.pminus, jbf.pplus, 1. / (ibf.Aminus * jbf.Aplus), -1);
            block(iqps_plus, jqps_minus, ibf.pplus, jbf.pminus, 1. / (ibf.Aplus * jbf.Aminus), -1);
            block(iqps_plus, jqps_plus, ibf.pplus, jbf.pplus, 1. / (ibf.Aplus * jbf.Aplus), +1);
            Z[jbf.matrix_index + ibf.matrix_index * N] = entry;
        }
    }
    return 0;
}

/*
 * Approximate (MOM3D centroid) impedance matrix: src/emw_solver.cpp:144-220.
 * Self pairs use the equivalent-disk term 0.25*(sqrt(4 pi)/sqrt(A) - jk)/pi (:185-214).
 */
int orc_fill_approx(orc_mesh* m, double freq, double* Zout)
{
    const int64_t N = (int64_t)m->bf.size();
    cd* Z = reinterpret_cast<cd*>(Zout);
    double omega = 2.0 * M_PI * freq;
    double k = omega * std::sqrt(MU0 * EPS0);
    double inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    double sqrt_4pi = std::sqrt(4 * M_PI);

#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ii = 0; ii < N; ii++) {
        const rwg& ibf = m->bf[ii];
        vec3 iTmbar = tri_bar(*m, ibf.itminus);
        vec3 iTpbar = tri_bar(*m, ibf.itplus);
        for (int64_t jj = 0; jj < N; jj++) {
            const rwg& jbf = m->bf[jj];
            vec3 jTmbar = tri_bar(*m, jbf.itminus);
            vec3 jTpbar = tri_bar(*m, jbf.itplus);
            bool selfmm = ibf.itminus == jbf.itminus;
            bool selfmp = ibf.itminus == jbf.itplus;
            bool selfpm = ibf.itplus == jbf.itminus;
            bool selfpp = ibf.itplus == jbf.itplus;
            double sqrt_Am = std::sqrt(tri_measure(*m, ibf.itminus));
            double sqrt_Ap = std::sqrt(tri_measure(*m, ibf.itplus));
            cd jk{0., k};
            cd entry = 0.0;
            auto term = [&](const vec3& ci, const vec3& pi, const vec3& cj, const vec3& pj,
                            bool self, double sqrtA) -> cd {
                double rho_c = dot(ci - pi, cj - pj);
                double R = norm(ci - cj);
                cd G = std::exp(cd{0, -k * R}) / (4.0 * M_PI * R);
                if (self)
                    G = 0.25 * (sqrt_4pi / sqrtA - jk) * M_1_PI;
                return ibf.length * jbf.length * (0.25 * rho_c * G - inv_ksq * G);
            };
            entry += term(iTmbar, ibf.pminus, jTmbar, jbf.pminus, selfmm, sqrt_Am);
            entry -= term(iTmbar, ibf.pminus, jTpbar, jbf.pplus, selfmp, sqrt_Am);
            entry -= term(iTpbar, ibf.pplus, jTmbar, jbf.pminus, selfpm, sqrt_Ap);
            entry += term(iTpbar, ibf.pplus, jTpbar, jbf.pplus, selfpp, sqrt_Ap);
            Z[jbf.matrix_index + ibf.matrix_index * N] = entry;
        }
    }
    return 0;
}

/*
 * Delta-gap edge reorientation: src/emw_solver.cpp:222-350.  Writes signs[N] (only the gap
 * entries are meaningful) and sets *applied = 1 when the reference would have reoriented.
 * Returns 0 on success or when no fix is needed, -1 if no edges / not connected.
 */
int orc_reorient_deltagap(const orc_mesh* m, const int32_t* ifaces, int nif, int force,
                          double* signs, int* applied)
{
    *applied = 0;
    bool fix_needed = false;
    std::list<std::pair<edge_t, int64_t>> edges;
    for (const auto& b : m->bf) {
        if (!b.has_iface)
            continue;
        for (int f = 0; f < nif; f++) {
            if (b.iface == ifaces[f]) {
                if (m->tag[b.itplus] <= m->tag[b.itminus])
                    fix_needed = true;
                edges.push_back({m->edges[b.edge_index], b.matrix_index});
            }
        }
    }
    if (!(fix_needed || force))
        return 0;
    if (edges.empty())
        return -1;
    struct nodepair {
        int64_t first, second, bf_index;
    };
    std::list<nodepair> seglist;
    auto eit = edges.begin();
    seglist.push_back({eit->first.iv0, eit->first.iv1, eit->second});
    edges.erase(eit);
    while (!edges.empty()) {
        size_t before = edges.size();
        for (eit = edges.begin(); eit != edges.end();) {
            auto f = seglist.front();
            auto b = seglist.back();
            auto cur = eit++;
            /* the reference erases `cur` and may test it again in the next `if`; a removed
             * element is never revisited here, which matches it whenever one edge matches once */
            bool erased = false;
            if (!erased && !edges.empty() && cur->first.iv0 == f.first) {
                seglist.push_front({cur->first.iv1, cur->first.iv0, cur->second});
                edges.erase(cur);
                erased = true;
            }
            if (!erased && !edges.empty() && cur->first.iv1 == f.first) {
                seglist.push_front({cur->first.iv0, cur->first.iv1, cur->second});
                edges.erase(cur);
                erased = true;
            }
            if (!erased && !edges.empty() && cur->first.iv0 == b.second) {
                seglist.push_back({cur->first.iv0, cur->first.iv1, cur->second});
                edges.erase(cur);
                erased = true;
            }
            if (!erased && !edges.empty() && cur->first.iv1 == b.second) {
                seglist.push_back({cur->first.iv1, cur->first.iv0, cur->second});
                edges.erase(cur);
                erased = true;
            }
        }
        if (edges.size() == before)
            return -1;
    }
    const auto& seg0 = seglist.front();
    const auto& bf0 = m->bf[seg0.bf_index];
    vec3 dir = tri_bar(*m, bf0.itplus) - edge_bar(*m, m->edges[bf0.edge_index]);
    double sign = +1;
    for (auto& seg : seglist) {
        const auto& b = m->bf[seg.bf_index];
        vec3 curdir = tri_bar(*m, b.itplus) - edge_bar(*m, m->edges[b.edge_index]);
        if (dot(dir, curdir) < 0)
            sign *= -1;
        signs[b.matrix_index] = sign;
        dir = curdir;
    }
    *applied = 1;
    return 0;
}

/* update_rhs(delta_gap): src/emw_solver.cpp:352-378.  V accumulates (+=).  signs may be NULL. */
int orc_rhs_deltagap(const orc_mesh* m, double freq, const int32_t* ifaces, int nif, double vre,
                     double vim, const double* signs, double* Vout)
{
    cd* V = reinterpret_cast<cd*>(Vout);
    double omega = 2.0 * M_PI * freq;
    cd volt{vre, vim};
    for (const auto& b : m->bf) {
        if (!b.has_iface)
            continue;
        for (int f = 0; f < nif; f++) {
            if (b.iface == ifaces[f]) {
                cd v{0.0, -b.length / (omega * MU0)};
                double sign = signs ? signs[b.matrix_index] : 1.0;
                V[b.matrix_index] += v * volt * sign;
            }
        }
    }
    return 0;
}

/*
 * plane_wave (src/sources.h:34-75): E(r) = (Eth theta_hat + Eph phi_hat) exp(+j k dir.r), with dir
 * the unit vector from the origin toward the source at (theta, phi) (:47-58, 67-74).
 */
struct plane_wave_field {
    vec3 dir;
    cvec3 E0;
    cd jk;
    plane_wave_field(double k, double theta, double phi, cd Eth, cd Eph)
    {
        dir = vec3{std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
        vec3 th{std::cos(theta) * std::cos(phi), std::cos(theta) * std::sin(phi), -std::sin(theta)};
        vec3 ph{-std::sin(phi), std::cos(phi), 0.0};
        E0 = cvec3{Eth * th.x + Eph * ph.x, Eth * th.y + Eph * ph.y, Eth * th.z + Eph * ph.z};
        jk = cd{0.0, k};
    }
    cvec3 operator()(const vec3& p) const
    {
        cd ph_ = std::exp(jk * dot(dir, p));
        return {E0.x * ph_, E0.y * ph_, E0.z * ph_};
    }
};

/* The incident field of plane_wave at npts points (pts npts x 3, E npts x 3 complex). */
void orc_plane_wave_field(double freq, double theta, double phi, double eth_re, double eth_im, double eph_re,
                          double eph_im, const double* pts, int64_t npts, double* Eout)
{
    const double k = 2.0 * M_PI * freq * std::sqrt(MU0 * EPS0);
    const plane_wave_field field(k, theta, phi, cd{eth_re, eth_im}, cd{eph_re, eph_im});
    cd* E = reinterpret_cast<cd*>(Eout);
    for (int64_t i = 0; i < npts; i++) {
        const cvec3 e = field(vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
        E[3 * i] = e.x;
        E[3 * i + 1] = e.y;
        E[3 * i + 2] = e.z;
    }
}

/*
 * update_rhs(plane_wave): src/emw_solver.cpp:380-416 with the plane_wave field above.
 * V accumulates (+=).
 */
int orc_rhs_planewave(orc_mesh* m, double freq, int degree, double theta, double phi, double eth_re,
                      double eth_im, double eph_re, double eph_im, double* Vout)
{
    if (degree < 0 || degree > FRICO_DUNAVANT_MAX_DEGREE)
        return -2;
    cache_rules(*m, degree);
    cd* V = reinterpret_cast<cd*>(Vout);
    double omega = 2.0 * M_PI * freq;
    double k = omega * std::sqrt(MU0 * EPS0);
    const plane_wave_field field(k, theta, phi, cd{eth_re, eth_im}, cd{eph_re, eph_im});
    for (const auto& b : m->bf) {
        cd v{0.0, 0.0};
        double wminus = 0.5 * b.length / b.Aminus;
        double wplus = 0.5 * b.length / b.Aplus;
        for (const auto& qp : m->q_reg[b.itminus]) {
            cvec3 E = field(qp.p);
            vec3 r = qp.p - b.pminus;
            /* Eigen's real.dot(complex): sum of r_i * E_i (no conjugation of a real vector) */
            v -= wminus * qp.w * (r.x * E.x + r.y * E.y + r.z * E.z);
        }
        for (const auto& qp : m->q_reg[b.itplus]) {
            cvec3 E = field(qp.p);
            vec3 r = qp.p - b.pplus;
            v += wplus * qp.w * (r.x * E.x + r.y * E.y + r.z * E.z);
        }
        V[b.matrix_index] += v / cd{0, -omega * MU0};
    }
    return 0;
}

/*
 * Dense complex LU with partial pivoting and solve -- the LAPACK zgetrf/zgetrs semantics that
 * solve_general uses (src/lapack_wrappers.cpp:53-133): pivot = first max of |re|+|im| (izamax),
 * row interchanges over the whole row, unit-lower L, B overwritten by X.  A (n x n, col-major)
 * is overwritten by the factors.  Returns 0, or k>0 if U(k,k) is exactly zero (1-based).
 */
int orc_lu_solve(double* Aout, int64_t n, double* Bout, int64_t nrhs)
{
    cd* A = reinterpret_cast<cd*>(Aout);
    cd* B = reinterpret_cast<cd*>(Bout);
    std::vector<int64_t> ipiv(n);
    int info = 0;
    auto a = [&](int64_t i, int64_t j) -> cd& { return A[i + j * n]; };
    for (int64_t j = 0; j < n; j++) {
        int64_t p = j;
        double best = -1;
        for (int64_t i = j; i < n; i++) {
            double v = std::abs(a(i, j).real()) + std::abs(a(i, j).imag());
            if (v > best) {
                best = v;
                p = i;
            }
        }
        ipiv[j] = p;
        if (a(p, j) != cd(0)) {
            if (p != j)
                for (int64_t c = 0; c < n; c++)
                    std::swap(a(j, c), a(p, c));
            cd rp = 1.0 / a(j, j);
            for (int64_t i = j + 1; i < n; i++)
                a(i, j) *= rp;
        } else if (info == 0) {
            info = (int)(j + 1);
        }
#pragma omp parallel for schedule(static)
        for (int64_t c = j + 1; c < n; c++) {
            cd u = a(j, c);
            if (u != cd(0))
                for (int64_t i = j + 1; i < n; i++)
                    a(i, c) -= a(i, j) * u;
        }
    }
    if (info)
        return info;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrhs; r++) {
        cd* b = B + r * n;
        for (int64_t j = 0; j < n; j++)
            if (ipiv[j] != j)
                std::swap(b[j], b[ipiv[j]]);
        for (int64_t j = 0; j < n; j++) {
            cd x = b[j];
            for (int64_t i = j + 1; i < n; i++)
                b[i] -= a(i, j) * x;
        }
        for (int64_t j = n - 1; j >= 0; j--) {
            b[j] /= a(j, j);
            cd x = b[j];
            for (int64_t i = 0; i < j; i++)
                b[i] -= a(i, j) * x;
        }
    }
    return 0;
}

/*
 * RWG -> triangle scatter of one current vector: src/emw_solver.h:137-164.
 * tri_AJ and tri_J are T x 3 complex (row-major: [t][xyz]); tri_AdivJ is T complex.
 */
int orc_scatter(const orc_mesh* m, const double* Iin, double* AJout, double* AdivJout, double* Jout)
{
    const cd* I = reinterpret_cast<const cd*>(Iin);
    size_t T = m->T.size();
    cd* AJ = reinterpret_cast<cd*>(AJout);
    cd* AdivJ = reinterpret_cast<cd*>(AdivJout);
    cd* J = reinterpret_cast<cd*>(Jout);
    std::fill(AJ, AJ + 3 * T, cd(0));
    std::fill(AdivJ, AdivJ + T, cd(0));
    if (J)
        std::fill(J, J + 3 * T, cd(0));
    for (const auto& b : m->bf) {
        vec3 bm = tri_bar(*m, b.itminus), bp = tri_bar(*m, b.itplus);
        double Am = tri_measure(*m, b.itminus), Ap = tri_measure(*m, b.itplus);
        cd Ie = I[b.matrix_index];
        vec3 fm = (bm - b.pminus) * (-b.length) / (2.0 * b.Aminus); /* eval_minus rwg_basis.h:76-81 */
        vec3 fp = (bp - b.pplus) * b.length / (2.0 * b.Aplus);      /* eval_plus  rwg_basis.h:83-88 */
        cd Jm[3] = {fm.x * Ie, fm.y * Ie, fm.z * Ie};
        cd Jp[3] = {fp.x * Ie, fp.y * Ie, fp.z * Ie};
        for (int c = 0; c < 3; c++) {
            AJ[3 * b.itminus + c] += Am * Jm[c];
            AJ[3 * b.itplus + c] += Ap * Jp[c];
            if (J) {
                J[3 * b.itminus + c] += Jm[c];
                J[3 * b.itplus + c] += Jp[c];
            }
        }
        AdivJ[b.itminus] += Am * (-b.length / b.Aminus) * Ie; /* div_minus rwg_basis.h:90-94 */
        AdivJ[b.itplus] += Ap * (b.length / b.Aplus) * Ie;    /* div_plus  rwg_basis.h:96-100 */
    }
    return 0;
}

/*
 * eval_fields at points: src/emw_postpro_common.cpp:31-66.  E uses exp(-jkR), H uses the
 * reference's exp(+jkR) (:60-61).  E, H: npts x 3 complex (row-major).
 */
int orc_eval_fields(const orc_mesh* m, double freq, const double* AJin, const double* AdivJin,
                    const double* pts, int64_t npts, double* Eout, double* Hout)
{
    const cd* AJ = reinterpret_cast<const cd*>(AJin);
    const cd* AdivJ = reinterpret_cast<const cd*>(AdivJin);
    cd* Eo = reinterpret_cast<cd*>(Eout);
    cd* Ho = reinterpret_cast<cd*>(Hout);
    const double omega = 2.0 * M_PI * freq;
    const double k = omega * std::sqrt(MU0 * EPS0);
    const cd jomega{0.0, omega};
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t ip = 0; ip < npts; ip++) {
        vec3 pt{pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2]};
        cd E[3] = {0, 0, 0}, H[3] = {0, 0, 0};
        for (size_t t = 0; t < m->T.size(); t++) {
            vec3 bar = tri_bar(*m, (int64_t)t);
            vec3 vR = pt - bar;
            double R = norm(vR);
            double _4piR = 4 * M_PI * R;
            double _4piR3 = _4piR * R * R;
            cd jkR{0, k * R};
            cd g = std::exp(-jkR) / _4piR;
            cd J[3] = {AJ[3 * t], AJ[3 * t + 1], AJ[3 * t + 2]};
            cd divJ = AdivJ[t];
            double kR2 = std::pow(k * R, 2);
            double vr[3] = {vR.x, vR.y, vR.z};
            for (int c = 0; c < 3; c++)
                E[c] += -jomega * MU0 * (J[c] - divJ * (1.0 + jkR) * vr[c] / kR2) * g;
            cd RxJ[3] = {vR.y * J[2] - vR.z * J[1], vR.z * J[0] - vR.x * J[2], vR.x * J[1] - vR.y * J[0]};
            cd gradg = -((1.0 + jkR) / _4piR3) * std::exp(jkR);
            for (int c = 0; c < 3; c++)
                H[c] += RxJ[c] * gradg;
        }
        for (int c = 0; c < 3; c++) {
            Eo[3 * ip + c] = E[c];
            Ho[3 * ip + c] = H[c];
        }
    }
    return 0;
}

/*
 * Port values of a delta-gap: src/emw_ana_antenna.cpp:76-105; gamma/swr
 * src/emw_postpro_common.cpp:99-123.  out[0..9] = I(re,im), P(re,im), Z(re,im), Gamma(re,im),
 * swr, unused.
 */
int orc_port_values(const orc_mesh* m, const double* Iin, const int32_t* ifaces, int nif,
                    double vre, double vim, const double* signs, double Z0, double* out)
{
    const cd* Ivec = reinterpret_cast<const cd*>(Iin);
    cd volt{vre, vim}, cur = 0.0, pwr = 0.0;
    for (const auto& b : m->bf) {
        if (!b.has_iface)
            continue;
        for (int f = 0; f < nif; f++) {
            if (b.iface == ifaces[f]) {
                double sign = signs ? signs[b.matrix_index] : 1.0;
                cd I = b.length * Ivec[b.matrix_index] * sign;
                cur += I;
                pwr += 0.5 * volt * std::conj(I);
            }
        }
    }
    cd Zin = volt / cur;
    cd gam = (Zin - Z0) / (Zin + Z0);
    double s = (1.0 + std::abs(gam)) / (1.0 - std::abs(gam));
    out[0] = cur.real();
    out[1] = cur.imag();
    out[2] = pwr.real();
    out[3] = pwr.imag();
    out[4] = Zin.real();
    out[5] = Zin.imag();
    out[6] = gam.real();
    out[7] = gam.imag();
    out[8] = s;
    out[9] = 0;
    return 0;
}

/*
 * Gain on the three reference cuts (xy, yz, xz), 360 x 1 degree at distance R:
 * src/emw_ana_antenna.cpp:33-65.  G = 4 pi |P|^2 |S| / pwr with S = E x conj(H) / 2 and
 * |S| = sqrt of Eigen's conjugating S.dot(S).  gains: 3 x 360 doubles.
 */
int orc_gain(const orc_mesh* m, double freq, const double* AJ, const double* AdivJ, double pwr,
             double R, double* gains)
{
    std::vector<double> pts(3 * 1080);
    for (int plane = 0; plane < 3; plane++)
        for (int deg = 0; deg < 360; deg++) {
            double th = deg * M_PI / 180;
            double c = std::cos(th), s = std::sin(th);
            double* P = &pts[3 * (plane * 360 + deg)];
            if (plane == 0) {
                P[0] = R * c; P[1] = R * s; P[2] = 0.0;
            } else if (plane == 1) {
                P[0] = 0.0; P[1] = R * c; P[2] = R * s;
            } else {
                P[0] = -R * c; P[1] = 0.0; P[2] = R * s;
            }
        }
    std::vector<double> E(6 * 1080), H(6 * 1080);
    orc_eval_fields(m, freq, AJ, AdivJ, pts.data(), 1080, E.data(), H.data());
    for (int i = 0; i < 1080; i++) {
        const cd* e = reinterpret_cast<const cd*>(&E[6 * i]);
        const cd* h = reinterpret_cast<const cd*>(&H[6 * i]);
        cd hc[3] = {std::conj(h[0]), std::conj(h[1]), std::conj(h[2])};
        cd S[3] = {0.5 * (e[1] * hc[2] - e[2] * hc[1]), 0.5 * (e[2] * hc[0] - e[0] * hc[2]),
                   0.5 * (e[0] * hc[1] - e[1] * hc[0])};
        cd SS = std::conj(S[0]) * S[0] + std::conj(S[1]) * S[1] + std::conj(S[2]) * S[2];
        double Prad = std::real(std::sqrt(SS));
        const double* P = &pts[3 * i];
        double Rn = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
        gains[i] = 4 * M_PI * Rn * Rn * Prad / pwr;
    }
    return 0;
}

/*
 * RCS of a scattered field at (R, theta_s, phi_s): bistatic_RCS_unit_Ei
 * src/emw_ana_bistatic_rcs.cpp:62-82.  out = {sigma_VV, sigma_HH}.
 */
void orc_rcs_unit_Ei(const double* Es_in, double theta_s, double phi_s, double R, double* out)
{
    const cd* Es = reinterpret_cast<const cd*>(Es_in);
    double th[3] = {std::cos(theta_s) * std::cos(phi_s), std::cos(theta_s) * std::sin(phi_s),
                    -std::sin(theta_s)};
    double ph[3] = {-std::sin(phi_s), std::cos(phi_s), 0.0};
    cd Et = th[0] * Es[0] + th[1] * Es[1] + th[2] * Es[2];
    cd Ep = ph[0] * Es[0] + ph[1] * Es[1] + ph[2] * Es[2];
    out[0] = 4.0 * M_PI * R * R * std::norm(Et);
    out[1] = 4.0 * M_PI * R * R * std::norm(Ep);
}

/* sph2rect: src/utils.h:53-60 */
void orc_sph2rect(double R, double theta, double phi, double* p)
{
    p[0] = R * std::cos(phi) * std::sin(theta);
    p[1] = R * std::sin(phi) * std::sin(theta);
    p[2] = R * std::cos(theta);
}

/*
 * Bistatic RCS table: src/emw_ana_bistatic_rcs.cpp:96-117.  359 rows (i = 0..358) of
 * {VV_theta, HH_theta, VV_phi, HH_phi}; the phi cut point is sph2rect(R, 0, i deg) (:110).
 */
int orc_bistatic(const orc_mesh* m, double freq, const double* AJ, const double* AdivJ, double R,
                 double* table)
{
    std::vector<double> pts(3 * 718);
    double deg2rad = M_PI / 180.0;
    for (int i = 0; i < 359; i++) {
        orc_sph2rect(R, deg2rad * i, 0.0, &pts[3 * (2 * i)]);
        orc_sph2rect(R, 0.0, deg2rad * i, &pts[3 * (2 * i + 1)]);
    }
    std::vector<double> E(6 * 718), H(6 * 718);
    orc_eval_fields(m, freq, AJ, AdivJ, pts.data(), 718, E.data(), H.data());
    for (int i = 0; i < 359; i++) {
        orc_rcs_unit_Ei(&E[6 * (2 * i)], deg2rad * i, 0.0, R, &table[4 * i]);
        orc_rcs_unit_Ei(&E[6 * (2 * i + 1)], 0.0, deg2rad * i, R, &table[4 * i + 2]);
    }
    return 0;
}


/* ---------------------------------------------------------------------------------------------
 * Adjoint support (no counterpart in the reference).  Derivatives of the quadrature-mode Z with
 * respect to vertex coordinates, differentiated by hand through the reference's per-RWG-pair
 * loop (src/emw_solver.cpp:56-141): quadrature points are linear in the vertices (barycentric),
 * weights scale with the triangle area (Kahan/Heron == cross-product area mathematically),
 * A-/A+ are cross-product areas and the edge length is hypot(dx, dy).
 * ------------------------------------------------------------------------------------------- */
} // extern "C"

namespace {

struct dqpoint {
    vec3 p;
    double w;
    double beta[3];      /* p = sum_k beta[k] * V[tri[k]] */
    vec3 dw[3];          /* d w / d V[tri[k]] */
};

/* gradient of the area of triangle (a, b, c): d/da = 0.5 n_hat x (c - b), cyclic */
void area_grad(const vec3& a, const vec3& b, const vec3& c, vec3 g[3])
{
    vec3 n = cross(b - a, c - a);
    double nn = norm(n);
    if (nn == 0) {
        g[0] = g[1] = g[2] = {0, 0, 0};
        return;
    }
    vec3 nh = n / nn;
    g[0] = cross(nh, c - b) * 0.5;
    g[1] = cross(nh, a - c) * 0.5;
    g[2] = cross(nh, b - a) * 0.5;
}

/* Dunavant rule on corners q_j = sum_k M[j][k] P_k of triangle P, with derivative data. */
void dunavant_d(int degree, const vec3 P[3], const double M[3][3], std::vector<dqpoint>& out)
{
    vec3 q[3];
    for (int j = 0; j < 3; j++)
        q[j] = P[0] * M[j][0] + P[1] * M[j][1] + P[2] * M[j][2];
    auto base = dunavant(degree, q[0], q[1], q[2]);
    double area = area_kahan(q[0], q[1], q[2]);
    vec3 ga[3];
    area_grad(q[0], q[1], q[2], ga);
    int rule = degree == 0 ? 0 : degree - 1;
    size_t idx = 0;
    for (int o = frico_dunavant_orbit_begin[rule]; o < frico_dunavant_orbit_begin[rule + 1]; o++) {
        const auto& d = frico_dunavant_orbits[o];
        double perm[6][3] = {{d.a, d.b, d.c}, {d.b, d.c, d.a}, {d.b, d.a, d.c}};
        int np = d.orbit;
        if (np == 6) {
            double p6[6][3] = {{d.a, d.b, d.c}, {d.a, d.c, d.b}, {d.b, d.a, d.c},
                               {d.b, d.c, d.a}, {d.c, d.a, d.b}, {d.c, d.b, d.a}};
            std::memcpy(perm, p6, sizeof p6);
        }
        for (int s = 0; s < np; s++, idx++) {
            dqpoint dp;
            dp.p = base[idx].p;
            dp.w = base[idx].w;
            double wref = d.w;
            for (int k = 0; k < 3; k++) {
                dp.beta[k] = perm[s][0] * M[0][k] + perm[s][1] * M[1][k] + perm[s][2] * M[2][k];
                vec3 g = ga[0] * M[0][k] + ga[1] * M[1][k] + ga[2] * M[2][k];
                dp.dw[k] = g * wref;
            }
            (void)area;
            out.push_back(dp);
        }
    }
}

void rules_d(const orc_mesh& m, int64_t t, int degree, bool split, std::vector<dqpoint>& out)
{
    const auto& tr = m.T[t];
    vec3 P[3] = {m.V[tr[0]], m.V[tr[1]], m.V[tr[2]]};
    out.clear();
    if (!split) {
        const double I3[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        dunavant_d(degree, P, I3, out);
        return;
    }
    const double t3 = 1.0 / 3.0;
    const double S0[3][3] = {{1, 0, 0}, {0, 1, 0}, {t3, t3, t3}};
    const double S1[3][3] = {{0, 1, 0}, {0, 0, 1}, {t3, t3, t3}};
    const double S2[3][3] = {{0, 0, 1}, {1, 0, 0}, {t3, t3, t3}};
    dunavant_d(degree, P, S0, out);
    dunavant_d(degree, P, S1, out);
    dunavant_d(degree, P, S2, out);
}

struct cvec3acc {
    cd x = 0, y = 0, z = 0;
    void add(const vec3& v, cd s)
    {
        x += v.x * s;
        y += v.y * s;
        z += v.z * s;
    }
};

} // namespace

extern "C" {

/*
 * grad[3*v + c] (complex) += sum_{i,j} W_j I_i dZ(j,i)/dx_{v,c} for the quadrature-mode Z.
 * grad is nv*3 complex (interleaved), overwritten.
 */
int orc_adjoint_contract(orc_mesh* m, double freq, int degree, const double* Iin, const double* Win, double* gout)
{
    if (degree < 0 || degree > FRICO_DUNAVANT_MAX_DEGREE)
        return -2;
    const int64_t N = (int64_t)m->bf.size();
    const cd* I = reinterpret_cast<const cd*>(Iin);
    const cd* W = reinterpret_cast<const cd*>(Win);
    const size_t nv = m->V.size();
    double omega = 2.0 * M_PI * freq;
    double k = omega * std::sqrt(MU0 * EPS0);
    double inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    const cd jk{0, k};
    std::vector<cd> total(3 * nv, cd(0));
#pragma omp parallel
    {
        std::vector<cd> G(3 * nv, cd(0));
        std::vector<dqpoint> qi, qj;
        auto addv = [&](int64_t v, const cvec3acc& a) {
            G[3 * v] += a.x;
            G[3 * v + 1] += a.y;
            G[3 * v + 2] += a.z;
        };
        auto edge_grad = [&](const rwg& b, cd s) {
            /* d hypot(dx, dy): d/d v1 = (dx, dy, 0)/l, d/d v0 = -(dx, dy, 0)/l */
            const edge_t& e = m->edges[b.edge_index];
            vec3 d = m->V[e.iv1] - m->V[e.iv0];
            if (b.length == 0)
                return;
            vec3 g{d.x / b.length, d.y / b.length, 0.0};
            cvec3acc a;
            a.add(g, s);
            addv(e.iv1, a);
            cvec3acc a2;
            a2.add(g, -s);
            addv(e.iv0, a2);
        };
        auto tri_area_grad = [&](int64_t t, cd s) {
            const auto& tr = m->T[t];
            vec3 g[3];
            area_grad(m->V[tr[0]], m->V[tr[1]], m->V[tr[2]], g);
            for (int kk = 0; kk < 3; kk++) {
                cvec3acc a;
                a.add(g[kk], s);
                addv(tr[kk], a);
            }
        };
        const int evmap[3] = {2, 0, 1};
#pragma omp for schedule(dynamic, 4)
        for (int64_t ii = 0; ii < N; ii++) {
            const rwg& ibf = m->bf[ii];
            for (int64_t jj = 0; jj < N; jj++) {
                const rwg& jbf = m->bf[jj];
                const cd om = W[jbf.matrix_index] * I[ibf.matrix_index];
                if (om == cd(0))
                    continue;
                double prodl = 0.25 * ibf.length * jbf.length * M_1_PI;
                for (int a = 0; a < 2; a++) {
                    int64_t P = a == 0 ? ibf.itminus : ibf.itplus;
                    const vec3 fi = a == 0 ? ibf.pminus : ibf.pplus;
                    int64_t fiv = m->T[P][evmap[a == 0 ? ibf.loc_eminus : ibf.loc_eplus]];
                    double Ai = a == 0 ? ibf.Aminus : ibf.Aplus;
                    rules_d(*m, P, degree, false, qi);
                    for (int b = 0; b < 2; b++) {
                        int64_t Q = b == 0 ? jbf.itminus : jbf.itplus;
                        bool split = (Q == ibf.itminus) || (Q == ibf.itplus);
                        const vec3 fj = b == 0 ? jbf.pminus : jbf.pplus;
                        int64_t fjv = m->T[Q][evmap[b == 0 ? jbf.loc_eminus : jbf.loc_eplus]];
                        double Aj = b == 0 ? jbf.Aminus : jbf.Aplus;
                        rules_d(*m, Q, degree, split, qj);
                        const double sg = (a == b) ? 1.0 : -1.0;
                        const double invAA = 1.0 / (Ai * Aj);
                        /* S = sum w w h g and the coefficient vectors of its derivative */
                        cd S = 0;
                        std::vector<cvec3acc> gP(3), gQ(3);
                        cvec3acc gfi, gfj;
                        for (const auto& p : qi) {
                            cvec3acc grp;
                            cd dwp = 0;
                            vec3 rhop = p.p - fi;
                            for (const auto& r : qj) {
                                vec3 rhor = r.p - fj;
                                vec3 dv = p.p - r.p;
                                double R = norm(dv);
                                cd g = std::exp(cd{0, -k * R}) / R;
                                cd gp = -(jk + 1.0 / R) * g; /* dg/dR */
                                double h = 0.25 * dot(rhop, rhor) - inv_ksq;
                                double ww = p.w * r.w;
                                S += ww * h * g;
                                /* d/dr_p: ww (0.25 rho_r g + h g' dv/R);  d/dr_r: ww (0.25 rho_p g - h g' dv/R) */
                                grp.add(rhor, ww * 0.25 * g);
                                grp.add(dv, ww * h * gp / R);
                                cvec3acc grr;
                                grr.add(rhop, ww * 0.25 * g);
                                grr.add(dv, -ww * h * gp / R);
                                gfi.add(rhor, -ww * 0.25 * g);
                                gfj.add(rhop, -ww * 0.25 * g);
                                dwp += r.w * h * g;
                                cd dwr = p.w * h * g;
                                for (int kk = 0; kk < 3; kk++) {
                                    gQ[kk].x += grr.x * r.beta[kk];
                                    gQ[kk].y += grr.y * r.beta[kk];
                                    gQ[kk].z += grr.z * r.beta[kk];
                                    gQ[kk].add(r.dw[kk], dwr);
                                }
                            }
                            for (int kk = 0; kk < 3; kk++) {
                                gP[kk].x += grp.x * p.beta[kk];
                                gP[kk].y += grp.y * p.beta[kk];
                                gP[kk].z += grp.z * p.beta[kk];
                                gP[kk].add(p.dw[kk], dwp);
                            }
                        }
                        /* T = sg prodl invAA S */
                        const cd c0 = om * sg * prodl * invAA;
                        auto scale_add = [&](int64_t v, const cvec3acc& a) {
                            cvec3acc s;
                            s.x = a.x * c0;
                            s.y = a.y * c0;
                            s.z = a.z * c0;
                            addv(v, s);
                        };
                        for (int kk = 0; kk < 3; kk++) {
                            scale_add(m->T[P][kk], gP[kk]);
                            scale_add(m->T[Q][kk], gQ[kk]);
                        }
                        scale_add(fiv, gfi);
                        scale_add(fjv, gfj);
                        /* areas: -T (dAi/Ai + dAj/Aj) */
                        const cd Tv = om * sg * prodl * invAA * S;
                        tri_area_grad(P, -Tv / Ai);
                        tri_area_grad(Q, -Tv / Aj);
                        /* lengths: prodl' = 0.25/pi (l_j dl_i + l_i dl_j) */
                        const cd Tl = om * sg * invAA * S * 0.25 * M_1_PI;
                        edge_grad(ibf, Tl * jbf.length);
                        edge_grad(jbf, Tl * ibf.length);
                    }
                }
            }
        }
#pragma omp critical
        for (size_t x = 0; x < 3 * nv; x++)
            total[x] += G[x];
    }
    std::memcpy(gout, total.data(), total.size() * sizeof(cd));
    return 0;
}


/*
 * grad[3*v + c] (complex) = sum_{i,j} W_j I_i dZ(j,i)/dx_{v,c} for the approximate-mode Z of
 * compute_matrix_approx (src/emw_solver.cpp:144-220), differentiated by hand through its per-RWG-
 * pair loop: each of the four terms s l_i l_j (0.25 rho_c G - G / k^2) depends on the two
 * barycenters (mean of the vertices), the two free vertices (rho_c = (c_i - p_i).(c_j - p_j)),
 * their distance R through G = exp(-jkR) / (4 pi R), or, for a pair of equal triangles, on the
 * area through G = 0.25 (sqrt(4 pi) / sqrt(A) - jk) / pi; the lengths are hypot(dx, dy).
 */
int orc_adjoint_contract_approx(orc_mesh* m, double freq, const double* Iin, const double* Win, double* gout)
{
    const int64_t N = (int64_t)m->bf.size();
    const cd* I = reinterpret_cast<const cd*>(Iin);
    const cd* W = reinterpret_cast<const cd*>(Win);
    const size_t nv = m->V.size();
    const double omega = 2.0 * M_PI * freq;
    const double k = omega * std::sqrt(MU0 * EPS0);
    const double inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    const double sqrt_4pi = std::sqrt(4 * M_PI);
    const cd jk{0., k};
    const int evmap[3] = {2, 0, 1};
    std::vector<cd> total(3 * nv, cd(0));
#pragma omp parallel
    {
        std::vector<cd> G(3 * nv, cd(0));
        auto addv = [&](int64_t v, const vec3& d, cd s) {
            G[3 * v] += d.x * s;
            G[3 * v + 1] += d.y * s;
            G[3 * v + 2] += d.z * s;
        };
        auto add_bar = [&](int64_t t, const vec3& d, cd s) { // d/d(barycenter) -> a third to each vertex
            for (int kk = 0; kk < 3; kk++)
                addv(m->T[t][kk], d, s / 3.0);
        };
        auto edge_grad = [&](const rwg& b, cd s) { // s dl/dx, l = hypot(dx, dy)
            const edge_t& e = m->edges[b.edge_index];
            vec3 d = m->V[e.iv1] - m->V[e.iv0];
            if (b.length == 0)
                return;
            const vec3 g{d.x / b.length, d.y / b.length, 0.0};
            addv(e.iv1, g, s);
            addv(e.iv0, g, -s);
        };
        auto area_grad_add = [&](int64_t t, cd s) {
            const auto& tr = m->T[t];
            vec3 g[3];
            area_grad(m->V[tr[0]], m->V[tr[1]], m->V[tr[2]], g);
            for (int kk = 0; kk < 3; kk++)
                addv(tr[kk], g[kk], s);
        };
#pragma omp for schedule(dynamic, 4)
        for (int64_t ii = 0; ii < N; ii++) {
            const rwg& ibf = m->bf[ii];
            for (int64_t jj = 0; jj < N; jj++) {
                const rwg& jbf = m->bf[jj];
                const cd om = W[jbf.matrix_index] * I[ibf.matrix_index];
                if (om == cd(0))
                    continue;
                const double ll = ibf.length * jbf.length;
                cd entry_over_ll = 0;
                for (int a = 0; a < 2; a++) {
                    for (int b = 0; b < 2; b++) {
                        const int64_t P = a == 0 ? ibf.itminus : ibf.itplus;
                        const int64_t Q = b == 0 ? jbf.itminus : jbf.itplus;
                        const double sg = a == b ? 1.0 : -1.0;
                        const vec3 ci = tri_bar(*m, P), cj = tri_bar(*m, Q);
                        const vec3 pi = a == 0 ? ibf.pminus : ibf.pplus, pj = b == 0 ? jbf.pminus : jbf.pplus;
                        const int64_t piv = m->T[P][evmap[a == 0 ? ibf.loc_eminus : ibf.loc_eplus]];
                        const int64_t pjv = m->T[Q][evmap[b == 0 ? jbf.loc_eminus : jbf.loc_eplus]];
                        const vec3 ui = ci - pi, uj = cj - pj;
                        const double rho_c = dot(ui, uj);
                        const bool self = P == Q;
                        cd Gv, dGdR = 0;
                        double A = 0;
                        vec3 dR{0, 0, 0};
                        if (self) {
                            A = tri_measure(*m, P);
                            Gv = 0.25 * (sqrt_4pi / std::sqrt(A) - jk) * M_1_PI;
                        } else {
                            const vec3 d = ci - cj;
                            const double R = norm(d);
                            Gv = std::exp(cd{0, -k * R}) / (4.0 * M_PI * R);
                            dGdR = -(jk + 1.0 / R) * Gv;
                            dR = d / R; // dR/dci = d / R, dR/dcj = -d / R
                        }
                        const cd h = 0.25 * rho_c - inv_ksq;
                        entry_over_ll += sg * h * Gv;
                        const cd c0 = om * sg * ll;
                        // rho_c: d/dci = uj, d/dpi = -uj, d/dcj = ui, d/dpj = -ui (times 0.25 G)
                        add_bar(P, uj, c0 * 0.25 * Gv);
                        addv(piv, uj, -c0 * 0.25 * Gv);
                        add_bar(Q, ui, c0 * 0.25 * Gv);
                        addv(pjv, ui, -c0 * 0.25 * Gv);
                        if (self) {
                            // dG/dA = -0.125 sqrt(4 pi) A^(-3/2) / pi
                            const double dGdA = -0.125 * sqrt_4pi / (A * std::sqrt(A)) * M_1_PI;
                            area_grad_add(P, c0 * h * dGdA);
                        } else {
                            add_bar(P, dR, c0 * h * dGdR);
                            add_bar(Q, dR, -c0 * h * dGdR);
                        }
                    }
                }
                // l_i l_j: d/dl_i = l_j (entry / ll), d/dl_j = l_i (entry / ll)
                edge_grad(ibf, om * entry_over_ll * jbf.length);
                edge_grad(jbf, om * entry_over_ll * ibf.length);
            }
        }
#pragma omp critical
        for (size_t x = 0; x < 3 * nv; x++)
            total[x] += G[x];
    }
    std::memcpy(gout, total.data(), total.size() * sizeof(cd));
    return 0;
}

/*
 * Vector-Jacobian product of the field sums at one point (src/emw_postpro_common.cpp:31-66 with
 * the scatter of src/emw_solver.h:137-164) for the current vector I:
 *   gx[3v+c] = sum_d eE_d dE_d/dx_{v,c} + eH_d dH_d/dx_{v,c}   (I held fixed)
 *   gI[n]    = sum_d eE_d dE_d/dI_n     + eH_d dH_d/dI_n
 * E_t = -j w mu0 (AJ_t g - D_t phi vR), H_t = (vR x AJ_t) psi with vR = pt - c_t,
 * g = e^{-jkR}/(4 pi R), phi = (1+jkR) e^{-jkR}/(4 pi k^2 R^3), psi = -(1+jkR) e^{+jkR}/(4 pi R^3).
 */
int orc_fields_vjp(const orc_mesh* m, double freq, const double* Iin, const double* ptin, const double* eEin,
                   const double* eHin, double* gxout, double* gIout)
{
    const cd* I = reinterpret_cast<const cd*>(Iin);
    const cd* eE = reinterpret_cast<const cd*>(eEin);
    const cd* eH = reinterpret_cast<const cd*>(eHin);
    cd* gx = reinterpret_cast<cd*>(gxout);
    cd* gI = reinterpret_cast<cd*>(gIout);
    const size_t nv = m->V.size(), T = m->T.size();
    std::fill(gx, gx + 3 * nv, cd(0));
    std::fill(gI, gI + m->bf.size(), cd(0));
    const double omega = 2.0 * M_PI * freq;
    const double k = omega * std::sqrt(MU0 * EPS0);
    const cd jwm{0.0, omega * MU0};
    const cd jk{0.0, k};
    vec3 pt{ptin[0], ptin[1], ptin[2]};
    std::vector<cd> AJ(3 * T, cd(0)), D(T, cd(0));
    for (const auto& b : m->bf) {
        for (int s = 0; s < 2; s++) {
            int64_t t = s == 0 ? b.itminus : b.itplus;
            double sg = s == 0 ? -1.0 : 1.0;
            vec3 c = tri_bar(*m, t), f = s == 0 ? b.pminus : b.pplus;
            vec3 al = c - f;
            cd In = I[b.matrix_index];
            AJ[3 * t] += sg * b.length * al.x * 0.5 * In;
            AJ[3 * t + 1] += sg * b.length * al.y * 0.5 * In;
            AJ[3 * t + 2] += sg * b.length * al.z * 0.5 * In;
            D[t] += sg * b.length * In;
        }
    }
    std::vector<cd> cAJ(3 * T), cD(T);
    std::vector<vec3> dummy;
    std::vector<cd> gc(3 * T, cd(0)); /* cotangent of the centroid c_t */
    for (size_t t = 0; t < T; t++) {
        vec3 c = tri_bar(*m, (int64_t)t);
        vec3 v = pt - c;
        double R = norm(v);
        cd e = std::exp(cd{0, -k * R});
        cd g = e / (4 * M_PI * R);
        cd gp = -(jk + 1.0 / R) * g;
        cd phi = (1.0 + jk * R) * e / (4 * M_PI * k * k * R * R * R);
        cd vphi = e / (4 * M_PI * k * k * R * R * R);
        cd phip = vphi * (-3.0 / R + k * k * R - 3.0 * jk);
        cd ep = std::exp(cd{0, k * R});
        cd psi = -(1.0 + jk * R) * ep / (4 * M_PI * R * R * R);
        cd vpsi = ep / (4 * M_PI * R * R * R);
        cd psip = vpsi * (3.0 / R + k * k * R + jk);
        cd A[3] = {AJ[3 * t], AJ[3 * t + 1], AJ[3 * t + 2]};
        double vr[3] = {v.x, v.y, v.z};
        cd eEv = eE[0] * v.x + eE[1] * v.y + eE[2] * v.z;
        cd eEA = eE[0] * A[0] + eE[1] * A[1] + eE[2] * A[2];
        /* eH x vR and eH . (vR x A) */
        cd eHxv[3] = {eH[1] * v.z - eH[2] * v.y, eH[2] * v.x - eH[0] * v.z, eH[0] * v.y - eH[1] * v.x};
        cd vxA[3] = {v.y * A[2] - v.z * A[1], v.z * A[0] - v.x * A[2], v.x * A[1] - v.y * A[0]};
        cd eHvxA = eH[0] * vxA[0] + eH[1] * vxA[1] + eH[2] * vxA[2];
        cd Axe[3] = {A[1] * eH[2] - A[2] * eH[1], A[2] * eH[0] - A[0] * eH[2], A[0] * eH[1] - A[1] * eH[0]};
        for (int d = 0; d < 3; d++)
            cAJ[3 * t + d] = -jwm * g * eE[d] + psi * eHxv[d];
        cD[t] = jwm * phi * eEv;
        for (int d = 0; d < 3; d++) {
            cd dE = -jwm * (eEA * gp * vr[d] / R - D[t] * (phip * eEv * vr[d] / R + phi * eE[d]));
            cd dH = psi * Axe[d] + psip * vr[d] / R * eHvxA;
            gc[3 * t + d] -= dE + dH; /* vR = pt - c */
        }
    }
    auto addv = [&](int64_t vv, const cd g3[3]) {
        gx[3 * vv] += g3[0];
        gx[3 * vv + 1] += g3[1];
        gx[3 * vv + 2] += g3[2];
    };
    for (const auto& b : m->bf) {
        cd In = I[b.matrix_index];
        cd dl = 0;
        for (int s = 0; s < 2; s++) {
            int64_t t = s == 0 ? b.itminus : b.itplus;
            double sg = s == 0 ? -1.0 : 1.0;
            vec3 c = tri_bar(*m, t), f = s == 0 ? b.pminus : b.pplus;
            vec3 al = c - f;
            cd ca = cAJ[3 * t] * al.x + cAJ[3 * t + 1] * al.y + cAJ[3 * t + 2] * al.z;
            gI[b.matrix_index] += sg * b.length * 0.5 * ca + sg * b.length * cD[t];
            dl += sg * 0.5 * ca * In + sg * In * cD[t];
            /* AJ += s l (c - f)/2 I: centroid and free vertex */
            cd gcc[3], gf[3];
            for (int d = 0; d < 3; d++) {
                gcc[d] = sg * b.length * 0.5 * In * cAJ[3 * t + d];
                gf[d] = -gcc[d];
                gc[3 * t + d] += gcc[d];
            }
            const int evmap[3] = {2, 0, 1};
            int lo = s == 0 ? b.loc_eminus : b.loc_eplus;
            addv(m->T[t][evmap[lo]], gf);
        }
        const edge_t& e = m->edges[b.edge_index];
        vec3 d = m->V[e.iv1] - m->V[e.iv0];
        if (b.length > 0) {
            cd g1[3] = {dl * (d.x / b.length), dl * (d.y / b.length), 0.0};
            cd g0[3] = {-g1[0], -g1[1], 0.0};
            addv(e.iv1, g1);
            addv(e.iv0, g0);
        }
    }
    for (size_t t = 0; t < T; t++) {
        cd g3[3] = {gc[3 * t] / 3.0, gc[3 * t + 1] / 3.0, gc[3 * t + 2] / 3.0};
        for (int kk = 0; kk < 3; kk++)
            addv(m->T[t][kk], g3);
    }
    return 0;
}

/* d length(edge of RWG n) / d x: writes gl[3*nv] real for one RWG (zero elsewhere). */
/*
 * g[3*v + c] (complex) = sum_n W_n dV_n/dx_{v,c} for the plane-wave moments of orc_rhs_planewave
 * (update_rhs(plane_wave), src/emw_solver.cpp:380-416): V_n = sum over T+, T- of
 * +-(l / 2A) sum_p w_p (r_p - p_T) . E(r_p) / (-j omega mu0), with every vertex dependence
 * differentiated: the edge length l = hypot(dx, dy), the area A = |e0 x e1| / 2, the rule weights
 * (reference weight times the triangle's area), the rule points r_p = sum_k beta_pk v_k, the free
 * vertex p_T, and the field E(r) = E0 exp(+j k dir.r), dE/dr_p = j k dir E^T.  g is nv*3 complex.
 */
int orc_rhs_planewave_vjp(orc_mesh* m, double freq, int degree, double theta, double phi, double eth_re,
                          double eth_im, double eph_re, double eph_im, const double* Win, double* gout)
{
    if (degree < 0 || degree > FRICO_DUNAVANT_MAX_DEGREE)
        return -2;
    const cd* W = reinterpret_cast<const cd*>(Win);
    cd* G = reinterpret_cast<cd*>(gout);
    const size_t nv = m->V.size();
    std::fill(G, G + 3 * nv, cd(0));
    const double omega = 2.0 * M_PI * freq;
    const double k = omega * std::sqrt(MU0 * EPS0);
    const plane_wave_field field(k, theta, phi, cd{eth_re, eth_im}, cd{eph_re, eph_im});
    const cd c0 = 1.0 / cd{0, -omega * MU0};
    const int evmap[3] = {2, 0, 1};
    auto addv = [&](int64_t v, const vec3& g, cd s) {
        G[3 * v] += g.x * s;
        G[3 * v + 1] += g.y * s;
        G[3 * v + 2] += g.z * s;
    };
    std::vector<dqpoint> q;
    for (const auto& b : m->bf) {
        const cd wn = W[b.matrix_index];
        if (wn == cd(0))
            continue;
        for (int a = 0; a < 2; a++) {
            const int64_t t = a == 0 ? b.itminus : b.itplus;
            const double sgn = a == 0 ? -1.0 : 1.0;
            const double A = a == 0 ? b.Aminus : b.Aplus;
            const vec3 pf = a == 0 ? b.pminus : b.pplus;
            const int64_t fv = m->T[t][evmap[a == 0 ? b.loc_eminus : b.loc_eplus]];
            rules_d(*m, t, degree, false, q);
            // F = sum_p w_p rho_p . E_p and its gradient per vertex of the triangle
            cd F = 0;
            cvec3acc dFv[3], dFfree;
            for (const auto& p : q) {
                const cvec3 E = field(p.p);
                const vec3 rho = p.p - pf;
                const cd rE = rho.x * E.x + rho.y * E.y + rho.z * E.z;
                F += p.w * rE;
                for (int kk = 0; kk < 3; kk++) {
                    // weight: dw_p/dv_k (rho . E); point: w_p beta_pk (E + (rho . E) j k dir)
                    dFv[kk].add(p.dw[kk], rE);
                    const cd s = p.w * p.beta[kk];
                    dFv[kk].x += s * (E.x + rE * field.jk * field.dir.x);
                    dFv[kk].y += s * (E.y + rE * field.jk * field.dir.y);
                    dFv[kk].z += s * (E.z + rE * field.jk * field.dir.z);
                }
                // free vertex: d rho / d p_T = -1
                dFfree.x -= p.w * E.x;
                dFfree.y -= p.w * E.y;
                dFfree.z -= p.w * E.z;
            }
            const cd pre = wn * sgn * c0;               // V_n += sgn c0 (l / 2A) F
            const double lA = b.length / (2.0 * A);
            // d(l / 2A) = dl / 2A - l dA / 2A^2
            {
                const edge_t& e = m->edges[b.edge_index];
                if (b.length > 0) {
                    const vec3 d = m->V[e.iv1] - m->V[e.iv0];
                    const vec3 gl{d.x / b.length, d.y / b.length, 0.0};
                    addv(e.iv1, gl, pre * F / (2.0 * A));
                    addv(e.iv0, gl, -pre * F / (2.0 * A));
                }
                const auto& tr = m->T[t];
                vec3 ga[3];
                area_grad(m->V[tr[0]], m->V[tr[1]], m->V[tr[2]], ga);
                for (int kk = 0; kk < 3; kk++)
                    addv(tr[kk], ga[kk], -pre * F * b.length / (2.0 * A * A));
                for (int kk = 0; kk < 3; kk++) {
                    G[3 * tr[kk]] += pre * lA * dFv[kk].x;
                    G[3 * tr[kk] + 1] += pre * lA * dFv[kk].y;
                    G[3 * tr[kk] + 2] += pre * lA * dFv[kk].z;
                }
                G[3 * fv] += pre * lA * dFfree.x;
                G[3 * fv + 1] += pre * lA * dFfree.y;
                G[3 * fv + 2] += pre * lA * dFfree.z;
            }
        }
    }
    return 0;
}

int orc_length_grad(const orc_mesh* m, int64_t n, double* gl)
{
    std::fill(gl, gl + 3 * m->V.size(), 0.0);
    const rwg& b = m->bf[n];
    const edge_t& e = m->edges[b.edge_index];
    vec3 d = m->V[e.iv1] - m->V[e.iv0];
    if (b.length == 0)
        return 0;
    gl[3 * e.iv1] = d.x / b.length;
    gl[3 * e.iv1 + 1] = d.y / b.length;
    gl[3 * e.iv0] = -d.x / b.length;
    gl[3 * e.iv0 + 1] = -d.y / b.length;
    return 0;
}
} // extern "C"
