// gmsh .msh (format 4.1, ASCII) ingestion reproducing load_from_gmsh (src/input_gmsh.cpp:36-406)
// without the gmsh library: the reference opens the file with gmsh and reads the model back through
// its API; this reader parses the file and returns what those API calls return, in their order.
//   gmsh_get_vertices   getNodes(-1, -1): every entity's nodes, entities by (dim, tag), each
//                       entity's nodes in stored (file) order; a node tag -> offset map (:36-66)
//   gmsh_get_triangles  getEntities(2) by tag, each surface's elements in stored order, the first
//                       three nodes of each (first-order corners) (:83-141)
//   gmsh_get_boundary_edges  getEntities(1) by tag, first two nodes of each line element (:143-196)
//   used-vertex compaction: vertices no triangle references are dropped, the others keep their
//                       relative order (:335-394)
//   gmsh_get_physical_groups: name -> (dim, tag, entity tags) (:198-212)
// The connectivity, edge order and RWG numbering follow in frico_mesh_load from these arrays.
#include "../../include/frico_gpu.h"
#include "common.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

struct frico_msh {
    std::vector<double> V;       // compacted vertices, 3 per vertex
    std::vector<int64_t> T;      // triangles, 3 vertex offsets each
    std::vector<int32_t> ttag;   // surface entity tag of each triangle
    std::vector<int64_t> B;      // line elements of curves, 2 vertex offsets each
    std::vector<int32_t> btag;   // curve entity tag of each
    struct group {
        std::string name;
        int dim, tag;
        std::vector<int32_t> entities;
    };
    std::vector<group> groups; // ordered by name, as the reference's std::map
};

namespace frico {

namespace {

struct block {
    int dim, tag;
    std::vector<int64_t> node_tags; // nodes: tags in file order
    std::vector<double> xyz;
    int etype = 0;                  // elements: type and node tags, nodes_per_elem each
    std::vector<int64_t> enodes;
};

// gmsh element types used here: nodes per element (MSH 4.1 specification, element type table).
int nodes_per_element(int type)
{
    switch (type) {
    case 1: return 2;   // 2-node line
    case 2: return 3;   // 3-node triangle
    case 3: return 4;   // 4-node quadrangle
    case 4: return 4;   // 4-node tetrahedron
    case 5: return 8;   // hexahedron
    case 6: return 6;   // prism
    case 7: return 5;   // pyramid
    case 8: return 3;   // 3-node line (second order)
    case 9: return 6;   // 6-node triangle
    case 10: return 9;  // 9-node quadrangle
    case 11: return 10; // 10-node tetrahedron
    case 15: return 1;  // point
    case 16: return 8;  // 8-node quadrangle
    case 20: return 9;  // 9-node triangle (incomplete third order)
    case 21: return 10; // 10-node triangle
    case 26: return 4;  // 4-node line
    case 27: return 5;  // 5-node line
    default: return -1;
    }
}

bool is_line(int type) { return type == 1 || type == 8 || type == 26 || type == 27; }
bool is_triangle(int type) { return type == 2 || type == 9 || type == 20 || type == 21; }

struct parse_error {
    std::string what;
};

std::string next_token(std::istream& in, const char* ctx)
{
    std::string s;
    if (!(in >> s))
        throw parse_error{std::string("unexpected end of file in ") + ctx};
    return s;
}

template <typename I> I next_int(std::istream& in, const char* ctx)
{
    const std::string s = next_token(in, ctx);
    try {
        size_t pos = 0;
        const long long v = std::stoll(s, &pos);
        if (pos != s.size())
            throw 0;
        return (I)v;
    } catch (...) {
        throw parse_error{"'" + s + "' is not an integer in " + ctx};
    }
}

double next_double(std::istream& in, const char* ctx)
{
    const std::string s = next_token(in, ctx);
    try {
        size_t pos = 0;
        const double v = std::stod(s, &pos);
        if (pos != s.size())
            throw 0;
        return v;
    } catch (...) {
        throw parse_error{"'" + s + "' is not a number in " + ctx};
    }
}

void expect_end(std::istream& in, const std::string& section)
{
    const std::string s = next_token(in, section.c_str());
    if (s != "$End" + section)
        throw parse_error{"expected $End" + section + ", found '" + s + "'"};
}

} // namespace

static frico_msh* parse_msh(std::istream& in)
{
    std::map<int, std::string> pnames[4];          // physical tag -> name, per dimension
    std::map<int, std::vector<int>> ent_phys[4];   // entity tag -> physical tags, per dimension
    std::vector<block> nblocks, eblocks;
    bool have_format = false, have_nodes = false, have_elements = false;
    std::string tok;
    while (in >> tok) {
        if (tok == "$MeshFormat") {
            const std::string ver = next_token(in, "$MeshFormat");
            const int ftype = next_int<int>(in, "$MeshFormat");
            (void)next_int<int>(in, "$MeshFormat"); // data size
            if (ver.rfind("4.1", 0) != 0)
                throw parse_error{"mesh format " + ver + " (only 4.1 is read)"};
            if (ftype != 0)
                throw parse_error{"binary mesh files are not read (save as ASCII)"};
            expect_end(in, "MeshFormat");
            have_format = true;
        } else if (tok == "$PhysicalNames") {
            const int n = next_int<int>(in, "$PhysicalNames");
            std::string rest;
            std::getline(in, rest);
            for (int i = 0; i < n; i++) {
                std::string line;
                if (!std::getline(in, line))
                    throw parse_error{"unexpected end of file in $PhysicalNames"};
                std::istringstream ls(line);
                const int dim = next_int<int>(ls, "$PhysicalNames"), tag = next_int<int>(ls, "$PhysicalNames");
                const size_t q0 = line.find('"'), q1 = line.rfind('"');
                if (q0 == std::string::npos || q1 == q0 || dim < 0 || dim > 3)
                    throw parse_error{"bad physical name line '" + line + "'"};
                pnames[dim][tag] = line.substr(q0 + 1, q1 - q0 - 1);
            }
            expect_end(in, "PhysicalNames");
        } else if (tok == "$Entities") {
            int cnt[4];
            for (int d = 0; d < 4; d++)
                cnt[d] = next_int<int>(in, "$Entities");
            for (int d = 0; d < 4; d++)
                for (int i = 0; i < cnt[d]; i++) {
                    const int tag = next_int<int>(in, "$Entities");
                    for (int q = 0; q < (d == 0 ? 3 : 6); q++)
                        (void)next_double(in, "$Entities");
                    const int np = next_int<int>(in, "$Entities");
                    std::vector<int> ph(np);
                    for (auto& p : ph)
                        p = next_int<int>(in, "$Entities");
                    ent_phys[d][tag] = ph;
                    if (d > 0) {
                        const int nbd = next_int<int>(in, "$Entities");
                        for (int q = 0; q < nbd; q++)
                            (void)next_int<int>(in, "$Entities");
                    }
                }
            expect_end(in, "Entities");
        } else if (tok == "$Nodes") {
            const int64_t nb = next_int<int64_t>(in, "$Nodes");
            (void)next_int<int64_t>(in, "$Nodes");
            (void)next_int<int64_t>(in, "$Nodes");
            (void)next_int<int64_t>(in, "$Nodes");
            for (int64_t b = 0; b < nb; b++) {
                block bl;
                bl.dim = next_int<int>(in, "$Nodes");
                bl.tag = next_int<int>(in, "$Nodes");
                const int param = next_int<int>(in, "$Nodes");
                const int64_t n = next_int<int64_t>(in, "$Nodes");
                bl.node_tags.resize((size_t)n);
                for (auto& t : bl.node_tags)
                    t = next_int<int64_t>(in, "$Nodes");
                bl.xyz.resize((size_t)(3 * n));
                const int extra = param ? bl.dim : 0;
                for (int64_t i = 0; i < n; i++) {
                    for (int q = 0; q < 3; q++)
                        bl.xyz[(size_t)(3 * i + q)] = next_double(in, "$Nodes");
                    for (int q = 0; q < extra; q++)
                        (void)next_double(in, "$Nodes");
                }
                nblocks.push_back(std::move(bl));
            }
            expect_end(in, "Nodes");
            have_nodes = true;
        } else if (tok == "$Elements") {
            const int64_t nb = next_int<int64_t>(in, "$Elements");
            (void)next_int<int64_t>(in, "$Elements");
            (void)next_int<int64_t>(in, "$Elements");
            (void)next_int<int64_t>(in, "$Elements");
            for (int64_t b = 0; b < nb; b++) {
                block bl;
                bl.dim = next_int<int>(in, "$Elements");
                bl.tag = next_int<int>(in, "$Elements");
                bl.etype = next_int<int>(in, "$Elements");
                const int64_t n = next_int<int64_t>(in, "$Elements");
                const int npe = nodes_per_element(bl.etype);
                if (npe < 0)
                    throw parse_error{"element type " + std::to_string(bl.etype) + " is not known"};
                bl.enodes.resize((size_t)(n * npe));
                for (int64_t i = 0; i < n; i++) {
                    (void)next_int<int64_t>(in, "$Elements"); // element tag
                    for (int q = 0; q < npe; q++)
                        bl.enodes[(size_t)(i * npe + q)] = next_int<int64_t>(in, "$Elements");
                }
                eblocks.push_back(std::move(bl));
            }
            expect_end(in, "Elements");
            have_elements = true;
        } else if (tok.size() > 1 && tok[0] == '$' && tok.rfind("$End", 0) != 0) {
            // a section this reader does not need ($Periodic, $NodeData, ...): skip to its end
            const std::string end = "$End" + tok.substr(1);
            std::string s;
            while (in >> s && s != end) {
            }
        }
    }
    if (!have_format || !have_nodes || !have_elements)
        throw parse_error{"missing $MeshFormat, $Nodes or $Elements"};

    auto* m = new frico_msh();
    // getNodes(-1, -1): entities by (dim, tag), nodes of each in stored order
    std::stable_sort(nblocks.begin(), nblocks.end(),
                     [](const block& a, const block& b) { return a.dim != b.dim ? a.dim < b.dim : a.tag < b.tag; });
    std::map<int64_t, int64_t> tag2ofs;
    std::vector<double> verts;
    for (const auto& bl : nblocks)
        for (size_t i = 0; i < bl.node_tags.size(); i++) {
            tag2ofs[bl.node_tags[i]] = (int64_t)(verts.size() / 3);
            verts.insert(verts.end(), bl.xyz.begin() + (long)(3 * i), bl.xyz.begin() + (long)(3 * i + 3));
        }
    auto ofs = [&](int64_t t) {
        const auto it = tag2ofs.find(t);
        if (it == tag2ofs.end())
            throw parse_error{"element references node " + std::to_string(t) + ", which is not defined"};
        return it->second;
    };
    // elements by entity tag (getEntities(d) is ordered by tag), each entity's blocks in stored order
    std::stable_sort(eblocks.begin(), eblocks.end(),
                     [](const block& a, const block& b) { return a.dim != b.dim ? a.dim < b.dim : a.tag < b.tag; });
    std::vector<int64_t> T, B;
    for (const auto& bl : eblocks) {
        const int npe = nodes_per_element(bl.etype);
        const size_t n = bl.enodes.size() / (size_t)npe;
        if (bl.dim == 2 && is_triangle(bl.etype)) {
            for (size_t i = 0; i < n; i++) {
                for (int q = 0; q < 3; q++)
                    T.push_back(ofs(bl.enodes[i * npe + q]));
                m->ttag.push_back(bl.tag);
            }
        } else if (bl.dim == 1 && is_line(bl.etype)) {
            for (size_t i = 0; i < n; i++) {
                for (int q = 0; q < 2; q++)
                    B.push_back(ofs(bl.enodes[i * npe + q]));
                m->btag.push_back(bl.tag);
            }
        } else if (bl.dim == 2) {
            throw parse_error{"surface " + std::to_string(bl.tag) + " holds elements of type " +
                              std::to_string(bl.etype) + ", not triangles"};
        }
    }
    // vertices referenced by triangles, renumbered in their original order
    const int64_t nv0 = (int64_t)(verts.size() / 3);
    std::vector<int64_t> used((size_t)nv0, -1);
    for (auto v : T)
        used[(size_t)v] = 1;
    int64_t ci = 0;
    for (auto& u : used)
        if (u >= 0)
            u = ci++;
    m->V.resize((size_t)(3 * ci));
    for (int64_t i = 0; i < nv0; i++)
        if (used[(size_t)i] >= 0)
            for (int q = 0; q < 3; q++)
                m->V[(size_t)(3 * used[(size_t)i] + q)] = verts[(size_t)(3 * i + q)];
    m->T.resize(T.size());
    for (size_t i = 0; i < T.size(); i++)
        m->T[i] = used[(size_t)T[i]];
    // a line element on vertices no triangle uses cannot be an edge of the mesh (the reference
    // reports such an edge as lonely before compacting, :177-185)
    for (size_t i = 0; i < B.size(); i++) {
        if (used[(size_t)B[i]] < 0) {
            const int tag = m->btag[i / 2];
            delete m;
            throw parse_error{"lonely edge on curve " + std::to_string(tag)};
        }
        m->B.push_back(used[(size_t)B[i]]);
    }
    // physical groups: getPhysicalGroups, named, with the entities carrying them
    std::map<std::string, frico_msh::group> byname;
    for (int d = 0; d < 4; d++) {
        std::map<int, std::vector<int32_t>> ents;
        for (const auto& [etag, phys] : ent_phys[d])
            for (int p : phys)
                ents[p].push_back(etag);
        for (const auto& [ptag, elist] : ents) {
            frico_msh::group g;
            const auto nm = pnames[d].find(ptag);
            g.name = nm != pnames[d].end() ? nm->second : "";
            g.dim = d;
            g.tag = ptag;
            g.entities = elist;
            byname[g.name] = g; // std::map by name: a later group of the same name replaces it
        }
    }
    for (auto& [name, g] : byname)
        m->groups.push_back(g);
    return m;
}

} // namespace frico

extern "C" {

int frico_msh_read(const char* path, frico_msh** out)
{
    using namespace frico;
    if (!path)
        return set_error("null path"), -1;
    if (!out)
        return set_error("null output"), -2;
    *out = nullptr;
    std::ifstream in(path);
    if (!in)
        return set_error("can't open '%s'", path), FRICO_ERR_MESH;
    try {
        *out = parse_msh(in);
    } catch (const parse_error& e) {
        return set_error("%s: %s", path, e.what.c_str()), FRICO_ERR_MESH;
    } catch (const std::bad_alloc&) {
        return set_error("%s: out of memory", path), FRICO_ERR_NOMEM;
    }
    return 0;
}

int frico_msh_free(frico_msh* m)
{
    delete m;
    return 0;
}

int frico_msh_sizes(const frico_msh* m, int64_t* nv, int64_t* nt, int64_t* nb, int32_t* ngroups)
{
    if (!m)
        return frico::set_error("null mesh"), -1;
    if (nv)
        *nv = (int64_t)(m->V.size() / 3);
    if (nt)
        *nt = (int64_t)m->ttag.size();
    if (nb)
        *nb = (int64_t)m->btag.size();
    if (ngroups)
        *ngroups = (int32_t)m->groups.size();
    return 0;
}

int frico_msh_arrays(const frico_msh* m, double* vertices, int64_t* triangles, int32_t* tags, int64_t* bedges,
                     int32_t* btags)
{
    if (!m)
        return frico::set_error("null mesh"), -1;
    if (vertices)
        std::copy(m->V.begin(), m->V.end(), vertices);
    if (triangles)
        std::copy(m->T.begin(), m->T.end(), triangles);
    if (tags)
        std::copy(m->ttag.begin(), m->ttag.end(), tags);
    if (bedges)
        std::copy(m->B.begin(), m->B.end(), bedges);
    if (btags)
        std::copy(m->btag.begin(), m->btag.end(), btags);
    return 0;
}

int frico_msh_group(const frico_msh* m, int32_t i, char* name, int64_t name_size, int32_t* dim, int32_t* tag,
                    int32_t* entities, int32_t max_entities, int32_t* nentities)
{
    if (!m)
        return frico::set_error("null mesh"), -1;
    if (i < 0 || i >= (int32_t)m->groups.size())
        return frico::set_error("group %d out of range", i), -2;
    const auto& g = m->groups[(size_t)i];
    if (name && name_size > 0)
        std::snprintf(name, (size_t)name_size, "%s", g.name.c_str());
    if (dim)
        *dim = g.dim;
    if (tag)
        *tag = g.tag;
    if (nentities)
        *nentities = (int32_t)g.entities.size();
    if (entities) {
        if (max_entities < (int32_t)g.entities.size())
            return frico::set_error("entity buffer holds %d, the group has %d", max_entities, (int)g.entities.size()), -8;
        std::copy(g.entities.begin(), g.entities.end(), entities);
    }
    return 0;
}

} // extern "C"
