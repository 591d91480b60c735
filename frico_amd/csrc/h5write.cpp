// HDF5 outputs in the layout the reference writes through HighFive/H5Easy:
//   antenna analysis:  write_hdf5 (src/emw_ana_antenna.cpp:207-234, man/frico.1 "HDF5 OUTPUT FORMAT")
//   -d dump:           /frico/Z, /frico/V, /frico/I (src/emw_ana_antenna.cpp:260-268; the bistatic
//                      sweep's dump, src/emw_solver.cpp:494-500, has no I)
// No HDF5 library is used.  The file is written with the format's original structures, the ones
// the reference's own output (share/hdf5/antenna_analysis/biquad.h5) uses: superblock version 0
// with 8-byte offsets and lengths; version-1 object headers; groups as symbol tables (a version-1
// B-tree of symbol table nodes, names in a local heap); datasets with a version-1 dataspace, a
// version-1 datatype (64-bit unsigned, IEEE double, or the {r, i} compound HighFive uses for
// std::complex<double>), a version-2 fill value message and version-3 contiguous storage.  The
// structures follow the HDF5 File Format Specification 2.0.  Everything is laid out first
// (addresses of every header, node, heap and data block), then written in one pass; data blocks
// come last and are streamed from their sources, so a dump of Z is never copied whole.
#include "../../include/frico_gpu.h"
#include "common.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace frico {
namespace {

constexpr uint64_t UNDEF = ~(uint64_t)0;
constexpr int INTERNAL_K = 16; // group B-tree node: up to 2K children

enum class H5T { U64, F64, C128 };

struct H5Node {
    std::string name;
    bool group = true;
    std::vector<std::unique_ptr<H5Node>> kids;
    // dataset
    H5T type = H5T::F64;
    std::vector<uint64_t> dims; // empty: scalar
    uint64_t nbytes = 0;
    std::function<bool(std::FILE*)> write; // writes exactly nbytes
    // layout
    uint64_t ohdr = 0, btree = 0, heap = 0, heap_data = 0, snod = 0, data = 0;
    uint64_t heap_size = 0;
    std::vector<uint64_t> name_off; // heap offsets of the kids' names (sorted order)
};

struct Buf {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void u16(uint16_t v) { put(&v, 2); }
    void u32(uint32_t v) { put(&v, 4); }
    void u64(uint64_t v) { put(&v, 8); }
    void put(const void* p, size_t n)
    {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        b.insert(b.end(), c, c + n);
    }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad8() { zeros((8 - b.size() % 8) % 8); }
};

uint64_t up8(uint64_t v) { return (v + 7) / 8 * 8; }

// ---- messages ---------------------------------------------------------------------------------
void dtype_f64(Buf& m)
{
    // class 1 (floating point), version 1; little endian, implied mantissa msb, sign at bit 63
    m.u8(0x11);
    m.u8(0x20);
    m.u8(63);
    m.u8(0);
    m.u32(8);
    m.u16(0);  // bit offset
    m.u16(64); // precision
    m.u8(52);  // exponent location
    m.u8(11);  // exponent size
    m.u8(0);   // mantissa location
    m.u8(52);  // mantissa size
    m.u32(1023);
}

void dtype_message(Buf& m, H5T t)
{
    if (t == H5T::U64) {
        // class 0 (fixed point), version 1; little endian, unsigned
        m.u8(0x10);
        m.u8(0);
        m.u8(0);
        m.u8(0);
        m.u32(8);
        m.u16(0);
        m.u16(64);
    } else if (t == H5T::F64) {
        dtype_f64(m);
    } else {
        // class 6 (compound), version 1, two members of 8 bytes: "r" at 0, "i" at 8
        m.u8(0x16);
        m.u8(2);
        m.u8(0);
        m.u8(0);
        m.u32(16);
        const char* names[2] = {"r", "i"};
        for (int k = 0; k < 2; k++) {
            Buf nm;
            nm.put(names[k], 2);
            nm.pad8();
            m.put(nm.b.data(), nm.b.size());
            m.u32(8 * k);   // byte offset
            m.u8(0);        // dimensionality
            m.zeros(3 + 4 + 4 + 16);
            dtype_f64(m);
        }
    }
}

void message(Buf& hdr, uint16_t type, uint8_t flags, const Buf& data)
{
    Buf d = data;
    d.pad8();
    hdr.u16(type);
    hdr.u16((uint16_t)d.b.size());
    hdr.u8(flags);
    hdr.zeros(3);
    hdr.put(d.b.data(), d.b.size());
}

// object header (version 1) from its message block
void object_header(Buf& out, uint16_t nmsg, const Buf& msgs)
{
    out.u8(1);
    out.u8(0);
    out.u16(nmsg);
    out.u32(1);
    out.u32((uint32_t)msgs.b.size());
    out.u32(0);
    out.put(msgs.b.data(), msgs.b.size());
}

Buf dataset_messages(const H5Node& d)
{
    Buf msgs;
    Buf sp;
    sp.u8(1);
    sp.u8((uint8_t)d.dims.size());
    sp.u8(d.dims.empty() ? 0 : 1); // maximum dimensions present
    sp.zeros(5);
    for (uint64_t v : d.dims)
        sp.u64(v);
    for (uint64_t v : d.dims)
        sp.u64(v);
    message(msgs, 0x0001, 0, sp);
    Buf ty;
    dtype_message(ty, d.type);
    message(msgs, 0x0003, 1, ty);
    Buf fill;
    fill.u8(2); // version
    fill.u8(2); // space allocation time: late
    fill.u8(2); // fill value write time: if set
    fill.u8(1); // fill value defined (default)
    fill.u32(0);
    message(msgs, 0x0005, 1, fill);
    Buf lay;
    lay.u8(3); // version
    lay.u8(1); // contiguous
    lay.u64(d.data);
    lay.u64(d.nbytes);
    message(msgs, 0x0008, 0, lay);
    return msgs;
}

uint64_t dataset_header_size(const H5Node& d)
{
    const uint64_t space = 8 + 16 * d.dims.size();
    const uint64_t type = d.type == H5T::U64 ? 16 : (d.type == H5T::F64 ? 24 : 128);
    return 16 + (8 + space) + (8 + type) + (8 + 8) + (8 + 24);
}

class H5Writer {
public:
    H5Writer() { root.name = ""; }

    // the group at path (created with its parents)
    H5Node* group(const std::string& path)
    {
        H5Node* g = &root;
        size_t p = 0;
        while (p < path.size()) {
            size_t q = path.find('/', p);
            if (q == std::string::npos)
                q = path.size();
            const std::string part = path.substr(p, q - p);
            p = q + 1;
            if (part.empty())
                continue;
            H5Node* next = nullptr;
            for (auto& k : g->kids)
                if (k->name == part)
                    next = k.get();
            if (!next) {
                g->kids.push_back(std::make_unique<H5Node>());
                next = g->kids.back().get();
                next->name = part;
            }
            g = next;
        }
        return g;
    }

    void dataset(const std::string& path, H5T type, std::vector<uint64_t> dims, std::function<bool(std::FILE*)> write)
    {
        const size_t s = path.rfind('/');
        H5Node* g = group(path.substr(0, s));
        g->kids.push_back(std::make_unique<H5Node>());
        H5Node* d = g->kids.back().get();
        d->name = path.substr(s + 1);
        d->group = false;
        d->type = type;
        d->dims = std::move(dims);
        uint64_t count = 1;
        for (uint64_t v : d->dims)
            count *= v;
        d->nbytes = count * (type == H5T::C128 ? 16 : 8);
        d->write = std::move(write);
    }

    // a small dataset from bytes held here
    void dataset_bytes(const std::string& path, H5T type, std::vector<uint64_t> dims, const void* p, size_t n)
    {
        auto data = std::make_shared<std::vector<uint8_t>>(static_cast<const uint8_t*>(p),
                                                          static_cast<const uint8_t*>(p) + n);
        dataset(path, type, std::move(dims),
                [data](std::FILE* f) { return std::fwrite(data->data(), 1, data->size(), f) == data->size(); });
    }

    int write(const char* path)
    {
        // leaf K: every group's symbol nodes must fit one B-tree node (2 INTERNAL_K children)
        size_t most = 0;
        count_kids(root, most);
        leaf_k = 4;
        while ((most + 2 * leaf_k - 1) / (2 * leaf_k) > 2 * INTERNAL_K)
            leaf_k *= 2;
        uint64_t at = 96; // superblock
        layout_meta(root, at);
        layout_data(root, at);
        eof = at;
        std::FILE* f = std::fopen(path, "wb");
        if (!f)
            return set_error("Can't open '%s' for writing", path), FRICO_ERR_STATE;
        Buf sb;
        superblock(sb);
        bool ok = std::fwrite(sb.b.data(), 1, sb.b.size(), f) == sb.b.size();
        uint64_t pos = 96;
        ok = ok && emit_meta(root, f, pos) && emit_data(root, f, pos) && pos == eof;
        ok = std::fclose(f) == 0 && ok;
        if (!ok)
            return set_error("writing '%s' failed", path), FRICO_ERR_STATE;
        return 0;
    }

private:
    H5Node root;
    int leaf_k = 4;
    uint64_t eof = 0;

    static void count_kids(const H5Node& g, size_t& most)
    {
        if (!g.group)
            return;
        most = std::max(most, g.kids.size());
        for (auto& k : g.kids)
            count_kids(*k, most);
    }

    uint64_t btree_size() const { return 24 + (2 * INTERNAL_K) * 8 + (2 * INTERNAL_K + 1) * 8; }
    uint64_t snod_size() const { return 8 + (uint64_t)(2 * leaf_k) * 40; }
    size_t nsnod(const H5Node& g) const { return std::max<size_t>(1, (g.kids.size() + 2 * leaf_k - 1) / (2 * leaf_k)); }

    void layout_meta(H5Node& g, uint64_t& at)
    {
        if (!g.group) {
            g.ohdr = at;
            at += up8(dataset_header_size(g));
            return;
        }
        std::sort(g.kids.begin(), g.kids.end(),
                  [](const std::unique_ptr<H5Node>& a, const std::unique_ptr<H5Node>& b) { return a->name < b->name; });
        g.ohdr = at;
        at += 16 + 8 + 16;
        g.btree = at;
        at += btree_size();
        g.heap = at;
        at += 32;
        g.heap_data = at;
        uint64_t hs = 8; // the empty name at offset 0
        g.name_off.clear();
        for (auto& k : g.kids) {
            g.name_off.push_back(hs);
            hs += up8(k->name.size() + 1);
        }
        g.heap_size = hs;
        at += hs;
        g.snod = at;
        at += nsnod(g) * snod_size();
        for (auto& k : g.kids)
            layout_meta(*k, at);
    }

    void layout_data(H5Node& g, uint64_t& at)
    {
        if (!g.group) {
            g.data = at;
            at += up8(g.nbytes);
            return;
        }
        for (auto& k : g.kids)
            layout_data(*k, at);
    }

    void entry(Buf& b, uint64_t name_off, const H5Node& n) const
    {
        b.u64(name_off);
        b.u64(n.ohdr);
        if (n.group) {
            b.u32(1);
            b.u32(0);
            b.u64(n.btree);
            b.u64(n.heap);
        } else {
            b.u32(0);
            b.u32(0);
            b.zeros(16);
        }
    }

    void superblock(Buf& b) const
    {
        const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        b.put(sig, 8);
        b.u8(0); // superblock version
        b.u8(0); // free-space version
        b.u8(0); // root group symbol table entry version
        b.u8(0);
        b.u8(0); // shared header message format version
        b.u8(8); // size of offsets
        b.u8(8); // size of lengths
        b.u8(0);
        b.u16((uint16_t)leaf_k);
        b.u16((uint16_t)INTERNAL_K);
        b.u32(0);
        b.u64(0);     // base address
        b.u64(UNDEF); // free-space info
        b.u64(eof);
        b.u64(UNDEF); // driver information
        entry(b, 0, root);
    }

    static bool put(std::FILE* f, const Buf& b, uint64_t& pos)
    {
        pos += b.b.size();
        return std::fwrite(b.b.data(), 1, b.b.size(), f) == b.b.size();
    }

    bool emit_meta(const H5Node& g, std::FILE* f, uint64_t& pos) const
    {
        if (!g.group) {
            if (pos != g.ohdr)
                return false;
            Buf h;
            object_header(h, 4, dataset_messages(g));
            h.pad8();
            return put(f, h, pos);
        }
        if (pos != g.ohdr)
            return false;
        Buf h;
        {
            Buf m, st;
            st.u64(g.btree);
            st.u64(g.heap);
            message(m, 0x0011, 0, st);
            object_header(h, 1, m);
        }
        // B-tree (group node, level 0): key 0 = "", key i+1 = last name of symbol node i
        const size_t ns = nsnod(g), per = 2 * leaf_k;
        h.put("TREE", 4);
        h.u8(0);
        h.u8(0);
        h.u16((uint16_t)ns);
        h.u64(UNDEF);
        h.u64(UNDEF);
        h.u64(0);
        for (size_t i = 0; i < ns; i++) {
            h.u64(g.snod + i * snod_size());
            const size_t last = std::min(g.kids.size(), (i + 1) * per);
            h.u64(last > 0 ? g.name_off[last - 1] : 0);
        }
        h.zeros(btree_size() - (24 + 8 + ns * 16));
        // local heap (no free block: offset 1 ends the free list)
        h.put("HEAP", 4);
        h.u8(0);
        h.zeros(3);
        h.u64(g.heap_size);
        h.u64(1);
        h.u64(g.heap_data);
        h.zeros(8);
        for (auto& k : g.kids) {
            Buf nm;
            nm.put(k->name.c_str(), k->name.size() + 1);
            nm.pad8();
            h.put(nm.b.data(), nm.b.size());
        }
        // symbol table nodes
        for (size_t i = 0; i < ns; i++) {
            const size_t a = i * per, e = std::min(g.kids.size(), a + per);
            h.put("SNOD", 4);
            h.u8(1);
            h.u8(0);
            h.u16((uint16_t)(e - a));
            for (size_t k = a; k < e; k++)
                entry(h, g.name_off[k], *g.kids[k]);
            h.zeros((per - (e - a)) * 40);
        }
        if (!put(f, h, pos))
            return false;
        for (auto& k : g.kids)
            if (!emit_meta(*k, f, pos))
                return false;
        return true;
    }

    bool emit_data(const H5Node& g, std::FILE* f, uint64_t& pos) const
    {
        if (!g.group) {
            if (pos != g.data || !g.write(f))
                return false;
            pos += g.nbytes;
            const uint64_t pad = up8(g.nbytes) - g.nbytes;
            static const uint8_t z[8] = {};
            pos += pad;
            return std::fwrite(z, 1, pad, f) == pad;
        }
        for (auto& k : g.kids)
            if (!emit_data(*k, f, pos))
                return false;
        return true;
    }
};

std::string num(int64_t v) { return std::to_string(v); }

} // namespace
} // namespace frico

extern "C" {

int frico_write_antenna_h5(const char* path, double rdiag_dist, double Z0, int64_t nsteps, const double* freqs,
                           double v_re, double v_im, const double* port, const double* gain)
{
    using namespace frico;
    if (!path)
        return -1;
    if (nsteps < 0)
        return -4;
    if (nsteps > 0 && (!freqs || !port || !gain))
        return !freqs ? -5 : (!port ? -8 : -9);
    H5Writer w;
    const std::string base = "/frico/antenna_analysis/";
    w.dataset_bytes(base + "rdiag_dist", H5T::F64, {}, &rdiag_dist, 8);
    const uint64_t steps = (uint64_t)nsteps;
    w.dataset_bytes(base + "sweep_steps", H5T::U64, {}, &steps, 8);
    w.dataset_bytes(base + "Z0", H5T::F64, {}, &Z0, 8);
    const double volt[2] = {v_re, v_im};
    for (int64_t s = 0; s < nsteps; s++) {
        const std::string p = base + num(s) + "/";
        const double* po = port + 10 * s; // I, P, Zin, Gamma (re, im), SWR, residual
        const double* g = gain + 1080 * s;  // xy, yz, xz
        w.dataset_bytes(p + "frequency", H5T::F64, {}, freqs + s, 8);
        w.dataset_bytes(p + "gain_XY", H5T::F64, {360}, g, 360 * 8);
        w.dataset_bytes(p + "gain_YZ", H5T::F64, {360}, g + 360, 360 * 8);
        w.dataset_bytes(p + "gain_XZ", H5T::F64, {360}, g + 720, 360 * 8);
        w.dataset_bytes(p + "port/voltage", H5T::C128, {}, volt, 16);
        w.dataset_bytes(p + "port/current", H5T::C128, {}, po, 16);
        w.dataset_bytes(p + "port/power", H5T::C128, {}, po + 2, 16);
        w.dataset_bytes(p + "port/impedance", H5T::C128, {}, po + 4, 16);
        w.dataset_bytes(p + "port/gamma", H5T::C128, {}, po + 6, 16);
        w.dataset_bytes(p + "port/swr", H5T::F64, {}, po + 8, 8);
    }
    return w.write(path);
}

int frico_write_linear_system_h5(const char* path, int64_t n, const double* Z, int64_t ldz, int64_t nrhs,
                                 const double* V, int64_t ldv, const double* I, int64_t ldi)
{
    using namespace frico;
    if (!path)
        return -1;
    if (n < 0)
        return -2;
    if (n > 0 && !Z)
        return -3;
    if (ldz < n)
        return -4;
    if (nrhs < 1)
        return -5;
    if (!V)
        return -6;
    if (ldv < n)
        return -7;
    if (I && ldi < n)
        return -9;
    H5Writer w;
    // Z(i, j) at Z[2 (i + j ldz)]: written row by row, entry [i][j] = Z(i, j)
    w.dataset("/frico/Z", H5T::C128, {(uint64_t)n, (uint64_t)n}, [=](std::FILE* f) {
        std::vector<double> row(2 * (size_t)n);
        for (int64_t i = 0; i < n; i++) {
            for (int64_t j = 0; j < n; j++) {
                row[2 * j] = Z[2 * (i + j * ldz)];
                row[2 * j + 1] = Z[2 * (i + j * ldz) + 1];
            }
            if (std::fwrite(row.data(), 16, (size_t)n, f) != (size_t)n)
                return false;
        }
        return true;
    });
    // V, I: n x nrhs (a column for a single right-hand side, as the reference's vectors), [i][q]
    auto columns = [n, nrhs](const double* X, int64_t ldx) {
        return [=](std::FILE* f) {
            std::vector<double> row(2 * (size_t)nrhs);
            for (int64_t i = 0; i < n; i++) {
                for (int64_t q = 0; q < nrhs; q++) {
                    row[2 * q] = X[2 * (i + q * ldx)];
                    row[2 * q + 1] = X[2 * (i + q * ldx) + 1];
                }
                if (std::fwrite(row.data(), 16, (size_t)nrhs, f) != (size_t)nrhs)
                    return false;
            }
            return true;
        };
    };
    w.dataset("/frico/V", H5T::C128, {(uint64_t)n, (uint64_t)nrhs}, columns(V, ldv));
    if (I)
        w.dataset("/frico/I", H5T::C128, {(uint64_t)n, (uint64_t)nrhs}, columns(I, ldi));
    return w.write(path);
}

int frico_dump_linear_system(frico_ctx* ctx, const char* path)
{
    using namespace frico;
    if (!ctx)
        return -1;
    if (!path)
        return -2;
    const int64_t n = frico_num_unknowns(ctx), nrhs = frico_num_rhs(ctx);
    if (n < 1 || nrhs < 1)
        return set_error("dump: the context has no matrix or no right-hand side"), FRICO_ERR_STATE;
    std::vector<double> Z(2 * (size_t)n * (size_t)n), V(2 * (size_t)n * (size_t)nrhs), I(V.size());
    int rc = frico_matrix_get(ctx, Z.data(), n);
    if (!rc)
        rc = frico_rhs_get(ctx, V.data(), n);
    if (rc)
        return rc;
    const bool have_i = frico_solution_get(ctx, I.data(), n) == 0;
    return frico_write_linear_system_h5(path, n, Z.data(), n, nrhs, V.data(), n, have_i ? I.data() : nullptr, n);
}

} // extern "C"
