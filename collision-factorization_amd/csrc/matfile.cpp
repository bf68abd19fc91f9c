// Reader for float64 datasets of HDF5 files, which is what a MATLAB v7.3 .mat file is (the
// reference's precalc tensors, src/precalc/*.mat, written by precompute_collision_operator.m:61).
//
// Supported: superblock versions 0 and 1 (after an optional user block, found at 0, 512, 1024, ...),
// groups stored as symbol tables (v1 B-tree of SNOD nodes + local heap), version-1 object headers
// with continuation blocks, dataspace v1/v2, data layout v3 (compact, contiguous, chunked with a v1
// B-tree) and the deflate and shuffle filters (zlib).  That is the subset MATLAB's v7.3 writer uses.
// Everything else (other superblocks, fractal-heap groups, other filters, non-float64 data) is
// reported as unsupported rather than guessed.  Every offset read from the file is bounds-checked
// against the file size.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfk.h"

namespace cfk {
int set_error(int code, const char *fmt, ...);  // api.hip

namespace {

struct H5Error {
    std::string msg;
};

[[noreturn]] void fail(const std::string &m) { throw H5Error{m}; }

struct File {
    std::vector<uint8_t> buf;
    uint64_t base = 0;       // absolute position of the superblock; file addresses are relative to it
    int so = 8, sl = 8;      // size of offsets / lengths

    void need(uint64_t abs, uint64_t n) const {
        if (abs > buf.size() || n > buf.size() - abs) fail("read past the end of the file");
    }
    uint64_t u(uint64_t abs, int n) const {
        need(abs, (uint64_t)n);
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) v |= (uint64_t)buf[abs + i] << (8 * i);
        return v;
    }
    uint64_t addr(uint64_t rel) const { return base + rel; }
    bool undef(uint64_t a) const { return so == 8 ? a == ~0ull : a == ((1ull << (8 * so)) - 1); }
};

struct Message {
    int type;
    uint64_t pos;  // absolute position of the message data
    uint64_t size;
};

// Version-1 object header: messages of the header block and of its continuation blocks.
std::vector<Message> object_header(const File &f, uint64_t rel) {
    uint64_t h = f.addr(rel);
    if (f.u(h, 1) != 1) fail("object header version other than 1");
    const uint64_t nmsg = f.u(h + 2, 2);
    const uint64_t hsize = f.u(h + 8, 4);
    std::vector<std::pair<uint64_t, uint64_t>> blocks{{h + 16, hsize}};
    std::vector<Message> out;
    for (size_t b = 0; b < blocks.size() && out.size() < nmsg; ++b) {
        uint64_t p = blocks[b].first;
        const uint64_t end = blocks[b].first + blocks[b].second;
        f.need(blocks[b].first, blocks[b].second);
        while (p + 8 <= end && out.size() < nmsg) {
            const int type = (int)f.u(p, 2);
            const uint64_t size = f.u(p + 2, 2);
            if (p + 8 + size > end) fail("object header message overruns its block");
            if (type == 0x10) {  // continuation
                blocks.push_back({f.addr(f.u(p + 8, f.so)), f.u(p + 8 + f.so, f.sl)});
                if (blocks.size() > 4096) fail("too many object header continuation blocks");
            }
            out.push_back({type, p + 8, size});
            p += 8 + size;
        }
    }
    return out;
}

const Message *find_msg(const std::vector<Message> &m, int type) {
    for (const auto &x : m)
        if (x.type == type) return &x;
    return nullptr;
}

std::string heap_string(const File &f, uint64_t heap_rel, uint64_t off) {
    const uint64_t h = f.addr(heap_rel);
    f.need(h, 8);
    if (std::memcmp(&f.buf[h], "HEAP", 4) != 0) fail("local heap signature missing");
    const uint64_t seg_size = f.u(h + 8, f.sl);
    const uint64_t seg = f.addr(f.u(h + 8 + 2 * f.sl, f.so));
    if (off >= seg_size) fail("link name outside the local heap");
    f.need(seg, seg_size);
    std::string s;
    for (uint64_t i = seg + off; i < seg + seg_size && f.buf[i]; ++i) s.push_back((char)f.buf[i]);
    return s;
}

// Object header address of the link `name` in a symbol-table group (v1 B-tree of SNOD nodes).
bool group_lookup(const File &f, uint64_t btree_rel, uint64_t heap_rel, const std::string &name, uint64_t &out,
                  int depth = 0) {
    if (depth > 64) fail("group B-tree too deep");
    const uint64_t n = f.addr(btree_rel);
    f.need(n, 8);
    if (std::memcmp(&f.buf[n], "TREE", 4) != 0) fail("group B-tree signature missing");
    if (f.u(n + 4, 1) != 0) fail("group B-tree of the wrong node type");
    const int level = (int)f.u(n + 5, 1);
    const int used = (int)f.u(n + 6, 2);
    uint64_t p = n + 8 + 2 * f.so + f.sl;  // past both siblings and key 0
    for (int i = 0; i < used; ++i, p += f.so + f.sl) {
        const uint64_t child = f.u(p, f.so);
        if (level > 0) {
            if (group_lookup(f, child, heap_rel, name, out, depth + 1)) return true;
            continue;
        }
        const uint64_t s = f.addr(child);
        f.need(s, 8);
        if (std::memcmp(&f.buf[s], "SNOD", 4) != 0) fail("symbol table node signature missing");
        const int nsym = (int)f.u(s + 6, 2);
        const uint64_t esize = 2 * (uint64_t)f.so + 24;
        for (int k = 0; k < nsym; ++k) {
            const uint64_t e = s + 8 + k * esize;
            if (heap_string(f, heap_rel, f.u(e, f.so)) == name) {
                out = f.u(e + f.so, f.so);
                return true;
            }
        }
    }
    return false;
}

// Follow a '/'-separated path from the root group to an object header address.
uint64_t resolve(const File &f, uint64_t root_oh, const std::string &path) {
    uint64_t oh = root_oh;
    size_t i = 0;
    while (i < path.size()) {
        while (i < path.size() && path[i] == '/') ++i;
        if (i >= path.size()) break;
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        const std::string part = path.substr(i, j - i);
        const auto msgs = object_header(f, oh);
        const Message *st = find_msg(msgs, 0x11);
        if (!st) fail("'" + part + "': parent is not a symbol-table group");
        uint64_t next = 0;
        if (!group_lookup(f, f.u(st->pos, f.so), f.u(st->pos + f.so, f.so), part, next))
            fail("no object '" + part + "' in the group");
        oh = next;
        i = j;
    }
    return oh;
}

struct Filter {
    int id;
    std::vector<uint32_t> client;
};

std::vector<Filter> filters(const File &f, const Message *m) {
    std::vector<Filter> out;
    if (!m) return out;
    const int ver = (int)f.u(m->pos, 1), nf = (int)f.u(m->pos + 1, 1);
    uint64_t p = m->pos + (ver == 1 ? 8 : 2);
    for (int i = 0; i < nf; ++i) {
        Filter flt;
        flt.id = (int)f.u(p, 2);
        uint64_t name_len = 0;
        if (ver == 1 || flt.id >= 256) {
            name_len = f.u(p + 2, 2);
            p += 4;
        } else {
            p += 2;
        }
        p += 2;  // flags
        const int nc = (int)f.u(p, 2);
        p += 2;
        if (ver == 1) name_len = (name_len + 7) / 8 * 8;
        p += name_len;
        for (int c = 0; c < nc; ++c, p += 4) flt.client.push_back((uint32_t)f.u(p, 4));
        if (ver == 1 && (nc & 1)) p += 4;
        out.push_back(flt);
    }
    return out;
}

// Undo the filter pipeline (applied in reverse order) unless the chunk's mask skips a filter.
std::vector<uint8_t> unfilter(const std::vector<Filter> &fl, uint32_t mask, std::vector<uint8_t> data,
                              size_t expect, int elem) {
    for (int i = (int)fl.size() - 1; i >= 0; --i) {
        if (mask & (1u << i)) continue;
        if (fl[i].id == 1) {  // deflate
            std::vector<uint8_t> out(expect);
            uLongf n = (uLongf)expect;
            if (uncompress(out.data(), &n, data.data(), (uLong)data.size()) != Z_OK || n != expect)
                fail("deflate chunk did not inflate to the chunk size");
            data.swap(out);
        } else if (fl[i].id == 2) {  // shuffle: byte planes back to elements
            const size_t ne = data.size() / elem;
            std::vector<uint8_t> out(data.size());
            for (size_t e = 0; e < ne; ++e)
                for (int b = 0; b < elem; ++b) out[e * elem + b] = data[b * ne + e];
            data.swap(out);
        } else if (fl[i].id == 3) {  // fletcher32: drop the trailing checksum
            if (data.size() < 4) fail("fletcher32 chunk too short");
            data.resize(data.size() - 4);
        } else {
            fail("unsupported HDF5 filter " + std::to_string(fl[i].id));
        }
    }
    if (data.size() != expect) fail("chunk size after filters does not match the chunk");
    return data;
}

void read_chunks(const File &f, uint64_t btree_rel, int rank, const std::vector<uint64_t> &dims,
                 const std::vector<uint64_t> &cdims, const std::vector<Filter> &fl, double *out, int depth = 0) {
    if (depth > 64) fail("chunk B-tree too deep");
    const uint64_t n = f.addr(btree_rel);
    f.need(n, 8);
    if (std::memcmp(&f.buf[n], "TREE", 4) != 0) fail("chunk B-tree signature missing");
    if (f.u(n + 4, 1) != 1) fail("chunk B-tree of the wrong node type");
    const int level = (int)f.u(n + 5, 1);
    const int used = (int)f.u(n + 6, 2);
    const uint64_t key_size = 8 + 8 * (uint64_t)(rank + 1);
    uint64_t p = n + 8 + 2 * f.so;
    size_t celems = 1;
    for (int d = 0; d < rank; ++d) celems *= cdims[d];
    for (int i = 0; i < used; ++i) {
        const uint64_t csize = f.u(p, 4);
        const uint32_t mask = (uint32_t)f.u(p + 4, 4);
        std::vector<uint64_t> off(rank);
        for (int d = 0; d < rank; ++d) off[d] = f.u(p + 8 + 8 * d, 8);
        const uint64_t child = f.u(p + key_size, f.so);
        p += key_size + f.so;
        if (level > 0) {
            read_chunks(f, child, rank, dims, cdims, fl, out, depth + 1);
            continue;
        }
        const uint64_t a = f.addr(child);
        f.need(a, csize);
        std::vector<uint8_t> raw(f.buf.begin() + a, f.buf.begin() + a + csize);
        const std::vector<uint8_t> data = unfilter(fl, mask, std::move(raw), celems * 8, 8);
        // scatter the chunk (C order, edge chunks padded) into the dataset
        std::vector<uint64_t> c(rank, 0);
        for (size_t e = 0; e < celems; ++e) {
            bool inside = true;
            uint64_t lin = 0;
            for (int d = 0; d < rank; ++d) {
                const uint64_t x = off[d] + c[d];
                if (x >= dims[d]) inside = false;
                lin = lin * dims[d] + x;
            }
            if (inside) std::memcpy(out + lin, &data[e * 8], 8);
            for (int d = rank - 1; d >= 0; --d) {
                if (++c[d] < cdims[d]) break;
                c[d] = 0;
            }
        }
    }
}

struct Dataset {
    std::vector<uint64_t> dims;
    uint64_t numel = 1;
};

Dataset read_dataset(const File &f, uint64_t oh, double *out, int64_t capacity) {
    const auto msgs = object_header(f, oh);
    const Message *ds = find_msg(msgs, 0x01), *dt = find_msg(msgs, 0x03), *ly = find_msg(msgs, 0x08);
    if (!ds || !dt || !ly) fail("object is not a dataset");
    Dataset D;
    const int dver = (int)f.u(ds->pos, 1), rank = (int)f.u(ds->pos + 1, 1);
    if (rank > 8) fail("dataset rank above 8");
    const uint64_t dpos = ds->pos + (dver == 1 ? 8 : 4);
    for (int d = 0; d < rank; ++d) {
        D.dims.push_back(f.u(dpos + (uint64_t)d * f.sl, f.sl));
        D.numel *= D.dims.back();
    }
    const int cls = (int)(f.u(dt->pos, 1) & 0x0f);
    const uint64_t esize = f.u(dt->pos + 4, 4);
    const uint64_t bits = f.u(dt->pos + 1, 1);
    if (cls != 1 || esize != 8 || (bits & 1)) fail("dataset is not little-endian float64");
    if (!out) return D;
    if ((int64_t)D.numel > capacity) fail("output buffer too small for the dataset");
    if (f.u(ly->pos, 1) != 3) fail("data layout message version other than 3");
    const int lclass = (int)f.u(ly->pos + 1, 1);
    if (lclass == 0) {  // compact
        const uint64_t size = f.u(ly->pos + 2, 2);
        if (size != D.numel * 8) fail("compact data size mismatch");
        f.need(ly->pos + 4, size);
        std::memcpy(out, &f.buf[ly->pos + 4], size);
    } else if (lclass == 1) {  // contiguous
        const uint64_t a = f.u(ly->pos + 2, f.so), size = f.u(ly->pos + 2 + f.so, f.sl);
        if (f.undef(a)) {
            std::memset(out, 0, D.numel * 8);
        } else {
            if (size < D.numel * 8) fail("contiguous data size mismatch");
            f.need(f.addr(a), D.numel * 8);
            std::memcpy(out, &f.buf[f.addr(a)], D.numel * 8);
        }
    } else if (lclass == 2) {  // chunked
        const int cr = (int)f.u(ly->pos + 2, 1);
        if (cr != rank + 1) fail("chunk dimensionality mismatch");
        const uint64_t bt = f.u(ly->pos + 3, f.so);
        std::vector<uint64_t> cd(rank);
        for (int d = 0; d < rank; ++d) {
            cd[d] = f.u(ly->pos + 3 + f.so + 4 * d, 4);
            if (cd[d] == 0) fail("zero chunk dimension");
        }
        std::memset(out, 0, D.numel * 8);
        if (!f.undef(bt)) read_chunks(f, bt, rank, D.dims, cd, filters(f, find_msg(msgs, 0x0b)), out);
    } else {
        fail("unsupported data layout class");
    }
    return D;
}

int h5_read(const char *path, const char *name, double *out, int64_t capacity, int64_t *numel, int32_t *rank,
            int64_t *dims) {
    File f;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(CFK_EINVAL, "cannot open %s", path);
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz <= 0) {
        std::fclose(fp);
        return set_error(CFK_EINVAL, "%s is empty", path);
    }
    f.buf.resize((size_t)sz);
    const size_t got = std::fread(f.buf.data(), 1, f.buf.size(), fp);
    std::fclose(fp);
    if (got != f.buf.size()) return set_error(CFK_EINVAL, "short read of %s", path);
    try {
        static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        uint64_t sb = ~0ull;
        for (uint64_t p = 0; p + 8 <= f.buf.size(); p = p ? 2 * p : 512)
            if (std::memcmp(&f.buf[p], sig, 8) == 0) {
                sb = p;
                break;
            }
        if (sb == ~0ull) fail("no HDF5 superblock (not a MATLAB v7.3 file?)");
        const int ver = (int)f.u(sb + 8, 1);
        if (ver > 1) fail("superblock version " + std::to_string(ver) + " unsupported");
        f.so = (int)f.u(sb + 13, 1);
        f.sl = (int)f.u(sb + 14, 1);
        if ((f.so != 4 && f.so != 8) || (f.sl != 4 && f.sl != 8)) fail("unsupported offset/length size");
        const uint64_t vars = sb + 24 + (ver == 1 ? 4 : 0);
        const uint64_t base = f.u(vars, f.so);
        f.base = base ? base : sb;
        const uint64_t root_entry = vars + 4 * (uint64_t)f.so;
        const uint64_t root_oh = f.u(root_entry + f.so, f.so);
        const Dataset D = read_dataset(f, resolve(f, root_oh, name), out, capacity);
        if (numel) *numel = (int64_t)D.numel;
        if (rank) *rank = (int32_t)D.dims.size();
        if (dims)
            for (size_t d = 0; d < D.dims.size(); ++d) dims[d] = (int64_t)D.dims[d];
    } catch (const H5Error &e) {
        return set_error(CFK_EUNSUP, "%s: %s: %s", path, name, e.msg.c_str());
    }
    return CFK_OK;
}

}  // namespace
}  // namespace cfk

extern "C" int cfk_h5_read_f64(const char *path, const char *dataset, double *out, int64_t capacity, int64_t *numel,
                               int32_t *rank, int64_t dims[8]) {
    if (!path || !dataset) return cfk::set_error(CFK_EINVAL, "null path or dataset name");
    return cfk::h5_read(path, dataset, out, capacity, numel, rank, dims);
}

// ------------------------------------------------------------------------------------------------
// Writer: float64 arrays as the top-level variables of a MATLAB v7.3 .mat file.
//
// Layout (what MATLAB's load reads for a plain double array): a 512-byte user block opening with the
// 128-byte MAT-file header (116 bytes of text, 8 bytes of subsystem offset, version 0x0200, "IM"),
// then an HDF5 file whose superblock (version 0, 8-byte offsets and lengths, base address 512) roots a
// symbol-table group: a local heap with the variable names, one v1 B-tree node whose children are
// symbol-table nodes of up to 8 entries sorted by name, and one dataset per variable (version-1 object
// header: dataspace with the MATLAB dimensions reversed, IEEE F64LE datatype, fill-value, contiguous
// layout, and the string attribute MATLAB_class = "double"), its data the array in column-major order.
// File addresses are relative to the superblock; the end-of-file address is the file's size.
// ------------------------------------------------------------------------------------------------
namespace cfk {
namespace {

constexpr uint64_t kUndef = ~0ull;
constexpr int kLeafK = 4, kNodeK = 16;  // symbol-table nodes hold 2*4 entries; a B-tree node 2*16 children

struct Out {
    std::vector<uint8_t> b;  // the HDF5 part (addresses are offsets into b)
    size_t pos() const { return b.size(); }
    void u(uint64_t v, int n) {
        for (int i = 0; i < n; ++i) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void bytes(const void *p, size_t n) { b.insert(b.end(), (const uint8_t *)p, (const uint8_t *)p + n); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void align8() {
        while (b.size() % 8) b.push_back(0);
    }
    void put(size_t off, uint64_t v, int n) {
        for (int i = 0; i < n; ++i) b[off + i] = (uint8_t)(v >> (8 * i));
    }
};

// Header message (version-1 object header): type, size (a multiple of 8), flags, 3 reserved bytes.
void msg_header(Out &o, int type, size_t size) {
    o.u((uint64_t)type, 2);
    o.u((uint64_t)size, 2);
    o.u(0, 4);
}

void f64_datatype(Out &o) {  // 20 bytes, padded to 24 by the caller's message size
    o.u(0x11, 1);                   // version 1, class 1 (floating point)
    o.u(0x20, 1);                   // little endian, implied mantissa MSB
    o.u(63, 1);                     // sign bit position
    o.u(0, 1);
    o.u(8, 4);                      // size in bytes
    o.u(0, 2);                      // bit offset
    o.u(64, 2);                     // bit precision
    o.u(52, 1);                     // exponent location
    o.u(11, 1);                     // exponent size
    o.u(0, 1);                      // mantissa location
    o.u(52, 1);                     // mantissa size
    o.u(1023, 4);                   // exponent bias
}

struct Var {
    std::string name;
    std::vector<int64_t> dims;  // MATLAB order
    const double *data;
    uint64_t numel;
    size_t heap_off = 0, ohdr = 0;
};

}  // namespace
}  // namespace cfk

extern "C" int cfk_mat_write_f64(const char *path, int32_t n_vars, const char *const *names, const int32_t *ranks,
                                 const int64_t *dims, const double *const *data) {
    using namespace cfk;
    if (!path || n_vars < 1 || !names || !ranks || !dims || !data) return set_error(CFK_EINVAL, "null or empty argument");
    if (n_vars > 2 * kLeafK * 2 * kNodeK)
        return set_error(CFK_EUNSUP, "%d variables (at most %d)", n_vars, 2 * kLeafK * 2 * kNodeK);
    std::vector<Var> vars((size_t)n_vars);
    size_t dpos = 0;
    for (int i = 0; i < n_vars; ++i) {
        Var &v = vars[(size_t)i];
        if (!names[i] || !data[i]) return set_error(CFK_EINVAL, "variable %d: null name or data", i);
        v.name = names[i];
        bool ok = !v.name.empty() && v.name.size() <= 63 && std::isalpha((unsigned char)v.name[0]);
        for (char c : v.name) ok = ok && (std::isalnum((unsigned char)c) || c == '_');
        if (!ok) return set_error(CFK_EINVAL, "variable %d: '%s' is not a MATLAB identifier", i, v.name.c_str());
        if (ranks[i] < 1 || ranks[i] > 8) return set_error(CFK_EINVAL, "variable %s: rank %d", v.name.c_str(), ranks[i]);
        v.numel = 1;
        for (int r = 0; r < ranks[i]; ++r) {
            const int64_t d = dims[dpos++];
            if (d < 1) return set_error(CFK_EINVAL, "variable %s: dimension %lld", v.name.c_str(), (long long)d);
            v.dims.push_back(d);
            v.numel *= (uint64_t)d;
        }
        if (v.dims.size() == 1) v.dims.push_back(1);  // MATLAB arrays have at least two dimensions
        v.data = data[i];
    }
    std::vector<int> order((size_t)n_vars);
    for (int i = 0; i < n_vars; ++i) order[(size_t)i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return vars[(size_t)a].name < vars[(size_t)b].name; });
    for (size_t i = 1; i < order.size(); ++i)
        if (vars[(size_t)order[i]].name == vars[(size_t)order[i - 1]].name)
            return set_error(CFK_EINVAL, "variable '%s' given twice", vars[(size_t)order[i]].name.c_str());

    Out o;
    // superblock (version 0); addresses patched below
    const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    o.bytes(sig, 8);
    o.u(0, 1); o.u(0, 1); o.u(0, 1); o.u(0, 1); o.u(0, 1);  // versions: superblock, free space, root entry, -, shared msgs
    o.u(8, 1); o.u(8, 1); o.u(0, 1);                        // size of offsets, of lengths, reserved
    o.u(kLeafK, 2); o.u(kNodeK, 2); o.u(0, 4);              // group leaf / internal node K, flags
    o.u(512, 8);                                            // base address (absolute)
    o.u(kUndef, 8);                                         // free-space info
    const size_t eof_at = o.pos();
    o.u(0, 8);                                              // end of file (patched)
    o.u(kUndef, 8);                                         // driver info
    // root group symbol-table entry
    o.u(0, 8);                                              // link name offset ("")
    const size_t root_ohdr_at = o.pos();
    o.u(0, 8);                                              // root object header (patched)
    o.u(1, 4); o.u(0, 4);                                   // cache type 1: scratch holds B-tree and heap
    const size_t root_scratch = o.pos();
    o.u(0, 8); o.u(0, 8);

    // local heap: "" then each name, null-terminated, padded to 8
    std::vector<uint8_t> heap(8, 0);
    for (Var &v : vars) {
        v.heap_off = heap.size();
        heap.insert(heap.end(), v.name.begin(), v.name.end());
        heap.push_back(0);
        while (heap.size() % 8) heap.push_back(0);
    }
    const size_t heap_hdr = o.pos();
    o.bytes("HEAP", 4); o.u(0, 1); o.zeros(3);
    o.u(heap.size(), 8); o.u(1, 8);  // data segment size; free list: none (HDF5 stores 1 for that)
    const size_t heap_data_at = o.pos();
    o.u(0, 8);
    o.put(heap_data_at, o.pos(), 8);
    o.bytes(heap.data(), heap.size());
    o.align8();

    // symbol-table nodes of up to 2 kLeafK entries, in name order
    const int per = 2 * kLeafK;
    const int n_snod = (n_vars + per - 1) / per;
    std::vector<size_t> snod_at((size_t)n_snod), entry_ohdr;  // patch positions of the object header addresses
    for (int s = 0; s < n_snod; ++s) {
        snod_at[(size_t)s] = o.pos();
        const int i0 = s * per, i1 = std::min(n_vars, i0 + per);
        o.bytes("SNOD", 4); o.u(1, 1); o.u(0, 1); o.u((uint64_t)(i1 - i0), 2);
        for (int e = 0; e < per; ++e) {
            if (i0 + e < i1) {
                const Var &v = vars[(size_t)order[(size_t)(i0 + e)]];
                o.u(v.heap_off, 8);
                entry_ohdr.push_back(o.pos());
                o.u(0, 8);
            } else {
                o.u(0, 8); o.u(0, 8);
            }
            o.u(0, 4); o.u(0, 4); o.zeros(16);  // cache type 0, reserved, scratch
        }
    }
    // one B-tree node (group nodes, level 0) over the symbol-table nodes; keys are heap offsets of the
    // last name of each child (key 0: the empty name)
    const size_t btree_at = o.pos();
    o.bytes("TREE", 4); o.u(0, 1); o.u(0, 1); o.u((uint64_t)n_snod, 2);
    o.u(kUndef, 8); o.u(kUndef, 8);
    o.u(0, 8);
    for (int s = 0; s < 2 * kNodeK; ++s) {
        if (s < n_snod) {
            const int last = std::min(n_vars, (s + 1) * per) - 1;
            o.u(snod_at[(size_t)s], 8);
            o.u(vars[(size_t)order[(size_t)last]].heap_off, 8);
        } else {
            o.u(0, 8); o.u(0, 8);
        }
    }
    o.put(root_scratch, btree_at, 8);
    o.put(root_scratch + 8, heap_hdr, 8);

    // root group object header: one symbol-table message
    o.align8();
    o.put(root_ohdr_at, o.pos(), 8);
    o.u(1, 1); o.u(0, 1); o.u(1, 2); o.u(1, 4); o.u(24, 4); o.u(0, 4);
    msg_header(o, 0x11, 16);
    o.u(btree_at, 8); o.u(heap_hdr, 8);

    // datasets
    std::vector<size_t> layout_addr_at((size_t)n_vars);
    for (int i = 0; i < n_vars; ++i) {
        Var &v = vars[(size_t)i];
        const int r = (int)v.dims.size();
        o.align8();
        v.ohdr = o.pos();
        const size_t ds_size = 8 + 16 * (size_t)r, dt_size = 24, fv_size = 8, ly_size = 24, at_size = 48;
        o.u(1, 1); o.u(0, 1); o.u(5, 2); o.u(1, 4);
        o.u(5 * 8 + ds_size + dt_size + fv_size + ly_size + at_size, 4); o.u(0, 4);
        // dataspace v1: current and maximum dimensions, HDF5 order = MATLAB order reversed
        msg_header(o, 0x01, ds_size);
        o.u(1, 1); o.u((uint64_t)r, 1); o.u(1, 1); o.u(0, 1); o.u(0, 4);
        for (int k = r - 1; k >= 0; --k) o.u((uint64_t)v.dims[(size_t)k], 8);
        for (int k = r - 1; k >= 0; --k) o.u((uint64_t)v.dims[(size_t)k], 8);
        msg_header(o, 0x03, dt_size);
        f64_datatype(o);
        o.zeros(4);
        // fill value v2: allocated early, written if set, no value defined
        msg_header(o, 0x05, fv_size);
        o.u(2, 1); o.u(1, 1); o.u(2, 1); o.u(0, 1); o.zeros(4);
        // layout v3: contiguous
        msg_header(o, 0x08, ly_size);
        o.u(3, 1); o.u(1, 1);
        layout_addr_at[(size_t)i] = o.pos();
        o.u(0, 8); o.u(v.numel * 8, 8); o.zeros(6);
        // attribute v1 MATLAB_class = "double" (string of 6 bytes, null-terminated padding, scalar)
        msg_header(o, 0x0C, at_size);
        o.u(1, 1); o.u(0, 1); o.u(13, 2); o.u(8, 2); o.u(8, 2);
        o.bytes("MATLAB_class\0\0\0\0", 16);
        o.u(0x13, 1); o.u(0, 1); o.u(0, 1); o.u(0, 1); o.u(6, 4);  // string, null terminated, ASCII
        o.u(1, 1); o.u(0, 1); o.u(0, 1); o.u(0, 1); o.u(0, 4);    // scalar dataspace v1
        o.bytes("double\0\0", 8);
    }
    for (size_t i = 0; i < entry_ohdr.size(); ++i) o.put(entry_ohdr[i], vars[(size_t)order[i]].ohdr, 8);
    for (int i = 0; i < n_vars; ++i) {
        o.align8();
        o.put(layout_addr_at[(size_t)i], o.pos(), 8);
        o.bytes(vars[(size_t)i].data, vars[(size_t)i].numel * 8);
    }
    o.put(eof_at, 512 + o.pos(), 8);

    // MAT-file header in the user block
    char head[512];
    std::memset(head, 0, sizeof(head));
    char text[117];
    std::snprintf(text, sizeof(text), "MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: libcfk HDF5 schema 1.00 .");
    std::memset(head, ' ', 116);
    std::memcpy(head, text, std::strlen(text));
    head[124] = 0x00;
    head[125] = 0x02;
    head[126] = 'I';
    head[127] = 'M';
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(CFK_EINVAL, "cannot open %s for writing", path);
    const bool ok = std::fwrite(head, 1, 512, fp) == 512 && std::fwrite(o.b.data(), 1, o.b.size(), fp) == o.b.size();
    if (std::fclose(fp) != 0 || !ok) return set_error(CFK_EINVAL, "writing %s failed", path);
    return CFK_OK;
}
