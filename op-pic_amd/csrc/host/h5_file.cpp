// Reader of the HDF5 files that OP-PIC's opp_decl_set_hdf5 / opp_decl_map_hdf5 / opp_decl_dat_hdf5 read
// (opp_lib/src/hdf5/opp_mpi_hdf5.cpp:126-517), written from the published HDF5 file-format
// specification, with no HDF5 library. It covers the layout those files use: superblock version 0 or 1
// with 8-byte offsets and lengths, groups stored as symbol tables (version-1 B-tree of SNOD nodes plus a
// local heap of names), version-1 object headers with continuation blocks, and datasets of
// little-endian 4- or 8-byte integers or IEEE floats, stored contiguously or compactly without filters.
// Any other construct is reported as unsupported instead of being guessed at. Every offset read from
// the file is checked against the file's size before it is dereferenced.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "oppx.h"

namespace oppx {
void set_error(const std::string& msg);
}

namespace {

constexpr uint64_t kUndef = ~0ull;

// The file, mapped read-only for the duration of one call.
struct Mapped {
    const uint8_t* p = nullptr;
    uint64_t n = 0;
    int fd = -1;
    explicit Mapped(const char* path) {
        fd = path ? open(path, O_RDONLY) : -1;
        if (fd < 0) return;
        struct stat st;
        if (fstat(fd, &st) != 0 || st.st_size <= 0) return;
        void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) return;
        p = (const uint8_t*)m;
        n = (uint64_t)st.st_size;
    }
    ~Mapped() {
        if (p) munmap((void*)p, (size_t)n);
        if (fd >= 0) close(fd);
    }
    Mapped(const Mapped&) = delete;
    Mapped& operator=(const Mapped&) = delete;
};

struct Reader {
    const Mapped& f;
    uint64_t base = 0;  // every address in the file is relative to the superblock's base address
    std::string err;

    explicit Reader(const Mapped& m) : f(m) {}

    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }
    bool in(uint64_t off, uint64_t len) const { return off <= f.n && len <= f.n - off; }
    template <typename T>
    bool get(uint64_t off, T& v) {
        if (!in(off, sizeof(T))) return fail("HDF5 structure points past the end of the file");
        memcpy(&v, f.p + off, sizeof(T));
        return true;
    }
    bool addr(uint64_t off, uint64_t& a) {  // an address field (relative to base)
        if (!get(off, a)) return false;
        if (a != kUndef) a += base;
        return true;
    }

    // superblock -> object header address of the root group
    bool root(uint64_t& oh) {
        static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        if (!in(0, 8) || memcmp(f.p, sig, 8) != 0) return fail("not an HDF5 file (no signature at offset 0)");
        uint8_t ver = 0, so = 0, sl = 0;
        if (!get(8, ver) || !get(13, so) || !get(14, sl)) return false;
        if (ver > 1) return fail("HDF5 superblock version " + std::to_string(ver) + " is not supported (0 or 1)");
        if (so != 8 || sl != 8) return fail("HDF5 offsets/lengths other than 8 bytes are not supported");
        const uint64_t p = ver == 0 ? 24 : 28;
        if (!get(p, base)) return false;
        // root group symbol table entry after base, free-space, end-of-file and driver addresses
        return addr(p + 32 + 8, oh);
    }

    struct Msg {
        uint16_t type;
        uint64_t data;
        uint16_t size;
    };

    bool messages(uint64_t oh, std::vector<Msg>& out) {
        uint8_t ver = 0;
        uint16_t nmsg = 0;
        uint32_t size = 0;
        if (!get(oh, ver)) return false;
        if (ver != 1) return fail("HDF5 object header version " + std::to_string(ver) + " is not supported (1)");
        if (!get(oh + 2, nmsg) || !get(oh + 8, size)) return false;
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{oh + 16, size}};
        for (size_t b = 0; b < blocks.size() && out.size() < nmsg; b++) {
            if (blocks.size() > 4096) return fail("HDF5 object header continuation chain too long");
            const uint64_t start = blocks[b].first, len = blocks[b].second;
            if (!in(start, len)) return fail("HDF5 object header block past the end of the file");
            uint64_t p = start;
            while (p + 8 <= start + len && out.size() < nmsg) {
                uint16_t t = 0, sz = 0;
                if (!get(p, t) || !get(p + 2, sz)) return false;
                const uint64_t d = p + 8;
                if (!in(d, sz) || d + sz > start + len) return fail("HDF5 header message past its block");
                if (t == 0x10) {  // continuation: another block of messages
                    uint64_t a = 0, l = 0;
                    if (!addr(d, a) || !get(d + 8, l)) return false;
                    blocks.push_back({a, l});
                } else {
                    out.push_back({t, d, sz});
                }
                p = d + sz;
            }
        }
        return true;
    }

    bool heap_name(uint64_t heap, uint64_t off, std::string& name) {
        if (!in(heap, 32) || memcmp(f.p + heap, "HEAP", 4) != 0) return fail("HDF5 local heap signature missing");
        uint64_t seg_size = 0, seg = 0;
        if (!get(heap + 8, seg_size) || !addr(heap + 24, seg)) return false;
        if (off >= seg_size || !in(seg, seg_size)) return fail("HDF5 link name outside its local heap");
        const char* s = (const char*)f.p + seg + off;
        const size_t room = (size_t)(seg_size - off);
        const size_t len = strnlen(s, room);
        if (len == room) return fail("HDF5 link name not terminated");
        name.assign(s, len);
        return true;
    }

    // Look `name` up among the links of the symbol-table group at `group` (B-tree walk).
    bool lookup(uint64_t group, const std::string& name, uint64_t& oh, bool& found) {
        std::vector<Msg> ms;
        if (!messages(group, ms)) return false;
        uint64_t btree = kUndef, heap = kUndef;
        for (const Msg& m : ms)
            if (m.type == 0x11 && (!addr(m.data, btree) || !addr(m.data + 8, heap))) return false;
        if (btree == kUndef) return fail("HDF5 group without a symbol table (new-style groups are not supported)");
        found = false;
        std::vector<std::pair<uint64_t, int>> stack{{btree, 0}};
        while (!stack.empty()) {
            const uint64_t node = stack.back().first;
            const int depth = stack.back().second;
            stack.pop_back();
            if (depth > 64 || stack.size() > 65536) return fail("HDF5 group B-tree too deep");
            if (!in(node, 24) || memcmp(f.p + node, "TREE", 4) != 0) return fail("HDF5 B-tree signature missing");
            uint8_t type = 0, level = 0;
            uint16_t used = 0;
            if (!get(node + 4, type) || !get(node + 5, level) || !get(node + 6, used)) return false;
            if (type != 0) return fail("HDF5 B-tree of a group is not a group node");
            for (uint16_t k = 0; k < used; k++) {
                uint64_t child = 0;
                if (!addr(node + 24 + 16 * (uint64_t)k + 8, child)) return false;
                if (level > 0) {
                    stack.push_back({child, depth + 1});
                    continue;
                }
                if (!in(child, 8) || memcmp(f.p + child, "SNOD", 4) != 0) return fail("HDF5 symbol node signature missing");
                uint16_t ns = 0;
                if (!get(child + 6, ns)) return false;
                for (uint16_t s = 0; s < ns; s++) {
                    const uint64_t e = child + 8 + 40 * (uint64_t)s;
                    uint64_t name_off = 0, obj = 0;
                    std::string n;
                    if (!get(e, name_off) || !addr(e + 8, obj) || !heap_name(heap, name_off, n)) return false;
                    if (n == name) {
                        oh = obj;
                        found = true;
                        return true;
                    }
                }
            }
        }
        return true;
    }

    // Object header of the path (names separated by '/', from the root group).
    bool find(const char* path, uint64_t& oh, bool& found) {
        uint64_t cur = 0;
        if (!root(cur)) return false;
        std::string p(path ? path : "");
        size_t i = 0;
        found = true;
        while (i < p.size()) {
            size_t j = p.find('/', i);
            if (j == std::string::npos) j = p.size();
            if (j > i) {
                bool f2 = false;
                if (!lookup(cur, p.substr(i, j - i), cur, f2)) return false;
                if (!f2) {
                    found = false;
                    return true;
                }
            }
            i = j + 1;
        }
        oh = cur;
        return true;
    }

    struct Dataset {
        int type = 0;  // OPPX_H5_* of the stored elements
        int elem = 0;  // bytes per element
        std::vector<uint64_t> dims;
        uint64_t data = kUndef;  // file offset of the elements (kUndef: never written, reads as zeros)
    };

    bool dataset(uint64_t oh, Dataset& ds) {
        std::vector<Msg> ms;
        if (!messages(oh, ms)) return false;
        bool have_type = false, have_space = false, have_layout = false;
        for (const Msg& m : ms) {
            if (m.type == 0x0B) return fail("filtered (compressed) HDF5 datasets are not supported");
            if (m.type == 0x03) {  // datatype
                uint8_t cv = 0, bits = 0;
                uint32_t size = 0;
                if (!get(m.data, cv) || !get(m.data + 1, bits) || !get(m.data + 4, size)) return false;
                const int cls = cv & 15;
                if (bits & 1) return fail("big-endian HDF5 data is not supported");
                if (cls == 0 && (size == 4 || size == 8)) ds.type = size == 4 ? OPPX_H5_INT32 : OPPX_H5_INT64;
                else if (cls == 1 && (size == 4 || size == 8)) ds.type = size == 4 ? OPPX_H5_FLOAT32 : OPPX_H5_FLOAT64;
                else return fail("HDF5 datatype class " + std::to_string(cls) + " of " + std::to_string(size) +
                                 " bytes is not supported (4- or 8-byte integers and floats)");
                if (cls == 0 && !(bits & 8) && size == 8) return fail("unsigned 8-byte HDF5 integers are not supported");
                ds.elem = (int)size;
                have_type = true;
            } else if (m.type == 0x01) {  // dataspace
                uint8_t ver = 0, nd = 0;
                if (!get(m.data, ver) || !get(m.data + 1, nd)) return false;
                if (ver != 1 && ver != 2) return fail("HDF5 dataspace version " + std::to_string(ver));
                const uint64_t p = m.data + (ver == 1 ? 8 : 4);
                ds.dims.assign(nd, 0);
                for (uint8_t d = 0; d < nd; d++)
                    if (!get(p + 8 * (uint64_t)d, ds.dims[d])) return false;
                have_space = true;
            } else if (m.type == 0x08) {  // data layout
                uint8_t ver = 0, cls = 0;
                if (!get(m.data, ver)) return false;
                if (ver == 3) {
                    if (!get(m.data + 1, cls)) return false;
                    if (cls == 1) {
                        if (!addr(m.data + 2, ds.data)) return false;
                    } else if (cls == 0) {
                        ds.data = m.data + 4;  // compact: the elements follow in the message
                    } else {
                        return fail("chunked HDF5 datasets are not supported");
                    }
                } else if (ver == 1 || ver == 2) {
                    if (!get(m.data + 2, cls)) return false;
                    if (cls != 1) return fail("HDF5 data layout class " + std::to_string(cls) + " is not supported");
                    if (!addr(m.data + 8, ds.data)) return false;
                } else {
                    return fail("HDF5 data layout version " + std::to_string(ver) + " is not supported");
                }
                have_layout = true;
            }
        }
        if (!have_type || !have_space || !have_layout) return fail("HDF5 object is not a dataset");
        return true;
    }
};

int64_t count_of(const std::vector<uint64_t>& dims) {
    int64_t n = 1;
    for (uint64_t d : dims) {
        if (d > (uint64_t)INT64_MAX || (d && n > INT64_MAX / (int64_t)d)) return -1;
        n *= (int64_t)d;
    }
    return n;
}

// Rows and the product of the trailing extents, as the reference takes them (opp_hdf5_common.cpp's
// get_dataset_properties: size = dims[0], dim = dims[1] or 1). A scalar is one row of one element.
void rows_dim(const std::vector<uint64_t>& dims, int64_t& rows, int64_t& dim) {
    if (dims.empty()) {
        rows = 1;
        dim = 1;
        return;
    }
    rows = (int64_t)dims[0];
    dim = dims.size() > 1 ? (int64_t)dims[1] : 1;
}

int open_dataset(const Mapped& m, Reader& r, const char* path, const char* name, Reader::Dataset& ds) {
    if (!m.p) {
        oppx::set_error(std::string("cannot open HDF5 file ") + (path ? path : "(null)"));
        return OPPX_ERR_ARG;
    }
    uint64_t oh = 0;
    bool found = false;
    if (!r.find(name, oh, found)) {
        oppx::set_error(std::string(path) + ": " + r.err);
        return OPPX_ERR_ARG;
    }
    if (!found) {
        oppx::set_error(std::string("no dataset '") + name + "' in " + path);
        return OPPX_ERR_NOTFOUND;
    }
    if (!r.dataset(oh, ds)) {
        oppx::set_error(std::string(path) + ": '" + name + "': " + r.err);
        return OPPX_ERR_ARG;
    }
    return OPPX_OK;
}

}  // namespace

extern "C" {

int oppx_h5_dataset_info(const char* path, const char* name, int32_t* type, int64_t* rows, int64_t* dim) {
    if (!path || !name || !type || !rows || !dim) {
        oppx::set_error("oppx_h5_dataset_info: null argument");
        return OPPX_ERR_ARG;
    }
    Mapped m(path);
    Reader r(m);
    Reader::Dataset ds;
    const int rc = open_dataset(m, r, path, name, ds);
    if (rc) return rc;
    *type = ds.type;
    rows_dim(ds.dims, *rows, *dim);
    return OPPX_OK;
}

int oppx_h5_read(const char* path, const char* name, int32_t type, int64_t rows, int64_t dim, void* out) {
    if (!path || !name || (!out && rows * dim > 0) || rows < 0 || dim < 0) {
        oppx::set_error("oppx_h5_read: invalid argument");
        return OPPX_ERR_ARG;
    }
    if (type != OPPX_H5_INT32 && type != OPPX_H5_INT64 && type != OPPX_H5_FLOAT32 && type != OPPX_H5_FLOAT64) {
        oppx::set_error("oppx_h5_read: unknown element type");
        return OPPX_ERR_ARG;
    }
    Mapped m(path);
    Reader r(m);
    Reader::Dataset ds;
    const int rc = open_dataset(m, r, path, name, ds);
    if (rc) return rc;
    int64_t frows = 0, fdim = 0;
    rows_dim(ds.dims, frows, fdim);
    const int64_t total = count_of(ds.dims);
    if (total < 0 || frows * fdim != total) {
        oppx::set_error(std::string("'") + name + "' has more than two dimensions");
        return OPPX_ERR_ARG;
    }
    if (rows > frows || dim != fdim) {
        oppx::set_error(std::string("'") + name + "' holds " + std::to_string(frows) + " x " + std::to_string(fdim) +
                        ", asked for " + std::to_string(rows) + " x " + std::to_string(dim));
        return OPPX_ERR_ARG;
    }
    const int64_t n = rows * dim;
    const int out_size = (type == OPPX_H5_INT32 || type == OPPX_H5_FLOAT32) ? 4 : 8;
    if (ds.data == kUndef) {  // never written: the default fill value
        memset(out, 0, (size_t)n * (size_t)out_size);
        return OPPX_OK;
    }
    if ((uint64_t)total > m.n / (uint64_t)ds.elem || !r.in(ds.data, (uint64_t)total * (uint64_t)ds.elem)) {
        oppx::set_error(std::string("'") + name + "' data past the end of " + path);
        return OPPX_ERR_ARG;
    }
    const uint8_t* src = m.p + ds.data;
    if (ds.type == type) {
        memcpy(out, src, (size_t)n * (size_t)out_size);
        return OPPX_OK;
    }
    // conversion as HDF5's to the memory type: integers widen or narrow (out of range is an error here,
    // where HDF5 would clip), integers to floats round, floats to integers truncate toward zero
    for (int64_t i = 0; i < n; i++) {
        double v = 0.0;
        long long iv = 0;
        const bool is_int = ds.type == OPPX_H5_INT32 || ds.type == OPPX_H5_INT64;
        if (ds.type == OPPX_H5_INT32) { int32_t t; memcpy(&t, src + 4 * i, 4); iv = t; }
        else if (ds.type == OPPX_H5_INT64) { int64_t t; memcpy(&t, src + 8 * i, 8); iv = t; }
        else if (ds.type == OPPX_H5_FLOAT32) { float t; memcpy(&t, src + 4 * i, 4); v = t; }
        else { double t; memcpy(&t, src + 8 * i, 8); v = t; }
        if (type == OPPX_H5_FLOAT64) ((double*)out)[i] = is_int ? (double)iv : v;
        else if (type == OPPX_H5_FLOAT32) ((float*)out)[i] = is_int ? (float)iv : (float)v;
        else {
            if (!is_int) {
                if (!(std::fabs(v) < 9.2e18)) {
                    oppx::set_error(std::string("'") + name + "': value out of integer range");
                    return OPPX_ERR_ARG;
                }
                iv = (long long)v;
            }
            if (type == OPPX_H5_INT32) {
                if (iv < INT32_MIN || iv > INT32_MAX) {
                    oppx::set_error(std::string("'") + name + "': value out of int32 range");
                    return OPPX_ERR_ARG;
                }
                ((int32_t*)out)[i] = (int32_t)iv;
            } else {
                ((int64_t*)out)[i] = (int64_t)iv;
            }
        }
    }
    return OPPX_OK;
}

}  // extern "C"
