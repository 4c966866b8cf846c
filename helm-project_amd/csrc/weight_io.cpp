// weight_io.cpp -- weight persistence on the host side of the boundary.
//
//   * AXWC blob (axis/solver/weight_cache.hpp:9-27, 64-157, 286-296): a
//     40-byte header [magic u32 0x43574841][version u32 = 1][endian u8][pad 7]
//     [n_src u64][n_dst u64][nnz u64] followed by factor_list (f64),
//     factor_row and factor_col (int64), frac_a, frac_b, area_a, area_b (f64),
//     all in native byte order.  The caller owns the buffer; no file is
//     opened.  A matrix is written in CSR (row, col) order, one of the COO
//     orders the format admits; reading sorts any COO order into CSR, as
//     to_csr does.
//   * ESMF weight file (io/esmf_weight_io.cpp:68-173): dimension num_wgts,
//     variables S (double), col_idx and row_idx (int, 1-based), global int
//     attributes n_a and n_b.  Written in the NetCDF classic 64-bit-offset
//     format (which the netCDF library reads like its other formats);
//     classic files of any of the three classic versions are read.  A
//     NetCDF-4 (HDF5) file is refused with an explicit error: reading HDF5
//     needs the netCDF/HDF5 libraries, which this library does not link.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "axis_hip.hpp"

namespace axis_hip {

namespace {

constexpr std::uint32_t kMagic = 0x43574841u;  // the reference's constant; stored bytes 41 48 57 43 on a little-endian host
constexpr std::uint32_t kVersion = 1;
constexpr std::uint8_t kLittle = 0x01, kBig = 0x02;
constexpr std::size_t kHeader = 40;

std::uint8_t native_endian() {
    const std::uint32_t probe = 1;
    std::uint8_t b;
    std::memcpy(&b, &probe, 1);
    return b == 1 ? kLittle : kBig;
}

std::string hex8(std::uint32_t v) {
    char buf[16];
    std::snprintf(buf, sizeof buf, "%08X", v);
    return buf;
}

std::string endian_name(std::uint8_t e) {
    if (e == kLittle) return "little-endian";
    if (e == kBig) return "big-endian";
    return "unknown(0x" + hex8(e) + ")";
}

struct HostMatrix {
    std::vector<index_t> row_ptr, col;
    std::vector<double> val, frac_a, frac_b, area_a, area_b;
};

HostMatrix download(const InterpolationMatrix &M) {
    HostMatrix h;
    hipStream_t s = nullptr;
    h.row_ptr.resize(M.n_dst + 1);
    h.col.resize(M.nnz);
    h.val.resize(M.nnz);
    h.frac_a.resize(M.n_src);
    h.area_a.resize(M.n_src);
    h.frac_b.resize(M.n_dst);
    h.area_b.resize(M.n_dst);
    M.row_ptr.download(h.row_ptr.data(), M.n_dst + 1, s);
    M.col.download(h.col.data(), M.nnz, s);
    M.val.download(h.val.data(), M.nnz, s);
    M.frac_a.download(h.frac_a.data(), M.n_src, s);
    M.area_a.download(h.area_a.data(), M.n_src, s);
    M.frac_b.download(h.frac_b.data(), M.n_dst, s);
    M.area_b.download(h.area_b.data(), M.n_dst, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    return h;
}

// ---- NetCDF classic encoding (big-endian throughout)
constexpr std::uint32_t kNcDimension = 10, kNcVariable = 11, kNcAttribute = 12;
constexpr std::uint32_t kNcByte = 1, kNcChar = 2, kNcShort = 3, kNcInt = 4, kNcFloat = 5, kNcDouble = 6;
constexpr std::uint32_t kNcUbyte = 7, kNcUshort = 8, kNcUint = 9, kNcInt64 = 10, kNcUint64 = 11;

int type_size(std::uint32_t t) {
    switch (t) {
        case kNcByte: case kNcChar: case kNcUbyte: return 1;
        case kNcShort: case kNcUshort: return 2;
        case kNcInt: case kNcFloat: case kNcUint: return 4;
        case kNcDouble: case kNcInt64: case kNcUint64: return 8;
        default: return 0;
    }
}

struct Writer {
    std::vector<std::uint8_t> b;
    void u32(std::uint32_t v) {
        for (int i = 3; i >= 0; --i) b.push_back(static_cast<std::uint8_t>(v >> (8 * i)));
    }
    void u64(std::uint64_t v) {
        for (int i = 7; i >= 0; --i) b.push_back(static_cast<std::uint8_t>(v >> (8 * i)));
    }
    void pad4() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        u32(static_cast<std::uint32_t>(s.size()));
        b.insert(b.end(), s.begin(), s.end());
        pad4();
    }
};

struct Reader {
    const std::uint8_t *p;
    std::size_t n, at = 0;
    int version;
    std::string path;
    [[noreturn]] void fail(const std::string &what) const { throw RuntimeError("read_esmf: " + path + ": " + what); }
    void need(std::size_t k) const {
        if (at + k > n) fail("truncated NetCDF header");
    }
    std::uint64_t be(int bytes) {
        need(static_cast<std::size_t>(bytes));
        std::uint64_t v = 0;
        for (int i = 0; i < bytes; ++i) v = (v << 8) | p[at + i];
        at += static_cast<std::size_t>(bytes);
        return v;
    }
    std::uint64_t count() { return be(version == 5 ? 8 : 4); }  // NON_NEG / nelems
    std::string name() {
        const std::uint64_t len = count();
        need(len);
        std::string s(reinterpret_cast<const char *>(p + at), len);
        at += len;
        at = (at + 3) & ~std::size_t(3);
        return s;
    }
};

struct NcVar {
    std::string name;
    std::vector<std::uint64_t> dims;
    std::uint32_t type = 0;
    std::uint64_t begin = 0;
};

double decode(const std::uint8_t *q, std::uint32_t type) {
    std::uint64_t v = 0;
    const int sz = type_size(type);
    for (int i = 0; i < sz; ++i) v = (v << 8) | q[i];
    switch (type) {
        case kNcByte: return static_cast<double>(static_cast<std::int8_t>(v));
        case kNcUbyte: return static_cast<double>(static_cast<std::uint8_t>(v));
        case kNcShort: return static_cast<double>(static_cast<std::int16_t>(v));
        case kNcUshort: return static_cast<double>(static_cast<std::uint16_t>(v));
        case kNcInt: return static_cast<double>(static_cast<std::int32_t>(v));
        case kNcUint: return static_cast<double>(static_cast<std::uint32_t>(v));
        case kNcInt64: return static_cast<double>(static_cast<std::int64_t>(v));
        case kNcUint64: return static_cast<double>(v);
        case kNcFloat: {
            const std::uint32_t u = static_cast<std::uint32_t>(v);
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        }
        case kNcDouble: {
            double d;
            std::memcpy(&d, &v, 8);
            return d;
        }
        default: return 0.0;
    }
}

std::int64_t decode_int(const std::uint8_t *q, std::uint32_t type) {
    std::uint64_t v = 0;
    const int sz = type_size(type);
    for (int i = 0; i < sz; ++i) v = (v << 8) | q[i];
    switch (type) {
        case kNcByte: return static_cast<std::int8_t>(v);
        case kNcUbyte: return static_cast<std::uint8_t>(v);
        case kNcShort: return static_cast<std::int16_t>(v);
        case kNcUshort: return static_cast<std::uint16_t>(v);
        case kNcInt: return static_cast<std::int32_t>(v);
        case kNcUint: return static_cast<std::uint32_t>(v);
        case kNcInt64: case kNcUint64: return static_cast<std::int64_t>(v);
        default: return static_cast<std::int64_t>(decode(q, type));
    }
}

}  // namespace

std::size_t axwc_size(index_t n_src, index_t n_dst, index_t nnz) {
    return kHeader + static_cast<std::size_t>(nnz) * (8 + 2 * sizeof(index_t)) + static_cast<std::size_t>(n_src) * 16 +
           static_cast<std::size_t>(n_dst) * 16;
}

std::size_t serialize_axwc(const InterpolationMatrix &M, void *buf, std::size_t buf_size) {
    const std::size_t need = axwc_size(M.n_src, M.n_dst, M.nnz);
    if (buf == nullptr) return need;
    if (buf_size < need)
        throw RuntimeError("axis::solver::WeightCache::serialize: buf_size (" + std::to_string(buf_size) + ") < required (" +
                           std::to_string(need) + ")");
    const HostMatrix h = download(M);
    auto *d = static_cast<std::uint8_t *>(buf);
    const std::uint32_t magic = kMagic, version = kVersion;
    const std::uint64_t dims[3] = {static_cast<std::uint64_t>(M.n_src), static_cast<std::uint64_t>(M.n_dst), static_cast<std::uint64_t>(M.nnz)};
    std::memcpy(d, &magic, 4);
    std::memcpy(d + 4, &version, 4);
    d[8] = native_endian();
    std::memset(d + 9, 0, 7);
    std::memcpy(d + 16, dims, 24);
    d += kHeader;
    auto put = [&](const void *src, std::size_t bytes) {
        if (bytes) std::memcpy(d, src, bytes);
        d += bytes;
    };
    // COO triplets in CSR order
    std::vector<index_t> rows(M.nnz);
    for (index_t r = 0; r < M.n_dst; ++r)
        for (index_t e = h.row_ptr[r]; e < h.row_ptr[r + 1]; ++e) rows[e] = r;
    put(h.val.data(), M.nnz * sizeof(double));
    put(rows.data(), M.nnz * sizeof(index_t));
    put(h.col.data(), M.nnz * sizeof(index_t));
    put(h.frac_a.data(), M.n_src * sizeof(double));
    put(h.frac_b.data(), M.n_dst * sizeof(double));
    put(h.area_a.data(), M.n_src * sizeof(double));
    put(h.area_b.data(), M.n_dst * sizeof(double));
    return need;
}

void check_axwc(const void *buf, std::size_t buf_size, index_t &n_src, index_t &n_dst, index_t &nnz) {
    if (buf == nullptr) throw RuntimeError("axis::solver::WeightCache::deserialize: buf is nullptr");
    if (buf_size < kHeader)
        throw RuntimeError("axis::solver::WeightCache::deserialize: buf_size (" + std::to_string(buf_size) + ") < minimum header size (40)");
    const auto *s = static_cast<const std::uint8_t *>(buf);
    std::uint32_t magic, version;
    std::uint64_t dims[3];
    std::memcpy(&magic, s, 4);
    std::memcpy(&version, s + 4, 4);
    const std::uint8_t endian = s[8];
    std::memcpy(dims, s + 16, 24);
    if (magic != kMagic)
        throw RuntimeError("axis::solver::WeightCache::deserialize: invalid magic bytes (expected 0x" + hex8(kMagic) + ", got 0x" + hex8(magic) +
                           ")");
    if (version != kVersion)
        throw RuntimeError("axis::solver::WeightCache::deserialize: format version mismatch (expected " + std::to_string(kVersion) + ", got " +
                           std::to_string(version) + ")");
    if (endian != native_endian())
        throw RuntimeError("axis::solver::WeightCache::deserialize: endianness mismatch (blob is " + endian_name(endian) + ", native is " +
                           endian_name(native_endian()) + ")");
    if (dims[0] >= (std::uint64_t(1) << 40) || dims[1] >= (std::uint64_t(1) << 40) || dims[2] >= (std::uint64_t(1) << 40))
        throw RuntimeError("axis::solver::WeightCache::deserialize: implausible dimensions in the header");
    n_src = static_cast<index_t>(dims[0]);
    n_dst = static_cast<index_t>(dims[1]);
    nnz = static_cast<index_t>(dims[2]);
    const std::size_t expected = axwc_size(n_src, n_dst, nnz);
    if (buf_size < expected)
        throw RuntimeError("axis::solver::WeightCache::deserialize: blob size mismatch (buf_size=" + std::to_string(buf_size) +
                           ", expected=" + std::to_string(expected) + " for n_src=" + std::to_string(n_src) + ", n_dst=" + std::to_string(n_dst) +
                           ", nnz=" + std::to_string(nnz) + ")");
}

std::unique_ptr<InterpolationMatrix> deserialize_axwc(const void *buf, std::size_t buf_size) {
    index_t n_src = 0, n_dst = 0, nnz = 0;
    check_axwc(buf, buf_size, n_src, n_dst, nnz);
    const auto *s = static_cast<const std::uint8_t *>(buf) + kHeader;
    std::vector<double> val(nnz), fa(n_src), fb(n_dst), aa(n_src), ab(n_dst);
    std::vector<index_t> rows(nnz), cols(nnz);
    auto get = [&](void *dst, std::size_t bytes) {
        if (bytes) std::memcpy(dst, s, bytes);
        s += bytes;
    };
    get(val.data(), nnz * sizeof(double));
    get(rows.data(), nnz * sizeof(index_t));
    get(cols.data(), nnz * sizeof(index_t));
    get(fa.data(), n_src * sizeof(double));
    get(fb.data(), n_dst * sizeof(double));
    get(aa.data(), n_src * sizeof(double));
    get(ab.data(), n_dst * sizeof(double));
    return matrix_from_coo(rows.data(), cols.data(), val.data(), nnz, n_src, n_dst, fa.data(), fb.data(), aa.data(), ab.data());
}

void write_esmf(const std::string &path, const InterpolationMatrix &M) {
    if (M.nnz >= (index_t(1) << 31) || M.n_src >= (index_t(1) << 31) || M.n_dst >= (index_t(1) << 31))
        throw InvalidArgument("write_esmf: the ESMF format stores int indices and counts; this matrix exceeds 2^31");
    const HostMatrix h = download(M);
    const std::uint64_t nnz = static_cast<std::uint64_t>(M.nnz);
    Writer w;
    w.b = {'C', 'D', 'F', 2};  // 64-bit offset format
    w.u32(0);                  // numrecs
    w.u32(kNcDimension);
    w.u32(1);
    w.name("num_wgts");
    w.u32(static_cast<std::uint32_t>(nnz));
    w.u32(kNcAttribute);
    w.u32(2);
    for (auto &[nm, v] : {std::pair<const char *, index_t>{"n_a", M.n_src}, {"n_b", M.n_dst}}) {
        w.name(nm);
        w.u32(kNcInt);
        w.u32(1);
        w.u32(static_cast<std::uint32_t>(static_cast<std::int32_t>(v)));
    }
    w.u32(kNcVariable);
    w.u32(3);
    const char *names[3] = {"S", "col_idx", "row_idx"};
    const std::uint32_t types[3] = {kNcDouble, kNcInt, kNcInt};
    std::vector<std::size_t> begin_at(3);
    std::uint64_t vsize[3];
    for (int v = 0; v < 3; ++v) {
        w.name(names[v]);
        w.u32(1);  // one dimension
        w.u32(0);  // dimid 0
        w.u32(0);  // no attributes (ABSENT = ZERO ZERO)
        w.u32(0);
        w.u32(types[v]);
        const std::uint64_t bytes = nnz * static_cast<std::uint64_t>(type_size(types[v]));
        vsize[v] = (bytes + 3) & ~std::uint64_t(3);
        w.u32(vsize[v] >= 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<std::uint32_t>(vsize[v]));
        begin_at[v] = w.b.size();
        w.u64(0);  // patched below
    }
    std::uint64_t off = w.b.size();
    for (int v = 0; v < 3; ++v) {
        for (int i = 0; i < 8; ++i) w.b[begin_at[v] + i] = static_cast<std::uint8_t>(off >> (8 * (7 - i)));
        off += vsize[v];
    }
    std::vector<index_t> rows(M.nnz);
    for (index_t r = 0; r < M.n_dst; ++r)
        for (index_t e = h.row_ptr[r]; e < h.row_ptr[r + 1]; ++e) rows[e] = r;
    w.b.reserve(off);
    for (index_t e = 0; e < M.nnz; ++e) {
        std::uint64_t u;
        std::memcpy(&u, &h.val[e], 8);
        w.u64(u);
    }
    w.pad4();
    for (index_t e = 0; e < M.nnz; ++e) w.u32(static_cast<std::uint32_t>(static_cast<std::int32_t>(h.col[e] + 1)));
    for (index_t e = 0; e < M.nnz; ++e) w.u32(static_cast<std::uint32_t>(static_cast<std::int32_t>(rows[e] + 1)));
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw RuntimeError("write_esmf: cannot open " + path + " for writing");
    f.write(reinterpret_cast<const char *>(w.b.data()), static_cast<std::streamsize>(w.b.size()));
    if (!f) throw RuntimeError("write_esmf: write to " + path + " failed");
}

EsmfTriplets parse_esmf(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw RuntimeError("read_esmf: cannot open " + path);
    std::vector<std::uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Reader r{data.data(), data.size(), 0, 0, path};
    if (data.size() >= 8 && std::memcmp(data.data(), "\x89HDF\r\n\x1a\n", 8) == 0)
        r.fail("a NetCDF-4 (HDF5) file; this library reads the NetCDF classic formats only (no netCDF/HDF5 library is linked)");
    if (data.size() < 4 || data[0] != 'C' || data[1] != 'D' || data[2] != 'F' || (data[3] != 1 && data[3] != 2 && data[3] != 5))
        r.fail("not a NetCDF classic file");
    r.version = data[3];
    r.at = 4;
    (void)r.count();  // numrecs
    std::vector<std::pair<std::string, std::uint64_t>> dims;
    std::uint64_t tag = r.be(4), ne = r.count();
    if (tag == kNcDimension) {
        for (std::uint64_t i = 0; i < ne; ++i) {
            std::string nm = r.name();
            dims.emplace_back(nm, r.count());
        }
    } else if (tag != 0 || ne != 0) {
        r.fail("malformed dimension list");
    }
    std::int64_t n_a = -1, n_b = -1;
    auto read_atts = [&](bool global) {
        std::uint64_t t = r.be(4), m = r.count();
        if (t == 0 && m == 0) return;
        if (t != kNcAttribute) r.fail("malformed attribute list");
        for (std::uint64_t i = 0; i < m; ++i) {
            std::string nm = r.name();
            const std::uint32_t ty = static_cast<std::uint32_t>(r.be(4));
            const std::uint64_t cnt = r.count();
            const int sz = type_size(ty);
            if (sz == 0) r.fail("unknown attribute type");
            r.need(cnt * sz);
            if (global && cnt >= 1 && ty != kNcChar && (nm == "n_a" || nm == "n_b")) (nm == "n_a" ? n_a : n_b) = decode_int(r.p + r.at, ty);
            r.at += cnt * sz;
            r.at = (r.at + 3) & ~std::size_t(3);
        }
    };
    read_atts(true);
    std::vector<NcVar> vars;
    tag = r.be(4);
    ne = r.count();
    if (tag == kNcVariable) {
        for (std::uint64_t i = 0; i < ne; ++i) {
            NcVar v;
            v.name = r.name();
            const std::uint64_t nd = r.count();
            for (std::uint64_t k = 0; k < nd; ++k) {
                const std::uint64_t id = r.count();
                if (id >= dims.size()) r.fail("variable " + v.name + " names an unknown dimension");
                v.dims.push_back(dims[id].second);
            }
            read_atts(false);
            v.type = static_cast<std::uint32_t>(r.be(4));
            (void)r.count();  // vsize
            v.begin = r.be(r.version == 1 ? 4 : 8);
            vars.push_back(v);
        }
    } else if (tag != 0 || ne != 0) {
        r.fail("malformed variable list");
    }
    if (n_a < 0 || n_b < 0) r.fail("global attributes n_a / n_b missing");
    auto find = [&](const char *nm) -> const NcVar & {
        for (const NcVar &v : vars)
            if (v.name == nm) return v;
        r.fail(std::string("variable ") + nm + " missing");
    };
    const NcVar &S = find("S"), &C = find("col_idx"), &R = find("row_idx");
    auto length = [&](const NcVar &v) {
        if (v.dims.size() != 1) r.fail("variable " + v.name + " is not one-dimensional");
        return v.dims[0];
    };
    const std::uint64_t nnz = length(S);
    if (length(C) != nnz || length(R) != nnz) r.fail("S, col_idx and row_idx differ in length");
    for (const NcVar *v : {&S, &C, &R}) {
        const int sz = type_size(v->type);
        if (sz == 0 || v->type == kNcChar) r.fail("variable " + v->name + " has an unsupported type");
        if (v->begin + nnz * sz > data.size()) r.fail("variable " + v->name + " runs past the end of the file");
    }
    EsmfTriplets t;
    t.n_src = n_a;
    t.n_dst = n_b;
    t.val.resize(nnz);
    t.rows.resize(nnz);
    t.cols.resize(nnz);
    const int ss = type_size(S.type), cs = type_size(C.type), rs = type_size(R.type);
    for (std::uint64_t k = 0; k < nnz; ++k) {
        t.val[k] = decode(data.data() + S.begin + k * ss, S.type);
        t.cols[k] = decode_int(data.data() + C.begin + k * cs, C.type) - 1;
        t.rows[k] = decode_int(data.data() + R.begin + k * rs, R.type) - 1;
    }
    for (std::uint64_t k = 0; k < nnz; ++k)
        if (t.rows[k] < 0 || t.rows[k] >= n_b || t.cols[k] < 0 || t.cols[k] >= n_a)
            r.fail("entry " + std::to_string(k) + " has a 1-based index outside [1, n_b] x [1, n_a]");
    return t;
}

std::unique_ptr<InterpolationMatrix> matrix_from_esmf(const EsmfTriplets &t) {
    // the reference builds the matrix with zero-filled fractions and areas
    std::vector<double> za(static_cast<std::size_t>(t.n_src), 0.0), zb(static_cast<std::size_t>(t.n_dst), 0.0);
    return matrix_from_coo(t.rows.data(), t.cols.data(), t.val.data(), static_cast<index_t>(t.val.size()), t.n_src, t.n_dst, za.data(), zb.data(),
                           za.data(), zb.data());
}

}  // namespace axis_hip
