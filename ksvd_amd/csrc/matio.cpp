// matio.cpp -- MATLAB level-5 MAT-file read/write without the MATLAB libraries.
//
// Replaces readMat / writeMat (S/FileManager.cu:93-160, built on the MATLAB
// MAT API matOpen/matGetVariable/matPutVariableAsGlobal).  Format (MATLAB
// "MAT-File Format", level 5):
//   128-byte header: 116 bytes of text, 8 bytes subsystem offset, version
//   0x0100 and the endian indicator "IM" (file written little-endian);
//   then data elements, each an 8-byte tag (uint32 type, uint32 byte count)
//   followed by the data padded to 8 bytes, or a 4-byte "small element" tag
//   (byte count in the upper 16 bits) with up to 4 bytes of data.
//   miCOMPRESSED (15) wraps one zlib stream holding a further element;
//   miMATRIX (14) holds array flags (miUINT32), dimensions (miINT32), the
//   name (miINT8) and the real part (any numeric type).
// Reading accepts real numeric 2-D arrays of any numeric storage type and
// converts to double; writing produces uncompressed double matrices.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ksvd_amd.h"

namespace {

enum : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6,
    miSINGLE = 7, miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15,
};
constexpr uint32_t mxDOUBLE_CLASS = 6;
constexpr uint32_t FLAG_GLOBAL = 0x0400;
constexpr uint32_t FLAG_COMPLEX = 0x0800;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct Reader {
    const uint8_t* p;
    size_t n, off = 0;
    bool swap = false;
    uint32_t u32(size_t at) const {
        uint32_t v;
        std::memcpy(&v, p + at, 4);
        if (swap) v = __builtin_bswap32(v);
        return v;
    }
};

// Reads one element tag at r.off: type, byte count, data offset, total size.
bool read_tag(const Reader& r, size_t off, uint32_t& type, uint32_t& nbytes, size_t& data, size_t& total) {
    if (off + 4 > r.n) return false;
    const uint32_t w0 = r.u32(off);
    if ((w0 >> 16) != 0) {  // small data element: count in the upper half
        type = w0 & 0xffff;
        nbytes = w0 >> 16;
        if (nbytes > 4) return false;
        data = off + 4;
        total = 8;
        return off + 8 <= r.n;
    }
    if (off + 8 > r.n) return false;
    type = w0;
    nbytes = r.u32(off + 4);
    data = off + 8;
    total = 8 + ((size_t)nbytes + 7) / 8 * 8;
    if (type == miCOMPRESSED) total = 8 + nbytes;  // compressed elements are not padded
    return data + nbytes <= r.n;
}

template <typename T>
T get(const uint8_t* p, bool swap) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    if (swap) {
        uint8_t b[sizeof(T)];
        std::memcpy(b, &v, sizeof(T));
        for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
        std::memcpy(&v, b, sizeof(T));
    }
    return v;
}

bool to_double(const uint8_t* src, uint32_t type, size_t count, bool swap, std::vector<double>& out) {
    out.resize(count);
    for (size_t i = 0; i < count; ++i) {
        switch (type) {
            case miDOUBLE: out[i] = get<double>(src + 8 * i, swap); break;
            case miSINGLE: out[i] = get<float>(src + 4 * i, swap); break;
            case miINT8: out[i] = (int8_t)src[i]; break;
            case miUINT8: out[i] = src[i]; break;
            case miINT16: out[i] = get<int16_t>(src + 2 * i, swap); break;
            case miUINT16: out[i] = get<uint16_t>(src + 2 * i, swap); break;
            case miINT32: out[i] = get<int32_t>(src + 4 * i, swap); break;
            case miUINT32: out[i] = get<uint32_t>(src + 4 * i, swap); break;
            case miINT64: out[i] = (double)get<int64_t>(src + 8 * i, swap); break;
            case miUINT64: out[i] = (double)get<uint64_t>(src + 8 * i, swap); break;
            default: return false;
        }
    }
    return true;
}

size_t type_size(uint32_t t) {
    switch (t) {
        case miINT8: case miUINT8: return 1;
        case miINT16: case miUINT16: return 2;
        case miINT32: case miUINT32: case miSINGLE: return 4;
        case miDOUBLE: case miINT64: case miUINT64: return 8;
    }
    return 0;
}

// Parses a miMATRIX element body [data, data + nbytes).  Returns 1 when it is
// the variable `name` (filling rows/cols/values), 0 when it is another
// variable, negative on a malformed or unsupported element.
int parse_matrix(const Reader& r, size_t data, uint32_t nbytes, const char* name, int64_t& rows, int64_t& cols,
                 std::vector<double>& values) {
    Reader m{r.p + data, nbytes, 0, r.swap};
    size_t off = 0;
    uint32_t t, nb;
    size_t d, tot;
    // array flags
    if (!read_tag(m, off, t, nb, d, tot) || t != miUINT32 || nb < 8)
        return fail(KSVD_ERR_INVALID, "MAT: bad array-flags subelement");
    const uint32_t flags = m.u32(d);
    off += tot;
    const uint32_t cls = flags & 0xff;
    // dimensions
    if (!read_tag(m, off, t, nb, d, tot) || t != miINT32 || nb < 8)
        return fail(KSVD_ERR_INVALID, "MAT: bad dimensions subelement");
    const int ndim = (int)(nb / 4);
    std::vector<int64_t> dims(ndim);
    for (int i = 0; i < ndim; ++i) dims[i] = get<int32_t>(m.p + d + 4 * i, m.swap);
    off += tot;
    // name
    if (!read_tag(m, off, t, nb, d, tot) || t != miINT8) return fail(KSVD_ERR_INVALID, "MAT: bad name subelement");
    const std::string vname(reinterpret_cast<const char*>(m.p + d), nb);
    off += tot;
    if (vname != name) return 0;
    if (cls < 6 || cls > 15)  // numeric classes: double, single, int8 .. uint64
        return fail(KSVD_ERR_INVALID, "MAT: variable '" + vname + "' is not a numeric array");
    if (flags & FLAG_COMPLEX) return fail(KSVD_ERR_INVALID, "MAT: variable '" + vname + "' is complex");
    if (ndim != 2) return fail(KSVD_ERR_INVALID, "MAT: variable '" + vname + "' is not 2-D");
    // real part
    if (!read_tag(m, off, t, nb, d, tot)) return fail(KSVD_ERR_INVALID, "MAT: bad real-part subelement");
    if (dims[0] < 0 || dims[1] < 0) return fail(KSVD_ERR_INVALID, "MAT: variable '" + vname + "' has a negative dimension");
    // both dimensions are below 2^31, so the count and its byte size fit 64 bits
    const uint64_t count = (uint64_t)dims[0] * (uint64_t)dims[1];
    const uint64_t ts = type_size(t);
    if (ts == 0 || (uint64_t)nb != count * ts) return fail(KSVD_ERR_INVALID, "MAT: real part has the wrong size");
    if (!to_double(m.p + d, t, count, m.swap, values)) return fail(KSVD_ERR_INVALID, "MAT: unsupported data type");
    rows = dims[0];
    cols = dims[1];
    return 1;
}

int find_variable(const std::vector<uint8_t>& file, const char* name, int64_t& rows, int64_t& cols,
                  std::vector<double>& values) {
    if (file.size() < 128) return fail(KSVD_ERR_INVALID, "MAT: file shorter than the 128-byte header");
    Reader r{file.data(), file.size()};
    if (file[126] == 'I' && file[127] == 'M') r.swap = false;
    else if (file[126] == 'M' && file[127] == 'I') r.swap = true;
    else return fail(KSVD_ERR_INVALID, "MAT: not a level-5 MAT-file (endian indicator)");
    size_t off = 128;
    while (off < r.n) {
        uint32_t t, nb;
        size_t d, tot;
        if (!read_tag(r, off, t, nb, d, tot)) return fail(KSVD_ERR_INVALID, "MAT: truncated data element");
        if (t == miCOMPRESSED) {
            std::vector<uint8_t> buf;
            z_stream zs;
            std::memset(&zs, 0, sizeof(zs));
            if (inflateInit(&zs) != Z_OK) return fail(KSVD_ERR_INVALID, "MAT: zlib init failed");
            zs.next_in = const_cast<Bytef*>(r.p + d);
            zs.avail_in = nb;
            int zr = Z_OK;
            uint8_t chunk[1 << 16];
            while (zr == Z_OK) {
                zs.next_out = chunk;
                zs.avail_out = sizeof(chunk);
                zr = inflate(&zs, Z_NO_FLUSH);
                if (zr != Z_OK && zr != Z_STREAM_END) break;
                buf.insert(buf.end(), chunk, chunk + (sizeof(chunk) - zs.avail_out));
            }
            inflateEnd(&zs);
            if (zr != Z_STREAM_END) return fail(KSVD_ERR_INVALID, "MAT: corrupt compressed element");
            Reader c{buf.data(), buf.size(), 0, r.swap};
            uint32_t t2, nb2;
            size_t d2, tot2;
            if (!read_tag(c, 0, t2, nb2, d2, tot2)) return fail(KSVD_ERR_INVALID, "MAT: bad compressed element");
            if (t2 == miMATRIX) {
                const int rc = parse_matrix(c, d2, nb2, name, rows, cols, values);
                if (rc != 0) return rc;
            }
        } else if (t == miMATRIX) {
            const int rc = parse_matrix(r, d, nb, name, rows, cols, values);
            if (rc != 0) return rc;
        }
        off += tot;
    }
    return fail(KSVD_ERR_INVALID, std::string("MAT: no variable named '") + name + "'");
}

void put_u32(std::vector<uint8_t>& b, uint32_t v) {
    uint8_t x[4];
    std::memcpy(x, &v, 4);
    b.insert(b.end(), x, x + 4);
}

void pad8(std::vector<uint8_t>& b) {
    while (b.size() % 8) b.push_back(0);
}

}  // namespace

extern "C" {

int ksvd_mat_read(const char* path, const char* name, double* out, int64_t capacity, int64_t* rows, int64_t* cols) {
    if (!path || !name || !rows || !cols) return fail(KSVD_ERR_NULL, "mat_read: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(KSVD_ERR_INVALID, std::string("MAT: cannot open ") + path);
    std::vector<uint8_t> file;
    uint8_t chunk[1 << 16];
    size_t k;
    while ((k = std::fread(chunk, 1, sizeof(chunk), f)) > 0) file.insert(file.end(), chunk, chunk + k);
    std::fclose(f);
    std::vector<double> values;
    int64_t r = 0, c = 0;
    const int rc = find_variable(file, name, r, c, values);
    if (rc < 0) return rc;
    *rows = r;
    *cols = c;
    if (out) {
        if (capacity < r * c) return fail(KSVD_ERR_INVALID, "mat_read: output buffer too small");
        std::memcpy(out, values.data(), sizeof(double) * (size_t)(r * c));
    }
    return KSVD_OK;
}

int ksvd_mat_write(const char* path, const char* name, const double* data, int64_t rows, int64_t cols, int append,
                   int global) {
    if (!path || !name || (!data && rows * cols > 0)) return fail(KSVD_ERR_NULL, "mat_write: null argument");
    if (rows < 0 || cols < 0 || rows > 0x7fffffff || cols > 0x7fffffff)
        return fail(KSVD_ERR_INVALID, "mat_write: bad dimensions");
    const size_t nname = std::strlen(name);
    if (nname == 0 || nname > 63) return fail(KSVD_ERR_INVALID, "mat_write: variable name must have 1..63 characters");
    const uint64_t nbytes_real = (uint64_t)rows * (uint64_t)cols * 8;
    if (nbytes_real > 0xffffff00ull) return fail(KSVD_ERR_LIMIT, "mat_write: array larger than a level-5 element");
    std::vector<uint8_t> el;  // miMATRIX body
    put_u32(el, miUINT32);
    put_u32(el, 8);
    put_u32(el, mxDOUBLE_CLASS | (global ? FLAG_GLOBAL : 0));
    put_u32(el, 0);
    put_u32(el, miINT32);
    put_u32(el, 8);
    put_u32(el, (uint32_t)rows);
    put_u32(el, (uint32_t)cols);
    if (nname <= 4) {
        put_u32(el, (uint32_t)(miINT8 | (nname << 16)));
        uint8_t nm[4] = {0, 0, 0, 0};
        std::memcpy(nm, name, nname);
        el.insert(el.end(), nm, nm + 4);
    } else {
        put_u32(el, miINT8);
        put_u32(el, (uint32_t)nname);
        el.insert(el.end(), name, name + nname);
        pad8(el);
    }
    put_u32(el, miDOUBLE);
    put_u32(el, (uint32_t)nbytes_real);
    const uint8_t* src = reinterpret_cast<const uint8_t*>(data);
    el.insert(el.end(), src, src + nbytes_real);
    pad8(el);
    FILE* f = std::fopen(path, append ? "ab" : "wb");
    if (!f) return fail(KSVD_ERR_INVALID, std::string("MAT: cannot create ") + path);
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    if (ok && std::ftell(f) == 0) {
        char hdr[128];
        std::memset(hdr, ' ', 116);
        const char* text = "MATLAB 5.0 MAT-file, written by ksvd_amd";
        std::memcpy(hdr, text, std::strlen(text));
        std::memset(hdr + 116, 0, 8);
        const uint16_t ver = 0x0100;
        std::memcpy(hdr + 124, &ver, 2);
        hdr[126] = 'I';
        hdr[127] = 'M';
        ok = std::fwrite(hdr, 1, 128, f) == 128;
    }
    std::vector<uint8_t> tag;
    put_u32(tag, miMATRIX);
    put_u32(tag, (uint32_t)el.size());
    ok = ok && std::fwrite(tag.data(), 1, tag.size(), f) == tag.size();
    ok = ok && std::fwrite(el.data(), 1, el.size(), f) == el.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(KSVD_ERR_INVALID, std::string("MAT: write failed for ") + path);
    return KSVD_OK;
}

const char* ksvd_mat_last_error(void) { return g_err.c_str(); }

}  // extern "C"
