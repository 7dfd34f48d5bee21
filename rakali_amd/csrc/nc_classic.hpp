// nc_classic.hpp -- the restart file as classic NetCDF (the 64-bit-offset
// format, "CDF\x02"), written and read without the NetCDF library, which
// this build does not have.  The layout follows the reference's restart
// (restart_write / restart_read, rki_output.F90:740-988): dimensions x, y;
// FP64 variables h, hu, hv, b on (y, x) (x fastest, the Fortran (x, y)
// order) holding the interior cells; a scalar FP64 time; units attributes
// and the global Conventions, title, source, dx, dy, nghost.  The
// reference writes NetCDF-4 (HDF5), which its NetCDF library reads as well
// as this classic form.  Not part of the ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace rk {
namespace nc {

// Classic-format tags and types (NetCDF file format specification).
constexpr uint32_t TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C;
constexpr uint32_t NC_CHAR = 2, NC_DOUBLE = 6;

inline void put_u32(std::vector<uint8_t>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
inline void put_u64(std::vector<uint8_t>& o, uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
inline void pad4(std::vector<uint8_t>& o) {
    while (o.size() % 4) o.push_back(0);
}
inline void put_name(std::vector<uint8_t>& o, const std::string& s) {
    put_u32(o, (uint32_t)s.size());
    o.insert(o.end(), s.begin(), s.end());
    pad4(o);
}
inline void put_text_att(std::vector<uint8_t>& o, const std::string& name, const std::string& text) {
    put_name(o, name);
    put_u32(o, NC_CHAR);
    put_u32(o, (uint32_t)text.size());
    o.insert(o.end(), text.begin(), text.end());
    pad4(o);
}
inline void put_double_be(std::vector<uint8_t>& o, double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    put_u64(o, u);
}

struct Field2D {
    std::string name, units;
    const double* data;   // interior, x fastest, leading dimension ld
    long long ld;
};

// Writes the restart; returns false on an I/O error.
inline bool write_restart(const std::string& path, int nx, int ny, const std::vector<Field2D>& fields,
                          double time, const std::string& dx, const std::string& dy, const std::string& nghost) {
    // header with placeholder offsets, patched once its size is known
    std::vector<uint8_t> h;
    h.insert(h.end(), {'C', 'D', 'F', 2});
    put_u32(h, 0);   // numrecs: no record dimension
    put_u32(h, TAG_DIMENSION);
    put_u32(h, 2);
    put_name(h, "x"); put_u32(h, (uint32_t)nx);
    put_name(h, "y"); put_u32(h, (uint32_t)ny);
    put_u32(h, TAG_ATTRIBUTE);
    put_u32(h, 6);
    put_text_att(h, "Conventions", "CF-1.11");
    put_text_att(h, "title", "Tidal restart file");
    put_text_att(h, "source", "rakali_amd");
    put_text_att(h, "dx", dx);
    put_text_att(h, "dy", dy);
    put_text_att(h, "nghost", nghost);
    put_u32(h, TAG_VARIABLE);
    put_u32(h, (uint32_t)fields.size() + 1);
    const uint64_t vsize2d = (uint64_t)nx * (uint64_t)ny * 8u;
    std::vector<size_t> begin_at;
    auto var_header = [&](const std::string& name, bool scalar, const std::string& units) {
        put_name(h, name);
        if (scalar) {
            put_u32(h, 0);
        } else {
            put_u32(h, 2);
            put_u32(h, 1);   // y (slowest)
            put_u32(h, 0);   // x
        }
        put_u32(h, TAG_ATTRIBUTE);
        put_u32(h, 1);
        put_text_att(h, "units", units);
        put_u32(h, NC_DOUBLE);
        const uint64_t vs = scalar ? 8u : vsize2d;
        put_u32(h, vs > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)vs);
        begin_at.push_back(h.size());
        put_u64(h, 0);
    };
    for (const Field2D& f : fields) var_header(f.name, false, f.units);
    var_header("time", true, "seconds since simulation start");
    uint64_t off = h.size();
    for (size_t k = 0; k < begin_at.size(); ++k) {
        for (int b = 0; b < 8; ++b) h[begin_at[k] + b] = (uint8_t)(off >> (56 - 8 * b));
        off += k + 1 < begin_at.size() ? vsize2d : 8u;
    }
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    bool ok = std::fwrite(h.data(), 1, h.size(), fp) == h.size();
    std::vector<uint8_t> row;
    for (const Field2D& f : fields) {
        for (int j = 0; ok && j < ny; ++j) {
            row.clear();
            for (int i = 0; i < nx; ++i) put_double_be(row, f.data[(size_t)j * (size_t)f.ld + (size_t)i]);
            ok = std::fwrite(row.data(), 1, row.size(), fp) == row.size();
        }
    }
    row.clear();
    put_double_be(row, time);
    ok = ok && std::fwrite(row.data(), 1, row.size(), fp) == row.size();
    return (std::fclose(fp) == 0) && ok;
}

// A reader of the header of a classic file (CDF-1 or CDF-2): the named
// FP64 variables, their dimensions and data offsets.
struct Var {
    std::string name;
    std::vector<uint32_t> dims;
    uint32_t type = 0;
    uint64_t begin = 0;
};

struct Reader {
    std::vector<uint8_t> buf;
    size_t pos = 0;
    bool fail = false;
    int version = 0;
    std::vector<std::pair<std::string, uint32_t>> dims;
    std::vector<Var> vars;

    uint32_t u32() {
        if (pos + 4 > buf.size()) { fail = true; return 0; }
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) v = (v << 8) | buf[pos++];
        return v;
    }
    uint64_t u64() {
        uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    std::string name() {
        const uint32_t n = u32();
        if (fail || pos + n > buf.size()) { fail = true; return std::string(); }
        std::string s(buf.begin() + (long)pos, buf.begin() + (long)(pos + n));
        pos += (n + 3) / 4 * 4;
        return s;
    }
    void skip_atts() {
        const uint32_t tag = u32(), n = u32();
        if (tag == 0 && n == 0) return;
        if (tag != TAG_ATTRIBUTE) { fail = true; return; }
        for (uint32_t a = 0; a < n && !fail; ++a) {
            name();
            const uint32_t type = u32(), cnt = u32();
            static const uint32_t size_of[7] = {0, 1, 1, 2, 4, 4, 8};
            if (type < 1 || type > 6) { fail = true; return; }
            pos += ((uint64_t)cnt * size_of[type] + 3) / 4 * 4;
        }
    }
    bool parse() {
        if (buf.size() < 8 || std::memcmp(buf.data(), "CDF", 3) != 0) return false;
        version = buf[3];
        if (version != 1 && version != 2) return false;
        pos = 4;
        u32();   // numrecs
        uint32_t tag = u32(), n = u32();
        if (tag == TAG_DIMENSION)
            for (uint32_t d = 0; d < n && !fail; ++d) {
                std::string s = name();
                dims.push_back({s, u32()});
            }
        skip_atts();
        tag = u32(); n = u32();
        if (tag == TAG_VARIABLE)
            for (uint32_t v = 0; v < n && !fail; ++v) {
                Var var;
                var.name = name();
                const uint32_t nd = u32();
                for (uint32_t d = 0; d < nd && !fail; ++d) var.dims.push_back(u32());
                skip_atts();
                var.type = u32();
                u32();   // vsize
                var.begin = version == 2 ? u64() : u32();
                vars.push_back(var);
            }
        return !fail;
    }
    const Var* find(const std::string& n) const {
        for (const Var& v : vars)
            if (v.name == n) return &v;
        return nullptr;
    }
    uint32_t dim_len(const std::string& n) const {
        for (const auto& d : dims)
            if (d.first == n) return d.second;
        return 0;
    }
    // Index of the named dimension, or -1.
    int dim_id(const std::string& n) const {
        for (size_t k = 0; k < dims.size(); ++k)
            if (dims[k].first == n) return (int)k;
        return -1;
    }
    bool read_doubles(const Var& v, size_t count, double* out) const {
        if (v.type != NC_DOUBLE || v.begin > buf.size() || count > (buf.size() - v.begin) / 8) return false;
        const uint8_t* p = buf.data() + v.begin;
        for (size_t k = 0; k < count; ++k) {
            uint64_t u = 0;
            for (int b = 0; b < 8; ++b) u = (u << 8) | p[8 * k + b];
            std::memcpy(&out[k], &u, 8);
        }
        return true;
    }
};

inline bool load_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    uint8_t chunk[65536];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof chunk, fp)) > 0) out.insert(out.end(), chunk, chunk + r);
    std::fclose(fp);
    return true;
}

}  // namespace nc
}  // namespace rk
