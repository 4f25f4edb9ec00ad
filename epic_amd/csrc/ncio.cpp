// NetCDF classic files of EPIC's parcel and field output (parcel_netcdf.f90,
// field_netcdf.f90, netcdf_writer.f90), written and read without a netCDF
// library.
//
// Format: the NetCDF classic binary format (CDF-1 "CDF\1", CDF-2 64-bit
// offset "CDF\2" and CDF-5 64-bit data "CDF\5"; big-endian, header of
// dimension, attribute and variable lists, fixed-size variables then
// interleaved records of the unlimited dimension).  Written files are CDF-2,
// or CDF-5 when a variable or a record would not fit CDF-2.
//
// The reference writes NetCDF-4 files, whose groups ('parameters',
// 'physical_quantities') the classic format lacks: their attributes are
// written as global attributes named "<group>.<attribute>".  The variables,
// dimensions, units and long names are the reference's.
//
// Host code on top of the library's C-ABI (parcels and fields are fetched
// with epic_get_parcels / epic_get_field and uploaded with
// epic_set_parcels).
#include "epic_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// records a failure message on the handle (abi.hip)
int epic_internal_fail(epic_handle h, int code, const std::string &what);

namespace {

enum NcType { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6, NC_INT64 = 10 };
constexpr uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;

int type_size(int t)
{
    switch (t) {
    case NC_BYTE:
    case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT:
    case NC_FLOAT: return 4;
    case NC_DOUBLE:
    case NC_INT64: return 8;
    default: return 0;
    }
}

int64_t pad4(int64_t n) { return (n + 3) & ~int64_t(3); }

// ---- big-endian encoding ----
struct Out {
    std::vector<uint8_t> b;
    void u32(uint32_t v)
    {
        for (int i = 3; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void u64(uint64_t v)
    {
        for (int i = 7; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void bytes(const void *p, size_t n)
    {
        const uint8_t *c = (const uint8_t *)p;
        b.insert(b.end(), c, c + n);
    }
    void pad()
    {
        while (b.size() % 4) b.push_back(0);
    }
};

void put_be(uint8_t *dst, const void *src, int size)
{
    const uint8_t *s = (const uint8_t *)src;
    for (int i = 0; i < size; ++i) dst[i] = s[size - 1 - i];   // the host is little-endian (x86-64)
}

struct Att {
    std::string name;
    int type;
    int64_t n;
    std::vector<uint8_t> data;   // big-endian values
};

Att att_text(const std::string &name, const std::string &v)
{
    return Att{name, NC_CHAR, (int64_t)v.size(), std::vector<uint8_t>(v.begin(), v.end())};
}
Att att_doubles(const std::string &name, const double *v, int n)
{
    Att a{name, NC_DOUBLE, n, std::vector<uint8_t>((size_t)n * 8)};
    for (int i = 0; i < n; ++i) put_be(a.data.data() + 8 * i, v + i, 8);
    return a;
}
Att att_ints(const std::string &name, const int *v, int n)
{
    Att a{name, NC_INT, n, std::vector<uint8_t>((size_t)n * 4)};
    for (int i = 0; i < n; ++i) put_be(a.data.data() + 4 * i, v + i, 4);
    return a;
}

struct Dim {
    std::string name;
    int64_t len;   // 0: the unlimited (record) dimension
};

struct Var {
    std::string name;
    int type;
    std::vector<int> dims;
    std::vector<Att> atts;
    int64_t vsize = 0, begin = 0;
    bool rec = false;
};

// ---- writer ----
class Writer {
  public:
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    int64_t numrecs = 0;

    int dim(const std::string &n, int64_t len)
    {
        dims.push_back(Dim{n, len});
        return (int)dims.size() - 1;
    }
    int var(const std::string &n, int type, std::vector<int> d, const std::string &units,
            const std::string &long_name, const std::string &std_name = "")
    {
        Var v;
        v.name = n;
        v.type = type;
        v.dims = std::move(d);
        v.atts.push_back(att_text("units", units));   // define_netcdf_dataset (netcdf_writer.f90:96-128)
        if (!long_name.empty()) v.atts.push_back(att_text("long_name", long_name));
        if (!std_name.empty()) v.atts.push_back(att_text("standard_name", std_name));
        vars.push_back(std::move(v));
        return (int)vars.size() - 1;
    }

    // sizes and offsets of the variables; CDF-5 when a size does not fit CDF-2
    void layout(int &version, int64_t &recsize)
    {
        for (Var &v : vars) {
            int64_t sz = type_size(v.type);
            v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
            for (size_t i = v.rec ? 1 : 0; i < v.dims.size(); ++i) sz *= dims[v.dims[i]].len;
            v.vsize = pad4(sz);
        }
        version = 2;
        for (const Var &v : vars)
            if (v.vsize >= (int64_t(1) << 32) - 4) version = 5;
        recsize = 0;
        for (const Var &v : vars)
            if (v.rec) recsize += v.vsize;   // every type used here is a multiple of 4 bytes
        int64_t off = header(version).size();
        for (Var &v : vars)
            if (!v.rec) {
                v.begin = off;
                off += v.vsize;
            }
        for (Var &v : vars)
            if (v.rec) {
                v.begin = off;
                off += v.vsize;
            }
    }

    std::vector<uint8_t> header(int version)
    {
        Out o;
        const bool w = (version == 5);
        auto count = [&](uint64_t n) { w ? o.u64(n) : o.u32((uint32_t)n); };
        auto name = [&](const std::string &s) {
            count(s.size());
            o.bytes(s.data(), s.size());
            o.pad();
        };
        auto atts = [&](const std::vector<Att> &as) {
            if (as.empty()) {
                o.u32(0);
                count(0);
                return;
            }
            o.u32(TAG_ATT);
            count(as.size());
            for (const Att &a : as) {
                name(a.name);
                o.u32((uint32_t)a.type);
                count((uint64_t)a.n);
                o.bytes(a.data.data(), a.data.size());
                o.pad();
            }
        };
        o.bytes("CDF", 3);
        o.b.push_back((uint8_t)version);
        count((uint64_t)numrecs);
        if (dims.empty()) {
            o.u32(0);
            count(0);
        } else {
            o.u32(TAG_DIM);
            count(dims.size());
            for (const Dim &d : dims) {
                name(d.name);
                count((uint64_t)d.len);
            }
        }
        atts(gatts);
        if (vars.empty()) {
            o.u32(0);
            count(0);
        } else {
            o.u32(TAG_VAR);
            count(vars.size());
            for (const Var &v : vars) {
                name(v.name);
                count(v.dims.size());
                for (int d : v.dims) count((uint64_t)d);
                atts(v.atts);
                o.u32((uint32_t)v.type);
                count((uint64_t)v.vsize);
                if (version == 1) o.u32((uint32_t)v.begin);
                else o.u64((uint64_t)v.begin);
            }
        }
        return o.b;
    }
};

// element conversion to big-endian bytes
void encode(const double *src, int type, int64_t n, std::vector<uint8_t> &out)
{
    const int sz = type_size(type);
    out.assign((size_t)pad4(n * sz), 0);
    for (int64_t i = 0; i < n; ++i) {
        if (type == NC_DOUBLE) put_be(out.data() + 8 * i, src + i, 8);
        else if (type == NC_INT) {
            const int32_t v = (int32_t)std::lround(src[i]);
            put_be(out.data() + 4 * i, &v, 4);
        } else if (type == NC_FLOAT) {
            const float v = (float)src[i];
            put_be(out.data() + 4 * i, &v, 4);
        }
    }
}

// a written file: header plus data placed at the variables' offsets
struct Image {
    FILE *f = nullptr;
    ~Image()
    {
        if (f) fclose(f);
    }
};

bool write_at(FILE *f, int64_t off, const std::vector<uint8_t> &b)
{
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
    return fwrite(b.data(), 1, b.size(), f) == b.size();
}

// ---- reader ----
struct In {
    const uint8_t *p, *e;
    bool ok = true;
    uint64_t u(int bytes)
    {
        if (p + bytes > e) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 0; i < bytes; ++i) v = (v << 8) | p[i];
        p += bytes;
        return v;
    }
    std::string str(uint64_t n)
    {
        if (p + n > e) {
            ok = false;
            return "";
        }
        std::string s((const char *)p, (size_t)n);
        p += pad4((int64_t)n);
        if (p > e) ok = false;
        return s;
    }
};

class Reader {
  public:
    int version = 0;
    int64_t numrecs = 0, recsize = 0;
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    FILE *f = nullptr;
    std::string err;

    ~Reader()
    {
        if (f) fclose(f);
    }

    bool open(const char *path)
    {
        f = fopen(path, "rb");
        if (!f) return fail("cannot open the file");
        std::vector<uint8_t> buf;
        // the header is at most a few hundred KB for these schemas; read growing chunks
        size_t want = 1 << 16;
        for (;;) {
            buf.resize(want);
            if (fseeko(f, 0, SEEK_SET) != 0) return fail("seek failed");
            const size_t got = fread(buf.data(), 1, want, f);
            buf.resize(got);
            if (parse(buf)) return true;
            if (got < want) return fail(err.empty() ? "truncated or invalid header" : err);
            want *= 4;
            err.clear();
        }
    }

    int find_var(const std::string &n) const
    {
        for (size_t i = 0; i < vars.size(); ++i)
            if (vars[i].name == n) return (int)i;
        return -1;
    }
    int find_dim(const std::string &n) const
    {
        for (size_t i = 0; i < dims.size(); ++i)
            if (dims[i].name == n) return (int)i;
        return -1;
    }
    const Att *gatt(const std::string &n) const
    {
        for (const Att &a : gatts)
            if (a.name == n) return &a;
        return nullptr;
    }
    static bool att_values(const Att *a, double *out, int n)
    {
        if (!a || a->n != n) return false;
        for (int i = 0; i < n; ++i) {
            uint8_t le[8];
            const int sz = type_size(a->type);
            for (int k = 0; k < sz; ++k) le[k] = a->data[(size_t)i * sz + sz - 1 - k];
            if (a->type == NC_DOUBLE) std::memcpy(&out[i], le, 8);
            else if (a->type == NC_INT) {
                int32_t v;
                std::memcpy(&v, le, 4);
                out[i] = v;
            } else if (a->type == NC_FLOAT) {
                float v;
                std::memcpy(&v, le, 4);
                out[i] = v;
            } else return false;
        }
        return true;
    }

    // elements of one record (or the whole fixed-size variable) as doubles
    bool read(int vi, int64_t rec, std::vector<double> &out)
    {
        const Var &v = vars[vi];
        int64_t n = 1;
        for (size_t i = v.rec ? 1 : 0; i < v.dims.size(); ++i) n *= dims[v.dims[i]].len;
        const int sz = type_size(v.type);
        if (sz == 0) return fail("variable '" + v.name + "' has an unsupported type");
        if (v.rec && (rec < 0 || rec >= numrecs)) return fail("record out of range");
        const int64_t off = v.begin + (v.rec ? rec * recsize : 0);
        std::vector<uint8_t> b((size_t)(n * sz));
        if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fread(b.data(), 1, b.size(), f) != b.size())
            return fail("cannot read variable '" + v.name + "'");
        out.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            uint8_t le[8];
            for (int k = 0; k < sz; ++k) le[k] = b[(size_t)(i * sz + sz - 1 - k)];
            switch (v.type) {
            case NC_DOUBLE: std::memcpy(&out[i], le, 8); break;
            case NC_FLOAT: {
                float x;
                std::memcpy(&x, le, 4);
                out[i] = x;
                break;
            }
            case NC_INT: {
                int32_t x;
                std::memcpy(&x, le, 4);
                out[i] = x;
                break;
            }
            case NC_SHORT: {
                int16_t x;
                std::memcpy(&x, le, 2);
                out[i] = x;
                break;
            }
            case NC_INT64: {
                int64_t x;
                std::memcpy(&x, le, 8);
                out[i] = (double)x;
                break;
            }
            default: out[i] = (int8_t)le[0];
            }
        }
        return true;
    }

  private:
    bool fail(const std::string &m)
    {
        err = m;
        return false;
    }

    bool parse(const std::vector<uint8_t> &buf)
    {
        In in{buf.data(), buf.data() + buf.size()};
        if (buf.size() < 4 || std::memcmp(buf.data(), "CDF", 3) != 0) return fail("not a NetCDF classic file");
        version = buf[3];
        if (version != 1 && version != 2 && version != 5)
            return fail("unsupported NetCDF format (NetCDF-4/HDF5 files are not readable here)");
        in.p += 4;
        const int cw = (version == 5) ? 8 : 4;
        numrecs = (int64_t)in.u(cw);
        dims.clear();
        gatts.clear();
        vars.clear();
        auto atts = [&](std::vector<Att> &as) {
            const uint32_t tag = (uint32_t)in.u(4);
            const uint64_t n = in.u(cw);
            if (tag != TAG_ATT && !(tag == 0 && n == 0)) in.ok = false;
            for (uint64_t i = 0; i < n && in.ok; ++i) {
                Att a;
                a.name = in.str(in.u(cw));
                a.type = (int)in.u(4);
                a.n = (int64_t)in.u(cw);
                const int64_t bytes = a.n * type_size(a.type);
                if (type_size(a.type) == 0 || in.p + bytes > in.e) {
                    in.ok = false;
                    break;
                }
                a.data.assign(in.p, in.p + bytes);
                in.p += pad4(bytes);
                as.push_back(std::move(a));
            }
        };
        uint32_t tag = (uint32_t)in.u(4);
        uint64_t n = in.u(cw);
        if (tag != TAG_DIM && !(tag == 0 && n == 0)) return fail("bad dimension list");
        for (uint64_t i = 0; i < n && in.ok; ++i) {
            Dim d;
            d.name = in.str(in.u(cw));
            d.len = (int64_t)in.u(cw);
            dims.push_back(d);
        }
        atts(gatts);
        tag = (uint32_t)in.u(4);
        n = in.u(cw);
        if (tag != TAG_VAR && !(tag == 0 && n == 0)) return fail("bad variable list");
        int nrec = 0;
        for (uint64_t i = 0; i < n && in.ok; ++i) {
            Var v;
            v.name = in.str(in.u(cw));
            const uint64_t nd = in.u(cw);
            for (uint64_t k = 0; k < nd && in.ok; ++k) {
                const uint64_t d = in.u(cw);
                if (d >= dims.size()) in.ok = false;
                else v.dims.push_back((int)d);
            }
            atts(v.atts);
            v.type = (int)in.u(4);
            v.vsize = (int64_t)in.u(cw);
            v.begin = (int64_t)in.u(version == 1 ? 4 : 8);
            v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
            nrec += v.rec;
            vars.push_back(std::move(v));
        }
        if (!in.ok) return false;
        recsize = 0;
        for (const Var &v : vars)
            if (v.rec) {
                int64_t sz = type_size(v.type);
                for (size_t k = 1; k < v.dims.size(); ++k) sz *= dims[v.dims[k]].len;
                recsize += (nrec == 1) ? sz : pad4(sz);   // a single record variable is not padded
            }
        return true;
    }
};

// ---- the reference's schema ----
const char *VERSION_NAME = "epic_version";   // package // '_version' (netcdf_utils.f90:19)
const char *VERSION_TAG = "epic_amd (EPIC 0.14.3 file layout)";
const char *CF_VERSION = "CF-1.8";

void common_header(Writer &w, const epic_grid_t &g, const epic_phys_t &ph, const epic_flags_t &fl,
                   const char *file_type)
{
    w.gatts.push_back(att_text(VERSION_NAME, VERSION_TAG));   // write_netcdf_info (netcdf_writer.f90:453-462)
    w.gatts.push_back(att_text("file_type", file_type));
    w.gatts.push_back(att_text("Conventions", CF_VERSION));
    const int nc[3] = {g.nx, g.ny, g.nz};
    w.gatts.push_back(att_ints("ncells", nc, 3));   // write_netcdf_box (:374-387)
    w.gatts.push_back(att_doubles("extent", g.extent, 3));
    w.gatts.push_back(att_doubles("origin", g.lower, 3));
    // group attributes (write_zeta_boundary_flag, parameters.f90:253-265;
    // write_physical_quantities, physics.f90:248-277), flattened
    w.gatts.push_back(att_text("parameters.l_lower_boundry_zeta_zero", fl.zeta_bndry_zero[0] ? "true" : "false"));
    w.gatts.push_back(att_text("parameters.l_upper_boundry_zeta_zero", fl.zeta_bndry_zero[1] ? "true" : "false"));
    w.gatts.push_back(att_doubles("physical_quantities.saturation_specific_humidity_at_ground_level", &ph.q0, 1));
    const double hc = (ph.lambda_c > 0.0) ? 1.0 / ph.lambda_c : 0.0;
    w.gatts.push_back(att_doubles("physical_quantities.scale_height", &hc, 1));
    w.gatts.push_back(att_doubles("physical_quantities.glat", &ph.glat, 1));
    w.gatts.push_back(att_doubles("physical_quantities.planetary_vorticity", ph.f_cor, 3));
}

int time_axis(Writer &w)
{   // define_netcdf_temporal_dimension (netcdf_writer.f90:569-594)
    const int t = w.dim("t", 0);
    const int v = w.var("t", NC_DOUBLE, {t}, "seconds since 1970-01-01", "time ", "time");
    w.vars[v].atts.push_back(att_text("axis", "T"));
    w.vars[v].atts.push_back(att_text("calendar", "proleptic_gregorian"));
    return t;
}

thread_local std::string g_err;   // message of the last failure of the handle-free calls

int fail(const std::string &what)
{
    g_err = what;
    return EPIC_E_INVALID;
}

// write the header and every variable (record 0 for record variables)
int write_file(const char *path, Writer &w, const std::vector<std::vector<double>> &data)
{
    int version;
    int64_t recsize;
    w.numrecs = 1;
    w.layout(version, recsize);
    FILE *f = fopen(path, "wb");
    if (!f) return fail(std::string("cannot create ") + path);
    Image img{f};
    if (!write_at(f, 0, w.header(version))) return fail(std::string("cannot write ") + path);
    std::vector<uint8_t> b;
    for (size_t i = 0; i < w.vars.size(); ++i) {
        const Var &v = w.vars[i];
        int64_t n = 1;
        for (size_t k = v.rec ? 1 : 0; k < v.dims.size(); ++k) n *= w.dims[v.dims[k]].len;
        if ((int64_t)data[i].size() != n) return fail("internal: data size of " + v.name);
        encode(data[i].data(), v.type, n, b);
        if (!write_at(f, v.begin, b)) return fail(std::string("cannot write ") + path);
    }
    return EPIC_OK;
}

}  // namespace

extern "C" {

const char *epic_nc_last_error(void) { return g_err.c_str(); }

int epic_nc_write_parcels(const char *path, const epic_grid_t *grid, const epic_phys_t *phys,
                          const epic_flags_t *flags, int64_t n, const double *position, const double *vorticity,
                          const double *B, const double *volume, const double *buoyancy, const double *humidity,
                          double t)
{
    if (!path || !grid || !phys || !flags || n < 0) return fail("epic_nc_write_parcels: invalid argument");
    if (n > 2147483647) return fail("epic_nc_write_parcels: the parcel count exceeds the integer limit");
    if (n > 0 && (!position || !vorticity || !B || !volume || !buoyancy || !humidity))
        return fail("epic_nc_write_parcels: null array");
    const epic_grid_t &g = *grid;
    Writer w;
    common_header(w, g, *phys, *flags, "parcels");
    const int dn = w.dim("n_parcels", n), dw = w.dim("world%size", 1);
    const int dt = time_axis(w);
    std::vector<std::vector<double>> data;
    data.push_back({t});   // t
    // per-rank datasets (parcel_netcdf.f90:170-180, 268-291), one rank
    const struct {
        const char *name, *long_name;
        double v;
    } rank_vars[] = {{"start_index", "MPI rank start index", 1.0},
                     {"xlo", "lower box boundary in x", 0.0},
                     {"xhi", "upper box boundary in x", (double)(g.nx - 1)},
                     {"ylo", "lower box boundary in y", 0.0},
                     {"yhi", "upper box boundary in y", (double)(g.ny - 1)}};
    for (const auto &rv : rank_vars) {
        w.var(rv.name, NC_INT, {dw}, "1", rv.long_name);
        data.push_back({rv.v});
    }
    // parcel datasets in the order of the file (parcel_netcdf.f90:856-975)
    auto column = [&](const double *a, int stride, int c) {
        std::vector<double> o((size_t)n);
        for (int64_t q = 0; q < n; ++q) o[q] = a[stride * q + c];
        return o;
    };
    const struct {
        const char *name, *long_name, *unit;
        const double *src;
        int stride, comp;
    } pv[] = {{"volume", "parcel volume", "m^3", volume, 1, 0},
              {"x_position", "x position component", "m", position, 3, 0},
              {"y_position", "y position component", "m", position, 3, 1},
              {"z_position", "z position component", "m", position, 3, 2},
              {"buoyancy", "parcel buoyancy", "m/s^2", buoyancy, 1, 0},
              {"x_vorticity", "x vorticity component", "1/s", vorticity, 3, 0},
              {"y_vorticity", "y vorticity component", "1/s", vorticity, 3, 1},
              {"z_vorticity", "z vorticity component", "1/s", vorticity, 3, 2},
              {"B11", "B11 element of shape matrix", "m^2", B, 5, 0},
              {"B12", "B12 element of shape matrix", "m^2", B, 5, 1},
              {"B13", "B13 element of shape matrix", "m^2", B, 5, 2},
              {"B22", "B22 element of shape matrix", "m^2", B, 5, 3},
              {"B23", "B23 element of shape matrix", "m^2", B, 5, 4},
              {"humidity", "parcel humidity", "1", humidity, 1, 0}};
    for (const auto &p : pv) {
        w.var(p.name, NC_DOUBLE, {dt, dn}, p.unit, p.long_name);
        data.push_back(column(p.src, p.stride, p.comp));
    }
    return write_file(path, w, data);
}

int epic_write_parcels_nc(epic_handle h, const char *path, double t)
{
    if (!h || !path) return EPIC_E_INVALID;
    const int64_t n = epic_num_parcels(h);
    epic_grid_t g;
    epic_phys_t ph;
    epic_flags_t fl;
    epic_get_config(h, &g, &ph, &fl);
    std::vector<double> pos(3 * n), vor(3 * n), B(5 * n), vol(n), buoy(n), hum(n);
    int rc = epic_get_parcels(h, n, pos.data(), vor.data(), B.data(), vol.data(), buoy.data(), hum.data());
    if (rc) return rc;
    rc = epic_nc_write_parcels(path, &g, &ph, &fl, n, pos.data(), vor.data(), B.data(), vol.data(), buoy.data(),
                               hum.data(), t);
    return rc ? epic_internal_fail(h, rc, g_err) : EPIC_OK;
}

int epic_nc_read_parcels(const char *path, int64_t capacity, int64_t *n_out, epic_grid_t *grid, double *position,
                         double *vorticity, double *B, double *volume, double *buoyancy, double *humidity, double *t)
{
    if (!path || !n_out) return fail("epic_nc_read_parcels: invalid argument");
    Reader r;
    if (!r.open(path)) return fail(std::string("epic_nc_read_parcels: ") + path + ": " + r.err);
    const int dn = r.find_dim("n_parcels");
    if (dn < 0) return fail("epic_nc_read_parcels: no 'n_parcels' dimension");
    const int64_t n = r.dims[dn].len;
    *n_out = n;
    if (grid) {   // write_netcdf_box attributes; a file without them reports nx = ny = nz = 0
        double nc[3] = {0, 0, 0}, ext[3] = {0, 0, 0}, org[3] = {0, 0, 0};
        if (!Reader::att_values(r.gatt("ncells"), nc, 3) || !Reader::att_values(r.gatt("extent"), ext, 3) ||
            !Reader::att_values(r.gatt("origin"), org, 3))
            nc[0] = nc[1] = nc[2] = ext[0] = ext[1] = ext[2] = org[0] = org[1] = org[2] = 0.0;
        grid->nx = (int)nc[0];
        grid->ny = (int)nc[1];
        grid->nz = (int)nc[2];
        for (int d = 0; d < 3; ++d) {
            grid->extent[d] = ext[d];
            grid->lower[d] = org[d];
        }
    }
    const int64_t rec = r.numrecs > 0 ? r.numrecs - 1 : 0;   // the last record
    if (t) {
        *t = 0.0;
        const int vt = r.find_var("t");
        std::vector<double> tv;
        if (vt >= 0 && r.read(vt, rec, tv) && !tv.empty()) *t = tv[0];
    }
    if (!position) return EPIC_OK;   // size query
    if (n > capacity) return fail("epic_nc_read_parcels: more parcels in the file than the arrays hold");
    // required and optional datasets (parcel_netcdf.f90:595-700)
    struct Ds {
        const char *name;
        bool required;
        double *dst;
        int stride, comp;
    } ds[] = {{"B11", true, B, 5, 0},          {"B12", true, B, 5, 1},          {"B13", true, B, 5, 2},
              {"B22", true, B, 5, 3},          {"B23", true, B, 5, 4},          {"x_position", true, position, 3, 0},
              {"y_position", true, position, 3, 1}, {"z_position", true, position, 3, 2}, {"volume", true, volume, 1, 0},
              {"x_vorticity", false, vorticity, 3, 0}, {"y_vorticity", false, vorticity, 3, 1},
              {"z_vorticity", false, vorticity, 3, 2}, {"buoyancy", false, buoyancy, 1, 0},
              {"humidity", false, humidity, 1, 0}};
    std::vector<double> v;
    for (const Ds &d : ds) {
        if (!d.dst) return fail("epic_nc_read_parcels: null array");
        const int vi = r.find_var(d.name);
        if (vi < 0) {
            if (d.required) return fail(std::string("epic_nc_read_parcels: The parcel ") + d.name + " must be provided.");
            for (int64_t q = 0; q < n; ++q) d.dst[d.stride * q + d.comp] = 0.0;
            continue;
        }
        if (!r.read(vi, rec, v) || (int64_t)v.size() != n)
            return fail(std::string("epic_nc_read_parcels: ") + (r.err.empty() ? d.name : r.err));
        for (int64_t q = 0; q < n; ++q) d.dst[d.stride * q + d.comp] = v[q];
    }
    return EPIC_OK;
}

int epic_read_parcels_nc(epic_handle h, const char *path, double *t)
{
    if (!h || !path) return EPIC_E_INVALID;
    int64_t n = 0;
    epic_grid_t gf, g;
    int rc = epic_nc_read_parcels(path, 0, &n, &gf, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, t);
    if (rc) return epic_internal_fail(h, rc, g_err);
    epic_get_config(h, &g, nullptr, nullptr);
    for (int d = 0; d < 3 && gf.nx > 0; ++d)   // a file with a box must match the member's grid
        if ((&gf.nx)[d] != (&g.nx)[d] || std::fabs(gf.extent[d] - g.extent[d]) > 1e-12 * std::fabs(g.extent[d]) ||
            std::fabs(gf.lower[d] - g.lower[d]) > 1e-12 * (std::fabs(g.lower[d]) + std::fabs(g.extent[d])))
            return epic_internal_fail(h, EPIC_E_INVALID, "epic_read_parcels_nc: the file's box differs from the member's grid");
    std::vector<double> pos(3 * n), vor(3 * n), B(5 * n), vol(n), buoy(n), hum(n);
    rc = epic_nc_read_parcels(path, n, &n, nullptr, pos.data(), vor.data(), B.data(), vol.data(), buoy.data(),
                              hum.data(), t);
    if (rc) return epic_internal_fail(h, rc, g_err);
    return epic_set_parcels(h, n, pos.data(), vor.data(), B.data(), vol.data(), buoy.data(), hum.data());
}

int epic_nc_write_fields(const char *path, const epic_grid_t *grid, const epic_phys_t *phys,
                         const epic_flags_t *flags, double t, const double *velog, const double *vtend,
                         const double *vortg, const double *tbuoyg, const double *dbuoyg, const double *humg,
                         const double *volg, const double *nparg, const double *nsparg)
{
    if (!path || !grid || !phys || !flags || !velog || !vtend || !vortg || !tbuoyg || !dbuoyg || !humg || !volg ||
        !nparg || !nsparg)
        return fail("epic_nc_write_fields: invalid argument");
    const epic_grid_t &g = *grid;
    const epic_phys_t &ph = *phys;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t H = (int64_t)(nz + 3) * (ny + 3) * (nx + 3), N = (int64_t)(nz + 1) * ny * nx;
    Writer w;
    common_header(w, g, ph, *flags, "fields");
    // define_netcdf_spatial_dimensions_3d (netcdf_writer.f90:508-567)
    const int dx = w.dim("x", nx), dy = w.dim("y", ny), dz = w.dim("z", nz + 1);   // ngps (nx, ny, nz+1)
    std::vector<std::vector<double>> data;
    const char *axn[3] = {"x", "y", "z"}, *axa[3] = {"X", "Y", "Z"};
    const int ad[3] = {dx, dy, dz}, an[3] = {nx, ny, nz + 1};
    const double h3[3] = {g.extent[0] / nx, g.extent[1] / ny, g.extent[2] / nz};
    for (int d = 0; d < 3; ++d) {
        const std::string ln = (d < 2) ? std::string(axn[d]) + "-coordinate in projected coordinate system"
                                       : "height coordinate in projected coordinate system";
        const std::string sn = (d < 2) ? std::string("projection_") + axn[d] + "_coordinate" : "height";
        const int v = w.var(axn[d], NC_DOUBLE, {ad[d]}, "m", ln, sn);
        w.vars[v].atts.push_back(att_text("axis", axa[d]));
        std::vector<double> c((size_t)an[d]);
        for (int i = 0; i < an[d]; ++i) c[i] = g.lower[d] + i * h3[d];   // write_netcdf_axis_3d (:409-451)
        data.push_back(c);
    }
    const int dt = time_axis(w);
    data.push_back({t});
    // a field on the nodes (x fastest, z planes 0..nz) from the halo layout
    auto interior = [&](const double *f, int comp) {
        std::vector<double> o((size_t)N);
        for (int iz = 0; iz <= nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    o[((size_t)iz * ny + iy) * nx + ix] =
                        f[(size_t)(comp * H + ((int64_t)(ix + 1) * (ny + 3) + (iy + 1)) * (nz + 3) + (iz + 1))];
        return o;
    };
    std::vector<int> dims4 = {dt, dz, dy, dx};
    auto add = [&](const char *name, const char *ln, const char *unit, int type, std::vector<double> v) {
        w.var(name, type, dims4, unit, ln);
        data.push_back(std::move(v));
    };
    // the datasets of write_netcdf_fields (field_netcdf.f90:226-256, names :465-561)
    add("x_velocity", "x velocity component", "m/s", NC_DOUBLE, interior(velog, 0));
    add("y_velocity", "y velocity component", "m/s", NC_DOUBLE, interior(velog, 1));
    add("z_velocity", "z velocity component", "m/s", NC_DOUBLE, interior(velog, 2));
    add("x_vorticity_tendency", "x vorticity tendency", "1/s", NC_DOUBLE, interior(vtend, 0));
    add("y_vorticity_tendency", "y vorticity tendency", "1/s", NC_DOUBLE, interior(vtend, 1));
    add("z_vorticity_tendency", "z vorticity tendency", "1/s", NC_DOUBLE, interior(vtend, 2));
    add("nparg", "number of parcels per cell", "1", NC_INT, interior(nparg, 0));
    add("nsparg", "number of small parcels per cell", "1", NC_INT, interior(nsparg, 0));
    add("x_vorticity", "x vorticity component", "1/s", NC_DOUBLE, interior(vortg, 0));
    add("y_vorticity", "y vorticity component", "1/s", NC_DOUBLE, interior(vortg, 1));
    add("z_vorticity", "z vorticity component", "1/s", NC_DOUBLE, interior(vortg, 2));
    add("buoyancy", "total buoyancy", "m/s^2", NC_DOUBLE, interior(tbuoyg, 0));
    add("dry_buoyancy", "dry buoyancy", "m/s^2", NC_DOUBLE, interior(dbuoyg, 0));
    {   // glati (tbuoyg - dbuoyg)
        std::vector<double> l = interior(tbuoyg, 0), d = interior(dbuoyg, 0);
        const double gi = (ph.glat != 0.0) ? 1.0 / ph.glat : 0.0;
        for (size_t i = 0; i < l.size(); ++i) l[i] = gi * (l[i] - d[i]);
        add("liquid_water_content", "liquid-water content", "1", NC_DOUBLE, l);
    }
    add("humidity", "specific humidity", "kg/kg", NC_DOUBLE, interior(humg, 0));
    add("volume", "volume", "m^3", NC_DOUBLE, interior(volg, 0));
    return write_file(path, w, data);
}

int epic_nc_file_info(const char *path, char *file_type, double *t, int64_t *n_steps, epic_grid_t *grid)
{
    if (!path) return fail("epic_nc_file_info: invalid argument");
    Reader r;
    if (!r.open(path)) return fail(std::string("epic_nc_file_info: ") + path + ": " + r.err);
    if (file_type) {   // get_file_type (netcdf_reader.f90): the global attribute 'file_type'
        file_type[0] = '\0';
        const Att *a = r.gatt("file_type");
        if (a && a->type == NC_CHAR) {
            const size_t n = std::min<size_t>(a->data.size(), 15);
            std::memcpy(file_type, a->data.data(), n);
            file_type[n] = '\0';
        }
    }
    if (n_steps) *n_steps = r.numrecs;   // get_num_steps
    if (t) {   // get_time: the last record's time
        *t = 0.0;
        const int vt = r.find_var("t");
        std::vector<double> tv;
        if (vt >= 0 && r.numrecs > 0 && r.read(vt, r.numrecs - 1, tv) && !tv.empty()) *t = tv[0];
    }
    if (grid) {   // get_netcdf_box; nx = ny = nz = 0 without the box attributes
        double nc[3] = {0, 0, 0}, ext[3] = {0, 0, 0}, org[3] = {0, 0, 0};
        if (!Reader::att_values(r.gatt("ncells"), nc, 3) || !Reader::att_values(r.gatt("extent"), ext, 3) ||
            !Reader::att_values(r.gatt("origin"), org, 3))
            nc[0] = nc[1] = nc[2] = ext[0] = ext[1] = ext[2] = org[0] = org[1] = org[2] = 0.0;
        grid->nx = (int)nc[0];
        grid->ny = (int)nc[1];
        grid->nz = (int)nc[2];
        for (int d = 0; d < 3; ++d) {
            grid->extent[d] = ext[d];
            grid->lower[d] = org[d];
        }
    }
    return EPIC_OK;
}

int epic_nc_read_fields(const char *path, int step, const epic_grid_t *grid, double *vortg, double *tbuoyg,
                        double *humg, double *t)
{
    if (!path || !grid || !vortg || !tbuoyg || !humg) return fail("epic_nc_read_fields: invalid argument");
    Reader r;
    if (!r.open(path)) return fail(std::string("epic_nc_read_fields: ") + path + ": " + r.err);
    const int64_t n_steps = r.numrecs;
    int64_t st = step;
    if (st == -1) st = n_steps;
    else if (st == 0 || st > n_steps) return fail("epic_nc_read_fields: Step number in NetCDF field file is out of bounds.");
    const int64_t rec = st - 1;
    const int nx = grid->nx, ny = grid->ny, nz = grid->nz;
    const int64_t H = (int64_t)(nz + 3) * (ny + 3) * (nx + 3), N = (int64_t)(nz + 1) * ny * nx;
    if (t) {
        *t = 0.0;
        const int vt = r.find_var("t");
        std::vector<double> tv;
        if (vt >= 0 && r.read(vt, rec, tv) && !tv.empty()) *t = tv[0];
    }
    // a present dataset fills planes 0..nz of its field; an absent one
    // leaves the field zero (field_default)
    struct Ds {
        const char *name;
        double *dst;
        int comp;
    } ds[] = {{"x_vorticity", vortg, 0}, {"y_vorticity", vortg, 1}, {"z_vorticity", vortg, 2},
              {"buoyancy", tbuoyg, 0}, {"humidity", humg, 0}};
    std::vector<double> v;
    for (const Ds &d : ds) {
        double *f = d.dst + d.comp * H;
        for (int64_t i = 0; i < H; ++i) f[i] = 0.0;
        const int vi = r.find_var(d.name);
        if (vi < 0) continue;
        if (!r.read(vi, r.vars[vi].rec ? rec : 0, v) || (int64_t)v.size() != N)
            return fail(std::string("epic_nc_read_fields: ") + (r.err.empty() ? std::string(d.name) + ": wrong size" : r.err));
        for (int iz = 0; iz <= nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    f[((int64_t)(ix + 1) * (ny + 3) + (iy + 1)) * (nz + 3) + (iz + 1)] =
                        v[((size_t)iz * ny + iy) * nx + ix];
    }
    return EPIC_OK;
}

int epic_read_fields_nc(epic_handle h, const char *path, int step, double *t)
{
    if (!h || !path) return EPIC_E_INVALID;
    epic_grid_t g, gf;
    epic_get_config(h, &g, nullptr, nullptr);
    int rc = epic_nc_file_info(path, nullptr, nullptr, nullptr, &gf);
    if (rc) return epic_internal_fail(h, rc, g_err);
    for (int d = 0; d < 3 && gf.nx > 0; ++d)   // a file with a box must match the member's grid
        if ((&gf.nx)[d] != (&g.nx)[d] || std::fabs(gf.extent[d] - g.extent[d]) > 1e-12 * std::fabs(g.extent[d]) ||
            std::fabs(gf.lower[d] - g.lower[d]) > 1e-12 * (std::fabs(g.lower[d]) + std::fabs(g.extent[d])))
            return epic_internal_fail(h, EPIC_E_INVALID, "epic_read_fields_nc: the file's box differs from the member's grid");
    const int64_t H = (int64_t)(g.nz + 3) * (g.ny + 3) * (g.nx + 3);
    std::vector<double> vor(3 * H), tb(H), hm(H);
    rc = epic_nc_read_fields(path, step, &g, vor.data(), tb.data(), hm.data(), t);
    if (rc) return epic_internal_fail(h, rc, g_err);
    if ((rc = epic_set_field(h, EPIC_FIELD_VORTG, vor.data())) || (rc = epic_set_field(h, EPIC_FIELD_TBUOYG, tb.data())) ||
        (rc = epic_set_field(h, EPIC_FIELD_HUMG, hm.data())))
        return rc;
    return EPIC_OK;
}

int epic_write_fields_nc(epic_handle h, const char *path, double t)
{
    if (!h || !path) return EPIC_E_INVALID;
    epic_grid_t g;
    epic_phys_t ph;
    epic_flags_t fl;
    epic_get_config(h, &g, &ph, &fl);
    const int64_t H = (int64_t)(g.nz + 3) * (g.ny + 3) * (g.nx + 3);
    auto field = [&](int which, int nc, std::vector<double> &f) -> int {
        f.assign((size_t)(nc * H), 0.0);
        return epic_get_field(h, which, f.data());
    };
    std::vector<double> velog, vtend, vortg, tb, db, hm, vol, np, nsp;
    int rc;
    if ((rc = field(EPIC_FIELD_VELOG, 3, velog)) || (rc = field(EPIC_FIELD_VTEND, 3, vtend)) ||
        (rc = field(EPIC_FIELD_VORTG, 3, vortg)) || (rc = field(EPIC_FIELD_TBUOYG, 1, tb)) ||
        (rc = field(EPIC_FIELD_DBUOYG, 1, db)) || (rc = field(EPIC_FIELD_HUMG, 1, hm)) ||
        (rc = field(EPIC_FIELD_VOLG, 1, vol)) || (rc = field(EPIC_FIELD_NPARG, 1, np)) ||
        (rc = field(EPIC_FIELD_NSPARG, 1, nsp)))
        return rc;
    rc = epic_nc_write_fields(path, &g, &ph, &fl, t, velog.data(), vtend.data(), vortg.data(), tb.data(), db.data(),
                              hm.data(), vol.data(), np.data(), nsp.data());
    return rc ? epic_internal_fail(h, rc, g_err) : EPIC_OK;
}

}  // extern "C"
