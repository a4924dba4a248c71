// ncfile.cpp -- ps3d's field and field_stats files in NetCDF classic format,
// written and read without libnetcdf: the 64-bit offset variant ("CDF-2") where
// every variable's record fits its 32-bit size field, else the 64-bit data
// variant ("CDF-5", the format ps3d's 1024^3 records need).
//
// Field file (src/fields/field_netcdf.f90:65-272, 507-575): dimensions t (unlimited), z
// (nz+1), y, x; coordinates x, y, z; record variables t and x/y/z_velocity,
// x/y/z_vorticity and, when requested, pressure and horizontal_divergence, each in
// C order (t, z, y, x) -- ps3d's transpose-on-write of its (z, y, x) arrays
// (src/netcdf/netcdf_writer.f90:250-272).  Global attributes: file_type,
// ncells, extent, origin (netcdf_writer.f90:355-369, 434-443).  ps3d writes
// NetCDF-4 with a `physical_quantities` group; the classic format has no groups,
// so those quantities are not written.
//
// field_stats file (src/fields/field_diagnostics_netcdf.f90:124-252): dimension t
// (unlimited), the record variable t and ps3d's 40 statistics (36-75, 453-798)
// in ps3d's order, then the extensions helicity, dt and prefactor.
//
// Format (all integers big-endian): magic "CDF" + version (2 or 5), numrecs,
// dimension list, global attribute list, variable list (name, dimids, attributes,
// type, vsize, 64-bit begin), non-record data, then records holding every record
// variable in definition order.  Counts, lengths, dimension ids and vsize are
// 4-byte in CDF-2 and 8-byte in CDF-5 ("NON_NEG" of the format specification).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ps3d_gpu.h"
#include "../../include/ps3d_nc.h"

namespace {

thread_local std::string nc_err;
int nfail(const std::string &m) {
    nc_err = m;
    return PS3D_E_ARG;
}

enum { NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

struct Att {
    std::string name;
    int type;
    std::string text;             // NC_CHAR
    std::vector<int32_t> ints;    // NC_INT
    std::vector<double> dbls;     // NC_DOUBLE
};
struct Var {
    std::string name;
    std::vector<int> dims;
    std::vector<Att> atts;
    int type = NC_DOUBLE;
    uint64_t vsize = 0, begin = 0;
    bool record = false;
};
struct Header {
    int ver = 2;                                          // 2: CDF-2, 5: CDF-5
    std::vector<std::pair<std::string, uint64_t>> dims;   // length 0 = unlimited
    std::vector<Att> gatts;
    std::vector<Var> vars;
    uint64_t numrecs = 0;
    uint64_t recsize = 0;
    uint64_t data_start = 0;
};

// ---- big-endian buffer ----
struct Out {
    std::vector<unsigned char> b;
    int ver = 2;
    void u32(uint32_t v) {
        for (int i = 3; i >= 0; --i) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void u64(uint64_t v) {
        for (int i = 7; i >= 0; --i) b.push_back((unsigned char)(v >> (8 * i)));
    }
    void f64(double d) {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        u64(v);
    }
    void nn(uint64_t v) {   // NON_NEG
        if (ver == 5) u64(v); else u32((uint32_t)v);
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        nn(s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct In {
    const unsigned char *p;
    size_t n, o = 0;
    bool ok = true;
    int ver = 2;
    uint32_t u32() {
        if (o + 4 > n) { ok = false; return 0; }
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) v = (v << 8) | p[o++];
        return v;
    }
    uint64_t u64() {
        if (o + 8 > n) { ok = false; return 0; }
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v = (v << 8) | p[o++];
        return v;
    }
    double f64() {
        uint64_t v = u64();
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
    uint64_t nn() { return ver == 5 ? u64() : u32(); }
    void skip_pad() { o = (o + 3) & ~(size_t)3; }
    std::string name() {
        const uint64_t len = nn();
        if (!ok || o + len > n) { ok = false; return {}; }
        std::string s((const char *)p + o, len);
        o += len;
        skip_pad();
        return s;
    }
};

void put_atts(Out &w, const std::vector<Att> &atts) {
    if (atts.empty()) {   // ABSENT
        w.u32(0);
        w.nn(0);
        return;
    }
    w.u32(TAG_ATT);
    w.nn(atts.size());
    for (const Att &a : atts) {
        w.name(a.name);
        w.u32((uint32_t)a.type);
        if (a.type == NC_CHAR) {
            w.nn(a.text.size());
            w.b.insert(w.b.end(), a.text.begin(), a.text.end());
            w.pad();
        } else if (a.type == NC_INT) {
            w.nn(a.ints.size());
            for (int32_t v : a.ints) w.u32((uint32_t)v);
        } else {
            w.nn(a.dbls.size());
            for (double v : a.dbls) w.f64(v);
        }
    }
}

std::vector<Att> get_atts(In &r) {
    std::vector<Att> out;
    const uint32_t tag = r.u32();
    const uint64_t n = r.nn();
    if (tag == 0) return out;
    if (tag != TAG_ATT) { r.ok = false; return out; }
    for (uint64_t i = 0; i < n && r.ok; ++i) {
        Att a;
        a.name = r.name();
        a.type = (int)r.u32();
        const uint64_t m = r.nn();
        if (a.type == NC_CHAR) {
            if (r.o + m > r.n) { r.ok = false; break; }
            a.text.assign((const char *)r.p + r.o, m);
            r.o += m;
            r.skip_pad();
        } else if (a.type == NC_INT) {
            for (uint64_t k = 0; k < m && r.ok; ++k) a.ints.push_back((int32_t)r.u32());
        } else if (a.type == NC_DOUBLE) {
            for (uint64_t k = 0; k < m && r.ok; ++k) a.dbls.push_back(r.f64());
        } else {
            r.ok = false;
        }
        out.push_back(a);
    }
    return out;
}

uint64_t var_bytes(const Header &h, const Var &v) {
    uint64_t n = 8;   // NC_DOUBLE only
    for (int d : v.dims)
        if (h.dims[d].second != 0) n *= h.dims[d].second;
    return n;
}

std::vector<unsigned char> encode(Header &h) {
    // two passes: the header size does not depend on the begin values
    for (int pass = 0; pass < 2; ++pass) {
        Out w;
        w.ver = h.ver;
        w.b = {'C', 'D', 'F', (unsigned char)h.ver};
        w.nn(h.numrecs);
        w.u32(TAG_DIM);
        w.nn(h.dims.size());
        for (auto &d : h.dims) {
            w.name(d.first);
            w.nn(d.second);
        }
        put_atts(w, h.gatts);
        w.u32(TAG_VAR);
        w.nn(h.vars.size());
        for (Var &v : h.vars) {
            w.name(v.name);
            w.nn(v.dims.size());
            for (int d : v.dims) w.nn((uint64_t)d);
            put_atts(w, v.atts);
            w.u32((uint32_t)v.type);
            // CDF-2 keeps 2^32 - 1 for a size its 32-bit field cannot hold
            w.nn(h.ver == 2 && v.vsize > 0xffffffffull ? 0xffffffffull : v.vsize);
            w.u64(v.begin);
        }
        if (pass == 1) return w.b;
        uint64_t off = w.b.size();
        h.data_start = off;
        for (Var &v : h.vars)
            if (!v.record) {
                v.vsize = var_bytes(h, v);
                v.begin = off;
                off += v.vsize;
            }
        h.recsize = 0;
        for (Var &v : h.vars)
            if (v.record) {
                v.vsize = var_bytes(h, v);
                v.begin = off + h.recsize;
                h.recsize += v.vsize;
            }
    }
    return {};
}

int decode(const std::vector<unsigned char> &buf, Header &h) {
    In r{buf.data(), buf.size()};
    if (buf.size() < 8 || std::memcmp(buf.data(), "CDF", 3) != 0 || (buf[3] != 2 && buf[3] != 5))
        return nfail("not a NetCDF 64-bit offset (CDF-2) or 64-bit data (CDF-5) file");
    h.ver = r.ver = buf[3];
    r.o = 4;
    h.numrecs = r.nn();
    uint32_t tag = r.u32();
    uint64_t n = r.nn();
    if (tag == TAG_DIM)
        for (uint64_t i = 0; i < n && r.ok; ++i) {
            std::string nm = r.name();
            h.dims.push_back({nm, r.nn()});
        }
    h.gatts = get_atts(r);
    tag = r.u32();
    n = r.nn();
    if (tag == TAG_VAR)
        for (uint64_t i = 0; i < n && r.ok; ++i) {
            Var v;
            v.name = r.name();
            const uint64_t nd = r.nn();
            for (uint64_t k = 0; k < nd && r.ok; ++k) v.dims.push_back((int)r.nn());
            v.atts = get_atts(r);
            v.type = (int)r.u32();
            v.vsize = r.nn();
            v.begin = r.u64();
            for (int d : v.dims)
                if (d < 0 || d >= (int)h.dims.size()) r.ok = false;
            v.record = r.ok && !v.dims.empty() && h.dims[v.dims[0]].second == 0;
            if (r.ok) v.vsize = var_bytes(h, v);
            h.vars.push_back(v);
        }
    if (!r.ok) return nfail("malformed NetCDF header");
    h.recsize = 0;
    for (Var &v : h.vars)
        if (v.record) h.recsize += v.vsize;
    return 0;
}

int read_header(const char *path, Header &h) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return nfail(std::string("cannot open ") + path);
    std::vector<unsigned char> buf(1 << 16);
    const size_t got = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    buf.resize(got);
    return decode(buf, h);
}

Att catt(const char *n, const std::string &s) { return Att{n, NC_CHAR, s, {}, {}}; }
Att iatt(const char *n, std::vector<int32_t> v) { return Att{n, NC_INT, {}, std::move(v), {}}; }
Att datt(const char *n, std::vector<double> v) { return Att{n, NC_DOUBLE, {}, {}, std::move(v)}; }

Var rec_var(const char *name, std::vector<int> dims, const char *long_name, const char *unit) {
    Var v;
    v.name = name;
    v.dims = std::move(dims);
    v.record = true;
    v.atts = {catt("units", unit), catt("long_name", long_name)};
    return v;
}

int write_file(const char *path, Header &h) {
    std::vector<unsigned char> hdr = encode(h);
    FILE *f = std::fopen(path, "wb");
    if (!f) return nfail(std::string("cannot create ") + path);
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    std::fclose(f);
    return ok ? 0 : nfail("write failed");
}

void be_doubles(const double *src, size_t n, std::vector<unsigned char> &out) {
    out.resize(n * 8);
    for (size_t i = 0; i < n; ++i) {
        uint64_t v;
        std::memcpy(&v, &src[i], 8);
        for (int k = 0; k < 8; ++k) out[i * 8 + k] = (unsigned char)(v >> (56 - 8 * k));
    }
}
double from_be(const unsigned char *p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) v = (v << 8) | p[k];
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

// numrecs + 1 in the file's header (4 bytes in CDF-2, 8 in CDF-5, after the magic)
bool bump_numrecs(FILE *f, const Header &h) {
    const uint64_t nr = h.numrecs + 1;
    const int nb = h.ver == 5 ? 8 : 4;
    unsigned char b[8];
    for (int i = 0; i < nb; ++i) b[i] = (unsigned char)(nr >> (8 * (nb - 1 - i)));
    return fseeko(f, 4, SEEK_SET) == 0 && std::fwrite(b, 1, nb, f) == (size_t)nb;
}

// append one record: values[k] = data of the k-th record variable (npts doubles each,
// already in file order)
int append_record(const char *path, const std::vector<std::vector<double>> &values) {
    Header h;
    int st = read_header(path, h);
    if (st) return st;
    std::vector<const Var *> rv;
    for (const Var &v : h.vars)
        if (v.record) rv.push_back(&v);
    if (rv.size() != values.size()) return nfail("record layout mismatch");
    FILE *f = std::fopen(path, "r+b");
    if (!f) return nfail(std::string("cannot open ") + path);
    bool ok = true;
    std::vector<unsigned char> buf;
    for (size_t k = 0; k < rv.size() && ok; ++k) {
        if (values[k].size() * 8 != rv[k]->vsize) {
            std::fclose(f);
            return nfail("record variable size mismatch: " + rv[k]->name);
        }
        const uint64_t off = rv[k]->begin + (uint64_t)h.numrecs * h.recsize;
        be_doubles(values[k].data(), values[k].size(), buf);
        ok = fseeko(f, (off_t)off, SEEK_SET) == 0 && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    if (ok) ok = bump_numrecs(f, h);
    std::fclose(f);
    return ok ? 0 : nfail("write failed");
}

const char *FIELD_NAMES[8] = {"x_velocity", "y_velocity", "z_velocity", "x_vorticity",
                              "y_vorticity", "z_vorticity", "pressure", "horizontal_divergence"};
const char *FIELD_LONG[8] = {"x velocity component", "y velocity component", "z velocity component",
                             "x vorticity component", "y vorticity component", "z vorticity component",
                             "pressure", "horizontal divergence"};
const char *FIELD_UNIT[8] = {"m/s", "m/s", "m/s", "1/s", "1/s", "1/s", "m^2/s^2", "1/s"};

// field_stats variables: t, ps3d's statistics in the order NC_KE .. NC_ROMAX with
// its names, long names and units (field_diagnostics_netcdf.f90:36-75, 453-798),
// then the extensions (a helicity series, the step and the prefactor)
struct StatInfo {
    const char *name, *long_name, *unit;
};
const StatInfo STATS[PS3D_NC_STATS_ROW] = {
    {"t", "time ", "seconds since 1970-01-01"},
    {"ke", "domain-averaged kinetic energy", "m^2/s^2"},
    {"en", "domain-averaged enstrophy", "1/s^2"},
    {"vortmax", "maximum vorticity magnitude", "1/s"},
    {"vortrms", "root-mean square vorticity", "1/s"},
    {"vorch", "characteristic vorticity", "1/s"},
    {"x_vormean", "mean x-vorticity", "1/s"},
    {"y_vormean", "mean y-vorticity", "1/s"},
    {"z_vormean", "mean z-vorticity", "1/s"},
    {"kexy", "domain-averaged horizontal kinetic energy", "m^2/s^2"},
    {"kez", "domain-averaged vertical kinetic energy", "m^2/s^2"},
    {"enxy", "domain-averaged horizontal enstrophy", "1/s^2"},
    {"enz", "domain-averaged vertical enstrophy", "1/s^2"},
    {"x_vormin", "min x-vorticity", "1/s"},
    {"y_vormin", "min y-vorticity", "1/s"},
    {"z_vormin", "min z-vorticity", "1/s"},
    {"x_vormax", "max x-vorticity", "1/s"},
    {"y_vormax", "max y-vorticity", "1/s"},
    {"z_vormax", "max z-vorticity", "1/s"},
    {"enxy_max", "max horizontal enstrophy", "1/s"},
    {"gmax", "maximum gamma", "1/s"},
    {"bfmax", "maximum buoyancy frequency", "1/s"},
    {"umax", "maximum x-velocity", "m/s"},
    {"vmax", "maximum y-velocity", "m/s"},
    {"wmax", "maximum z-velocity", "m/s"},
    {"uoxmax", "upper surface maximum x-vorticity", "1/s"},
    {"loxmax", "lower surface maximum x-vorticity", "1/s"},
    {"uoymax", "upper surface maximum y-vorticity", "1/s"},
    {"loymax", "lower surface maximum y-vorticity", "1/s"},
    {"uozmax", "upper surface maximum z-vorticity", "1/s"},
    {"lozmax", "lower surface maximum z-vorticity", "1/s"},
    {"usuhmax", "upper surface maximum horizontal speed", "m/s"},
    {"usgmax", "upper surface maximum strain", "1/s"},
    {"lsgmax", "lower surface maximum strain", "1/s"},
    {"uszrms", "rms of upper surface z-vorticity", "1/s"},
    {"usdeltarms", "rms of upper surface horizontal divergence", "1/s"},
    {"rolling_mean_gmax", "rolling mean maximum gamma", "1/s"},
    {"rolling_mean_bfmax", "rolling mean maximum buoyancy frequency", "1/s"},
    {"ri_min", "minimum Richardson number", "1"},
    {"ro_min", "minimum Rossby number", "1"},
    {"ro_max", "maximum Rossby number", "1"},
    {"helicity", "domain-averaged helicity", "m/s^2"},
    {"dt", "time step", "s"},
    {"prefactor", "diffusion prefactor", "1"},
};

}  // namespace

extern "C" {

const char *ps3d_nc_last_error(void) { return nc_err.c_str(); }

int ps3d_nc_create_fields(const char *path, int nx, int ny, int nz, const double *lower, const double *extent,
                          int flags) {
    const bool with_derived = (flags & PS3D_NC_DERIVED) != 0;
    if (!path || !lower || !extent || nx < 1 || ny < 1 || nz < 1) return nfail("invalid argument");
    const uint64_t vol = 8ull * nx * ny * (nz + 1);
    Header h;
    // a variable's record must fit CDF-2's 32-bit size field (less 4 bytes of
    // alignment slack), else the file is CDF-5
    h.ver = (vol >= 0xfffffffcull || (flags & PS3D_NC_CDF5)) ? 5 : 2;
    h.dims = {{"t", 0}, {"z", (uint64_t)(nz + 1)}, {"y", (uint64_t)ny}, {"x", (uint64_t)nx}};
    h.gatts = {catt("file_type", "fields"), catt("Conventions", "CF-1.9"),
               iatt("ncells", {nx, ny, nz}), datt("extent", {extent[0], extent[1], extent[2]}),
               datt("origin", {lower[0], lower[1], lower[2]})};
    const char *axn[3] = {"x", "y", "z"};
    const int axd[3] = {3, 2, 1};
    const char *axu[3] = {"X", "Y", "Z"};
    for (int a = 0; a < 3; ++a) {
        Var v;
        v.name = axn[a];
        v.dims = {axd[a]};
        v.atts = {catt("units", "m"),
                  catt("long_name", a < 2 ? std::string(axn[a]) + "-coordinate in projected coordinate system"
                                          : std::string("height coordinate in projected coordinate system")),
                  catt("standard_name", a < 2 ? std::string("projection_") + axn[a] + "_coordinate"
                                              : std::string("height")),
                  catt("axis", axu[a])};
        h.vars.push_back(v);
    }
    Var tv = rec_var("t", {0}, "time ", "seconds since 1970-01-01");
    tv.atts.push_back(catt("standard_name", "time"));
    tv.atts.push_back(catt("axis", "T"));
    h.vars.push_back(tv);
    for (int k = 0; k < (with_derived ? 8 : 6); ++k) h.vars.push_back(rec_var(FIELD_NAMES[k], {0, 1, 2, 3}, FIELD_LONG[k], FIELD_UNIT[k]));
    int st = write_file(path, h);
    if (st) return st;
    // coordinates x_i = lower + i dx (periodic, no end point), z_k = lower + k dz
    FILE *f = std::fopen(path, "r+b");
    if (!f) return nfail(std::string("cannot open ") + path);
    const int n[3] = {nx, ny, nz};
    bool ok = true;
    std::vector<unsigned char> buf;
    for (int a = 0; a < 3 && ok; ++a) {
        const int m = a < 2 ? n[a] : nz + 1;
        std::vector<double> c(m);
        const double d = extent[a] / (double)n[a];
        for (int i = 0; i < m; ++i) c[i] = lower[a] + d * (double)i;
        be_doubles(c.data(), c.size(), buf);
        ok = fseeko(f, (off_t)h.vars[a].begin, SEEK_SET) == 0 && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    std::fclose(f);
    return ok ? 0 : nfail("write failed");
}

int ps3d_nc_append_fields(const char *path, double t, int nx, int ny, int nz, const double *vel, const double *vor,
                          const double *pres, const double *delta) {
    if (!path || !vel || !vor) return nfail("invalid argument");
    Header h;
    int st = read_header(path, h);
    if (st) return st;
    std::vector<const Var *> rv;
    for (const Var &v : h.vars)
        if (v.record) rv.push_back(&v);
    const bool derived = rv.size() == 9;
    if (rv.size() != 7 && !derived) return nfail("record layout mismatch");
    if (derived && (!pres || !delta)) return nfail("this field file holds pressure and divergence: pass them");
    const size_t hv = (size_t)nx * ny * (nz + 1), plane = (size_t)nx * ny;
    if (rv[1]->vsize != hv * 8) return nfail("grid differs from the file's");
    const double *src[8] = {vel, vel + hv, vel + 2 * hv, vor, vor + hv, vor + 2 * hv, pres, delta};
    FILE *f = std::fopen(path, "r+b");
    if (!f) return nfail(std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    bool ok;
    {
        be_doubles(&t, 1, buf);
        ok = fseeko(f, (off_t)(rv[0]->begin + h.numrecs * h.recsize), SEEK_SET) == 0 &&
             std::fwrite(buf.data(), 1, 8, f) == 8;
    }
    // each variable in file order (t, z, y, x), one z level at a time
    std::vector<double> lev(plane);
    for (int k = 0; k < (derived ? 8 : 6) && ok; ++k) {
        const uint64_t base = rv[1 + k]->begin + h.numrecs * h.recsize;
        ok = fseeko(f, (off_t)base, SEEK_SET) == 0;
        for (int iz = 0; iz <= nz && ok; ++iz) {
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) lev[(size_t)iy * nx + ix] = src[k][((size_t)ix * ny + iy) * (nz + 1) + iz];
            be_doubles(lev.data(), plane, buf);
            ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
        }
    }
    if (ok) ok = bump_numrecs(f, h);
    std::fclose(f);
    return ok ? 0 : nfail("write failed");
}

int ps3d_nc_read_fields(const char *path, int step, double *t, int nx, int ny, int nz, double *vel, double *vor) {
    if (!path || !vor) return nfail("invalid argument");
    Header h;
    int st = read_header(path, h);
    if (st) return st;
    if (h.numrecs == 0) return nfail("no records");
    // field_netcdf.f90:286-299: step counts from 1; -1 or a step past the end reads the last
    if (step == 0 || step < -1) return nfail("step must be >= 1 or -1");
    const int rec = (step == -1 || step > (int)h.numrecs) ? (int)h.numrecs - 1 : step - 1;
    auto dim = [&](const char *n) -> int {
        for (auto &d : h.dims)
            if (d.first == n) return (int)d.second;
        return -1;
    };
    if (dim("x") != nx || dim("y") != ny || dim("z") != nz + 1) return nfail("grid differs from the file's");
    FILE *f = std::fopen(path, "rb");
    if (!f) return nfail(std::string("cannot open ") + path);
    const size_t hv = (size_t)nx * ny * (nz + 1);
    std::vector<unsigned char> buf;
    bool ok = true;
    auto read_var = [&](const char *name, double *dst, bool scalar) {
        for (const Var &v : h.vars) {
            if (v.name != name || !v.record) continue;
            const uint64_t off = v.begin + (uint64_t)rec * h.recsize;
            ok = ok && fseeko(f, (off_t)off, SEEK_SET) == 0;
            if (!ok) return false;
            if (scalar) {
                buf.resize(8);
                ok = std::fread(buf.data(), 1, 8, f) == 8;
                if (ok) *dst = from_be(buf.data());
                return ok;
            }
            // one z level (ny, nx) at a time
            buf.resize((size_t)nx * ny * 8);
            for (int iz = 0; iz <= nz && ok; ++iz) {
                ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
                for (int iy = 0; iy < ny && ok; ++iy)
                    for (int ix = 0; ix < nx; ++ix)
                        dst[((size_t)ix * ny + iy) * (nz + 1) + iz] = from_be(&buf[((size_t)iy * nx + ix) * 8]);
            }
            return ok;
        }
        return false;
    };
    double tt = 0.0;
    bool found = read_var("t", &tt, true);
    for (int c = 0; c < 3 && found; ++c) found = read_var(FIELD_NAMES[3 + c], vor + c * hv, false);
    for (int c = 0; c < 3 && found && vel; ++c) found = read_var(FIELD_NAMES[c], vel + c * hv, false);
    std::fclose(f);
    if (!ok) return nfail("read failed");
    if (!found) return nfail("field file lacks a vorticity or velocity component");
    if (t) *t = tt;
    return 0;
}

int ps3d_nc_records(const char *path, int *nrecords, double *t_last) {
    if (!path || !nrecords) return nfail("invalid argument");
    Header h;
    int st = read_header(path, h);
    if (st) return st;
    *nrecords = (int)h.numrecs;
    if (!t_last) return 0;
    *t_last = 0.0;
    for (const Var &v : h.vars) {
        if (v.name != "t" || !v.record || h.numrecs == 0) continue;
        FILE *f = std::fopen(path, "rb");
        if (!f) return nfail(std::string("cannot open ") + path);
        unsigned char b[8];
        const uint64_t off = v.begin + (uint64_t)(h.numrecs - 1) * h.recsize;
        const bool ok = fseeko(f, (off_t)off, SEEK_SET) == 0 && std::fread(b, 1, 8, f) == 8;
        std::fclose(f);
        if (!ok) return nfail("read failed");
        *t_last = from_be(b);
    }
    return 0;
}

int ps3d_nc_grid(const char *path, int *ncells, double *extent, double *origin) {
    if (!path || !ncells || !extent || !origin) return nfail("invalid argument");
    Header h;
    int st = read_header(path, h);
    if (st) return st;
    bool got[3] = {false, false, false};
    for (const Att &a : h.gatts) {
        if (a.name == "ncells" && a.type == NC_INT && a.ints.size() == 3) {
            for (int k = 0; k < 3; ++k) ncells[k] = a.ints[k];
            got[0] = true;
        } else if (a.name == "extent" && a.type == NC_DOUBLE && a.dbls.size() == 3) {
            for (int k = 0; k < 3; ++k) extent[k] = a.dbls[k];
            got[1] = true;
        } else if (a.name == "origin" && a.type == NC_DOUBLE && a.dbls.size() == 3) {
            for (int k = 0; k < 3; ++k) origin[k] = a.dbls[k];
            got[2] = true;
        }
    }
    if (!got[0] || !got[1] || !got[2]) return nfail("file lacks the ncells / extent / origin attributes");
    return 0;
}

int ps3d_nc_create_stats(const char *path, int nx, int ny, int nz, const double *lower, const double *extent) {
    if (!path || !lower || !extent || nx < 1 || ny < 1 || nz < 1) return nfail("invalid argument");
    Header h;
    h.dims = {{"t", 0}};
    h.gatts = {catt("file_type", "field_stats"), catt("Conventions", "CF-1.9"),
               iatt("ncells", {nx, ny, nz}), datt("extent", {extent[0], extent[1], extent[2]}),
               datt("origin", {lower[0], lower[1], lower[2]})};
    for (int k = 0; k < PS3D_NC_STATS_ROW; ++k) {
        h.vars.push_back(rec_var(STATS[k].name, {0}, STATS[k].long_name, STATS[k].unit));
        if (k == 0) {
            h.vars.back().atts.push_back(catt("standard_name", "time"));
            h.vars.back().atts.push_back(catt("axis", "T"));
        }
    }
    return write_file(path, h);
}

int ps3d_nc_stats_row(const double *diag, const double *stats, double *row) {
    if (!diag || !stats || !row) return nfail("invalid argument");
    row[0] = diag[PS3D_D_T];
    for (int k = 0; k < PS3D_NSTAT; ++k) row[1 + k] = stats[k];
    // the values adapt hands to set_netcdf_field_diagnostic (advance.f90:188-193,
    // 315-321, 366) come from the adapt record of the same time
    const int ad[][2] = {{PS3D_ST_OMAX, PS3D_D_VORTMAX},     {PS3D_ST_ORMS, PS3D_D_VORTRMS},
                         {PS3D_ST_OCHAR, PS3D_D_VORCH},      {PS3D_ST_OXMEAN, PS3D_D_VORMEAN_X},
                         {PS3D_ST_OYMEAN, PS3D_D_VORMEAN_Y}, {PS3D_ST_OZMEAN, PS3D_D_VORMEAN_Z},
                         {PS3D_ST_GMAX, PS3D_D_GGMAX},       {PS3D_ST_UMAX, PS3D_D_UMAX},
                         {PS3D_ST_VMAX, PS3D_D_VMAX},        {PS3D_ST_WMAX, PS3D_D_WMAX},
                         {PS3D_ST_USGMAX, PS3D_D_USGGMAX},   {PS3D_ST_LSGMAX, PS3D_D_LSGGMAX},
                         {PS3D_ST_RGMAX, PS3D_D_RGMAX}};
    for (const auto &m : ad) row[1 + m[0]] = diag[m[1]];
    row[1 + PS3D_NSTAT] = diag[PS3D_D_HEL];
    row[2 + PS3D_NSTAT] = diag[PS3D_D_DT];
    row[3 + PS3D_NSTAT] = diag[PS3D_D_PREFACTOR];
    return 0;
}

int ps3d_nc_append_stats(const char *path, const double *rows, int nrows) {
    if (!path || !rows || nrows < 0) return nfail("invalid argument");
    for (int r = 0; r < nrows; ++r) {
        std::vector<std::vector<double>> vals;
        for (int k = 0; k < PS3D_NC_STATS_ROW; ++k) vals.push_back({rows[(size_t)r * PS3D_NC_STATS_ROW + k]});
        int st = append_record(path, vals);
        if (st) return st;
    }
    return 0;
}

}  // extern "C"
