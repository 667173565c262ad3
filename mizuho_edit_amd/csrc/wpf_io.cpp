// wpf_io.cpp -- native input pipeline: the .cfg + well file + wPFAWHFG field
// file that loadConfig / loadField / saveField read and write
// (wPFAW3D.f90:264-557, 2096-2141), for callers that drive the C ABI from
// compiled code (the Fortran drop-in driver).
//
// wpf_config_load parses a configuration into a per-thread staging area;
// wpf_config_wells and wpf_config_fields then copy it into caller-owned
// arrays sized from the returned parameters.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wpf.h"

namespace wpf {
void set_error(const char* fmt, ...);
}

namespace {

struct Staged {
    wpf_params p{};
    double time_start = 0.0, time_total = 0.0;
    int32_t n_res_out = 0, n_dat_out = 0;  // gid_output, scr_output
    std::vector<wpf_well> wells;
    std::string init_file;
    double init_k = 0.0, init_phase = -1.0, init_conc = 0.0;
    bool loaded = false;
};
thread_local Staged g_cfg;

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::string dir_of(const std::string& path)
{
    size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

std::string resolve(const std::string& base, const std::string& f)
{
    if (f.empty() || f[0] == '/') return f;
    return base + f;
}

bool read_field_body(FILE* fp, size_t n, std::vector<double>& out)
{
    out.resize(n);
    return fread(out.data(), sizeof(double), n, fp) == n;
}

}  // namespace

extern "C" {

// loadConfig, wPFAW3D.f90:264-505: key = value lines, '#' comments, the 24
// keys of 1:305-385, derived dx and alpha (1:390-399), the well file
// (1:401-432).  Returns WPF_ERR_IO where loadConfig calls exit(1).
int wpf_config_load(const char* path, wpf_params* out, int32_t* nwell, double* time_start,
                    double* time_total)
{
    if (!path || !out || !nwell) {
        wpf::set_error("wpf_config_load: null argument");
        return WPF_ERR_ARG;
    }
    Staged S;
    S.p.dtime = 0.1000e-06;  // default of 1:286
    FILE* fp = fopen(path, "r");
    if (!fp) {
        wpf::set_error("can not open config file: %s", path);
        return WPF_ERR_IO;
    }
    double dxw = 0.0;
    std::string well_file;
    char line[2048];
    int rc = WPF_OK;
    while (fgets(line, sizeof(line), fp)) {
        std::string buf(line);
        if (!buf.empty() && buf[0] == '#') continue;
        size_t h = buf.find('#');
        if (h != std::string::npos) buf = buf.substr(0, h);
        if (trim(buf).empty()) continue;
        size_t e = buf.find('=');
        if (e != std::string::npos) buf[e] = ' ';
        for (char& ch : buf)
            if (ch == ',') ch = ' ';
        std::istringstream is(buf);
        std::string key;
        is >> key;
        bool ok = true;
        auto rd = [&](double& v) { ok = ok && static_cast<bool>(is >> v); };
        auto ri = [&](int32_t& v) { ok = ok && static_cast<bool>(is >> v); };
        double dummy;
        if (key == "grid") { ri(S.p.nx); ri(S.p.ny); ri(S.p.nz); }
        else if (key == "dt") rd(S.p.dtime);
        else if (key == "tfac") rd(S.p.tfac);
        else if (key == "time_total") rd(S.time_total);
        else if (key == "gid_output") ri(S.n_res_out);
        else if (key == "scr_output") ri(S.n_dat_out);
        else if (key == "w") rd(S.p.epsilon);
        else if (key == "dxw") rd(dxw);
        else if (key == "lambda") rd(S.p.lambda);
        else if (key == "poro") rd(S.p.porosity);
        else if (key == "rhos") rd(S.p.rhos);
        else if (key == "rhoa") rd(S.p.rhoa);
        else if (key == "visco") rd(S.p.viscosity);
        else if (key == "cwtp") rd(S.p.cwtp);
        else if (key == "D") rd(S.p.diffusion);
        else if (key == "b100") rd(S.p.b100);
        else if (key == "ral") rd(S.p.ral);
        else if (key == "rat") rd(S.p.rat);
        else if (key == "kwh") rd(S.p.permeab_hole);
        else if (key == "pini") rd(S.p.pini);
        else if (key == "ref") { ri(S.p.ix_ref); ri(S.p.iy_ref); ri(S.p.iz_ref); }
        else if (key == "well") { ok = static_cast<bool>(is >> well_file); }
        else if (key == "init") {
            // three numbers, else a field-file name (1:372-377)
            std::string rest;
            std::getline(is, rest);
            std::istringstream i2(rest);
            double a, b, c;
            if (i2 >> a >> b >> c) {
                S.init_k = a; S.init_phase = b; S.init_conc = c;
            } else {
                std::istringstream i3(rest);
                ok = static_cast<bool>(i3 >> S.init_file);
            }
        }
        else if (key == "last") { std::string s; ok = static_cast<bool>(is >> s); }
        else {
            wpf::set_error("unknown key in config file %s: %s", path, key.c_str());
            rc = WPF_ERR_IO;
            break;
        }
        (void)dummy;
        if (!ok) {
            wpf::set_error("broken line in config file %s: %s", path, trim(line).c_str());
            rc = WPF_ERR_IO;
            break;
        }
    }
    fclose(fp);
    if (rc) return rc;
    // derived quantities (1:390-399)
    S.p.dx = dxw * S.p.epsilon;
    const double b = S.p.rhoa / S.p.rhos;
    const double kc = S.p.cwtp * S.p.b100 / (1.0 - S.p.porosity);
    S.p.alpha = 0.5 / b / kc;
    // well file (1:401-432): line 1 ignored, line 2 = count, then rows
    const std::string base = dir_of(path);
    const std::string wpath = resolve(base, well_file);
    FILE* wf = fopen(wpath.c_str(), "r");
    if (!wf) {
        wpf::set_error("can not open well file: %s", wpath.c_str());
        return WPF_ERR_IO;
    }
    int nw = 0;
    if (!fgets(line, sizeof(line), wf) || !fgets(line, sizeof(line), wf) ||
        sscanf(line, "%d", &nw) != 1 || nw < 0) {
        fclose(wf);
        wpf::set_error("invalid data in well file: %s", wpath.c_str());
        return WPF_ERR_IO;
    }
    S.wells.resize(nw);
    for (int i = 0; i < nw; i++) {
        int ix, iy, iz, wpq;
        double wval, ph, cc;
        if (!fgets(line, sizeof(line), wf)) { rc = WPF_ERR_IO; break; }
        for (char* q = line; *q; q++)
            if (*q == ',') *q = ' ';
        if (sscanf(line, "%d %d %d %d %lf %lf %lf", &ix, &iy, &iz, &wpq, &wval, &ph, &cc) != 7 ||
            (wpq != 1 && wpq != -1)) {
            rc = WPF_ERR_IO;
            break;
        }
        wpf_well w{};
        w.ix_g = ix; w.iy_g = iy; w.iz_g = iz;
        w.pressure = (wpq == 1) ? wval : 0.0;
        w.injection = (wpq == -1) ? wval : 0.0;
        w.phase = ph;
        w.concent = cc;
        S.wells[i] = w;
    }
    fclose(wf);
    if (rc) {
        wpf::set_error("invalid data in well file: %s", wpath.c_str());
        return rc;
    }
    if (!S.init_file.empty()) S.init_file = resolve(base, S.init_file);
    S.loaded = true;
    g_cfg = S;
    *out = S.p;
    *nwell = (int32_t)S.wells.size();
    if (time_total) *time_total = S.time_total;
    if (time_start) {
        *time_start = 0.0;
        if (!S.init_file.empty()) {
            FILE* ff = fopen(S.init_file.c_str(), "rb");
            if (ff) {
                char hdr[32];
                if (fread(hdr, 1, 32, ff) == 32) memcpy(time_start, hdr + 16, 8);
                fclose(ff);
            }
        }
    }
    return WPF_OK;
}

int wpf_config_wells(wpf_well* wells)
{
    if (!g_cfg.loaded) {
        wpf::set_error("wpf_config_wells before wpf_config_load");
        return WPF_ERR_STATE;
    }
    if (!g_cfg.wells.empty()) memcpy(wells, g_cfg.wells.data(), sizeof(wpf_well) * g_cfg.wells.size());
    return WPF_OK;
}

// The state loadConfig leaves behind (1:440-496): lw with halos, permeab_rock,
// phase and concent initialised, wells' phase/concent applied, the field file
// read over them when given (loadField, 1:509-557), halos clamped.
int wpf_config_fields(int32_t* lw, double* krock, double* phase, double* concent)
{
    if (!g_cfg.loaded) {
        wpf::set_error("wpf_config_fields before wpf_config_load");
        return WPF_ERR_STATE;
    }
    const wpf_params& p = g_cfg.p;
    const int64_t sx = p.nx + 2, sy = p.ny + 2, sz = p.nz + 2;
    const int64_t nh = sx * sy * sz;
    auto H = [&](int64_t i, int64_t j, int64_t k) { return i + sx * (j + sy * k); };
    for (int64_t q = 0; q < nh; q++) {
        lw[q] = 0;
        krock[q] = g_cfg.init_k;
        phase[q] = g_cfg.init_phase;
        concent[q] = g_cfg.init_conc;
    }
    for (size_t iw = 0; iw < g_cfg.wells.size(); iw++) {
        const wpf_well& w = g_cfg.wells[iw];
        if (w.ix_g >= 0 && w.ix_g <= p.nx + 1 && w.iy_g >= 0 && w.iy_g <= p.ny + 1 && w.iz_g >= 0 &&
            w.iz_g <= p.nz + 1)
            lw[H(w.ix_g, w.iy_g, w.iz_g)] = (int32_t)(iw + 1);
    }
    for (const wpf_well& w : g_cfg.wells) {
        if (w.ix_g >= 1 && w.ix_g <= p.nx && w.iy_g >= 1 && w.iy_g <= p.ny && w.iz_g >= 1 &&
            w.iz_g <= p.nz) {
            phase[H(w.ix_g, w.iy_g, w.iz_g)] = w.phase;
            concent[H(w.ix_g, w.iy_g, w.iz_g)] = w.concent;
        }
    }
    if (!g_cfg.init_file.empty()) {
        FILE* fp = fopen(g_cfg.init_file.c_str(), "rb");
        if (!fp) {
            wpf::set_error("can not open field file: %s", g_cfg.init_file.c_str());
            return WPF_ERR_IO;
        }
        char hdr[32];
        int16_t dims[4];
        if (fread(hdr, 1, 32, fp) != 32 || memcmp(hdr, "wPFAWHFG", 8) != 0) {
            fclose(fp);
            wpf::set_error("invalid magic in field file: %s", g_cfg.init_file.c_str());
            return WPF_ERR_IO;
        }
        memcpy(dims, hdr + 24, 8);
        if (dims[0] != p.nx || dims[1] != p.ny || dims[2] != p.nz) {
            fclose(fp);
            wpf::set_error("invalid grid in field file: %s", g_cfg.init_file.c_str());
            return WPF_ERR_IO;
        }
        const size_t n = (size_t)p.nx * p.ny * p.nz;
        std::vector<double> body;
        double* dst[3] = {krock, phase, concent};
        for (int f = 0; f < 3; f++) {
            if (!read_field_body(fp, n, body)) {
                fclose(fp);
                wpf::set_error("short field file: %s", g_cfg.init_file.c_str());
                return WPF_ERR_IO;
            }
            size_t q = 0;
            for (int k = 1; k <= p.nz; k++)
                for (int j = 1; j <= p.ny; j++)
                    for (int i = 1; i <= p.nx; i++) dst[f][H(i, j, k)] = body[q++];
        }
        fclose(fp);
    }
    // updateHalo on one rank: x, then y, then z (1:1651-1656)
    double* fl[3] = {krock, phase, concent};
    for (double* a : fl) {
        for (int64_t k = 0; k < sz; k++)
            for (int64_t j = 0; j < sy; j++) {
                a[H(0, j, k)] = a[H(1, j, k)];
                a[H(sx - 1, j, k)] = a[H(sx - 2, j, k)];
            }
        for (int64_t k = 0; k < sz; k++)
            for (int64_t i = 0; i < sx; i++) {
                a[H(i, 0, k)] = a[H(i, 1, k)];
                a[H(i, sy - 1, k)] = a[H(i, sy - 2, k)];
            }
        for (int64_t j = 0; j < sy; j++)
            for (int64_t i = 0; i < sx; i++) {
                a[H(i, j, 0)] = a[H(i, j, 1)];
                a[H(i, j, sz - 1)] = a[H(i, j, sz - 2)];
            }
    }
    return WPF_OK;
}

// saveField, wPFAW3D.f90:2096-2141: header then the interior of permeab_rock,
// phase and concent (haloed host arrays in, global arrays out).
int wpf_field_write(const char* path, double time, const wpf_params* p, const double* krock,
                    const double* phase, const double* concent)
{
    if (!path || !p || !krock || !phase || !concent) {
        wpf::set_error("wpf_field_write: null argument");
        return WPF_ERR_ARG;
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) {
        wpf::set_error("can not create field file: %s", path);
        return WPF_ERR_IO;
    }
    char hdr[32];
    memcpy(hdr, "wPFAWHFG", 8);
    memcpy(hdr + 8, "20250304", 8);
    memcpy(hdr + 16, &time, 8);
    int16_t dims[4] = {(int16_t)p->nx, (int16_t)p->ny, (int16_t)p->nz, 0};
    memcpy(hdr + 24, dims, 8);
    bool ok = fwrite(hdr, 1, 32, fp) == 32;
    const int64_t sx = p->nx + 2, sy = p->ny + 2;
    const double* src[3] = {krock, phase, concent};
    std::vector<double> row(p->nx);
    for (int f = 0; f < 3 && ok; f++)
        for (int k = 1; k <= p->nz && ok; k++)
            for (int j = 1; j <= p->ny && ok; j++) {
                const double* r = src[f] + 1 + sx * (j + sy * k);
                ok = fwrite(r, sizeof(double), p->nx, fp) == (size_t)p->nx;
            }
    fclose(fp);
    if (!ok) {
        wpf::set_error("write failed: %s", path);
        return WPF_ERR_IO;
    }
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// GiD post-processing result file (saveResult, wPFAW3D.f90:1842-2091)
// ---------------------------------------------------------------------------
namespace {

struct GidOut {
    FILE* fp;
    bool ok = true;
    void raw(const void* d, size_t n) { ok = ok && fwrite(d, 1, n, fp) == n; }
    void i32(int32_t v) { raw(&v, 4); }
    // write_string (1:2045-2051): length + 1, then the text and a NUL
    void str(const std::string& t)
    {
        i32((int32_t)t.size() + 1);
        raw(t.c_str(), t.size() + 1);
    }
};

// Fortran's ES12.5 edit of the time: d.ddddd, then E+xx, or +xxx (no E) for
// a three-digit exponent, right-justified in 12 characters (C++ linkage:
// it returns a std::string)
extern "C++" std::string es12_5(double v)
{
    char b[64];
    snprintf(b, sizeof b, "%.5E", v);
    std::string t(b);
    const size_t e = t.find('E');
    if (e != std::string::npos && t.size() - e - 2 > 2) t.erase(e, 1);
    if (t.size() < 12) t.insert(0, 12 - t.size(), ' ');
    return t;
}

}  // namespace

int wpf_config_output(int32_t* n_res_out, int32_t* n_dat_out)
{
    if (!g_cfg.loaded) {
        wpf::set_error("wpf_config_output before wpf_config_load");
        return WPF_ERR_STATE;
    }
    if (n_res_out) *n_res_out = g_cfg.n_res_out;
    if (n_dat_out) *n_dat_out = g_cfg.n_dat_out;
    return WPF_OK;
}

int wpf_result_write(const char* path, double time, const wpf_params* p, int32_t first, const double* pressure,
                     const double* phase, const double* concent, const double* permeab0)
{
    if (!path || !p || !pressure || !phase || !concent || !permeab0) {
        wpf::set_error("wpf_result_write: null argument");
        return WPF_ERR_ARG;
    }
    FILE* fp = fopen(path, first ? "wb" : "ab");
    if (!fp) {
        wpf::set_error("can not create result file: %s", path);
        return WPF_ERR_IO;
    }
    GidOut o{fp};
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    // serial indices of mesh points (with the extra point per axis) and of
    // elements, 1-based, x fastest (mindex3D, eindex3D: 1:2071-2089)
    auto mindex = [&](int i, int j, int k) {
        return (int32_t)((i - 1) + (int64_t)(j - 1) * (nx + 1) + (int64_t)(k - 1) * (nx + 1) * (ny + 1) + 1);
    };
    auto eindex = [&](int i, int j, int k) {
        return (int32_t)((i - 1) + (int64_t)(j - 1) * nx + (int64_t)(k - 1) * nx * ny + 1);
    };
    if (first) {
        o.i32(0x091d);  // ByteOrderCheck
        o.str("GiDPostEx1.1");
        o.str("MESH \"quadmesh\" dimension 3 ElemType Hexahedra Nnode 8");
        o.str("# color 0.7 0.7 0.4");
        o.str("Unit \"m\"");
        o.str("Coordinates -1 Indexed");
        struct Coord {
            int32_t index;
            float x, y, z;
        };
        std::vector<Coord> cr(nx + 1);
        for (int k = 1; k <= nz + 1; k++)
            for (int j = 1; j <= ny + 1; j++) {
                for (int i = 1; i <= nx + 1; i++)
                    cr[i - 1] = Coord{mindex(i, j, k), (float)((i - 1) * p->dx), (float)((j - 1) * p->dx),
                                      (float)((k - 1) * p->dx)};
                o.raw(cr.data(), sizeof(Coord) * cr.size());
            }
        o.str("End Coordinates");
        o.str("Elements -1 Indexed");
        struct Elem {
            int32_t index, nid[8], conn;
        };
        std::vector<Elem> er(nx);
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                for (int i = 1; i <= nx; i++)
                    er[i - 1] = Elem{eindex(i, j, k),
                                     {mindex(i, j, k), mindex(i + 1, j, k), mindex(i + 1, j + 1, k),
                                      mindex(i, j + 1, k), mindex(i, j, k + 1), mindex(i + 1, j, k + 1),
                                      mindex(i + 1, j + 1, k + 1), mindex(i, j + 1, k + 1)},
                                     1};
                o.raw(er.data(), sizeof(Elem) * er.size());
            }
        o.str("End Elements");
        o.str("GaussPoints \"Cell_Center\" ElemType Hexahedra");
        o.str("Number Of Gauss Points: 1");
        o.str("Natural Coordinates: Given");
        const float zeros[3] = {0.0f, 0.0f, 0.0f};
        o.raw(zeros, sizeof zeros);
        o.str("End GaussPoints");
    }
    // the four results of each call: haloed Fortran-order host arrays
    struct Value {
        int32_t index;
        float value;
    };
    const char* names[4] = {"Pressure", "Phase-field", "Concentration", "Permeability"};
    const char* units[4] = {"Pa", "-", "-", "-"};
    const double* src[4] = {pressure, phase, concent, permeab0};
    const int64_t sx = nx + 2, sy = ny + 2;
    std::vector<Value> vr(nx);
    for (int f = 0; f < 4; f++) {
        o.str(std::string("Result \"") + names[f] + "\" \"Analysis\" " + es12_5(time) +
              " Scalar OnGaussPoints \"Cell_Center\"");
        o.str(std::string("Unit \"") + units[f] + "\"");
        o.str("Values -1 Indexed");
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                for (int i = 1; i <= nx; i++)
                    vr[i - 1] = Value{eindex(i, j, k), (float)src[f][i + sx * (j + sy * k)]};
                o.raw(vr.data(), sizeof(Value) * vr.size());
            }
        o.i32(-1);
        o.str("End Values");
    }
    const bool ok = o.ok;
    if (fclose(fp) != 0 || !ok) {
        wpf::set_error("write failed: %s", path);
        return WPF_ERR_IO;
    }
    return WPF_OK;
}

}  // extern "C"
