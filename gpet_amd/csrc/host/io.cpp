/*
 * io.cpp - host-only parsers and writers of the drop-in file formats.
 *
 * Each parser restates the reference's positional fgets/fscanf sequence
 * (cited per function) with error codes instead of exit().  Values are
 * parsed with strtof/strtod, which round the same way as fscanf("%f").
 */
#include "../../../include/gpet.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace {

bool read_lines(const char *path, std::vector<std::string> &lines)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string l;
    while (std::getline(f, l)) {
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
    }
    return true;
}

bool blank(const std::string &s)
{
    for (char c : s)
        if (!isspace((unsigned char)c)) return false;
    return true;
}

std::vector<std::string> tokens(const std::string &s)
{
    std::vector<std::string> t;
    std::istringstream is(s);
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

bool is_number(const std::string &s)
{
    if (s.empty()) return false;
    char *end = nullptr;
    strtod(s.c_str(), &end);
    return end && *end == '\0';
}

/* Line cursor that skips blank lines, as the reference's fscanf("...\n")
 * calls consume all whitespace before the next fgets. */
struct Cursor {
    std::vector<std::string> L;
    size_t i = 0;
    bool next_nonblank(std::string &out)
    {
        while (i < L.size() && blank(L[i])) ++i;
        if (i >= L.size()) return false;
        out = L[i++];
        return true;
    }
    /* k numeric tokens, possibly spanning lines */
    bool values(size_t k, std::vector<std::string> &out)
    {
        out.clear();
        while (out.size() < k) {
            std::string l;
            if (!next_nonblank(l)) return false;
            for (auto &t : tokens(l)) {
                if (out.size() < k) out.push_back(t);
            }
        }
        return true;
    }
};

float f32(const std::string &s) { return strtof(s.c_str(), nullptr); }
double f64(const std::string &s) { return strtod(s.c_str(), nullptr); }
long long i64(const std::string &s) { return (long long)strtod(s.c_str(), nullptr); }

struct OwnedTables {
    std::vector<float> mu[4];
    std::vector<float> surf[2];
};

/* Tables returned by gpet_tables_load own library memory, found again by the
 * address of their first array. */
std::mutex g_owned_mu;
std::map<const float *, OwnedTables *> g_owned;

/* One 1-D table file (rlamph/rcompt/rphote/rrayle, initialize.cu:370-662). */
gpet_status read_table1d(const std::string &path, int nmat, std::vector<float> &vals, float &e0, float &idl)
{
    std::vector<std::string> L;
    if (!read_lines(path.c_str(), L)) return GPET_ERR_IO;
    vals.assign((size_t)nmat * GPET_NTAB, 0.0f);
    size_t i = 2;
    std::vector<float> energies(GPET_NTAB);
    for (int m = 0; m < nmat; ++m) {
        while (i < L.size() && L[i].find("ndata") == std::string::npos) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        ++i;
        while (i < L.size() && blank(L[i])) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        auto h = tokens(L[i++]);
        if (h.empty() || i64(h[0]) != GPET_NTAB) return GPET_ERR_FORMAT;
        ++i; /* column label */
        for (int k = 0; k < GPET_NTAB; ++k) {
            while (i < L.size() && blank(L[i])) ++i;
            if (i >= L.size()) return GPET_ERR_FORMAT;
            auto t = tokens(L[i++]);
            if (t.size() < 2) return GPET_ERR_FORMAT;
            float e = f32(t[0]);
            if (m == 0) energies[k] = e;
            vals[(size_t)m * GPET_NTAB + k] = f32(t[1]);
        }
    }
    e0 = energies[0];
    idl = (float)(GPET_NTAB - 1) / (energies[GPET_NTAB - 1] - energies[0]);
    return GPET_OK;
}

/* One surface file (rcmpsf/rrayff, initialize.cu:483-566, 664-748). */
gpet_status read_surface(const std::string &path, int nmat, std::vector<float> &vals, float &idcp, float &ide)
{
    std::vector<std::string> L;
    if (!read_lines(path.c_str(), L)) return GPET_ERR_IO;
    vals.assign((size_t)nmat * GPET_NCP * GPET_NE, 0.0f);
    size_t i = 2;
    for (int m = 0; m < nmat; ++m) {
        while (i < L.size() && L[i].find("ndata") == std::string::npos) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        ++i;
        while (i < L.size() && blank(L[i])) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        auto h = tokens(L[i++]);
        if (h.empty()) return GPET_ERR_FORMAT;
        long long nd = i64(h[0]);
        ++i;
        for (long long k = 0; k < nd; ++k) {
            while (i < L.size() && blank(L[i])) ++i;
            ++i;
        }
        while (i < L.size() && L[i].find("nCP") == std::string::npos) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        ++i;
        while (i < L.size() && blank(L[i])) ++i;
        if (i >= L.size()) return GPET_ERR_FORMAT;
        auto g = tokens(L[i++]);
        if (g.size() < 8) return GPET_ERR_FORMAT;
        if (i64(g[0]) != GPET_NCP || i64(g[4]) != GPET_NE) return GPET_ERR_FORMAT;
        idcp = 1.0f / f32(g[3]);
        ide = 1.0f / f32(g[7]);
        i += 2; /* CP row, E row */
        for (int icp = 0; icp < GPET_NCP; ++icp) {
            while (i < L.size() && blank(L[i])) ++i;
            if (i >= L.size()) return GPET_ERR_FORMAT;
            auto t = tokens(L[i++]);
            if ((int)t.size() < GPET_NE) return GPET_ERR_FORMAT;
            for (int ie = 0; ie < GPET_NE; ++ie)
                vals[((size_t)m * GPET_NCP + icp) * GPET_NE + ie] = f32(t[ie]);
        }
    }
    return GPET_OK;
}

} // namespace

extern "C" {

const char *gpet_status_string(gpet_status s)
{
    switch (s) {
    case GPET_OK: return "ok";
    case GPET_ERR_ARG: return "bad argument";
    case GPET_ERR_IO: return "file missing or unreadable";
    case GPET_ERR_FORMAT: return "format or dimension mismatch";
    case GPET_ERR_DEVICE: return "device error";
    case GPET_ERR_CAPACITY: return "capacity exceeded";
    case GPET_ERR_STATE: return "call order";
    case GPET_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown";
}

/* rmater (initialize.cu:279-368): only [Emin_ph, Emax] and nmat are used;
 * then the four 1-D tables and the two surfaces. */
gpet_status gpet_tables_load(const char *data_dir, gpet_tables *out)
{
    if (!data_dir || !out) return GPET_ERR_ARG;
    memset(out, 0, sizeof(*out));
    std::string d(data_dir);
    std::vector<std::string> L;
    if (!read_lines((d + "/input4gPET.matter").c_str(), L)) return GPET_ERR_IO;
    float eminph = 0, emax = 0;
    int nmat = -1;
    for (size_t i = 0; i < L.size(); ++i) {
        if (L[i].find("[Emin_ph") != std::string::npos) {
            size_t j = i + 1;
            while (j < L.size() && blank(L[j])) ++j;
            if (j >= L.size()) return GPET_ERR_FORMAT;
            auto t = tokens(L[j]);
            if (t.size() < 3) return GPET_ERR_FORMAT;
            eminph = f32(t[0]);
            emax = f32(t[2]);
        }
        if (L[i].find("No of materials") != std::string::npos) {
            size_t j = i + 1;
            while (j < L.size() && blank(L[j])) ++j;
            if (j >= L.size()) return GPET_ERR_FORMAT;
            nmat = (int)i64(tokens(L[j])[0]);
            break;
        }
    }
    if (nmat <= 0 || nmat > GPET_MAXMAT) return GPET_ERR_FORMAT;
    OwnedTables *own = new OwnedTables();
    static const char *names[4] = {"lamph", "compt", "phote", "rayle"};
    gpet_status st = GPET_OK;
    for (int k = 0; k < 4 && st == GPET_OK; ++k)
        st = read_table1d(d + "/input4gPET." + names[k], nmat, own->mu[k], out->e0[k], out->idl[k]);
    if (st == GPET_OK) st = read_surface(d + "/input4gPET.cmpsf", nmat, own->surf[0], out->idcp[0], out->ide[0]);
    if (st == GPET_OK) st = read_surface(d + "/input4gPET.rayff", nmat, own->surf[1], out->idcp[1], out->ide[1]);
    if (st != GPET_OK) {
        delete own;
        memset(out, 0, sizeof(*out));
        return st;
    }
    out->nmat = nmat;
    out->eminph = eminph;
    out->emax = emax;
    out->n1d = GPET_NTAB;
    for (int k = 0; k < 4; ++k) out->mu[k] = own->mu[k].data();
    out->ncp = GPET_NCP;
    out->ne = GPET_NE;
    for (int k = 0; k < 2; ++k) out->surf[k] = own->surf[k].data();
    std::lock_guard<std::mutex> lk(g_owned_mu);
    g_owned[out->mu[0]] = own;
    return GPET_OK;
}

void gpet_tables_free(gpet_tables *t)
{
    if (!t || !t->mu[0]) return;
    OwnedTables *own = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_owned_mu);
        auto it = g_owned.find(t->mu[0]);
        if (it != g_owned.end()) {
            own = it->second;
            g_owned.erase(it);
        }
    }
    delete own;
    memset(t, 0, sizeof(*t));
}

gpet_status gpet_panel_counts(const gpet_panel *p0, int32_t *module_n, int32_t *crystal_n, int32_t *module_ny,
                              int32_t *crystal_ny)
{
    if (!p0) return GPET_ERR_ARG;
    int mny = (int)std::floor(p0->leny / (p0->mody + p0->mspy)) + 1;
    int cny = (int)std::floor(p0->mody / (p0->cry + p0->cspy)) + 1;
    int mn = mny * ((int)std::floor(p0->lenz / (p0->modz + p0->mspz)) + 1);
    int cn = cny * ((int)std::floor(p0->modz / (p0->crz + p0->cspz)) + 1);
    if (module_n) *module_n = mn;
    if (crystal_n) *crystal_n = cn;
    if (module_ny) *module_ny = mny;
    if (crystal_ny) *crystal_ny = cny;
    return GPET_OK;
}

/* read_file_ro (detector.cu:64-285).  Panel 0 is read; panels 1..n-1 are
 * rotated copies about `rot` by i*angle with the reference's formula, which
 * is a rotation only for an axis-aligned unit axis (Q14): other axes are
 * rejected.  Extension: an optional trailing label/value pair
 * "rings spacing" replicates the ring along the rotation axis (spacing in
 * cm, rings centred on the original ring); panel ids are ring-major. */
gpet_status gpet_read_geometry(const char *path, gpet_panel *panels, int32_t cap, int32_t *n_panels,
                               gpet_panel_material mats[2])
{
    if (!path || !n_panels) return GPET_ERR_ARG;
    Cursor c;
    if (!read_lines(path, c.L)) return GPET_ERR_IO;
    std::string lab;
    std::vector<std::string> v;
    if (!c.next_nonblank(lab) || !c.values(1, v)) return GPET_ERR_FORMAT;
    int count = (int)i64(v[0]);
    if (!c.next_nonblank(lab) || !c.values(3, v)) return GPET_ERR_FORMAT;
    float rot[3] = {f32(v[0]), f32(v[1]), f32(v[2])};
    if (!c.next_nonblank(lab) || !c.values(1, v)) return GPET_ERR_FORMAT;
    float rotAng = f32(v[0]);
    if (!c.next_nonblank(lab)) return GPET_ERR_FORMAT;
    gpet_panel_material pm[2];
    for (int i = 0; i < 2; ++i) {
        if (!c.values(2, v)) return GPET_ERR_FORMAT;
        pm[i].material = (int32_t)i64(v[0]);
        pm[i].density = f32(v[1]);
    }
    if (!c.next_nonblank(lab)) return GPET_ERR_FORMAT; /* separator */
    gpet_panel p0;
    memset(&p0, 0, sizeof(p0));
    float *dst3[] = {&p0.lenx, &p0.modx, &p0.mspx, &p0.crx, &p0.cspx, &p0.dirx, &p0.ox, p0.ux, p0.uy, p0.uz};
    if (!c.next_nonblank(lab) || !c.values(1, v)) return GPET_ERR_FORMAT;
    p0.panel = (int32_t)i64(v[0]);
    for (float *d3 : dst3) {
        if (!c.next_nonblank(lab) || !c.values(3, v)) return GPET_ERR_FORMAT;
        d3[0] = f32(v[0]);
        d3[1] = f32(v[1]);
        d3[2] = f32(v[2]);
    }
    int nrings = 1;
    float ring_dz = 0.0f;
    {
        size_t save = c.i;
        std::string l2;
        if (c.next_nonblank(l2)) { /* separator after panel 0 */
            std::string l3;
            std::vector<std::string> rv;
            if (c.next_nonblank(l3) && c.values(2, rv) && is_number(rv[0]) && is_number(rv[1])) {
                nrings = (int)i64(rv[0]);
                ring_dz = f32(rv[1]);
            }
        }
        c.i = save;
    }
    if (count < 1 || nrings < 1) return GPET_ERR_FORMAT;
    int nonzero = 0;
    for (int k = 0; k < 3; ++k) nonzero += (rot[k] != 0.0f);
    if (count > 1 && (nonzero != 1 || std::fabs(std::fabs(rot[0] + rot[1] + rot[2]) - 1.0f) > 1e-6f))
        return GPET_ERR_UNSUPPORTED;
    int total = count * nrings;
    *n_panels = total;
    if (mats) {
        mats[0] = pm[0];
        mats[1] = pm[1];
    }
    if (!panels) return GPET_OK;
    if (total > cap) return GPET_ERR_CAPACITY;
    const float PIF = 3.1415926535897932384626433f;
    for (int ring = 0; ring < nrings; ++ring) {
        float shift = ((float)ring - 0.5f * (float)(nrings - 1)) * ring_dz;
        for (int i = 0; i < count; ++i) {
            gpet_panel q = p0;
            q.panel = ring * count + i;
            if (i == 0 && ring == 0) q.panel = p0.panel;
            if (i > 0) {
                float ang = rotAng * PIF / 180.0f * (float)i;
                float ca = cosf(ang), sa = sinf(ang);
                auto rot3 = [&](const float *vec, float *o) {
                    o[0] = (1 - ca) * (vec[0] * rot[0]) * rot[0] + ca * vec[0] + sa * (rot[1] * vec[2] - rot[2] * vec[1]);
                    o[1] = (1 - ca) * (vec[1] * rot[1]) * rot[1] + ca * vec[1] + sa * (rot[2] * vec[0] - rot[0] * vec[2]);
                    o[2] = (1 - ca) * (vec[2] * rot[2]) * rot[2] + ca * vec[2] + sa * (rot[0] * vec[1] - rot[1] * vec[0]);
                };
                float o3[3] = {p0.ox, p0.oy, p0.oz}, r3[3];
                rot3(o3, r3);
                q.ox = r3[0]; q.oy = r3[1]; q.oz = r3[2];
                rot3(p0.ux, q.ux);
                rot3(p0.uy, q.uy);
                rot3(p0.uz, q.uz);
            }
            q.ox += shift * rot[0];
            q.oy += shift * rot[1];
            q.oz += shift * rot[2];
            panels[ring * count + i] = q;
        }
    }
    return GPET_OK;
}

/* readSource (initialize.cu:116-144): N, a header ending in '#', N rows. */
gpet_status gpet_read_sources(const char *path, gpet_source *out, int32_t cap, int32_t *n)
{
    if (!path || !n) return GPET_ERR_ARG;
    std::ifstream f(path);
    if (!f) return GPET_ERR_IO;
    int ns = 0;
    f >> ns;
    f.ignore(512, '#');
    if (!f || ns < 0) return GPET_ERR_FORMAT;
    *n = ns;
    if (!out) return GPET_OK;
    if (ns > cap) return GPET_ERR_CAPACITY;
    for (int i = 0; i < ns; ++i) {
        std::string a;
        double natom;
        f >> a;
        natom = f64(a);
        out[i].natom = natom;
        f >> out[i].isotope >> out[i].shape;
        for (int j = 0; j < 6; ++j) {
            std::string s;
            f >> s;
            out[i].coef[j] = f32(s);
        }
        if (!f) return GPET_ERR_FORMAT;
    }
    return GPET_OK;
}

/* loadIsotopes (initialize.cu:10-31). */
gpet_status gpet_read_isotopes(const char *path, gpet_isotope *out, int32_t cap, int32_t *n)
{
    if (!path || !n) return GPET_ERR_ARG;
    std::ifstream f(path);
    if (!f) return GPET_ERR_IO;
    int ni = 0;
    f >> ni;
    f.ignore(512, '#');
    if (!f || ni < 0) return GPET_ERR_FORMAT;
    *n = ni;
    if (!out) return GPET_OK;
    if (ni > cap) return GPET_ERR_CAPACITY;
    for (int i = 0; i < ni; ++i) {
        std::string s;
        f >> s; out[i].half_life_s = f32(s);
        f >> s; out[i].beta_ratio = f32(s);
        for (int j = 0; j < 8; ++j) {
            f >> s;
            out[i].coef[j] = f32(s);
        }
        if (!f) return GPET_ERR_FORMAT;
    }
    return GPET_OK;
}

/* readParticle (initialize.cu:76-115): 64-byte records of 8 doubles. */
gpet_status gpet_read_psf(const char *path, double *out, int64_t cap, int64_t *n)
{
    if (!path || !n) return GPET_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return GPET_ERR_IO;
    fseek(f, 0, SEEK_END);
    long long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    int64_t nrec = sz / 64;
    if (out) {
        if (nrec > cap) nrec = cap;
        size_t got = fread(out, 64, (size_t)nrec, f);
        nrec = (int64_t)got;
    }
    fclose(f);
    *n = nrec;
    return GPET_OK;
}

/* main.cu:50-184 field order; trailing optional extension pairs. */
gpet_status gpet_read_input(const char *path, gpet_input *in)
{
    if (!path || !in) return GPET_ERR_ARG;
    memset(in, 0, sizeof(*in));
    Cursor c;
    if (!read_lines(path, c.L)) return GPET_ERR_IO;
    std::string lab;
    std::vector<std::string> v;
    auto field = [&](size_t k) { return c.next_nonblank(lab) && c.values(k, v); };
    auto copy_str = [](char *dst, const std::string &s) {
        strncpy(dst, s.c_str(), GPET_PATH_MAX - 1);
        dst[GPET_PATH_MAX - 1] = 0;
    };
    if (!field(1)) return GPET_ERR_FORMAT;
    in->device = (int32_t)i64(v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->noncol_sigma = f32(v[0]);
    if (!field(3)) return GPET_ERR_FORMAT;
    for (int k = 0; k < 3; ++k) in->dims[k] = (int32_t)i64(v[k]);
    if (!field(3)) return GPET_ERR_FORMAT;
    for (int k = 0; k < 3; ++k) in->offset[k] = f32(v[k]);
    if (!field(3)) return GPET_ERR_FORMAT;
    for (int k = 0; k < 3; ++k) in->size[k] = f32(v[k]);
    if (!field(1)) return GPET_ERR_FORMAT;
    copy_str(in->mat_file, v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    copy_str(in->den_file, v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->n_histories = i64(v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->use_psf = (int32_t)i64(v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    copy_str(in->source_file, v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->ptype = (int32_t)i64(v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->prange = (int32_t)i64(v[0]);
    if (!field(2)) return GPET_ERR_FORMAT;
    in->t0 = f32(v[0]);
    in->t1 = f32(v[1]);
    if (!field(4)) return GPET_ERR_FORMAT;
    for (int k = 0; k < 4; ++k) in->record_sphere[k] = f32(v[k]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->eabs = f32(v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    copy_str(in->geo_file, v[0]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->n_surfaces = (int32_t)i64(v[0]);
    if (in->n_surfaces < 0 || in->n_surfaces > GPET_MAXSURFACE) return GPET_ERR_CAPACITY; /* main.cu:186-190 */
    if (in->n_surfaces > 0) {
        if (!c.values(10 * (size_t)in->n_surfaces, v)) return GPET_ERR_FORMAT;
        for (int k = 0; k < 10 * in->n_surfaces; ++k) in->surfaces[k] = f32(v[k]);
    }
    if (!field(2)) return GPET_ERR_FORMAT;
    in->readout_depth = (int32_t)i64(v[0]);
    in->readout_policy = (int32_t)i64(v[1]);
    if (!field(1)) return GPET_ERR_FORMAT;
    in->e_threshold = f32(v[0]);
    if (!field(5)) return GPET_ERR_FORMAT;
    in->blur_policy = (int32_t)i64(v[0]);
    in->e_ref = f32(v[1]);
    in->r_ref = f32(v[2]);
    in->slope = f32(v[3]);
    in->space_blur = f32(v[4]);
    if (!field(3)) return GPET_ERR_FORMAT;
    in->dead_level = (int32_t)i64(v[0]);
    in->dead_type = (int32_t)i64(v[1]);
    in->dead_time = f32(v[2]);
    if (!field(2)) return GPET_ERR_FORMAT;
    in->win_min = f32(v[0]);
    in->win_max = f32(v[1]);
    in->seed = 1;
    /* optional extension pairs */
    if (field(1)) in->time_blur_fwhm_us = f64(v[0]);
    if (field(2)) {
        in->coinc_window_us = f64(v[0]);
        in->coinc_policy = (int32_t)i64(v[1]);
    }
    if (field(1)) in->seed = (uint64_t)strtoull(v[0].c_str(), nullptr, 10);
    if (field(3)) {
        in->noise_rate = f64(v[0]);
        in->noise_e_mean = f32(v[1]);
        in->noise_e_sigma = f32(v[2]);
    }
    if (field(1)) in->delay_us = f64(v[0]);
    return GPET_OK;
}

void gpet_input_digitizer(const gpet_input *in, gpet_digitizer *d)
{
    memset(d, 0, sizeof(*d));
    d->readout_depth = in->readout_depth;
    d->readout_policy = in->readout_policy;
    d->e_threshold = in->e_threshold;
    d->e_upholder = 2000000.0f;
    d->blur_policy = in->blur_policy;
    d->e_ref = in->e_ref;
    d->r_ref = in->r_ref;
    d->slope = in->slope;
    d->space_blur = in->space_blur;
    d->dead_level = in->dead_level;
    d->dead_type = in->dead_type;
    d->dead_time_us = in->dead_time;
    d->win_min = in->win_min;
    d->win_max = in->win_max;
    d->time_blur_fwhm_us = in->time_blur_fwhm_us;
    d->coinc_window_us = in->coinc_window_us;
    d->coinc_policy = in->coinc_policy;
    d->record_rayleigh = 0;
    d->noise_rate = in->noise_rate;
    d->noise_e_mean = in->noise_e_mean;
    d->noise_e_sigma = in->noise_e_sigma;
    d->delay_us = in->delay_us;
}

gpet_status gpet_write_events(const char *path, const gpet_event *ev, int64_t n, int32_t append)
{
    if (!path || (n > 0 && !ev)) return GPET_ERR_ARG;
    FILE *f = fopen(path, append ? "ab" : "wb");
    if (!f) return GPET_ERR_IO;
    size_t w = n > 0 ? fwrite(ev, sizeof(gpet_event), (size_t)n, f) : 0;
    fclose(f);
    return (int64_t)w == n ? GPET_OK : GPET_ERR_IO;
}

gpet_status gpet_write_hits(const char *id_path, const char *f_path, const int32_t *hit_id, const float *hit_f,
                            int64_t n, int32_t append)
{
    if (!id_path || !f_path) return GPET_ERR_ARG;
    FILE *a = fopen(id_path, append ? "ab" : "wb");
    if (!a) return GPET_ERR_IO;
    FILE *b = fopen(f_path, append ? "ab" : "wb");
    if (!b) {
        fclose(a);
        return GPET_ERR_IO;
    }
    if (n > 0) {
        fwrite(hit_id, sizeof(int32_t) * 5, (size_t)n, a);
        fwrite(hit_f, sizeof(float) * 5, (size_t)n, b);
    }
    fclose(a);
    fclose(b);
    return GPET_OK;
}

gpet_status gpet_write_psf(const char *dir, const char *name, const gpet_psf_record *rec, int64_t n,
                           int32_t append)
{
    if (!dir || !name || (n > 0 && !rec)) return GPET_ERR_ARG;
    const std::string d(dir), nm(name);
    const char *mode = append ? "ab" : "wb";
    FILE *f[3] = {fopen((d + "/out" + nm + ".dat").c_str(), mode), fopen((d + "/id" + nm + ".dat").c_str(), mode),
                  fopen((d + "/time" + nm + ".dat").c_str(), mode)};
    bool ok = f[0] && f[1] && f[2];
    for (int64_t i = 0; ok && i < n; ++i) {
        ok = fwrite(&rec[i].x, sizeof(float), 7, f[0]) == 7 && fwrite(&rec[i].eventid, sizeof(int32_t), 1, f[1]) == 1 &&
             fwrite(&rec[i].t, sizeof(double), 1, f[2]) == 1;
    }
    for (FILE *x : f)
        if (x && fclose(x) != 0) ok = false;
    return ok ? GPET_OK : GPET_ERR_IO;
}

gpet_status gpet_listmode_records(const gpet_event *pairs, int64_t n_pairs, int32_t module_n, int32_t crystal_n,
                                  uint32_t flags, gpet_lm_record *out)
{
    if (n_pairs < 0 || (n_pairs > 0 && (!pairs || !out)) || module_n <= 0 || crystal_n <= 0) return GPET_ERR_ARG;
    for (int64_t k = 0; k < n_pairs; ++k) {
        const gpet_event &a = pairs[2 * k], &b = pairs[2 * k + 1];
        gpet_lm_record &r = out[k];
        r.crystal_a = (uint32_t)(((int64_t)a.pann * module_n + a.modn) * crystal_n + a.cryn);
        r.crystal_b = (uint32_t)(((int64_t)b.pann * module_n + b.modn) * crystal_n + b.cryn);
        r.t_us = a.t;
        r.dt_ps = (float)((b.t - a.t) * 1.0e6);
        r.e_a_kev = a.E * 1.0e-3f;
        r.e_b_kev = b.E * 1.0e-3f;
        r.flags = flags;
        if (a.parn < 0 || b.parn < 0) r.flags |= GPET_LM_BACKGROUND;
        else if (a.eventid == b.eventid && !(flags & GPET_LM_DELAYED)) r.flags |= GPET_LM_TRUE;
    }
    return GPET_OK;
}

gpet_status gpet_write_listmode(const char *path, const gpet_event *prompts, int64_t n_prompts,
                                const gpet_event *delayed, int64_t n_delayed, int32_t module_n, int32_t crystal_n,
                                int32_t n_panels, int32_t append)
{
    if (!path || n_prompts < 0 || n_delayed < 0 || (n_prompts > 0 && !prompts) || (n_delayed > 0 && !delayed) ||
        module_n <= 0 || crystal_n <= 0 || n_panels <= 0)
        return GPET_ERR_ARG;
    /* counts of what the file already holds (append mode) */
    int64_t old_p = 0, old_d = 0;
    const std::string hpath = std::string(path) + "h";
    if (append) {
        FILE *h = fopen(hpath.c_str(), "r");
        if (h) {
            char line[512];
            while (fgets(line, sizeof(line), h)) {
                long long v = 0;
                if (sscanf(line, "number of prompts := %lld", &v) == 1) old_p = v;
                if (sscanf(line, "number of delayed := %lld", &v) == 1) old_d = v;
            }
            fclose(h);
        }
    }
    FILE *f = fopen(path, append ? "ab" : "wb");
    if (!f) return GPET_ERR_IO;
    bool ok = true;
    std::vector<gpet_lm_record> buf;
    for (int pass = 0; pass < 2 && ok; ++pass) {
        const gpet_event *ev = pass == 0 ? prompts : delayed;
        const int64_t n = pass == 0 ? n_prompts : n_delayed;
        const int64_t chunk = 1 << 20;
        for (int64_t k0 = 0; k0 < n && ok; k0 += chunk) {
            const int64_t m = std::min(chunk, n - k0);
            buf.resize((size_t)m);
            gpet_listmode_records(ev + 2 * k0, m, module_n, crystal_n, pass == 0 ? 0u : (uint32_t)GPET_LM_DELAYED,
                                  buf.data());
            ok = fwrite(buf.data(), sizeof(gpet_lm_record), (size_t)m, f) == (size_t)m;
        }
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) return GPET_ERR_IO;
    FILE *h = fopen(hpath.c_str(), "w");
    if (!h) return GPET_ERR_IO;
    const char *base = strrchr(path, '/');
    base = base ? base + 1 : path;
    fprintf(h, "!GPET LIST MODE :=\n");
    fprintf(h, "name of data file := %s\n", base);
    fprintf(h, "record size (bytes) := %d\n", (int)sizeof(gpet_lm_record));
    fprintf(h, "record fields := uint32 crystal_a, uint32 crystal_b, float64 t_us, float32 dt_ps, "
               "float32 e_a_kev, float32 e_b_kev, uint32 flags\n");
    fprintf(h, "flags := 1 true (both singles from one decay), 2 delayed window, 4 background single\n");
    fprintf(h, "crystal index := (panel * modules per panel + module) * crystals per module + crystal\n");
    fprintf(h, "number of panels := %d\n", n_panels);
    fprintf(h, "modules per panel := %d\n", module_n);
    fprintf(h, "crystals per module := %d\n", crystal_n);
    fprintf(h, "number of prompts := %lld\n", (long long)(old_p + n_prompts));
    fprintf(h, "number of delayed := %lld\n", (long long)(old_d + n_delayed));
    fprintf(h, "record order := %s\n", append ? "runs appended, each run prompts then delayed"
                                               : "prompts (time order), then delayed");
    fprintf(h, "!END OF GPET LIST MODE :=\n");
    return fclose(h) == 0 ? GPET_OK : GPET_ERR_IO;
}

gpet_status gpet_write_sinogram(const char *prefix, const int64_t *sino, int32_t n_s, int32_t n_phi, float s_max)
{
    if (!prefix || !sino || n_s <= 0 || n_phi <= 0 || !(s_max > 0.0f)) return GPET_ERR_ARG;
    const std::string p(prefix);
    std::vector<float> v((size_t)n_s * n_phi);
    for (size_t k = 0; k < v.size(); ++k) v[k] = (float)sino[k];
    FILE *f = fopen((p + ".s").c_str(), "wb");
    if (!f) return GPET_ERR_IO;
    bool ok = fwrite(v.data(), sizeof(float), v.size(), f) == v.size();
    if (fclose(f) != 0) ok = false;
    if (!ok) return GPET_ERR_IO;
    FILE *h = fopen((p + ".hs").c_str(), "w");
    if (!h) return GPET_ERR_IO;
    const char *base = strrchr(prefix, '/');
    base = base ? base + 1 : prefix;
    fprintf(h, "!INTERFILE :=\n");
    fprintf(h, "!imaging modality := PT\n");
    fprintf(h, "name of data file := %s.s\n", base);
    fprintf(h, "!GENERAL DATA :=\n");
    fprintf(h, "!GENERAL IMAGE DATA :=\n");
    fprintf(h, "!type of data := PET\n");
    fprintf(h, "imagedata byte order := LITTLEENDIAN\n");
    fprintf(h, "!PET STUDY (General) :=\n");
    fprintf(h, "!PET data type := Emission\n");
    fprintf(h, "applied corrections := {None}\n");
    fprintf(h, "!number format := float\n");
    fprintf(h, "!number of bytes per pixel := 4\n");
    fprintf(h, "number of dimensions := 2\n");
    fprintf(h, "matrix axis label [2] := view\n");
    fprintf(h, "!matrix size [2] := %d\n", n_phi);
    fprintf(h, "matrix axis label [1] := tangential coordinate\n");
    fprintf(h, "!matrix size [1] := %d\n", n_s);
    fprintf(h, "scale factor (mm/pixel) [1] := %.9g\n", 20.0 * (double)s_max / n_s);
    fprintf(h, "scale factor (rad/pixel) [2] := %.9g\n", 3.14159265358979323846 / n_phi);
    fprintf(h, "tangential range (cm) := %.9g %.9g\n", -(double)s_max, (double)s_max);
    fprintf(h, "view range (rad) := 0 3.14159265358979\n");
    fprintf(h, "number of time frames := 1\n");
    fprintf(h, "!END OF INTERFILE :=\n");
    return fclose(h) == 0 ? GPET_OK : GPET_ERR_IO;
}

} /* extern "C" */
