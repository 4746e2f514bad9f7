/*
 * gpet_oracle.cpp - CPU oracle for the photon transport and detection path.
 *
 * TEST INFRASTRUCTURE.  This is a scalar, host-compiled restatement of the
 * reference's algorithm (utaresearch/gPET) used only to check the HIP path:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library never links or calls it.
 *
 * It exports the C ABI of include/gpet.h with the prefix gpet_cpu_ so tests
 * can run the same calls against either implementation.  Random numbers and
 * elementary functions come from the shared headers gp_philox.h / gp_math.h
 * (built with -ffp-contract=off), so per-history results are comparable bit
 * for bit with the GPU path.
 *
 * Every physics function cites the reference function it restates.  Where
 * the restatement departs from the reference on purpose (a quirk fixed) the
 * comment says so; the quirk list is Appendix B of the survey (Q1..Q22).
 *
 * Parity pins (what checks this file besides the GPU path):
 *   philox / math     - KATs and error bounds (tests/test_rng_math.py)
 *   table lerp        - node values of input4gPET.* (tests/test_tables.py)
 *   surface read      - node / mid-cell values and a chi-square against the
 *                       angular distribution the surface encodes
 *                       (tests/test_oracle_pins.py)
 *   majorant          - invariant mu*rho_max*lambda <= 1 on a dense grid
 *   samplers          - chi2/KS against closed forms (tests/test_oracle_physics.py)
 *   decay counts/times- chi-square against Binomial(n, p); per-source counts
 *                       in a run; KS of emission times against the truncated
 *                       exponential (tests/test_oracle_pins.py)
 *   positron walk     - Maxwell KS in homogeneous media; two-slab closed
 *                       form and brute-force crossing lists in heterogeneous
 *                       media, exits into air (tests/test_oracle_pins.py)
 *   geometry          - brute-force crystal index, closed-form panel entry
 *   transport         - narrow-beam exp(-mu L) in water and LSO
 *   adder / readout   - rebuilt from the hit records at every readout depth
 *                       and policy (tests/test_oracle_pins.py)
 *   digitizer         - brute-force O(n^2) restatements on small lists
 */
#include "../include/gpet.h"
#include "../gpet_amd/csrc/common/gp_math.h"
#include "../gpet_amd/csrc/common/gp_binomial.h"
#include "../gpet_amd/csrc/common/gp_walkmap.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Working precision of the transport.  The default build is FP32 (the GPU
 * path's precision, bit-compatible with it); -DGPET_ORACLE_FP64 builds the
 * same physics in FP64 with libm transcendentals, used as the reference of
 * the FP32 error report.  Tables, densities and output records stay FP32. */
#ifndef GPET_ORACLE_FP64
typedef float real;
#define R_LOG gp_logf
#define R_EXP gp_expf
#define R_SINCOS2PI gp_sincos2pi
#define R_SINCOS gp_sincosf
#define R_CBRT gp_cbrtf
#define R_NORMAL gp_rng_normal
#else
typedef double real;
#define R_LOG std::log
#define R_EXP std::exp
#define R_CBRT std::cbrt
static inline void R_SINCOS2PI(double u, double *s, double *c)
{
    *s = std::sin(6.283185307179586 * u);
    *c = std::cos(6.283185307179586 * u);
}
static inline void R_SINCOS(double x, double *s, double *c)
{
    *s = std::sin(x);
    *c = std::cos(x);
}
static inline double R_NORMAL(gp_rng *r)
{
    double u1 = gp_rng_uniform(r), u2 = gp_rng_uniform(r);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
#endif

namespace {

constexpr double C_LIGHT = 29979.2458;   /* cm/us, gPET_kernals.cu:264        */
constexpr double INV_C = 1.0 / 29979.2458;
constexpr float MC2F = 510999.1f;        /* constants.h:34                    */
constexpr float IMC2F = 1.95695061e-06f; /* constants.h:35                    */
constexpr float SZERO = 1.0e-4f;         /* constants.h:37                    */
constexpr float ZEROF = 1.0e-20f;        /* constants.h:36                    */
constexpr double MAXT = 1e20;            /* constants.h:18                    */
constexpr float RHO_OUTSIDE = 0.0012905f;/* gPET_kernals.cu:403               */
constexpr int MAX_PH_EVENTS = 4;         /* reference events[4]; bounded, drops counted (Q5) */
constexpr int STEP_CAP = 1000000;

struct Wck {
    float e0 = 0, idl = 0;
    std::vector<float> lam;
};

struct Tables {
    int nmat = 0;
    float eminph = 0, emax = 0;
    int n1d = 0;
    float e0[4] = {0}, idl[4] = {0};
    std::vector<float> mu[4];
    int ncp = 0, ne = 0;
    float idcp[2] = {0}, ide[2] = {0};
    std::vector<float> surf[2];
};

struct Phantom {
    int nx = 0, ny = 0, nz = 0;
    float offx = 0, offy = 0, offz = 0;
    float dx = 0, dy = 0, dz = 0, idx = 0, idy = 0, idz = 0;
    std::vector<int32_t> mat;
    std::vector<float> dens;
    std::vector<uint32_t> wmap;      /* uniform boxes per padded brick (gp_walkmap.h) */
    int wbx = 0, wby = 0;
    Wck wck;
};

struct Detector {
    std::vector<gpet_panel> p;
    gpet_panel_material mats[2];
    std::vector<float> quad;
    int nquad = 0;
    int moduleN = 0, crystalN = 0, moduleNy = 0, crystalNy = 0;
    Wck wck;
};

struct Counters {
    int64_t photons = 0, panel_entries = 0, hits = 0, dropped_events = 0, majorant = 0;
    int64_t phantom_steps = 0, range_steps = 0, panel_steps = 0, range_jumps = 0;
    void add(const Counters &o)
    {
        photons += o.photons; panel_entries += o.panel_entries; hits += o.hits;
        dropped_events += o.dropped_events; majorant += o.majorant;
        phantom_steps += o.phantom_steps; range_steps += o.range_steps; panel_steps += o.panel_steps;
        range_jumps += o.range_jumps;
    }
};

struct HitRec {
    int32_t id[5];
    float f[5];
};

/* Acquisition window (s) and first history of a job: where background
 * singles are generated. */
struct NoiseWin {
    double t0, t1;
    int64_t first;
};

struct Result {
    std::vector<gpet_psf_record> psf[2];
    std::vector<NoiseWin> noise_win;
    std::vector<gpet_event> delayed;  /* delayed-window pairs                     */
    int64_t n_noise = 0;
    std::vector<int64_t> job_first;   /* batched runs: first history of each job */
    std::vector<int64_t> job_counts;  /* batched runs: per job {histories, singles, pairs} */
    std::vector<gpet_event> adder, singles, coinc;
    std::vector<HitRec> hits;
    std::vector<gpet_trace> trace;
    int64_t n_histories = 0, n_threshold = 0, n_deadtime = 0, n_coinc_true = 0;
    Counters c;
    double sec_transport = 0, sec_digi = 0;
};

} // namespace

struct gpet_ctx {
    gpet_options opt;
    Tables tab;
    bool has_tab = false;
    std::vector<Phantom> slots = std::vector<Phantom>(1); /* slot 0: gpet_cpu_set_phantom */
    std::vector<char> slot_set = std::vector<char>(1, 0);
    bool has_ph = false;
    Detector det;
    bool has_det = false;
    gpet_digitizer dig;
    bool has_dig = false;
    Result last;
    std::string err;
    int threads = 0;
    gpet_launch launch = gpet_launch();
    std::vector<float> mlp_w;  /* surrogate parameters (set_params)        */
    int mlp_exact = 0;         /* 1: no bf16 rounding (finite differences) */
};

namespace {

gpet_status fail(gpet_ctx *c, gpet_status s, const char *msg)
{
    if (c) c->err = msg;
    return s;
}

/* ---- table interpolation ------------------------------------------------
 * The reference binds the tables to linear-filtered textures and reads them at
 * i = idl*(E-e0)+0.5 (gPET_kernals.cu:31-64); here the same lerp is explicit
 * in FP32 (texture filtering's 9-bit weights are not reproduced). */
inline real lerp1(const float *T, int n, real e0, real idl, real E)
{
    real x = (E - e0) * idl;
    x = std::fmin(std::fmax(x, 0.0f), (real)(n - 1));
    int j = (int)x;
    if (j > n - 2) j = n - 2;
    real f = x - (real)j;
    return T[j] + f * (T[j + 1] - T[j]);
}

inline real mu_tab(const Tables &t, int k, int mat, real E)
{
    return lerp1(t.mu[k].data() + (size_t)mat * t.n1d, t.n1d, t.e0[k], t.idl[k], E);
}

inline real lam_wck(const Wck &w, real E)
{
    return lerp1(w.lam.data(), GPET_NWCK, w.e0, w.idl, E);
}

/* Bilinear read of an inverse-CDF surface: tex3D(s_tex, inde+0.5, indcp+0.5,
 * mat+0.5) of gPET_kernals.cu:66-88,128-146 with clamp addressing. */
inline real surf_read(const Tables &t, int k, int mat, real E, real u)
{
    const int ne = t.ne, ncp = t.ncp;
    real x = E * t.ide[k];
    x = std::fmin(std::fmax(x, 0.0f), (real)(ne - 1));
    real y = u * t.idcp[k];
    y = std::fmin(std::fmax(y, 0.0f), (real)(ncp - 1));
    int ix = (int)x;
    if (ix > ne - 2) ix = ne - 2;
    int iy = (int)y;
    if (iy > ncp - 2) iy = ncp - 2;
    real fx = x - (real)ix, fy = y - (real)iy;
    const float *S = t.surf[k].data() + (size_t)mat * ncp * ne;
    real v00 = S[iy * ne + ix], v01 = S[iy * ne + ix + 1];
    real v10 = S[(iy + 1) * ne + ix], v11 = S[(iy + 1) * ne + ix + 1];
    real a = v00 + fx * (v01 - v00);
    real b = v10 + fx * (v11 - v10);
    real v = a + fy * (b - a);
    return std::fmin(std::fmax(v, -1.0f), 1.0f);
}

/* ---- Woodcock majorant: iniwck (initialize.cu:773-829, 919-966) ---------
 * Same node grid as the reference (NWCK nodes from eminph with spacing
 * (emax-eminph)/NWCK).  Q18 fix: the reference fits a cubic spline and reads
 * the majorant with linear filtering, so lambda*mu can exceed 1.  Here the
 * node value is 1/max of the linearly interpolated mu*rho_max over the two
 * adjacent intervals (exact maximum of a piecewise-linear function: interval
 * ends and the data nodes inside), shrunk by 1e-5 against rounding. */
double mu_lin_max(const Tables &t, const float *rho_max, double E)
{
    double best = 0.0;
    for (int m = 0; m < t.nmat; ++m) {
        if (rho_max[m] <= 0.0f) continue;
        double v = (double)(float)mu_tab(t, GPET_TAB_LAMPH, m, (float)E) * rho_max[m];
        if (v > best) best = v;
    }
    return best;
}

void build_wck(const Tables &t, const float *rho_max, Wck &w)
{
    w.e0 = t.eminph;
    real de = (t.emax * (1.0f - 1.0e-10f) - w.e0) / (real)GPET_NWCK;
    w.idl = 1.0f / de;
    w.lam.assign(GPET_NWCK, 0.0f);
    const double d_node = 1.0 / (double)t.idl[GPET_TAB_LAMPH];
    std::vector<double> imax(GPET_NWCK + 1, 0.0); /* max over [E_i, E_i+1] */
    for (int i = 0; i <= GPET_NWCK; ++i) {
        double lo = (double)w.e0 + (double)de * i;
        double hi = (double)w.e0 + (double)de * (i + 1);
        double m = std::max(mu_lin_max(t, rho_max, lo), mu_lin_max(t, rho_max, hi));
        double k0 = std::ceil((lo - t.e0[GPET_TAB_LAMPH]) / d_node);
        double k1 = std::floor((hi - t.e0[GPET_TAB_LAMPH]) / d_node);
        for (double k = k0; k <= k1; k += 1.0)
            m = std::max(m, mu_lin_max(t, rho_max, (double)t.e0[GPET_TAB_LAMPH] + k * d_node));
        imax[i] = m;
    }
    for (int i = 0; i < GPET_NWCK; ++i) {
        double m = imax[i];
        if (i > 0) m = std::max(m, imax[i - 1]);
        w.lam[i] = (m > 0.0) ? (real)(1.0 / (m * (1.0 + 1.0e-5))) : 1.0e20f;
    }
}

/* ---- geometry helpers -------------------------------------------------- */

/* getAbsVox (gPET_kernals.cu:19-29), Q1 fixed: voxel 0 is inside and
 * truncation toward zero no longer maps (-1,0) into the grid. */
inline bool vox_index(const Phantom &p, real x, real y, real z, int &ix, int &iy, int &iz)
{
    real fx = (x - p.offx) * p.idx;
    real fy = (y - p.offy) * p.idy;
    real fz = (z - p.offz) * p.idz;
    if (!(fx >= 0.0f && fx < (real)p.nx && fy >= 0.0f && fy < (real)p.ny && fz >= 0.0f &&
          fz < (real)p.nz))
        return false;
    ix = (int)fx; iy = (int)fy; iz = (int)fz;
    return true;
}

/* rotate (gPET_kernals.cu:172-254).  The azimuth arrives as a uniform u with
 * phi = 2 pi u.  rho2 is recomputed after a renormalisation (the reference
 * keeps the stale value). */
inline void rotate_dir(real &u, real &v, real &w, real costh, real uphi)
{
    real rho2 = u * u + v * v;
    real norm = rho2 + w * w;
    if (std::fabs(norm - 1.0f) > SZERO) {
        norm = 1.0f / std::sqrt(norm);
        u = u * norm; v = v * norm; w = w * norm;
        rho2 = u * u + v * v;
    }
    real sinphi, cosphi;
    R_SINCOS2PI(uphi, &sinphi, &cosphi);
    real temp = costh * costh;
    if (rho2 > ZEROF) {
        real sthrho = (temp < 1.0f) ? std::sqrt((1.0f - temp) / rho2) : 0.0f;
        real urho = u * sthrho;
        real vrho = v * sthrho;
        real nu = u * costh - vrho * sinphi + (w * urho) * cosphi;
        real nv = v * costh + urho * sinphi + (w * vrho) * cosphi;
        real nw = w * costh - (rho2 * sthrho) * cosphi;
        u = nu; v = nv; w = nw;
    } else {
        real sinth = (temp < 1.0f) ? std::sqrt(1.0f - temp) : 0.0f;
        v = sinth * sinphi;
        if (w > 0.0f) { u = sinth * cosphi; w = costh; }
        else { u = -sinth * cosphi; w = -costh; }
    }
}

/* rotate with sin(theta) given: used for the pair's noncollinearity, where
 * sqrt(1 - cos^2) in FP32 would round every |delta| < 3.5e-4 rad to a
 * collinear pair (the reference computes it that way). */
inline void rotate_dir_cs(real &u, real &v, real &w, real costh, real sinth, real uphi)
{
    real rho2 = u * u + v * v;
    real norm = rho2 + w * w;
    if (std::fabs(norm - 1.0f) > SZERO) {
        norm = 1.0f / std::sqrt(norm);
        u = u * norm; v = v * norm; w = w * norm;
        rho2 = u * u + v * v;
    }
    real sinphi, cosphi;
    R_SINCOS2PI(uphi, &sinphi, &cosphi);
    if (rho2 > ZEROF) {
        real sthrho = sinth / std::sqrt(rho2);
        real urho = u * sthrho;
        real vrho = v * sthrho;
        real nu = u * costh - vrho * sinphi + (w * urho) * cosphi;
        real nv = v * costh + urho * sinphi + (w * vrho) * cosphi;
        real nw = w * costh - (rho2 * sthrho) * cosphi;
        u = nu; v = nv; w = nw;
    } else {
        v = sinth * sinphi;
        if (w > 0.0f) { u = sinth * cosphi; w = costh; }
        else { u = -sinth * cosphi; w = -costh; }
    }
}

/* Analytic Klein-Nishina sampling (gPET_kernals.cu:90-126), FP32. */
inline void kn_sample(real E, gp_rng &r, real &efrac, real &costhe)
{
    real e0 = E * IMC2F;
    real twoe = 2.0f * e0;
    real kmin2 = 1.0f / ((1.0f + twoe) * (1.0f + twoe));
    real loge = R_LOG(1.0f + twoe);
    real a1 = loge + (twoe * (1.0f + e0)) * kmin2;
    for (int it = 0; it < STEP_CAP; ++it) {
        real u1 = gp_rng_uniform(&r);
        real u2 = gp_rng_uniform(&r);
        if (u1 * a1 < loge) efrac = R_EXP(-u2 * loge);
        else efrac = std::sqrt(kmin2 + u2 * (1.0f - kmin2));
        real mess = ((e0 * e0) * efrac) * (1.0f + efrac * efrac);
        real u3 = gp_rng_uniform(&r);
        if (u3 * mess <= mess - (1.0f - efrac) * ((1.0f + twoe) * efrac - 1.0f)) break;
    }
    costhe = 1.0f - (1.0f - efrac) / (efrac * e0);
}

/* crystalSearch (gPET_kernals.cu:1236-1279): m_id 0 crystal, 1 gap/excluded. */
inline int crystal_search(const Detector &d, const gpet_panel &P, real x, real y, real z,
                          int &M_id, int &L_id)
{
    for (int q = 0; q < d.nquad; ++q) {
        const float *s = &d.quad[q * 10];
        real v = s[0] * x * x + s[1] * y * y + s[2] * z * z + s[3] * x * y + s[4] * x * z +
                  s[5] * y * z + s[6] * x + s[7] * y + s[8] * z + s[9];
        if (v < 0.0f) return 1;
    }
    real yy = P.leny / 2.0f + y;
    real zz = P.lenz / 2.0f + z;
    real mpy = P.mody + P.mspy, mpz = P.modz + P.mspz;
    int my = std::floor(yy / mpy) > 0.0f ? (int)(yy / mpy) : 0;
    int mz = std::floor(zz / mpz) > 0.0f ? (int)(zz / mpz) : 0;
    M_id = mz * d.moduleNy + my;
    yy = yy - (real)my * mpy;
    zz = zz - (real)mz * mpz;
    if (yy > P.mody || zz > P.modz) return 1;
    real cpy = P.cry + P.cspy, cpz = P.crz + P.cspz;
    int ly = std::floor(yy / cpy) > 0.0f ? (int)(yy / cpy) : 0;
    int lz = std::floor(zz / cpz) > 0.0f ? (int)(zz / cpz) : 0;
    L_id = lz * d.crystalNy + ly;
    yy = yy - (real)ly * cpy;
    zz = zz - (real)lz * cpz;
    if (yy > P.cry || zz > P.crz) return 1;
    return 0;
}

/* adder (gPET_kernals.cu:737-755); capacity-checked (Q5). */
inline void adder(gpet_event *ev, int &n, const gpet_event &e, Counters &c)
{
    for (int i = 0; i < n; ++i) {
        if (ev[i].siten == e.siten) {
            real sE = ev[i].E + e.E;
            ev[i].x = (ev[i].x * ev[i].E + e.x * e.E) / sE;
            ev[i].y = (ev[i].y * ev[i].E + e.y * e.E) / sE;
            ev[i].z = (ev[i].z * ev[i].E + e.z * e.E) / sE;
            ev[i].E = sE;
            return;
        }
    }
    if (n < MAX_PH_EVENTS) ev[n++] = e;
    else c.dropped_events++;
}

/* readout (gPET_kernals.cu:756-813).  depth 3 keeps the crystal-level list
 * (the reference returns without setting the output count). */
inline int readout(gpet_event *ev, int n, int depth, int policy, int moduleN)
{
    if (depth == 3) return n;
    if (policy == 1) depth = 2;
    for (int i = 0; i < n; ++i) {
        if (depth == 0) ev[i].siten = 0;
        else if (depth == 1) ev[i].siten = ev[i].pann;
        else ev[i].siten = ev[i].pann * moduleN + ev[i].modn;
    }
    int out = 0;
    for (int i = 0; i < n; ++i) {
        gpet_event e0 = ev[i];
        if (e0.t > MAXT * 0.1) continue;
        for (int j = i + 1; j < n; ++j) {
            gpet_event e = ev[j];
            if (e.parn == e0.parn && e.siten == e0.siten) {
                if (policy == 1) {
                    real sE = e0.E + e.E;
                    e0.x = (e0.x * e0.E + e.x * e.E) / sE;
                    e0.y = (e0.y * e0.E + e.y * e.E) / sE;
                    e0.z = (e0.z * e0.E + e.z * e.E) / sE;
                    e0.E = sE;
                    ev[j].t = MAXT;
                    continue;
                }
                e0 = (e0.E > e.E) ? e0 : e;
                ev[j].t = MAXT;
            }
        }
        ev[out++] = e0;
    }
    return out;
}

/* ---- per-photon transport ---------------------------------------------- */

struct PhotonOut {
    gpet_event ev[MAX_PH_EVENTS];
    int nev = 0;
    std::vector<HitRec> hits;
    gpet_psf_record psf[2];    /* source, phantom exit                     */
    bool has_psf[2] = {false, false};
};

/* photon (gPET_kernals.cu:256-345): Woodcock tracking in the phantom.
 * Returns false if the photon is absorbed.  Q10 fixed: a Compton event that
 * leaves less than eabs absorbs the photon. */
bool phantom_transport(const gpet_ctx &C, const Phantom &p, real &x, real &y, real &z, real &vx, real &vy,
                       real &vz, real &E, double &tof, gp_rng &r, gpet_trace *tr, Counters &cnt)
{
    const Tables &t = C.tab;
    for (int it = 0; it < STEP_CAP; ++it) {
        /* one 4-word block per step: length, selector, scattering CDF, azimuth */
        gp_u4 b = gp_rng_block(&r);
        real lammin = lam_wck(p.wck, E);
        real s = -lammin * R_LOG(gp_u01(b.x));
        x = x + s * vx;
        y = y + s * vy;
        z = z + s * vz;
        tof = tof + (double)s * INV_C;
        int ix, iy, iz;
        if (!vox_index(p, x, y, z, ix, iy, iz)) return true; /* escaped (Q9: landing point) */
        if (tr) tr->n_steps++;
        cnt.phantom_steps++;
        size_t vid = ((size_t)iz * p.ny + iy) * p.nx + ix;
        real den = p.dens[vid];
        int mat = p.mat[vid];
        real lamden = lammin * den;
        real prob = 1.0f - lamden * mu_tab(t, GPET_TAB_LAMPH, mat, E);
        if (prob < 0.0f) { prob = 0.0f; cnt.majorant++; }
        real rn = gp_u01(b.y);
        if (rn < prob) continue;
        prob = prob + lamden * mu_tab(t, GPET_TAB_COMPT, mat, E);
        if (rn < prob) {
            real cth = surf_read(t, GPET_SURF_COMPTON, mat, E, gp_u01(b.z));
            real efrac = 1.0f / (1.0f + (E * IMC2F) * (1.0f - cth));
            real uphi = gp_u01(b.w);
            if (tr) tr->n_compton++;
            E = E * efrac;
            if (E < C.opt.eabs) return false;
            rotate_dir(vx, vy, vz, cth, uphi);
            continue;
        }
        prob = prob + lamden * mu_tab(t, GPET_TAB_RAYLE, mat, E);
        if (rn < prob) {
            real cth = surf_read(t, GPET_SURF_RAYLEIGH, mat, E, gp_u01(b.z));
            real uphi = gp_u01(b.w);
            if (tr) tr->n_rayleigh++;
            rotate_dir(vx, vy, vz, cth, uphi);
            continue;
        }
        return false; /* photoelectric */
    }
    return false;
}

/* photonde (gPET_kernals.cu:839-1233) for one photon: panel entry, Woodcock
 * in the panel, hit records, adder, readout. */
void panel_transport(const gpet_ctx &C, real x, real y, real z, real vx, real vy, real vz,
                     real E, double tof, int32_t parn, int32_t eid, gp_rng &r, gpet_trace *tr,
                     PhotonOut &out, Counters &cnt, bool want_hits)
{
    const Detector &d = C.det;
    const Tables &t = C.tab;
    int pa = -1;
    real x2 = 0, y2 = 0, z2 = 0, vx2 = 0, vy2 = 0, vz2 = 0;
    const int np = (int)d.p.size();
    for (int i = 0; i < np; ++i) {
        const gpet_panel &P = d.p[i];
        real rx = x - P.ox, ry = y - P.oy, rz = z - P.oz;
        real tx = rx * P.ux[0] + ry * P.ux[1] + rz * P.ux[2];
        real ty = rx * P.uy[0] + ry * P.uy[1] + rz * P.uy[2];
        real tz = rx * P.uz[0] + ry * P.uz[1] + rz * P.uz[2];
        real tvx = vx * P.ux[0] + vy * P.ux[1] + vz * P.ux[2];
        real tvy = vx * P.uy[0] + vy * P.uy[1] + vz * P.uy[2];
        real tvz = vx * P.uz[0] + vy * P.uz[1] + vz * P.uz[2];
        if (tvx * P.dirx >= 0.0f) {
            real ty2 = ty - (tx / tvx) * tvy;
            real tz2 = tz - (tx / tvx) * tvz;
            if (std::fabs(ty2) < P.leny / 2.0f && std::fabs(tz2) < P.lenz / 2.0f) {
                x2 = 0.0f; y2 = ty2; z2 = tz2;
                tof = tof + (double)(-tx) / (C_LIGHT * (double)tvx);
                vx2 = tvx; vy2 = tvy; vz2 = tvz;
                pa = i;
                break;
            }
        }
    }
    if (tr) tr->panel = pa;
    if (pa < 0) return;
    cnt.panel_entries++;
    const gpet_panel &P = d.p[pa];
    const int pann = P.panel;
    const real hx = P.lenx, hy = P.leny * 0.5f, hz = P.lenz * 0.5f;
    int nev = 0;
    gpet_event *ev = out.ev;
    auto record = [&](int type, real edep, int M_id, int L_id, bool to_adder) {
        cnt.hits++;
        if (want_hits) {
            HitRec h;
            h.id[0] = parn; h.id[1] = pann; h.id[2] = M_id; h.id[3] = L_id; h.id[4] = type;
            h.f[0] = edep; h.f[1] = (real)tof; h.f[2] = x2; h.f[3] = y2; h.f[4] = z2;
            out.hits.push_back(h);
        }
        if (!to_adder) return;
        gpet_event e;
        e.parn = parn; e.pann = pann; e.modn = M_id; e.cryn = L_id;
        e.siten = pann * d.moduleN * d.crystalN + M_id * d.crystalN + L_id;
        e.eventid = eid; e.t = tof; e.E = edep; e.x = x2; e.y = y2; e.z = z2;
        adder(ev, nev, e, cnt);
    };
    for (int it = 0; it < STEP_CAP; ++it) {
        /* block per step: length, selector, Rayleigh CDF, azimuth; the KN
         * rejection loop draws after the block */
        gp_u4 b = gp_rng_block(&r);
        real lammin = lam_wck(d.wck, E);
        real s = -lammin * R_LOG(gp_u01(b.x));
        x2 = x2 + s * vx2;
        y2 = y2 + s * vy2;
        z2 = z2 + s * vz2;
        tof = tof + (double)s * INV_C;
        if (std::fabs(y2) > hy || std::fabs(z2) > hz || x2 * P.dirx < 0.0f || x2 * P.dirx > hx) break;
        if (tr) tr->n_steps_panel++;
        cnt.panel_steps++;
        int M_id = -1, L_id = -1;
        int m_id = crystal_search(d, P, x2, y2, z2, M_id, L_id);
        real den = d.mats[m_id].density;
        int mat = d.mats[m_id].material;
        real lamden = lammin * den;
        real prob = 1.0f - lamden * mu_tab(t, GPET_TAB_LAMPH, mat, E);
        if (prob < 0.0f) { prob = 0.0f; cnt.majorant++; }
        real rn = gp_u01(b.y);
        if (rn < prob) continue;
        if (tr) tr->n_int_panel++;
        prob = prob + lamden * mu_tab(t, GPET_TAB_COMPT, mat, E);
        if (rn < prob) {
            real efrac, cth;
            kn_sample(E, r, efrac, cth);
            real de = E * (1.0f - efrac);
            real uphi = gp_u01(b.w);
            if (m_id == 0) record(1, de, M_id, L_id, true);
            E = E - de;
            if (E < C.opt.eabs) {
                if (m_id == 0) record(2, E, M_id, L_id, true);
                break;
            }
            rotate_dir(vx2, vy2, vz2, cth, uphi);
            continue;
        }
        prob = prob + lamden * mu_tab(t, GPET_TAB_RAYLE, mat, E);
        if (rn < prob) {
            real cth = surf_read(t, GPET_SURF_RAYLEIGH, mat, E, gp_u01(b.z));
            real uphi = gp_u01(b.w);
            if (C.dig.record_rayleigh && m_id == 0) record(3, 0.0f, M_id, L_id, false);
            rotate_dir(vx2, vy2, vz2, cth, uphi);
            continue;
        }
        if (m_id == 0) record(4, E, M_id, L_id, true);
        break;
    }
    if (nev > 0) nev = readout(ev, nev, C.dig.readout_depth, C.dig.readout_policy, d.moduleN);
    out.nev = nev;
    if (tr) tr->n_events = nev;
}

/* Phantom + panel for one photon, filling the trace record. */
void photon_full(const gpet_ctx &C, const Phantom *ph, int64_t hist, int k, real x, real y, real z, real vx,
                 real vy, real vz, real E, double t, int32_t parn, gpet_trace *tr,
                 PhotonOut &out, Counters &cnt, bool want_hits)
{
    gp_rng r;
    gp_rng_init(&r, C.opt.seed, (uint64_t)hist, (uint32_t)(GP_TAG_PHOTON + k), 0u);
    cnt.photons++;
    if (tr) {
        tr->status |= GPET_TR_ALIVE;
        tr->x0 = x; tr->y0 = y; tr->z0 = z; tr->E0 = E; tr->t0 = t;
        tr->panel = -1;
    }
    bool alive = true;
    auto record = [&](int which) {
        gpet_psf_record &p = out.psf[which];
        p.x = (float)x; p.y = (float)y; p.z = (float)z; p.vx = (float)vx; p.vy = (float)vy; p.vz = (float)vz;
        p.E = (float)E; p.parn = parn; p.eventid = (int32_t)hist; p.pad = 0; p.t = t;
        out.has_psf[which] = true;
    };
    if (C.opt.flags & GPET_FLAG_PSF_SOURCE) record(0);
    if (ph) alive = phantom_transport(C, *ph, x, y, z, vx, vy, vz, E, t, r, tr, cnt);
    if (alive && C.opt.record_sphere[3] > 0.0f) {
        /* getDistance (gPET_kernals.cu:148-171, 288-294): move onto the
         * recording sphere; far root from inside, near root when
         * approaching from outside, the receding root otherwise */
        const float *sp = C.opt.record_sphere;
        const real cx = x - sp[0], cy = y - sp[1], cz = z - sp[2];
        const real a = vx * vx + vy * vy + vz * vz;
        const real b = 2.0f * (vx * cx + vy * cy + vz * cz);
        const real c = (cx * cx + cy * cy + cz * cz) - sp[3] * sp[3];
        const real disc = b * b - 4.0f * a * c;
        real sd = 0.0f;
        if (!(disc < 0.0f)) {
            if (c < 0.0f) sd = (-b + std::sqrt(disc)) / (2.0f * a);
            else if (b < 0.0f) sd = (-b - std::sqrt(disc)) / (2.0f * a);
            else sd = (-b + std::sqrt(disc)) / (2.0f * a);
        }
        x = x + sd * vx;
        y = y + sd * vy;
        z = z + sd * vz;
        t = t + (double)sd * INV_C;
    }
    if (alive && (C.opt.flags & GPET_FLAG_PSF_PHANTOM)) record(1);
    if (tr) {
        tr->status |= alive ? GPET_TR_ESCAPED : GPET_TR_ABSORBED;
        tr->x = x; tr->y = y; tr->z = z; tr->vx = vx; tr->vy = vy; tr->vz = vz; tr->E = E; tr->t = t;
    }
    if (alive && C.has_det) {
        panel_transport(C, x, y, z, vx, vy, vz, E, t, parn, (int32_t)hist, r, tr, out, cnt, want_hits);
        if (tr && tr->panel >= 0) tr->status |= GPET_TR_PANEL;
    }
    if (tr) tr->draws = r.n;
}

/* ---- source ------------------------------------------------------------ */

/* sampleEkPositron (gPET_kernals.cu:420-443): rejection from the fitted
 * spectrum; Horner form of the degree-5 polynomial; returns kinetic eV. */
real sample_ek(const gpet_isotope &iso, gp_rng &r)
{
    real E = 0.511f;
    for (int it = 0; it < 100000; ++it) {
        E = gp_rng_uniform(&r) * (iso.coef[0] - 0.511f) + 0.511f;
        real u = iso.coef[1] * gp_rng_uniform(&r);
        real p = iso.coef[2];
        for (int i = 3; i < 8; ++i) p = p * E + iso.coef[i];
        if (u <= p) break;
    }
    return (E - 0.511f) * 1.0e6f;
}

/* getPositionFromShape (gPET_kernals.cu:445-482). */
void sample_shape(const gpet_source &s, gp_rng &r, real &x, real &y, real &z)
{
    const float *c = s.coef;
    int shape = (s.shape < 0 || s.shape > 3) ? 0 : s.shape;
    if (shape == 0) {
        x = c[0] + (c[3] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
        y = c[1] + (c[4] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
        z = c[2] + (c[5] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
    } else if (shape == 1) {
        real sp, cp;
        R_SINCOS2PI(gp_rng_uniform(&r), &sp, &cp);
        real rr = c[3] * std::sqrt(gp_rng_uniform(&r));
        x = c[0] + rr * cp;
        y = c[1] + rr * sp;
        z = c[2] + (c[4] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
    } else {
        real sp, cp;
        R_SINCOS2PI(gp_rng_uniform(&r), &sp, &cp);
        real ct = 2.0f * gp_rng_uniform(&r) - 1.0f;
        real rr = R_CBRT(gp_rng_uniform(&r));
        real st = std::sqrt(1.0f - ct * ct);
        if (shape == 2) {
            rr = c[3] * rr;
            x = c[0] + (rr * st) * cp;
            y = c[1] + (rr * st) * sp;
            z = c[2] + rr * ct;
        } else { /* 3: ellipsoid with semi-axes c3, c4, c5 (extension) */
            x = c[0] + ((c[3] * rr) * st) * cp;
            y = c[1] + ((c[4] * rr) * st) * sp;
            z = c[2] + (c[5] * rr) * ct;
        }
    }
}

/* setPositronRange (gPET_kernals.cu:347-418).  The range r = |sigma N3| is a
 * mass thickness; the positron advances until the accumulated sum of
 * segment length x density reaches r.  The voxel walk is restated as an
 * incremental 3-D DDA (Amanatides-Woo): per axis the distance to the next
 * boundary and the boundary spacing along the unit direction, advanced by
 * the smallest (ties: x, then y, then z, as the reference's sequential
 * `step > b` tests).  Kept (Q8): a segment is charged at the density of
 * the voxel it enters.  Fixes: a crossing out of the grid charges
 * RHO_OUTSIDE and the remaining range is covered at RHO_OUTSIDE (the
 * reference reads the clamped border voxel and then adds the remainder to
 * a leftover 1000-unit step). */
inline void dda_axis(real x, real off, real d, int i, real u, real &tmax, real &tdelta, int &step)
{
    if (u > 0.0f) {
        real inv = 1.0f / u;
        tmax = (off + (real)(i + 1) * d - x) * inv;
        tdelta = d * inv;
        step = 1;
    } else if (u < 0.0f) {
        real inv = 1.0f / u;
        tmax = (off + (real)i * d - x) * inv;
        tdelta = -(d * inv);
        step = -1;
    } else {
        tmax = INFINITY;
        tdelta = INFINITY;
        step = 0;
    }
}

/* Next-crossing parameter along one axis from voxel i (the DDA's first
 * boundary, as dda_axis computes it). */
inline real dda_next(real x, real off, real d, int i, real u, int step)
{
    if (step > 0) return (off + (real)(i + 1) * d - x) * (1.0f / u);
    if (step < 0) return (off + (real)i * d - x) * (1.0f / u);
    return INFINITY;
}

/* One jump of the walk across a uniform box of the phantom's walk map
 * (gp_walkmap.h).  The walk sits in voxel (ix, iy, iz) at parameter tcur
 * with accumulated charge s; every crossing before the box's exit enters a
 * voxel of the box's density rho, so the walk moves in one step to the
 * crossing into the box's last voxel layer on its exit side (the exit of
 * the box shrunk by one layer there: the crossing after it, possibly out of
 * the box, is left to the voxel steps, which charge it at the density of
 * the voxel it enters).  Ties between axes go to x, then y, then z, as in
 * the voxel steps.  The new voxel along the other axes is the one holding
 * the point at the new parameter, kept inside the box.  Returns true when
 * the range runs out inside the box (t set). */
bool range_jump(const Phantom &p, real x, real y, real z, real ux, real uy, real uz, real rr, int &ix, int &iy,
                int &iz, real &tmx, real &tmy, real &tmz, int sx, int sy, int sz, real &tcur, real &s, real &t,
                int64_t &cnt_jumps)
{
    const size_t b = ((size_t)((iz + 1) >> 2) * p.wby + (size_t)((iy + 1) >> 2)) * p.wbx + (size_t)((ix + 1) >> 2);
    const uint32_t *w = &p.wmap[4 * b];
    cnt_jumps++;
    float rho_f;
    memcpy(&rho_f, &w[3], 4);
    if (!(rho_f >= 0.0f)) return false;
    const real rho = rho_f;
    const real pos[3] = {x, y, z}, u[3] = {ux, uy, uz}, off[3] = {p.offx, p.offy, p.offz}, d[3] = {p.dx, p.dy, p.dz};
    const real id[3] = {p.idx, p.idy, p.idz};
    const int st[3] = {sx, sy, sz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = (int)(w[a] & 0xffffu) - 1;      /* unpadded voxel bounds */
        hi[a] = (int)(w[a] >> 16) - 1;
    }
    real T = INFINITY;
    int ax = -1;
    for (int a = 0; a < 3; ++a) {
        if (st[a] == 0) continue;
        /* s > 0: the crossing into voxel hi (boundary hi); s < 0: into voxel
         * lo (boundary lo + 1) */
        const int k = st[a] > 0 ? hi[a] : lo[a] + 1;
        const real ta = (off[a] + (real)k * d[a] - pos[a]) * (1.0f / u[a]);
        if (ta < T) { T = ta; ax = a; }
    }
    if (!(T > tcur)) return false;
    const real sn = s + (T - tcur) * rho;
    if (sn >= rr) {
        t = tcur + (rr - s) / rho;
        return true;
    }
    s = sn;
    tcur = T;
    int idx3[3];
    for (int a = 0; a < 3; ++a) {
        int i;
        if (a == ax) {
            i = st[a] > 0 ? hi[a] : lo[a];
        } else {
            /* along an axis the walk does not move this is its voxel */
            i = (int)std::floor(((pos[a] + T * u[a]) - off[a]) * id[a]);
            i = std::min(std::max(i, lo[a]), hi[a]);
        }
        idx3[a] = i;
    }
    ix = idx3[0]; iy = idx3[1]; iz = idx3[2];
    tmx = dda_next(x, p.offx, p.dx, ix, ux, sx);
    tmy = dda_next(y, p.offy, p.dy, iy, uy, sy);
    tmz = dda_next(z, p.offz, p.dz, iz, uz, sz);
    return false;
}

/* The walk of a positron with mass range rr (g/cm^2 over the reference's
 * unit density: rr cm of water) from (x, y, z) along the unit direction
 * (ux, uy, uz): the loop of setPositronRange, gPET_kernals.cu:380-415, in
 * groups of WALK_GROUP voxel crossings, each group preceded by a jump
 * across a uniform box when the walk sits in one (range_jump). */
constexpr int WALK_GROUP = 4;
void range_walk(const Phantom &p, real &x, real &y, real &z, real ux, real uy, real uz, real rr, Counters &cnt)
{
    int ix = 0, iy = 0, iz = 0;
    real t;
    if (!(p.nx > 0 && vox_index(p, x, y, z, ix, iy, iz))) {
        t = rr / RHO_OUTSIDE;
    } else {
        real tmx, tmy, tmz, tdx, tdy, tdz;
        int sx, sy, sz;
        dda_axis(x, p.offx, p.dx, ix, ux, tmx, tdx, sx);
        dda_axis(y, p.offy, p.dy, iy, uy, tmy, tdy, sy);
        dda_axis(z, p.offz, p.dz, iz, uz, tmz, tdz, sz);
        real tcur = 0.0f, s = 0.0f;
        t = 0.0f;
        bool done = false;
        for (int it = 0; it < STEP_CAP && !done; it += WALK_GROUP) {
            if (!p.wmap.empty() && range_jump(p, x, y, z, ux, uy, uz, rr, ix, iy, iz, tmx, tmy, tmz, sx, sy, sz, tcur,
                                              s, t, cnt.range_jumps))
                break;
            for (int k = 0; k < WALK_GROUP; ++k) {
                real tnext;
                if (tmx <= tmy && tmx <= tmz) { tnext = tmx; tmx = tmx + tdx; ix += sx; }
                else if (tmy <= tmz) { tnext = tmy; tmy = tmy + tdy; iy += sy; }
                else { tnext = tmz; tmz = tmz + tdz; iz += sz; }
                bool in = ix >= 0 && ix < p.nx && iy >= 0 && iy < p.ny && iz >= 0 && iz < p.nz;
                real rho = RHO_OUTSIDE;
                if (in) {
                    rho = p.dens[((size_t)iz * p.ny + iy) * p.nx + ix];
                    cnt.range_steps++;
                }
                s = s + (tnext - tcur) * rho;
                if (s >= rr) {
                    t = (s > rr) ? tnext + (rr - s) / rho : tnext;
                    done = true;
                    break;
                }
                if (!in) {
                    t = tnext + (rr - s) / RHO_OUTSIDE;
                    done = true;
                    break;
                }
                tcur = tnext;
            }
        }
    }
    x = x + ux * t;
    y = y + uy * t;
    z = z + uz * t;
}

/* setPositronRange (gPET_kernals.cu:347-418): a Gaussian 3-D displacement
 * of sigma = Rex/2 per axis (Rex from the positron's kinetic energy) gives
 * the range rr = |displacement| and, unless usedir, the direction; then the
 * walk above. */
void positron_range(const Phantom &p, real &x, real &y, real &z, real ek_ev, int usedir,
                    real vx, real vy, real vz, gp_rng &r, Counters &cnt)
{
    real ekin = ek_ev / 1.0e6f;
    real rex = ((0.1f * 5.44040782f) * ekin * ekin) / (0.369516529f + ekin);
    real sigma = rex / 2.0f;
    real dx = sigma * R_NORMAL(&r);
    real dy = sigma * R_NORMAL(&r);
    real dz = sigma * R_NORMAL(&r);
    real rr = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (usedir) {
        real tmp = std::sqrt(vx * vx + vy * vy + vz * vz);
        dx = rr * vx / tmp;
        dy = rr * vy / tmp;
        dz = rr * vz / tmp;
    }
    if (!(rr > 0.0f)) return;
    real inv_rr = 1.0f / rr;
    range_walk(p, x, y, z, dx * inv_rr, dy * inv_rr, dz * inv_rr, rr, cnt);
}

/* Pair emission (gPET_kernals.cu:536-555): isotropic direction, Gaussian
 * noncollinearity delta; photon 1 along d with E = mc2 + delta mc2/2,
 * photon 2 rotated by (cos = -cos delta, phi) with E = mc2 - delta mc2/2. */
void emit_pair(real sigma, gp_rng &r, real v[2][3], real e[2])
{
    real ct = 2.0f * gp_rng_uniform(&r) - 1.0f;
    real sp, cp;
    R_SINCOS2PI(gp_rng_uniform(&r), &sp, &cp);
    real st = std::sqrt(1.0f - ct * ct);
    v[0][0] = st * cp; v[0][1] = st * sp; v[0][2] = ct;
    real uphi = gp_rng_uniform(&r);
    real delta = R_NORMAL(&r) * sigma;
    e[0] = MC2F + (delta * MC2F) / 2.0f;
    e[1] = MC2F - (delta * MC2F) / 2.0f;
    v[1][0] = v[0][0]; v[1][1] = v[0][1]; v[1][2] = v[0][2];
    real sd, cd;
    R_SINCOS(delta, &sd, &cd);
    rotate_dir_cs(v[1][0], v[1][1], v[1][2], -cd, std::fabs(sd), uphi);
}

/* Per-source decay plan shared by all histories of a run. */
struct SourcePlan {
    std::vector<int64_t> first;   /* first history of each source (prefix) */
    std::vector<double> tau, pdec; /* T1/2/ln2 (s) and 1-2^(-(t1-t0)/T1/2) */
    std::vector<double> t0;       /* acquisition start of each source (s)  */
    std::vector<int> phantom;     /* phantom slot of each source           */
    std::vector<const gpet_source *> src;
    std::vector<int64_t> job_first; /* batched runs: first history per job */
    int64_t total = 0;
};

/* Decay counts of one acquisition's sources, appended to the plan.
 * src_base is the index of src[0] in the run's source array (it keys the
 * binomial stream).  Counts: the exact Binomial(navail, pdec) of
 * gp_binomial.h (shared with the library's run planning; pinned on its own
 * by a chi-square against the binomial distribution). */
gpet_status plan_job(const gpet_ctx &C, const gpet_source *src, int ns, const gpet_isotope *iso, int niso, double t0,
                     double t1, int64_t n_decays, int src_base, int phantom, SourcePlan &pl)
{
    if (pl.first.empty()) pl.first.assign(1, 0);
    std::vector<double> expect(ns, 0.0), navail(ns, 0.0), tau(ns), pdec(ns);
    for (int i = 0; i < ns; ++i) {
        if (src[i].isotope < 0 || src[i].isotope >= niso) return GPET_ERR_ARG;
        double th = iso[src[i].isotope].half_life_s;
        if (!(th > 0.0)) return GPET_ERR_ARG;
        tau[i] = th / std::log(2.0);
        pdec[i] = -std::expm1(-(t1 - t0) / tau[i]);
        navail[i] = std::floor(src[i].natom * std::exp2(-t0 / th));
        expect[i] = navail[i] * pdec[i];
    }
    std::vector<int64_t> cnt(ns, 0);
    if (n_decays > 0) {
        double tot = 0;
        for (double e : expect) tot += e;
        if (tot <= 0) return GPET_ERR_ARG;
        int64_t assigned = 0;
        std::vector<std::pair<double, int>> rem;
        for (int i = 0; i < ns; ++i) {
            double share = (double)n_decays * expect[i] / tot;
            cnt[i] = (int64_t)std::floor(share);
            assigned += cnt[i];
            rem.push_back({share - (double)cnt[i], i});
        }
        std::stable_sort(rem.begin(), rem.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
            return a.first > b.first;
        });
        for (int k = 0; assigned < n_decays; ++k, ++assigned) cnt[rem[k % ns].second]++;
    } else {
        for (int i = 0; i < ns; ++i) cnt[i] = gp_binomial_draw(navail[i], pdec[i], C.opt.seed, src_base + i);
    }
    for (int i = 0; i < ns; ++i) {
        pl.first.push_back(pl.first.back() + cnt[i]);
        pl.tau.push_back(tau[i]);
        pl.pdec.push_back(pdec[i]);
        pl.t0.push_back(t0);
        pl.phantom.push_back(phantom);
        pl.src.push_back(&src[i]);
    }
    pl.total = pl.first.back();
    return GPET_OK;
}

/* One decay history (setPosition, gPET_kernals.cu:483-561, sampled per
 * decay instead of per atom, F8/Q17). */
void decay_history(const gpet_ctx &C, const SourcePlan &pl, const gpet_isotope *iso, int64_t h, gpet_trace *tr2,
                   PhotonOut out[2], Counters &cnt, bool want_hits)
{
    int si = (int)(std::upper_bound(pl.first.begin(), pl.first.end(), h) - pl.first.begin()) - 1;
    const gpet_source &S = *pl.src[si];
    const gpet_isotope &I = iso[S.isotope];
    const int slot = pl.phantom[si];
    const Phantom *ph = C.slot_set[slot] ? &C.slots[slot] : nullptr;
    static const Phantom none;
    gp_rng r;
    gp_rng_init(&r, C.opt.seed, (uint64_t)h, GP_TAG_DECAY, 0u);
    double u = gp_rng_uniform_d(&r);
    double ts = pl.t0[si] + pl.tau[si] * (-gp_log1p_d(-u * pl.pdec[si]));
    double tus = ts * 1.0e6;
    real rb = gp_rng_uniform(&r);
    if (tr2) {
        for (int k = 0; k < 2; ++k) {
            tr2[k].history = h; tr2[k].photon = k; tr2[k].panel = -1; tr2[k].t0 = tus;
        }
    }
    if (!(rb < I.beta_ratio)) return;
    real x, y, z;
    gp_rng_align(&r);
    sample_shape(S, r, x, y, z);
    if (C.opt.positron_range) {
        gp_rng_align(&r);
        real ek = sample_ek(I, r);
        gp_rng_align(&r);
        positron_range(ph ? *ph : none, x, y, z, ek, 0, 0.0f, 0.0f, 0.0f, r, cnt);
    }
    real v[2][3], e[2];
    gp_rng_align(&r);
    emit_pair(C.opt.noncol_sigma, r, v, e);
    for (int k = 0; k < 2; ++k)
        photon_full(C, ph, h, k, x, y, z, v[k][0], v[k][1], v[k][2], e[k], tus, (int32_t)(2 * h + k),
                    tr2 ? &tr2[k] : nullptr, out[k], cnt, want_hits);
}

/* One PSF history (simulateParticle / setPositionForPhoton,
 * gPET_kernals.cu:563-604).  Q3/Q11: the pair's photons both carry the
 * record time; records with t <= 0 are still transported. */
void psf_history(const gpet_ctx &C, const double *rec, int ptype, int64_t h, gpet_trace *tr2,
                 PhotonOut out[2], Counters &cnt, bool want_hits)
{
    real x = (real)rec[0], y = (real)rec[1], z = (real)rec[2];
    double t = rec[3];
    real vx = (real)rec[4], vy = (real)rec[5], vz = (real)rec[6], E = (real)rec[7];
    if (tr2) {
        for (int k = 0; k < 2; ++k) {
            tr2[k].history = h; tr2[k].photon = k; tr2[k].panel = -1; tr2[k].t0 = t;
        }
    }
    const Phantom *ph = C.slot_set[0] ? &C.slots[0] : nullptr;
    static const Phantom none;
    if (ptype == 1) {
        photon_full(C, ph, h, 0, x, y, z, vx, vy, vz, E, t, (int32_t)h, tr2 ? &tr2[0] : nullptr, out[0],
                    cnt, want_hits);
        return;
    }
    gp_rng r;
    gp_rng_init(&r, C.opt.seed, (uint64_t)h, GP_TAG_DECAY, 0u);
    if (C.opt.positron_range) positron_range(ph ? *ph : none, x, y, z, E, 1, vx, vy, vz, r, cnt);
    real v[2][3], e[2];
    gp_rng_align(&r);
    emit_pair(C.opt.noncol_sigma, r, v, e);
    for (int k = 0; k < 2; ++k)
        photon_full(C, ph, h, k, x, y, z, v[k][0], v[k][1], v[k][2], e[k], t, (int32_t)(2 * h + k),
                    tr2 ? &tr2[k] : nullptr, out[k], cnt, want_hits);
}

/* ---- digitizer (gPET.cu:380-424, gPET_kernals.cu:607-698,814-837) ------ */

inline bool ev_less_t(const gpet_event &a, const gpet_event &b)
{
    if (a.t != b.t) return a.t < b.t;
    if (a.parn != b.parn) return a.parn < b.parn;
    return a.siten < b.siten;
}

inline int site_at(const gpet_event &e, int level, const Detector &d)
{
    if (level == 0) return 0;
    if (level == 1) return e.pann;
    if (level == 2) return e.pann * d.moduleN + e.modn;
    return e.pann * d.moduleN * d.crystalN + e.modn * d.crystalN + e.cryn;
}

/* blur (gPET_kernals.cu:814-837) plus the time blur the reference lacks.
 * Stream: id = parn, aux = GP_TAG_DIGI | siten << 4 (unique per event). */
void blur_event(const gpet_ctx &C, gpet_event &e)
{
    const gpet_digitizer &g = C.dig;
    gp_rng r;
    gp_rng_init(&r, C.opt.seed, (uint64_t)(uint32_t)e.parn, (uint32_t)GP_TAG_DIGI | ((uint32_t)e.siten << 4), 0u);
    real R = 0.0f;
    if (g.blur_policy == 0) R = std::sqrt(g.e_ref / e.E) * g.r_ref;
    if (g.blur_policy == 1) R = g.r_ref + g.slope * (e.E - g.e_ref) / 1.0e6f;
    if (R < 0.0f) R = 0.0f;
    real n = R_NORMAL(&r);
    e.E = e.E + ((n * R) * e.E) / 2.35482f;
    if (g.space_blur > 0.0f) {
        e.x = e.x + g.space_blur * R_NORMAL(&r);
        e.y = e.y + g.space_blur * R_NORMAL(&r);
        e.z = e.z + g.space_blur * R_NORMAL(&r);
    }
    if (g.time_blur_fwhm_us > 0.0) {
        real nt = R_NORMAL(&r);
        e.t = e.t + (double)nt * (g.time_blur_fwhm_us / 2.35482);
    }
}

/* deadtime (gPET_kernals.cu:657-698) on (site,t)-sorted events.  Chain
 * starts are taken from the unmodified times (the reference reads times
 * other threads may already have overwritten).  Returns alive flags. */
void deadtime(std::vector<gpet_event> &ev, const std::vector<int64_t> &site, double tau, int type,
              std::vector<char> &alive)
{
    size_t n = ev.size();
    alive.assign(n, 1);
    size_t i = 0;
    while (i < n) {
        size_t j = i + 1;
        double tdead = ev[i].t;
        while (j < n && site[j] == site[i]) {
            if (ev[j].t < tdead + tau) {
                alive[j] = 0;
                if (type == 0) tdead = ev[j].t;
            } else {
                tdead = ev[j].t;
            }
            ++j;
        }
        i = j;
    }
}

/* Job of a history: the largest j with job_first[j] <= h (0 for one job). */
inline int64_t job_of(const std::vector<int64_t> &jf, int64_t h)
{
    if (jf.size() <= 2) return 0;
    return (int64_t)(std::upper_bound(jf.begin(), jf.end(), h) - jf.begin()) - 1;
}

/* Readout site of a crystal (gPET_kernals.cu:756-813, as readout above). */
inline int readout_site(int depth, int policy, int pann, int modn, int cryn, const Detector &d)
{
    if (depth == 3) return pann * d.moduleN * d.crystalN + modn * d.crystalN + cryn;
    if (policy == 1) depth = 2;
    if (depth == 0) return 0;
    if (depth == 1) return pann;
    return pann * d.moduleN + modn;
}

/* Crystal centre in panel-local coordinates. */
inline void crystal_local(const Detector &d, int pann, int modn, int cryn, float &xl, float &yl, float &zl)
{
    const gpet_panel &P = d.p[pann];
    int my = modn % d.moduleNy, mz = modn / d.moduleNy;
    int ly = cryn % d.crystalNy, lz = cryn / d.crystalNy;
    xl = (P.dirx * P.lenx) * 0.5f;
    yl = (-P.leny * 0.5f + (float)my * (P.mody + P.mspy)) + ((float)ly * (P.cry + P.cspy) + P.cry * 0.5f);
    zl = (-P.lenz * 0.5f + (float)mz * (P.modz + P.mspz)) + ((float)lz * (P.crz + P.cspz) + P.crz * 0.5f);
}

/* Background singles: the reference's addnoise (gPET_kernals.cu:699-735,
 * never launched) draws exponential arrival gaps of mean lambda per thread
 * interval, an energy Emean + sigma N(0,1), a uniform panel/module/crystal
 * and placeholder positions, with parn -1.  Restated sequentially: the
 * window [t0, t1] of job j (in us) is cut into chunks of 64 mean gaps (the
 * last ends at t1); chunk c's arrivals come from the Philox stream
 * (id c, aux 5 | j << 4), at most 4096 per chunk, and arrival m's crystal and
 * energy from (id c, aux 6 | j << 4) draws 8m..8m+7.  The k-th single (in
 * job, chunk, arrival order) has parn -1-k, eventid = the job's first
 * history, the crystal centre as position and the readout site. */
void gen_noise(const gpet_ctx &C, const std::vector<NoiseWin> &win, std::vector<gpet_event> &out)
{
    const gpet_digitizer &g = C.dig;
    const Detector &d = C.det;
    const double gap = 1.0e6 / g.noise_rate;
    const double L = 64.0 * gap;
    const int np = (int)d.p.size();
    int64_t k = 0;
    for (size_t j = 0; j < win.size(); ++j) {
        const double t0 = win[j].t0 * 1.0e6, t1 = win[j].t1 * 1.0e6;
        const double T = t1 - t0;
        const int64_t nch = T > 0.0 ? (int64_t)std::ceil(T / L) : 0;
        for (int64_t c = 0; c < nch; ++c) {
            double t = t0 + (double)c * L;
            const double end = (c + 1 == nch) ? t1 : t0 + (double)(c + 1) * L;
            gp_rng rt;
            gp_rng_init(&rt, C.opt.seed, (uint64_t)c, (uint32_t)GP_TAG_NOISE_T | ((uint32_t)j << 4), 0u);
            for (unsigned m = 0; m < 4096u; ++m) {
                t = t + (-gp_log_d(gp_rng_uniform_d(&rt))) * gap;
                if (!(t < end)) break;
                gp_rng ra;
                gp_rng_init(&ra, C.opt.seed, (uint64_t)c, (uint32_t)GP_TAG_NOISE_A | ((uint32_t)j << 4), 8u * m);
                const gp_u4 b0 = gp_rng_block(&ra), b1 = gp_rng_block(&ra);
                gpet_event e;
                e.pann = (int)(((uint64_t)b0.z * (uint64_t)np) >> 32);
                e.modn = (int)(((uint64_t)b0.w * (uint64_t)d.moduleN) >> 32);
                e.cryn = (int)(((uint64_t)b1.x * (uint64_t)d.crystalN) >> 32);
                e.siten = readout_site(g.readout_depth, g.readout_policy, e.pann, e.modn, e.cryn, d);
                e.parn = -1 - (int)k;
                e.eventid = (int32_t)win[j].first;
                e.t = t;
                e.E = fmaxf(g.noise_e_mean + g.noise_e_sigma * gp_normal(gp_u01(b0.x), gp_u01(b0.y)), 1.0f);
                crystal_local(d, e.pann, e.modn, e.cryn, e.x, e.y, e.z);
                out.push_back(e);
                ++k;
            }
        }
    }
}

/* Delayed-window randoms estimate (new): the prompt windows' openers (the
 * sequential rule below) each look at the singles of their job with
 * delay <= t - t_open <= delay + W; one such single makes a delayed pair
 * with the opener, more make the pairs of {opener, those singles} under the
 * multiples policy. */
void delayed_pairs(const gpet_digitizer &g, const std::vector<int64_t> &jf, const std::vector<gpet_event> &sg,
                   std::vector<gpet_event> &out)
{
    out.clear();
    const double W = g.coinc_window_us, D = g.delay_us;
    const size_t n = sg.size();
    size_t i = 0;
    while (i < n) {
        const int64_t ji = job_of(jf, sg[i].eventid);
        size_t j0 = i + 1;
        while (j0 < n && job_of(jf, sg[j0].eventid) == ji && sg[j0].t - sg[i].t < D) ++j0;
        size_t e = j0;
        while (e < n && job_of(jf, sg[e].eventid) == ji && sg[e].t - sg[i].t <= D + W) ++e;
        const size_t m = e - j0;
        if (m == 1 || (m > 1 && g.coinc_policy == GPET_COINC_TAKE_ALL_GOODS)) {
            std::vector<size_t> set(1, i);
            for (size_t q = j0; q < e; ++q) set.push_back(q);
            for (size_t a = 0; a < set.size(); ++a)
                for (size_t b = a + 1; b < set.size(); ++b) {
                    out.push_back(sg[set[a]]);
                    out.push_back(sg[set[b]]);
                }
        }
        size_t j = i + 1;
        while (j < n && sg[j].t - sg[i].t <= W && job_of(jf, sg[j].eventid) == ji) ++j;
        i = j;
    }
}

void run_digitizer(gpet_ctx &C, Result &R)
{
    const gpet_digitizer &g = C.dig;
    const Detector &d = C.det;
    const std::vector<int64_t> &jf = R.job_first;
    std::vector<gpet_event> noise;
    if (g.noise_rate > 0.0) gen_noise(C, R.noise_win, noise);
    R.n_noise = (int64_t)noise.size();
    std::vector<gpet_event> ev;
    ev.reserve(R.adder.size() + noise.size());
    std::vector<gpet_event> all(R.adder);
    all.insert(all.end(), noise.begin(), noise.end());
    for (gpet_event e : all) {
        blur_event(C, e);
        if (e.E < g.e_threshold || e.E > g.e_upholder) continue; /* energywindow */
        ev.push_back(e);
    }
    R.n_threshold = (int64_t)ev.size();
    /* setSitenum runs whenever dlevel != 3, deadtime or not (gPET.cu:164-169);
     * the rewritten site index is what singles.dat carries */
    if (g.dead_level != 3)
        for (auto &e : ev) e.siten = site_at(e, g.dead_level, d);
    if (g.dead_time_us > 0.0f && !ev.empty()) {
        std::vector<size_t> idx(ev.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        /* deadtime sites are per job: key (job, siten) */
        std::vector<int64_t> key(ev.size());
        for (size_t i = 0; i < ev.size(); ++i)
            key[i] = (job_of(jf, ev[i].eventid) << 32) | (int64_t)(uint32_t)ev[i].siten;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            if (key[a] != key[b]) return key[a] < key[b];
            return ev_less_t(ev[a], ev[b]);
        });
        std::vector<gpet_event> s(ev.size());
        std::vector<int64_t> site(ev.size());
        for (size_t i = 0; i < idx.size(); ++i) { s[i] = ev[idx[i]]; site[i] = key[idx[i]]; }
        std::vector<char> alive;
        deadtime(s, site, (double)g.dead_time_us, g.dead_type, alive);
        ev.clear();
        for (size_t i = 0; i < s.size(); ++i)
            if (alive[i]) ev.push_back(s[i]);
    }
    R.n_deadtime = (int64_t)ev.size();
    std::vector<gpet_event> sg;
    sg.reserve(ev.size());
    for (auto &e : ev)
        if (!(e.E < g.win_min || e.E > g.win_max)) sg.push_back(e);
    std::sort(sg.begin(), sg.end(), [&](const gpet_event &a, const gpet_event &b) {
        int64_t ja = job_of(jf, a.eventid), jb = job_of(jf, b.eventid);
        if (ja != jb) return ja < jb;
        return ev_less_t(a, b);
    });
    R.singles = sg;
    R.coinc.clear();
    R.n_coinc_true = 0;
    if (g.coinc_window_us > 0.0) {
        const double W = g.coinc_window_us;
        size_t n = sg.size(), i = 0;
        while (i < n) {
            size_t j = i + 1;
            const int64_t ji = job_of(jf, sg[i].eventid);
            while (j < n && sg[j].t - sg[i].t <= W && job_of(jf, sg[j].eventid) == ji) ++j;
            size_t m = j - i;
            if (m == 2 || (m > 2 && g.coinc_policy == GPET_COINC_TAKE_ALL_GOODS)) {
                for (size_t a = i; a < j; ++a)
                    for (size_t b = a + 1; b < j; ++b) {
                        R.coinc.push_back(sg[a]);
                        R.coinc.push_back(sg[b]);
                        if (sg[a].eventid == sg[b].eventid && sg[a].parn >= 0 && sg[b].parn >= 0)
                            R.n_coinc_true++;
                    }
            }
            i = j;
        }
    }
    R.delayed.clear();
    if (g.coinc_window_us > 0.0 && g.delay_us > 0.0) delayed_pairs(g, jf, sg, R.delayed);
    if (jf.size() >= 2) {
        const size_t nj = jf.size() - 1;
        R.job_counts.assign(3 * nj, 0);
        for (size_t j = 0; j < nj; ++j) R.job_counts[3 * j] = jf[j + 1] - jf[j];
        for (const auto &e : R.singles) R.job_counts[3 * job_of(jf, e.eventid) + 1]++;
        for (size_t k = 0; k + 1 < R.coinc.size(); k += 2) R.job_counts[3 * job_of(jf, R.coinc[k].eventid) + 2]++;
    }
}

/* Sinogram bin of a coincidence (new; the reference has no coincidences):
 * crystal centres of both singles in global coordinates, the line through
 * them in the transaxial (x, y) plane as (phi in [0, pi), signed distance s
 * = x1 sin(phi) - y1 cos(phi)), binned uniformly; -1 outside the grid. */
inline void crystal_centre(const Detector &d, const gpet_event &e, float &gx, float &gy)
{
    const gpet_panel &P = d.p[e.pann];
    int my = e.modn % d.moduleNy, mz = e.modn / d.moduleNy;
    int ly = e.cryn % d.crystalNy, lz = e.cryn / d.crystalNy;
    float xl = (P.dirx * P.lenx) * 0.5f;
    float yl = (-P.leny * 0.5f + (float)my * (P.mody + P.mspy)) + ((float)ly * (P.cry + P.cspy) + P.cry * 0.5f);
    float zl = (-P.lenz * 0.5f + (float)mz * (P.modz + P.mspz)) + ((float)lz * (P.crz + P.cspz) + P.crz * 0.5f);
    gx = P.ox + xl * P.ux[0] + yl * P.uy[0] + zl * P.uz[0];
    gy = P.oy + xl * P.ux[1] + yl * P.uy[1] + zl * P.uz[1];
}

inline int sino_bin(const Detector &d, const gpet_event &a, const gpet_event &b, int n_s, int n_phi, float s_max)
{
    if (a.pann < 0 || a.pann >= (int)d.p.size() || b.pann < 0 || b.pann >= (int)d.p.size()) return -1;
    float x1, y1, x2, y2;
    crystal_centre(d, a, x1, y1);
    crystal_centre(d, b, x2, y2);
    float dx = x2 - x1, dy = y2 - y1;
    if (dx == 0.0f && dy == 0.0f) return -1;
    const float PIF = 3.14159274f;
    float phi = gp_atan2f(dy, dx);
    if (phi < 0.0f) phi = phi + PIF;
    if (phi >= PIF) phi = phi - PIF;
    float sp, cp;
    gp_sincosf(phi, &sp, &cp);
    float sd = x1 * sp - y1 * cp;
    int ip = (int)(phi * ((float)n_phi / PIF));
    if (ip > n_phi - 1) ip = n_phi - 1;
    float fs = (sd + s_max) * ((float)n_s / (2.0f * s_max));
    if (!(fs >= 0.0f && fs < (float)n_s)) return -1;
    return ip * n_s + (int)fs;
}

bool ev_less_id(const gpet_event &a, const gpet_event &b)
{
    if (a.parn != b.parn) return a.parn < b.parn;
    return a.siten < b.siten;
}

/* Run the transport over n histories with OpenMP, collect, digitize. */
template <class F>
void run_histories(gpet_ctx &C, int64_t n, F hist_fn, gpet_sink *sink, Result &R,
                   const std::vector<int64_t> *job_first = nullptr, const std::vector<NoiseWin> *noise_win = nullptr)
{
    R = Result();
    if (job_first) R.job_first = *job_first;
    if (noise_win) R.noise_win = *noise_win;
    R.n_histories = n;
    const bool want_trace = sink && sink->trace && sink->trace_cap > 0;
    const bool want_hits = C.opt.record_hits != 0;
    auto t_begin = std::chrono::steady_clock::now();
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = C.threads > 0 ? C.threads : omp_get_max_threads();
#endif
    std::vector<std::vector<gpet_event>> tev(nthreads);
    std::vector<std::vector<HitRec>> thit(nthreads);
    std::vector<std::vector<gpet_psf_record>> tpsf[2] = {std::vector<std::vector<gpet_psf_record>>(nthreads),
                                                         std::vector<std::vector<gpet_psf_record>>(nthreads)};
    std::vector<Counters> tcnt(nthreads);
    std::vector<gpet_trace> trace;
    if (want_trace) trace.assign((size_t)std::min<int64_t>(2 * n, sink->trace_cap), gpet_trace());
    const int64_t ntr = (int64_t)trace.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int64_t h = 0; h < n; ++h) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        PhotonOut out[2];
        gpet_trace tr2[2];
        memset(tr2, 0, sizeof(tr2));
        bool tracing = want_trace && 2 * h < ntr;
        hist_fn(h, tracing ? tr2 : nullptr, out, tcnt[tid], want_hits);
        for (int k = 0; k < 2; ++k) {
            for (int i = 0; i < out[k].nev; ++i) tev[tid].push_back(out[k].ev[i]);
            for (auto &hr : out[k].hits) thit[tid].push_back(hr);
            for (int w = 0; w < 2; ++w)
                if (out[k].has_psf[w]) tpsf[w][tid].push_back(out[k].psf[w]);
        }
        if (tracing) {
            trace[2 * h] = tr2[0];
            if (2 * h + 1 < ntr) trace[2 * h + 1] = tr2[1];
        }
    }
    for (int t = 0; t < nthreads; ++t) {
        R.adder.insert(R.adder.end(), tev[t].begin(), tev[t].end());
        R.hits.insert(R.hits.end(), thit[t].begin(), thit[t].end());
        for (int w = 0; w < 2; ++w) R.psf[w].insert(R.psf[w].end(), tpsf[w][t].begin(), tpsf[w][t].end());
        R.c.add(tcnt[t]);
    }
    std::sort(R.adder.begin(), R.adder.end(), ev_less_id);
    for (int w = 0; w < 2; ++w)
        std::sort(R.psf[w].begin(), R.psf[w].end(),
                  [](const gpet_psf_record &a, const gpet_psf_record &b) { return a.parn < b.parn; });
    /* per photon, hits stay in the order they happened */
    std::stable_sort(R.hits.begin(), R.hits.end(), [](const HitRec &a, const HitRec &b) { return a.id[0] < b.id[0]; });
    R.trace.swap(trace);
    auto t_mid = std::chrono::steady_clock::now();
    run_digitizer(C, R);
    auto t_end = std::chrono::steady_clock::now();
    R.sec_transport = std::chrono::duration<double>(t_mid - t_begin).count();
    R.sec_digi = std::chrono::duration<double>(t_end - t_mid).count();
}

template <class T>
void copy_out(const std::vector<T> &v, T *dst, int64_t cap, int64_t &n, int64_t per = 1)
{
    n = (int64_t)v.size() / per;
    if (dst && cap > 0 && n > 0) {
        int64_t m = std::min<int64_t>(n, cap);
        memcpy(dst, v.data(), (size_t)(m * per) * sizeof(T));
    }
}

void fill_sink(const Result &R, gpet_sink *s)
{
    if (!s) return;
    copy_out(R.singles, s->singles, s->singles_cap, s->n_singles);
    copy_out(R.coinc, s->coinc, s->coinc_cap, s->n_coinc, 2);
    copy_out(R.delayed, s->delayed, s->delayed_cap, s->n_delayed, 2);
    copy_out(R.adder, s->adder, s->adder_cap, s->n_adder);
    s->n_hits = (int64_t)R.hits.size();
    if (s->hit_id && s->hit_f && s->hits_cap > 0) {
        int64_t m = std::min<int64_t>(s->n_hits, s->hits_cap);
        for (int64_t i = 0; i < m; ++i) {
            memcpy(s->hit_id + 5 * i, R.hits[i].id, 20);
            memcpy(s->hit_f + 5 * i, R.hits[i].f, 20);
        }
    }
    copy_out(R.trace, s->trace, s->trace_cap, s->n_trace);
    copy_out(R.psf[0], s->psf_source, s->psf_source_cap, s->n_psf_source);
    copy_out(R.psf[1], s->psf_phantom, s->psf_phantom_cap, s->n_psf_phantom);
    if (s->job_counts && !R.job_counts.empty())
        memcpy(s->job_counts, R.job_counts.data(), sizeof(int64_t) * R.job_counts.size());
}

gpet_status check_ready(gpet_ctx *c)
{
    if (!c) return GPET_ERR_ARG;
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables not set");
    if (!c->has_dig) return fail(c, GPET_ERR_STATE, "digitizer not set");
    return GPET_OK;
}

} // namespace

/* ======================= C ABI (gpet_cpu_*) ============================== */
extern "C" {

gpet_status gpet_cpu_create(gpet_ctx **out, const gpet_options *opt)
{
    if (!out || !opt) return GPET_ERR_ARG;
    gpet_ctx *c = new gpet_ctx();
    c->opt = *opt;
    memset(&c->dig, 0, sizeof(c->dig));
    *out = c;
    return GPET_OK;
}

gpet_status gpet_cpu_destroy(gpet_ctx *c)
{
    delete c;
    return GPET_OK;
}

const char *gpet_cpu_last_error(gpet_ctx *c) { return c ? c->err.c_str() : "null context"; }

gpet_status gpet_cpu_set_threads(gpet_ctx *c, int32_t n)
{
    if (!c) return GPET_ERR_ARG;
    c->threads = n;
    return GPET_OK;
}

gpet_status gpet_cpu_set_tables(gpet_ctx *c, const gpet_tables *t)
{
    if (!c || !t) return GPET_ERR_ARG;
    if (t->nmat <= 0 || t->nmat > GPET_MAXMAT || t->n1d < 2 || t->ncp < 2 || t->ne < 2)
        return fail(c, GPET_ERR_FORMAT, "table dimensions");
    Tables &T = c->tab;
    T.nmat = t->nmat; T.eminph = t->eminph; T.emax = t->emax; T.n1d = t->n1d;
    for (int k = 0; k < 4; ++k) {
        T.e0[k] = t->e0[k]; T.idl[k] = t->idl[k];
        T.mu[k].assign(t->mu[k], t->mu[k] + (size_t)t->nmat * t->n1d);
    }
    T.ncp = t->ncp; T.ne = t->ne;
    for (int k = 0; k < 2; ++k) {
        T.idcp[k] = t->idcp[k]; T.ide[k] = t->ide[k];
        T.surf[k].assign(t->surf[k], t->surf[k] + (size_t)t->nmat * t->ncp * t->ne);
    }
    c->has_tab = true;
    return GPET_OK;
}

gpet_status gpet_cpu_set_phantom_slot(gpet_ctx *c, int32_t slot, const int32_t *mat, const float *dens,
                                      const int32_t dims[3], const float off[3], const float size[3])
{
    if (!c || !mat || !dens || !dims || !off || !size) return GPET_ERR_ARG;
    if (slot < 0 || slot >= GPET_MAXPHANTOMS) return fail(c, GPET_ERR_ARG, "phantom slot out of range");
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables must be set before the phantom");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return fail(c, GPET_ERR_ARG, "phantom dims");
    if ((int)c->slots.size() <= slot) {
        c->slots.resize((size_t)slot + 1);
        c->slot_set.resize((size_t)slot + 1, 0);
    }
    Phantom p;
    p.nx = dims[0]; p.ny = dims[1]; p.nz = dims[2];
    p.offx = off[0]; p.offy = off[1]; p.offz = off[2];
    p.dx = size[0] / (float)p.nx; p.dy = size[1] / (float)p.ny; p.dz = size[2] / (float)p.nz;
    p.idx = 1.0f / p.dx; p.idy = 1.0f / p.dy; p.idz = 1.0f / p.dz;
    size_t nv = (size_t)p.nx * p.ny * p.nz;
    p.mat.assign(mat, mat + nv);
    p.dens.assign(dens, dens + nv);
    float rmax[GPET_MAXMAT] = {0};
    for (size_t i = 0; i < nv; ++i) {
        int m = p.mat[i];
        if (m < 0 || m >= c->tab.nmat) return fail(c, GPET_ERR_ARG, "phantom material out of range");
        if (p.dens[i] > rmax[m]) rmax[m] = p.dens[i];
    }
    build_wck(c->tab, rmax, p.wck);
    (void)gp_build_walkmap(p.nx, p.ny, p.nz, p.dens.data(), p.wmap);
    p.wbx = gp_walkmap_bricks(p.nx);
    p.wby = gp_walkmap_bricks(p.ny);
    c->slots[slot] = std::move(p);
    c->slot_set[slot] = 1;
    if (slot == 0) c->has_ph = true;
    return GPET_OK;
}

gpet_status gpet_cpu_set_phantom(gpet_ctx *c, const int32_t *mat, const float *dens,
                                 const int32_t dims[3], const float off[3], const float size[3])
{
    return gpet_cpu_set_phantom_slot(c, 0, mat, dens, dims, off, size);
}

gpet_status gpet_cpu_set_detector(gpet_ctx *c, const gpet_panel *panels, int32_t n,
                                  const gpet_panel_material mats[2], const float *quad, int32_t nq)
{
    if (!c || !panels || n <= 0 || !mats) return GPET_ERR_ARG;
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables must be set before the detector");
    if (nq < 0 || nq > GPET_MAXSURFACE) return fail(c, GPET_ERR_CAPACITY, "too many surfaces");
    Detector &d = c->det;
    d.p.assign(panels, panels + n);
    d.mats[0] = mats[0]; d.mats[1] = mats[1];
    if (quad && nq > 0) d.quad.assign(quad, quad + 10 * nq);
    else d.quad.clear();
    d.nquad = (int)d.quad.size() / 10;
    const gpet_panel &p0 = panels[0];
    int mny = (int)std::floor(p0.leny / (p0.mody + p0.mspy)) + 1;
    int cny = (int)std::floor(p0.mody / (p0.cry + p0.cspy)) + 1;
    d.moduleNy = mny; d.crystalNy = cny;
    d.moduleN = mny * ((int)std::floor(p0.lenz / (p0.modz + p0.mspz)) + 1);
    d.crystalN = cny * ((int)std::floor(p0.modz / (p0.crz + p0.cspz)) + 1);
    float rmax[GPET_MAXMAT] = {0};
    for (int i = 0; i < 2; ++i) {
        int m = mats[i].material;
        if (m < 0 || m >= c->tab.nmat) return fail(c, GPET_ERR_ARG, "panel material out of range");
        if (mats[i].density > rmax[m]) rmax[m] = mats[i].density;
    }
    build_wck(c->tab, rmax, d.wck);
    c->has_det = true;
    return GPET_OK;
}

gpet_status gpet_cpu_set_digitizer(gpet_ctx *c, const gpet_digitizer *d)
{
    if (!c || !d) return GPET_ERR_ARG;
    c->dig = *d;
    c->has_dig = true;
    return GPET_OK;
}

gpet_status gpet_cpu_run_source(gpet_ctx *c, const gpet_source *src, int32_t ns, const gpet_isotope *iso,
                                int32_t niso, double t0, double t1, int64_t n_decays, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!src || ns <= 0 || !iso || niso <= 0 || !(t1 >= t0)) return fail(c, GPET_ERR_ARG, "source arguments");
    SourcePlan pl;
    st = plan_job(*c, src, ns, iso, niso, t0, t1, n_decays, 0, 0, pl);
    if (st) return fail(c, st, "bad source/isotope");
    if (pl.total > GPET_MAX_HISTORIES) return fail(c, GPET_ERR_CAPACITY, "more than 2^30 histories in one run");
    const gpet_ctx &C = *c;
    const std::vector<NoiseWin> nw(1, NoiseWin{t0, t1, 0});
    run_histories(*c, pl.total,
                  [&](int64_t h, gpet_trace *tr, PhotonOut *out, Counters &cnt, bool wh) {
                      decay_history(C, pl, iso, h, tr, out, cnt, wh);
                  },
                  sink, c->last, nullptr, &nw);
    fill_sink(c->last, sink);
    return GPET_OK;
}

gpet_status gpet_cpu_run_batch(gpet_ctx *c, const gpet_job *jobs, int32_t njobs, const gpet_source *src,
                               int32_t nsrc, const gpet_isotope *iso, int32_t niso, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!jobs || njobs <= 0 || !src || nsrc <= 0 || !iso || niso <= 0) return fail(c, GPET_ERR_ARG, "batch arguments");
    SourcePlan pl;
    pl.job_first.assign(1, 0);
    for (int j = 0; j < njobs; ++j) {
        const gpet_job &J = jobs[j];
        if (J.first_source < 0 || J.n_sources <= 0 || J.first_source + J.n_sources > nsrc || !(J.t1 >= J.t0))
            return fail(c, GPET_ERR_ARG, "job source range or time window");
        if (J.phantom < 0 || J.phantom >= (int)c->slots.size() || !c->slot_set[J.phantom])
            return fail(c, GPET_ERR_STATE, "job phantom slot not set");
        st = plan_job(*c, src + J.first_source, J.n_sources, iso, niso, J.t0, J.t1, J.n_decays, J.first_source,
                      J.phantom, pl);
        if (st) return fail(c, st, "bad source/isotope");
        pl.job_first.push_back(pl.total);
    }
    if (pl.total > GPET_MAX_HISTORIES) return fail(c, GPET_ERR_CAPACITY, "more than 2^30 histories in one run");
    const gpet_ctx &C = *c;
    std::vector<NoiseWin> nw;
    for (int j = 0; j < njobs; ++j) nw.push_back(NoiseWin{jobs[j].t0, jobs[j].t1, pl.job_first[j]});
    run_histories(*c, pl.total,
                  [&](int64_t h, gpet_trace *tr, PhotonOut *out, Counters &cnt, bool wh) {
                      decay_history(C, pl, iso, h, tr, out, cnt, wh);
                  },
                  sink, c->last, &pl.job_first, &nw);
    fill_sink(c->last, sink);
    return GPET_OK;
}

/* The CPU build has no launch shape; these keep the ABI symmetric. */
gpet_status gpet_cpu_set_launch(gpet_ctx *c, const gpet_launch *l)
{
    if (!c || !l) return GPET_ERR_ARG;
    c->launch = *l;
    return GPET_OK;
}

gpet_status gpet_cpu_get_launch(gpet_ctx *c, gpet_launch *l)
{
    if (!c || !l) return GPET_ERR_ARG;
    *l = c->launch;
    return GPET_OK;
}

gpet_status gpet_cpu_run_psf(gpet_ctx *c, const double *rec8, int64_t n, int32_t ptype, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!rec8 || n < 0 || (ptype != 0 && ptype != 1)) return fail(c, GPET_ERR_ARG, "psf arguments");
    if (c->dig.noise_rate > 0.0)
        return fail(c, GPET_ERR_UNSUPPORTED, "background singles need a decay run's acquisition window");
    const gpet_ctx &C = *c;
    run_histories(*c, n,
                  [&](int64_t h, gpet_trace *tr, PhotonOut *out, Counters &cnt, bool wh) {
                      psf_history(C, rec8 + 8 * h, ptype, h, tr, out, cnt, wh);
                  },
                  sink, c->last);
    fill_sink(c->last, sink);
    return GPET_OK;
}

gpet_status gpet_cpu_get_tallies(gpet_ctx *c, gpet_tallies *t)
{
    if (!c || !t) return GPET_ERR_ARG;
    const Result &R = c->last;
    t->n_histories = R.n_histories;
    t->n_photons = R.c.photons;
    t->n_panel_entries = R.c.panel_entries;
    t->n_hits = R.c.hits;
    t->n_adder = (int64_t)R.adder.size();
    t->n_threshold = R.n_threshold;
    t->n_deadtime = R.n_deadtime;
    t->n_singles = (int64_t)R.singles.size();
    t->n_coinc = (int64_t)R.coinc.size() / 2;
    t->n_coinc_true = R.n_coinc_true;
    t->dropped_events = R.c.dropped_events;
    t->dropped_hits = 0;
    t->majorant_violations = R.c.majorant;
    t->phantom_steps = R.c.phantom_steps;
    t->range_steps = R.c.range_steps;
    t->range_jumps = R.c.range_jumps;
    t->panel_steps = R.c.panel_steps;
    t->seconds_transport = R.sec_transport;
    t->seconds_digitizer = R.sec_digi;
    t->n_noise = R.n_noise;
    t->n_delayed = (int64_t)R.delayed.size() / 2;
    const Detector &d = c->det;
    if (t->crystal_singles && t->n_crystals > 0) {
        memset(t->crystal_singles, 0, sizeof(int64_t) * t->n_crystals);
        for (auto &e : R.singles) {
            int64_t g = (int64_t)e.pann * d.moduleN * d.crystalN + (int64_t)e.modn * d.crystalN + e.cryn;
            if (g >= 0 && g < t->n_crystals) t->crystal_singles[g]++;
        }
    }
    for (int k = 0; k < 2; ++k) {
        int64_t *dst = k == 0 ? t->sinogram : t->sinogram_delayed;
        const std::vector<gpet_event> &pairs = k == 0 ? R.coinc : R.delayed;
        if (!(dst && t->n_s > 0 && t->n_phi > 0 && t->s_max > 0.0f && c->has_det)) continue;
        memset(dst, 0, sizeof(int64_t) * (size_t)t->n_s * t->n_phi);
        for (size_t i = 0; i + 1 < pairs.size(); i += 2) {
            int b = sino_bin(d, pairs[i], pairs[i + 1], t->n_s, t->n_phi, t->s_max);
            if (b >= 0) dst[b]++;
        }
    }
    if (t->energy_hist && t->n_ebins > 0 && t->e_hi > t->e_lo) {
        memset(t->energy_hist, 0, sizeof(int64_t) * t->n_ebins);
        float inv = (float)t->n_ebins / (t->e_hi - t->e_lo);
        for (auto &e : R.singles) {
            float x = (e.E - t->e_lo) * inv;
            if (x >= 0.0f && x < (float)t->n_ebins) t->energy_hist[(int)x]++;
        }
    }
    return GPET_OK;
}

/* Majorant table of the last set_phantom (which=0) or set_detector (1),
 * for the invariant test. */
gpet_status gpet_cpu_get_majorant(gpet_ctx *c, int32_t which, float *out, float *e0, float *idl)
{
    if (!c || !out) return GPET_ERR_ARG;
    if (which < 0 || (which >= 2 && which - 2 >= (int)c->slots.size())) return GPET_ERR_ARG;
    const Wck &w = which == 1 ? c->det.wck : c->slots[which == 0 ? 0 : which - 2].wck;
    if (w.lam.empty()) return GPET_ERR_STATE;
    memcpy(out, w.lam.data(), sizeof(float) * GPET_NWCK);
    if (e0) *e0 = w.e0;
    if (idl) *idl = w.idl;
    return GPET_OK;
}

#ifndef GPET_ORACLE_FP64
/* Single-function probes used by the CPU physics tests. */
float gpet_cpu_lerp_mu(gpet_ctx *c, int32_t table, int32_t mat, float E)
{
    return mu_tab(c->tab, table, mat, E);
}

float gpet_cpu_surface(gpet_ctx *c, int32_t which, int32_t mat, float E, float u)
{
    return surf_read(c->tab, which, mat, E, u);
}

/* The same read for n cumulative probabilities u[i] at one energy. */
void gpet_cpu_surface_n(gpet_ctx *c, int32_t which, int32_t mat, float E, const float *u, int64_t n, float *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = surf_read(c->tab, which, mat, E, u[i]);
}

void gpet_cpu_sample_kn(float E, uint64_t seed, int64_t n, float *efrac, float *costh)
{
    gp_rng r;
    gp_rng_init(&r, seed, 0, 9u, 0u);
    for (int64_t i = 0; i < n; ++i) kn_sample(E, r, efrac[i], costh[i]);
}

void gpet_cpu_rotate(float *uvw, float costh, float uphi)
{
    rotate_dir(uvw[0], uvw[1], uvw[2], costh, uphi);
}

void gpet_cpu_sample_ek(const gpet_isotope *iso, uint64_t seed, int64_t n, float *out)
{
    gp_rng r;
    gp_rng_init(&r, seed, 0, 9u, 0u);
    for (int64_t i = 0; i < n; ++i) out[i] = sample_ek(*iso, r);
}

void gpet_cpu_sample_shape(const gpet_source *s, uint64_t seed, int64_t n, float *xyz)
{
    gp_rng r;
    gp_rng_init(&r, seed, 0, 9u, 0u);
    for (int64_t i = 0; i < n; ++i) sample_shape(*s, r, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

void gpet_cpu_emit_pairs(float sigma, uint64_t seed, int64_t n, float *v6, float *e2)
{
    gp_rng r;
    gp_rng_init(&r, seed, 0, 9u, 0u);
    for (int64_t i = 0; i < n; ++i) {
        float v[2][3], e[2];
        emit_pair(sigma, r, v, e);
        for (int k = 0; k < 2; ++k) {
            v6[6 * i + 3 * k] = v[k][0]; v6[6 * i + 3 * k + 1] = v[k][1]; v6[6 * i + 3 * k + 2] = v[k][2];
            e2[2 * i + k] = e[k];
        }
    }
}

/* Positron range displacement for n positrons of kinetic energy ek starting
 * at xyz0 (needs a phantom). */
gpet_status gpet_cpu_positron_range(gpet_ctx *c, const float *xyz0, float ek, uint64_t seed, int64_t n, float *xyz)
{
    if (!c || !c->has_ph) return GPET_ERR_STATE;
    gp_rng r;
    gp_rng_init(&r, seed, 0, 9u, 0u);
    Counters cnt;
    for (int64_t i = 0; i < n; ++i) {
        float x = xyz0[0], y = xyz0[1], z = xyz0[2];
        positron_range(c->slots[0], x, y, z, ek, 0, 0, 0, 0, r, cnt);
        xyz[3 * i] = x; xyz[3 * i + 1] = y; xyz[3 * i + 2] = z;
    }
    return GPET_OK;
}

/* The positron walk alone for n given ranges rr[i] (g/cm^2) and unit
 * directions dir[3i..3i+2] from xyz0[3i..3i+2] (needs a phantom); returns
 * the end points and, if steps is non-NULL, the voxel crossings inside the
 * phantom per walk. */
gpet_status gpet_cpu_range_walk(gpet_ctx *c, const float *xyz0, const float *dir, const float *rr, int64_t n,
                                float *xyz, int64_t *steps)
{
    if (!c || !c->has_ph) return GPET_ERR_STATE;
    for (int64_t i = 0; i < n; ++i) {
        Counters cnt;
        float x = xyz0[3 * i], y = xyz0[3 * i + 1], z = xyz0[3 * i + 2];
        if (rr[i] > 0.0f)
            range_walk(c->slots[0], x, y, z, dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], rr[i], cnt);
        xyz[3 * i] = x; xyz[3 * i + 1] = y; xyz[3 * i + 2] = z;
        if (steps) steps[i] = cnt.range_steps;
    }
    return GPET_OK;
}

/* Binomial count probe: the draw of source `src` of a run with this seed. */
int64_t gpet_cpu_binomial(double n, double p, uint64_t seed, int32_t src)
{
    return gp_binomial_draw(n, p, seed, src);
}

/* Crystal index probe: returns m_id and writes module/crystal ids. */
int32_t gpet_cpu_crystal_search(gpet_ctx *c, int32_t panel, float x, float y, float z, int32_t *M_id, int32_t *L_id)
{
    int M = -1, L = -1;
    int m = crystal_search(c->det, c->det.p[panel], x, y, z, M, L);
    *M_id = M; *L_id = L;
    return m;
}

#endif /* GPET_ORACLE_FP64 */

/* Digitizer stages on a caller list (already adder/readout output). */
gpet_status gpet_cpu_digitize(gpet_ctx *c, const gpet_event *ev, int64_t n, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    c->last = Result();
    c->last.adder.assign(ev, ev + n);
    std::sort(c->last.adder.begin(), c->last.adder.end(), ev_less_id);
    run_digitizer(*c, c->last);
    fill_sink(c->last, sink);
    return GPET_OK;
}

} /* extern "C" */

/* ---- probes of the shared RNG / math headers (tests only) -------------- */
extern "C" {

void gpet_cpu_philox(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4)
{
    gp_u4 c = {ctr4[0], ctr4[1], ctr4[2], ctr4[3]};
    gp_u4 r = gp_philox4x32_10(c, key2[0], key2[1]);
    out4[0] = r.x; out4[1] = r.y; out4[2] = r.z; out4[3] = r.w;
}

/* kind: 0 logf, 1 expf, 2 sin2pi, 3 cos2pi, 4 sinf, 5 cosf, 6 cbrtf */
void gpet_cpu_mathf(int32_t kind, const float *in, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        float x = in[i], s, c;
        switch (kind) {
        case 0: out[i] = gp_logf(x); break;
        case 1: out[i] = gp_expf(x); break;
        case 2: gp_sincos2pi(x, &s, &c); out[i] = s; break;
        case 3: gp_sincos2pi(x, &s, &c); out[i] = c; break;
        case 4: gp_sincosf(x, &s, &c); out[i] = s; break;
        case 5: gp_sincosf(x, &s, &c); out[i] = c; break;
        case 6: out[i] = gp_cbrtf(x); break;
        case 7: out[i] = gp_atan2f(x, in[n + i]); break;
        default: out[i] = 0.0f;
        }
    }
}

/* kind: 0 log, 1 log1p */
void gpet_cpu_mathd(int32_t kind, const double *in, double *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = kind == 0 ? gp_log_d(in[i]) : gp_log1p_d(in[i]);
}

/* Draws of one stream: uniform floats, normals or 53-bit doubles. */
void gpet_cpu_stream(uint64_t seed, uint64_t id, uint32_t aux, int32_t kind, int64_t n, void *out)
{
    gp_rng r;
    gp_rng_init(&r, seed, id, aux, 0u);
    for (int64_t i = 0; i < n; ++i) {
        if (kind == 0) ((float *)out)[i] = gp_rng_uniform(&r);
        else if (kind == 1) ((float *)out)[i] = gp_rng_normal(&r);
        else if (kind == 2) ((double *)out)[i] = gp_rng_uniform_d(&r);
        else ((uint32_t *)out)[i] = gp_rng_u32(&r);
    }
}

} /* extern "C" */

/* ======================= surrogate network (CPU restatement) =============
 * The detector-response MLP of the HIP library restated in FP64 for the
 * gradient check: 6 -> 256 -> 256 -> 256 -> 256 -> 190, SiLU, weights
 * [outputs x (inputs + 8)] with the bias in column `inputs`, head stored
 * with 192 rows.  By default it rounds what the GPU path stores in BF16
 * (weights, inputs, activations, dL/dh, dL/dz) to BF16 and accumulates in
 * FP64; exact mode (gpet_cpu_surrogate_set_exact) rounds nothing, for
 * finite-difference checks.  Loss: BCE on the event logit plus, for events,
 * the softmax cross-entropy of the deposited-energy class (h[1 .. 61]: full
 * absorption for fE >= 0.999, else 48 bins of fE over [0, edge) and 12 of
 * [edge, 0.999), edge = 2a / (1 + 2a) the Compton edge of an entry of
 * a = E / 511 keV, the first feature) and the negative log-likelihood of
 * the 16-component diagonal Gaussian mixture over (depth, dy, dz), whose
 * weight logits are h[62 + k] for a full absorption and h[78 + k] otherwise
 * (means h[94 + 3k + d], log sigmas h[142 + 3k + d] clamped to [-5, 2]),
 * averaged over rows. */
namespace {

constexpr int M_NL = 5, M_HID = 256, M_NH = 190, M_K = 16, M_D = 3, M_NX = 6, M_NY = 5;
constexpr int M_NLO = 48, M_NHI = 12, M_NC = 1 + M_NLO + M_NHI;
constexpr int M_CLS = 1, M_WPP = 62, M_WCT = 78, M_MU = 94, M_LS = 142;
constexpr float M_PP_MIN = 0.999f;  /* fE >= this: full absorption                */
constexpr int M_LN[M_NL] = {256, 256, 256, 256, 192};
constexpr int M_LK[M_NL] = {6, 256, 256, 256, 256};
constexpr double M_LSMIN = -5.0, M_LSMAX = 2.0;

int m_kp(int l) { return l == 0 ? 8 : M_LK[l] + 8; }
int64_t m_off(int l)
{
    int64_t o = 0;
    for (int i = 0; i < l; ++i) o += (int64_t)M_LN[i] * m_kp(i);
    return o;
}

double bfr(double x, int exact)
{
    if (exact) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

double m_sig(double z) { return 1.0 / (1.0 + std::exp(-z)); }

struct MlpRow {
    std::vector<double> a[M_NL];   /* inputs of layer l (length kp(l)) */
    std::vector<double> z[4];      /* pre-activations of hidden layers  */
    double h[M_NH];
};

void mlp_forward_row(const std::vector<double> &W, const float *x, int exact, MlpRow &R)
{
    R.a[0].assign(8, 0.0);
    for (int k = 0; k < M_NX; ++k) R.a[0][k] = bfr(x[k], exact);
    R.a[0][M_NX] = 1.0;
    for (int l = 0; l < 4; ++l) {
        const int kp = m_kp(l);
        R.z[l].assign(M_HID, 0.0);
        R.a[l + 1].assign(M_HID + 8, 0.0);
        for (int n = 0; n < M_HID; ++n) {
            double s = 0.0;
            for (int k = 0; k < kp; ++k) s += W[(size_t)(m_off(l) + (int64_t)n * kp + k)] * R.a[l][k];
            R.z[l][n] = s;
            R.a[l + 1][n] = bfr(s * m_sig(s), exact);
        }
        R.a[l + 1][M_HID] = 1.0;
    }
    for (int n = 0; n < M_NH; ++n) {
        double s = 0.0;
        for (int k = 0; k < m_kp(4); ++k) s += W[(size_t)(m_off(4) + (int64_t)n * m_kp(4) + k)] * R.a[4][k];
        R.h[n] = s;
    }
}

double softplus_d(double x) { return x > 0 ? x + std::log1p(std::exp(-x)) : std::log1p(std::exp(x)); }

/* Energy class of a deposited fraction fe for an entry of e511 = E / 511
 * keV (float arithmetic, as the library computes it). */
int m_class(float e511, float fe)
{
    if (fe >= M_PP_MIN) return 0;
    const float a2 = 2.0f * e511;
    const float ed = a2 / (1.0f + a2);
    if (fe < ed) return 1 + std::min(std::max((int)(fe / ed * (float)M_NLO), 0), M_NLO - 1);
    return 1 + M_NLO + std::min(std::max((int)((fe - ed) / (M_PP_MIN - ed) * (float)M_NHI), 0), M_NHI - 1);
}

/* loss of one row and dL/dh (unscaled) */
double mlp_loss_row(const double *h, const float *x, const float *y, double *g)
{
    const double ev = y[0];
    double loss = ev > 0.5 ? softplus_d(-h[0]) : softplus_d(h[0]);
    for (int c = 0; c < M_NH; ++c) g[c] = 0.0;
    g[0] = m_sig(h[0]) - ev;
    if (ev > 0.5) {
        const int cls = m_class(x[0], y[1]);
        double mc = -1e300;
        for (int c = 0; c < M_NC; ++c) mc = std::max(mc, h[M_CLS + c]);
        double sc = 0.0;
        for (int c = 0; c < M_NC; ++c) sc += std::exp(h[M_CLS + c] - mc);
        const double lse_c = mc + std::log(sc);
        loss += lse_c - h[M_CLS + cls];
        for (int c = 0; c < M_NC; ++c) g[M_CLS + c] = std::exp(h[M_CLS + c] - lse_c) - (c == cls ? 1.0 : 0.0);
        const int wo = cls == 0 ? M_WPP : M_WCT;
        double mw = -1e300;
        for (int k = 0; k < M_K; ++k) mw = std::max(mw, h[wo + k]);
        double sw = 0.0;
        for (int k = 0; k < M_K; ++k) sw += std::exp(h[wo + k] - mw);
        const double lse_w = mw + std::log(sw);
        double lp[M_K], mx = -1e300;
        for (int k = 0; k < M_K; ++k) {
            double s = h[wo + k] - lse_w;
            for (int d = 0; d < M_D; ++d) {
                double ls = std::min(std::max(h[M_LS + 3 * k + d], M_LSMIN), M_LSMAX);
                double zz = (y[2 + d] - h[M_MU + 3 * k + d]) * std::exp(-ls);
                s += -0.5 * zz * zz - ls - 0.91893853320467274;
            }
            lp[k] = s;
            mx = std::max(mx, s);
        }
        double se = 0.0;
        for (int k = 0; k < M_K; ++k) se += std::exp(lp[k] - mx);
        const double lse = mx + std::log(se);
        loss += -lse;
        for (int k = 0; k < M_K; ++k) {
            const double gam = std::exp(lp[k] - lse), pik = std::exp(h[wo + k] - lse_w);
            g[wo + k] = pik - gam;
            for (int d = 0; d < M_D; ++d) {
                const double raw = h[M_LS + 3 * k + d];
                const double ls = std::min(std::max(raw, M_LSMIN), M_LSMAX);
                const double is = std::exp(-ls);
                const double zz = (y[2 + d] - h[M_MU + 3 * k + d]) * is;
                g[M_MU + 3 * k + d] = -gam * zz * is;
                g[M_LS + 3 * k + d] = (raw > M_LSMIN && raw < M_LSMAX) ? gam * (1.0 - zz * zz) : 0.0;
            }
        }
    }
    return loss;
}

std::vector<double> mlp_weights(const gpet_ctx &c)
{
    std::vector<double> W(c.mlp_w.size());
    for (size_t i = 0; i < W.size(); ++i) W[i] = bfr(c.mlp_w[i], c.mlp_exact);
    return W;
}

} // namespace

extern "C" {

gpet_status gpet_cpu_surrogate_set_params(gpet_ctx *c, const float *w, int64_t n)
{
    if (!c || !w || n != m_off(M_NL)) return GPET_ERR_ARG;
    c->mlp_w.assign(w, w + n);
    return GPET_OK;
}

gpet_status gpet_cpu_surrogate_get_params(gpet_ctx *c, float *out, int64_t cap, int64_t *n)
{
    if (!c) return GPET_ERR_ARG;
    if (n) *n = m_off(M_NL);
    if (!out) return GPET_OK;
    if (c->mlp_w.empty()) return GPET_ERR_STATE;
    if (cap < m_off(M_NL)) return GPET_ERR_CAPACITY;
    memcpy(out, c->mlp_w.data(), sizeof(float) * c->mlp_w.size());
    return GPET_OK;
}

gpet_status gpet_cpu_surrogate_set_exact(gpet_ctx *c, int32_t exact)
{
    if (!c) return GPET_ERR_ARG;
    c->mlp_exact = exact != 0;
    return GPET_OK;
}

gpet_status gpet_cpu_surrogate_forward(gpet_ctx *c, const float *x, int64_t n, float *h)
{
    if (!c || !x || !h || n <= 0) return GPET_ERR_ARG;
    if (c->mlp_w.empty()) return fail(c, GPET_ERR_STATE, "no surrogate parameters");
    const std::vector<double> W = mlp_weights(*c);
    for (int64_t i = 0; i < n; ++i) {
        MlpRow R;
        mlp_forward_row(W, x + i * M_NX, c->mlp_exact, R);
        for (int k = 0; k < M_NH; ++k) h[i * M_NH + k] = (float)R.h[k];
    }
    return GPET_OK;
}

/* Mean loss and its gradient w.r.t. the flat parameter vector. */
gpet_status gpet_cpu_surrogate_grad_d(gpet_ctx *c, const float *x, const float *y, int64_t n, double *loss,
                                      double *grad)
{
    if (!c || !x || !y || !grad || n <= 0) return GPET_ERR_ARG;
    if (c->mlp_w.empty()) return fail(c, GPET_ERR_STATE, "no surrogate parameters");
    const int ex = c->mlp_exact;
    const std::vector<double> W = mlp_weights(*c);
    const int64_t np = m_off(M_NL);
    for (int64_t i = 0; i < np; ++i) grad[i] = 0.0;
    double L = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        MlpRow R;
        mlp_forward_row(W, x + i * M_NX, ex, R);
        double gh[M_NH];
        L += mlp_loss_row(R.h, x + i * M_NX, y + i * M_NY, gh);
        std::vector<double> dn(M_HID + 8, 0.0);
        for (int k = 0; k < M_NH; ++k) dn[k] = bfr(gh[k] / (double)n, ex);
        int nout = M_NH;
        for (int l = 4; l >= 0; --l) {
            const int kp = m_kp(l);
            const int64_t o = m_off(l);
            for (int r = 0; r < nout; ++r)
                for (int k = 0; k < kp; ++k) grad[o + (int64_t)r * kp + k] += dn[r] * R.a[l][k];
            if (l == 0) break;
            std::vector<double> dz(M_HID, 0.0);
            for (int k = 0; k < M_HID; ++k) {
                double s = 0.0;
                for (int r = 0; r < nout; ++r) s += dn[r] * W[(size_t)(o + (int64_t)r * kp + k)];
                const double zz = R.z[l - 1][k], sg = m_sig(zz);
                dz[k] = bfr(s * (sg * (1.0 + zz * (1.0 - sg))), ex);
            }
            dn.assign(dz.begin(), dz.end());
            nout = M_HID;
        }
    }
    if (loss) *loss = L / (double)n;
    return GPET_OK;
}

/* Same as gpet_surrogate_grad (float outputs). */
gpet_status gpet_cpu_surrogate_grad(gpet_ctx *c, const float *x, const float *y, int64_t n, float *loss, float *grad)
{
    if (!grad) return GPET_ERR_ARG;
    std::vector<double> g((size_t)m_off(M_NL));
    double l = 0;
    gpet_status st = gpet_cpu_surrogate_grad_d(c, x, y, n, &l, g.data());
    if (st) return st;
    for (size_t i = 0; i < g.size(); ++i) grad[i] = (float)g[i];
    if (loss) *loss = (float)l;
    return GPET_OK;
}

/* The CPU build does not harvest, train or run the network in a transport
 * run; these keep the ABI symmetric. */
gpet_status gpet_cpu_surrogate_harvest(gpet_ctx *c, const gpet_source *, int32_t, const gpet_isotope *, int32_t, double,
                                       double, int64_t, int64_t)
{
    return fail(c, GPET_ERR_UNSUPPORTED, "the CPU oracle does not harvest");
}
gpet_status gpet_cpu_surrogate_train(gpet_ctx *c, const gpet_surrogate_config *, gpet_surrogate_report *)
{
    return fail(c, GPET_ERR_UNSUPPORTED, "the CPU oracle does not train");
}
gpet_status gpet_cpu_surrogate_enable(gpet_ctx *c, int32_t on)
{
    return on ? fail(c, GPET_ERR_UNSUPPORTED, "the CPU oracle runs plain transport only") : GPET_OK;
}
gpet_status gpet_cpu_surrogate_save(gpet_ctx *c, const char *)
{
    return fail(c, GPET_ERR_UNSUPPORTED, "the CPU oracle does not save networks");
}
gpet_status gpet_cpu_surrogate_records(gpet_ctx *c, float *, int64_t, int64_t *n)
{
    if (n) *n = 0;
    return c ? GPET_OK : GPET_ERR_ARG;
}
gpet_status gpet_cpu_surrogate_load(gpet_ctx *c, const char *)
{
    return fail(c, GPET_ERR_UNSUPPORTED, "the CPU oracle does not load networks");
}

} /* extern "C" */

