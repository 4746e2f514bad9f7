/*
 * Host sanitizer driver (AddressSanitizer + UndefinedBehaviorSanitizer).
 *
 * Built by tests/sanitize/Makefile from the parsers and writers
 * (gpet_amd/csrc/host/io.cpp) and the CPU oracle (oracle/gpet_oracle.cpp)
 * with -fsanitize=address,undefined; tests/test_sanitize.py runs it.  It
 *   1. parses the fixture files (input_PET.in, config8.geo, source.txt,
 *      pointsource.txt, isotopes.txt, the cross-section tables) and every
 *      prefix cut of each text fixture at line and mid-line boundaries, so
 *      that the positional readers meet truncated input;
 *   2. runs the oracle's whole path (decays with positron range, phantom and
 *      panel transport with hits, digitizer, sorter, background singles,
 *      delayed window, PSF runs and a batched run) on a small cylinder
 *      phantom in the 8-panel scanner of config8.geo;
 *   3. writes every output format into a temporary directory.
 * Any sanitizer report aborts with a non-zero status.
 *
 * usage: driver <repo root> <scratch dir>
 */
#include "../../include/gpet.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

extern "C" {
gpet_status gpet_cpu_create(gpet_ctx **ctx, const gpet_options *opt);
gpet_status gpet_cpu_destroy(gpet_ctx *ctx);
gpet_status gpet_cpu_set_threads(gpet_ctx *c, int32_t n);
gpet_status gpet_cpu_set_tables(gpet_ctx *ctx, const gpet_tables *t);
gpet_status gpet_cpu_set_phantom(gpet_ctx *ctx, const int32_t *mat, const float *dens, const int32_t dims[3],
                                 const float offset[3], const float size[3]);
gpet_status gpet_cpu_set_phantom_slot(gpet_ctx *ctx, int32_t slot, const int32_t *mat, const float *dens,
                                      const int32_t dims[3], const float offset[3], const float size[3]);
gpet_status gpet_cpu_set_detector(gpet_ctx *ctx, const gpet_panel *panels, int32_t n_panels,
                                  const gpet_panel_material mats[2], const float *quadrics, int32_t n_quadrics);
gpet_status gpet_cpu_set_digitizer(gpet_ctx *ctx, const gpet_digitizer *d);
gpet_status gpet_cpu_run_source(gpet_ctx *ctx, const gpet_source *src, int32_t n_src, const gpet_isotope *iso,
                                int32_t n_iso, double t0, double t1, int64_t n_decays, gpet_sink *sink);
gpet_status gpet_cpu_run_psf(gpet_ctx *ctx, const double *rec8, int64_t n, int32_t ptype, gpet_sink *sink);
gpet_status gpet_cpu_run_batch(gpet_ctx *ctx, const gpet_job *jobs, int32_t n_jobs, const gpet_source *src,
                               int32_t n_src, const gpet_isotope *iso, int32_t n_iso, gpet_sink *sink);
gpet_status gpet_cpu_get_tallies(gpet_ctx *ctx, gpet_tallies *t);
const char *gpet_cpu_last_error(gpet_ctx *ctx);
}

static int g_fail = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++g_fail;                                      \
        }                                                  \
    } while (0)

static std::string slurp(const std::string &p)
{
    std::ifstream f(p, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

static void spit(const std::string &p, const std::string &s)
{
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f << s;
}

/* Cut points of a text: every line end, and the middle of every line. */
static std::vector<size_t> cuts(const std::string &s)
{
    std::vector<size_t> c{0};
    size_t start = 0;
    for (size_t i = 0; i < s.size(); ++i)
        if (s[i] == '\n') {
            c.push_back(start + (i - start) / 2);
            c.push_back(i + 1);
            start = i + 1;
        }
    c.push_back(s.size());
    return c;
}

/* Parse every truncation of a fixture: a status (any) but no sanitizer report. */
template <class F>
static int truncations(const std::string &src, const std::string &tmp, F parse)
{
    const std::string text = slurp(src);
    int n = 0;
    for (size_t k : cuts(text)) {
        spit(tmp, text.substr(0, k));
        parse(tmp.c_str());
        ++n;
    }
    /* a copy with every number replaced by a huge one */
    std::string big = text;
    for (auto &ch : big)
        if (ch >= '0' && ch <= '9') ch = '9';
    spit(tmp, big);
    parse(tmp.c_str());
    return n + 1;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <repo root> <scratch dir>\n", argv[0]);
        return 2;
    }
    const std::string root = argv[1], tmp = argv[2];
    const std::string gold = root + "/tests/golden/";

    /* ---- 1. parsers on the fixtures and their truncations -------------- */
    gpet_input in;
    CHECK(gpet_read_input((gold + "input_PET.in").c_str(), &in) == GPET_OK, "input_PET.in");
    gpet_digitizer dig;
    gpet_input_digitizer(&in, &dig);
    static gpet_panel panels[GPET_MAXPANELS];
    gpet_panel_material mats[2];
    int32_t npan = 0;
    CHECK(gpet_read_geometry((gold + "config8.geo").c_str(), panels, GPET_MAXPANELS, &npan, mats) == GPET_OK &&
              npan == 8, "config8.geo");
    gpet_source src[64];
    int32_t nsrc = 0;
    CHECK(gpet_read_sources((gold + "source.txt").c_str(), src, 64, &nsrc) == GPET_OK && nsrc > 0, "source.txt");
    CHECK(gpet_read_sources((gold + "pointsource.txt").c_str(), src, 64, &nsrc) == GPET_OK && nsrc > 0,
          "pointsource.txt");
    gpet_isotope iso[64];
    int32_t niso = 0;
    CHECK(gpet_read_isotopes((root + "/gpet_amd/data/isotopes.txt").c_str(), iso, 64, &niso) == GPET_OK && niso > 0,
          "isotopes.txt");
    gpet_tables tabs;
    CHECK(gpet_tables_load((root + "/gpet_amd/data/tables").c_str(), &tabs) == GPET_OK, "tables");

    int n_cut = 0;
    const std::string t_in = tmp + "/trunc.txt";
    n_cut += truncations(gold + "input_PET.in", t_in, [](const char *p) {
        gpet_input x;
        (void)gpet_read_input(p, &x);
    });
    n_cut += truncations(gold + "config8.geo", t_in, [](const char *p) {
        static gpet_panel pp[GPET_MAXPANELS];
        gpet_panel_material mm[2];
        int32_t n = 0;
        (void)gpet_read_geometry(p, pp, GPET_MAXPANELS, &n, mm);
        (void)gpet_read_geometry(p, pp, 3, &n, mm);   /* capacity below the repetition count */
    });
    n_cut += truncations(gold + "source.txt", t_in, [](const char *p) {
        gpet_source s[4];
        int32_t n = 0;
        (void)gpet_read_sources(p, s, 4, &n);
        (void)gpet_read_sources(p, s, 1, &n);
    });
    n_cut += truncations(root + "/gpet_amd/data/isotopes.txt", t_in, [](const char *p) {
        gpet_isotope s[4];
        int32_t n = 0;
        (void)gpet_read_isotopes(p, s, 4, &n);
    });
    {
        /* a PSF file whose size is not a whole number of records */
        std::string rec(64 * 3 + 17, '\0');
        spit(t_in, rec);
        double out[8 * 3];
        int64_t n = 0;
        (void)gpet_read_psf(t_in.c_str(), out, 3, &n);
        CHECK(n <= 3, "psf count");
    }
    std::printf("parsers: %d truncated or altered inputs parsed\n", n_cut);

    /* ---- 2. the oracle's path on a small cylinder phantom --------------- */
    gpet_options opt;
    std::memset(&opt, 0, sizeof(opt));
    opt.seed = 99;
    opt.eabs = 1e3f;
    opt.noncol_sigma = 0.0037f;
    opt.positron_range = 1;
    opt.record_hits = 1;
    opt.flags = GPET_FLAG_PSF_SOURCE | GPET_FLAG_PSF_PHANTOM;
    gpet_ctx *c = nullptr;
    CHECK(gpet_cpu_create(&c, &opt) == GPET_OK, "create");
    gpet_cpu_set_threads(c, 2);
    CHECK(gpet_cpu_set_tables(c, &tabs) == GPET_OK, "tables: %s", gpet_cpu_last_error(c));

    const int n = 24;
    const float size = 1.0f, d = size / n;
    std::vector<int32_t> mat((size_t)n * n * n);
    std::vector<float> den(mat.size());
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const float x = (i + 0.5f) * d - 0.5f, y = (j + 0.5f) * d - 0.5f;
                const float r2 = x * x + y * y;
                const size_t v = ((size_t)k * n + j) * n + i;
                mat[v] = r2 <= 0.16f ? (r2 <= 0.01f ? 3 : 1) : 0;
                den[v] = r2 <= 0.16f ? (r2 <= 0.01f ? 1.85f : 1.0f) : 0.0012f;
            }
    const int32_t dims[3] = {n, n, n};
    const float off[3] = {-0.5f, -0.5f, -0.5f}, sz[3] = {size, size, size};
    CHECK(gpet_cpu_set_phantom(c, mat.data(), den.data(), dims, off, sz) == GPET_OK, "phantom");
    const float quad[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
    CHECK(gpet_cpu_set_detector(c, panels, npan, mats, quad, 1) == GPET_OK, "detector: %s", gpet_cpu_last_error(c));
    dig.time_blur_fwhm_us = 300e-6;
    dig.coinc_window_us = 4e-3;
    dig.coinc_policy = GPET_COINC_TAKE_ALL_GOODS;
    dig.noise_rate = 2e4;
    dig.noise_e_mean = 450e3f;
    dig.noise_e_sigma = 50e3f;
    dig.delay_us = 0.05;
    dig.dead_time_us = 0.5f;
    CHECK(gpet_cpu_set_digitizer(c, &dig) == GPET_OK, "digitizer: %s", gpet_cpu_last_error(c));

    gpet_source cyl;
    std::memset(&cyl, 0, sizeof(cyl));
    cyl.natom = 1e9;
    cyl.shape = 1;
    cyl.coef[3] = 0.4f;
    cyl.coef[4] = 1.0f;
    const int64_t nd = 3000;
    std::vector<gpet_event> singles(nd), coinc(2 * nd), adder(4 * nd), delayed(2 * nd);
    std::vector<int32_t> hit_id(5 * 20 * nd);
    std::vector<float> hit_f(5 * 20 * nd);
    std::vector<gpet_trace> trace(2 * nd);
    std::vector<gpet_psf_record> psf_s(2 * nd), psf_p(2 * nd);
    gpet_sink sink;
    std::memset(&sink, 0, sizeof(sink));
    sink.singles = singles.data(); sink.singles_cap = nd;
    sink.coinc = coinc.data(); sink.coinc_cap = nd;
    sink.adder = adder.data(); sink.adder_cap = 4 * nd;
    sink.hit_id = hit_id.data(); sink.hit_f = hit_f.data(); sink.hits_cap = 20 * nd;
    sink.trace = trace.data(); sink.trace_cap = 2 * nd;
    sink.psf_source = psf_s.data(); sink.psf_source_cap = 2 * nd;
    sink.psf_phantom = psf_p.data(); sink.psf_phantom_cap = 2 * nd;
    sink.delayed = delayed.data(); sink.delayed_cap = nd;
    const std::string outdir = tmp + "/out";
    sink.out_dir = outdir.c_str();
    CHECK(gpet_cpu_run_source(c, &cyl, 1, iso, 1, 0.0, 10.0, nd, &sink) == GPET_OK, "run_source: %s",
          gpet_cpu_last_error(c));
    CHECK(sink.n_adder > 0 && sink.n_singles > 0, "no events");
    /* the same run with every capacity too small (counts still reported) */
    gpet_sink small = sink;
    small.singles_cap = 3; small.coinc_cap = 1; small.adder_cap = 5; small.hits_cap = 7; small.trace_cap = 11;
    small.psf_source_cap = 2; small.psf_phantom_cap = 2; small.delayed_cap = 1;
    small.out_dir = nullptr;
    CHECK(gpet_cpu_run_source(c, &cyl, 1, iso, 1, 0.0, 10.0, 500, &small) == GPET_OK, "run_source small caps");
    /* decay counts from the activity (n_decays = 0) */
    gpet_source weak = cyl;
    weak.natom = 2e6;
    gpet_sink none;
    std::memset(&none, 0, sizeof(none));
    CHECK(gpet_cpu_run_source(c, &weak, 1, iso, 1, 0.0, 10.0, 0, &none) == GPET_OK, "binomial run");

    std::vector<int64_t> cs(8 * 400), eh(100), sg(48 * 64), sgd(48 * 64);
    gpet_tallies t;
    std::memset(&t, 0, sizeof(t));
    t.crystal_singles = cs.data(); t.n_crystals = 8 * 400;
    t.energy_hist = eh.data(); t.n_ebins = 100; t.e_lo = 0; t.e_hi = 1e6f;
    t.sinogram = sg.data(); t.n_s = 64; t.n_phi = 48; t.s_max = 8.0f;
    t.sinogram_delayed = sgd.data();
    CHECK(gpet_cpu_get_tallies(c, &t) == GPET_OK, "tallies");

    /* PSF runs: positron and photon records */
    gpet_digitizer dig2 = dig;
    dig2.noise_rate = 0.0;
    CHECK(gpet_cpu_set_digitizer(c, &dig2) == GPET_OK, "digitizer 2");
    std::vector<double> rec(8 * 400);
    for (int i = 0; i < 400; ++i) {
        double *r = &rec[8 * i];
        const double ph = 0.0157 * i, ct = std::cos(0.031 * i), st = std::sqrt(1 - ct * ct);
        r[0] = 0.01 * (i % 7); r[1] = -0.01 * (i % 5); r[2] = 0.02 * (i % 3); r[3] = 0.001 * i;
        r[4] = st * std::cos(ph); r[5] = st * std::sin(ph); r[6] = ct; r[7] = i % 2 ? 511e3 : 3e5;
    }
    gpet_sink ps = sink;
    ps.out_dir = nullptr;
    CHECK(gpet_cpu_run_psf(c, rec.data(), 400, 0, &ps) == GPET_OK, "psf positrons: %s", gpet_cpu_last_error(c));
    CHECK(gpet_cpu_run_psf(c, rec.data(), 400, 1, &ps) == GPET_OK, "psf photons: %s", gpet_cpu_last_error(c));

    /* a batched run over two phantom slots of different size */
    const int n2 = 9;
    std::vector<int32_t> mat2((size_t)n2 * n2 * n2, 1);
    std::vector<float> den2(mat2.size(), 1.0f);
    const int32_t dims2[3] = {n2, n2, n2};
    CHECK(gpet_cpu_set_phantom_slot(c, 1, mat2.data(), den2.data(), dims2, off, sz) == GPET_OK, "slot 1");
    gpet_job jobs[2];
    std::memset(jobs, 0, sizeof(jobs));
    jobs[0].phantom = 0; jobs[0].first_source = 0; jobs[0].n_sources = 1; jobs[0].t1 = 10; jobs[0].n_decays = 700;
    jobs[1] = jobs[0];
    jobs[1].phantom = 1; jobs[1].first_source = 1; jobs[1].n_decays = 900;
    gpet_source two[2] = {cyl, cyl};
    std::vector<int64_t> jc(6);
    gpet_sink bs = sink;
    bs.out_dir = nullptr;
    bs.job_counts = jc.data();
    CHECK(gpet_cpu_run_batch(c, jobs, 2, two, 2, iso, 1, &bs) == GPET_OK, "batch: %s", gpet_cpu_last_error(c));

    /* ---- 3. writers -------------------------------------------------- */
    const int64_t ns = std::min<int64_t>(sink.n_singles, nd), nc = std::min<int64_t>(sink.n_coinc, nd);
    const int64_t nl = std::min<int64_t>(sink.n_delayed, nd);
    CHECK(gpet_write_events((tmp + "/ev.dat").c_str(), singles.data(), ns, 0) == GPET_OK, "write events");
    CHECK(gpet_write_events((tmp + "/ev.dat").c_str(), singles.data(), ns, 1) == GPET_OK, "append events");
    CHECK(gpet_write_hits((tmp + "/hid.dat").c_str(), (tmp + "/hf.dat").c_str(), hit_id.data(), hit_f.data(),
                          std::min<int64_t>(sink.n_hits, 20 * nd), 0) == GPET_OK, "write hits");
    CHECK(gpet_write_psf(tmp.c_str(), "source", psf_s.data(), std::min<int64_t>(sink.n_psf_source, 2 * nd), 0) ==
              GPET_OK, "write psf");
    int32_t mod_n = 0, cry_n = 0, mod_ny = 0, cry_ny = 0;
    CHECK(gpet_panel_counts(&panels[0], &mod_n, &cry_n, &mod_ny, &cry_ny) == GPET_OK, "panel counts");
    std::vector<gpet_lm_record> lm((size_t)std::max<int64_t>(nc, 1));
    CHECK(gpet_listmode_records(coinc.data(), nc, mod_n, cry_n, GPET_LM_TRUE, lm.data()) == GPET_OK, "lm records");
    CHECK(gpet_write_listmode((tmp + "/c.lm").c_str(), coinc.data(), nc, delayed.data(), nl, mod_n, cry_n, npan, 0) ==
              GPET_OK, "write listmode");
    CHECK(gpet_write_sinogram((tmp + "/sino").c_str(), sg.data(), 64, 48, 8.0f) == GPET_OK, "write sinogram");
    CHECK(gpet_write_events((tmp + "/no/such/dir/ev.dat").c_str(), singles.data(), ns, 0) == GPET_ERR_IO,
          "write into a missing directory");

    gpet_cpu_destroy(c);
    gpet_tables_free(&tabs);
    std::printf("oracle: %lld decays, %lld events, %lld singles, %lld coincidences, %lld delayed, %lld hits\n",
                (long long)nd, (long long)sink.n_adder, (long long)sink.n_singles, (long long)sink.n_coinc,
                (long long)sink.n_delayed, (long long)sink.n_hits);
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("sanitize driver ok\n");
    return 0;
}
