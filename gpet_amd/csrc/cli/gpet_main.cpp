/*
 * gpet_main.cpp - drop-in command line: `gpet input_PET.in`.
 *
 * Restates main.cu:26-229: parse input_PET.in, load the phantom, the tables
 * (data/input4gPET.* relative to the working directory, as the reference),
 * the detector geometry, then run the source or the phase-space file.
 * Outputs go to ./output/ (singles.dat, adder.dat, coincidences.dat and,
 * with --hits, HitsID.dat/Hits.dat), truncated per run unless --append.
 *
 * Options: --data DIR (tables), --out DIR, --hits, --append, --batch N,
 * --psf-source / --psf-phantom (outsource/idsource/timesource.dat and
 * outphantom/idphantom/timephantom.dat: OUTPUTPSF 1 / 2, gPET.cu:63-114),
 * --record-sphere (move photons leaving the phantom onto the input file's
 * record sphere: RECORDPSF -1), --autotune (pick launch shapes by timing),
 * --listmode (coincidences.lm + .lmh), --sinogram NS NPHI SMAX_CM
 * (sinogram.s + .hs, and sinogram_delayed.* with a delayed window).
 */
#include "../../../include/gpet.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

static int die(const char *what, gpet_status s, gpet_ctx *c)
{
    fprintf(stderr, "gpet: %s failed: %s%s%s\n", what, gpet_status_string(s), c ? " - " : "", c ? gpet_last_error(c) : "");
    if (c) gpet_destroy(c);
    return EXIT_FAILURE;
}

static bool read_bin(const char *path, void *dst, size_t bytes)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    size_t got = fread(dst, 1, bytes, f);
    fclose(f);
    return got == bytes;
}

int main(int argc, char **argv)
{
    const char *input = nullptr;
    std::string data_dir = "data", out_dir = "output";
    bool hits = false, append = false, sphere = false;
    long long batch = 0;
    int32_t flags = 0;
    int sino_s = 0, sino_phi = 0;
    float sino_smax = 0.0f;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--data") && i + 1 < argc) data_dir = argv[++i];
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) out_dir = argv[++i];
        else if (!strcmp(argv[i], "--hits")) hits = true;
        else if (!strcmp(argv[i], "--append")) append = true;
        else if (!strcmp(argv[i], "--batch") && i + 1 < argc) batch = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--psf-source")) flags |= GPET_FLAG_PSF_SOURCE;
        else if (!strcmp(argv[i], "--psf-phantom")) flags |= GPET_FLAG_PSF_PHANTOM;
        else if (!strcmp(argv[i], "--autotune")) flags |= GPET_FLAG_AUTOTUNE;
        else if (!strcmp(argv[i], "--record-sphere")) sphere = true;
        else if (!strcmp(argv[i], "--listmode")) flags |= GPET_FLAG_LISTMODE;
        else if (!strcmp(argv[i], "--sinogram") && i + 3 < argc) {
            sino_s = atoi(argv[++i]);
            sino_phi = atoi(argv[++i]);
            sino_smax = (float)atof(argv[++i]);
        }
        else if (!input) input = argv[i];
    }
    if (!input) {
        printf("Please execute ./gpet input_file [--data DIR] [--out DIR] [--hits] [--append] [--batch N] "
               "[--psf-source] [--psf-phantom] [--record-sphere] [--autotune] [--listmode] "
               "[--sinogram NS NPHI SMAX_CM]\n");
        return EXIT_FAILURE;
    }
    gpet_input in;
    gpet_status st = gpet_read_input(input, &in);
    if (st) return die("reading input file", st, nullptr);

    gpet_options opt;
    memset(&opt, 0, sizeof(opt));
    opt.device = in.device;
    opt.seed = in.seed;
    opt.eabs = in.eabs;
    opt.noncol_sigma = in.noncol_sigma;
    opt.positron_range = in.prange;
    opt.record_hits = hits ? 1 : 0;
    opt.batch = batch;
    opt.flags = flags;
    if (sphere)
        for (int k = 0; k < 4; ++k) opt.record_sphere[k] = in.record_sphere[k];
    gpet_ctx *c = nullptr;
    if ((st = gpet_create(&c, &opt))) return die("device init", st, nullptr);

    gpet_tables tab;
    if ((st = gpet_tables_load(data_dir.c_str(), &tab))) return die("loading tables", st, c);
    if (in.eabs < tab.eminph) {
        fprintf(stderr, "gpet: Eabs out of range.\n"); /* initialize.cu:898-902 */
        gpet_tables_free(&tab);
        gpet_destroy(c);
        return EXIT_FAILURE;
    }
    st = gpet_set_tables(c, &tab);
    gpet_tables_free(&tab);
    if (st) return die("uploading tables", st, c);

    size_t nv = (size_t)in.dims[0] * in.dims[1] * in.dims[2];
    std::vector<int32_t> mat(nv);
    std::vector<float> den(nv);
    if (!read_bin(in.mat_file, mat.data(), nv * 4) || !read_bin(in.den_file, den.data(), nv * 4))
        return die("reading phantom files", GPET_ERR_IO, c);
    if ((st = gpet_set_phantom(c, mat.data(), den.data(), in.dims, in.offset, in.size))) return die("phantom", st, c);

    std::vector<gpet_panel> panels(GPET_MAXPANELS);
    gpet_panel_material pm[2];
    int32_t np = 0;
    if ((st = gpet_read_geometry(in.geo_file, panels.data(), GPET_MAXPANELS, &np, pm))) return die("geometry", st, c);
    if ((st = gpet_set_detector(c, panels.data(), np, pm, in.surfaces, in.n_surfaces))) return die("detector", st, c);
    gpet_digitizer dg;
    gpet_input_digitizer(&in, &dg);
    if ((st = gpet_set_digitizer(c, &dg))) return die("digitizer", st, c);

    mkdir(out_dir.c_str(), 0755);
    gpet_sink sink;
    memset(&sink, 0, sizeof(sink));
    std::string tmp_dir = out_dir;
    sink.out_dir = tmp_dir.c_str();
    sink.out_append = append ? 1 : 0;
    auto t0 = std::chrono::steady_clock::now();
    if (in.use_psf) {
        int64_t n = 0;
        if ((st = gpet_read_psf(in.source_file, nullptr, 0, &n))) return die("reading PSF", st, c);
        if (in.n_histories < n) n = in.n_histories;
        std::vector<double> rec(8 * (size_t)n);
        if ((st = gpet_read_psf(in.source_file, rec.data(), n, &n))) return die("reading PSF", st, c);
        if ((st = gpet_run_psf(c, rec.data(), n, in.ptype, &sink))) return die("PSF run", st, c);
    } else {
        std::vector<gpet_isotope> iso(64);
        std::vector<gpet_source> src(4096);
        int32_t ni = 0, ns = 0;
        if ((st = gpet_read_isotopes((data_dir + "/isotopes.txt").c_str(), iso.data(), 64, &ni)))
            return die("reading isotopes", st, c);
        if ((st = gpet_read_sources(in.source_file, src.data(), 4096, &ns))) return die("reading sources", st, c);
        if ((st = gpet_run_source(c, src.data(), ns, iso.data(), ni, in.t0, in.t1, 0, &sink)))
            return die("source run", st, c);
    }
    auto t1 = std::chrono::steady_clock::now();
    gpet_tallies t;
    memset(&t, 0, sizeof(t));
    std::vector<int64_t> sino, sino_d;
    if (sino_s > 0 && sino_phi > 0 && sino_smax > 0.0f) {
        sino.assign((size_t)sino_s * sino_phi, 0);
        sino_d.assign(sino.size(), 0);
        t.sinogram = sino.data();
        t.sinogram_delayed = sino_d.data();
        t.n_s = sino_s;
        t.n_phi = sino_phi;
        t.s_max = sino_smax;
    }
    if ((st = gpet_get_tallies(c, &t))) return die("tallies", st, c);
    if (!sino.empty()) {
        if ((st = gpet_write_sinogram((out_dir + "/sinogram").c_str(), sino.data(), sino_s, sino_phi, sino_smax)))
            return die("writing the sinogram", st, c);
        if (dg.delay_us > 0.0 &&
            (st = gpet_write_sinogram((out_dir + "/sinogram_delayed").c_str(), sino_d.data(), sino_s, sino_phi,
                                      sino_smax)))
            return die("writing the delayed sinogram", st, c);
    }
    printf("{\"histories\": %lld, \"photons\": %lld, \"panel_entries\": %lld, \"hits\": %lld, \"adder\": %lld, "
           "\"threshold\": %lld, \"deadtime\": %lld, \"singles\": %lld, \"coincidences\": %lld, \"true_coincidences\": %lld, "
           "\"dropped_events\": %lld, \"majorant_violations\": %lld, \"background_singles\": %lld, "
           "\"delayed_coincidences\": %lld, \"seconds_transport\": %.6f, "
           "\"seconds_digitizer\": %.6f, \"seconds_wall\": %.6f}\n",
           (long long)t.n_histories, (long long)t.n_photons, (long long)t.n_panel_entries, (long long)t.n_hits,
           (long long)t.n_adder, (long long)t.n_threshold, (long long)t.n_deadtime, (long long)t.n_singles,
           (long long)t.n_coinc, (long long)t.n_coinc_true, (long long)t.dropped_events,
           (long long)t.majorant_violations, (long long)t.n_noise, (long long)t.n_delayed, t.seconds_transport,
           t.seconds_digitizer,
           std::chrono::duration<double>(t1 - t0).count());
    gpet_destroy(c);
    return EXIT_SUCCESS;
}
