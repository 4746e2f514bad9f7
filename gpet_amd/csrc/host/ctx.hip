/*
 * ctx.hip - the C ABI context: device state, run orchestration, outputs.
 *
 * Restates the host side of the reference (init/iniPanel/iniwck in
 * initialize.cu, simulateParticle/sampleParticle in gPET.cu) around the HIP
 * kernels of transport.hip and digitizer.hip.  Differences by design:
 *   - status codes instead of exit(); every HIP call is checked;
 *   - decays are sampled directly (O(decays), not O(atoms): F8/Q17);
 *   - all events of a run stay in HBM and the digitizer runs once over them,
 *     so deadtime and coincidences see one global time order instead of
 *     per-batch orders (gPET.cu:260-428 digitizes each epoch separately);
 *   - outputs are truncated, not appended (Q21).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/gpet.h"
#include "../common/gp_math.h"
#include "../common/gp_binomial.h"
#include "../common/gp_walkmap.h"
#include "internal.h"

#include "context.h"

namespace {

/* ---- Woodcock majorant (iniwck, initialize.cu:773-829, 919-966) ---------
 * Same node grid as the reference; the node value is 1/max of the linearly
 * interpolated mu*rho_max over both adjacent intervals (exact maximum of a
 * piecewise-linear function) shrunk by 1e-5, so that lambda*mu <= 1 holds
 * everywhere for the lerped tables the kernels use (Q18 fixed). */
float host_lerp(const float *T, int n, float e0, float idl, float E)
{
    float x = (E - e0) * idl;
    x = std::fmin(std::fmax(x, 0.0f), (float)(n - 1));
    int j = (int)x;
    if (j > n - 2) j = n - 2;
    float f = x - (float)j;
    return T[j] + f * (T[j + 1] - T[j]);
}

double mu_lin_max(const gpet_ctx *c, const float *rho_max, double E)
{
    double best = 0.0;
    for (int m = 0; m < c->nmat; ++m) {
        if (rho_max[m] <= 0.0f) continue;
        double v = (double)host_lerp(c->h_lamph.data() + (size_t)m * GPET_NTAB, GPET_NTAB, c->e0, c->idl, (float)E) *
                   rho_max[m];
        if (v > best) best = v;
    }
    return best;
}

void build_majorant(const gpet_ctx *c, const float *rho_max, std::vector<float> &lam, float &we0, float &widl)
{
    we0 = c->eminph;
    float de = (c->emax * (1.0f - 1.0e-10f) - we0) / (float)GPET_NWCK;
    widl = 1.0f / de;
    lam.assign(GPET_NWCK, 0.0f);
    const double d_node = 1.0 / (double)c->idl;
    std::vector<double> imax(GPET_NWCK + 1, 0.0);
    for (int i = 0; i <= GPET_NWCK; ++i) {
        double lo = (double)we0 + (double)de * i;
        double hi = (double)we0 + (double)de * (i + 1);
        double m = std::max(mu_lin_max(c, rho_max, lo), mu_lin_max(c, rho_max, hi));
        double k0 = std::ceil((lo - c->e0) / d_node);
        double k1 = std::floor((hi - c->e0) / d_node);
        for (double k = k0; k <= k1; k += 1.0) m = std::max(m, mu_lin_max(c, rho_max, (double)c->e0 + k * d_node));
        imax[i] = m;
    }
    for (int i = 0; i < GPET_NWCK; ++i) {
        double m = imax[i];
        if (i > 0) m = std::max(m, imax[i - 1]);
        lam[i] = (m > 0.0) ? (float)(1.0 / (m * (1.0 + 1.0e-5))) : 1.0e20f;
    }
}

/* ---- source plan (restates the sampling rule documented in gpet.h) ----- */

/* Decay counts of one acquisition's sources (appended to cnt and ds).
 * src_base is the index of src[0] in the run's source array: it keys the
 * binomial stream, so a source draws the same count alone or in a batch. */
gpet_status plan_job(gpet_ctx *c, const gpet_source *src, int ns, const gpet_isotope *iso, int niso, double t0,
                     double t1, int64_t n_decays, int src_base, int phantom, std::vector<int64_t> &cnt_out,
                     std::vector<DevSource> &ds_out)
{
    std::vector<DevSource> ds(ns);
    std::vector<double> expect(ns, 0.0), navail(ns, 0.0);
    for (int i = 0; i < ns; ++i) {
        if (src[i].isotope < 0 || src[i].isotope >= niso) return fail(c, GPET_ERR_ARG, "source isotope out of range");
        double th = iso[src[i].isotope].half_life_s;
        if (!(th > 0.0)) return fail(c, GPET_ERR_ARG, "half life must be positive");
        memset(&ds[i], 0, sizeof(DevSource));
        ds[i].tau = th / std::log(2.0);
        ds[i].pdec = -std::expm1(-(t1 - t0) / ds[i].tau);
        ds[i].t0 = t0;
        ds[i].shape = src[i].shape;
        ds[i].iso = src[i].isotope;
        ds[i].phantom = phantom;
        for (int k = 0; k < 6; ++k) ds[i].coef[k] = src[i].coef[k];
        navail[i] = std::floor(src[i].natom * std::exp2(-t0 / th));
        expect[i] = navail[i] * ds[i].pdec;
    }
    std::vector<int64_t> cnt(ns, 0);
    if (n_decays > 0) {
        double tot = 0;
        for (double e : expect) tot += e;
        if (tot <= 0) return fail(c, GPET_ERR_ARG, "sources have no activity");
        int64_t assigned = 0;
        std::vector<std::pair<double, int>> rem;
        for (int i = 0; i < ns; ++i) {
            double share = (double)n_decays * expect[i] / tot;
            cnt[i] = (int64_t)std::floor(share);
            assigned += cnt[i];
            rem.push_back({share - (double)cnt[i], i});
        }
        std::stable_sort(rem.begin(), rem.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) { return a.first > b.first; });
        for (int k = 0; assigned < n_decays; ++k, ++assigned) cnt[rem[k % ns].second]++;
    } else {
        for (int i = 0; i < ns; ++i) cnt[i] = gp_binomial_draw(navail[i], ds[i].pdec, c->opt.seed, src_base + i);
    }
    cnt_out.insert(cnt_out.end(), cnt.begin(), cnt.end());
    ds_out.insert(ds_out.end(), ds.begin(), ds.end());
    return GPET_OK;
}

bool ev_less_id(const gpet_event &a, const gpet_event &b)
{
    if (a.parn != b.parn) return a.parn < b.parn;
    return a.siten < b.siten;
}

gpet_status check_ready(gpet_ctx *c)
{
    if (!c) return GPET_ERR_ARG;
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables not set");
    if (!c->has_dig) return fail(c, GPET_ERR_STATE, "digitizer not set");
    return GPET_OK;
}

/* Blocks until the last run's asynchronous sink copies have ended.  They
 * read the digitizer's output buffers and write the caller's. */
gpet_status copy_wait(gpet_ctx *c)
{
    if (!c->copy_pending) return GPET_OK;
    c->copy_pending = false;
    HIPOK(c, hipEventSynchronize(c->copy_done));
    return GPET_OK;
}

gpet_status write_outputs(gpet_ctx *c, const std::string &dir, const std::vector<gpet_event> &adder,
                          const std::vector<gpet_event> &singles, const std::vector<gpet_event> &coinc,
                          const std::vector<int32_t> &hid, const std::vector<float> &hf,
                          const std::vector<gpet_event> &delayed, int32_t app)
{
    gpet_status st = gpet_write_events((dir + "/adder.dat").c_str(), adder.data(), (int64_t)adder.size(), app);
    if (!st) st = gpet_write_events((dir + "/singles.dat").c_str(), singles.data(), (int64_t)singles.size(), app);
    if (!st) st = gpet_write_events((dir + "/coincidences.dat").c_str(), coinc.data(), (int64_t)coinc.size(), app);
    if (!st && c->dig.delay_us > 0.0)
        st = gpet_write_events((dir + "/delayed.dat").c_str(), delayed.data(), (int64_t)delayed.size(), app);
    if (!st && (c->opt.flags & GPET_FLAG_LISTMODE))
        st = gpet_write_listmode((dir + "/coincidences.lm").c_str(), coinc.data(), (int64_t)coinc.size() / 2,
                                 delayed.data(), (int64_t)delayed.size() / 2, c->ddet.moduleN, c->ddet.crystalN,
                                 c->ddet.npanels, app);
    if (!st && c->opt.record_hits)
        st = gpet_write_hits((dir + "/HitsID.dat").c_str(), (dir + "/Hits.dat").c_str(), hid.data(), hf.data(),
                             (int64_t)hid.size() / 5, app);
    if (st) return fail(c, st, "cannot write outputs to " + dir);
    return GPET_OK;
}

/* The batch loop shared by source and PSF runs. */
gpet_status upload_run_sources(gpet_ctx *c, const std::vector<int64_t> &first, const std::vector<DevSource> &ds,
                               const gpet_isotope *iso, int niso)
{
    std::vector<DevIsotope> di(niso);
    for (int i = 0; i < niso; ++i) {
        di[i].ratio = iso[i].beta_ratio;
        for (int k = 0; k < 8; ++k) di[i].coef[k] = iso[i].coef[k];
    }
    gpet_status st;
    if ((st = upload(c, c->d_first, first.data(), (int64_t)first.size())) ||
        (st = upload(c, c->d_src, ds.data(), (int64_t)ds.size())) || (st = upload(c, c->d_iso, di.data(), niso)))
        return st;
    return GPET_OK;
}

/* What a batched run adds: history ranges per source, their phantom slots
 * and the first history of each job. */
struct BatchInfo {
    std::vector<int64_t> first;      /* [nsrc + 1]  */
    std::vector<int> src_phantom;    /* [nsrc]      */
    std::vector<int64_t> job_first;  /* [njobs + 1] */
};

/* Acquisition windows (s) and first history of each job: where background
 * singles are generated. */
struct NoiseWindows {
    std::vector<double> t0, t1;
    std::vector<int64_t> first;
};

/* Segments of the launch covering histories [h0, h0 + nh): maximal runs of
 * consecutive sources on one phantom slot, each starting a new virtual
 * block of `block` histories. */
void build_segments(const BatchInfo &bi, int64_t h0, int64_t nh, int block, std::vector<DevSegment> &out,
                    int64_t &nvb)
{
    const int nsrc = (int)bi.src_phantom.size();
    nvb = 0;
    int64_t h = h0;
    const int64_t end = h0 + nh;
    int s = (int)(std::upper_bound(bi.first.begin(), bi.first.end(), h) - bi.first.begin()) - 1;
    while (h < end) {
        while (s < nsrc - 1 && bi.first[s + 1] <= h) ++s;
        const int ph = bi.src_phantom[s];
        int t = s + 1;
        while (t < nsrc && bi.src_phantom[t] == ph) ++t;
        const int64_t e = std::min(end, bi.first[t]);
        DevSegment g;
        g.h0 = h;
        g.nh = e - h;
        g.vb0 = nvb;
        g.phantom = ph;
        g.pad = 0;
        nvb += (g.nh + block - 1) / block;
        out.push_back(g);
        h = e;
        s = t;
    }
}

LaunchCfg cfg_of(const gpet_launch &L, int k)
{
    LaunchCfg c;
    c.block = L.block[k];
    c.wpe = L.wpe[k];
    c.blocks_per_cu = L.blocks_per_cu[k];
    if (k == 0)
        for (int i = 0; i < 3; ++i) c.rounds[i] = L.transit[i + 1];
    return c;
}

gpet_launch default_launch()
{
    gpet_launch L;
    memset(&L, 0, sizeof(L));
    L.block[0] = 256; L.block[1] = 256;
    L.wpe[0] = 0; L.wpe[1] = 0;
    L.blocks_per_cu[0] = 16; L.blocks_per_cu[1] = 8;  /* the tuner's pick on the cylinder workload */
    L.batch = DEFAULT_BATCH;
    return L;
}

gpet_status check_launch(gpet_ctx *c, const gpet_launch &L)
{
    for (int k = 0; k < 2; ++k) {
        if (L.block[k] < 64 || L.block[k] > 256 || L.block[k] % 64) return fail(c, GPET_ERR_ARG, "launch block size");
        bool ok = false;
        for (int v : gpet_wpe_variants) ok = ok || v == L.wpe[k];
        if (!ok) return fail(c, GPET_ERR_ARG, "launch register variant (wpe) must be 0, 5 or 6");
        if (L.blocks_per_cu[k] < 0 || L.blocks_per_cu[k] > 64) return fail(c, GPET_ERR_ARG, "launch blocks per CU");
        /* the phantom kernel's grid sets the panel queue's regions: 1..16 per CU */
        if (k == 0 && (L.blocks_per_cu[0] < 1 || L.blocks_per_cu[0] > 16))
            return fail(c, GPET_ERR_ARG, "phantom-kernel blocks per CU must be 1..16");
    }
    if (L.batch < 256 || L.batch > ((int64_t)1 << 26)) return fail(c, GPET_ERR_ARG, "launch batch size");
    if (L.transit[0] < 0 || L.transit[0] > 1024) return fail(c, GPET_ERR_ARG, "launch transit steps");
    for (int k = 1; k < 4; ++k)
        if (L.transit[k] < 0 || L.transit[k] > 1 << 20 || (L.transit[k] > 0 && k > 1 && L.transit[k - 1] == 0))
            return fail(c, GPET_ERR_ARG, "launch transit rounds (0 ends the list)");
    return GPET_OK;
}

gpet_status alloc_queue(gpet_ctx *c, int64_t qcap)
{
    gpet_status st;
    if ((st = dalloc(c, c->d_rcount, 64 * (int64_t)c->n_cu + 1))) return st;
    if ((st = dalloc(c, c->d_lw_count, 64 * (int64_t)c->n_cu + 1))) return st;
    if ((st = dalloc(c, c->d_lw, qcap / 8 + 1))) return st;
    const gpet_launch &L = c->launch;
    for (int k = 0; k < 2; ++k) {
        if (L.transit[0] > 0 && (k == 0 || L.transit[1] > 0)) {
            if ((st = dalloc(c, c->d_tq[k], qcap))) return st;
            if ((st = dalloc(c, c->d_tq_count[k], 64 * (int64_t)c->n_cu + 1))) return st;
        }
    }
    if (qcap > c->q_cap) {
        for (int k = 0; k < 6; ++k)
            if ((st = dalloc(c, c->q_f[k], qcap))) return st;
        if ((st = dalloc(c, c->q_t, qcap)) || (st = dalloc(c, c->q_info, qcap))) return st;
        c->q_cap = qcap;
    }
    return GPET_OK;
}

DevQueue make_queue(gpet_ctx *c, int64_t qcap, bool want_trace)
{
    DevQueue q;
    memset(&q, 0, sizeof(q));
    q.rcount = c->d_rcount.p;
    q.y = c->q_f[0].p; q.z = c->q_f[1].p; q.vx = c->q_f[2].p; q.vy = c->q_f[3].p; q.vz = c->q_f[4].p; q.E = c->q_f[5].p;
    q.t = c->q_t.p;
    q.info = c->q_info.p;
    q.trace_slot = want_trace ? c->q_slot.p : nullptr;
    q.count = (int *)c->d_small.p;
    q.cap = qcap;
    q.lw = c->d_lw.p;
    q.lw_count = c->d_lw_count.p;
    const gpet_launch &L = c->launch;
    if (L.transit[0] > 0) {
        q.tq = c->d_tq[0].p;
        q.tq_count = c->d_tq_count[0].p;
        q.tq_steps = L.transit[0];
        if (L.transit[1] > 0) {
            q.tq2 = c->d_tq[1].p;
            q.tq2_count = c->d_tq_count[1].p;
        }
    }
    return q;
}

/* Launch arguments of one transport batch; the panel queue has G regions
 * of `region` slots (one per phantom-kernel block). */
struct BatchLaunch {
    RunParams rp;
    int64_t nh;
    int G;
    int64_t region;
};

/* Queue shape of a phantom launch over `items` histories (virtual blocks x
 * block size for batched runs): a block handles at most
 * ceil(items / (G * block)) rounds of `block` histories, 2 photons each. */
void queue_shape(int64_t items, const gpet_launch &L, int n_cu, int &G, int64_t &region)
{
    const LaunchCfg cfg = cfg_of(L, 0);
    G = gpet_phantom_grid(std::max<int64_t>(items, 1), cfg, n_cu);
    const int64_t rounds = (std::max<int64_t>(items, 1) + (int64_t)G * cfg.block - 1) / ((int64_t)G * cfg.block);
    region = 2 * rounds * cfg.block;
}

gpet_status launch_batch(gpet_ctx *c, const BatchLaunch &B, const DevQueue &q0, const PanelParams &pp,
                         const gpet_launch &L, bool batched, hipEvent_t *ev, int pmode)
{
    DevQueue q = q0;
    q.nregions = B.G;
    q.region = B.region;
    if ((int64_t)B.G * B.region > q.cap) return fail(c, GPET_ERR_CAPACITY, "panel queue shape exceeds its buffer");
    if (ev) HIPOK(c, hipEventRecord(ev[0], c->stream));
    int e = gpet_launch_phantom(B.rp, c->dtab, c->dph, c->ddet, q, batched, cfg_of(L, 0), c->n_cu, c->stream);
    if (e) return fail(c, GPET_ERR_DEVICE, std::string("phantom kernel: ") + hipGetErrorString((hipError_t)e));
    if (ev) HIPOK(c, hipEventRecord(ev[1], c->stream));
    if (ev) HIPOK(c, hipEventRecord(ev[2], c->stream));
    e = gpet_launch_panel(pp, c->dtab, c->ddet, q, 2 * B.nh, cfg_of(L, 1), c->n_cu, c->stream, pmode);
    if (e) return fail(c, GPET_ERR_DEVICE, std::string("panel kernel: ") + hipGetErrorString((hipError_t)e));
    if (ev) HIPOK(c, hipEventRecord(ev[3], c->stream));
    return GPET_OK;
}

template <class T>
void key_add(std::vector<unsigned char> &k, const T &v)
{
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

void drop_graph(gpet_ctx *c)
{
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    c->graph.exec = nullptr;
    c->graph.key.clear();
}

/* Replays batches 1..n-1 as one hipGraph (captured on first use, reused
 * while every launch argument is unchanged).  Returns false, leaving the
 * stream out of capture mode, if capture is not possible; the caller then
 * launches the batches one by one. */
bool run_graph(gpet_ctx *c, const std::vector<BatchLaunch> &B, const DevQueue &q, const PanelParams &pp,
               const gpet_launch &L, bool batched, int pmode)
{
    std::vector<unsigned char> key;
    for (size_t b = 1; b < B.size(); ++b) {
        key_add(key, B[b].rp);
        key_add(key, B[b].nh);
    }
    key_add(key, pp);
    key_add(key, q);
    key_add(key, c->dtab);
    key_add(key, c->dph);
    key_add(key, c->ddet);
    key_add(key, L);
    key_add(key, batched);
    key_add(key, pmode);
    if (!(c->graph.exec && key == c->graph.key)) {
        drop_graph(c);
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        bool ok = true;
        for (size_t b = 1; b < B.size() && ok; ++b)
            ok = launch_batch(c, B[b], q, pp, L, batched, &c->kev[4 * b], pmode) == GPET_OK;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!ok || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return false;
        }
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        c->graph.exec = x;
        c->graph.key.swap(key);
    }
    return hipGraphLaunch(c->graph.exec, c->stream) == hipSuccess;
}

/* Launch-shape tuner: times the variants of each kernel on a prefix of the
 * run's own histories (coordinate search: register variant, then grid
 * shape, then block size) and keeps the fastest.  Every variant computes
 * the same histories, so the choice cannot change results; the prefix's
 * outputs are discarded. */
gpet_status autotune(gpet_ctx *c, const RunParams &rp0, int64_t n_hist, const PanelParams &pp0, const BatchInfo *bi,
                     double *seconds)
{
    auto t0 = std::chrono::steady_clock::now();
    gpet_status st;
    /* trials on one launch's worth of the run's histories (at least 2M):
     * the best grid depends on how many histories a launch covers */
    int64_t ns = std::min<int64_t>(n_hist, std::max<int64_t>((int64_t)1 << 21, c->launch.batch));
    /* queue large enough for every grid the search tries (>= 1 block per CU,
     * <= 256 threads per block) */
    auto qcap_of = [&](int64_t n) { return 2 * n + 2 * (int64_t)64 * c->n_cu * 256; };
    /* the trials hold the panel queue, the long-walk buffer, two transit
     * buffers and an event buffer of that size: halve the trial (down to
     * 2M histories) until they take at most half of the free memory */
    {
        size_t free_b = 0, total_b = 0;
        HIPOK(c, hipMemGetInfo(&free_b, &total_b));
        auto bytes = [&](int64_t n) {
            const double q = (double)qcap_of(n);
            return q * (48.0 + sizeof(LongWalkItem) / 8.0 + 2.0 * sizeof(TransitItem)
                        + GPET_MAX_PH_EVENTS * sizeof(gpet_event));
        };
        while (ns > ((int64_t)1 << 21) && bytes(ns) > 0.5 * (double)free_b) ns /= 2;
    }
    const int64_t qcap = qcap_of(ns);
    if ((st = alloc_queue(c, qcap))) return st;
    if ((st = dgrow(c, c->d_events, GPET_MAX_PH_EVENTS * qcap, 0))) return st;
    DevQueue q = make_queue(c, qcap, false);
    PanelParams pp = pp0;
    pp.hit_id = nullptr;
    pp.hit_f = nullptr;
    pp.trace = nullptr;
    pp.events = c->d_events.p;
    pp.events_cap = c->d_events.cap;
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    gpet_launch L = c->launch;
    std::vector<DevSegment> segs;

    auto time_phantom = [&](const gpet_launch &T, float &ms) -> gpet_status {
        RunParams rp = rp0;
        rp.h0 = 0;
        rp.nh = ns;
        rp.trace = nullptr;
        rp.trace_cap = 0;
        rp.psf_src = rp.psf_ph = nullptr;
        if (bi) {
            segs.clear();
            int64_t nvb = 0;
            build_segments(*bi, 0, ns, T.block[0], segs, nvb);
            gpet_status s2 = upload(c, c->d_seg, segs.data(), (int64_t)segs.size());
            if (s2) return s2;
            rp.seg = c->d_seg.p;
            rp.nseg = (int)segs.size();
            rp.n_vblocks = nvb;
        }
        queue_shape(bi ? rp.n_vblocks * T.block[0] : ns, T, c->n_cu, q.nregions, q.region);
        if ((int64_t)q.nregions * q.region > q.cap) return fail(c, GPET_ERR_CAPACITY, "tuner queue shape");
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            HIPOK(c, hipEventRecord(e0, c->stream));
            int e = gpet_launch_phantom(rp, c->dtab, c->dph, c->ddet, q, bi != nullptr, cfg_of(T, 0), c->n_cu, c->stream);
            if (e) return fail(c, GPET_ERR_DEVICE, std::string("phantom kernel: ") + hipGetErrorString((hipError_t)e));
            HIPOK(c, hipEventRecord(e1, c->stream));
            HIPOK(c, hipEventSynchronize(e1));
            float m = 0;
            HIPOK(c, hipEventElapsedTime(&m, e0, e1));
            best = std::min(best, m);
        }
        ms = best;
        return GPET_OK;
    };
    auto time_panel = [&](const gpet_launch &T, float &ms) -> gpet_status {
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            int e = gpet_launch_zero(c->d_small.p + 1, 1, c->stream);
            if (e) return fail(c, GPET_ERR_DEVICE, "zero kernel");
            HIPOK(c, hipEventRecord(e0, c->stream));
            e = gpet_launch_panel(pp, c->dtab, c->ddet, q, 2 * ns, cfg_of(T, 1), c->n_cu, c->stream);
            if (e) return fail(c, GPET_ERR_DEVICE, std::string("panel kernel: ") + hipGetErrorString((hipError_t)e));
            HIPOK(c, hipEventRecord(e1, c->stream));
            HIPOK(c, hipEventSynchronize(e1));
            float m = 0;
            HIPOK(c, hipEventElapsedTime(&m, e0, e1));
            best = std::min(best, m);
        }
        ms = best;
        return GPET_OK;
    };
    /* coordinate search over one kernel's knobs */
    auto search = [&](int k, auto timer) -> gpet_status {
        float best_ms = 0;
        if ((st = timer(L, best_ms))) return st;
        const int wpes[] = {0, 5, 6};
        const int bpcs[] = {1, 2, 4, 8, 16};
        const int blocks[] = {128, 256};
        for (int stage = 0; stage < 3; ++stage) {
            gpet_launch base = L;
            const int *vals = stage == 0 ? wpes : (stage == 1 ? bpcs : blocks);
            int nv = stage == 0 ? 3 : (stage == 1 ? 5 : 2);
            for (int v = 0; v < nv; ++v) {
                gpet_launch T = base;
                if (stage == 0) T.wpe[k] = vals[v];
                else if (stage == 1) T.blocks_per_cu[k] = vals[v];
                else T.block[k] = vals[v];
                if (memcmp(&T, &base, sizeof(T)) == 0) continue;
                float ms = 0;
                if ((st = timer(T, ms))) return st;
                if (ms < best_ms * 0.98f) {
                    best_ms = ms;
                    L = T;
                }
            }
        }
        return GPET_OK;
    };
    if ((st = search(0, time_phantom))) return st;
    /* photon regrouping: the current schedule, off, or one for phantoms
     * where photons take many steps (kept when 2% faster) */
    {
        auto set_tq = [&](const gpet_launch &T) {
            q.tq = nullptr; q.tq_count = nullptr; q.tq_steps = 0; q.tq2 = nullptr; q.tq2_count = nullptr;
            if (T.transit[0] > 0) {
                q.tq = c->d_tq[0].p; q.tq_count = c->d_tq_count[0].p; q.tq_steps = T.transit[0];
                if (T.transit[1] > 0) { q.tq2 = c->d_tq[1].p; q.tq2_count = c->d_tq_count[1].p; }
            }
        };
        for (int k = 0; k < 2; ++k)
            if ((st = dalloc(c, c->d_tq[k], qcap)) || (st = dalloc(c, c->d_tq_count[k], 64 * (int64_t)c->n_cu + 1)))
                return st;
        const int32_t presets[3][4] = {{0, 0, 0, 0}, {3, 6, 12, 0}, {4, 8, 16, 0}};
        float best_ms = 0;
        set_tq(L);
        if ((st = time_phantom(L, best_ms))) return st;
        /* each schedule with each register budget (the regrouped kernel
         * holds more state, so its best budget differs) */
        const gpet_launch base = L;
        for (const auto &pr : presets) {
            for (int w : gpet_wpe_variants) {
                gpet_launch T = base;
                memcpy(T.transit, pr, sizeof(T.transit));
                T.wpe[0] = w;
                if (memcmp(&T, &base, sizeof(T)) == 0) continue;
                set_tq(T);
                float ms = 0;
                if ((st = time_phantom(T, ms))) return st;
                if (ms < best_ms * 0.98f) {
                    best_ms = ms;
                    L = T;
                }
            }
        }
        set_tq(L);
    }
    /* leave a queue from the chosen phantom variant for the panel trials */
    float dummy = 0;
    if ((st = time_phantom(L, dummy))) return st;
    if ((st = search(1, time_panel))) return st;
    c->launch = L;
    /* transit buffers the chosen schedule does not use */
    for (int k = 0; k < 2; ++k)
        if (L.transit[0] == 0 || (k == 1 && L.transit[1] == 0)) { dfree(c->d_tq[k]); dfree(c->d_tq_count[k]); }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return GPET_OK;
}

gpet_status run_common(gpet_ctx *c, RunParams rp, int64_t n_hist, gpet_sink *sink, const BatchInfo *bi,
                       const NoiseWindows *nw)
{
    gpet_status st;
    if (n_hist > GPET_MAX_HISTORIES) return fail(c, GPET_ERR_CAPACITY, "more than 2^30 histories in one run");
    if (c->dig.noise_rate > 0.0 && !nw)
        return fail(c, GPET_ERR_UNSUPPORTED, "background singles need a decay run's acquisition window");
    const bool batched = bi != nullptr;
    c->last_surr_items = 0;
    c->last_surr_seconds = 0;
    if ((st = dalloc(c, c->d_small, 4)) || (st = dalloc(c, c->d_cnt, CNT_NUM * CNT_STRIPES))) return st;

    rp.seed = c->opt.seed;
    rp.eabs = c->opt.eabs;
    rp.sigma = c->opt.noncol_sigma;
    rp.prange = c->opt.positron_range;
    rp.cnt = c->d_cnt.p;
    rp.phs = batched ? c->d_phs.p : nullptr;
    for (int k = 0; k < 4; ++k) rp.sphere[k] = c->opt.record_sphere[k];
    const bool want_psf_src = (c->opt.flags & GPET_FLAG_PSF_SOURCE) != 0;
    const bool want_psf_ph = (c->opt.flags & GPET_FLAG_PSF_PHANTOM) != 0;
    const int64_t psf_cap = 2 * std::max<int64_t>(n_hist, 1);
    if (want_psf_src || want_psf_ph) {
        if ((st = dalloc(c, c->d_npsf, 2))) return st;
        if (want_psf_src && (st = dalloc(c, c->d_psf_src, psf_cap))) return st;
        if (want_psf_ph && (st = dalloc(c, c->d_psf_ph, psf_cap))) return st;
    }
    rp.psf_src = want_psf_src ? c->d_psf_src.p : nullptr;
    rp.psf_ph = want_psf_ph ? c->d_psf_ph.p : nullptr;
    rp.n_psf = (want_psf_src || want_psf_ph) ? c->d_npsf.p : nullptr;
    rp.psf_cap = psf_cap;

    PanelParams pp;
    memset(&pp, 0, sizeof(pp));
    pp.seed = c->opt.seed;
    pp.eabs = c->opt.eabs;
    pp.readout_depth = c->dig.readout_depth;
    pp.readout_policy = c->dig.readout_policy;
    pp.record_rayleigh = c->dig.record_rayleigh;
    pp.n_events = c->d_small.p + 1;
    pp.n_hits = c->d_small.p + 2;
    pp.cnt = c->d_cnt.p;

    double tune_s = 0;
    if ((c->opt.flags & GPET_FLAG_AUTOTUNE) && !c->tuned[batched ? 1 : 0] && n_hist > 0) {
        if ((st = autotune(c, rp, n_hist, pp, bi, &tune_s))) return st;
        c->tuned[batched ? 1 : 0] = true;
    }
    const gpet_launch L = c->launch;
    const int64_t batch = L.batch;
    /* per-batch launch arguments (and the segment tables of a batched run),
     * and the panel-queue size their shapes need */
    const int64_t nbatch = (n_hist + batch - 1) / batch;
    std::vector<BatchLaunch> B((size_t)nbatch);
    int64_t qcap = 2;
    {
        std::vector<DevSegment> segs;
        std::vector<int64_t> seg_off((size_t)nbatch, 0), nvb((size_t)nbatch, 0);
        std::vector<int> nseg((size_t)nbatch, 0);
        for (int64_t b = 0; b < nbatch; ++b) {
            const int64_t h0 = b * batch, nh = std::min(batch, n_hist - h0);
            if (batched) {
                seg_off[b] = (int64_t)segs.size();
                build_segments(*bi, h0, nh, L.block[0], segs, nvb[b]);
                nseg[b] = (int)(segs.size() - seg_off[b]);
            }
        }
        if (batched && (st = upload(c, c->d_seg, segs.data(), (int64_t)segs.size()))) return st;
        for (int64_t b = 0; b < nbatch; ++b) {
            BatchLaunch &X = B[b];
            X.rp = rp;
            X.rp.h0 = b * batch;
            X.nh = X.rp.nh = std::min(batch, n_hist - X.rp.h0);
            if (rp.psf) X.rp.psf = rp.psf + 8 * X.rp.h0;
            if (batched) {
                X.rp.seg = c->d_seg.p + seg_off[b];
                X.rp.nseg = nseg[b];
                X.rp.n_vblocks = nvb[b];
            }
            queue_shape(batched ? nvb[b] * L.block[0] : X.nh, L, c->n_cu, X.G, X.region);
            qcap = std::max<int64_t>(qcap, (int64_t)X.G * X.region);
        }
    }
    if ((st = alloc_queue(c, qcap))) return st;
    const bool want_trace = sink && sink->trace && sink->trace_cap > 0;
    int64_t ntr = want_trace ? std::min<int64_t>(2 * n_hist, sink->trace_cap) : 0;
    if (want_trace)
        if ((st = dalloc(c, c->d_trace, ntr)) || (st = dalloc(c, c->q_slot, qcap))) return st;
    const bool want_hits = c->opt.record_hits != 0;
    int64_t hits_cap = 0;
    if (want_hits) {
        hits_cap = (sink && sink->hit_id && sink->hits_cap > 0) ? sink->hits_cap : (int64_t)1 << 24;
        if ((st = dalloc(c, c->d_hit_id, 5 * hits_cap)) || (st = dalloc(c, c->d_hit_f, 5 * hits_cap))) return st;
    }
    DevQueue q = make_queue(c, qcap, want_trace);
    const int pmode = c->panel_mode == PANEL_HARVEST ? PANEL_HARVEST : (c->surr.enabled ? PANEL_SURROGATE : PANEL_PLAIN);
    if (pmode == PANEL_SURROGATE) {
        if (want_hits) return fail(c, GPET_ERR_UNSUPPORTED, "surrogate mode does not produce hit records");
        if ((st = surr_run_prepare(c, n_hist, qcap, pp.sq))) return st;
    }
    if (pmode == PANEL_HARVEST) {
        pp.harvest = c->surr.set.p + SURR_REC * c->surr.set_n;
        pp.harvest_count = c->surr.set_count.p;
        pp.harvest_cap = c->surr.set.cap / SURR_REC - c->surr.set_n;
    }
    rp.trace = want_trace ? c->d_trace.p : nullptr;
    rp.trace_cap = ntr;
    for (BatchLaunch &X : B) {
        X.rp.trace = rp.trace;
        X.rp.trace_cap = ntr;
    }
    pp.hit_id = want_hits ? c->d_hit_id.p : nullptr;
    pp.hit_f = want_hits ? c->d_hit_f.p : nullptr;
    pp.hits_cap = hits_cap;
    pp.trace = rp.trace;

    while ((int64_t)c->kev.size() < 4 * std::max<int64_t>(nbatch, 1)) {
        hipEvent_t e;
        HIPOK(c, hipEventCreate(&e));
        c->kev.push_back(e);
    }

    /* Event capacity: from the last run's rate when known, else measured on
     * batch 0.  A run whose events overflow the estimate is run again with
     * the exact capacity (same histories, same results). */
    const int64_t per_batch_max = (int64_t)GPET_MAX_PH_EVENTS * qcap;
    int64_t ev_need = per_batch_max;
    if (c->ev_per_hist >= 0) ev_need = (int64_t)(c->ev_per_hist * 1.05 * (double)n_hist) + per_batch_max;
    if ((st = dgrow(c, c->d_events, std::max<int64_t>(ev_need, 1 << 20), 0))) return st;

    unsigned small[4] = {0, 0, 0, 0};
    unsigned long long cnt[CNT_NUM];
    std::vector<unsigned long long> cnt_s((size_t)CNT_NUM * CNT_STRIPES);
    bool used_graph = false;
    for (int attempt = 0;; ++attempt) {
        HIPOK(c, hipMemsetAsync(c->d_small.p, 0, sizeof(unsigned) * 4, c->stream));
        HIPOK(c, hipMemsetAsync(c->d_cnt.p, 0, sizeof(unsigned long long) * CNT_NUM * CNT_STRIPES, c->stream));
        if (pmode == PANEL_SURROGATE) HIPOK(c, hipMemsetAsync(pp.sq.count, 0, sizeof(int), c->stream));
        if (pmode == PANEL_HARVEST) HIPOK(c, hipMemsetAsync(pp.harvest_count, 0, sizeof(unsigned long long), c->stream));
        if (rp.n_psf) HIPOK(c, hipMemsetAsync(rp.n_psf, 0, 2 * sizeof(unsigned long long), c->stream));
        if (want_trace) HIPOK(c, hipMemsetAsync(c->d_trace.p, 0, sizeof(gpet_trace) * (size_t)ntr, c->stream));
        pp.events = c->d_events.p;
        pp.events_cap = c->d_events.cap;
        HIPOK(c, hipEventRecord(c->ev0, c->stream));
        used_graph = false;
        if (nbatch > 0) {
            if ((st = launch_batch(c, B[0], q, pp, L, batched, &c->kev[0], pmode))) return st;
            if (nbatch > 1 && c->ev_per_hist < 0 && attempt == 0) {
                HIPOK(c, hipMemcpyAsync(small, c->d_small.p, sizeof(small), hipMemcpyDeviceToHost, c->stream));
                HIPOK(c, hipStreamSynchronize(c->stream));
                const double rate = (double)small[1] / (double)B[0].nh;
                const int64_t need = (int64_t)small[1] + (int64_t)(rate * 1.1 * (double)(n_hist - B[0].nh)) + per_batch_max;
                if ((st = dgrow(c, c->d_events, need, (int64_t)small[1]))) return st;
                pp.events = c->d_events.p;
                pp.events_cap = c->d_events.cap;
            }
            if (nbatch > 1) {
                if (!(c->opt.flags & GPET_FLAG_NO_GRAPH)) used_graph = run_graph(c, B, q, pp, L, batched, pmode);
                if (!used_graph)
                    for (int64_t b = 1; b < nbatch; ++b)
                        if ((st = launch_batch(c, B[b], q, pp, L, batched, &c->kev[4 * b], pmode))) return st;
            }
        }
        HIPOK(c, hipEventRecord(c->ev1, c->stream));
        HIPOK(c, hipMemcpyAsync(small, c->d_small.p, sizeof(small), hipMemcpyDeviceToHost, c->stream));
        HIPOK(c, hipMemcpyAsync(cnt_s.data(), c->d_cnt.p, sizeof(unsigned long long) * cnt_s.size(),
                                hipMemcpyDeviceToHost, c->stream));
        int n_surr = 0;
        if (pmode == PANEL_SURROGATE)
            HIPOK(c, hipMemcpyAsync(&n_surr, pp.sq.count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPOK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < CNT_NUM; ++k) {
            cnt[k] = 0;
            for (int j = 0; j < CNT_STRIPES; ++j) cnt[k] += cnt_s[(size_t)k * CNT_STRIPES + j];
        }
        const bool ev_ok = (int64_t)small[1] <= c->d_events.cap;
        const bool sq_ok = (int64_t)n_surr <= pp.sq.cap;
        if (ev_ok && sq_ok) {
            if (pmode == PANEL_SURROGATE) {
                /* the network's events: at most one per listed photon */
                if ((st = dgrow(c, c->d_events, (int64_t)small[1] + n_surr, (int64_t)small[1]))) return st;
                pp.events = c->d_events.p;
                pp.events_cap = c->d_events.cap;
                if ((st = surr_run_infer(c, n_surr, pp))) return st;
                HIPOK(c, hipMemcpyAsync(small, c->d_small.p, sizeof(small), hipMemcpyDeviceToHost, c->stream));
                HIPOK(c, hipStreamSynchronize(c->stream));
                if (n_hist > 0) c->surr.items_per_hist = (double)n_surr / (double)n_hist;
                c->last_surr_items = n_surr;
            }
            break;
        }
        if (attempt >= 1) return fail(c, GPET_ERR_CAPACITY, "event buffer overflow after resize");
        if (!ev_ok && (st = dgrow(c, c->d_events, (int64_t)small[1] + per_batch_max, 0))) return st;
        if (!sq_ok) {
            c->surr.items_per_hist = (double)n_surr / (double)std::max<int64_t>(n_hist, 1);
            if ((st = surr_run_prepare(c, n_hist, qcap, pp.sq))) return st;
        }
    }
    c->n_events = small[1];
    if (n_hist > 0 && pmode != PANEL_SURROGATE) c->ev_per_hist = (double)small[1] / (double)n_hist;
    if (pmode == PANEL_HARVEST) {
        unsigned long long nh = 0;
        HIPOK(c, hipMemcpy(&nh, c->surr.set_count.p, sizeof(nh), hipMemcpyDeviceToHost));
        const int64_t added = std::min<int64_t>((int64_t)nh, pp.harvest_cap);
        /* the panel kernel appends records in scheduling order; put them in
         * photon order so that training on them is reproducible */
        char serr[256] = {0};
        if (gpet_sort_records(c->dw, pp.harvest, SURR_REC, 11, added, c->stream, serr, sizeof(serr)))
            return fail(c, GPET_ERR_DEVICE, std::string("sorting harvested records: ") + serr);
        HIPOK(c, hipStreamSynchronize(c->stream));
        c->surr.set_n += added;
    }

    /* background singles join the adder output before blur */
    int64_t n_noise = 0;
    if (c->dig.noise_rate > 0.0 && nw && c->has_det) {
        NoiseParams np;
        memset(&np, 0, sizeof(np));
        std::vector<NoiseJob> nj;
        gpet_noise_layout(c->dig.noise_rate, nw->t0.data(), nw->t1.data(), nw->first.data(), (int)nw->t0.size(), nj,
                          np.gap_us, np.chunk_us, np.nchunks);
        np.seed = c->opt.seed;
        np.e_mean = c->dig.noise_e_mean;
        np.e_sigma = c->dig.noise_e_sigma;
        np.jobs = nj.data();
        np.njobs = (int)nj.size();
        np.panels = c->ddet.p;
        np.npanels = c->ddet.npanels;
        np.moduleN = c->ddet.moduleN;
        np.crystalN = c->ddet.crystalN;
        np.moduleNy = c->ddet.moduleNy;
        np.crystalNy = c->ddet.crystalNy;
        np.depth = c->dig.readout_depth;
        np.policy = c->dig.readout_policy;
        char nerr[256] = {0};
        if (gpet_digi_noise_count(c->dw, np, &n_noise, c->stream, nerr, sizeof(nerr)))
            return fail(c, GPET_ERR_CAPACITY, std::string("background singles: ") + nerr);
        if (n_noise + c->n_events > (int64_t)INT32_MAX)
            return fail(c, GPET_ERR_CAPACITY, "more than 2^31 events with background singles");
        if ((st = dgrow(c, c->d_events, c->n_events + n_noise, c->n_events))) return st;
        if (gpet_digi_noise_write(c->dw, np, c->d_events.p + c->n_events, c->stream, nerr, sizeof(nerr)))
            return fail(c, GPET_ERR_DEVICE, std::string("background singles: ") + nerr);
    }

    DigiParams dp;
    memset(&dp, 0, sizeof(dp));
    dp.seed = c->opt.seed;
    dp.d = c->dig;
    dp.moduleN = c->ddet.moduleN;
    dp.crystalN = c->ddet.crystalN;
    const int njobs = batched ? (int)bi->job_first.size() - 1 : 1;
    if (batched) {
        if ((st = upload(c, c->d_job_first, bi->job_first.data(), (int64_t)bi->job_first.size()))) return st;
        if ((st = dalloc(c, c->d_job_cnt, 2 * (int64_t)(njobs + 1)))) return st;
        dp.job_first = c->d_job_first.p;
        dp.njobs = njobs;
        dp.job_bounds = c->d_job_cnt.p;
    }
    char err[256] = {0};
    /* the previous run's asynchronous sink copies read the buffers the
     * digitizer writes (or reallocates) next; the transport above has
     * overlapped them */
    if ((st = copy_wait(c))) return st;
    if (gpet_digitize(c->dw, dp, c->d_events.p, c->n_events + n_noise, &c->dout, c->stream, err, sizeof(err)))
        return fail(c, GPET_ERR_DEVICE, std::string("digitizer: ") + err);
    HIPOK(c, hipEventRecord(c->ev2, c->stream));
    HIPOK(c, hipEventSynchronize(c->ev2));
    float ms_t = 0, ms_d = 0;
    HIPOK(c, hipEventElapsedTime(&ms_t, c->ev0, c->ev1));
    HIPOK(c, hipEventElapsedTime(&ms_d, c->ev1, c->ev2));

    gpet_tallies &T = c->last;
    memset(&T, 0, sizeof(T));
    T.n_histories = n_hist;
    T.n_photons = (int64_t)cnt[CNT_PHOTONS];
    T.n_panel_entries = (int64_t)cnt[CNT_PANEL_ENTRIES];
    T.n_hits = (int64_t)cnt[CNT_HITS];
    T.n_adder = c->n_events;
    T.n_threshold = c->dout.n_threshold;
    T.n_deadtime = c->dout.n_deadtime;
    T.n_singles = c->dout.n_singles;
    T.n_coinc = c->dout.n_coinc;
    T.n_coinc_true = c->dout.n_coinc_true;
    T.dropped_events = (int64_t)cnt[CNT_DROPPED_EVENTS];
    T.dropped_hits = (int64_t)cnt[CNT_DROPPED_HITS];
    T.majorant_violations = (int64_t)cnt[CNT_MAJORANT];
    T.phantom_steps = (int64_t)cnt[CNT_PHANTOM_STEPS];
    T.range_steps = (int64_t)cnt[CNT_RANGE_STEPS];
    T.range_jumps = (int64_t)cnt[CNT_RANGE_JUMPS];
    T.panel_steps = (int64_t)cnt[CNT_PANEL_STEPS];
    T.seconds_transport = ms_t * 1e-3;
    T.seconds_digitizer = ms_d * 1e-3;
    /* eager launches carry event timing; graph replays do not */
    const int64_t n_timed = used_graph ? std::min<int64_t>(nbatch, 1) : nbatch;
    for (int64_t i = 0; i < n_timed; ++i) {
        float a = 0, b = 0;
        HIPOK(c, hipEventElapsedTime(&a, c->kev[4 * i], c->kev[4 * i + 1]));
        HIPOK(c, hipEventElapsedTime(&b, c->kev[4 * i + 2], c->kev[4 * i + 3]));
        T.kernel_seconds[0] += a * 1e-3;
        T.kernel_seconds[1] += b * 1e-3;
        T.kernel_histories += B[i].nh;
    }
    T.kernel_launches[0] = n_timed;
    T.kernel_launches[1] = n_timed;
    T.launches = nbatch;
    T.launch = L;
    T.graph = used_graph ? 1 : 0;
    T.n_jobs = njobs;
    T.seconds_tune = tune_s;
    T.n_surrogate = c->last_surr_items;
    T.seconds_surrogate = c->last_surr_seconds;
    T.n_noise = n_noise;
    T.n_delayed = c->dout.n_delayed;

    if (!sink) return GPET_OK;
    if (batched && sink->job_counts) {
        std::vector<unsigned long long> jb(2 * (size_t)(njobs + 1));
        HIPOK(c, hipMemcpy(jb.data(), c->d_job_cnt.p, sizeof(unsigned long long) * jb.size(), hipMemcpyDeviceToHost));
        for (int j = 0; j < njobs; ++j) {
            sink->job_counts[3 * j] = bi->job_first[j + 1] - bi->job_first[j];
            sink->job_counts[3 * j + 1] = (int64_t)(jb[j + 1] - jb[j]);
            sink->job_counts[3 * j + 2] = (int64_t)(jb[njobs + 1 + j + 1] - jb[njobs + 1 + j]);
        }
    }
    const bool to_files = sink->out_dir && sink->out_dir[0];
    std::vector<gpet_event> adder, singles, coinc, delayed;
    std::vector<int32_t> hid;
    std::vector<float> hf;
    sink->n_adder = c->n_events;
    sink->n_singles = c->dout.n_singles;
    sink->n_coinc = c->dout.n_coinc;
    sink->n_delayed = c->dout.n_delayed;
    sink->n_hits = std::min<int64_t>(small[2], hits_cap);
    if ((sink->adder && sink->adder_cap > 0) || to_files) {
        adder.resize((size_t)c->n_events);
        if (c->n_events)
            HIPOK(c, hipMemcpy(adder.data(), c->d_events.p, sizeof(gpet_event) * adder.size(), hipMemcpyDeviceToHost));
        std::sort(adder.begin(), adder.end(), ev_less_id);
        if (sink->adder)
            memcpy(sink->adder, adder.data(), sizeof(gpet_event) * (size_t)std::min<int64_t>(sink->adder_cap, c->n_events));
    }
    /* singles and pairs go straight into the caller's buffers (one copy;
     * with pinned buffers at full PCIe rate); files need a host copy.  With
     * async_copy the two copies go on the copy stream and the call returns
     * at once: they run beside the next run's transport */
    const bool async = sink->async_copy && !to_files &&
                       ((sink->singles && sink->singles_cap > 0 && c->dout.n_singles > 0) ||
                        (sink->coinc && sink->coinc_cap > 0 && c->dout.n_coinc > 0));
    if (async) {
        if (!c->copy_stream) {
            HIPOK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
            HIPOK(c, hipEventCreateWithFlags(&c->copy_done, hipEventDisableTiming));
        }
        HIPOK(c, hipStreamWaitEvent(c->copy_stream, c->ev2, 0));
    }
    if (to_files) {
        singles.resize((size_t)c->dout.n_singles);
        if (!singles.empty())
            HIPOK(c, hipMemcpy(singles.data(), c->dout.singles, sizeof(gpet_event) * singles.size(), hipMemcpyDeviceToHost));
        if (sink->singles && sink->singles_cap > 0)
            memcpy(sink->singles, singles.data(),
                   sizeof(gpet_event) * (size_t)std::min<int64_t>(sink->singles_cap, c->dout.n_singles));
    } else if (sink->singles && sink->singles_cap > 0) {
        const int64_t m = std::min<int64_t>(sink->singles_cap, c->dout.n_singles);
        if (m && async)
            HIPOK(c, hipMemcpyAsync(sink->singles, c->dout.singles, sizeof(gpet_event) * (size_t)m,
                                    hipMemcpyDeviceToHost, c->copy_stream));
        else if (m)
            HIPOK(c, hipMemcpy(sink->singles, c->dout.singles, sizeof(gpet_event) * (size_t)m, hipMemcpyDeviceToHost));
    }
    if (to_files) {
        coinc.resize((size_t)(2 * c->dout.n_coinc));
        if (!coinc.empty())
            HIPOK(c, hipMemcpy(coinc.data(), c->dout.coinc, sizeof(gpet_event) * coinc.size(), hipMemcpyDeviceToHost));
        if (sink->coinc && sink->coinc_cap > 0)
            memcpy(sink->coinc, coinc.data(),
                   sizeof(gpet_event) * (size_t)(2 * std::min<int64_t>(sink->coinc_cap, c->dout.n_coinc)));
    } else if (sink->coinc && sink->coinc_cap > 0) {
        const int64_t m = std::min<int64_t>(sink->coinc_cap, c->dout.n_coinc);
        if (m && async)
            HIPOK(c, hipMemcpyAsync(sink->coinc, c->dout.coinc, sizeof(gpet_event) * (size_t)(2 * m),
                                    hipMemcpyDeviceToHost, c->copy_stream));
        else if (m)
            HIPOK(c, hipMemcpy(sink->coinc, c->dout.coinc, sizeof(gpet_event) * (size_t)(2 * m), hipMemcpyDeviceToHost));
    }
    if (async) {
        HIPOK(c, hipEventRecord(c->copy_done, c->copy_stream));
        c->copy_pending = true;
    }
    if ((sink->delayed && sink->delayed_cap > 0) || to_files) {
        int64_t m = to_files ? c->dout.n_delayed : std::min<int64_t>(sink->delayed_cap, c->dout.n_delayed);
        delayed.resize((size_t)(2 * m));
        if (m)
            HIPOK(c, hipMemcpy(delayed.data(), c->dout.delayed, sizeof(gpet_event) * (size_t)(2 * m),
                               hipMemcpyDeviceToHost));
        if (sink->delayed)
            memcpy(sink->delayed, delayed.data(),
                   sizeof(gpet_event) * (size_t)(2 * std::min<int64_t>(sink->delayed_cap, m)));
    }
    if (want_hits && sink->n_hits > 0 && ((sink->hit_id && sink->hit_f) || to_files)) {
        hid.resize((size_t)(5 * sink->n_hits));
        hf.resize((size_t)(5 * sink->n_hits));
        HIPOK(c, hipMemcpy(hid.data(), c->d_hit_id.p, sizeof(int32_t) * hid.size(), hipMemcpyDeviceToHost));
        HIPOK(c, hipMemcpy(hf.data(), c->d_hit_f.p, sizeof(float) * hf.size(), hipMemcpyDeviceToHost));
        if (sink->hit_id && sink->hit_f) {
            int64_t m = std::min<int64_t>(sink->hits_cap, sink->n_hits);
            memcpy(sink->hit_id, hid.data(), sizeof(int32_t) * (size_t)(5 * m));
            memcpy(sink->hit_f, hf.data(), sizeof(float) * (size_t)(5 * m));
        }
    }
    if (want_trace) {
        sink->n_trace = ntr;
        HIPOK(c, hipMemcpy(sink->trace, c->d_trace.p, sizeof(gpet_trace) * (size_t)ntr, hipMemcpyDeviceToHost));
    }
    /* phase-space records, ordered by photon id */
    if (rp.n_psf) {
        unsigned long long np[2] = {0, 0};
        HIPOK(c, hipMemcpy(np, rp.n_psf, sizeof(np), hipMemcpyDeviceToHost));
        gpet_psf_record *dev[2] = {rp.psf_src, rp.psf_ph};
        gpet_psf_record *dst[2] = {sink->psf_source, sink->psf_phantom};
        const int64_t caps[2] = {sink->psf_source_cap, sink->psf_phantom_cap};
        int64_t *ns[2] = {&sink->n_psf_source, &sink->n_psf_phantom};
        const char *names[2] = {"source", "phantom"};
        for (int k = 0; k < 2; ++k) {
            if (!dev[k]) continue;
            const int64_t m = std::min<int64_t>((int64_t)np[k], psf_cap);
            *ns[k] = m;
            if (!((dst[k] && caps[k] > 0) || to_files)) continue;
            std::vector<gpet_psf_record> v((size_t)m);
            if (m) HIPOK(c, hipMemcpy(v.data(), dev[k], sizeof(gpet_psf_record) * (size_t)m, hipMemcpyDeviceToHost));
            std::sort(v.begin(), v.end(),
                      [](const gpet_psf_record &a, const gpet_psf_record &b) { return a.parn < b.parn; });
            if (dst[k]) memcpy(dst[k], v.data(), sizeof(gpet_psf_record) * (size_t)std::min<int64_t>(m, caps[k]));
            if (to_files) {
                gpet_status s2 = gpet_write_psf(sink->out_dir, names[k], v.data(), m, sink->out_append);
                if (s2) return fail(c, s2, std::string("cannot write PSF files in ") + sink->out_dir);
            }
        }
    }
    if (to_files) return write_outputs(c, sink->out_dir, adder, singles, coinc, hid, hf, delayed, sink->out_append);
    return GPET_OK;
}

/* ---- scratch memory of gpet_run_psf_noalloc ------------------------------
 * A file-scope device array holds everything that entry point's kernels need
 * beyond its inputs and outputs: the transport counters, the panel queue's
 * region counts and the panel queue itself.  Each device that loads the
 * library holds one copy until the process ends; every call on a device
 * shares it (calls that overlap on one device race).  Host code uses the
 * address that hipGetSymbolAddress returns for the current device, never
 * the array's name. */
constexpr int64_t NA_MAX = GPET_NOALLOC_MAX_RECORDS;
/* regions: one per phantom-kernel block, at most ceil(n / 64) + 1 */
constexpr int64_t NA_REGIONS = NA_MAX / 64 + 2;
/* queue slots: G x region < 2 (n + G x block) <= 4 n + 2 x 256 */
constexpr int64_t NA_QCAP = 4 * NA_MAX + 2 * 256;
constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_OFF_CNT = 0;
constexpr size_t NA_OFF_SMALL = NA_OFF_CNT + na_round(sizeof(unsigned long long) * CNT_NUM * CNT_STRIPES);
constexpr size_t NA_OFF_RCOUNT = NA_OFF_SMALL + 256;
constexpr size_t NA_OFF_QF = NA_OFF_RCOUNT + na_round(sizeof(unsigned) * (size_t)NA_REGIONS);
constexpr size_t NA_OFF_QT = NA_OFF_QF + 6 * na_round(sizeof(float) * (size_t)NA_QCAP);
constexpr size_t NA_OFF_QINFO = NA_OFF_QT + na_round(sizeof(double) * (size_t)NA_QCAP);
constexpr size_t NA_BYTES = NA_OFF_QINFO + na_round(sizeof(int4) * (size_t)NA_QCAP);

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
void *na_ptr[NA_MAX_DEVICES] = {nullptr};

/* Device address of `scratch` on the current device `dev`, looked up once
 * per device. */
hipError_t na_scratch_address(int dev, void **out)
{
    if (dev < 0 || dev >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        na_ptr[dev] = p;
    }
    *out = na_ptr[dev];
    return hipSuccess;
}

} // namespace

extern "C" {

gpet_status gpet_create(gpet_ctx **out, const gpet_options *opt)
{
    if (!out || !opt) return GPET_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GPET_ERR_DEVICE;
    if (opt->device < 0 || opt->device >= ndev) return GPET_ERR_ARG;
    gpet_ctx *c = new gpet_ctx();
    c->opt = *opt;
    c->device = opt->device;
    memset(&c->dig, 0, sizeof(c->dig));
    memset(&c->last, 0, sizeof(c->last));
    memset(&c->dtab, 0, sizeof(c->dtab));
    memset(&c->dph, 0, sizeof(c->dph));
    memset(&c->ddet, 0, sizeof(c->ddet));
    c->launch = default_launch();
    if (opt->batch > 0) c->launch.batch = std::min<int64_t>(std::max<int64_t>(opt->batch, 256), (int64_t)1 << 26);
    c->slots.resize(1);
    memset(&c->slots[0].dp, 0, sizeof(DevPhantom));
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess) {
        delete c;
        return GPET_ERR_DEVICE;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0)
        c->n_cu = ncu;
    c->dw = gpet_digi_work_create();
    /* the allocation-free entry point's scratch address on this device */
    if (na_scratch_address(c->device, &c->na_scratch) != hipSuccess) {
        (void)hipGetLastError();
        c->na_scratch = nullptr;
    }
    *out = c;
    return GPET_OK;
}

gpet_status gpet_destroy(gpet_ctx *c)
{
    if (!c) return GPET_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_pending) (void)hipEventSynchronize(c->copy_done);
    dfree(c->d_mu4); dfree(c->d_surf[0]); dfree(c->d_surf[1]);
    for (PhSlot &sl : c->slots) {
        dfree(sl.vox); dfree(sl.dens_b); dfree(sl.wck); dfree(sl.wmap);
    }
    dfree(c->d_phs); dfree(c->d_seg); dfree(c->d_job_first); dfree(c->d_job_cnt);
    drop_graph(c);
    surr_free(c);
    dfree(c->d_panels); dfree(c->d_wck_det); dfree(c->d_pgroup_sph); dfree(c->d_pgroup_mask);
    for (auto &b : c->q_f) dfree(b);
    dfree(c->q_t); dfree(c->q_info); dfree(c->q_slot); dfree(c->d_small); dfree(c->d_cnt);
    dfree(c->d_rcount); dfree(c->d_lw); dfree(c->d_lw_count);
    for (int k = 0; k < 2; ++k) { dfree(c->d_tq[k]); dfree(c->d_tq_count[k]); }
    dfree(c->d_events); dfree(c->d_trace); dfree(c->d_hit_id); dfree(c->d_hit_f);
    dfree(c->d_psf_src); dfree(c->d_psf_ph); dfree(c->d_npsf);
    dfree(c->d_first); dfree(c->d_src); dfree(c->d_iso); dfree(c->d_psf);
    gpet_digi_work_destroy(c->dw);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    for (hipEvent_t e : c->kev) (void)hipEventDestroy(e);
    if (c->copy_done) (void)hipEventDestroy(c->copy_done);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GPET_OK;
}

const char *gpet_last_error(gpet_ctx *c) { return c ? c->err.c_str() : "null context"; }

gpet_status gpet_set_tables(gpet_ctx *c, const gpet_tables *t)
{
    if (!c || !t) return GPET_ERR_ARG;
    if (t->nmat <= 0 || t->nmat > GPET_MAXMAT || t->n1d != GPET_NTAB || t->ncp != GPET_NCP || t->ne != GPET_NE)
        return fail(c, GPET_ERR_FORMAT, "table dimensions");
    for (int k : {GPET_TAB_COMPT, GPET_TAB_RAYLE})
        if (t->e0[k] != t->e0[GPET_TAB_LAMPH] || t->idl[k] != t->idl[GPET_TAB_LAMPH])
            return fail(c, GPET_ERR_FORMAT, "lamph/compt/rayle must share one energy grid");
    HIPOK(c, hipSetDevice(c->device));
    c->nmat = t->nmat;
    c->eminph = t->eminph;
    c->emax = t->emax;
    c->e0 = t->e0[GPET_TAB_LAMPH];
    c->idl = t->idl[GPET_TAB_LAMPH];
    const size_t n = (size_t)t->nmat * t->n1d;
    c->h_lamph.assign(t->mu[GPET_TAB_LAMPH], t->mu[GPET_TAB_LAMPH] + n);
    std::vector<float4> mu4(n);
    for (size_t i = 0; i < n; ++i)
        mu4[i] = make_float4(t->mu[GPET_TAB_LAMPH][i], t->mu[GPET_TAB_COMPT][i], t->mu[GPET_TAB_RAYLE][i], 0.0f);
    gpet_status st = upload(c, c->d_mu4, mu4.data(), (int64_t)n);
    if (st) return st;
    const int64_t ns = (int64_t)t->nmat * t->ncp * t->ne;
    for (int k = 0; k < 2; ++k)
        if ((st = upload(c, c->d_surf[k], t->surf[k], ns))) return st;
    DevTables &d = c->dtab;
    d.nmat = t->nmat; d.n1d = t->n1d; d.ncp = t->ncp; d.ne = t->ne;
    d.e0 = c->e0; d.idl = c->idl;
    d.mu4 = c->d_mu4.p;
    for (int k = 0; k < 2; ++k) {
        d.idcp[k] = t->idcp[k];
        d.ide[k] = t->ide[k];
        d.surf[k] = c->d_surf[k].p;
    }
    c->has_tab = true;
    return GPET_OK;
}

gpet_status gpet_load_tables(gpet_ctx *c, const char *dir)
{
    if (!c || !dir) return GPET_ERR_ARG;
    gpet_tables t;
    gpet_status st = gpet_tables_load(dir, &t);
    if (st) return fail(c, st, std::string("cannot load tables from ") + dir);
    st = gpet_set_tables(c, &t);
    gpet_tables_free(&t);
    return st;
}

gpet_status gpet_set_phantom_slot(gpet_ctx *c, int32_t slot, const int32_t *mat, const float *dens,
                                  const int32_t dims[3], const float off[3], const float size[3])
{
    if (!c || !mat || !dens || !dims || !off || !size) return GPET_ERR_ARG;
    if (slot < 0 || slot >= GPET_MAXPHANTOMS) return fail(c, GPET_ERR_ARG, "phantom slot out of range");
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables must be set before the phantom");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return fail(c, GPET_ERR_ARG, "phantom dims");
    HIPOK(c, hipSetDevice(c->device));
    if ((int)c->slots.size() <= slot) {
        size_t n0 = c->slots.size();
        c->slots.resize((size_t)slot + 1);
        for (size_t k = n0; k < c->slots.size(); ++k) memset(&c->slots[k].dp, 0, sizeof(DevPhantom));
    }
    {
        /* the positron walk addresses the bricked densities in 32 bits */
        const int64_t nb = (int64_t)gpet_padded_bricks(dims[0]) * gpet_padded_bricks(dims[1]) *
                           gpet_padded_bricks(dims[2]);
        if (nb >= ((int64_t)1 << 24)) return fail(c, GPET_ERR_CAPACITY, "phantom above 2^24 bricks of 4^3 voxels");
    }
    PhSlot &S = c->slots[slot];
    const size_t nv = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<int2> vox(nv);
    float rmax[GPET_MAXMAT] = {0};
    for (size_t i = 0; i < nv; ++i) {
        int m = mat[i];
        if (m < 0 || m >= c->nmat) return fail(c, GPET_ERR_ARG, "phantom material out of range");
        if (dens[i] < 0.0f) return fail(c, GPET_ERR_ARG, "negative phantom density");
        if (dens[i] > rmax[m]) rmax[m] = dens[i];
        int b;
        memcpy(&b, &dens[i], 4);
        vox[i] = make_int2(b, m);
    }
    gpet_status st = upload(c, S.vox, vox.data(), (int64_t)nv);
    if (st) return st;
    const int nbx = gpet_padded_bricks(dims[0]), nby = gpet_padded_bricks(dims[1]), nbz = gpet_padded_bricks(dims[2]);
    {
        /* padded bricked copy for the positron walk: the border (and the
         * bricks' slack beyond it) holds the negative outside density */
        std::vector<float> db((size_t)nbx * nby * nbz * 64, -D_RHO_OUTSIDE);
        for (int iz = 0; iz < dims[2]; ++iz)
            for (int iy = 0; iy < dims[1]; ++iy)
                for (int ix = 0; ix < dims[0]; ++ix)
                    db[gpet_brick_offset(ix + 1, iy + 1, iz + 1, nbx, nby)] =
                        dens[((size_t)iz * dims[1] + iy) * dims[0] + ix];
        if ((st = upload(c, S.dens_b, db.data(), (int64_t)db.size()))) return st;
    }
    /* uniform boxes of the positron walk (none: no map) */
    std::vector<uint32_t> wm;
    const bool has_map = gp_build_walkmap(dims[0], dims[1], dims[2], dens, wm);
    if (has_map) {
        if ((st = upload(c, S.wmap, reinterpret_cast<const uint4 *>(wm.data()), (int64_t)(wm.size() / 4)))) return st;
    } else {
        dfree(S.wmap);
    }
    DevPhantom &p = S.dp;
    memset(&p, 0, sizeof(p));
    p.wmap = has_map ? S.wmap.p : nullptr;
    p.wmap_bytes = has_map ? (unsigned)(wm.size() * sizeof(uint32_t)) : 0u;
    p.dens_b = S.dens_b.p;
    p.nbx = nbx;
    p.nby = nby;
    p.dens_bytes = (unsigned)((size_t)nbx * nby * nbz * 64 * sizeof(float));
    build_majorant(c, rmax, S.h_wck, p.wck.e0, p.wck.idl);
    if ((st = upload(c, S.wck, S.h_wck.data(), GPET_NWCK))) return st;
    p.wck.lam = S.wck.p;
    p.nx = dims[0]; p.ny = dims[1]; p.nz = dims[2];
    p.offx = off[0]; p.offy = off[1]; p.offz = off[2];
    p.dx = size[0] / (float)p.nx; p.dy = size[1] / (float)p.ny; p.dz = size[2] / (float)p.nz;
    p.idx = 1.0f / p.dx; p.idy = 1.0f / p.dy; p.idz = 1.0f / p.dz;
    p.vox = S.vox.p;
    p.present = 1;
    S.present = true;
    c->phs_dirty = true;
    if (slot == 0) {
        c->dph = p;
        c->has_ph = true;
    }
    return GPET_OK;
}

gpet_status gpet_set_phantom(gpet_ctx *c, const int32_t *mat, const float *dens, const int32_t dims[3],
                             const float off[3], const float size[3])
{
    return gpet_set_phantom_slot(c, 0, mat, dens, dims, off, size);
}

/* Groups of panels for the entry test of large scanners (16 < n <= 64):
 * panels sorted by the azimuth of their face centre (then z, then index)
 * are cut into runs of s, and each group's sphere encloses its members'
 * entry spheres (|c_g - c_i| + r_i, in double, rounded up).  A line that
 * misses a group sphere misses every member sphere, so skipping the group
 * skips only panels the per-panel prefilter would skip.  s is chosen to
 * minimise groups + sum over groups of members x (r_g / |c_g|)^2, the
 * expected tests for directions spread evenly over the scanner. */
static void panel_groups(const std::vector<DevPanel> &dp, std::vector<float4> &gs,
                         std::vector<unsigned long long> &gm)
{
    const int n = (int)dp.size();
    gs.clear();
    gm.clear();
    if (n <= 16 || n > 64) return;
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::vector<double> az(n);
    for (int i = 0; i < n; ++i) az[i] = std::atan2((double)dp[i].oy, (double)dp[i].ox);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (std::fabs(az[a] - az[b]) > 1e-6) return az[a] < az[b];
        if (dp[a].oz != dp[b].oz) return dp[a].oz < dp[b].oz;
        return a < b;
    });
    auto build = [&](int s, std::vector<float4> &S, std::vector<unsigned long long> &M) {
        S.clear();
        M.clear();
        double cost = 0;
        for (int g0 = 0; g0 < n; g0 += s) {
            const int g1 = std::min(n, g0 + s);
            double cx = 0, cy = 0, cz = 0;
            for (int k = g0; k < g1; ++k) {
                cx += dp[order[k]].ox; cy += dp[order[k]].oy; cz += dp[order[k]].oz;
            }
            cx /= (g1 - g0); cy /= (g1 - g0); cz /= (g1 - g0);
            double r = 0;
            unsigned long long m = 0;
            for (int k = g0; k < g1; ++k) {
                const DevPanel &P = dp[order[k]];
                const double dx = P.ox - cx, dy = P.oy - cy, dz = P.oz - cz;
                r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz) + std::sqrt((double)P.bound2));
                m |= 1ull << order[k];
            }
            r = r * (1.0 + 1e-6) + 1e-4;
            S.push_back(make_float4((float)cx, (float)cy, (float)cz, (float)(r * r * (1.0 + 1e-6))));
            M.push_back(m);
            const double c2 = std::max(cx * cx + cy * cy + cz * cz, r * r);
            cost += (double)(g1 - g0) * (r * r) / c2;
        }
        return (double)S.size() + cost;
    };
    double best = 1e300;
    std::vector<float4> S;
    std::vector<unsigned long long> M;
    for (int s = 2; s <= 8; ++s) {
        const double cost = build(s, S, M);
        if (cost < best) {
            best = cost;
            gs = S;
            gm = M;
        }
    }
    /* no grouping unless it beats testing every panel */
    if (!(best < (double)n)) {
        gs.clear();
        gm.clear();
    }
}

gpet_status gpet_set_detector(gpet_ctx *c, const gpet_panel *panels, int32_t n, const gpet_panel_material mats[2],
                              const float *quad, int32_t nq)
{
    if (!c || !panels || n <= 0 || !mats) return GPET_ERR_ARG;
    if (!c->has_tab) return fail(c, GPET_ERR_STATE, "tables must be set before the detector");
    if (n > GPET_MAXPANELS) return fail(c, GPET_ERR_CAPACITY, "too many panels");
    if (nq < 0 || nq > GPET_MAXSURFACE) return fail(c, GPET_ERR_CAPACITY, "too many surfaces");
    HIPOK(c, hipSetDevice(c->device));
    c->panels.assign(panels, panels + n);
    std::vector<DevPanel> dp(n);
    for (int i = 0; i < n; ++i) {
        const gpet_panel &s = panels[i];
        DevPanel &d = dp[i];
        d.panel = s.panel;
        d.lenx = s.lenx; d.leny = s.leny; d.lenz = s.lenz;
        d.mody = s.mody; d.modz = s.modz; d.mspy = s.mspy; d.mspz = s.mspz;
        d.cry = s.cry; d.crz = s.crz; d.cspy = s.cspy; d.cspz = s.cspz;
        d.ox = s.ox; d.oy = s.oy; d.oz = s.oz; d.dirx = s.dirx;
        for (int k = 0; k < 3; ++k) { d.ux[k] = s.ux[k]; d.uy[k] = s.uy[k]; d.uz[k] = s.uz[k]; }
        double r2 = 0.25 * ((double)s.leny * s.leny + (double)s.lenz * s.lenz);
        d.bound2 = (float)(r2 * 1.002 + 0.01);
    }
    gpet_status st = upload(c, c->d_panels, dp.data(), n);
    if (st) return st;
    DevDetector &d = c->ddet;
    d.npanels = n;
    d.p = c->d_panels.p;
    d.ngroups = 0;
    d.gsph = nullptr;
    d.gmask = nullptr;
    {
        std::vector<float4> gs;
        std::vector<unsigned long long> gm;
        panel_groups(dp, gs, gm);
        if (!gs.empty()) {
            if ((st = upload(c, c->d_pgroup_sph, gs.data(), (int64_t)gs.size()))) return st;
            if ((st = upload(c, c->d_pgroup_mask, gm.data(), (int64_t)gm.size()))) return st;
            d.ngroups = (int)gs.size();
            d.gsph = c->d_pgroup_sph.p;
            d.gmask = c->d_pgroup_mask.p;
        }
    }
    float rmax[GPET_MAXMAT] = {0};
    for (int i = 0; i < 2; ++i) {
        int m = mats[i].material;
        if (m < 0 || m >= c->nmat) return fail(c, GPET_ERR_ARG, "panel material out of range");
        d.mat[i] = m;
        d.dens[i] = mats[i].density;
        if (mats[i].density > rmax[m]) rmax[m] = mats[i].density;
    }
    d.nquad = (quad && nq > 0) ? nq : 0;
    memset(d.quad, 0, sizeof(d.quad));
    if (d.nquad) memcpy(d.quad, quad, sizeof(float) * 10 * nq);
    gpet_panel_counts(&panels[0], &d.moduleN, &d.crystalN, &d.moduleNy, &d.crystalNy);
    build_majorant(c, rmax, c->h_wck_det, d.wck.e0, d.wck.idl);
    if ((st = upload(c, c->d_wck_det, c->h_wck_det.data(), GPET_NWCK))) return st;
    d.wck.lam = c->d_wck_det.p;
    d.present = 1;
    c->has_det = true;
    return GPET_OK;
}

gpet_status gpet_set_digitizer(gpet_ctx *c, const gpet_digitizer *d)
{
    if (!c || !d) return GPET_ERR_ARG;
    if (d->readout_depth < 0 || d->readout_depth > 3 || d->dead_level < 0 || d->dead_level > 3)
        return fail(c, GPET_ERR_ARG, "digitizer depth/level out of range");
    c->dig = *d;
    c->has_dig = true;
    return GPET_OK;
}

gpet_status gpet_run_source(gpet_ctx *c, const gpet_source *src, int32_t ns, const gpet_isotope *iso, int32_t niso,
                            double t0, double t1, int64_t n_decays, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!src || ns <= 0 || !iso || niso <= 0 || !(t1 >= t0)) return fail(c, GPET_ERR_ARG, "source arguments");
    HIPOK(c, hipSetDevice(c->device));
    std::vector<int64_t> cnt;
    std::vector<DevSource> ds;
    if ((st = plan_job(c, src, ns, iso, niso, t0, t1, n_decays, 0, 0, cnt, ds))) return st;
    std::vector<int64_t> first(ns + 1, 0);
    for (int i = 0; i < ns; ++i) first[i + 1] = first[i] + cnt[i];
    if ((st = upload_run_sources(c, first, ds, iso, niso))) return st;
    RunParams rp;
    memset(&rp, 0, sizeof(rp));
    rp.mode = RUN_DECAY;
    rp.nsrc = ns;
    rp.first = c->d_first.p;
    rp.src = c->d_src.p;
    rp.iso = c->d_iso.p;
    rp.psf = nullptr;
    NoiseWindows nw;
    nw.t0.assign(1, t0);
    nw.t1.assign(1, t1);
    nw.first.assign(1, 0);
    return run_common(c, rp, first[ns], sink, nullptr, &nw);
}

gpet_status gpet_run_batch(gpet_ctx *c, const gpet_job *jobs, int32_t njobs, const gpet_source *src, int32_t nsrc,
                           const gpet_isotope *iso, int32_t niso, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!jobs || njobs <= 0 || !src || nsrc <= 0 || !iso || niso <= 0) return fail(c, GPET_ERR_ARG, "batch arguments");
    HIPOK(c, hipSetDevice(c->device));
    BatchInfo bi;
    std::vector<int64_t> cnt;
    std::vector<DevSource> ds;
    bi.job_first.assign(1, 0);
    for (int j = 0; j < njobs; ++j) {
        const gpet_job &J = jobs[j];
        if (J.first_source < 0 || J.n_sources <= 0 || J.first_source + J.n_sources > nsrc || !(J.t1 >= J.t0))
            return fail(c, GPET_ERR_ARG, "job source range or time window");
        if (J.phantom < 0 || J.phantom >= (int)c->slots.size() || !c->slots[J.phantom].present)
            return fail(c, GPET_ERR_STATE, "job phantom slot not set");
        const size_t before = cnt.size();
        if ((st = plan_job(c, src + J.first_source, J.n_sources, iso, niso, J.t0, J.t1, J.n_decays, J.first_source,
                           J.phantom, cnt, ds)))
            return st;
        int64_t tot = 0;
        for (size_t k = before; k < cnt.size(); ++k) {
            tot += cnt[k];
            bi.src_phantom.push_back(J.phantom);
        }
        bi.job_first.push_back(bi.job_first.back() + tot);
    }
    const int ns = (int)cnt.size();
    bi.first.assign(ns + 1, 0);
    for (int i = 0; i < ns; ++i) bi.first[i + 1] = bi.first[i] + cnt[i];
    if ((st = upload_run_sources(c, bi.first, ds, iso, niso))) return st;
    if (c->phs_dirty) {
        std::vector<DevPhantom> dp(c->slots.size());
        for (size_t k = 0; k < dp.size(); ++k) dp[k] = c->slots[k].dp;
        if ((st = upload(c, c->d_phs, dp.data(), (int64_t)dp.size()))) return st;
        c->phs_dirty = false;
    }
    RunParams rp;
    memset(&rp, 0, sizeof(rp));
    rp.mode = RUN_DECAY;
    rp.nsrc = ns;
    rp.first = c->d_first.p;
    rp.src = c->d_src.p;
    rp.iso = c->d_iso.p;
    rp.psf = nullptr;
    NoiseWindows nw;
    for (int j = 0; j < njobs; ++j) {
        nw.t0.push_back(jobs[j].t0);
        nw.t1.push_back(jobs[j].t1);
        nw.first.push_back(bi.job_first[j]);
    }
    return run_common(c, rp, bi.first[ns], sink, &bi, &nw);
}

gpet_status gpet_set_launch(gpet_ctx *c, const gpet_launch *l)
{
    if (!c || !l) return GPET_ERR_ARG;
    gpet_status st = check_launch(c, *l);
    if (st) return st;
    c->launch = *l;
    return GPET_OK;
}

gpet_status gpet_get_launch(gpet_ctx *c, gpet_launch *l)
{
    if (!c || !l) return GPET_ERR_ARG;
    *l = c->launch;
    return GPET_OK;
}

gpet_status gpet_run_psf(gpet_ctx *c, const double *rec8, int64_t n, int32_t ptype, gpet_sink *sink)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!rec8 || n < 0 || (ptype != 0 && ptype != 1)) return fail(c, GPET_ERR_ARG, "psf arguments");
    HIPOK(c, hipSetDevice(c->device));
    if ((st = upload(c, c->d_psf, rec8, 8 * n))) return st;
    RunParams rp;
    memset(&rp, 0, sizeof(rp));
    rp.mode = ptype == 0 ? RUN_PSF_POSITRON : RUN_PSF_PHOTON;
    rp.psf = c->d_psf.p;
    return run_common(c, rp, n, sink, nullptr, nullptr);
}

gpet_status gpet_run_psf_noalloc(gpet_ctx *c, const double *d_rec8, int64_t n, int32_t ptype, gpet_event *d_events,
                                 int64_t events_cap, uint32_t *d_n_events, void *stream)
{
    gpet_status st = check_ready(c);
    if (st) return st;
    if (!d_rec8 || !d_events || !d_n_events || n <= 0 || events_cap <= 0 || (ptype != 0 && ptype != 1))
        return fail(c, GPET_ERR_ARG, "psf_noalloc arguments");
    if (n > NA_MAX)
        return fail(c, GPET_ERR_CAPACITY, "psf_noalloc: more than GPET_NOALLOC_MAX_RECORDS (65536) records");
    if (events_cap > (int64_t)UINT32_MAX) events_cap = (int64_t)UINT32_MAX;
    HIPOK(c, hipSetDevice(c->device));
    if (!c->na_scratch) {
        /* the first context on this device could not look it up */
        const hipError_t e = na_scratch_address(c->device, &c->na_scratch);
        if (e != hipSuccess) return fail(c, GPET_ERR_DEVICE, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    }
    unsigned char *base = static_cast<unsigned char *>(c->na_scratch);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;

    /* launch shape: the context's, without photon regrouping; the queue
     * shape it implies must fit the scratch queue */
    gpet_launch L = c->launch;
    memset(L.transit, 0, sizeof(L.transit));
    int G = 0;
    int64_t region = 0;
    queue_shape(n, L, c->n_cu, G, region);
    if ((int64_t)G + 1 > NA_REGIONS || (int64_t)G * region > NA_QCAP)
        return fail(c, GPET_ERR_CAPACITY, "psf_noalloc: queue shape exceeds the scratch queue");

    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base + NA_OFF_CNT);
    unsigned *small = reinterpret_cast<unsigned *>(base + NA_OFF_SMALL);
    DevQueue q;
    memset(&q, 0, sizeof(q));
    float *qf = reinterpret_cast<float *>(base + NA_OFF_QF);
    const size_t fstride = na_round(sizeof(float) * (size_t)NA_QCAP) / sizeof(float);
    q.y = qf; q.z = qf + fstride; q.vx = qf + 2 * fstride; q.vy = qf + 3 * fstride; q.vz = qf + 4 * fstride;
    q.E = qf + 5 * fstride;
    q.t = reinterpret_cast<double *>(base + NA_OFF_QT);
    q.info = reinterpret_cast<int4 *>(base + NA_OFF_QINFO);
    q.count = reinterpret_cast<int *>(small);
    q.cap = NA_QCAP;
    q.rcount = reinterpret_cast<unsigned *>(base + NA_OFF_RCOUNT);
    q.nregions = G;
    q.region = region;

    RunParams rp;
    memset(&rp, 0, sizeof(rp));
    rp.mode = ptype == 0 ? RUN_PSF_POSITRON : RUN_PSF_PHOTON;
    rp.psf = d_rec8;
    rp.h0 = 0;
    rp.nh = n;
    rp.seed = c->opt.seed;
    rp.eabs = c->opt.eabs;
    rp.sigma = c->opt.noncol_sigma;
    rp.prange = c->opt.positron_range;
    rp.cnt = cnt;
    for (int k = 0; k < 4; ++k) rp.sphere[k] = c->opt.record_sphere[k];

    PanelParams pp;
    memset(&pp, 0, sizeof(pp));
    pp.seed = c->opt.seed;
    pp.eabs = c->opt.eabs;
    pp.readout_depth = c->dig.readout_depth;
    pp.readout_policy = c->dig.readout_policy;
    pp.record_rayleigh = c->dig.record_rayleigh;
    pp.events = d_events;
    pp.events_cap = events_cap;
    pp.n_events = d_n_events;
    pp.n_hits = small + 2;
    pp.cnt = cnt;

    /* clear the counters with kernels (no memset nodes if the caller
     * captures the stream) */
    int e = gpet_launch_zero(reinterpret_cast<unsigned *>(cnt), 2 * CNT_NUM * CNT_STRIPES, s);
    if (!e) e = gpet_launch_zero(small, 4, s);
    if (!e) e = gpet_launch_zero(d_n_events, 1, s);
    if (e) return fail(c, GPET_ERR_DEVICE, "psf_noalloc: zero kernel");
    e = gpet_launch_phantom(rp, c->dtab, c->dph, c->ddet, q, false, cfg_of(L, 0), c->n_cu, s);
    if (e) return fail(c, GPET_ERR_DEVICE, std::string("phantom kernel: ") + hipGetErrorString((hipError_t)e));
    e = gpet_launch_panel(pp, c->dtab, c->ddet, q, 2 * n, cfg_of(L, 1), c->n_cu, s, PANEL_PLAIN);
    if (e) return fail(c, GPET_ERR_DEVICE, std::string("panel kernel: ") + hipGetErrorString((hipError_t)e));
    return GPET_OK;
}

gpet_status gpet_sink_wait(gpet_ctx *c)
{
    if (!c) return GPET_ERR_ARG;
    HIPOK(c, hipSetDevice(c->device));
    return copy_wait(c);
}

gpet_status gpet_get_tallies(gpet_ctx *c, gpet_tallies *t)
{
    if (!c || !t) return GPET_ERR_ARG;
    int64_t *cs = t->crystal_singles, *eh = t->energy_hist, *sg = t->sinogram, *sgd = t->sinogram_delayed;
    int32_t nc = t->n_crystals, nb = t->n_ebins, nsb = t->n_s, nph = t->n_phi;
    float elo = t->e_lo, ehi = t->e_hi, smax = t->s_max;
    *t = c->last;
    t->crystal_singles = cs; t->n_crystals = nc;
    t->energy_hist = eh; t->n_ebins = nb; t->e_lo = elo; t->e_hi = ehi;
    t->sinogram = sg; t->n_s = nsb; t->n_phi = nph; t->s_max = smax;
    t->sinogram_delayed = sgd;
    if ((cs && nc > 0) || (eh && nb > 0)) {
        HIPOK(c, hipSetDevice(c->device));
        DBuf<unsigned long long> dc, dh;
        gpet_status st;
        if (cs && nc > 0) {
            if ((st = dalloc(c, dc, nc))) return st;
            HIPOK(c, hipMemsetAsync(dc.p, 0, sizeof(unsigned long long) * nc, c->stream));
        }
        if (eh && nb > 0) {
            if ((st = dalloc(c, dh, nb))) return st;
            HIPOK(c, hipMemsetAsync(dh.p, 0, sizeof(unsigned long long) * nb, c->stream));
        }
        int e = gpet_digi_tallies(c->dw, c->dout.singles, c->dout.n_singles, c->ddet.moduleN, c->ddet.crystalN,
                                  dc.p, (cs && nc > 0) ? nc : 0, dh.p, (eh && nb > 0) ? nb : 0, elo, ehi, c->stream);
        if (e) return fail(c, GPET_ERR_DEVICE, "tally kernel");
        if (dc.p) HIPOK(c, hipMemcpyAsync(cs, dc.p, sizeof(int64_t) * nc, hipMemcpyDeviceToHost, c->stream));
        if (dh.p) HIPOK(c, hipMemcpyAsync(eh, dh.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, c->stream));
        HIPOK(c, hipStreamSynchronize(c->stream));
        dfree(dc);
        dfree(dh);
    }
    for (int k = 0; k < 2; ++k) {
        int64_t *dst = k == 0 ? sg : sgd;
        if (!(dst && nsb > 0 && nph > 0 && smax > 0.0f && c->has_det)) continue;
        DBuf<unsigned long long> ds;
        gpet_status st;
        const int64_t nb = (int64_t)nsb * nph;
        if ((st = dalloc(c, ds, nb))) return st;
        HIPOK(c, hipMemsetAsync(ds.p, 0, sizeof(unsigned long long) * nb, c->stream));
        int e = gpet_digi_sinogram(k == 0 ? c->dout.coinc : c->dout.delayed, k == 0 ? c->dout.n_coinc : c->dout.n_delayed,
                                   c->ddet.p, c->ddet.npanels, c->ddet.moduleNy, c->ddet.crystalNy, ds.p, nsb, nph,
                                   smax, c->stream);
        if (e) return fail(c, GPET_ERR_DEVICE, "sinogram kernel");
        HIPOK(c, hipMemcpyAsync(dst, ds.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, c->stream));
        HIPOK(c, hipStreamSynchronize(c->stream));
        dfree(ds);
    }
    return GPET_OK;
}

/* Host-only: the majorant gpet_set_phantom/gpet_set_detector would build
 * for these tables and per-material maximum densities. */
gpet_status gpet_build_majorant(const gpet_tables *t, const float *rho_max, float *out, float *e0, float *idl)
{
    if (!t || !rho_max || !out || t->nmat <= 0 || t->nmat > GPET_MAXMAT) return GPET_ERR_ARG;
    gpet_ctx tmp;
    tmp.nmat = t->nmat;
    tmp.eminph = t->eminph;
    tmp.emax = t->emax;
    tmp.e0 = t->e0[GPET_TAB_LAMPH];
    tmp.idl = t->idl[GPET_TAB_LAMPH];
    tmp.h_lamph.assign(t->mu[GPET_TAB_LAMPH], t->mu[GPET_TAB_LAMPH] + (size_t)t->nmat * t->n1d);
    std::vector<float> lam;
    float a = 0, b = 0;
    build_majorant(&tmp, rho_max, lam, a, b);
    memcpy(out, lam.data(), sizeof(float) * GPET_NWCK);
    if (e0) *e0 = a;
    if (idl) *idl = b;
    return GPET_OK;
}

/* Majorant node values of the phantom (which = 0) or detector (1). */
gpet_status gpet_get_majorant(gpet_ctx *c, int32_t which, float *out, float *e0, float *idl)
{
    if (!c || !out) return GPET_ERR_ARG;
    if (which < 0 || (which >= 2 && which - 2 >= (int)c->slots.size())) return GPET_ERR_ARG;
    const int slot = which == 0 ? 0 : which - 2;
    const std::vector<float> &v = which == 1 ? c->h_wck_det : c->slots[slot].h_wck;
    if (v.empty()) return GPET_ERR_STATE;
    memcpy(out, v.data(), sizeof(float) * GPET_NWCK);
    const DevWck &w = which == 1 ? c->ddet.wck : c->slots[slot].dp.wck;
    if (e0) *e0 = w.e0;
    if (idl) *idl = w.idl;
    return GPET_OK;
}

} /* extern "C" */
