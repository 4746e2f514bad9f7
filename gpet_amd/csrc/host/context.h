/*
 * context.h - the context behind the C ABI (device buffers, phantom slots,
 * launch shape, graph cache, surrogate state) and the buffer helpers shared
 * by the host translation units (ctx.hip, surrogate.hip).
 */
#ifndef GPET_CONTEXT_H
#define GPET_CONTEXT_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/gpet.h"
#include "../kernels/surr_common.h"
#include "internal.h"

constexpr int64_t DEFAULT_BATCH = 1 << 23;

template <class T>
struct DBuf {
    T *p = nullptr;
    int64_t cap = 0;
};

/* One phantom slot: voxels, bricked densities and its own majorant. */
struct PhSlot {
    DBuf<int2> vox;
    DBuf<float> dens_b;
    DBuf<uint4> wmap;
    DBuf<float> wck;
    std::vector<float> h_wck;
    DevPhantom dp;
    bool present = false;
};

/* Captured transport batches of the last graph run and the launch
 * arguments they were captured with. */
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
};

/* Detector-response surrogate: network parameters (flat fp32 master copy,
 * Adam moments, gradient, bf16 copy), the harvested training set, the
 * run-wide queue of in-range panel entries, and work buffers sized for
 * `work_rows` rows. */
struct SurrState {
    bool has_model = false, enabled = false;
    int train_depth = -1, train_policy = -1, train_modules = -1;
    DBuf<float> w, m, v, g;
    DBuf<uint16_t> wb;
    int64_t adam_t = 0;
    DBuf<float> set;
    int64_t set_n = 0;
    DBuf<unsigned long long> set_count;
    DBuf<float> sq_f[6];
    DBuf<double> sq_t;
    DBuf<int4> sq_info;
    DBuf<int> sq_count;
    int64_t sq_cap = 0;
    double items_per_hist = -1.0;
    int64_t work_rows = 0;
    DBuf<uint16_t> X, A[4], dH, dZ[2];
    DBuf<float> Z[4], H, Y, part, lpart;
    /* training: the inputs, activations and gradients transposed ([features][ldt],
     * ldt = work rows rounded up to 8) and the weights transposed per layer */
    DBuf<uint16_t> XT, AT[4], dHT, dZT[2], wbT;
    DBuf<float> HT;
    int64_t ldt = 0;
    DBuf<int64_t> idx;
};

struct gpet_ctx {
    gpet_options opt;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::vector<hipEvent_t> kev; /* per-launch timing events */
    /* asynchronous sink copies (gpet_sink.async_copy): issued on copy_stream
     * once the digitizer has ended; copy_done follows the last of them */
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_done = nullptr;
    bool copy_pending = false;

    /* tables */
    bool has_tab = false;
    int nmat = 0;
    float eminph = 0, emax = 0;
    float e0 = 0, idl = 0, lam_idl = 0;
    std::vector<float> h_lamph; /* host copy for majorant construction */
    DBuf<float4> d_mu4;
    DBuf<float> d_surf[2];
    DevTables dtab;

    /* phantom slots; slot 0 is the phantom of gpet_set_phantom */
    bool has_ph = false;
    std::vector<PhSlot> slots;
    DevPhantom dph;              /* slot 0's descriptor (or absent)          */
    DBuf<DevPhantom> d_phs;      /* descriptors of all slots, for batches    */
    bool phs_dirty = true;

    /* launch shape, tuner state, graph cache */
    gpet_launch launch;
    int n_cu = 256;
    bool tuned[2] = {false, false}; /* single-phantom runs, batched runs    */
    double tune_seconds = 0;
    GraphCache graph;
    double ev_per_hist = -1.0;   /* adder events per history of the last run */
    DBuf<DevSegment> d_seg;
    DBuf<int64_t> d_job_first;
    DBuf<unsigned long long> d_job_cnt;

    /* detector */
    bool has_det = false;
    DBuf<DevPanel> d_panels;
    DBuf<float4> d_pgroup_sph;   /* panel groups of the entry test           */
    DBuf<unsigned long long> d_pgroup_mask;
    DBuf<float> d_wck_det;
    std::vector<float> h_wck_det;
    std::vector<gpet_panel> panels;
    DevDetector ddet;

    /* digitizer */
    bool has_dig = false;
    gpet_digitizer dig;

    /* run buffers */
    DBuf<float> q_f[6];
    DBuf<double> q_t;
    DBuf<int4> q_info;
    DBuf<int> q_slot;
    int64_t q_cap = 0;
    DBuf<unsigned> d_small;      /* [0] queue count, [1] events, [2] hits */
    void *na_scratch = nullptr;  /* device address of gpet_run_psf_noalloc's scratch */
    DBuf<unsigned long long> d_cnt;
    DBuf<gpet_event> d_events;
    DBuf<gpet_trace> d_trace;
    DBuf<unsigned> d_rcount;     /* panel-queue region counts                */
    DBuf<LongWalkItem> d_lw;     /* deferred positron walks (queue / 8)      */
    DBuf<unsigned> d_lw_count;
    DBuf<TransitItem> d_tq[2];   /* photons in transit (two per history slot) */
    DBuf<unsigned> d_tq_count[2];
    DBuf<gpet_psf_record> d_psf_src, d_psf_ph;
    DBuf<unsigned long long> d_npsf;
    DBuf<int32_t> d_hit_id;
    DBuf<float> d_hit_f;
    DBuf<int64_t> d_first;
    DBuf<DevSource> d_src;
    DBuf<DevIsotope> d_iso;
    DBuf<double> d_psf;
    DigiWork *dw = nullptr;

    /* surrogate */
    SurrState surr;
    int panel_mode = PANEL_PLAIN; /* of the next run (harvest runs set it)  */
    int64_t last_surr_items = 0;
    double last_surr_seconds = 0;

    /* last run */
    gpet_tallies last;
    DigiOut dout;
    int64_t n_events = 0;
};


inline gpet_status fail(gpet_ctx *c, gpet_status s, const std::string &msg)
{
    if (c) c->err = msg;
    return s;
}

#define HIPOK(c, x)                                                                                    \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            return fail((c), GPET_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

template <class T>
gpet_status dalloc(gpet_ctx *c, DBuf<T> &b, int64_t n)
{
    if (n <= b.cap && b.p) return GPET_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    int64_t m = std::max<int64_t>(n, 1);
    hipError_t e = hipMalloc((void **)&b.p, sizeof(T) * (size_t)m);
    if (e != hipSuccess) return fail(c, GPET_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.cap = m;
    return GPET_OK;
}

template <class T>
void dfree(DBuf<T> &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

/* Grow keeping the first `keep` elements (stream-ordered copy). */
template <class T>
gpet_status dgrow(gpet_ctx *c, DBuf<T> &b, int64_t need, int64_t keep)
{
    if (need <= b.cap && b.p) return GPET_OK;
    DBuf<T> nb;
    gpet_status st = dalloc(c, nb, std::max<int64_t>(need, b.cap + b.cap / 2));
    if (st) return st;
    if (b.p && keep > 0) HIPOK(c, hipMemcpyAsync(nb.p, b.p, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice, c->stream));
    HIPOK(c, hipStreamSynchronize(c->stream));
    dfree(b);
    b = nb;
    return GPET_OK;
}

template <class T>
gpet_status upload(gpet_ctx *c, DBuf<T> &b, const T *h, int64_t n)
{
    gpet_status st = dalloc(c, b, n);
    if (st) return st;
    if (n > 0) HIPOK(c, hipMemcpyAsync(b.p, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPOK(c, hipStreamSynchronize(c->stream));
    return GPET_OK;
}

/* Surrogate entry points used by the run orchestration (surrogate.hip). */
gpet_status surr_run_prepare(gpet_ctx *c, int64_t n_hist, int64_t qcap, DevQueue &sq);
gpet_status surr_run_infer(gpet_ctx *c, int64_t n_items, const PanelParams &pp);
void surr_free(gpet_ctx *c);

#endif
