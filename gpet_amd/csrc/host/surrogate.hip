/*
 * surrogate.hip - host side of the detector-response surrogate: parameter
 * layout, the forward and backward passes as chains of the BF16 MFMA GEMM
 * (mlp.hip), Adam training on harvested panel crossings, inference inside a
 * run, save/load, and the C ABI of gpet.h's surrogate section.
 *
 * Network: 6 -> 256 -> 256 -> 256 -> 256 -> 190 (SiLU).  Layer l keeps its
 * weights as [N_l x Kp_l] with Kp = K + 8 (bias in column K, zeros after);
 * the head is stored with 192 rows (rows 190.. stay zero) so every GEMM
 * dimension is a multiple of 8.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../common/gp_philox.h"
#include "context.h"

namespace {

constexpr int NL = 5;
constexpr int LN[NL] = {256, 256, 256, 256, 192};      /* outputs (stored)       */
constexpr int LK[NL] = {SURR_NX, 256, 256, 256, 256};  /* inputs                 */
constexpr int HID = 256, HLD = HID + 8, HEAD_LD = 192;
constexpr int64_t SPLIT_ROWS = 1024;                   /* split-K chunk (rows)   */

int kp(int l) { return l == 0 ? SURR_LDX : LK[l] + 8; }

int64_t param_offset(int l)
{
    int64_t o = 0;
    for (int i = 0; i < l; ++i) o += (int64_t)LN[i] * kp(i);
    return o;
}

int64_t n_params() { return param_offset(NL); }

LayerShapes layer_shapes()
{
    LayerShapes L;
    for (int l = 0; l <= NL; ++l) L.off[l] = param_offset(l);
    for (int l = 0; l < NL; ++l) {
        L.ln[l] = LN[l];
        L.kp[l] = kp(l);
    }
    return L;
}

/* algorithmic FLOPs of one forward pass per row (stored shapes) */
double fwd_flops_per_row()
{
    double f = 0;
    for (int l = 0; l < NL; ++l) f += 2.0 * LN[l] * kp(l);
    return f;
}

uint16_t host_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

#define KOK(c, e, what)                                                                                \
    do {                                                                                               \
        int e_ = (e);                                                                                  \
        if (e_) return fail((c), GPET_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString((hipError_t)e_)); \
    } while (0)

gpet_status alloc_params(gpet_ctx *c)
{
    SurrState &S = c->surr;
    const int64_t n = n_params();
    gpet_status st;
    if ((st = dalloc(c, S.w, n)) || (st = dalloc(c, S.m, n)) || (st = dalloc(c, S.v, n)) || (st = dalloc(c, S.g, n)) ||
        (st = dalloc(c, S.wb, n)) || (st = dalloc(c, S.wbT, n)))
        return st;
    return GPET_OK;
}

/* Upload fp32 weights (host) as master + bf16 copy; zero Adam moments. */
gpet_status put_params(gpet_ctx *c, const std::vector<float> &w)
{
    SurrState &S = c->surr;
    gpet_status st = alloc_params(c);
    if (st) return st;
    std::vector<uint16_t> wb(w.size());
    for (size_t i = 0; i < w.size(); ++i) wb[i] = host_bf16(w[i]);
    HIPOK(c, hipMemcpyAsync(S.w.p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, c->stream));
    HIPOK(c, hipMemcpyAsync(S.wb.p, wb.data(), sizeof(uint16_t) * wb.size(), hipMemcpyHostToDevice, c->stream));
    KOK(c, gpet_transpose_params(S.wb.p, S.wbT.p, layer_shapes(), c->stream), "transpose weights");
    HIPOK(c, hipMemsetAsync(S.m.p, 0, sizeof(float) * w.size(), c->stream));
    HIPOK(c, hipMemsetAsync(S.v.p, 0, sizeof(float) * w.size(), c->stream));
    HIPOK(c, hipStreamSynchronize(c->stream));
    S.adam_t = 0;
    S.has_model = true;
    return GPET_OK;
}

/* Uniform(-a, a) weights with a = sqrt(3 / fan_in) (unit-variance
 * pre-activations), the head scaled by 0.1; zero biases and padding except
 * the head's: the position components' means start spread over the depths
 * ((k + 1/2) / 16, no lateral offset) with log sigma -2, so every component
 * starts on some data. */
std::vector<float> init_params(uint64_t seed)
{
    std::vector<float> w((size_t)n_params(), 0.0f);
    for (int l = 0; l < NL; ++l) {
        const float a = std::sqrt(3.0f / (float)LK[l]) * (l == NL - 1 ? 0.1f : 1.0f);
        const int rows = l == NL - 1 ? SURR_NH : LN[l];
        gp_rng r;
        gp_rng_init(&r, seed, (uint64_t)l, 13u, 0u);
        for (int n = 0; n < rows; ++n)
            for (int k = 0; k < LK[l]; ++k)
                w[(size_t)(param_offset(l) + (int64_t)n * kp(l) + k)] = a * (2.0f * gp_rng_uniform(&r) - 1.0f);
    }
    const int64_t hb = param_offset(NL - 1) + LK[NL - 1]; /* head bias column */
    for (int k = 0; k < SURR_K; ++k) {
        w[(size_t)(hb + (int64_t)(SURR_MU + SURR_D * k) * kp(NL - 1))] = ((float)k + 0.5f) / (float)SURR_K;
        for (int d = 0; d < SURR_D; ++d)
            w[(size_t)(hb + (int64_t)(SURR_LS + SURR_D * k + d) * kp(NL - 1))] = -2.0f;
    }
    return w;
}

/* Mirror symmetries of every panel's response (bit 0: local y, bit 1:
 * local z): the modules fill the panel and the crystals fill each module
 * exactly, so the crystal grid is its own mirror image and a photon's
 * response to the mirrored entry is the mirrored response. */
int surr_reflections(const gpet_ctx *c)
{
    const DevDetector &d = c->ddet;
    if (d.moduleNy <= 0 || d.crystalNy <= 0 || c->panels.empty()) return 0;
    const int mnz = d.moduleN / d.moduleNy, cnz = d.crystalN / d.crystalNy;
    auto fills = [](int k, double len, double gap, double total) {
        return std::fabs((double)k * len + (double)(k - 1) * gap - total) <= 1e-5 * total;
    };
    int bits = 3;
    for (const gpet_panel &p : c->panels) {
        if (!(fills(d.moduleNy, p.mody, p.mspy, p.leny) && fills(d.crystalNy, p.cry, p.cspy, p.mody))) bits &= ~1;
        if (!(fills(mnz, p.modz, p.mspz, p.lenz) && fills(cnz, p.crz, p.cspz, p.modz))) bits &= ~2;
    }
    return bits;
}

gpet_status alloc_work(gpet_ctx *c, int64_t rows, bool train)
{
    SurrState &S = c->surr;
    gpet_status st;
    const bool grow = rows > S.work_rows;
    if (grow || !S.X.p) {
        if ((st = dalloc(c, S.X, rows * SURR_LDX)) || (st = dalloc(c, S.H, rows * HEAD_LD))) return st;
        for (int l = 0; l < 4; ++l) {
            if ((st = dalloc(c, S.A[l], rows * HLD))) return st;
            KOK(c, gpet_bias_columns(S.A[l].p, rows, HLD, HID, c->stream), "bias columns");
        }
        S.work_rows = rows;
        S.ldt = (rows + 7) / 8 * 8;
        for (int l = 0; l < 4; ++l) {
            dfree(S.Z[l]);
            dfree(S.AT[l]);
        }
        dfree(S.dH); dfree(S.dZ[0]); dfree(S.dZ[1]); dfree(S.Y); dfree(S.part); dfree(S.lpart);
        dfree(S.XT); dfree(S.dHT); dfree(S.dZT[0]); dfree(S.dZT[1]); dfree(S.HT);
    }
    if (train && !S.Y.p) {
        for (int l = 0; l < 4; ++l)
            if ((st = dalloc(c, S.Z[l], S.work_rows * HID))) return st;
        const int64_t splits = (S.work_rows + SPLIT_ROWS - 1) / SPLIT_ROWS;
        if ((st = dalloc(c, S.dH, S.work_rows * HEAD_LD)) || (st = dalloc(c, S.dZ[0], S.work_rows * HID)) ||
            (st = dalloc(c, S.dZ[1], S.work_rows * HID)) || (st = dalloc(c, S.Y, S.work_rows * SURR_NY)) ||
            (st = dalloc(c, S.part, splits * (int64_t)HID * HLD)) || (st = dalloc(c, S.lpart, (S.work_rows + 255) / 256 + 1)))
            return st;
        if ((st = dalloc(c, S.XT, SURR_LDX * S.ldt)) || (st = dalloc(c, S.dHT, HEAD_LD * S.ldt)) ||
            (st = dalloc(c, S.HT, HEAD_LD * S.ldt)) ||
            (st = dalloc(c, S.dZT[0], HID * S.ldt)) || (st = dalloc(c, S.dZT[1], HID * S.ldt)))
            return st;
        for (int l = 0; l < 4; ++l) {
            if ((st = dalloc(c, S.AT[l], HLD * S.ldt))) return st;
            KOK(c, gpet_bias_rows(S.AT[l].p, S.ldt, HID, HLD, c->stream), "bias rows");
        }
    }
    return GPET_OK;
}

GemmArgs gargs(const uint16_t *A, int64_t lda, const uint16_t *B, int64_t ldb, int64_t M, int64_t N, int64_t K)
{
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.M = M; g.N = N; g.K = K;
    return g;
}

/* Forward over M rows of S.X; train keeps pre-activations for backward. */
gpet_status forward(gpet_ctx *c, int64_t M, bool train)
{
    SurrState &S = c->surr;
    const uint16_t *in = S.X.p;
    int64_t ldin = SURR_LDX;
    for (int l = 0; l < 4; ++l) {
        GemmArgs g = gargs(in, ldin, S.wb.p + param_offset(l), kp(l), M, HID, kp(l));
        g.Cb = S.A[l].p;
        g.ldcb = HLD;
        if (train) {
            g.C = S.Z[l].p;
            g.ldc = HID;
            g.CbT = S.AT[l].p;
            g.ldt = S.ldt;
        }
        KOK(c, gpet_gemm(train ? EPI_SILU : EPI_ACT, g, 1, c->stream), "gemm forward");
        in = S.A[l].p;
        ldin = HLD;
    }
    GemmArgs g = gargs(in, HLD, S.wb.p + param_offset(4), kp(4), M, HEAD_LD, kp(4));
    if (train) {
        g.CT = S.HT.p;   /* the loss reads the head outputs transposed */
        g.ldt = S.ldt;
    } else {
        g.C = S.H.p;
        g.ldc = HEAD_LD;
    }
    KOK(c, gpet_gemm(EPI_F32, g, 1, c->stream), "gemm head");
    return GPET_OK;
}

/* dW_l = dZ^T . A_in over M rows from the transposed copies dZ^T
 * ([LN][ldt]) and A_in^T ([kp][ldt]) (split-K partials, fixed-order sum). */
gpet_status weight_grad(gpet_ctx *c, int l, const uint16_t *dzT, const uint16_t *ainT, int64_t M)
{
    SurrState &S = c->surr;
    const int64_t splits = (M + SPLIT_ROWS - 1) / SPLIT_ROWS;
    GemmArgs g = gargs(dzT, S.ldt, ainT, S.ldt, LN[l], kp(l), M);
    g.k_chunk = SPLIT_ROWS;
    g.C = S.part.p;
    KOK(c, gpet_gemm(EPI_PARTIAL, g, (int)splits, c->stream), "gemm weight grad");
    KOK(c, gpet_sum_partials(S.part.p, (int)splits, (int64_t)LN[l] * kp(l), S.g.p + param_offset(l), c->stream),
        "partials");
    return GPET_OK;
}

/* Backward from S.dH (set by the loss kernel) to S.g. */
gpet_status backward(gpet_ctx *c, int64_t M)
{
    SurrState &S = c->surr;
    gpet_status st;
    /* head */
    if ((st = weight_grad(c, 4, S.dHT.p, S.AT[3].p, M))) return st;
    const uint16_t *dnext = S.dH.p;
    int64_t ldnext = HEAD_LD;
    int knext = HEAD_LD; /* reduction length of the input-gradient GEMM */
    for (int l = 3; l >= 0; --l) {
        uint16_t *dz = S.dZ[l & 1].p;
        /* dZ_l = (dnext . W_{l+1})[:, :256] * silu'(Z_l), with W_{l+1}
         * transposed ([kp][LN]: row k holds input k's weights); the
         * row-major copy only where another input gradient reads it */
        GemmArgs g = gargs(dnext, ldnext, S.wbT.p + param_offset(l + 1), LN[l + 1], M, HID, knext);
        g.Cb = l > 0 ? dz : nullptr;
        g.ldcb = HID;
        g.Z = S.Z[l].p;
        g.ldz = HID;
        g.CbT = S.dZT[l & 1].p;
        g.ldt = S.ldt;
        KOK(c, gpet_gemm(EPI_DSILU, g, 1, c->stream), "gemm input grad");
        const uint16_t *a_inT = l == 0 ? S.XT.p : S.AT[l - 1].p;
        if ((st = weight_grad(c, l, S.dZT[l & 1].p, a_inT, M))) return st;
        dnext = dz;
        ldnext = HID;
        knext = HID;
    }
    return GPET_OK;
}

gpet_status check_model(gpet_ctx *c)
{
    if (!c->surr.has_model) return fail(c, GPET_ERR_STATE, "no surrogate model (train, load or set_params first)");
    return GPET_OK;
}

void pack_host_rows(const float *x, int64_t n, std::vector<uint16_t> &X)
{
    X.assign((size_t)n * SURR_LDX, 0);
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < SURR_NX; ++k) X[(size_t)(i * SURR_LDX + k)] = host_bf16(x[i * SURR_NX + k]);
        X[(size_t)(i * SURR_LDX + SURR_NX)] = 0x3f80;
    }
}

} // namespace

/* Run-wide surrogate queue with a capacity from the last run's rate (or the
 * plain queue bound for a first run, capped at 1/4 of all photons). */
gpet_status surr_run_prepare(gpet_ctx *c, int64_t n_hist, int64_t qcap, DevQueue &sq)
{
    SurrState &S = c->surr;
    gpet_status st;
    if ((st = check_model(c))) return st;
    if (c->dig.readout_depth != S.train_depth || c->dig.readout_policy != S.train_policy ||
        c->ddet.moduleN != S.train_modules)
        return fail(c, GPET_ERR_UNSUPPORTED, "surrogate was trained with another readout or module layout");
    int64_t need = S.items_per_hist >= 0 ? (int64_t)(S.items_per_hist * 1.1 * (double)n_hist) + qcap
                                         : std::max<int64_t>(qcap, n_hist / 2);
    need = std::max<int64_t>(need, 1024);
    if (need > S.sq_cap) {
        for (int k = 0; k < 6; ++k)
            if ((st = dalloc(c, S.sq_f[k], need))) return st;
        if ((st = dalloc(c, S.sq_t, need)) || (st = dalloc(c, S.sq_info, need))) return st;
        S.sq_cap = need;
    }
    if ((st = dalloc(c, S.sq_count, 1))) return st;
    memset(&sq, 0, sizeof(sq));
    sq.y = S.sq_f[0].p; sq.z = S.sq_f[1].p; sq.vx = S.sq_f[2].p; sq.vy = S.sq_f[3].p; sq.vz = S.sq_f[4].p;
    sq.E = S.sq_f[5].p;
    sq.t = S.sq_t.p;
    sq.info = S.sq_info.p;
    sq.trace_slot = nullptr;
    sq.count = S.sq_count.p;
    sq.cap = S.sq_cap;
    return GPET_OK;
}

/* Network over the n listed entries in chunks; events appended via pp. */
/* Layer weights (bf16) of the fused inference kernel. */
void layer_weights(const SurrState &S, const uint16_t *w[NL])
{
    for (int l = 0; l < NL; ++l) w[l] = S.wb.p + param_offset(l);
}

gpet_status surr_run_infer(gpet_ctx *c, int64_t n, const PanelParams &pp)
{
    SurrState &S = c->surr;
    auto t0 = std::chrono::steady_clock::now();
    /* the whole network and the sampling in one kernel: no work buffers */
    const uint16_t *w[NL];
    layer_weights(S, w);
    KOK(c, gpet_mlp_infer(w, nullptr, pp.sq, 0, n, c->ddet, pp, nullptr, 0, c->stream), "fused inference");
    HIPOK(c, hipStreamSynchronize(c->stream));
    c->last_surr_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return GPET_OK;
}

void surr_free(gpet_ctx *c)
{
    SurrState &S = c->surr;
    dfree(S.w); dfree(S.m); dfree(S.v); dfree(S.g); dfree(S.wb); dfree(S.wbT); dfree(S.set); dfree(S.set_count);
    for (auto &b : S.sq_f) dfree(b);
    dfree(S.sq_t); dfree(S.sq_info); dfree(S.sq_count);
    dfree(S.X); dfree(S.dH); dfree(S.dZ[0]); dfree(S.dZ[1]); dfree(S.H); dfree(S.Y); dfree(S.part); dfree(S.lpart);
    for (int l = 0; l < 4; ++l) {
        dfree(S.A[l]);
        dfree(S.Z[l]);
    }
    dfree(S.idx);
}

extern "C" {

gpet_status gpet_surrogate_harvest(gpet_ctx *c, const gpet_source *src, int32_t ns, const gpet_isotope *iso,
                                   int32_t niso, double t0, double t1, int64_t n_decays, int64_t capacity)
{
    if (!c || capacity <= 0) return GPET_ERR_ARG;
    SurrState &S = c->surr;
    HIPOK(c, hipSetDevice(c->device));
    gpet_status st;
    if (capacity * SURR_REC > S.set.cap) {
        if ((st = dgrow(c, S.set, capacity * SURR_REC, S.set_n * SURR_REC))) return st;
    }
    if ((st = dalloc(c, S.set_count, 1))) return st;
    if (S.set_n >= S.set.cap / SURR_REC) return fail(c, GPET_ERR_CAPACITY, "training set is full");
    S.train_depth = c->dig.readout_depth;
    S.train_policy = c->dig.readout_policy;
    S.train_modules = c->ddet.moduleN;
    const bool was = S.enabled;
    S.enabled = false;
    c->panel_mode = PANEL_HARVEST;
    st = gpet_run_source(c, src, ns, iso, niso, t0, t1, n_decays, nullptr);
    c->panel_mode = PANEL_PLAIN;
    S.enabled = was;
    return st;
}

gpet_status gpet_surrogate_train(gpet_ctx *c, const gpet_surrogate_config *cfg, gpet_surrogate_report *rep)
{
    if (!c || !cfg || cfg->steps <= 0 || cfg->batch <= 0 || !(cfg->lr > 0.0f)) return GPET_ERR_ARG;
    SurrState &S = c->surr;
    if (S.set_n <= 0) return fail(c, GPET_ERR_STATE, "no harvested training records");
    HIPOK(c, hipSetDevice(c->device));
    gpet_status st;
    if (!S.has_model || cfg->reinit) {
        if ((st = put_params(c, init_params(cfg->seed)))) return st;
    }
    const int64_t M = cfg->batch;
    if ((st = alloc_work(c, M, true))) return st;
    DBuf<float> losses, clip;
    if ((st = dalloc(c, losses, cfg->steps)) || (st = dalloc(c, clip, 256))) return st;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t warm = std::max<int64_t>(1, cfg->steps / 50);
    const int nblk = (int)((M + 255) / 256);
    const int refl = surr_reflections(c);
    HIPOK(c, hipStreamSynchronize(c->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (int64_t step = 0; step < cfg->steps; ++step) {
        float lr = cfg->lr;
        if (step < warm) lr = cfg->lr * (float)(step + 1) / (float)warm;
        else {
            double p = (double)(step - warm) / (double)std::max<int64_t>(1, cfg->steps - warm);
            lr = (float)(cfg->lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * p))));
        }
        KOK(c, gpet_pack_train(S.set.p, S.set_n, cfg->seed, S.adam_t, M, S.X.p, S.Y.p, S.XT.p, S.ldt, refl, c->stream),
            "pack");
        if ((st = forward(c, M, true))) return st;
        KOK(c, gpet_head_loss_grad(S.HT.p, S.ldt, S.Y.p, M, 1.0f / (float)M, S.dH.p, HEAD_LD, S.dHT.p, S.lpart.p,
                                  c->stream),
            "loss");
        KOK(c, gpet_loss_reduce(S.lpart.p, nblk, losses.p + step, c->stream), "loss reduce");
        if ((st = backward(c, M))) return st;
        S.adam_t++;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)S.adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)S.adam_t)));
        /* global-norm clipping at 1 (the loss is a mean over rows) */
        KOK(c, gpet_clip_factor(S.g.p, n_params(), 1.0f, clip.p, clip.p + 255, c->stream), "clip");
        KOK(c, gpet_adam(S.w.p, S.m.p, S.v.p, S.g.p, clip.p + 255, S.wb.p, n_params(), lr, b1, b2, c1, c2, eps,
                         c->stream),
            "adam");
        KOK(c, gpet_transpose_params(S.wb.p, S.wbT.p, layer_shapes(), c->stream), "transpose weights");
    }
    HIPOK(c, hipStreamSynchronize(c->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<float> L((size_t)cfg->steps);
    HIPOK(c, hipMemcpy(L.data(), losses.p, sizeof(float) * L.size(), hipMemcpyDeviceToHost));
    dfree(losses);
    dfree(clip);
    if (rep) {
        const int64_t k = std::max<int64_t>(1, cfg->steps / 50);
        double a = 0, b = 0;
        for (int64_t i = 0; i < k; ++i) {
            a += L[(size_t)i];
            b += L[(size_t)(cfg->steps - 1 - i)];
        }
        rep->records = S.set_n;
        rep->steps = cfg->steps;
        rep->seconds = secs;
        rep->flops = 3.0 * fwd_flops_per_row() * (double)M * (double)cfg->steps;
        rep->loss_first = (float)(a / (double)k / (double)M);
        rep->loss_last = (float)(b / (double)k / (double)M);
    }
    return GPET_OK;
}

gpet_status gpet_surrogate_enable(gpet_ctx *c, int32_t on)
{
    if (!c) return GPET_ERR_ARG;
    gpet_status st;
    if (on && (st = check_model(c))) return st;
    c->surr.enabled = on != 0;
    return GPET_OK;
}

gpet_status gpet_surrogate_get_params(gpet_ctx *c, float *out, int64_t cap, int64_t *n)
{
    if (!c) return GPET_ERR_ARG;
    if (n) *n = n_params();
    if (!out) return GPET_OK;
    gpet_status st;
    if ((st = check_model(c))) return st;
    if (cap < n_params()) return fail(c, GPET_ERR_CAPACITY, "parameter buffer too small");
    HIPOK(c, hipSetDevice(c->device));
    HIPOK(c, hipMemcpy(out, c->surr.w.p, sizeof(float) * (size_t)n_params(), hipMemcpyDeviceToHost));
    return GPET_OK;
}

gpet_status gpet_surrogate_set_params(gpet_ctx *c, const float *w, int64_t n)
{
    if (!c || !w) return GPET_ERR_ARG;
    if (n != n_params()) return fail(c, GPET_ERR_ARG, "parameter count");
    HIPOK(c, hipSetDevice(c->device));
    if (c->surr.train_depth < 0) {
        c->surr.train_depth = c->dig.readout_depth;
        c->surr.train_policy = c->dig.readout_policy;
        c->surr.train_modules = c->ddet.moduleN;
    }
    return put_params(c, std::vector<float>(w, w + n));
}

gpet_status gpet_surrogate_save(gpet_ctx *c, const char *path)
{
    if (!c || !path) return GPET_ERR_ARG;
    gpet_status st;
    if ((st = check_model(c))) return st;
    std::vector<float> w((size_t)n_params());
    if ((st = gpet_surrogate_get_params(c, w.data(), (int64_t)w.size(), nullptr))) return st;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(c, GPET_ERR_IO, std::string("cannot write ") + path);
    const char magic[8] = {'G', 'P', 'E', 'T', 'S', 'U', 'R', '2'};
    int32_t meta[4] = {c->surr.train_depth, c->surr.train_policy, c->surr.train_modules, 0};
    int64_t n = (int64_t)w.size();
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(meta, 4, 4, f) == 4 && fwrite(&n, 8, 1, f) == 1 &&
              fwrite(w.data(), 4, w.size(), f) == w.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? GPET_OK : fail(c, GPET_ERR_IO, std::string("short write to ") + path);
}

gpet_status gpet_surrogate_load(gpet_ctx *c, const char *path)
{
    if (!c || !path) return GPET_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(c, GPET_ERR_IO, std::string("cannot open ") + path);
    char magic[8];
    int32_t meta[4];
    int64_t n = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "GPETSUR2", 8) == 0 && fread(meta, 4, 4, f) == 4 &&
              fread(&n, 8, 1, f) == 1 && n == n_params();
    std::vector<float> w(ok ? (size_t)n : 0);
    ok = ok && fread(w.data(), 4, w.size(), f) == w.size();
    fclose(f);
    if (!ok) return fail(c, GPET_ERR_FORMAT, std::string("not a surrogate parameter file: ") + path);
    HIPOK(c, hipSetDevice(c->device));
    c->surr.train_depth = meta[0];
    c->surr.train_policy = meta[1];
    c->surr.train_modules = meta[2];
    return put_params(c, w);
}

gpet_status gpet_surrogate_records(gpet_ctx *c, float *out, int64_t cap, int64_t *n)
{
    if (!c) return GPET_ERR_ARG;
    if (n) *n = c->surr.set_n;
    if (!out || c->surr.set_n == 0) return GPET_OK;
    HIPOK(c, hipSetDevice(c->device));
    const int64_t m = std::min<int64_t>(cap, c->surr.set_n);
    HIPOK(c, hipMemcpy(out, c->surr.set.p, sizeof(float) * (size_t)(m * SURR_REC), hipMemcpyDeviceToHost));
    return GPET_OK;
}

gpet_status gpet_surrogate_forward(gpet_ctx *c, const float *x, int64_t n, float *h)
{
    if (!c || !x || !h || n <= 0) return GPET_ERR_ARG;
    gpet_status st;
    if ((st = check_model(c))) return st;
    HIPOK(c, hipSetDevice(c->device));
    if ((st = alloc_work(c, n, false))) return st;
    std::vector<uint16_t> X;
    pack_host_rows(x, n, X);
    HIPOK(c, hipMemcpy(c->surr.X.p, X.data(), sizeof(uint16_t) * X.size(), hipMemcpyHostToDevice));
    /* the inference kernel (the one surrogate runs use), head outputs to H */
    const uint16_t *w[NL];
    layer_weights(c->surr, w);
    DevQueue none;
    memset(&none, 0, sizeof(none));
    PanelParams pp;
    memset(&pp, 0, sizeof(pp));
    KOK(c, gpet_mlp_infer(w, c->surr.X.p, none, 0, n, c->ddet, pp, c->surr.H.p, HEAD_LD, c->stream), "fused forward");
    std::vector<float> H((size_t)n * HEAD_LD);
    HIPOK(c, hipStreamSynchronize(c->stream));
    HIPOK(c, hipMemcpy(H.data(), c->surr.H.p, sizeof(float) * H.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) memcpy(h + i * SURR_NH, &H[(size_t)(i * HEAD_LD)], sizeof(float) * SURR_NH);
    return GPET_OK;
}

gpet_status gpet_surrogate_grad(gpet_ctx *c, const float *x, const float *y, int64_t n, float *loss, float *grad)
{
    if (!c || !x || !y || !grad || n <= 0) return GPET_ERR_ARG;
    SurrState &S = c->surr;
    gpet_status st;
    if ((st = check_model(c))) return st;
    HIPOK(c, hipSetDevice(c->device));
    if (n > S.work_rows || !S.Y.p) {
        if ((st = alloc_work(c, std::max<int64_t>(n, S.work_rows), true))) return st;
    }
    std::vector<uint16_t> X;
    pack_host_rows(x, n, X);
    std::vector<uint16_t> XT((size_t)(SURR_LDX * S.ldt), 0);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < SURR_LDX; ++k) XT[(size_t)(k * S.ldt + i)] = X[(size_t)(i * SURR_LDX + k)];
    HIPOK(c, hipMemcpy(S.X.p, X.data(), sizeof(uint16_t) * X.size(), hipMemcpyHostToDevice));
    HIPOK(c, hipMemcpy(S.XT.p, XT.data(), sizeof(uint16_t) * XT.size(), hipMemcpyHostToDevice));
    /* the loss takes the energy class in place of the deposited fraction */
    std::vector<float> Y(y, y + n * SURR_NY);
    for (int64_t i = 0; i < n; ++i) Y[(size_t)(i * SURR_NY + 1)] = (float)surr_class(x[i * SURR_NX], y[i * SURR_NY + 1]);
    HIPOK(c, hipMemcpy(S.Y.p, Y.data(), sizeof(float) * Y.size(), hipMemcpyHostToDevice));
    if ((st = forward(c, n, true))) return st;
    const int nblk = (int)((n + 255) / 256);
    KOK(c, gpet_head_loss_grad(S.HT.p, S.ldt, S.Y.p, n, 1.0f / (float)n, S.dH.p, HEAD_LD, S.dHT.p, S.lpart.p,
                              c->stream),
        "loss");
    DBuf<float> l1;
    if ((st = dalloc(c, l1, 1))) return st;
    KOK(c, gpet_loss_reduce(S.lpart.p, nblk, l1.p, c->stream), "loss reduce");
    if ((st = backward(c, n))) return st;
    HIPOK(c, hipStreamSynchronize(c->stream));
    float lsum = 0;
    HIPOK(c, hipMemcpy(&lsum, l1.p, sizeof(float), hipMemcpyDeviceToHost));
    HIPOK(c, hipMemcpy(grad, S.g.p, sizeof(float) * (size_t)n_params(), hipMemcpyDeviceToHost));
    dfree(l1);
    if (loss) *loss = lsum / (float)n;
    return GPET_OK;
}

} /* extern "C" */
