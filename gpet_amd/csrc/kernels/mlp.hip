/*
 * mlp.hip - kernels of the detector-response surrogate: a BF16 MFMA GEMM
 * with FP32 accumulation and fused epilogues, the head's loss and its
 * gradient, Adam, and the inference-side input packing and sampling.
 *
 * The network (6 -> 256 -> 256 -> 256 -> 256 -> 190) maps a photon's panel
 * entry (energy, local direction, entry point) to P(event), a histogram of
 * the deposited-energy fraction over 61 classes (full absorption, 48 bins
 * up to the entry's Compton edge, 12 above it) and, for each of the two
 * energy types (full absorption, continuum), a 16-component diagonal
 * Gaussian mixture over the depth of the first deposit and its lateral
 * offsets from the straight entry ray.  It stands
 * in for the panel transport of photonde (gPET_kernals.cu:1011-1193) plus the
 * adder/readout (:737-813) inside its validity range; nothing in the reference
 * corresponds to the network itself.
 *
 * Layout conventions (all row-major):
 *   activations  bf16 [M x Kp], Kp = K + 8: column K holds the constant 1
 *                that carries the bias, columns K+1.. are zero;
 *   weights      bf16 [N x Kp] (one output neuron per row, bias in column K);
 *   pre-activations fp32 [M x N] (kept for the backward pass).
 *
 * GEMM: C[M x N] = op(A)[M x K] op(B)[K x N] on 64 x 64 block tiles, four
 * waves of 32 x 32, v_mfma_f32_32x32x16_bf16.  Operands are staged through
 * LDS k-contiguous, so the transposed forms (A^T for weight gradients, B^T
 * for input gradients) are transposed on the way into LDS.  Every operand's
 * contiguous dimension is a multiple of 8 (16-byte loads); M and the
 * reduction length are arbitrary (zero-filled tails).
 */
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "../common/gp_math.h"
#include "../host/internal.h"
#include "surr_common.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;


__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

/* round to nearest even */
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + gp_expf(-z)); }
__device__ __forceinline__ float silu_grad(float z)
{
    float s = 1.0f / (1.0f + gp_expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

} // namespace

/* ---- training GEMMs ----------------------------------------------------------
 * C[m][n] = sum_k A[m][k] B[n][k] with k contiguous in both operands (rows
 * of lda / ldb elements).  Every product of a training step has this form
 * because the forward pass also writes its activations transposed, the
 * backward pass its pre-activation gradients transposed, and the weights
 * are kept in both layouts: no operand is transposed on the way into the
 * MFMAs.  One block of 256 threads (4 waves, 2 x 2 sub-tiles of 64 x 64,
 * four v_mfma_f32_32x32x16_bf16 accumulators each) per 128 x 128 tile of C;
 * the k loop stages 64-wide slices of both operands in LDS, double
 * buffered: the next slice's global loads are in flight while the current
 * one multiplies, one barrier per slice.  blockIdx.z selects a chunk of k
 * (split-K partials, summed in a fixed order by k_sum_partials).  k must
 * be a multiple of 8 at every row start; a vector that straddles the end of
 * k is masked element by element.  Epilogues: EPI_F32 C; EPI_SILU Z = C
 * (fp32, optional), silu(C) bf16 (optional) and transposed (optional);
 * EPI_ACT silu(C) bf16; EPI_DSILU C * silu'(Z) bf16 (optional) and
 * transposed (optional); EPI_F32 C fp32 and/or transposed (CT); EPI_PARTIAL
 * the chunk's partial.  Transposed copies go through LDS so that their rows
 * are written 16 bytes per lane. */
constexpr int TM = 128, TN = 128, TK = 64;
constexpr int TLD = TK + 8;   /* staged row: 144 B, 16-B fragment reads without bank conflicts */
constexpr int TTL = TM + 8;   /* transposed epilogue row (bf16) */
constexpr int TTL32 = TM + 4; /* transposed epilogue row (fp32; 128 x 132 x 4 B fits the staging buffers) */

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) uint16_t sm[2][TM + TN][TLD];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * TM, n0 = (int64_t)blockIdx.y * TN;
    const int64_t kb = (int64_t)blockIdx.z * g.k_chunk;
    const int64_t ke = min(g.K, kb + g.k_chunk);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x16_t{};
    /* vector j of a thread: operand j / 4, element v = t + 256 (j % 4) of
     * the 128 rows x 8 vectors of the slice */
    uint4 st[8];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int v = t + 256 * (j & 3);
            const int r = v >> 3, kv = (v & 7) * 8;
            const bool isb = j >= 4;
            const int64_t row = (isb ? n0 : m0) + r;
            const int64_t k = k0 + kv;
            uint4 x = make_uint4(0u, 0u, 0u, 0u);
            if (row < (isb ? g.N : g.M) && k < ke) {
                x = *reinterpret_cast<const uint4 *>((isb ? g.B : g.A) + row * (isb ? g.ldb : g.lda) + k);
                if (k + 8 > ke) {
                    const int nv = (int)(ke - k);   /* 1..7 elements inside */
                    uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        if (e >= nv) w[e >> 1] &= (e & 1) ? 0x0000ffffu : 0xffff0000u;
                    x = make_uint4(w[0], w[1], w[2], w[3]);
                }
            }
            st[j] = x;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int v = t + 256 * (j & 3);
            *reinterpret_cast<uint4 *>(&sm[b][(j >= 4 ? TM : 0) + (v >> 3)][(v & 7) * 8]) = st[j];
        }
    };
    /* EPI_DSILU: the pre-activations of the tile, loaded before the k loop
     * so that their latency hides behind it */
    float zpre[2][2][16];
    if (EPI == EPI_DSILU) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int64_t col = n0 + wn + 32 * j + (lane & 31);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t row = m0 + wm + 32 * i + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    zpre[i][j][r] = (row < g.M && col < g.N) ? g.Z[row * g.ldz + col] : 0.0f;
                }
            }
    }
    const int nk = ke > kb ? (int)((ke - kb + TK - 1) / TK) : 0;
    if (nk > 0) {
        load(kb);
        store(0);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int b = kt & 1;
        if (kt + 1 < nk) load(kb + (int64_t)(kt + 1) * TK);
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            /* lanes 0-31 hold k 0-7 of the step, lanes 32-63 k 8-15 */
            const int kk = ks + (lane >> 5) * 8;
            bf16x8_t a[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = *reinterpret_cast<const bf16x8_t *>(&sm[b][wm + 32 * i + (lane & 31)][kk]);
                bv[i] = *reinterpret_cast<const bf16x8_t *>(&sm[b][TM + wn + 32 * i + (lane & 31)][kk]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) store(b ^ 1);
        __syncthreads();
    }
    /* accumulator r of lane l in sub-tile (i, j): row wm + 32 i + (r / 4) 8 +
     * (l / 32) 4 + r % 4, column wn + 32 j + l % 32 */
    constexpr bool TR = EPI == EPI_SILU || EPI == EPI_DSILU;
    uint16_t (*T)[TTL] = reinterpret_cast<uint16_t (*)[TTL]>(&sm[0][0][0]);
    float (*T32)[TTL32] = reinterpret_cast<float (*)[TTL32]>(&sm[0][0][0]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int cl = wn + 32 * j + (lane & 31);
            const int64_t col = n0 + cl;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rl = wm + 32 * i + q * 8 + (lane >> 5) * 4;
                uint16_t o[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t row = m0 + rl + e;
                    const float v = acc[i][j][4 * q + e];
                    const bool in = row < g.M && col < g.N;
                    if (EPI == EPI_F32) {
                        if (in && g.C) g.C[row * g.ldc + col] = v;
                    } else if (EPI == EPI_PARTIAL) {
                        if (in) g.C[(int64_t)blockIdx.z * g.M * g.N + row * g.N + col] = v;
                    } else if (EPI == EPI_SILU || EPI == EPI_ACT) {
                        o[e] = f2bf(silu(v));
                        if (in) {
                            if (EPI == EPI_SILU && g.C) g.C[row * g.ldc + col] = v;
                            if (g.Cb) g.Cb[row * g.ldcb + col] = o[e];
                        }
                    } else {
                        o[e] = in ? f2bf(v * silu_grad(zpre[i][j][4 * q + e])) : (uint16_t)0;
                        if (in && g.Cb) g.Cb[row * g.ldcb + col] = o[e];
                    }
                }
                if (TR && g.CbT)
                    *reinterpret_cast<uint2 *>(&T[cl][rl]) =
                        make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
                if (EPI == EPI_F32 && g.CT)
                    *reinterpret_cast<float4 *>(&T32[cl][rl]) =
                        make_float4(acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]);
            }
        }
    if (EPI == EPI_F32 && g.CT) {
        __syncthreads();
        for (int v = t; v < TN * (TM / 4); v += 256) {
            const int cl = v >> 5, rv = (v & 31) * 4;
            const int64_t col = n0 + cl, row = m0 + rv;
            if (col >= g.N || row >= g.M) continue;
            const float4 x = *reinterpret_cast<const float4 *>(&T32[cl][rv]);
            float *dst = g.CT + col * g.ldt + row;
            if (row + 4 <= g.M) {
                *reinterpret_cast<float4 *>(dst) = x;
            } else {
                const float w[4] = {x.x, x.y, x.z, x.w};
                for (int e = 0; e < (int)(g.M - row); ++e) dst[e] = w[e];
            }
        }
    }
    if (TR && g.CbT) {
        __syncthreads();
        /* row n of the transposed tile: the block's 128 m values of column n */
        for (int v = t; v < TN * (TM / 8); v += 256) {
            const int cl = v >> 4, rv = (v & 15) * 8;
            const int64_t col = n0 + cl, row = m0 + rv;
            if (col >= g.N || row >= g.M) continue;
            const uint4 x = *reinterpret_cast<const uint4 *>(&T[cl][rv]);
            uint16_t *dst = g.CbT + col * g.ldt + row;
            if (row + 8 <= g.M) {
                *reinterpret_cast<uint4 *>(dst) = x;
            } else {
                const uint32_t w[4] = {x.x, x.y, x.z, x.w};
                for (int e = 0; e < (int)(g.M - row); ++e) dst[e] = (uint16_t)(w[e >> 1] >> ((e & 1) * 16));
            }
        }
    }
}

/* Sum of split-K partials: out[i] = sum_z part[z][i].  Block: 64 elements
 * x 4 groups; group q sums z = q, q + 4, ... in order, then the four group
 * sums are added in order q = 0..3 (a fixed order: deterministic). */
__global__ __launch_bounds__(256) void k_sum_partials(const float *part, int nz, int64_t n, float *out)
{
    __shared__ float red[4][64];
    const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + e;
    float s = 0.0f;
    if (i < n)
        for (int z = q; z < nz; z += 4) s += part[(int64_t)z * n + i];
    red[q][e] = s;
    __syncthreads();
    if (q == 0 && i < n) out[i] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

/* Every layer's bf16 weights transposed: [kp][LN] at the same offset. */
__global__ void k_transpose_params(const uint16_t *wb, uint16_t *wbT, LayerShapes L)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.off[5]) return;
    int l = 0;
    while (l < 4 && i >= L.off[l + 1]) ++l;
    const int64_t r = i - L.off[l];
    const int n = (int)(r / L.kp[l]), k = (int)(r % L.kp[l]);
    wbT[L.off[l] + (int64_t)k * L.ln[l] + n] = wb[i];
}

/* Rows [r0, r1) of a [rows][ld] bf16 matrix: 1 in row r0, 0 after. */
__global__ void k_bias_rows(uint16_t *a, int64_t ld, int r0, int r1)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)(r1 - r0) * ld) return;
    const int r = r0 + (int)(i / ld);
    a[(int64_t)r * ld + i % ld] = r == r0 ? (uint16_t)0x3f80 : (uint16_t)0;
}

/* Columns [c0, ld) of an activation buffer: 1 in column c0, 0 after. */
__global__ void k_bias_columns(uint16_t *a, int64_t M, int64_t ld, int c0)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int w = (int)(ld - c0);
    if (i >= M * w) return;
    int64_t r = i / w;
    int c = (int)(i % w);
    a[r * ld + c0 + c] = c == 0 ? (uint16_t)0x3f80 : (uint16_t)0;
}

/* ---- head: energy classes and a position mixture -------------------------
 * Layout in surr_common.h.  Training targets Y (one row of SURR_NY floats):
 * {event, energy class (surr_class of the entry energy and the deposited
 * fraction), depth, dy, dz}.  An event's likelihood is
 *   softmax(h[SURR_CLS ..])[class]
 *   * sum_k pi_k N(depth) N(dy) N(dz | mean and sigma of component k)
 * with pi the softmax of the full-absorption weight logits for class 0 and
 * of the continuum ones otherwise, so the deposited energy comes from a
 * histogram whose low bins follow the Compton edge of each entry energy
 * (a Gaussian mixture smears that edge and the full-absorption point mass),
 * and the first-deposit point from the mixture of its energy type. */

__device__ __forceinline__ float softplus(float x) { return x > 0.0f ? x + gp_logf(1.0f + gp_expf(-x)) : gp_logf(1.0f + gp_expf(x)); }

/* Per-sample loss and gradient dL/dh / n_total (bf16, cols >= SURR_NH zero)
 * from the head outputs transposed (HT: [output][ldt], a wave reads 64
 * consecutive rows of one output); one thread per row.  Each gradient
 * value is stored as it is computed into dH^T ([output][ldt], a wave's
 * stores cover 64 consecutive rows), so no row of gradients is held in
 * registers; k_rows_from_cols then writes the row-major dH.  Loss partial
 * sums per block to part[blockIdx.x] (fixed order). */
constexpr int LOSS_LD = 192;
__global__ __launch_bounds__(256) void k_head_loss_grad(const float *HT, int64_t ldt, const float *Y, int64_t M,
                                                        float inv_n, uint16_t *dHT, float *part)
{
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.0f;
    if (i < M) {
        auto hv = [&](int c) -> float { return HT[(int64_t)c * ldt + i]; };
        auto put = [&](int c, float v) { dHT[(int64_t)c * ldt + i] = f2bf(v * inv_n); };
        const float *y = Y + i * SURR_NY;
        const float ev = y[0];
        const float h0 = hv(0);
        const float p = 1.0f / (1.0f + gp_expf(-h0));
        loss = ev > 0.5f ? softplus(-h0) : softplus(h0);
        put(0, p - ev);
        if (ev > 0.5f) {
            /* energy class: softmax cross-entropy */
            const int cls = (int)y[1];
            float mc = -1e30f;
            for (int c = 0; c < SURR_NC; ++c) mc = fmaxf(mc, hv(SURR_CLS + c));
            float sc = 0.0f;
            for (int c = 0; c < SURR_NC; ++c) sc += gp_expf(hv(SURR_CLS + c) - mc);
            const float lse_c = mc + gp_logf(sc);
            loss += lse_c - hv(SURR_CLS + cls);
            for (int c = 0; c < SURR_NC; ++c)
                put(SURR_CLS + c, gp_expf(hv(SURR_CLS + c) - lse_c) - (c == cls ? 1.0f : 0.0f));
            /* first-deposit point: the mixture of the class's type; the
             * other type's weights get no gradient */
            const int wo = cls == 0 ? SURR_WPP : SURR_WCT;
            const int wz = cls == 0 ? SURR_WCT : SURR_WPP;
            float lw[SURR_K], lp[SURR_K];
            float mw = -1e30f, mx = -1e30f;
#pragma unroll
            for (int k = 0; k < SURR_K; ++k) {
                lw[k] = hv(wo + k);
                mw = fmaxf(mw, lw[k]);
            }
            float sw = 0.0f;
#pragma unroll
            for (int k = 0; k < SURR_K; ++k) sw += gp_expf(lw[k] - mw);
            const float lse_w = mw + gp_logf(sw);
#pragma unroll
            for (int k = 0; k < SURR_K; ++k) {
                float s = lw[k] - lse_w;
#pragma unroll
                for (int d = 0; d < SURR_D; ++d) {
                    const float mu = hv(SURR_MU + SURR_D * k + d);
                    const float ls = fminf(fmaxf(hv(SURR_LS + SURR_D * k + d), SURR_LOGSIG_MIN), SURR_LOGSIG_MAX);
                    const float zz = (y[2 + d] - mu) * gp_expf(-ls);
                    s += -0.5f * zz * zz - ls - 0.918938533f;
                }
                lp[k] = s;
                mx = fmaxf(mx, s);
            }
            float se = 0.0f;
#pragma unroll
            for (int k = 0; k < SURR_K; ++k) se += gp_expf(lp[k] - mx);
            const float lse = mx + gp_logf(se);
            loss += -lse;
            for (int k = 0; k < SURR_K; ++k) {
                const float gam = gp_expf(lp[k] - lse);
                put(wo + k, gp_expf(lw[k] - lse_w) - gam);
                put(wz + k, 0.0f);
#pragma unroll
                for (int d = 0; d < SURR_D; ++d) {
                    const float mu = hv(SURR_MU + SURR_D * k + d);
                    const float raw = hv(SURR_LS + SURR_D * k + d);
                    const float ls = fminf(fmaxf(raw, SURR_LOGSIG_MIN), SURR_LOGSIG_MAX);
                    const float is = gp_expf(-ls);
                    const float zz = (y[2 + d] - mu) * is;
                    put(SURR_MU + SURR_D * k + d, -gam * zz * is);
                    const bool free_ = raw > SURR_LOGSIG_MIN && raw < SURR_LOGSIG_MAX;
                    put(SURR_LS + SURR_D * k + d, free_ ? gam * (1.0f - zz * zz) : 0.0f);
                }
            }
        } else {
            /* no event: only the event logit has a gradient */
            for (int c = 1; c < SURR_NH; ++c) put(c, 0.0f);
        }
        for (int c = SURR_NH; c < LOSS_LD; ++c) put(c, 0.0f);
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* dH [M x LOSS_LD] (row-major, 16-byte stores) from dH^T [LOSS_LD x ldt]:
 * 64 rows per block through LDS. */
__global__ __launch_bounds__(256) void k_rows_from_cols(const uint16_t *in, int64_t ldt, int64_t M, uint16_t *out)
{
    __shared__ uint32_t t[64][LOSS_LD / 2 + 1];
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const int r = (int)threadIdx.x & 63;
    const bool ok = i0 + r < M;
    for (int c2 = (int)threadIdx.x >> 6; c2 < LOSS_LD / 2; c2 += 4) {
        const uint32_t lo = ok ? in[(int64_t)(2 * c2) * ldt + i0 + r] : 0u;
        const uint32_t hi = ok ? in[(int64_t)(2 * c2 + 1) * ldt + i0 + r] : 0u;
        t[r][c2] = lo | (hi << 16);
    }
    __syncthreads();
    for (int e = (int)threadIdx.x; e < 64 * (LOSS_LD / 8); e += 256) {
        const int rr = e / (LOSS_LD / 8), v = e % (LOSS_LD / 8);
        if (i0 + rr < M)
            *reinterpret_cast<uint4 *>(out + (i0 + rr) * LOSS_LD + 8 * v) =
                make_uint4(t[rr][4 * v], t[rr][4 * v + 1], t[rr][4 * v + 2], t[rr][4 * v + 3]);
    }
}

/* Squared norm of g: per-block partial sums (fixed order), then one block
 * turns them into the clipping factor min(1, clip / |g|). */
__global__ __launch_bounds__(256) void k_sqnorm_part(const float *g, int64_t n, float *part)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += g[i] * g[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_clip_factor(const float *part, int nb, float clip, float *factor)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = (int)threadIdx.x; i < nb; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float nrm = sqrtf(red[0]);
        factor[0] = nrm > clip ? clip / nrm : 1.0f;
    }
}

/* Adam on fp32 master weights (gradient scaled by *gscale); refreshes the
 * bf16 copy. */
__global__ void k_adam(float *w, float *m, float *v, const float *g, const float *gscale, uint16_t *wb, int64_t n,
                       float lr, float b1, float b2, float c1, float c2, float eps)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i] * gscale[0];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

/* Sum of n loss partials into out[0] (one block, fixed order). */
__global__ void k_loss_reduce(const float *part, int n, float *out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

/* Training minibatch: M rows drawn with replacement from the n harvested
 * records (Philox stream (step, 14), one draw per row) -> input matrix
 * (bf16 [M x 8]: 6 features, 1, 0) and targets (fp32 [M x SURR_NY], the
 * deposited fraction replaced by its energy class). */
__global__ void k_pack_train(const float *set, int64_t n, uint64_t seed, int64_t step, int64_t M, uint16_t *X,
                             float *Y, uint16_t *XT, int64_t ldt, int refl)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    gp_u4 c = {(uint32_t)i, 14u, (uint32_t)step, (uint32_t)(step >> 32)};
    gp_u4 rr = gp_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t u = ((uint64_t)rr.x << 32) | rr.y;
    const int64_t j = (int64_t)(u % (uint64_t)n);
    const float *r = set + j * SURR_REC;
    /* mirror images of the record (half the rows each) where the panels'
     * crystal grid is its own mirror image: (y, vy, dy) -> (-y, -vy, -dy),
     * and the same in z */
    const bool ry = (refl & 1) && (rr.z & 1u), rz = (refl & 2) && (rr.z & 2u);
#pragma unroll
    for (int c = 0; c < SURR_LDX; ++c) {
        float v = c < SURR_NX ? r[c] : 0.0f;
        if ((c == 2 || c == 4) && ry) v = -v;
        if ((c == 3 || c == 5) && rz) v = -v;
        const uint16_t x = c < SURR_NX ? f2bf(v) : (c == SURR_NX ? (uint16_t)0x3f80 : (uint16_t)0);
        X[i * SURR_LDX + c] = x;
        if (XT) XT[(int64_t)c * ldt + i] = x;
    }
#pragma unroll
    for (int c = 0; c < SURR_NY; ++c) Y[i * SURR_NY + c] = r[SURR_NX + c];
    Y[i * SURR_NY + 1] = (float)surr_class(r[0], r[SURR_NX + 1]);
    if (ry) Y[i * SURR_NY + 3] = -r[SURR_NX + 3];
    if (rz) Y[i * SURR_NY + 4] = -r[SURR_NX + 4];
}

/* Elementary functions of the sampling: the hardware approximations
 * (v_exp_f32 = 2^x, v_log_f32 = log2, v_sin/v_cos_f32 of x turns,
 * v_sqrt_f32, v_rcp_f32; about 1 ulp).  The sampling is a statistical
 * model checked against plain MC, not restated by the oracle, so it needs
 * no correctly rounded library routine. */
__device__ __forceinline__ float s_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ __forceinline__ float s_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
__device__ __forceinline__ float s_sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + s_exp(-x)); }

/* Sampling from the head outputs h of surrogate-queue row j: P(event) on
 * the photon's own stream, an energy class by its weight and a fraction
 * uniform in its bin, a position component by its weight among those of
 * the class's type, three normals; the event at the first deposit point,
 * with the readout site of the digitizer settings (one event per photon).
 * Returns whether the photon gives an event (ev filled). */
typedef _Float16 h16x4_t __attribute__((ext_vector_type(4)));
static_assert(SURR_CLS + SURR_NC <= 64, "event and class logits in the first 64 outputs");
static_assert((SURR_WPP & 3) == (SURR_WCT & 3) && (SURR_WCT & ~3) + 20 <= SURR_NH,
              "both weight sets at the same offset within five 8-byte reads of the row");
__device__ __forceinline__ bool d_surr_sample_row(const _Float16 *h, const DevQueue &q, int64_t j,
                                                  const DevDetector &det, const PanelParams &pp, gpet_event &ev)
{
    const int4 info = q.info[j];
    const DevPanel &P = det.p[info.x];
    const unsigned draws = (unsigned)info.w & 0x7fffffffu;
    const unsigned k = (unsigned)info.w >> 31;
    gp_rng r;
    gp_rng_init(&r, pp.seed, (uint64_t)(int64_t)info.z, GP_TAG_PHOTON + k, draws);
    gp_u4 b = gp_rng_block(&r);
    /* outputs 0 .. 63 (event and class logits) in registers: 16 8-byte LDS
     * reads (the row starts 8-byte aligned) */
    h16x4_t hv[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) hv[e] = *reinterpret_cast<const h16x4_t *>(h + 4 * e);
    auto out = [&](int c) -> float { return (float)hv[c >> 2][c & 3]; };
    const float pe = s_sigmoid(out(0));
    if (!(gp_u01(b.x) <= pe)) return false;
    /* energy class: the first whose running weight reaches u (the last if
     * rounding leaves u above the total) */
    float mc = -1e30f;
#pragma unroll
    for (int c = 0; c < SURR_NC; ++c) mc = fmaxf(mc, out(SURR_CLS + c));
    float ew[SURR_NC];
    float sc = 0.0f;
#pragma unroll
    for (int c = 0; c < SURR_NC; ++c) {
        ew[c] = s_exp(out(SURR_CLS + c) - mc);
        sc += ew[c];
    }
    const float uc = gp_u01(b.y) * sc;
    float acc = 0.0f;
    int cls = SURR_NC - 1;
    bool found = false;
#pragma unroll
    for (int c = 0; c < SURR_NC; ++c) {
        acc += ew[c];
        if (!found && acc >= uc) {
            cls = c;
            found = true;
        }
    }
    const float E = q.E[j];
    const float fe = surr_class_fe(E * (1.0f / 511.0e3f), cls, gp_u01(b.z));
    /* position component among those of the class's type: its 16 weight
     * logits from five 8-byte LDS reads at the 4-aligned output below them */
    const int wo = cls == 0 ? SURR_WPP : SURR_WCT;
    const int wa = wo & ~3;
    constexpr int ws = SURR_WPP & 3;
    h16x4_t wv[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) wv[e] = *reinterpret_cast<const h16x4_t *>(h + wa + 4 * e);
    float lw[SURR_K];
    float mw = -1e30f;
#pragma unroll
    for (int c = 0; c < SURR_K; ++c) {
        lw[c] = (float)wv[(c + ws) >> 2][(c + ws) & 3];
        mw = fmaxf(mw, lw[c]);
    }
    float sw = 0.0f;
#pragma unroll
    for (int c = 0; c < SURR_K; ++c) {
        lw[c] = s_exp(lw[c] - mw);
        sw += lw[c];
    }
    const float u = gp_u01(b.w) * sw;
    acc = 0.0f;
    int kc = SURR_K - 1;
    found = false;
#pragma unroll
    for (int c = 0; c < SURR_K; ++c) {
        acc += lw[c];
        if (!found && u <= acc) {
            kc = c;
            found = true;
        }
    }
    /* the component's Gaussian restricted to where first deposits happen:
     * a depth in [0, 1] and a point inside a crystal of the panel (a photon
     * that crosses a gap or leaves through a side does not deposit there);
     * up to 8 draws of three normals, else no event */
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
    int M_id = -1, L_id = -1;
    bool has = false;
    for (int tries = 0; tries < 8 && !has; ++tries) {
        gp_u4 b2 = gp_rng_block(&r);
        float y[SURR_D], z[4];
        const float a0 = gp_u01(b2.y), a1 = gp_u01(b2.w);
        const float s0 = __builtin_amdgcn_sinf(a0), c0 = __builtin_amdgcn_cosf(a0);
        const float c1 = __builtin_amdgcn_cosf(a1);
        const float r0 = __builtin_amdgcn_sqrtf(-2.0f * s_log(gp_u01(b2.x)));
        const float r1 = __builtin_amdgcn_sqrtf(-2.0f * s_log(gp_u01(b2.z)));
        z[0] = r0 * c0; z[1] = r0 * s0; z[2] = r1 * c1;
#pragma unroll
        for (int d = 0; d < SURR_D; ++d) {
            const float ls = fminf(fmaxf((float)h[SURR_LS + SURR_D * kc + d], SURR_LOGSIG_MIN), SURR_LOGSIG_MAX);
            y[d] = (float)h[SURR_MU + SURR_D * kc + d] + s_exp(ls) * z[d];
        }
        if (y[0] >= 0.0f && y[0] <= 1.0f &&
            surr_position(P, q.vx[j], q.vy[j], q.vz[j], q.y[j], q.z[j], y[0], y[1], y[2], lx, ly, lz))
            has = d_crystal(det, P, lx, ly, lz, M_id, L_id) == 0;
    }
    if (!(has && fe > 0.0f)) return false;
    const int pann = P.panel;
    float dx = lx, dy = ly - q.y[j], dz = lz - q.z[j];
    ev.parn = info.y;
    ev.pann = pann;
    ev.modn = M_id;
    ev.cryn = L_id;
    ev.siten = surr_site(pp.readout_depth, pp.readout_policy, pann, M_id, L_id, det.moduleN, det.crystalN);
    ev.eventid = info.z;
    ev.t = q.t[j] + (double)sqrtf(dx * dx + dy * dy + dz * dz) * D_INV_C;
    ev.E = fe * E;
    ev.x = lx;
    ev.y = ly;
    ev.z = lz;
    return true;
}

/* Wave-aggregated append of the lanes' events (every lane of the wave
 * calls it). */
__device__ __forceinline__ void d_append_event(const PanelParams &pp, bool has, const gpet_event &ev)
{
    unsigned long long m = __ballot(has);
    unsigned base = 0;
    if (m) {
        int leader = __ffsll((unsigned long long)m) - 1;
        if ((int)__lane_id() == leader) base = atomicAdd(pp.n_events, (unsigned)__popcll(m));
        base = __shfl(base, leader);
    }
    unsigned long long lt = __lane_id() == 0 ? 0ull : (~0ull >> (64 - __lane_id()));
    if (has) {
        int64_t dst = (int64_t)base + __popcll(m & lt);
        if (dst < pp.events_cap) pp.events[dst] = ev;
    }
}

/* ---- fused inference ------------------------------------------------------
 * One block of 256 threads (4 waves) runs the whole network over IB = 128
 * rows: the activations stay in LDS (bf16 [128][264]; column 256 holds the
 * constant 1 that carries the bias, 257..263 zero), the weights stream
 * through the L2 into registers, and the products run on
 * v_mfma_f32_32x32x16_bf16 with the weights as the first operand (32
 * neurons) and the activations as the second (32 rows).  A lane's
 * accumulators are then 4 consecutive neurons of one row, so the SiLU and
 * the bf16 rounding happen in registers and every 4 outputs go to LDS in
 * one 8-byte write.  Hidden layers: wave w computes neurons [64w, 64w + 64)
 * of all 128 rows (2 x 4 tiles, 128 accumulators).  Head: wave w computes
 * outputs [48w, 48w + 48) of the 192 stored (190 used) for all 128 rows on
 * v_mfma_f32_16x16x32_bf16, so each head weight is loaded once per tile;
 * the outputs replace the activations in LDS, all rows at once as fp16
 * when the rows are sampled (one lane per row on two waves), or half the
 * rows at a time as fp32 when they are copied to H.  Inputs: the
 * surrogate queue (features packed in the prologue) or, without SAMPLE, a
 * packed bf16 [n x 8] matrix.  The SiLU here is z / (1 + 2^(-z log2 e))
 * with the hardware exp2 and reciprocal (training uses the exact form; the
 * difference is below the bf16 rounding of the activations). */
constexpr int IB = 128;
constexpr int ILD = 264;
constexpr int IHO = 3;            /* head: 16-output tiles per wave (4 x 48 = 192 >= SURR_NH) */
constexpr int IPD = 3;            /* hidden layers: weight loads this many steps ahead */
constexpr int IHPD = 2;           /* head: the same */
constexpr int IHLD = 196;         /* head row in LDS: >= SURR_NH, 16-byte rows (fp32), 4 banks apart */
static_assert(IHLD >= SURR_NH && IHLD % 4 == 0 && 4 * 48 >= SURR_NH, "head tiles and LDS rows");
/* LDS of a block of rg row groups: the activations, which the head outputs
 * of half of each group's rows at a time replace */
constexpr int infer_lds(int rg)
{
    return rg * (IB / 2) * IHLD * 4 > rg * IB * ILD * 2 ? rg * (IB / 2) * IHLD * 4 : rg * IB * ILD * 2;
}
/* Row groups per block (GPET_INFER_RG): 2 gives 256-row tiles on 8 waves,
 * so the two waves that compute the same neurons for different rows load
 * the same weights (the second from L1), halving the weight stream from
 * L2.  Measured on MI355X it is slower (11.1 vs 9.9 ms per cylinder
 * step, identical outputs): with 135 KB of LDS a CU holds one such block,
 * and its prologue, barriers and sampling no longer overlap with another
 * block's products.  One group (two 4-wave blocks per CU) is kept. */
#ifndef GPET_INFER_RG
#define GPET_INFER_RG 1
#endif

struct InferWeights {
    const uint16_t *w[5];      /* layer l: [outputs][kp(l)] bf16              */
};

__device__ __forceinline__ float silu_fast(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}

/* two floats to packed bf16 (round to nearest even, a in the low half):
 * one v_cvt_pk_bf16_f32 */
__device__ __forceinline__ uint32_t f2bf2(float a, float b)
{
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}

__device__ __forceinline__ bf16x8_t load_w(const uint16_t *p, bool ok)
{
    uint4 v = ok ? *reinterpret_cast<const uint4 *>(p) : make_uint4(0, 0, 0, 0);
    return __builtin_bit_cast(bf16x8_t, v);
}

template <bool SAMPLE, int RG>
__global__ __launch_bounds__(256 * RG) __attribute__((amdgpu_waves_per_eu(2))) void k_mlp_infer(InferWeights W, const uint16_t *X, DevQueue q, int64_t row0,
                                                   int64_t n, DevDetector det, PanelParams pp, float *H, int64_t ldh)
{
    __shared__ __attribute__((aligned(16))) unsigned char s_buf[infer_lds(RG)];
    uint16_t *act = reinterpret_cast<uint16_t *>(s_buf);
    float *head = reinterpret_cast<float *>(s_buf);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    /* row group rg: rows [IB rg, IB rg + IB) of the tile; wv: the wave's
     * share of the neurons (hidden layers) or outputs (head) */
    const int rg = wave >> 2, wv = wave & 3;
    constexpr int IBR = IB * RG;
    const int64_t b0 = (int64_t)blockIdx.x * IBR;

    /* prologue: inputs in columns 0..7, the bias column 256 and zeros */
    if (threadIdx.x < IBR) {
        const int m = threadIdx.x;
        const int64_t i = b0 + m;
        uint16_t in[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (i < n) {
            if (SAMPLE) {
                const int64_t j = row0 + i;
                const DevPanel &P = det.p[q.info[j].x];
                float f[SURR_NX];
                surr_features(P, q.E[j], q.vx[j], q.vy[j], q.vz[j], q.y[j], q.z[j], f);
#pragma unroll
                for (int c = 0; c < SURR_NX; ++c) in[c] = f2bf(f[c]);
                in[SURR_NX] = 0x3f80;
            } else {
#pragma unroll
                for (int c = 0; c < 8; ++c) in[c] = X[i * SURR_LDX + c];
            }
        }
        uint16_t *row = act + m * ILD;
#pragma unroll
        for (int c = 0; c < 8; ++c) row[c] = in[c];
        row[256] = 0x3f80;
#pragma unroll
        for (int c = 257; c < ILD; ++c) row[c] = 0;
    }
    __syncthreads();

    const int lr = lane & 31, kh = (lane >> 5) * 8;
    /* hidden layers */
#pragma unroll 1
    for (int l = 0; l < 4; ++l) {
        const uint16_t *w = W.w[l];
        f32x16_t acc[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[t][mt] = f32x16_t{};
        const int nrow0 = 64 * wv + lr, nrow1 = nrow0 + 32;
        if (l == 0) {
            /* 8 input columns: one step, lanes 32..63 (k 8..15) zero */
            const bool ok = kh < SURR_LDX;
            const bf16x8_t a0 = load_w(w + (int64_t)nrow0 * SURR_LDX + kh, ok);
            const bf16x8_t a1 = load_w(w + (int64_t)nrow1 * SURR_LDX + kh, ok);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                bf16x8_t bv = *reinterpret_cast<const bf16x8_t *>(act + (IB * rg + mt * 32 + lr) * ILD);
                if (!ok) bv = bf16x8_t{};
                acc[0][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bv, acc[0][mt], 0, 0, 0);
                acc[1][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bv, acc[1][mt], 0, 0, 0);
            }
        } else {
            /* 17 steps of 16 columns (the last: the bias column, zeros, and
             * lanes 32..63 past the row set to zero); the weights of step
             * st + IPD are loaded while step st multiplies */
            constexpr int NS = (ILD + 15) / 16;
            bf16x8_t ab[IPD][2];
#pragma unroll
            for (int p = 0; p < IPD; ++p) {
                const int k = p * 16 + kh;
                ab[p][0] = load_w(w + (int64_t)nrow0 * ILD + k, k < ILD);
                ab[p][1] = load_w(w + (int64_t)nrow1 * ILD + k, k < ILD);
            }
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                const int sl = st % IPD;
                const bf16x8_t a0 = ab[sl][0], a1 = ab[sl][1];
                if (st + IPD < NS) {
                    const int kf = (st + IPD) * 16 + kh;
                    ab[sl][0] = load_w(w + (int64_t)nrow0 * ILD + kf, kf < ILD);
                    ab[sl][1] = load_w(w + (int64_t)nrow1 * ILD + kf, kf < ILD);
                }
                const int k = st * 16 + kh;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    bf16x8_t bv =
                        *reinterpret_cast<const bf16x8_t *>(act + (IB * rg + mt * 32 + lr) * ILD + (k < ILD ? k : 0));
                    if (k >= ILD) bv = bf16x8_t{};
                    acc[0][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bv, acc[0][mt], 0, 0, 0);
                    acc[1][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bv, acc[1][mt], 0, 0, 0);
                }
            }
        }
        __syncthreads();
        /* accumulator i of lane: neuron (i/4)*8 + (lane/32)*4 + i%4 of the
         * tile, row lane%32 of the tile */
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int nb = 64 * wv + 32 * t + g * 8 + (lane >> 5) * 4;
                    const int m = IB * rg + mt * 32 + lr;
                    const uint32_t lo = f2bf2(silu_fast(acc[t][mt][4 * g]), silu_fast(acc[t][mt][4 * g + 1]));
                    const uint32_t hi = f2bf2(silu_fast(acc[t][mt][4 * g + 2]), silu_fast(acc[t][mt][4 * g + 3]));
                    *reinterpret_cast<uint2 *>(act + m * ILD + nb) = make_uint2(lo, hi);
                }
        __syncthreads();
    }

    /* head: outputs [48 wave, 48 wave + 48) of all 128 rows on
     * v_mfma_f32_16x16x32_bf16 (3 output tiles x 8 row tiles of 16), so
     * each head weight is loaded by one wave; lanes hold 8 consecutive k of
     * output (or row) lane % 16, k block lane / 16 */
    {
        const uint16_t *w = W.w[4];
        const int l16 = lane & 15, kq = (lane >> 4) * 8;
        f32x4_t acc[IHO][8];
#pragma unroll
        for (int t = 0; t < IHO; ++t)
#pragma unroll
            for (int rt = 0; rt < 8; ++rt) acc[t][rt] = f32x4_t{};
        constexpr int NS = (ILD + 31) / 32;
        const uint16_t *wr = w + (int64_t)(48 * wv + l16) * ILD;
        bf16x8_t ab[IHPD][IHO];
#pragma unroll
        for (int p = 0; p < IHPD; ++p)
#pragma unroll
            for (int t = 0; t < IHO; ++t) {
                const int k = p * 32 + kq;
                ab[p][t] = load_w(wr + 16 * t * ILD + k, k < ILD);
            }
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const int sl = st % IHPD;
            const int k = st * 32 + kq;
            const bool ok = k < ILD;
            bf16x8_t a[IHO];
#pragma unroll
            for (int t = 0; t < IHO; ++t) {
                a[t] = ab[sl][t];
                if (st + IHPD < NS) {
                    const int kf = (st + IHPD) * 32 + kq;
                    ab[sl][t] = load_w(wr + 16 * t * ILD + kf, kf < ILD);
                }
            }
#pragma unroll
            for (int rt = 0; rt < 8; ++rt) {
                bf16x8_t bv = *reinterpret_cast<const bf16x8_t *>(act + (IB * rg + rt * 16 + l16) * ILD + (ok ? k : 0));
                if (!ok) bv = bf16x8_t{};
#pragma unroll
                for (int t = 0; t < IHO; ++t)
                    acc[t][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bv, acc[t][rt], 0, 0, 0);
            }
        }
        if (SAMPLE) {
            /* every row's outputs at once, as fp16 (2^-11 relative: a
             * deposited-energy mean moves by < 0.3 keV), replace the
             * activations (accumulator i: output 4 (lane / 16) + i of the
             * tile, row lane % 16); then waves 0 .. 2 RG - 1 (one per SIMD)
             * sample 64 rows each, one per lane */
            typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));
            _Float16 *head16 = reinterpret_cast<_Float16 *>(s_buf);
            __syncthreads();
#pragma unroll
            for (int t = 0; t < IHO; ++t) {
                const int nb = 48 * wv + 16 * t + (lane >> 4) * 4;
                if (nb < IHLD) {
#pragma unroll
                    for (int rt = 0; rt < 8; ++rt) {
                        f32x4_t v = acc[t][rt];
#pragma unroll
                        for (int e = 0; e < 4; ++e) v[e] = fminf(fmaxf(v[e], -65000.0f), 65000.0f);
                        *reinterpret_cast<f16x4_t *>(head16 + (IB * rg + rt * 16 + l16) * IHLD + nb) =
                            __builtin_convertvector(v, f16x4_t);
                    }
                }
            }
            __syncthreads();
            if (wave < 2 * RG) {
                const int row = 64 * wave + lane;
                const int64_t i = b0 + row;
                bool has = false;
                gpet_event ev;
                if (i < n) has = d_surr_sample_row(head16 + row * IHLD, q, row0 + i, det, pp, ev);
                d_append_event(pp, has, ev);
            }
            return;
        }
        /* outputs to H (fp32): rows [0, 64) (row tiles 0-3), then rows
         * [64, 128) (4-7): every wave writes its outputs of the half's rows
         * to LDS and wave g copies row group g's 64 rows, one per lane */
#pragma unroll 1
        for (int half = 0; half < 2; ++half) {
            __syncthreads();
#pragma unroll
            for (int t = 0; t < IHO; ++t) {
                const int nb = 48 * wv + 16 * t + (lane >> 4) * 4;
                if (nb < IHLD) {
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const f32x4_t v = half ? acc[t][4 + r4] : acc[t][r4];
                        *reinterpret_cast<float4 *>(head + (64 * rg + r4 * 16 + l16) * IHLD + nb) =
                            make_float4(v[0], v[1], v[2], v[3]);
                    }
                }
            }
            __syncthreads();
            if (wave < RG) {
                /* wave g (on SIMD g) handles row group g */
                const int64_t i = b0 + IB * wave + 64 * half + lane;
                const float *h = head + (64 * wave + lane) * IHLD;
                if (i < n)
                    for (int c = 0; c < SURR_NH; ++c) H[i * ldh + c] = h[c];
            }
        }
    }
}

/* ---- launchers ------------------------------------------------------------ */

int gpet_gemm(int kind, const GemmArgs &g, int splits, hipStream_t s)
{
    dim3 grid((unsigned)((g.M + TM - 1) / TM), (unsigned)((g.N + TN - 1) / TN), (unsigned)(splits > 0 ? splits : 1));
    GemmArgs a = g;
    if (a.k_chunk <= 0) a.k_chunk = a.K;
    if ((a.lda | a.ldb) % 8 != 0 || a.k_chunk % TK != 0 && a.k_chunk < a.K) return (int)hipErrorInvalidValue;
    switch (kind) {
    case EPI_F32: hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, grid, dim3(256), 0, s, a); break;
    case EPI_SILU: hipLaunchKernelGGL(k_gemm_nt<EPI_SILU>, grid, dim3(256), 0, s, a); break;
    case EPI_ACT: hipLaunchKernelGGL(k_gemm_nt<EPI_ACT>, grid, dim3(256), 0, s, a); break;
    case EPI_DSILU: hipLaunchKernelGGL(k_gemm_nt<EPI_DSILU>, grid, dim3(256), 0, s, a); break;
    case EPI_PARTIAL: hipLaunchKernelGGL(k_gemm_nt<EPI_PARTIAL>, grid, dim3(256), 0, s, a); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

int gpet_sum_partials(const float *part, int nz, int64_t n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, part, nz, n, out);
    return (int)hipGetLastError();
}

int gpet_bias_columns(uint16_t *a, int64_t M, int64_t ld, int c0, hipStream_t s)
{
    int64_t n = M * (ld - c0);
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_bias_columns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, M, ld, c0);
    return (int)hipGetLastError();
}

int gpet_head_loss_grad(const float *HT, int64_t ldt, const float *Y, int64_t M, float inv_n, uint16_t *dH,
                       int64_t lddh, uint16_t *dHT, float *part, hipStream_t s)
{
    if (lddh != LOSS_LD) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_head_loss_grad, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, HT, ldt, Y, M, inv_n,
                       dHT, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_rows_from_cols, dim3((unsigned)((M + 63) / 64)), dim3(256), 0, s, dHT, ldt, M, dH);
    return (int)hipGetLastError();
}

int gpet_transpose_params(const uint16_t *wb, uint16_t *wbT, const LayerShapes &L, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose_params, dim3((unsigned)((L.off[5] + 255) / 256)), dim3(256), 0, s, wb, wbT, L);
    return (int)hipGetLastError();
}

int gpet_bias_rows(uint16_t *a, int64_t ld, int r0, int r1, hipStream_t s)
{
    const int64_t n = (int64_t)(r1 - r0) * ld;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_bias_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, ld, r0, r1);
    return (int)hipGetLastError();
}

int gpet_clip_factor(const float *g, int64_t n, float clip, float *part, float *factor, hipStream_t s)
{
    const int nb = 240;
    hipLaunchKernelGGL(k_sqnorm_part, dim3(nb), dim3(256), 0, s, g, n, part);
    hipLaunchKernelGGL(k_clip_factor, dim3(1), dim3(256), 0, s, part, nb, clip, factor);
    return (int)hipGetLastError();
}

int gpet_adam(float *w, float *m, float *v, const float *g, const float *gscale, uint16_t *wb, int64_t n, float lr,
              float b1, float b2, float c1, float c2, float eps, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, m, v, g, gscale, wb, n, lr, b1,
                       b2, c1, c2, eps);
    return (int)hipGetLastError();
}

int gpet_pack_train(const float *set, int64_t n, uint64_t seed, int64_t step, int64_t M, uint16_t *X, float *Y,
                    uint16_t *XT, int64_t ldt, int refl, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_train, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, set, n, seed, step, M, X, Y,
                       XT, ldt, refl);
    return (int)hipGetLastError();
}

int gpet_loss_reduce(const float *part, int n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, s, part, n, out);
    return (int)hipGetLastError();
}

int gpet_mlp_infer(const uint16_t *const w[5], const uint16_t *X, const DevQueue &sq, int64_t row0, int64_t n,
                   const DevDetector &det, const PanelParams &pp, float *H, int64_t ldh, hipStream_t s)
{
    if (n <= 0) return 0;
    InferWeights W;
    for (int l = 0; l < 5; ++l) W.w[l] = w[l];
    constexpr int RG = GPET_INFER_RG;
    const dim3 grid((unsigned)((n + IB * RG - 1) / (IB * RG)));
    if (X)
        hipLaunchKernelGGL((k_mlp_infer<false, RG>), grid, dim3(256 * RG), 0, s, W, X, sq, row0, n, det, pp, H, ldh);
    else
        hipLaunchKernelGGL((k_mlp_infer<true, RG>), grid, dim3(256 * RG), 0, s, W, X, sq, row0, n, det, pp, H, ldh);
    return (int)hipGetLastError();
}

