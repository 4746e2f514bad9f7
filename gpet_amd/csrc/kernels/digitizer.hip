/*
 * digitizer.hip - digitizer chain and coincidence sorter on the GPU.
 *
 * The reference runs blur/threshold/deadtime/window as kernels but compacts
 * and orders events by copying the whole event array to the host and calling
 * std::sort four times per batch (gPET.cu:380-424, detector.cu:354-385).
 * Here the chain stays in HBM:
 *
 *   blur + thresholder      one lane per event, Philox stream (parn, siten)
 *   compaction              block scan + scatter (stable)
 *   time order              LSD radix sort of (sortable t bits, index):
 *                           8 passes of 8 bits, ranks inside a wavefront
 *                           from 8 ballots per digit, tiles staged in LDS
 *                           in digit order before the coalesced write
 *   deadtime                stable radix sort by site over the time order,
 *                           then one lane per chain walks it
 *                           (gPET_kernals.cu:657-698)
 *   energy window           flags + compaction in time order -> singles
 *   coincidences            clusters of singles with gaps <= window are
 *                           independent; one lane per cluster replays the
 *                           single-window rule, counts, scan, write.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>
#include "../common/gp_math.h"
#include "../host/internal.h"

namespace {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) {                                                                        \
            snprintf(err, errlen, "%s:%d %s", __FILE__, __LINE__, hipGetErrorString(e_));              \
            return -1;                                                                                 \
        }                                                                                              \
    } while (0)

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;

__device__ __forceinline__ unsigned long long lanemask_lt()
{
    unsigned int l = __lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

/* ---- exclusive scan of uint32 ------------------------------------------ */

__device__ unsigned block_exclusive_scan(unsigned v, unsigned *sh, unsigned &total)
{
    /* wave inclusive scan via shuffles, then across 4 waves */
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
        unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            unsigned t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[8] = acc;
    }
    __syncthreads();
    unsigned res = sh[wid] + x - v;
    total = sh[8];
    __syncthreads();
    return res;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const unsigned *in, unsigned *out, unsigned *tile_sums, int64_t n)
{
    __shared__ unsigned sh[16];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    unsigned v[SCAN_ITEMS];
    unsigned s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    unsigned tot;
    unsigned pre = block_exclusive_scan(s, sh, tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

/* Single block scans the tile sums in chunks, carrying the running total. */
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(unsigned *sums, int64_t n, unsigned *grand)
{
    __shared__ unsigned sh[16];
    unsigned carry = 0;
    for (int64_t c = 0; c < n; c += SCAN_BLOCK) {
        int64_t i = c + threadIdx.x;
        unsigned v = (i < n) ? sums[i] : 0u;
        unsigned tot;
        unsigned pre = block_exclusive_scan(v, sh, tot);
        if (i < n) sums[i] = pre + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && grand) *grand = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(unsigned *out, const unsigned *tile_sums, int64_t n)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    unsigned add = tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) out[base + k] += add;
}

/* ---- radix sort of (uint64 key, uint32 value), stable ------------------ */

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const unsigned long long *keys, int64_t n, int shift, unsigned *hist,
                                                      int nblocks)
{
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        int64_t i = base + (int64_t)r * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(unsigned)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

/* One tile per block: items are ranked stably by digit (8 ballots per
 * round of 256), staged in LDS in digit order, then written out so that
 * consecutive threads write consecutive addresses of each digit's run. */
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const unsigned long long *kin, const unsigned *vin,
                                                         unsigned long long *kout, unsigned *vout, int64_t n, int shift,
                                                         const unsigned *offs, int nblocks)
{
    __shared__ unsigned long long s_key[RS_TILE];
    __shared__ unsigned s_val[RS_TILE];
    __shared__ unsigned run[256], gstart[256], lstart[256], wsum[RS_BLOCK / 64];
    __shared__ unsigned wcnt[RS_BLOCK / 64][256];
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    /* this tile's count of digit threadIdx.x from the scanned histogram
     * (entry d * nblocks + tile is the digit's global start) */
    const int64_t hn = (int64_t)256 * nblocks;
    const int64_t e = (int64_t)threadIdx.x * nblocks + blockIdx.x;
    const unsigned g0 = offs[e];
    const unsigned c = ((e + 1 < hn) ? offs[e + 1] : (unsigned)n) - g0;
    gstart[threadIdx.x] = g0;
    unsigned v = c;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    if (lane == 63) wsum[wid] = v;
    for (int w = 0; w < RS_BLOCK / 64; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    {
        unsigned add = 0;
        for (int w = 0; w < wid; ++w) add += wsum[w];
        lstart[threadIdx.x] = v - c + add;
        run[threadIdx.x] = v - c + add;
    }
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const unsigned long long lt = lanemask_lt();
    for (int r = 0; r < RS_ROUNDS; ++r) {
        int64_t i = base + (int64_t)r * RS_BLOCK + threadIdx.x;
        bool valid = i < n;
        unsigned long long key = valid ? kin[i] : 0ull;
        unsigned val = valid ? vin[i] : 0u;
        unsigned d = (unsigned)(key >> shift) & 255u;
        unsigned long long match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            unsigned long long bb = __ballot(valid && bit);
            match &= bit ? bb : ~bb;
        }
        unsigned rank = (unsigned)__popcll(match & lt);
        if (valid && rank == 0) wcnt[wid][d] = (unsigned)__popcll(match);
        __syncthreads();
        unsigned pos = 0;
        if (valid) {
            pos = run[d] + rank;
            for (int w = 0; w < wid; ++w) pos += wcnt[w][d];
        }
        __syncthreads();
        {
            unsigned add = 0;
            for (int w = 0; w < RS_BLOCK / 64; ++w) {
                add += wcnt[w][threadIdx.x];
                wcnt[w][threadIdx.x] = 0;
            }
            run[threadIdx.x] += add;
        }
        if (valid) {
            s_key[pos] = key;
            s_val[pos] = val;
        }
        __syncthreads();
    }
    const int tn = (int)min<int64_t>((int64_t)RS_TILE, n - base);
    for (int j = threadIdx.x; j < tn; j += RS_BLOCK) {
        const unsigned long long key = s_key[j];
        const unsigned d = (unsigned)(key >> shift) & 255u;
        const unsigned g = gstart[d] + ((unsigned)j - lstart[d]);
        kout[g] = key;
        vout[g] = s_val[j];
    }
}

/* ---- digitizer stages --------------------------------------------------- */

__device__ __forceinline__ unsigned long long sortable_t(double t)
{
    unsigned long long b = gp_d2u(t);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

/* blur (gPET_kernals.cu:814-837) + time blur + thresholder flags. */
/* Blur and thresholder; flag = passes the thresholder.  With nthr set (no
 * deadtime: the energy window does not depend on time order, so it is
 * applied here, before the time sort) flag = passes the thresholder and the
 * window, and nthr counts the thresholder's passes (one atomic per block). */
__global__ __launch_bounds__(256) void k_blur_threshold(const gpet_event *in, gpet_event *out, unsigned *flag, int64_t n,
                                                        uint64_t seed, gpet_digitizer g, unsigned *nthr)
{
    __shared__ unsigned s_thr;
    if (threadIdx.x == 0) s_thr = 0u;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool thr = false;
    if (i < n) {
        gpet_event e = in[i];
        gp_rng r;
        gp_rng_init(&r, seed, (uint64_t)(uint32_t)e.parn, (uint32_t)GP_TAG_DIGI | ((uint32_t)e.siten << 4), 0u);
        float R = 0.0f;
        if (g.blur_policy == 0) R = sqrtf(g.e_ref / e.E) * g.r_ref;
        if (g.blur_policy == 1) R = g.r_ref + g.slope * (e.E - g.e_ref) / 1.0e6f;
        if (R < 0.0f) R = 0.0f;
        float nr = gp_rng_normal(&r);
        e.E = e.E + ((nr * R) * e.E) / 2.35482f;
        if (g.space_blur > 0.0f) {
            e.x = e.x + g.space_blur * gp_rng_normal(&r);
            e.y = e.y + g.space_blur * gp_rng_normal(&r);
            e.z = e.z + g.space_blur * gp_rng_normal(&r);
        }
        if (g.time_blur_fwhm_us > 0.0) {
            float nt = gp_rng_normal(&r);
            e.t = e.t + (double)nt * (g.time_blur_fwhm_us / 2.35482);
        }
        out[i] = e;
        thr = !(e.E < g.e_threshold || e.E > g.e_upholder);
        flag[i] = (thr && (!nthr || !(e.E < g.win_min || e.E > g.win_max))) ? 1u : 0u;
    }
    if (nthr) {
        const unsigned long long b = __ballot(thr);
        if (__lane_id() == 0 && b) atomicAdd(&s_thr, (unsigned)__popcll(b));
        __syncthreads();
        if (threadIdx.x == 0 && s_thr) atomicAdd(nthr, s_thr);
    }
}

/* Compaction (input order kept); with key set it also writes each kept
 * event's time key and its index, the time sort's input. */
__global__ __launch_bounds__(256) void k_compact_events(const gpet_event *in, const unsigned *flag, const unsigned *pos,
                                                        gpet_event *out, int64_t n, unsigned long long *key,
                                                        unsigned *val)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) {
        const gpet_event e = in[i];
        const unsigned j = pos[i];
        out[j] = e;
        if (key) {
            key[j] = sortable_t(e.t);
            val[j] = j;
        }
    }
}

/* Job of a history in a batched run: the largest j with jf[j] <= h. */
__device__ __forceinline__ int d_job_of(const int64_t *jf, int nj, int64_t h)
{
    if (nj <= 1) return 0;
    int lo = 0, hi = nj;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (jf[mid] <= h) lo = mid; else hi = mid;
    }
    return lo;
}

/* Job of the event at perm[j] as the key of a stable pass after the time
 * sort: the order becomes (job, t). */
__global__ __launch_bounds__(256) void k_job_keys(const gpet_event *ev, const unsigned *perm, unsigned long long *key,
                                                  int64_t n, const int64_t *jf, int nj)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) key[j] = (unsigned long long)d_job_of(jf, nj, ev[perm[j]].eventid);
}

/* First single and first pair of every job (lower bounds over the
 * (job, t)-ordered lists); entry nj is the list length. */
__global__ void k_job_bounds(const gpet_event *singles, int64_t ns, const gpet_event *pairs, int64_t np,
                             const int64_t *jf, int nj, unsigned long long *out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nj) return;
    int64_t lo = 0, hi = ns;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (d_job_of(jf, nj, singles[mid].eventid) < j) lo = mid + 1; else hi = mid;
    }
    out[j] = (unsigned long long)lo;
    lo = 0;
    hi = np;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (d_job_of(jf, nj, pairs[2 * mid].eventid) < j) lo = mid + 1; else hi = mid;
    }
    out[nj + 1 + j] = (unsigned long long)lo;
}

__global__ __launch_bounds__(256) void k_time_keys(const gpet_event *ev, unsigned long long *key, unsigned *val, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        key[i] = sortable_t(ev[i].t);
        val[i] = (unsigned)i;
    }
}

/* setSitenum (gPET_kernals.cu:607-640): rewrite siten at the deadtime
 * level; run whenever dlevel != 3 (dlevel 3 keeps the readout site, Q7). */
__global__ __launch_bounds__(256) void k_set_site(gpet_event *ev, int64_t n, int level, int moduleN, int crystalN)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gpet_event &e = ev[i];
    int s;
    if (level == 0) s = 0;
    else if (level == 1) s = e.pann;
    else if (level == 2) s = e.pann * moduleN + e.modn;
    else s = e.pann * moduleN * crystalN + e.modn * crystalN + e.cryn;
    e.siten = s;
}

__global__ __launch_bounds__(256) void k_site_keys(const gpet_event *ev, const unsigned *perm, unsigned long long *key,
                                                   unsigned *val, int64_t n, const int64_t *jf, int nj)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    unsigned idx = perm[j];
    key[j] = ((unsigned long long)d_job_of(jf, nj, ev[idx].eventid) << 32) | (unsigned long long)(unsigned)ev[idx].siten;
    val[j] = idx;
}

/* One lane per chain: a chain starts at j == 0, at a site change, or where
 * t_j > t_{j-1} + tau (from the unmodified times). */
__global__ __launch_bounds__(256) void k_deadtime(const gpet_event *ev, const unsigned *perm, const unsigned long long *site,
                                                  unsigned *alive, int64_t n, double tau, int type)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    unsigned long long sj = site[j];
    double tj = ev[perm[j]].t;
    bool start = (j == 0) || site[j - 1] != sj || tj > ev[perm[j - 1]].t + tau;
    if (!start) return;
    alive[perm[j]] = 1u;
    double tdead = tj;
    for (int64_t k = j + 1; k < n && site[k] == sj; ++k) {
        double tk = ev[perm[k]].t;
        if (tk < tdead + tau) {
            alive[perm[k]] = 0u;
            if (type == 0) tdead = tk;
        } else {
            if (tk > ev[perm[k - 1]].t + tau) break; /* next chain's own lane */
            alive[perm[k]] = 1u;
            tdead = tk;
        }
    }
}

__global__ __launch_bounds__(256) void k_window_flags(const gpet_event *ev, const unsigned *perm, const unsigned *alive,
                                                      unsigned *flag, int64_t n, float wmin, float wmax)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    unsigned idx = perm[j];
    float E = ev[idx].E;
    flag[j] = (alive ? alive[idx] : 1u) && !(E < wmin || E > wmax);
}

__global__ __launch_bounds__(256) void k_gather_singles(const gpet_event *ev, const unsigned *perm, const unsigned *flag,
                                                        const unsigned *pos, gpet_event *out, int64_t n)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (!flag) out[j] = ev[perm[j]];   /* every event, in permutation order */
    else if (flag[j]) out[pos[j]] = ev[perm[j]];
}

/* Coincidence windows inside one cluster, as the sequential rule: a window
 * opens at the first single not inside the previous window and holds the
 * singles with t - t_open <= W; exactly two -> a pair, more -> policy. */
template <bool WRITE>
__global__ __launch_bounds__(256) void k_coinc(const gpet_event *__restrict__ s, int64_t n, double W, int policy,
                                               unsigned *__restrict__ cnt, const unsigned *__restrict__ pos,
                                               unsigned *__restrict__ out, unsigned long long *ntrue,
                                               const int64_t *__restrict__ jf, int nj)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    /* singles are in (job, t) order; a job change closes every window */
    auto same_job = [&](int64_t a, int64_t b) {
        return nj <= 1 || d_job_of(jf, nj, s[a].eventid) == d_job_of(jf, nj, s[b].eventid);
    };
    const bool start = i < n && ((i == 0) || (s[i].t - s[i - 1].t > W) || !same_job(i, i - 1));
    unsigned np = 0, nt = 0;
    const unsigned o = (WRITE && start) ? pos[i] : 0u;
    int64_t a = start ? i : n;
    while (a < n) {
        int64_t b = a + 1;
        while (b < n && s[b].t - s[a].t <= W && same_job(b, a)) ++b;
        int64_t m = b - a;
        if (m == 2 || (m > 2 && policy == GPET_COINC_TAKE_ALL_GOODS)) {
            for (int64_t p = a; p < b; ++p)
                for (int64_t q = p + 1; q < b; ++q) {
                    if (WRITE) {
                        /* the pair as two indices; k_gather_pairs copies the records */
                        if (s[p].eventid == s[q].eventid && s[p].parn >= 0 && s[q].parn >= 0) nt++;
                        out[2 * (int64_t)(o + np)] = (unsigned)p;
                        out[2 * (int64_t)(o + np) + 1] = (unsigned)q;
                    }
                    np++;
                }
        }
        a = b;
        if (a >= n || s[a].t - s[a - 1].t > W || !same_job(a, a - 1)) break;
    }
    if (!WRITE) {
        if (i < n) cnt[i] = np;
    } else {
        /* true pairs: one atomic per block into one of CNT_STRIPES words
         * (atomics on one address serialize when many blocks end together) */
        __shared__ unsigned long long s_nt[4];
        unsigned long long v = nt;
        for (int k = 32; k > 0; k >>= 1) v += __shfl_xor(v, k);
        if (__lane_id() == 0) s_nt[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t = s_nt[0] + s_nt[1] + s_nt[2] + s_nt[3];
            if (t) atomicAdd(ntrue + (blockIdx.x % CNT_STRIPES), t);
        }
    }
}

__global__ __launch_bounds__(256) void k_tally(const gpet_event *s, int64_t n, int moduleN, int crystalN,
                                               unsigned long long *crystal, int ncrystal, unsigned long long *ehist,
                                               int nebins, float elo, float inv)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gpet_event &e = s[i];
    if (crystal) {
        int64_t g = (int64_t)e.pann * moduleN * crystalN + (int64_t)e.modn * crystalN + e.cryn;
        if (g >= 0 && g < ncrystal) atomicAdd(&crystal[g], 1ull);
    }
    if (ehist) {
        float x = (e.E - elo) * inv;
        if (x >= 0.0f && x < (float)nebins) atomicAdd(&ehist[(int)x], 1ull);
    }
}

/* The same tallies with block-private counts in LDS (crystal bins first,
 * then energy bins; 32-bit, a block sees fewer than 2^32 singles) added to
 * the global counts once per block: a few hundred energy bins no longer
 * take every single's atomic on one address. */
constexpr int TALLY_LDS_BINS = 16000;
__global__ __launch_bounds__(256) void k_tally_lds(const gpet_event *s, int64_t n, int moduleN, int crystalN,
                                                   unsigned long long *crystal, int ncrystal,
                                                   unsigned long long *ehist, int nebins, float elo, float inv)
{
    extern __shared__ unsigned tl_bins[];
    const int nc = crystal ? ncrystal : 0, ne = ehist ? nebins : 0;
    for (int b = (int)threadIdx.x; b < nc + ne; b += 256) tl_bins[b] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const gpet_event &e = s[i];
        if (nc) {
            const int64_t g = (int64_t)e.pann * moduleN * crystalN + (int64_t)e.modn * crystalN + e.cryn;
            if (g >= 0 && g < nc) atomicAdd(&tl_bins[g], 1u);
        }
        if (ne) {
            const float x = (e.E - elo) * inv;
            if (x >= 0.0f && x < (float)ne) atomicAdd(&tl_bins[nc + (int)x], 1u);
        }
    }
    __syncthreads();
    for (int b = (int)threadIdx.x; b < nc + ne; b += 256) {
        const unsigned v = tl_bins[b];
        if (v) atomicAdd(b < nc ? &crystal[b] : &ehist[b - nc], (unsigned long long)v);
    }
}

/* Sinogram of coincidence pairs: same rule as the oracle (crystal centres,
 * transaxial line as (phi, s), uniform bins). */
__device__ __forceinline__ void d_crystal_local(const DevPanel &P, int modn, int cryn, int moduleNy, int crystalNy,
                                                float &xl, float &yl, float &zl)
{
    int my = modn % moduleNy, mz = modn / moduleNy;
    int ly = cryn % crystalNy, lz = cryn / crystalNy;
    xl = (P.dirx * P.lenx) * 0.5f;
    yl = (-P.leny * 0.5f + (float)my * (P.mody + P.mspy)) + ((float)ly * (P.cry + P.cspy) + P.cry * 0.5f);
    zl = (-P.lenz * 0.5f + (float)mz * (P.modz + P.mspz)) + ((float)lz * (P.crz + P.cspz) + P.crz * 0.5f);
}

__device__ __forceinline__ void d_crystal_centre(const DevPanel &P, int modn, int cryn, int moduleNy, int crystalNy,
                                                 float &gx, float &gy)
{
    float xl, yl, zl;
    d_crystal_local(P, modn, cryn, moduleNy, crystalNy, xl, yl, zl);
    gx = P.ox + xl * P.ux[0] + yl * P.uy[0] + zl * P.uz[0];
    gy = P.oy + xl * P.ux[1] + yl * P.uy[1] + zl * P.uz[1];
}

__global__ __launch_bounds__(256) void k_sinogram(const gpet_event *pairs, int64_t npairs, const DevPanel *panels,
                                                  int npanels, int moduleNy, int crystalNy, unsigned long long *sino,
                                                  int n_s, int n_phi, float s_max)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const gpet_event &a = pairs[2 * i], &b = pairs[2 * i + 1];
    if (a.pann < 0 || a.pann >= npanels || b.pann < 0 || b.pann >= npanels) return;
    float x1, y1, x2, y2;
    d_crystal_centre(panels[a.pann], a.modn, a.cryn, moduleNy, crystalNy, x1, y1);
    d_crystal_centre(panels[b.pann], b.modn, b.cryn, moduleNy, crystalNy, x2, y2);
    float dx = x2 - x1, dy = y2 - y1;
    if (dx == 0.0f && dy == 0.0f) return;
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
    if (!(fs >= 0.0f && fs < (float)n_s)) return;
    atomicAdd(&sino[(int64_t)ip * n_s + (int)fs], 1ull);
}

/* ---- background singles (addnoise, gPET_kernals.cu:699-735) -------------
 * The reference's kernel (never launched) gives each thread a time interval,
 * draws exponential gaps of mean lambda with curand, an energy
 * Emean + sigma N(0,1), a uniform panel, module and crystal, placeholder
 * positions, parn -1, and appends in groups of six through an atomic.  Here
 * a lane owns one chunk of a job's window and draws from Philox streams keyed
 * by (job, chunk); a count pass and a write pass at scanned offsets give a
 * fixed output order.  Positions are the crystal centre in panel-local
 * coordinates; the site index is the one the readout gives that crystal. */
struct NoiseDev {
    uint64_t seed;
    double gap_us, chunk_us;
    float e_mean, e_sigma;
    const NoiseJob *jobs;
    int njobs;
    int64_t nchunks;
    const DevPanel *panels;
    int npanels, moduleN, crystalN, moduleNy, crystalNy, depth, policy;
};

__device__ __forceinline__ int d_readout_site(int depth, int policy, int pann, int modn, int cryn, int moduleN,
                                              int crystalN)
{
    if (depth == 3) return pann * moduleN * crystalN + modn * crystalN + cryn;
    if (policy == 1) depth = 2;
    if (depth == 0) return 0;
    if (depth == 1) return pann;
    return pann * moduleN + modn;
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_noise(NoiseDev p, unsigned *cnt, const unsigned *pos, gpet_event *out)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nchunks) return;
    int lo = 0, hi = p.njobs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (p.jobs[mid].chunk0 <= c) lo = mid; else hi = mid;
    }
    const NoiseJob J = p.jobs[lo];
    const int64_t lc = c - J.chunk0;
    const int64_t nch = (lo + 1 < p.njobs ? p.jobs[lo + 1].chunk0 : p.nchunks) - J.chunk0;
    double t = J.t0_us + (double)lc * p.chunk_us;
    const double end = (lc + 1 == nch) ? J.t1_us : J.t0_us + (double)(lc + 1) * p.chunk_us;
    gp_rng rt;
    gp_rng_init(&rt, p.seed, (uint64_t)lc, (uint32_t)GP_TAG_NOISE_T | ((uint32_t)lo << 4), 0u);
    const unsigned base = WRITE ? pos[c] : 0u;
    unsigned m = 0;
    while (m < (unsigned)NOISE_CHUNK_MAX) {
        const double u = gp_rng_uniform_d(&rt);
        t = t + (-gp_log_d(u)) * p.gap_us;
        if (!(t < end)) break;
        if (WRITE) {
            gp_rng ra;
            gp_rng_init(&ra, p.seed, (uint64_t)lc, (uint32_t)GP_TAG_NOISE_A | ((uint32_t)lo << 4), 8u * m);
            const gp_u4 b0 = gp_rng_block(&ra), b1 = gp_rng_block(&ra);
            gpet_event e;
            e.pann = (int)(((uint64_t)b0.z * (uint64_t)p.npanels) >> 32);
            e.modn = (int)(((uint64_t)b0.w * (uint64_t)p.moduleN) >> 32);
            e.cryn = (int)(((uint64_t)b1.x * (uint64_t)p.crystalN) >> 32);
            e.siten = d_readout_site(p.depth, p.policy, e.pann, e.modn, e.cryn, p.moduleN, p.crystalN);
            e.parn = -1 - (int)(base + m);
            e.eventid = J.eventid;
            e.t = t;
            e.E = fmaxf(p.e_mean + p.e_sigma * gp_normal(gp_u01(b0.x), gp_u01(b0.y)), 1.0f);
            d_crystal_local(p.panels[e.pann], e.modn, e.cryn, p.moduleNy, p.crystalNy, e.x, e.y, e.z);
            out[base + m] = e;
        }
        ++m;
    }
    if (!WRITE) cnt[c] = m;
}

/* ---- delayed-window randoms estimate ------------------------------------
 * The prompt windows of k_coinc are opened by the same singles; a window
 * opened at t_a also collects the singles of its job with
 * delay <= t - t_a <= delay + W.  With exactly one such single the opener and
 * it are a delayed pair; with more, the multiples policy decides as for a
 * prompt window holding the opener and those singles. */
template <bool WRITE>
__global__ __launch_bounds__(256) void k_delayed(const gpet_event *__restrict__ s, int64_t n, double W, double D,
                                                 int policy, unsigned *__restrict__ cnt,
                                                 const unsigned *__restrict__ pos, unsigned *__restrict__ out,
                                                 const int64_t *__restrict__ jf, int nj)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto same_job = [&](int64_t a, int64_t b) {
        return nj <= 1 || d_job_of(jf, nj, s[a].eventid) == d_job_of(jf, nj, s[b].eventid);
    };
    bool start = (i == 0) || (s[i].t - s[i - 1].t > W) || !same_job(i, i - 1);
    if (!start) {
        if (!WRITE) cnt[i] = 0;
        return;
    }
    const double DW = D + W;
    unsigned np = 0;
    const unsigned o = WRITE ? pos[i] : 0u;
    int64_t a = i;
    while (a < n) {
        /* first single of a's job at or after t_a + D: galloping, then bisection */
        auto before = [&](int64_t j) { return j < n && same_job(j, a) && s[j].t - s[a].t < D; };
        int64_t lo = a, hi = a + 1, step = 1;
        while (before(hi)) {
            lo = hi;
            step *= 2;
            hi = a + step < n ? a + step : n;
        }
        while (hi - lo > 1) {
            int64_t mid = lo + ((hi - lo) >> 1);
            if (before(mid)) lo = mid; else hi = mid;
        }
        int64_t e = hi;
        while (e < n && same_job(e, a) && s[e].t - s[a].t <= DW) ++e;
        const int64_t m = e - hi;
        if (m == 1 || (m > 1 && policy == GPET_COINC_TAKE_ALL_GOODS)) {
            for (int64_t p = -1; p < m; ++p)
                for (int64_t q = p + 1; q < m; ++q) {
                    if (WRITE) {
                        out[2 * (int64_t)(o + np)] = (unsigned)(p < 0 ? a : hi + p);
                        out[2 * (int64_t)(o + np) + 1] = (unsigned)(hi + q);
                    }
                    np++;
                }
        }
        /* next prompt window opener, as in k_coinc */
        int64_t b = a + 1;
        while (b < n && s[b].t - s[a].t <= W && same_job(b, a)) ++b;
        a = b;
        if (a >= n || s[a].t - s[a - 1].t > W || !same_job(a, a - 1)) break;
    }
    if (!WRITE) cnt[i] = np;
}

/* Records of `stride` floats keyed by the int in field `kf`: sort keys. */
__global__ __launch_bounds__(256) void k_rec_keys(const float *rec, int stride, int kf, int64_t n,
                                                  unsigned long long *key, unsigned *val)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        key[i] = (unsigned long long)__float_as_uint(rec[i * stride + kf]);
        val[i] = (unsigned)i;
    }
}

__global__ __launch_bounds__(256) void k_rec_gather(const float *rec, int stride, const unsigned *perm, float *out,
                                                    int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * stride) {
        const int64_t r = i / stride, f = i - r * stride;
        out[i] = rec[(int64_t)perm[r] * stride + f];
    }
}

/* Pair records from the sorters' index pairs: one thread per output record. */
__global__ __launch_bounds__(256) void k_gather_pairs(const gpet_event *__restrict__ s, const unsigned *__restrict__ idx,
                                                      gpet_event *__restrict__ out, int64_t m)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) out[k] = s[idx[k]];
}

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

} // namespace

struct DigiWork {
    gpet_event *ev_a = nullptr, *ev_b = nullptr;  /* compacted, blurred */
    int64_t ev_a_cap = 0, ev_b_cap = 0;
    unsigned *flag = nullptr, *pos = nullptr, *alive = nullptr;
    unsigned long long *key_a = nullptr, *key_b = nullptr;
    unsigned *val_a = nullptr, *val_b = nullptr;
    int64_t cap = 0;
    unsigned *tile_sums = nullptr, *hist = nullptr, *grand = nullptr;
    int64_t tile_cap = 0, hist_cap = 0;
    gpet_event *singles = nullptr, *coinc = nullptr, *delayed = nullptr;
    int64_t singles_cap = 0, coinc_cap = 0, delayed_cap = 0;
    unsigned long long *ntrue = nullptr;
    unsigned *nthr = nullptr;
    unsigned *pidx = nullptr;
    int64_t pidx_cap = 0;
    float *rec_tmp = nullptr;
    int64_t rec_tmp_cap = 0;
    unsigned *noise_cnt = nullptr, *noise_pos = nullptr;
    int64_t noise_cnt_cap = 0, noise_pos_cap = 0;
    NoiseJob *noise_jobs = nullptr;
    int64_t noise_jobs_cap = 0;
};

namespace {

template <class T>
int ensure(T *&p, int64_t &cap, int64_t need, char *err, int errlen)
{
    if (need <= cap && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    int64_t c = std::max<int64_t>(need + need / 4, 1024);
    HIPCHK(hipMalloc((void **)&p, sizeof(T) * (size_t)c));
    cap = c;
    return 0;
}

/* Exclusive scan of n uint32 into out; returns the total in *total (host). */
int scan(DigiWork *w, const unsigned *in, unsigned *out, int64_t n, unsigned *total, hipStream_t s, char *err, int errlen)
{
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 0) {
        *total = 0;
        return 0;
    }
    if (ensure(w->tile_sums, w->tile_cap, tiles, err, errlen)) return -1;
    if (!w->grand) HIPCHK(hipMalloc((void **)&w->grand, 64));
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, in, out, w->tile_sums, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, w->tile_sums, tiles, w->grand);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, out, w->tile_sums, n);
    HIPCHK(hipGetLastError());
    if (total) {
        HIPCHK(hipMemcpyAsync(total, w->grand, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return 0;
}

/* Stable LSD sort of (key, val) over the low `bits` bits; result ends in
 * key_a/val_a. */
int radix_sort(DigiWork *w, int64_t n, int bits, hipStream_t s, char *err, int errlen)
{
    if (n <= 1) return 0;
    int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
    int64_t hn = (int64_t)256 * nblocks;
    if (ensure(w->hist, w->hist_cap, hn, err, errlen)) return -1;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(RS_BLOCK), 0, s, w->key_a, n, shift, w->hist, nblocks);
        HIPCHK(hipGetLastError());
        if (scan(w, w->hist, w->hist, hn, nullptr, s, err, errlen)) return -1;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(RS_BLOCK), 0, s, w->key_a, w->val_a, w->key_b, w->val_b, n,
                           shift, w->hist, nblocks);
        HIPCHK(hipGetLastError());
        std::swap(w->key_a, w->key_b);
        std::swap(w->val_a, w->val_b);
    }
    return 0;
}

} // namespace

DigiWork *gpet_digi_work_create() { return new DigiWork(); }

void gpet_digi_work_destroy(DigiWork *w)
{
    if (!w) return;
    void *ptrs[] = {w->ev_a,  w->ev_b,  w->flag,      w->pos,  w->alive,   w->key_a,   w->key_b,
                    w->val_a, w->val_b, w->tile_sums, w->hist, w->grand,   w->singles, w->coinc,
                    w->ntrue, w->nthr, w->delayed, w->noise_cnt, w->noise_pos, w->noise_jobs, w->pidx, w->rec_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete w;
}

int gpet_digitize(DigiWork *w, const DigiParams &p, const gpet_event *events, int64_t n, DigiOut *out, hipStream_t s,
                  char *err, int errlen)
{
    const gpet_digitizer &g = p.d;
    *out = DigiOut();
    const int64_t *jf = p.job_first;
    const int nj = (p.job_first && p.njobs > 1) ? p.njobs : 0;
    const int job_bits = nj > 65536 ? 24 : (nj > 256 ? 16 : 8);
    int64_t c1 = w->cap, c2 = w->cap, c3 = w->cap, c4 = w->cap, c5 = w->cap, c6 = w->cap, c7 = w->cap, c8 = w->cap;
    int64_t need = std::max<int64_t>(n, 1);
    if (need > w->cap) {
        if (ensure(w->flag, c1, need, err, errlen) || ensure(w->pos, c2, need, err, errlen) ||
            ensure(w->alive, c3, need, err, errlen) || ensure(w->key_a, c4, need, err, errlen) ||
            ensure(w->key_b, c5, need, err, errlen) || ensure(w->val_a, c6, need, err, errlen) ||
            ensure(w->val_b, c7, need, err, errlen))
            return -1;
        w->cap = std::min({c1, c2, c3, c4, c5, c6, c7});
        (void)c8;
    }
    if (ensure(w->ev_b, w->ev_b_cap, need, err, errlen)) return -1;
    if (!w->ntrue) HIPCHK(hipMalloc((void **)&w->ntrue, sizeof(unsigned long long) * CNT_STRIPES));
    HIPCHK(hipMemsetAsync(w->ntrue, 0, sizeof(unsigned long long) * CNT_STRIPES, s));

    /* blur + thresholder, compaction (input order kept).  Without deadtime
     * the energy window is applied here too (early): the window keeps the
     * same events in any order, so sorting only the window's events by
     * time gives the same singles in the same order, and the compaction
     * writes the time keys as it goes. */
    const bool early = !(g.dead_time_us > 0.0f);
    unsigned nT = 0, nsel = 0;
    if (n > 0) {
        if (early) {
            if (!w->nthr) HIPCHK(hipMalloc((void **)&w->nthr, sizeof(unsigned)));
            HIPCHK(hipMemsetAsync(w->nthr, 0, sizeof(unsigned), s));
        }
        hipLaunchKernelGGL(k_blur_threshold, dim3(grid_for(n, 256)), dim3(256), 0, s, events, w->ev_b, w->flag, n, p.seed, g,
                           early ? w->nthr : (unsigned *)nullptr);
        HIPCHK(hipGetLastError());
        if (scan(w, w->flag, w->pos, n, &nsel, s, err, errlen)) return -1;
        /* the compacted events: the thresholder's (and, early, the window's)
         * passes only, often well under half of the input */
        if (ensure(w->ev_a, w->ev_a_cap, std::max<int64_t>(nsel, 1), err, errlen)) return -1;
        hipLaunchKernelGGL(k_compact_events, dim3(grid_for(n, 256)), dim3(256), 0, s, w->ev_b, w->flag, w->pos, w->ev_a, n,
                           early ? w->key_a : (unsigned long long *)nullptr, early ? w->val_a : (unsigned *)nullptr);
        HIPCHK(hipGetLastError());
        if (early)
            HIPCHK(hipMemcpy(&nT, w->nthr, sizeof(unsigned), hipMemcpyDeviceToHost));
        else
            nT = nsel;
    }
    out->n_threshold = nT;
    const gpet_event *ev = w->ev_a;
    const int64_t m = nsel;
    if (g.dead_level != 3 && m > 0) {
        hipLaunchKernelGGL(k_set_site, dim3(grid_for(m, 256)), dim3(256), 0, s, w->ev_a, m, g.dead_level, p.moduleN,
                           p.crystalN);
        HIPCHK(hipGetLastError());
    }

    /* time order: perm_t in val_a */
    if (m > 0) {
        if (!early) {
            hipLaunchKernelGGL(k_time_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, w->key_a, w->val_a, m);
            HIPCHK(hipGetLastError());
        }
        if (radix_sort(w, m, 64, s, err, errlen)) return -1;
        if (nj > 1) {
            hipLaunchKernelGGL(k_job_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, w->val_a, w->key_a, m, jf, nj);
            HIPCHK(hipGetLastError());
            if (radix_sort(w, m, job_bits, s, err, errlen)) return -1;
        }
    }
    /* val_a = perm_t.  Keep it in `pos`-independent storage: copy to val_b? we
     * sort sites on (key_b,val_b) after moving perm_t aside into `alive` ... */
    unsigned *perm_t = w->val_a;
    unsigned *alive = nullptr;
    int64_t n_dead = nT;
    if (g.dead_time_us > 0.0f && m > 0) {
        /* move perm_t out of the sort buffers */
        HIPCHK(hipMemcpyAsync(w->pos, perm_t, sizeof(unsigned) * m, hipMemcpyDeviceToDevice, s));
        perm_t = w->pos;
        hipLaunchKernelGGL(k_site_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, w->key_a, w->val_a, m,
                           jf, nj);
        HIPCHK(hipGetLastError());
        if (radix_sort(w, m, nj > 1 ? 32 + job_bits : 32, s, err, errlen)) return -1;
        HIPCHK(hipMemsetAsync(w->alive, 0, sizeof(unsigned) * m, s));
        hipLaunchKernelGGL(k_deadtime, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, w->val_a, w->key_a, w->alive, m,
                           (double)g.dead_time_us, g.dead_type);
        HIPCHK(hipGetLastError());
        alive = w->alive;
        /* perm_t lives in pos; window flags go to flag, scan into key_b (as uint) */
    }
    /* window flags in time order, compaction -> singles */
    unsigned ns = 0;
    if (m > 0 && early) {
        /* the window was applied before the sort: every event is a single */
        ns = (unsigned)m;
        if (ensure(w->singles, w->singles_cap, std::max<int64_t>(ns, 1), err, errlen)) return -1;
        hipLaunchKernelGGL(k_gather_singles, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, (const unsigned *)nullptr,
                           (const unsigned *)nullptr, w->singles, m);
        HIPCHK(hipGetLastError());
    } else if (m > 0) {
        if (alive) {
            unsigned long long *tmp = w->key_b; /* reuse as a uint buffer of >= m entries */
            unsigned *flag2 = reinterpret_cast<unsigned *>(tmp);
            unsigned *pos2 = flag2 + m;
            if (m * 2 > w->cap * 2) return -1;
            unsigned nalive = 0;
            if (scan(w, alive, flag2, m, &nalive, s, err, errlen)) return -1;
            n_dead = nalive;
            hipLaunchKernelGGL(k_window_flags, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, alive, w->flag, m,
                               g.win_min, g.win_max);
            HIPCHK(hipGetLastError());
            if (scan(w, w->flag, pos2, m, &ns, s, err, errlen)) return -1;
            if (ensure(w->singles, w->singles_cap, std::max<int64_t>(ns, 1), err, errlen)) return -1;
            hipLaunchKernelGGL(k_gather_singles, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, w->flag, pos2,
                               w->singles, m);
        } else {
            hipLaunchKernelGGL(k_window_flags, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, (const unsigned *)nullptr,
                               w->flag, m, g.win_min, g.win_max);
            HIPCHK(hipGetLastError());
            unsigned *pos2 = reinterpret_cast<unsigned *>(w->key_b);
            if (scan(w, w->flag, pos2, m, &ns, s, err, errlen)) return -1;
            if (ensure(w->singles, w->singles_cap, std::max<int64_t>(ns, 1), err, errlen)) return -1;
            hipLaunchKernelGGL(k_gather_singles, dim3(grid_for(m, 256)), dim3(256), 0, s, ev, perm_t, w->flag, pos2,
                               w->singles, m);
        }
        HIPCHK(hipGetLastError());
    }
    out->n_deadtime = n_dead;
    out->n_singles = ns;
    out->singles = w->singles;

    /* coincidences */
    unsigned npairs = 0;
    if (g.coinc_window_us > 0.0 && ns > 1) {
        unsigned *cnt = w->flag;
        unsigned *cpos = w->pos;
        hipLaunchKernelGGL(k_coinc<false>, dim3(grid_for(ns, 256)), dim3(256), 0, s, w->singles, (int64_t)ns,
                           g.coinc_window_us, g.coinc_policy, cnt, (const unsigned *)nullptr, (unsigned *)nullptr,
                           (unsigned long long *)nullptr, jf, nj);
        HIPCHK(hipGetLastError());
        if (scan(w, cnt, cpos, ns, &npairs, s, err, errlen)) return -1;
        const int64_t m = 2 * (int64_t)npairs;
        if (ensure(w->coinc, w->coinc_cap, std::max<int64_t>(m, 2), err, errlen) ||
            ensure(w->pidx, w->pidx_cap, std::max<int64_t>(m, 2), err, errlen))
            return -1;
        hipLaunchKernelGGL(k_coinc<true>, dim3(grid_for(ns, 256)), dim3(256), 0, s, w->singles, (int64_t)ns,
                           g.coinc_window_us, g.coinc_policy, cnt, cpos, w->pidx, w->ntrue, jf, nj);
        HIPCHK(hipGetLastError());
        if (m > 0) {
            hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(m, 256)), dim3(256), 0, s, w->singles, w->pidx, w->coinc, m);
            HIPCHK(hipGetLastError());
        }
    }
    /* delayed window: same openers, window shifted by delay_us */
    unsigned nd = 0;
    if (g.coinc_window_us > 0.0 && g.delay_us > 0.0 && ns > 1) {
        unsigned *cnt = w->flag;
        unsigned *cpos = w->pos;
        hipLaunchKernelGGL(k_delayed<false>, dim3(grid_for(ns, 256)), dim3(256), 0, s, w->singles, (int64_t)ns,
                           g.coinc_window_us, g.delay_us, g.coinc_policy, cnt, (const unsigned *)nullptr,
                           (unsigned *)nullptr, jf, nj);
        HIPCHK(hipGetLastError());
        if (scan(w, cnt, cpos, ns, &nd, s, err, errlen)) return -1;
        const int64_t m = 2 * (int64_t)nd;
        if (ensure(w->delayed, w->delayed_cap, std::max<int64_t>(m, 2), err, errlen) ||
            ensure(w->pidx, w->pidx_cap, std::max<int64_t>(m, 2), err, errlen))
            return -1;
        hipLaunchKernelGGL(k_delayed<true>, dim3(grid_for(ns, 256)), dim3(256), 0, s, w->singles, (int64_t)ns,
                           g.coinc_window_us, g.delay_us, g.coinc_policy, cnt, cpos, w->pidx, jf, nj);
        HIPCHK(hipGetLastError());
        if (m > 0) {
            hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(m, 256)), dim3(256), 0, s, w->singles, w->pidx,
                               w->delayed, m);
            HIPCHK(hipGetLastError());
        }
    }
    out->delayed = w->delayed;
    out->n_delayed = nd;
    if (nj > 1 && p.job_bounds) {
        hipLaunchKernelGGL(k_job_bounds, dim3(grid_for(nj + 1, 64)), dim3(64), 0, s, w->singles, (int64_t)ns, w->coinc,
                           (int64_t)npairs, jf, nj, p.job_bounds);
        HIPCHK(hipGetLastError());
    }
    unsigned long long nt = 0;
    unsigned long long ntv[CNT_STRIPES];
    HIPCHK(hipMemcpyAsync(ntv, w->ntrue, sizeof(ntv), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int k = 0; k < CNT_STRIPES; ++k) nt += ntv[k];
    out->n_coinc = npairs;
    out->n_coinc_true = (int64_t)nt;
    out->coinc = w->coinc;
    return 0;
}

void gpet_noise_layout(double rate, const double *t0_s, const double *t1_s, const int64_t *first_hist, int njobs,
                       std::vector<NoiseJob> &jobs, double &gap_us, double &chunk_us, int64_t &nchunks)
{
    jobs.assign((size_t)std::max(njobs, 0), NoiseJob());
    gap_us = 1.0e6 / rate;
    chunk_us = NOISE_PER_CHUNK * gap_us;
    nchunks = 0;
    for (int j = 0; j < njobs; ++j) {
        NoiseJob &J = jobs[j];
        J.t0_us = t0_s[j] * 1.0e6;
        J.t1_us = t1_s[j] * 1.0e6;
        J.chunk0 = nchunks;
        J.eventid = (int32_t)first_hist[j];
        J.pad = 0;
        const double T = J.t1_us - J.t0_us;
        if (T > 0.0) nchunks += (int64_t)std::ceil(T / chunk_us);
    }
}

namespace {

NoiseDev noise_dev(DigiWork *w, const NoiseParams &p)
{
    NoiseDev d;
    d.seed = p.seed;
    d.gap_us = p.gap_us;
    d.chunk_us = p.chunk_us;
    d.e_mean = p.e_mean;
    d.e_sigma = p.e_sigma;
    d.jobs = w->noise_jobs;
    d.njobs = p.njobs;
    d.nchunks = p.nchunks;
    d.panels = p.panels;
    d.npanels = p.npanels;
    d.moduleN = p.moduleN;
    d.crystalN = p.crystalN;
    d.moduleNy = p.moduleNy;
    d.crystalNy = p.crystalNy;
    d.depth = p.depth;
    d.policy = p.policy;
    return d;
}

} // namespace

int gpet_digi_noise_count(DigiWork *w, const NoiseParams &p, int64_t *total, hipStream_t s, char *err, int errlen)
{
    *total = 0;
    if (p.nchunks <= 0 || p.njobs <= 0) return 0;
    if (p.nchunks > ((int64_t)1 << 25)) {
        snprintf(err, errlen, "background singles: rate x window above 2^31 singles");
        return -1;
    }
    if (ensure(w->noise_jobs, w->noise_jobs_cap, p.njobs, err, errlen) ||
        ensure(w->noise_cnt, w->noise_cnt_cap, p.nchunks, err, errlen) ||
        ensure(w->noise_pos, w->noise_pos_cap, p.nchunks, err, errlen))
        return -1;
    HIPCHK(hipMemcpy(w->noise_jobs, p.jobs, sizeof(NoiseJob) * (size_t)p.njobs, hipMemcpyHostToDevice));
    const NoiseDev d = noise_dev(w, p);
    hipLaunchKernelGGL(k_noise<false>, dim3(grid_for(p.nchunks, 256)), dim3(256), 0, s, d, w->noise_cnt,
                       (const unsigned *)nullptr, (gpet_event *)nullptr);
    HIPCHK(hipGetLastError());
    unsigned n = 0;
    if (scan(w, w->noise_cnt, w->noise_pos, p.nchunks, &n, s, err, errlen)) return -1;
    *total = n;
    return 0;
}

int gpet_digi_noise_write(DigiWork *w, const NoiseParams &p, gpet_event *out, hipStream_t s, char *err, int errlen)
{
    if (p.nchunks <= 0 || p.njobs <= 0) return 0;
    const NoiseDev d = noise_dev(w, p);
    hipLaunchKernelGGL(k_noise<true>, dim3(grid_for(p.nchunks, 256)), dim3(256), 0, s, d, w->noise_cnt,
                       (const unsigned *)w->noise_pos, out);
    HIPCHK(hipGetLastError());
    return 0;
}

int gpet_sort_records(DigiWork *w, float *rec, int stride, int key_field, int64_t n, hipStream_t s, char *err,
                      int errlen)
{
    if (n <= 1) return 0;
    if (n > w->cap) {
        int64_t c4 = w->cap, c5 = w->cap, c6 = w->cap, c7 = w->cap, c1 = w->cap, c2 = w->cap, c3 = w->cap;
        if (ensure(w->flag, c1, n, err, errlen) || ensure(w->pos, c2, n, err, errlen) ||
            ensure(w->alive, c3, n, err, errlen) || ensure(w->key_a, c4, n, err, errlen) ||
            ensure(w->key_b, c5, n, err, errlen) || ensure(w->val_a, c6, n, err, errlen) ||
            ensure(w->val_b, c7, n, err, errlen))
            return -1;
        w->cap = std::min({c1, c2, c3, c4, c5, c6, c7});
    }
    if (ensure(w->rec_tmp, w->rec_tmp_cap, n * stride, err, errlen)) return -1;
    hipLaunchKernelGGL(k_rec_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, rec, stride, key_field, n, w->key_a,
                       w->val_a);
    HIPCHK(hipGetLastError());
    if (radix_sort(w, n, 32, s, err, errlen)) return -1;
    hipLaunchKernelGGL(k_rec_gather, dim3(grid_for(n * stride, 256)), dim3(256), 0, s, rec, stride, w->val_a,
                       w->rec_tmp, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rec, w->rec_tmp, sizeof(float) * (size_t)(n * stride), hipMemcpyDeviceToDevice, s));
    return 0;
}

int gpet_digi_sinogram(const gpet_event *pairs, int64_t npairs, const DevPanel *panels, int npanels, int moduleNy,
                       int crystalNy, unsigned long long *sino, int n_s, int n_phi, float s_max, hipStream_t s)
{
    if (npairs <= 0) return 0;
    hipLaunchKernelGGL(k_sinogram, dim3(grid_for(npairs, 256)), dim3(256), 0, s, pairs, npairs, panels, npanels,
                       moduleNy, crystalNy, sino, n_s, n_phi, s_max);
    return (int)hipGetLastError();
}

int gpet_digi_tallies(DigiWork *w, const gpet_event *singles, int64_t n, int moduleN, int crystalN,
                      unsigned long long *crystal, int ncrystal, unsigned long long *ehist, int nebins, float elo,
                      float ehi, hipStream_t s)
{
    (void)w;
    if (n <= 0) return 0;
    float inv = (nebins > 0 && ehi > elo) ? (float)nebins / (ehi - elo) : 0.0f;
    const int64_t bins = (crystal ? (int64_t)ncrystal : 0) + (ehist ? (int64_t)nebins : 0);
    if (bins > 0 && bins <= TALLY_LDS_BINS) {
        const int64_t blocks = std::min<int64_t>(grid_for(n, 256), 2048);
        hipLaunchKernelGGL(k_tally_lds, dim3((unsigned)blocks), dim3(256), (size_t)bins * sizeof(unsigned), s, singles,
                           n, moduleN, crystalN, crystal, ncrystal, ehist, nebins, elo, inv);
    } else {
        hipLaunchKernelGGL(k_tally, dim3(grid_for(n, 256)), dim3(256), 0, s, singles, n, moduleN, crystalN, crystal,
                           ncrystal, ehist, nebins, elo, inv);
    }
    return (int)hipGetLastError();
}
