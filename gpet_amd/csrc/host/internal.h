/*
 * internal.h - device data layouts and kernel launchers of the HIP path
 * (not part of the public C ABI).
 *
 * HBM layout (MI355X, 288 GB): everything a run touches stays resident.
 *   phantom    int2 per voxel {density bits, material}, x fastest: one 8-byte
 *              gather per Woodcock step instead of two texture fetches
 *              (initialize.cu:831-884 uses two 3-D textures).
 *   tables     float4 per (material, energy node) {mu_tot, mu_compt, mu_rayl, 0}
 *              so one lerp reads two 16-byte nodes; inverse-CDF surfaces
 *              [mat][icp][ie]; majorants 4096 floats each.
 *   queue      SoA arrays of photons that entered a panel (regrouped by the
 *              phantom kernel so the panel kernel runs dense waves).
 *   events     gpet_event (48 B) appended with one atomic per wave.
 */
#ifndef GPET_INTERNAL_H
#define GPET_INTERNAL_H

#include <stdint.h>
#include <vector>
#include "../../../include/gpet.h"

#define GPET_MAX_PH_EVENTS 4   /* per-photon adder list (reference: events[4]) */

struct DevWck {
    float e0, idl;
    const float *lam;          /* [GPET_NWCK] */
};

struct DevTables {
    int nmat, n1d, ncp, ne;
    float e0, idl;             /* common grid of lamph/compt/rayle            */
    const float4 *mu4;         /* [nmat][n1d] {tot, compt, rayl, 0}           */
    float idcp[2], ide[2];
    const float *surf[2];      /* [nmat][ncp][ne]                             */
};

/* density outside the phantom for the positron walk (gPET_kernals.cu:403) */
#define D_RHO_OUTSIDE 0.0012905f

struct DevPhantom {
    int nx, ny, nz;
    float offx, offy, offz, dx, dy, dz, idx, idy, idz;
    const int2 *vox;           /* {float bits of density, material}           */
    const float *dens_b;       /* padded density in 4x4x4 bricks (positron
                                  walk); the border layer holds -D_RHO_OUTSIDE */
    int nbx, nby;              /* bricks along x and y (padded grid)          */
    unsigned dens_bytes;       /* size of dens_b (fewer than 2^24 bricks)     */
    const uint4 *wmap;         /* uniform box per padded brick (gp_walkmap.h),
                                  or null when the phantom has no uniform brick */
    unsigned wmap_bytes;
    DevWck wck;
    int present;
};

/* Offset (in floats) of voxel (ix, iy, iz) of the padded grid in the bricked
 * density array: 64 voxels per brick, x fastest inside a brick, bricks x
 * fastest.  A 128-byte line holds a 4 x 4 x 2 block, so the short random
 * walks of positrons reuse lines.  The padded grid has one extra layer on
 * every side (phantom voxel (i, j, k) is padded voxel (i+1, j+1, k+1)); the
 * layer holds -D_RHO_OUTSIDE, so a walk learns that it left the phantom from
 * the sign of the density it reads. */
__host__ __device__ __forceinline__ size_t gpet_brick_offset(int ix, int iy, int iz, int nbx, int nby)
{
    size_t b = ((size_t)(iz >> 2) * nby + (size_t)(iy >> 2)) * nbx + (size_t)(ix >> 2);
    return (b << 6) | (size_t)(((iz & 3) << 4) | ((iy & 3) << 2) | (ix & 3));
}

/* Bricks of the padded grid along an axis of n voxels. */
__host__ __device__ __forceinline__ int gpet_padded_bricks(int n) { return (n + 2 + 3) / 4; }

struct DevPanel {
    int panel;
    float lenx, leny, lenz, mody, modz, mspy, mspz, cry, crz, cspy, cspz;
    float ox, oy, oz, dirx;
    float ux[3], uy[3], uz[3];
    float bound2;              /* squared radius (with margin) of a sphere about
                                  the face centre that contains the face      */
};

struct DevDetector {
    int npanels;
    const DevPanel *p;
    int mat[2];
    float dens[2];
    int nquad;
    float quad[10 * GPET_MAXSURFACE];
    int moduleN, crystalN, moduleNy, crystalNy;
    DevWck wck;
    int present;
    /* panel groups for the entry test (16 < npanels <= 64, else 0): a
     * sphere enclosing the members' entry spheres, and the members as bits */
    int ngroups;
    const float4 *gsph;
    const unsigned long long *gmask;
};

/* LDS bytes of the panel tables the transport kernels stage: entry spheres,
 * group spheres and masks, panels. */
__host__ __device__ __forceinline__ size_t gpet_panel_lds_bytes(const DevDetector &d)
{
    if (!d.present) return 0;
    return (sizeof(float4) + sizeof(DevPanel)) * (size_t)d.npanels +
           (sizeof(float4) + sizeof(unsigned long long)) * (size_t)d.ngroups;
}

struct DevSource {
    double tau, pdec;          /* T1/2 / ln 2 (s); 1 - 2^(-(t1-t0)/T1/2)      */
    double t0;                 /* start of the source's acquisition (s)       */
    int shape, iso;
    int phantom;               /* phantom slot (batched runs)                 */
    float coef[6];
};

struct DevIsotope {
    float ratio;
    float coef[8];
};

/* A history range of one phantom slot inside a batched launch; vb0 is its
 * first virtual block (blocks of the launch's block size). */
struct DevSegment {
    int64_t h0, nh, vb0;
    int phantom, pad;
};

/* What the phantom kernel produces for the panel kernel.  The panel queue
 * is split into one region per phantom-kernel block (no global atomics):
 * region r holds slots [r * region, r * region + rcount[r]).  The run-wide
 * surrogate queue uses the single counter `count` instead (region = 0). */
/* A positron-range walk handed from the phantom kernel to k_longwalk: the
 * walk state after its first crossing groups and what the history needs after it. */
struct LongWalkItem {
    float ux, uy, uz, rr, tmx, tmy, tmz, tdx, tdy, tdz, tcur, s;
    unsigned o;                /* byte offset of the current voxel in dens_b  */
    int it;
    int i;                     /* launch index of the history                 */
    float px, py, pz;
    unsigned n;                /* aligned decay-stream position of the pair   */
    double tstart;
    int slot, pad;             /* phantom slot (batched runs)                 */
};

/* A photon still in the phantom after the phantom kernel's first Woodcock
 * steps, handed to the transit kernel (72 B; a traced photon's emission
 * fields go to its trace record at the hand-off). */
struct TransitItem {
    double t;                  /* time now (us)                               */
    int64_t h;                 /* history                                     */
    float x, y, z, vx, vy, vz, E;
    int ns, nc, nr, it;        /* steps inside, Compton, Rayleigh; all steps  */
    unsigned n;                /* stream position, photon index in bit 31     */
    int parn, slot;            /* photon id; phantom slot (batched runs)      */
};

struct DevQueue {
    float *y, *z, *vx, *vy, *vz, *E;
    double *t;
    int4 *info;                /* {panel index, parn, eventid, draws}         */
    int *trace_slot;           /* trace slot or -1                            */
    int *count;
    int64_t cap;
    unsigned *rcount;          /* [nregions] entries per region               */
    int64_t region;            /* slots per region                            */
    int nregions;
    /* long positron walks: region r holds items [r * region / 8, ...) and
     * lw_count[r] of them are used; null: walks are not deferred */
    LongWalkItem *lw;
    unsigned *lw_count;
    /* photons in transit: region r holds items [r * region, ...) and
     * tq_count[r] of them are used; tq_steps Woodcock steps are taken in the
     * phantom kernel before a photon is put here (0: never) */
    TransitItem *tq;
    unsigned *tq_count;
    int tq_steps;
    TransitItem *tq2;          /* second buffer: rounds alternate between them */
    unsigned *tq2_count;
};

/* Device-side counters, 64-bit, accumulated with atomics. */
enum {
    CNT_PHOTONS = 0, CNT_PANEL_ENTRIES, CNT_HITS, CNT_DROPPED_EVENTS, CNT_DROPPED_HITS,
    CNT_MAJORANT, CNT_PHANTOM_STEPS, CNT_RANGE_STEPS, CNT_PANEL_STEPS, CNT_RANGE_JUMPS, CNT_NUM
};
/* Each counter is striped over CNT_STRIPES words (counter k of block b goes
 * to cnt[k * CNT_STRIPES + b % CNT_STRIPES]; the host sums the stripes):
 * atomics on one address serialize, and thousands of blocks finishing
 * together would queue on it. */
#define CNT_STRIPES 64

enum { RUN_DECAY = 0, RUN_PSF_POSITRON = 1, RUN_PSF_PHOTON = 2 };

struct RunParams {
    int mode;
    int64_t h0, nh;            /* histories [h0, h0+nh)                        */
    uint64_t seed;
    float eabs, sigma;
    int prange;
    int nsrc;
    const int64_t *first;      /* [nsrc+1] first history per source            */
    const DevSource *src;
    const DevIsotope *iso;
    const DevPhantom *phs;     /* batched runs: phantom slots                 */
    const struct DevSegment *seg; /* batched runs: segments of this launch    */
    int nseg;
    int64_t n_vblocks;         /* batched runs: virtual blocks of this launch */
    const double *psf;         /* mode 1/2: records of 8 doubles for [h0,...)  */
    gpet_trace *trace;         /* device trace buffer or null                  */
    int64_t trace_cap;         /* slots: 2*(h - 0) < trace_cap are written     */
    unsigned long long *cnt;   /* [CNT_NUM * CNT_STRIPES]                      */
    float sphere[4];           /* recording sphere (r = sphere[3] > 0: on)     */
    gpet_psf_record *psf_src, *psf_ph; /* phase-space outputs or null         */
    unsigned long long *n_psf;  /* [2] appended records                         */
    int64_t psf_cap;
};

struct PanelParams {
    uint64_t seed;
    float eabs;
    int readout_depth, readout_policy, record_rayleigh;
    gpet_event *events;
    unsigned int *n_events;
    int64_t events_cap;
    int32_t *hit_id;
    float *hit_f;
    unsigned int *n_hits;
    int64_t hits_cap;
    gpet_trace *trace;
    unsigned long long *cnt;
    /* surrogate: queue entries inside the validity range are copied here
     * (run-wide) and left to the network instead of being transported */
    DevQueue sq;
    /* harvest: one record (features, targets) per in-range queue item */
    float *harvest;
    unsigned long long *harvest_count;
    int64_t harvest_cap;
};

/* Panel kernel modes. */
enum { PANEL_PLAIN = 0, PANEL_HARVEST = 1, PANEL_SURROGATE = 2 };

/* Launch shape of the two transport kernels: block size, a register
 * budget variant (minimum waves per SIMD, 0 = compiler default) and the
 * number of resident blocks per CU the grid-stride grid is sized for. */
struct LaunchCfg {
    int block = 256;
    int wpe = 0;
    int blocks_per_cu = 0;     /* 0: one lane per item (no grid-stride reuse) */
    int rounds[3] = {0, 0, 0}; /* phantom stage: Woodcock steps per transit
                                  round (0: not launched; the last launched
                                  round finishes every photon)              */
};
#define GPET_N_WPE 3
extern const int gpet_wpe_variants[GPET_N_WPE];

/* Grid of a phantom-kernel launch over `items` histories (virtual blocks x
 * block for batched runs): the queue has one region per block. */
int gpet_phantom_grid(int64_t items, const LaunchCfg &cfg, int n_cu);

/* Launchers (transport.hip).  batch: sources carry phantom slots (rp.phs).
 * The panel kernel reads its item count from q.count on the device, so the
 * pair can be captured in a graph without a host round trip. */
int gpet_launch_phantom(const RunParams &rp, const DevTables &tab, const DevPhantom &ph,
                        const DevDetector &det, const DevQueue &q, bool batch, const LaunchCfg &cfg,
                        int n_cu, hipStream_t s);
int gpet_launch_panel(const PanelParams &pp, const DevTables &tab, const DevDetector &det,
                      const DevQueue &q, int64_t max_items, const LaunchCfg &cfg, int n_cu, hipStream_t s,
                      int mode = PANEL_PLAIN);

/* Surrogate kernels (mlp.hip). */
struct GemmArgs {
    const uint16_t *A, *B;
    int64_t lda, ldb;
    int64_t M, N, K;
    int64_t k_chunk;
    float *C;
    int64_t ldc;
    uint16_t *Cb;
    int64_t ldcb;
    const float *Z;
    int64_t ldz;
    uint16_t *CbT;      /* EPI_SILU / EPI_DSILU: bf16 output transposed, [N][ldt] */
    int64_t ldt;
    float *CT;          /* EPI_F32: output transposed, [N][ldt] */
};
/* EPI_F32: C fp32; EPI_SILU: Z = C fp32 and silu(C) bf16; EPI_ACT: silu(C)
 * bf16 only; EPI_DSILU: C * silu'(Z) bf16; EPI_PARTIAL: split-K partial. */
enum { EPI_F32 = 0, EPI_SILU = 1, EPI_DSILU = 2, EPI_PARTIAL = 3, EPI_ACT = 4 };
/* C[m][n] = sum_k A[m][k] B[n][k] (k contiguous in both; lda, ldb and
 * k_chunk multiples of 8 and 64); splits > 1: k in chunks of k_chunk, one
 * partial per chunk (EPI_PARTIAL). */
int gpet_gemm(int kind, const GemmArgs &g, int splits, hipStream_t s);
/* wbT = the bf16 weights with every layer transposed ([kp][LN] instead of
 * [LN][kp], same offsets). */
struct LayerShapes {
    int64_t off[6];     /* parameter offset of layer l; off[5] = total */
    int ln[5], kp[5];   /* stored outputs, stored inputs (row length) */
};
int gpet_transpose_params(const uint16_t *wb, uint16_t *wbT, const LayerShapes &L, hipStream_t s);
/* rows [r0, r1) of a [rows][ld] bf16 matrix: row r0 = 1, the rest 0 (the
 * transposed activations' bias row and padding) */
int gpet_bias_rows(uint16_t *a, int64_t ld, int r0, int r1, hipStream_t s);
int gpet_sum_partials(const float *part, int nz, int64_t n, float *out, hipStream_t s);
int gpet_bias_columns(uint16_t *a, int64_t M, int64_t ld, int c0, hipStream_t s);
/* HT: head outputs transposed ([output][ldt]); Y: targets with the energy
 * class in place of the deposited fraction; dH row-major (lddh = 192),
 * dH^T [output][ldt]; one loss partial per 256 rows in part. */
int gpet_head_loss_grad(const float *HT, int64_t ldt, const float *Y, int64_t M, float inv_n, uint16_t *dH,
                       int64_t lddh, uint16_t *dHT, float *part, hipStream_t s);
int gpet_adam(float *w, float *m, float *v, const float *g, const float *gscale, uint16_t *wb, int64_t n, float lr,
              float b1, float b2, float c1, float c2, float eps, hipStream_t s);
/* factor[0] = min(1, clip / |g|) (240 partials in part). */
int gpet_clip_factor(const float *g, int64_t n, float clip, float *part, float *factor, hipStream_t s);
/* refl: bit 0 / bit 1 mirror half the rows in local y / z (features,
 * direction and lateral offset negated). */
int gpet_pack_train(const float *set, int64_t n, uint64_t seed, int64_t step, int64_t M, uint16_t *X, float *Y,
                    uint16_t *XT, int64_t ldt, int refl, hipStream_t s);
int gpet_loss_reduce(const float *part, int n, float *out, hipStream_t s);
/* The whole network in one kernel (layer l's bf16 weights w[l]): with X
 * (packed bf16 [n x 8] inputs) the SURR_NH head outputs of each row go to H
 * (row stride ldh); without X the inputs are surrogate-queue rows
 * [row0, row0 + n) of sq and each row's event is sampled and appended
 * (pp.events / pp.n_events). */
int gpet_mlp_infer(const uint16_t *const w[5], const uint16_t *X, const DevQueue &sq, int64_t row0, int64_t n,
                   const DevDetector &det, const PanelParams &pp, float *H, int64_t ldh, hipStream_t s);
/* Sets the first `n` entries of `p` to zero (graph-safe: a kernel, not a
 * memset node). */
int gpet_launch_zero(unsigned *p, int n, hipStream_t s);

/* Digitizer (digitizer.hip).  Works on the device event list; returns
 * counts through the pointers. */
struct DigiParams {
    uint64_t seed;
    gpet_digitizer d;
    int moduleN, crystalN;
    /* batched runs: job j owns histories [job_first[j], job_first[j+1]);
     * deadtime and coincidences stay inside a job, singles come out in
     * (job, t) order and job_bounds receives 2 x (njobs + 1) prefix offsets
     * {singles..., pairs...}.  njobs 0 or 1: one acquisition. */
    const int64_t *job_first;
    int njobs;
    unsigned long long *job_bounds;
};
struct DigiOut {
    gpet_event *singles;       /* device, owned by the caller of digitize     */
    int64_t n_singles;
    gpet_event *coinc;         /* device, 2 per pair                          */
    int64_t n_coinc, n_coinc_true;
    int64_t n_threshold, n_deadtime;
    gpet_event *delayed;       /* device, 2 per delayed-window pair           */
    int64_t n_delayed;
};

/* Background singles.  Job j's window [t0_us, t1_us) is cut into chunks of
 * chunk_us (the last one ends at t1_us); chunk c of job j draws exponential
 * gaps of mean gap_us from its own Philox stream and keeps the arrivals
 * before its end, at most NOISE_CHUNK_MAX.  Singles come out in (job, chunk,
 * arrival) order; the k-th has parn = -1 - k and eventid = the job's first
 * history (so job lookups by eventid place it in its job). */
#define NOISE_PER_CHUNK 64.0
#define NOISE_CHUNK_MAX 4096
struct NoiseJob {
    double t0_us, t1_us;
    int64_t chunk0;            /* first global chunk of the job               */
    int32_t eventid, pad;
};
struct NoiseParams {
    uint64_t seed;
    double gap_us, chunk_us;
    float e_mean, e_sigma;
    const NoiseJob *jobs;      /* host array, njobs entries                   */
    int njobs;
    int64_t nchunks;
    const DevPanel *panels;    /* device                                      */
    int npanels, moduleN, crystalN, moduleNy, crystalNy, depth, policy;
};
/* Job windows and chunk layout for a rate (singles/s) over windows in s. */
void gpet_noise_layout(double rate, const double *t0_s, const double *t1_s, const int64_t *first_hist, int njobs,
                       std::vector<NoiseJob> &jobs, double &gap_us, double &chunk_us, int64_t &nchunks);
struct DigiWork;               /* scratch owned by the context                */
DigiWork *gpet_digi_work_create();
void gpet_digi_work_destroy(DigiWork *w);
int gpet_digitize(DigiWork *w, const DigiParams &p, const gpet_event *events, int64_t n,
                  DigiOut *out, hipStream_t s, char *err, int errlen);
/* Stable sort of n records of `stride` floats by the non-negative int in
 * field key_field (ascending). */
int gpet_sort_records(DigiWork *w, float *rec, int stride, int key_field, int64_t n, hipStream_t s, char *err,
                      int errlen);
/* Count (into *total, host) then write the background singles. */
int gpet_digi_noise_count(DigiWork *w, const NoiseParams &p, int64_t *total, hipStream_t s, char *err, int errlen);
int gpet_digi_noise_write(DigiWork *w, const NoiseParams &p, gpet_event *out, hipStream_t s, char *err, int errlen);
int gpet_digi_sinogram(const gpet_event *pairs, int64_t npairs, const DevPanel *panels, int npanels, int moduleNy,
                       int crystalNy, unsigned long long *sino, int n_s, int n_phi, float s_max, hipStream_t s);
int gpet_digi_tallies(DigiWork *w, const gpet_event *singles, int64_t n, int moduleN, int crystalN,
                      unsigned long long *crystal, int ncrystal, unsigned long long *ehist, int nebins,
                      float elo, float ehi, hipStream_t s);

#endif
