/*
 * transport.hip - photon transport kernels for gfx950 (MI355X).
 *
 * Two stages per batch of histories, regrouped through a queue:
 *
 *   k_phantom  one lane per history: decay sampling (setPosition,
 *              gPET_kernals.cu:483-561) or PSF record (setPositionForPhoton,
 *              :563-604), positron range (:347-418), the annihilation pair,
 *              Woodcock tracking in the voxel phantom (photon, :256-345) and
 *              the panel entry test (photonde, :958-1009).  Only photons that
 *              enter a panel (about a tenth) are appended to the block's
 *              queue region (LDS counter, one atomic per wavefront).
 *              Positron walks longer than a few crossing groups go on in
 *              k_longwalk; optionally, photons still in the phantom after a
 *              step budget go on in k_transit rounds.
 *   k_panel    one lane per queued photon: Woodcock tracking in the crystal
 *              panel with analytic Klein-Nishina, Rayleigh and photoelectric
 *              (:1011-1193), hit records, adder and readout (:737-813), and
 *              the event append (staged in LDS, one reservation per block).
 *
 * Splitting at the panel face keeps the panel kernel's waves full: in the
 * reference one kernel per stage runs every photon slot, and ~90% of the
 * lanes of photonde idle on photons that miss the detector.
 *
 * Random numbers: every photon owns Philox stream (history, 1 + photon), the
 * decay owns (history, 0); the queue carries the draw counter, so regrouping
 * does not change any history's numbers.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include "dev_physics.h"
#include "surr_common.h"

namespace {

__device__ __forceinline__ unsigned long long lanemask_lt()
{
    unsigned int l = __lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

/* Wave-aggregated append to a 64-bit or unsigned counter, every calling
 * lane appending one item (one atomic per wave instead of one per lane). */
template <class T>
__device__ __forceinline__ T wave_append1(T *counter)
{
    const unsigned long long m = __ballot(1);
    const int leader = __ffsll((unsigned long long)m) - 1;
    T base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (T)__popcll(m));
    base = __shfl(base, leader);
    return base + (T)__popcll(m & lanemask_lt());
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* Block-wide sums of N per-lane counters, one atomic per counter and
 * block (called by every thread of the block, after its work loop). */
template <int N>
__device__ __forceinline__ void flush_counters(unsigned long long *const dst[N], const unsigned long long v[N])
{
    __shared__ unsigned long long s_red[4][N];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const unsigned long long s = wave_sum(v[k]);
        if (__lane_id() == 0) s_red[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        unsigned long long s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += s_red[i][threadIdx.x];
        if (s) atomicAdd(dst[threadIdx.x] + (blockIdx.x % CNT_STRIPES), s);
    }
}

/* ---- source ------------------------------------------------------------ */

/* Positron kinetic energy by rejection from the isotope's spectrum
 * polynomial: candidate i is (E, u) from draws 2i, 2i+1, the first accepted
 * one (at most 100000 candidates) is kept.  Called at a block boundary and
 * followed by a re-alignment, so one generator block yields two candidates
 * in draw order and a wave's lanes, which accept after different numbers of
 * tries, run half as many loop trips; the result and the stream position
 * after the alignment are those of one candidate per trip. */
__device__ float d_sample_ek(const DevIsotope &iso, gp_rng &r)
{
    float E = 0.511f;
    for (int it = 0; it < 100000; it += 2) {
        const gp_u4 b = gp_rng_block(&r);
        float E1 = gp_u01(b.x) * (iso.coef[0] - 0.511f) + 0.511f;
        float u1 = iso.coef[1] * gp_u01(b.y);
        float p1 = iso.coef[2];
        for (int i = 3; i < 8; ++i) p1 = p1 * E1 + iso.coef[i];
        E = E1;
        if (u1 <= p1) break;
        float E2 = gp_u01(b.z) * (iso.coef[0] - 0.511f) + 0.511f;
        float u2 = iso.coef[1] * gp_u01(b.w);
        float p2 = iso.coef[2];
        for (int i = 3; i < 8; ++i) p2 = p2 * E2 + iso.coef[i];
        E = E2;
        if (u2 <= p2) break;
    }
    return (E - 0.511f) * 1.0e6f;
}

__device__ void d_sample_shape(const DevSource &s, gp_rng &r, float &x, float &y, float &z)
{
    const float *c = s.coef;
    int shape = (s.shape < 0 || s.shape > 3) ? 0 : s.shape;
    if (shape == 0) {
        x = c[0] + (c[3] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
        y = c[1] + (c[4] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
        z = c[2] + (c[5] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
    } else if (shape == 1) {
        float sp, cp;
        gp_sincos2pi(gp_rng_uniform(&r), &sp, &cp);
        float rr = c[3] * sqrtf(gp_rng_uniform(&r));
        x = c[0] + rr * cp;
        y = c[1] + rr * sp;
        z = c[2] + (c[4] * (2.0f * gp_rng_uniform(&r) - 1.0f)) * 0.5f;
    } else {
        float sp, cp;
        gp_sincos2pi(gp_rng_uniform(&r), &sp, &cp);
        float ct = 2.0f * gp_rng_uniform(&r) - 1.0f;
        float rr = gp_cbrtf(gp_rng_uniform(&r));
        float st = sqrtf(1.0f - ct * ct);
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

/* Positron range: 3-D DDA voxel walk (same rule as the oracle, see
 * setPositronRange, gPET_kernals.cu:347-418); crossings gather only the
 * density half of the voxel record. */
__device__ __forceinline__ void d_dda_axis(float x, float off, float d, int i, float u, float inv, float &tmax,
                                           float &tdelta, int &step)
{
    /* inv = 1 / u (correctly rounded), computed once by the caller */
    if (u > 0.0f) {
        tmax = (off + (float)(i + 1) * d - x) * inv;
        tdelta = d * inv;
        step = 1;
    } else if (u < 0.0f) {
        tmax = (off + (float)i * d - x) * inv;
        tdelta = -(d * inv);
        step = -1;
    } else {
        tmax = INFINITY;
        tdelta = INFINITY;
        step = 0;
    }
}

/* The walk as a state: begin() samples the displacement and sets up the DDA
 * (or finishes at once outside the phantom), step4() advances four
 * crossings, end() moves the point. */
/* crossings per walk step (gathers in flight per lane) */
#ifndef WALK_K
#define WALK_K 4
#endif

struct RangeWalk {
    float ux, uy, uz, rr;
    float tmx, tmy, tmz, tdx, tdy, tdz;
    /* o: byte offset of the current voxel in the padded bricked densities;
     * d?i / d?c: change of o for a crossing along an axis that stays inside
     * the brick / enters the next brick; flip: the local-coordinate bits of
     * the axes walked in the + direction (a crossing enters the next brick
     * when the axis' local coordinate is 3 going up or 0 going down). */
    unsigned o, flip;
    int dxi, dyi, dzi, dxc, dyc, dzc;
    float ivx, ivy, ivz;       /* 1/u per axis (0 for a zero component)  */
    float tcur, s, t;
    int it;
    bool done;

    __device__ __forceinline__ void begin(const DevPhantom &p, float x, float y, float z, float ek_ev, int usedir,
                                          float vx, float vy, float vz, gp_rng &r)
    {
        if (sample(ek_ev, usedir, vx, vy, vz, r)) setup(p, x, y, z);
    }

    /* displacement: range rr (mass thickness) and unit direction; false
     * (nothing to walk) unless rr > 0 */
    __device__ __forceinline__ bool sample(float ek_ev, int usedir, float vx, float vy, float vz, gp_rng &r)
    {
        float ekin = ek_ev / 1.0e6f;
        float rex = ((0.1f * 5.44040782f) * ekin * ekin) / (0.369516529f + ekin);
        float sigma = rex / 2.0f;
        float dx = sigma * gp_rng_normal(&r);
        float dy = sigma * gp_rng_normal(&r);
        float dz = sigma * gp_rng_normal(&r);
        rr = sqrtf(dx * dx + dy * dy + dz * dz);
        if (usedir) {
            float tmp = sqrtf(vx * vx + vy * vy + vz * vz);
            dx = rr * vx / tmp;
            dy = rr * vy / tmp;
            dz = rr * vz / tmp;
        }
        t = 0.0f;
        ux = uy = uz = 0.0f;
        done = true;
        if (!(rr > 0.0f)) return false;
        float inv_rr = 1.0f / rr;
        ux = dx * inv_rr;
        uy = dy * inv_rr;
        uz = dz * inv_rr;
        return true;
    }

    /* offset steps of the walk direction (signs of ux, uy, uz, as the DDA's
     * steps) in a padded grid of nbx x nby bricks per z layer */
    __device__ __forceinline__ void deltas(int nbx, int nby)
    {
        const int sx = ux > 0.0f ? 1 : (ux < 0.0f ? -1 : 0);
        const int sy = uy > 0.0f ? 1 : (uy < 0.0f ? -1 : 0);
        const int sz = uz > 0.0f ? 1 : (uz < 0.0f ? -1 : 0);
        dxi = sx * 4;
        dyi = sy * 16;
        dzi = sz * 64;
        dxc = sx * (256 - 12);
        dyc = sy * (256 * nbx - 48);
        dzc = sz * (256 * nbx * nby - 192);
        flip = (sx > 0 ? 12u : 0u) | (sy > 0 ? 48u : 0u) | (sz > 0 ? 192u : 0u);
    }

    __device__ __forceinline__ void inverses()
    {
        ivx = ux != 0.0f ? 1.0f / ux : 0.0f;
        ivy = uy != 0.0f ? 1.0f / uy : 0.0f;
        ivz = uz != 0.0f ? 1.0f / uz : 0.0f;
    }

    /* One jump across a uniform box of the phantom's walk map
     * (gp_walkmap.h; the oracle's range_jump, same arithmetic): every
     * crossing before the box's exit enters a voxel of the box's density, so
     * the walk moves in one step to the crossing into the box's last voxel
     * layer on its exit side (ties between axes: x, then y, then z); the
     * crossings after it are left to the voxel steps.  The new voxel along
     * the other axes is the one holding the point at the new parameter,
     * kept inside the box.  (x0, y0, z0): the walk's start.  UNIFORM: one
     * phantom per wave (buffer load bounded by the map's size); otherwise a
     * plain load (the brick index is always inside the map). */
    /* walk-map record of the brick holding byte offset O of the bricked
     * densities (UNIFORM: a buffer load bounded by the map's size; else the
     * brick index clamped into the map) */
    template <bool UNIFORM>
    __device__ __forceinline__ uint4 map_at(const DevPhantom &p, unsigned O) const
    {
        if (UNIFORM) {
            const uint64_t base = (uint64_t)(uintptr_t)p.wmap;
            const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
            const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(uintptr_t)(((uint64_t)hi << 32) | lo), (short)0,
                (int)__builtin_amdgcn_readfirstlane(p.wmap_bytes), 0x00020000);
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (O >> 8) << 4, 0, 0);
            return make_uint4(v[0], v[1], v[2], v[3]);
        }
        return p.wmap[min(O >> 8, (p.wmap_bytes >> 4) - 1u)];
    }

    template <bool UNIFORM>
    __device__ __forceinline__ void jump(const DevPhantom &p, float x0, float y0, float z0, unsigned &njump)
    {
        if (done || !p.wmap) return;
        ++njump;
        jump_with(p, map_at<UNIFORM>(p, o), x0, y0, z0);
    }

    /* the jump with the record m of the current brick */
    __device__ __forceinline__ void jump_with(const DevPhantom &p, const uint4 m, float x0, float y0, float z0)
    {
        const float rho = __uint_as_float(m.w);
        if (!(rho >= 0.0f)) return;
        const int lx = (int)(m.x & 0xffffu), hx = (int)(m.x >> 16);
        const int ly = (int)(m.y & 0xffffu), hy = (int)(m.y >> 16);
        const int lz = (int)(m.z & 0xffffu), hz = (int)(m.z >> 16);
        /* padded bounds: the crossing into the exit layer is boundary hi - 1
         * going up, lo going down (phantom voxel i is padded i + 1) */
        const float Tx = ux > 0.0f ? (p.offx + (float)(hx - 1) * p.dx - x0) * ivx
                                   : (ux < 0.0f ? (p.offx + (float)lx * p.dx - x0) * ivx : INFINITY);
        const float Ty = uy > 0.0f ? (p.offy + (float)(hy - 1) * p.dy - y0) * ivy
                                   : (uy < 0.0f ? (p.offy + (float)ly * p.dy - y0) * ivy : INFINITY);
        const float Tz = uz > 0.0f ? (p.offz + (float)(hz - 1) * p.dz - z0) * ivz
                                   : (uz < 0.0f ? (p.offz + (float)lz * p.dz - z0) * ivz : INFINITY);
        const bool ax = (Tx <= Ty) & (Tx <= Tz);
        const bool ay = !ax & (Ty <= Tz);
        const bool az = !ax & !ay;
        const float T = ax ? Tx : (ay ? Ty : Tz);
        if (!(T > tcur)) return;
        const float sn = s + (T - tcur) * rho;
        if (sn >= rr) {
            t = tcur + (rr - s) / rho;
            done = true;
            return;
        }
        s = sn;
        tcur = T;
        int px = (int)floorf(((x0 + T * ux) - p.offx) * p.idx) + 1;
        int py = (int)floorf(((y0 + T * uy) - p.offy) * p.idy) + 1;
        int pz = (int)floorf(((z0 + T * uz) - p.offz) * p.idz) + 1;
        px = ax ? (ux > 0.0f ? hx : lx) : min(max(px, lx), hx);
        py = ay ? (uy > 0.0f ? hy : ly) : min(max(py, ly), hy);
        pz = az ? (uz > 0.0f ? hz : lz) : min(max(pz, lz), hz);
        if (ux != 0.0f) tmx = (p.offx + (float)(ux > 0.0f ? px : px - 1) * p.dx - x0) * ivx;
        if (uy != 0.0f) tmy = (p.offy + (float)(uy > 0.0f ? py : py - 1) * p.dy - y0) * ivy;
        if (uz != 0.0f) tmz = (p.offz + (float)(uz > 0.0f ? pz : pz - 1) * p.dz - z0) * ivz;
        o = ((((unsigned)(pz >> 2) * (unsigned)p.nby + (unsigned)(py >> 2)) * (unsigned)p.nbx + (unsigned)(px >> 2))
             << 8) | ((unsigned)(((pz & 3) << 4) | ((py & 3) << 2) | (px & 3)) << 2);
    }

    /* the DDA from (x, y, z) along (ux, uy, uz); outside the phantom the
     * walk ends at once in air */
    __device__ __forceinline__ void setup(const DevPhantom &p, float x, float y, float z)
    {
        t = 0.0f;
        done = true;
        int ix = 0, iy = 0, iz = 0, sx, sy, sz;
        if (!(p.present && d_vox(p, x, y, z, ix, iy, iz))) {
            t = rr / D_RHO_OUTSIDE;
            return;
        }
        inverses();
        d_dda_axis(x, p.offx, p.dx, ix, ux, ivx, tmx, tdx, sx);
        d_dda_axis(y, p.offy, p.dy, iy, uy, ivy, tmy, tdy, sy);
        d_dda_axis(z, p.offz, p.dz, iz, uz, ivz, tmz, tdz, sz);
        o = (unsigned)gpet_brick_offset(ix + 1, iy + 1, iz + 1, p.nbx, p.nby) << 2;
        deltas(p.nbx, p.nby);
        tcur = 0.0f;
        s = 0.0f;
        it = 0;
        done = false;
    }

    /* The crossing sequence does not depend on the densities: advance the DDA
     * four crossings, issue their gathers together, then accumulate in order
     * (identical arithmetic, 4x memory parallelism).  Written with selects,
     * not branches: lanes of a wave take different axes at every crossing.
     * The walk ends at the crossing where the accumulated density-weighted
     * length reaches rr, or at the first crossing out of the phantom, which
     * reads the border's -D_RHO_OUTSIDE; there t = tn + (rr - s)/rho, or tn
     * when s == rr exactly (rho = |density|, so one rule covers both ends).
     * Crossings after the end in the same call read whatever lies there and
     * are ignored.  If the walk goes on, the group ends with the jump across
     * the uniform box it now sits in, whose walk-map record is loaded with
     * the densities (the walk starts with one jump, so the order is that of
     * the oracle: jump, four crossings, jump, ...). */
    template <bool UNIFORM = true, int K = WALK_K>
    __device__ __forceinline__ void step4(const DevPhantom &p, unsigned &nsteps, unsigned &njump, float x0, float y0,
                                          float z0)
    {
        float tn[K];
        float g[K];
        /* UNIFORM (one phantom per wave): the densities are read through a
         * buffer descriptor with the array's size as its range (a read past
         * the range, possible only after the end, returns 0) and 32-bit
         * offsets.  Otherwise plain loads at offsets clamped into the
         * array. */
        __amdgpu_buffer_rsrc_t rs;
        if (UNIFORM) {
            const uint64_t base = (uint64_t)(uintptr_t)p.dens_b;
            const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
            const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
            rs = __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)(((uint64_t)hi << 32) | lo), (short)0,
                                                   (int)__builtin_amdgcn_readfirstlane(p.dens_bytes), 0x00020000);
        }
        /* values, not members, in the selects below: a select between
         * members becomes a load through a selected address, and the walk
         * state then lives in scratch memory */
        const int Dxi = dxi, Dyi = dyi, Dzi = dzi, Dxc = dxc, Dyc = dyc, Dzc = dzc;
        const unsigned F = flip;
        unsigned O = o;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const bool bx = (tmx <= tmy) & (tmx <= tmz);
            const bool by = !bx & (tmy <= tmz);
            const bool bz = !bx & !by;
            tn[k] = bx ? tmx : (by ? tmy : tmz);
            const float nx_ = tmx + tdx, ny_ = tmy + tdy, nz_ = tmz + tdz;
            tmx = bx ? nx_ : tmx;
            tmy = by ? ny_ : tmy;
            tmz = bz ? nz_ : tmz;
            const unsigned msk = bx ? 12u : (by ? 48u : 192u);
            const bool wrap = ((O ^ F) & msk) == 0u;
            const int di = bx ? Dxi : (by ? Dyi : Dzi);
            const int dc = bx ? Dxc : (by ? Dyc : Dzc);
            O += (unsigned)(wrap ? dc : di);
            if (UNIFORM) {
                g[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, O, 0, 0));
            } else {
                const unsigned oc = min(O, p.dens_bytes - 4u);
                g[k] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p.dens_b) + oc);
            }
        }
        o = O;
        /* the record for the jump after these crossings, loaded with their
         * densities (used only if the walk goes on, when O is inside) */
        const bool use_map = p.wmap != nullptr;
        uint4 mnext = make_uint4(0u, 0u, 0u, 0u);
        if (use_map) mnext = map_at<UNIFORM>(p, O);
        const bool was = done;
        float ftn = 0.0f, fs = 0.0f, frho = 1.0f;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float rho = fabsf(g[k]);
            const bool inb = !(g[k] < 0.0f);
            const float sn = s + (tn[k] - tcur) * rho;
            const bool end = (sn >= rr) | !inb;
            const bool first = end & !done;
            nsteps += (!done & inb) ? 1u : 0u;
            ftn = first ? tn[k] : ftn;
            fs = first ? sn : fs;
            frho = first ? rho : frho;
            done = done | end;
            s = sn;
            tcur = tn[k];
        }
        if (done & !was) t = (fs == rr) ? ftn : ftn + (rr - fs) / frho;
        it += K;
        if (it >= D_STEP_CAP) done = true;
        if (use_map && !done) {
            ++njump;
            jump_with(p, mnext, x0, y0, z0);
        }
    }

    __device__ __forceinline__ void end(float &x, float &y, float &z) const
    {
        if (!(rr > 0.0f)) return;
        x = x + ux * t;
        y = y + uy * t;
        z = z + uz * t;
    }
};

__device__ void d_positron_range(const DevPhantom &p, float &x, float &y, float &z, float ek_ev, int usedir,
                                 float vx, float vy, float vz, gp_rng &r, unsigned long long &nsteps,
                                 unsigned long long &njumps)
{
    RangeWalk w;
    w.begin(p, x, y, z, ek_ev, usedir, vx, vy, vz, r);
    unsigned n = 0, nj = 0;
    w.jump<true>(p, x, y, z, nj);
    while (!w.done) w.step4(p, n, nj, x, y, z);
    nsteps += n;
    njumps += nj;
    w.end(x, y, z);
}

__device__ void d_emit_pair(float sigma, gp_rng &r, float v[2][3], float e[2])
{
    float ct = 2.0f * gp_rng_uniform(&r) - 1.0f;
    float sp, cp;
    gp_sincos2pi(gp_rng_uniform(&r), &sp, &cp);
    float st = sqrtf(1.0f - ct * ct);
    v[0][0] = st * cp;
    v[0][1] = st * sp;
    v[0][2] = ct;
    float uphi = gp_rng_uniform(&r);
    float delta = gp_rng_normal(&r) * sigma;
    e[0] = D_MC2F + (delta * D_MC2F) / 2.0f;
    e[1] = D_MC2F - (delta * D_MC2F) / 2.0f;
    v[1][0] = v[0][0];
    v[1][1] = v[0][1];
    v[1][2] = v[0][2];
    float sd, cd;
    gp_sincosf(delta, &sd, &cd);
    d_rotate_cs(v[1][0], v[1][1], v[1][2], -cd, fabsf(sd), uphi);
}

/* Woodcock tracking in the phantom; returns false when absorbed. */
enum { WALK_ABSORBED = 0, WALK_ESCAPED = 1, WALK_IN_FLIGHT = 2 };

/* it counts the photon's Woodcock steps so far (across calls); the call
 * takes at most maxit more and reports WALK_IN_FLIGHT if the photon is
 * still in the phantom then.  A photon reaching D_STEP_CAP steps is
 * absorbed, as in one uninterrupted walk. */
__device__ int d_phantom_walk(const DevTables &tab, const DevPhantom &p, float eabs, float &x, float &y, float &z,
                              float &vx, float &vy, float &vz, float &E, double &tof, gp_rng &r, int &nsteps,
                              int &ncomp, int &nrayl, unsigned long long &nmaj, int &it, int maxit)
{
    const int stop = (maxit >= D_STEP_CAP - it) ? D_STEP_CAP : it + maxit;
    for (; it < stop; ++it) {
        /* one 4-word block per step: length, selector, scattering CDF, azimuth */
        gp_u4 b = gp_rng_block(&r);
        float lammin = d_lam(p.wck, E);
        float s = -lammin * gp_logf(gp_u01(b.x));
        x = x + s * vx;
        y = y + s * vy;
        z = z + s * vz;
        tof = tof + (double)s * D_INV_C;
        int ix, iy, iz;
        if (!d_vox(p, x, y, z, ix, iy, iz)) { ++it; return WALK_ESCAPED; }
        nsteps++;
        int2 vv = p.vox[((size_t)iz * p.ny + iy) * p.nx + ix];
        float den = __int_as_float(vv.x);
        int mat = vv.y;
        float3 mu = d_mu3(tab, mat, E);
        float lamden = lammin * den;
        float prob = 1.0f - lamden * mu.x;
        if (prob < 0.0f) { prob = 0.0f; nmaj++; }
        float rn = gp_u01(b.y);
        if (rn < prob) continue;
        /* Compton and Rayleigh share one path (surface lookup, rotation), so
         * a wave that holds both pays for it once */
        const float pc = prob + lamden * mu.y;
        const bool comp = rn < pc;
        if (!comp && !(rn < pc + lamden * mu.z)) { ++it; return WALK_ABSORBED; }   /* photoelectric */
        const float cth = d_surf(tab, comp ? GPET_SURF_COMPTON : GPET_SURF_RAYLEIGH, mat, E, gp_u01(b.z));
        const float uphi = gp_u01(b.w);
        if (comp) {
            const float efrac = 1.0f / (1.0f + (E * D_IMC2F) * (1.0f - cth));
            ncomp++;
            E = E * efrac;
            if (E < eabs) { ++it; return WALK_ABSORBED; }
        } else {
            nrayl++;
        }
        d_rotate(vx, vy, vz, cth, uphi);
    }
    return it >= D_STEP_CAP ? WALK_ABSORBED : WALK_IN_FLIGHT;
}

/* Distance along the direction to the recording sphere (getDistance,
 * gPET_kernals.cu:148-171): the far root from inside; from outside the near
 * root when approaching, the far (negative) one when receding; 0 when the
 * line misses. */
__device__ __forceinline__ float d_sphere_dist(const float *sp, float x, float y, float z, float vx, float vy,
                                               float vz)
{
    const float cx = x - sp[0], cy = y - sp[1], cz = z - sp[2];
    const float a = vx * vx + vy * vy + vz * vz;
    const float b = 2.0f * (vx * cx + vy * cy + vz * cz);
    const float c = (cx * cx + cy * cy + cz * cz) - sp[3] * sp[3];
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return 0.0f;
    if (c < 0.0f) return (-b + sqrtf(disc)) / (2.0f * a);
    if (b < 0.0f) return (-b - sqrtf(disc)) / (2.0f * a);
    return (-b + sqrtf(disc)) / (2.0f * a);
}

__device__ __forceinline__ void d_psf_append(gpet_psf_record *buf, unsigned long long *count, int64_t cap, float x,
                                             float y, float z, float vx, float vy, float vz, float E, int parn,
                                             int eid, double t)
{
    const unsigned long long i = wave_append1(count);
    if ((int64_t)i < cap) {
        gpet_psf_record r;
        r.x = x; r.y = y; r.z = z; r.vx = vx; r.vy = vy; r.vz = vz; r.E = E;
        r.parn = parn; r.eventid = eid; r.pad = 0; r.t = t;
        buf[i] = r;
    }
}

struct PhCounters {
    unsigned long long ph = 0, ent = 0, maj = 0, ps = 0, rs = 0, rj = 0;
};

/* Source of history h: the largest si with first[si] <= h. */
__device__ __forceinline__ int d_source_of(const RunParams &rp, int64_t h)
{
    int lo = 0, hi = rp.nsrc;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (rp.first[mid] <= h) lo = mid; else hi = mid;
    }
    return lo;
}

/* Decay of history h of source si up to the positron range: decay time
 * (us), whether it emits a pair, the emission point and the positron's
 * kinetic energy (eV; sampled only with positron range).  The stream is left
 * where the range walk starts. */
__device__ __forceinline__ bool d_decay_begin(const RunParams &rp, int64_t h, int si, gp_rng &r, double &tstart,
                                              float &px, float &py, float &pz, float &ek)
{
    const DevSource S = rp.src[si];
    const DevIsotope I = rp.iso[S.iso];
    gp_rng_init(&r, rp.seed, (uint64_t)h, GP_TAG_DECAY, 0u);
    double u = gp_rng_uniform_d(&r);
    double ts = S.t0 + S.tau * (-gp_log1p_d(-u * S.pdec));
    tstart = ts * 1.0e6;
    float rb = gp_rng_uniform(&r);
    if (!(rb < I.ratio)) return false;
    gp_rng_align(&r);
    d_sample_shape(S, r, px, py, pz);
    if (rp.prange) {
        gp_rng_align(&r);
        ek = d_sample_ek(I, r);
        gp_rng_align(&r);
    }
    return true;
}

template <int MODE, bool TQ>
__device__ __forceinline__ void d_photons(const RunParams &rp, const DevTables &tab, const DevPhantom &ph,
                                          const DevDetector &det, const float4 *s_pf, const DevPanel *s_p,
                                          const DevQueue &q, unsigned *s_qn, unsigned *s_tn, int64_t h, float px,
                                          float py, float pz, const float (&v)[2][3], const float (&e)[2],
                                          double tstart, int nph, int slot, PhCounters &C);

/* One history: decay (or PSF record i), positron range, pair, phantom walk
 * of both photons, panel entry and queue append. */
/* Crossing groups a positron walk may take in the phantom kernel before it
 * is handed to k_longwalk.  Walk lengths have a long tail (a positron that
 * leaves the water crosses the air to the grid's edge), and a wave runs
 * until its longest walk ends; deferring the tail lets the short walks'
 * waves move on, and k_longwalk runs the long ones together.  With uniform
 * boxes (a walk map) half the walks end in their first jump and the wave
 * maxima of the rest are long, so the hand-off comes after two groups;
 * without a map after eight (32 crossings). */
#ifndef LW_GROUPS_MAP
#define LW_GROUPS_MAP 2
#endif
#ifndef LW_GROUPS_NOMAP
#define LW_GROUPS_NOMAP 8
#endif
/* crossings per walk step in k_longwalk when no walk of the wave has a map
 * (without jumps the grouping does not change results) */
#ifndef LW_WALK_K
#define LW_WALK_K 8
#endif

template <int MODE, bool TQ>
__device__ __forceinline__ void d_history(const RunParams &rp, const DevTables &tab, const DevPhantom &ph,
                                          const DevDetector &det, const float4 *s_pf, const DevPanel *s_p,
                                          const DevQueue &q, unsigned *s_qn, unsigned *s_ln, unsigned *s_tn,
                                          int64_t i, int64_t h, int si, int slot, PhCounters &C)
{
    unsigned long long &c_rs = C.rs;
    float px = 0.f, py = 0.f, pz = 0.f;
    float v[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float e[2] = {0.f, 0.f};
    double tstart = 0.0;
    int nph = 0;
    if (MODE == RUN_DECAY) {
        gp_rng r;
        float ek = 0.0f;
        if (d_decay_begin(rp, h, si, r, tstart, px, py, pz, ek)) {
            if (rp.prange) {
                RangeWalk w;
                w.begin(ph, px, py, pz, ek, 0, 0.f, 0.f, 0.f, r);
                unsigned ns = 0, nj = 0;
                w.jump<true>(ph, px, py, pz, nj);
                if (q.lw) {
                    const int groups = ph.wmap ? LW_GROUPS_MAP : LW_GROUPS_NOMAP;
                    for (int k = 0; k < groups && !w.done; ++k) w.step4(ph, ns, nj, px, py, pz);
                    if (!w.done) {
                        const unsigned lcap = (unsigned)(q.region >> 3);
                        const unsigned li = atomicAdd(s_ln, 1u);
                        if (li < lcap) {
                            LongWalkItem L;
                            L.ux = w.ux; L.uy = w.uy; L.uz = w.uz; L.rr = w.rr;
                            L.tmx = w.tmx; L.tmy = w.tmy; L.tmz = w.tmz;
                            L.tdx = w.tdx; L.tdy = w.tdy; L.tdz = w.tdz;
                            L.tcur = w.tcur; L.s = w.s;
                            L.o = w.o;
                            L.it = w.it;
                            L.i = (int)i;
                            L.px = px; L.py = py; L.pz = pz;
                            L.n = (r.n + 3u) & ~3u;
                            L.tstart = tstart;
                            L.slot = slot;
                            L.pad = 0;
                            q.lw[(int64_t)blockIdx.x * lcap + li] = L;
                            c_rs += ns;
                            C.rj += nj;
                            return;   /* k_longwalk finishes this history */
                        }
                    }
                }
                while (!w.done) w.step4(ph, ns, nj, px, py, pz);
                c_rs += ns;
                C.rj += nj;
                w.end(px, py, pz);
            }
            gp_rng_align(&r);
            d_emit_pair(rp.sigma, r, v, e);
            nph = 2;
        }
    } else {
        const double *rec = rp.psf + 8 * i;
        px = (float)rec[0];
        py = (float)rec[1];
        pz = (float)rec[2];
        tstart = rec[3];
        float rvx = (float)rec[4], rvy = (float)rec[5], rvz = (float)rec[6], rE = (float)rec[7];
        if (MODE == RUN_PSF_PHOTON) {
            v[0][0] = rvx; v[0][1] = rvy; v[0][2] = rvz; e[0] = rE;
            nph = 1;
        } else {
            gp_rng r;
            gp_rng_init(&r, rp.seed, (uint64_t)h, GP_TAG_DECAY, 0u);
            if (rp.prange) d_positron_range(ph, px, py, pz, rE, 1, rvx, rvy, rvz, r, c_rs, C.rj);
            gp_rng_align(&r);
            d_emit_pair(rp.sigma, r, v, e);
            nph = 2;
        }
    }
    d_photons<MODE, TQ>(rp, tab, ph, det, s_pf, s_p, q, s_qn, s_tn, h, px, py, pz, v, e, tstart, nph, slot, C);
}

/* Panel tables in LDS, laid out as [entry spheres][group spheres][group
 * masks][panels] (gpet_panel_lds_bytes); filled by the whole block, read
 * after a barrier.  Returns the panel array. */
__device__ __forceinline__ DevPanel *d_panel_lds(unsigned char *smem, const DevDetector &det)
{
    float4 *pf = reinterpret_cast<float4 *>(smem);
    float4 *gf = pf + det.npanels;
    unsigned long long *gm = reinterpret_cast<unsigned long long *>(gf + det.ngroups);
    DevPanel *sp = reinterpret_cast<DevPanel *>(gm + det.ngroups);
    if (det.present) {
        for (int t = threadIdx.x; t < det.npanels; t += blockDim.x) {
            const DevPanel P = det.p[t];
            pf[t] = make_float4(P.ox, P.oy, P.oz, P.bound2);
            sp[t] = P;
        }
        for (int t = threadIdx.x; t < det.ngroups; t += blockDim.x) {
            gf[t] = det.gsph[t];
            gm[t] = det.gmask[t];
        }
    }
    return sp;
}

/* The line through (x, y, z) along v (|v|^2 = vv) passes within the sphere
 * (f.xyz, f.w = squared radius), with a bound that grows with the squared
 * distance w2 to the centre: the squared distance loses about 2^-22 w2 to
 * rounding, and the exact entry test's own rounding moves its crossing
 * point by a similar relative amount (1e-5 w2: nothing at scanner
 * distances; photons many metres away -- after a long flight through an
 * empty phantom -- are not skipped where the exact test would still accept
 * a face).  An overflow to inf or NaN passes. */
__device__ __forceinline__ bool d_line_near(const float4 f, float x, float y, float z, float vx, float vy, float vz,
                                            float vv)
{
    float wx = f.x - x, wy = f.y - y, wz = f.z - z;
    float a = vx * wx + vy * wy + vz * wz;
    float w2 = wx * wx + wy * wy + wz * wz;
    return !(w2 * vv - a * a > (f.w + 1.0e-5f * w2) * vv);
}

/* What follows a photon's phantom walk: recording sphere, phase-space
 * record, trace record, panel entry and the append to this block's queue
 * region (LDS counter, one atomic per wave among the lanes that call it).
 * draws: the photon's stream position, where the panel kernel goes on. */
__device__ __forceinline__ void d_photon_tail(const RunParams &rp, const DevDetector &det, const float4 *s_pf,
                                              const DevPanel *s_p, const DevQueue &q, unsigned *s_qn, int64_t h,
                                              int k, int parn, bool active, bool alive, float x, float y, float z,
                                              float vx, float vy, float vz, float E, double t, int ns, int nc,
                                              int nr, unsigned draws, float px, float py, float pz, float E0,
                                              double tstart, PhCounters &C, bool head = true)
{
    const int64_t slot = 2 * h + k;
    gpet_trace *tr = (rp.trace && slot < rp.trace_cap) ? rp.trace + slot : nullptr;
    if (active) {
        if (alive && rp.sphere[3] > 0.0f) {
            const float sd = d_sphere_dist(rp.sphere, x, y, z, vx, vy, vz);
            x = x + sd * vx;
            y = y + sd * vy;
            z = z + sd * vz;
            t = t + (double)sd * D_INV_C;
        }
        if (alive && rp.psf_ph)
            d_psf_append(rp.psf_ph, rp.n_psf + 1, rp.psf_cap, x, y, z, vx, vy, vz, E, parn, (int)h, t);
    }
    C.ps += (unsigned long long)ns;
    if (tr) {
        tr->history = h;
        tr->photon = k;
        tr->panel = -1;
        if (head) tr->t0 = tstart;
        if (active) {
            tr->status = GPET_TR_ALIVE | (alive ? GPET_TR_ESCAPED : GPET_TR_ABSORBED);
            tr->n_steps = ns; tr->n_compton = nc; tr->n_rayleigh = nr;
            if (head) { tr->x0 = px; tr->y0 = py; tr->z0 = pz; tr->E0 = E0; }
            tr->x = x; tr->y = y; tr->z = z; tr->vx = vx; tr->vy = vy; tr->vz = vz; tr->E = E;
            tr->t = t;
            tr->draws = draws;
        }
    }
    int pa = -1;
    float y2 = 0.f, z2 = 0.f, vx2 = 0.f, vy2 = 0.f, vz2 = 0.f;
    if (alive && det.present) {
        const float vv = vx * vx + vy * vy + vz * vz;
        /* the exact entry test of panel pi (conservative prefilter first:
         * the line through the photon must pass within the face's bounding
         * sphere, which covers the reference's backward entries too;
         * skipping a panel there never changes which panel the exact test
         * picks) */
        auto try_panel = [&](int pi) -> bool {
            if (!d_line_near(s_pf[pi], x, y, z, vx, vy, vz, vv)) return false;
            const DevPanel &P = s_p[pi];
            float rx = x - P.ox, ry = y - P.oy, rz = z - P.oz;
            float tx = rx * P.ux[0] + ry * P.ux[1] + rz * P.ux[2];
            float ty = rx * P.uy[0] + ry * P.uy[1] + rz * P.uy[2];
            float tz = rx * P.uz[0] + ry * P.uz[1] + rz * P.uz[2];
            float tvx = vx * P.ux[0] + vy * P.ux[1] + vz * P.ux[2];
            float tvy = vx * P.uy[0] + vy * P.uy[1] + vz * P.uy[2];
            float tvz = vx * P.uz[0] + vy * P.uz[1] + vz * P.uz[2];
            if (!(tvx * P.dirx >= 0.0f)) return false;
            float ty2 = ty - (tx / tvx) * tvy;
            float tz2 = tz - (tx / tvx) * tvz;
            if (!(fabsf(ty2) < P.leny / 2.0f && fabsf(tz2) < P.lenz / 2.0f)) return false;
            y2 = ty2;
            z2 = tz2;
            t = t + (double)(-tx) / (D_C_LIGHT * (double)tvx);
            vx2 = tvx; vy2 = tvy; vz2 = tvz;
            pa = pi;
            return true;
        };
        const int G = det.ngroups;
        if (G > 0) {
            /* candidates: the members of the groups whose sphere the line
             * passes near, tried in index order, so the first panel the
             * exact test accepts is the reference's */
            const float4 *s_gf = s_pf + det.npanels;
            const unsigned long long *s_gm = reinterpret_cast<const unsigned long long *>(s_gf + G);
            unsigned long long cand = 0;
            for (int g = 0; g < G; ++g)
                if (d_line_near(s_gf[g], x, y, z, vx, vy, vz, vv)) cand |= s_gm[g];
            while (cand) {
                const int pi = __ffsll(cand) - 1;
                cand &= cand - 1;
                if (try_panel(pi)) break;
            }
        } else {
            for (int pi = 0; pi < det.npanels; ++pi)
                if (try_panel(pi)) break;
        }
    }
    if (tr && pa >= 0) {
        tr->panel = pa;
        tr->status |= GPET_TR_PANEL;
    }
    const bool want = pa >= 0;
    const unsigned long long m = __ballot(want);
    int64_t idx = q.region;
    if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        unsigned b = 0;
        if ((int)__lane_id() == leader) b = atomicAdd(s_qn, (unsigned)__popcll(m));
        b = __shfl(b, leader);
        idx = (int64_t)b + __popcll(m & lanemask_lt());
    }
    if (want) {
        C.ent++;
        if (idx < q.region) {
            idx += (int64_t)blockIdx.x * q.region;
            q.y[idx] = y2; q.z[idx] = z2;
            q.vx[idx] = vx2; q.vy[idx] = vy2; q.vz[idx] = vz2;
            q.E[idx] = E;
            q.t[idx] = t;
            q.info[idx] = make_int4(pa, parn, (int)h, (int)(draws | ((unsigned)k << 31)));
            if (q.trace_slot) q.trace_slot[idx] = tr ? (int)slot : -1;
        }
    }
}

/* Slot of a photon that goes to this block's transit region (LDS counter,
 * one atomic per wave among the lanes that call it), or null.  A region
 * holds two photons per history of its block, so it is never full. */
__device__ __forceinline__ TransitItem *d_transit_slot(TransitItem *tq, int64_t region, unsigned *s_tn, bool push)
{
    const unsigned long long m = __ballot(push);
    if (!m) return nullptr;
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned b = 0;
    if ((int)__lane_id() == leader) b = atomicAdd(s_tn, (unsigned)__popcll(m));
    b = __shfl(b, leader);
    const int64_t idx = (int64_t)b + __popcll(m & lanemask_lt());
    return (push && idx < region) ? tq + (int64_t)blockIdx.x * region + idx : nullptr;
}

__device__ __forceinline__ void d_transit_write(TransitItem *T, double t, int64_t h, float x, float y, float z,
                                                float vx, float vy, float vz, float E, int ns, int nc, int nr, int it,
                                                unsigned n, int parn, int slot)
{
    T->t = t; T->h = h;
    T->x = x; T->y = y; T->z = z; T->vx = vx; T->vy = vy; T->vz = vz; T->E = E;
    T->ns = ns; T->nc = nc; T->nr = nr; T->it = it;
    T->n = n; T->parn = parn; T->slot = slot;
}

/* Both photons of history h (nph of them active) from (px, py, pz) at
 * tstart: phantom walk (TQ: a photon still in the phantom after q.tq_steps
 * steps is handed to the transit kernel), then d_photon_tail. */
template <int MODE, bool TQ>
__device__ __forceinline__ void d_photons(const RunParams &rp, const DevTables &tab, const DevPhantom &ph,
                                          const DevDetector &det, const float4 *s_pf, const DevPanel *s_p,
                                          const DevQueue &q, unsigned *s_qn, unsigned *s_tn, int64_t h, float px,
                                          float py, float pz, const float (&v)[2][3], const float (&e)[2],
                                          double tstart, int nph, int slot, PhCounters &C)
{
    auto photon = [&](const int k) {
        const bool active = k < nph;
        float x = px, y = py, z = pz, vx = v[k][0], vy = v[k][1], vz = v[k][2], E = e[k];
        double t = tstart;
        gp_rng r;
        gp_rng_init(&r, rp.seed, (uint64_t)h, (uint32_t)(GP_TAG_PHOTON + k), 0u);
        int ns = 0, nc = 0, nr = 0, it = 0;
        bool alive = active, defer = false;
        const int parn = (MODE == RUN_PSF_PHOTON) ? (int)h : (int)(2 * h + k);
        if (active) {
            C.ph++;
            if (rp.psf_src)
                d_psf_append(rp.psf_src, rp.n_psf, rp.psf_cap, x, y, z, vx, vy, vz, E, parn, (int)h, t);
            if (ph.present) {
                const int st = d_phantom_walk(tab, ph, rp.eabs, x, y, z, vx, vy, vz, E, t, r, ns, nc, nr, C.maj, it,
                                              TQ ? q.tq_steps : D_STEP_CAP);
                alive = st != WALK_ABSORBED;
                defer = st == WALK_IN_FLIGHT;
            }
        }
        if (TQ) {
            TransitItem *T = d_transit_slot(q.tq, q.region, s_tn, defer);
            if (T) {
                d_transit_write(T, t, h, x, y, z, vx, vy, vz, E, ns, nc, nr, it, r.n | ((unsigned)k << 31), parn,
                                slot);
                /* emission fields of a traced photon: written now */
                const int64_t ts = 2 * h + k;
                if (rp.trace && ts < rp.trace_cap) {
                    gpet_trace *tr = rp.trace + ts;
                    tr->t0 = tstart;
                    tr->x0 = px; tr->y0 = py; tr->z0 = pz; tr->E0 = e[k];
                }
            }
        }
        if (!defer)
            d_photon_tail(rp, det, s_pf, s_p, q, s_qn, h, k, parn, active, alive, x, y, z, vx, vy, vz, E, t, ns, nc,
                          nr, r.n, px, py, pz, e[k], tstart, C);
    };
    /* the regrouping variant keeps one photon's state at a time (unrolled,
     * its hand-off would keep both and cost a wave per SIMD) */
    if (TQ) {
#pragma unroll 1
        for (int k = 0; k < 2; ++k) photon(k);
    } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) photon(k);
    }
}

/* The phantom's Woodcock majorant (GPET_NWCK floats, 16 KB) staged in LDS
 * per block for single-phantom launches when GPET_PHANTOM_LDS_WCK is set:
 * the kernels then read it through phl, a copy of the phantom descriptor
 * whose majorant pointer is the LDS copy (same nodes, same arithmetic).
 * Off by default: on MI355X the phantom stage ran 1-3% slower with it (body
 * phantom 3.11-3.20 -> 3.21 ms per 2^23-history launch, cylinder 4.27 ->
 * 4.30-4.33 ms per 2^25), since the table is L1/L2-resident anyway and the
 * stage is bound by instruction issue, not by these lookups.  The material
 * tables (16 B per node and material) do not fit: the nodes up to 530 keV
 * of the body phantom's four materials take 139 KB, one block per CU. */
#ifndef GPET_PHANTOM_LDS_WCK
#define GPET_PHANTOM_LDS_WCK 0
#endif
#if GPET_PHANTOM_LDS_WCK
#define PHANTOM_LDS_WCK(BATCH, ph, phl)                                                            \
    __shared__ float s_wck[GPET_NWCK];                                                             \
    DevPhantom phl = ph;                                                                           \
    if (!(BATCH) && ph.wck.lam) {                                                                  \
        for (int t_ = threadIdx.x; t_ < GPET_NWCK; t_ += blockDim.x) s_wck[t_] = ph.wck.lam[t_];   \
        phl.wck.lam = s_wck;                                                                       \
    }
#else
#define PHANTOM_LDS_WCK(BATCH, ph, phl) const DevPhantom &phl = ph;
#endif

/* The phantom kernel, grid-stride over histories.
 *
 * BATCH: the launch covers a list of segments, each a history range of one
 * phantom slot.  Work is dealt in virtual blocks of blockDim histories that
 * never straddle a segment, so the slot is block-uniform: its descriptor is
 * read with scalar loads and held in SGPRs, and no wave mixes phantoms.
 * WPE > 0 asks the compiler for at least WPE waves per SIMD (a register
 * budget), one of the launch variants the tuner picks from. */
template <int MODE, bool BATCH, int WPE, bool TQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_phantom(RunParams rp, DevTables tab, DevPhantom ph, DevDetector det, DevQueue q)
{
    PhCounters C;
    __shared__ unsigned s_qn, s_ln, s_tn;
    if (threadIdx.x == 0) {
        s_qn = 0u;
        s_ln = 0u;
        s_tn = 0u;
    }
    /* panel table in LDS: the entry test scans every panel for every photon,
     * and from LDS the scan is not a chain of memory round trips */
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const float4 *s_pf = reinterpret_cast<const float4 *>(smem);
    const DevPanel *s_p = d_panel_lds(smem, det);
    PHANTOM_LDS_WCK(BATCH, ph, phl)
    __syncthreads();
    if (!BATCH) {
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rp.nh; i += stride) {
            const int64_t h = rp.h0 + i;
            const int si = (MODE == RUN_DECAY) ? d_source_of(rp, h) : 0;
            d_history<MODE, TQ>(rp, tab, phl, det, s_pf, s_p, q, &s_qn, &s_ln, &s_tn, i, h, si, 0, C);
        }
    } else {
        for (int64_t vb = blockIdx.x; vb < rp.n_vblocks; vb += gridDim.x) {
            int lo = 0, hi = rp.nseg; /* largest segment with vb0 <= vb */
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (rp.seg[mid].vb0 <= vb) lo = mid; else hi = mid;
            }
            const DevSegment S = rp.seg[lo];
            const int64_t off = (vb - S.vb0) * blockDim.x + threadIdx.x;
            const DevPhantom P = rp.phs[S.phantom];
            if (off < S.nh) {
                const int64_t h = S.h0 + off;
                const int si = d_source_of(rp, h);
                d_history<MODE, TQ>(rp, tab, P, det, s_pf, s_p, q, &s_qn, &s_ln, &s_tn, h - rp.h0, h, si, S.phantom,
                                    C);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        q.rcount[blockIdx.x] = min((unsigned)q.region, s_qn);
        if (q.lw) q.lw_count[blockIdx.x] = min((unsigned)(q.region >> 3), s_ln);
        if (TQ) q.tq_count[blockIdx.x] = min((unsigned)q.region, s_tn);
    }
    {
        unsigned long long *const dst[6] = {
            rp.cnt + CNT_PHOTONS * CNT_STRIPES, rp.cnt + CNT_PANEL_ENTRIES * CNT_STRIPES,
            rp.cnt + CNT_MAJORANT * CNT_STRIPES, rp.cnt + CNT_PHANTOM_STEPS * CNT_STRIPES,
            rp.cnt + CNT_RANGE_STEPS * CNT_STRIPES, rp.cnt + CNT_RANGE_JUMPS * CNT_STRIPES};
        const unsigned long long v[6] = {C.ph, C.ent, C.maj, C.ps, C.rs, C.rj};
        flush_counters<6>(dst, v);
    }
}

/* The long positron walks of one phantom-kernel block (same grid): finish
 * each walk, emit the pair and transport both photons, appending to that
 * block's queue region after the entries the phantom kernel left there. */
template <bool BATCH, bool TQ>
__global__ __launch_bounds__(256) void k_longwalk(RunParams rp, DevTables tab, DevPhantom ph, DevDetector det,
                                                  DevQueue q)
{
    PhCounters C;
    __shared__ unsigned s_qn, s_tn;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const float4 *s_pf = reinterpret_cast<const float4 *>(smem);
    if (threadIdx.x == 0) {
        s_qn = q.rcount[blockIdx.x];
        s_tn = TQ ? q.tq_count[blockIdx.x] : 0u;
    }
    const DevPanel *s_p = d_panel_lds(smem, det);
    PHANTOM_LDS_WCK(BATCH, ph, phl)
    __syncthreads();
    const unsigned lcap = (unsigned)(q.region >> 3);
    const unsigned n = min(q.lw_count[blockIdx.x], lcap);
    for (unsigned j = threadIdx.x; j < n; j += blockDim.x) {
        const LongWalkItem L = q.lw[(int64_t)blockIdx.x * lcap + j];
        const DevPhantom P = BATCH ? rp.phs[L.slot] : phl;
        RangeWalk w;
        w.ux = L.ux; w.uy = L.uy; w.uz = L.uz; w.rr = L.rr;
        w.tmx = L.tmx; w.tmy = L.tmy; w.tmz = L.tmz;
        w.tdx = L.tdx; w.tdy = L.tdy; w.tdz = L.tdz;
        w.tcur = L.tcur; w.s = L.s;
        w.o = L.o;
        w.deltas(P.nbx, P.nby);
        w.inverses();
        w.it = L.it;
        w.t = 0.0f;
        w.done = false;
        unsigned ns = 0, nj = 0;
        float px = L.px, py = L.py, pz = L.pz;
        if (__ballot(P.wmap != nullptr) != 0ull) {
            while (!w.done) w.step4<!BATCH, WALK_K>(P, ns, nj, px, py, pz);
        } else {
            while (!w.done) w.step4<!BATCH, LW_WALK_K>(P, ns, nj, px, py, pz);
        }
        C.rs += ns;
        C.rj += nj;
        w.end(px, py, pz);
        const int64_t h = rp.h0 + L.i;
        gp_rng r;
        gp_rng_init(&r, rp.seed, (uint64_t)h, GP_TAG_DECAY, L.n);
        float v[2][3], e[2];
        d_emit_pair(rp.sigma, r, v, e);
        d_photons<RUN_DECAY, TQ>(rp, tab, P, det, s_pf, s_p, q, &s_qn, &s_tn, h, px, py, pz, v, e, L.tstart, 2,
                                 L.slot, C);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        q.rcount[blockIdx.x] = min((unsigned)q.region, s_qn);
        if (TQ) q.tq_count[blockIdx.x] = min((unsigned)q.region, s_tn);
    }
    {
        unsigned long long *const dst[6] = {
            rp.cnt + CNT_PHOTONS * CNT_STRIPES, rp.cnt + CNT_PANEL_ENTRIES * CNT_STRIPES,
            rp.cnt + CNT_MAJORANT * CNT_STRIPES, rp.cnt + CNT_PHANTOM_STEPS * CNT_STRIPES,
            rp.cnt + CNT_RANGE_STEPS * CNT_STRIPES, rp.cnt + CNT_RANGE_JUMPS * CNT_STRIPES};
        const unsigned long long v[6] = {C.ph, C.ent, C.maj, C.ps, C.rs, C.rj};
        flush_counters<6>(dst, v);
    }
}

/* Photons in transit, one round: block b continues the photons of its
 * region of `in` (same grid as the phantom kernel) by at most maxit
 * Woodcock steps; those still in the phantom go to its region of `out`
 * (null: the round finishes every photon), the rest take d_photon_tail and
 * append to the block's panel-queue region after the entries already there.
 * Every photon is continued from its own stream position, so the rounds
 * change nothing in the results. */
template <bool BATCH>
__global__ __launch_bounds__(256) void k_transit(RunParams rp, DevTables tab, DevPhantom ph, DevDetector det,
                                                 DevQueue q, const TransitItem *in, const unsigned *in_count,
                                                 TransitItem *out, unsigned *out_count, int maxit)
{
    PhCounters C;
    __shared__ unsigned s_qn, s_tn;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const float4 *s_pf = reinterpret_cast<const float4 *>(smem);
    if (threadIdx.x == 0) {
        s_qn = q.rcount[blockIdx.x];
        s_tn = 0u;
    }
    const DevPanel *s_p = d_panel_lds(smem, det);
    PHANTOM_LDS_WCK(BATCH, ph, phl)
    __syncthreads();
    const unsigned n = min(in_count[blockIdx.x], (unsigned)q.region);
    for (unsigned j = threadIdx.x; j < n; j += blockDim.x) {
        const TransitItem T = in[(int64_t)blockIdx.x * q.region + j];
        const DevPhantom P = BATCH ? rp.phs[T.slot] : phl;
        const int k = (int)(T.n >> 31);
        gp_rng r;
        gp_rng_init(&r, rp.seed, (uint64_t)T.h, (uint32_t)(GP_TAG_PHOTON + k), T.n & 0x7fffffffu);
        float x = T.x, y = T.y, z = T.z, vx = T.vx, vy = T.vy, vz = T.vz, E = T.E;
        double t = T.t;
        int ns = T.ns, nc = T.nc, nr = T.nr, it = T.it;
        const int st = d_phantom_walk(tab, P, rp.eabs, x, y, z, vx, vy, vz, E, t, r, ns, nc, nr, C.maj, it,
                                      out ? maxit : D_STEP_CAP);
        const bool defer = st == WALK_IN_FLIGHT;
        if (out) {
            TransitItem *U = d_transit_slot(out, q.region, &s_tn, defer);
            if (U)
                d_transit_write(U, t, T.h, x, y, z, vx, vy, vz, E, ns, nc, nr, it, r.n | ((unsigned)k << 31), T.parn,
                                T.slot);
        }
        if (!defer)
            d_photon_tail(rp, det, s_pf, s_p, q, &s_qn, T.h, k, T.parn, true, st != WALK_ABSORBED, x, y, z, vx, vy,
                          vz, E, t, ns, nc, nr, r.n, 0.0f, 0.0f, 0.0f, 0.0f, 0.0, C, false);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        q.rcount[blockIdx.x] = min((unsigned)q.region, s_qn);
        if (out) out_count[blockIdx.x] = min((unsigned)q.region, s_tn);
    }
    {
        unsigned long long *const dst[6] = {
            rp.cnt + CNT_PHOTONS * CNT_STRIPES, rp.cnt + CNT_PANEL_ENTRIES * CNT_STRIPES,
            rp.cnt + CNT_MAJORANT * CNT_STRIPES, rp.cnt + CNT_PHANTOM_STEPS * CNT_STRIPES,
            rp.cnt + CNT_RANGE_STEPS * CNT_STRIPES, rp.cnt + CNT_RANGE_JUMPS * CNT_STRIPES};
        const unsigned long long v[6] = {C.ph, C.ent, C.maj, C.ps, C.rs, C.rj};
        flush_counters<6>(dst, v);
    }
}

/* ---- panel -------------------------------------------------------------- */

/* A photon's adder list in registers.  cs is the crystal site,
 * (panel * moduleN + module) * crystalN + crystal: the adder's merge key,
 * from which the module, the crystal and the readout site follow, so the
 * list carries one word for the three. */
struct EvList {
    int cs[GPET_MAX_PH_EVENTS];
    double t[GPET_MAX_PH_EVENTS];
    float E[GPET_MAX_PH_EVENTS], x[GPET_MAX_PH_EVENTS], y[GPET_MAX_PH_EVENTS], z[GPET_MAX_PH_EVENTS];
};

__device__ __forceinline__ void d_adder(EvList &L, int &n, int site, double t, float E, float x, float y, float z,
                                        unsigned long long &dropped)
{
    bool merged = false;
#pragma unroll
    for (int i = 0; i < GPET_MAX_PH_EVENTS; ++i) {
        if (!merged && i < n && L.cs[i] == site) {
            float sE = L.E[i] + E;
            L.x[i] = (L.x[i] * L.E[i] + x * E) / sE;
            L.y[i] = (L.y[i] * L.E[i] + y * E) / sE;
            L.z[i] = (L.z[i] * L.E[i] + z * E) / sE;
            L.E[i] = sE;
            merged = true;
        }
    }
    if (merged) return;
    if (n >= GPET_MAX_PH_EVENTS) {
        dropped++;
        return;
    }
#pragma unroll
    for (int i = 0; i < GPET_MAX_PH_EVENTS; ++i) {
        if (i == n) {
            L.cs[i] = site; L.t[i] = t;
            L.E[i] = E; L.x[i] = x; L.y[i] = y; L.z[i] = z;
        }
    }
    n++;
}

/* readout (gPET_kernals.cu:756-813) in place on the register list: merged
 * entries are flagged dead; survivors keep their order. */
__device__ __forceinline__ void d_readout(EvList &L, int n, int depth, int policy, int pann, int crystalN,
                                          bool (&alive)[GPET_MAX_PH_EVENTS], int (&site)[GPET_MAX_PH_EVENTS])
{
#pragma unroll
    for (int i = 0; i < GPET_MAX_PH_EVENTS; ++i) {
        alive[i] = i < n;
        site[i] = L.cs[i];
    }
    if (depth == 3) return;
    if (policy == 1) depth = 2;
    /* panel * moduleN + module = cs / crystalN */
#pragma unroll
    for (int i = 0; i < GPET_MAX_PH_EVENTS; ++i)
        site[i] = (depth == 0) ? 0 : (depth == 1 ? pann : L.cs[i] / crystalN);
#pragma unroll
    for (int i = 0; i < GPET_MAX_PH_EVENTS; ++i) {
        if (alive[i]) {
#pragma unroll
            for (int j = i + 1; j < GPET_MAX_PH_EVENTS; ++j) {
                if (alive[j] && site[j] == site[i]) {
                    if (policy == 1) {
                        float sE = L.E[i] + L.E[j];
                        L.x[i] = (L.x[i] * L.E[i] + L.x[j] * L.E[j]) / sE;
                        L.y[i] = (L.y[i] * L.E[i] + L.y[j] * L.E[j]) / sE;
                        L.z[i] = (L.z[i] * L.E[i] + L.z[j] * L.E[j]) / sE;
                        L.E[i] = sE;
                    } else if (!(L.E[i] > L.E[j])) {
                        L.cs[i] = L.cs[j]; L.t[i] = L.t[j];
                        L.E[i] = L.E[j]; L.x[i] = L.x[j]; L.y[i] = L.y[j]; L.z[i] = L.z[j];
                    }
                    alive[j] = false;
                }
            }
        }
    }
}

/* events staged in LDS per panel-kernel block (48 B each) */
#define EV_STAGE 512u
/* detector majorant in LDS for the panel kernel (nodes; 8.7 KB).  Off by
 * default: the body workload's panel kernel ran 5% slower with it (1.65 ->
 * 1.74 ms per launch; the cylinder's unchanged), the extra LDS leaving
 * fewer blocks per CU for no saving on L1-resident loads. */
#ifndef GPET_PANEL_LDS_WCK
#define GPET_PANEL_LDS_WCK 0
#endif
#define WCK_LDS_NODES 2176

/* MODE: PANEL_PLAIN transports every queued photon; PANEL_SURROGATE lists
 * the photons inside the network's validity range for the surrogate and
 * transports the rest; PANEL_HARVEST transports every photon and also writes
 * a training record (features, first-deposit targets) per in-range photon. */
/* register budget of the default panel-kernel variant (waves per SIMD the
 * compiler must allow).  The kernel needs 134 VGPRs (3 waves per SIMD); a
 * budget of 4 waves (127 VGPRs, 24 spilled to scratch) ran no faster on
 * either workload, so the compiler's own allocation is kept. */
#ifndef GPET_PANEL_WPE_DEFAULT
#define GPET_PANEL_WPE_DEFAULT 1
#endif
template <int WPE, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : GPET_PANEL_WPE_DEFAULT)))
void k_panel(PanelParams pp, DevTables tab, DevDetector det, DevQueue q)
{
    unsigned long long c_hits = 0, c_drop = 0, c_droph = 0, c_maj = 0, c_steps = 0;
    /* the block's events are staged in LDS and written out together at the
     * end with one reservation in the global event array (many waves adding
     * to one global counter serialize); a wave whose events do not fit
     * writes the rest straight to the array */
    __shared__ gpet_event s_ev[EV_STAGE];
    __shared__ unsigned s_evn, s_evbase;
    if (threadIdx.x == 0) s_evn = 0u;
#if GPET_PANEL_LDS_WCK
    /* the detector majorant's nodes up to ~530 keV (every photon that
     * enters a panel is below) in LDS */
    __shared__ float s_lam[WCK_LDS_NODES];
    for (int t = threadIdx.x; t < WCK_LDS_NODES; t += blockDim.x) s_lam[t] = det.wck.lam[t];
#endif
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    DevPanel *s_p = reinterpret_cast<DevPanel *>(smem);
    unsigned *s_pre = reinterpret_cast<unsigned *>(smem + sizeof(DevPanel) * (size_t)det.npanels);
    __shared__ unsigned s_part[256];
    for (int t = threadIdx.x; t < det.npanels; t += blockDim.x) s_p[t] = det.p[t];
    /* prefix over the phantom kernel's region counts: s_pre[r] = entries
     * before region r, s_pre[nregions] = total */
    const int G = q.nregions;
    const int chunk = (G + (int)blockDim.x - 1) / (int)blockDim.x;
    unsigned part = 0;
    for (int r = threadIdx.x * chunk; r < min(G, (int)(threadIdx.x + 1) * chunk); ++r) part += q.rcount[r];
    s_part[threadIdx.x] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const unsigned v = s_part[t];
            s_part[t] = acc;
            acc += v;
        }
    }
    __syncthreads();
    {
        unsigned acc = s_part[threadIdx.x];
        for (int r = threadIdx.x * chunk; r < min(G, (int)(threadIdx.x + 1) * chunk); ++r) {
            s_pre[r] = acc;
            acc += q.rcount[r];
        }
        if ((int)threadIdx.x == (int)blockDim.x - 1) s_pre[G] = acc;
    }
    __syncthreads();
    const int nq = (int)s_pre[G];
    const int stride = gridDim.x * blockDim.x;
    /* queue item of global index gi: the largest region r with s_pre[r] <= gi */
    auto item_of = [&](int gi) -> int64_t {
        int lo = 0, hi = G;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((int)s_pre[mid] <= gi) lo = mid; else hi = mid;
        }
        return (int64_t)lo * q.region + (gi - (int)s_pre[lo]);
    };
    if (MODE == PANEL_SURROGATE) {
        /* the in-range photons go to the network's queue with one global
         * reservation per block (one per wave on a single counter took
         * most of this kernel's time): count the block's photons, reserve,
         * then place them in passes whose slots come from a scan of the
         * waves' counts (every thread of the block makes the same passes) */
        __shared__ unsigned s_wc[4], s_sqbase;
        unsigned cnt = 0;
        for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < nq; gi += stride) {
            const int64_t i = item_of(gi);
            cnt += surr_valid(q.E[i], q.vx[i]) ? 1u : 0u;
        }
        s_part[threadIdx.x] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned tot = 0;
            for (int t = 0; t < (int)blockDim.x; ++t) tot += s_part[t];
            s_sqbase = tot ? (unsigned)atomicAdd(pp.sq.count, (int)tot) : 0u;
        }
        __syncthreads();
        const int wv = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
        unsigned off = s_sqbase;
        for (int b0 = blockIdx.x * blockDim.x; b0 < nq; b0 += stride) {
            const int gi = b0 + (int)threadIdx.x;
            int64_t i = 0;
            bool sv = false;
            if (gi < nq) {
                i = item_of(gi);
                sv = surr_valid(q.E[i], q.vx[i]);
            }
            const unsigned long long m = __ballot(sv);
            if (__lane_id() == 0) s_wc[wv] = (unsigned)__popcll(m);
            __syncthreads();
            unsigned before = off, pass = 0;
            for (int w = 0; w < nw; ++w) {
                before += w < wv ? s_wc[w] : 0u;
                pass += s_wc[w];
            }
            const unsigned slot = before + (unsigned)__popcll(m & lanemask_lt());
            if (sv && (int64_t)slot < (int64_t)pp.sq.cap) {
                pp.sq.y[slot] = q.y[i]; pp.sq.z[slot] = q.z[i];
                pp.sq.vx[slot] = q.vx[i]; pp.sq.vy[slot] = q.vy[i]; pp.sq.vz[slot] = q.vz[i];
                pp.sq.E[slot] = q.E[i];
                pp.sq.t[slot] = q.t[i];
                pp.sq.info[slot] = q.info[i];
            }
            off += pass;
            __syncthreads();
        }
    }
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < nq; gi += stride) {
        const int64_t i = item_of(gi);
        /* in-range photons were handed to the network above */
        if (MODE == PANEL_SURROGATE && surr_valid(q.E[i], q.vx[i])) continue;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        bool got = false;
        float x2 = 0.0f, y2 = q.y[i], z2 = q.z[i];
        float vx2 = q.vx[i], vy2 = q.vy[i], vz2 = q.vz[i];
        float E = q.E[i];
        double tof = q.t[i];
        int4 info = q.info[i];
        const int pa = info.x, parn = info.y, eid = info.z;
        const unsigned draws = (unsigned)info.w & 0x7fffffffu;
        const unsigned k = (unsigned)info.w >> 31;
        gp_rng r;
        gp_rng_init(&r, pp.seed, (uint64_t)(int64_t)eid, GP_TAG_PHOTON + k, draws);
        const DevPanel &P = s_p[pa];
        const int pann = P.panel;
        const float hx = P.lenx, hy = P.leny * 0.5f, hz = P.lenz * 0.5f;
        EvList L;
        int nev = 0;
        int nsteps = 0, nint = 0;
        for (int it = 0; it < D_STEP_CAP; ++it) {
            /* block per step: length, selector, Rayleigh CDF, azimuth; the KN
             * rejection loop draws after the block */
            gp_u4 b = gp_rng_block(&r);
#if GPET_PANEL_LDS_WCK
            float lammin = d_lam_lds(s_lam, WCK_LDS_NODES, det.wck, E);
#else
            float lammin = d_lam(det.wck, E);
#endif
            float s = -lammin * gp_logf(gp_u01(b.x));
            x2 = x2 + s * vx2;
            y2 = y2 + s * vy2;
            z2 = z2 + s * vz2;
            tof = tof + (double)s * D_INV_C;
            if (fabsf(y2) > hy || fabsf(z2) > hz || x2 * P.dirx < 0.0f || x2 * P.dirx > hx) break;
            nsteps++;
            int M_id = -1, L_id = -1;
            int m_id = d_crystal(det, P, x2, y2, z2, M_id, L_id);
            float den = det.dens[m_id];
            int mat = det.mat[m_id];
            float3 mu = d_mu3(tab, mat, E);
            float lamden = lammin * den;
            float prob = 1.0f - lamden * mu.x;
            if (prob < 0.0f) { prob = 0.0f; c_maj++; }
            float rn = gp_u01(b.y);
            if (rn < prob) continue;
            nint++;
            prob = prob + lamden * mu.y;
            int type = 0;
            float edep = 0.0f;
            bool stop = false, rot = false;
            float cth = 0.0f, uphi = 0.0f;
            float de2 = 0.0f;
            bool second = false;
            if (rn < prob) {
                float efrac;
                d_kn(E, r, efrac, cth);
                float de = E * (1.0f - efrac);
                uphi = gp_u01(b.w);
                type = 1;
                edep = de;
                E = E - de;
                if (E < pp.eabs) {
                    second = true;
                    de2 = E;
                    stop = true;
                } else {
                    rot = true;
                }
            } else {
                prob = prob + lamden * mu.z;
                if (rn < prob) {
                    cth = d_surf(tab, GPET_SURF_RAYLEIGH, mat, E, gp_u01(b.z));
                    uphi = gp_u01(b.w);
                    type = pp.record_rayleigh ? 3 : 0;
                    rot = true;
                } else {
                    type = 4;
                    edep = E;
                    stop = true;
                }
            }
            if (m_id == 0 && type != 0) {
                const int site = pann * det.moduleN * det.crystalN + M_id * det.crystalN + L_id;
#pragma unroll
                for (int rec = 0; rec < 2; ++rec) {
                    if (rec == 1 && !second) break;
                    const int ty = rec ? 2 : type;
                    const float ed = rec ? de2 : (type == 3 ? 0.0f : edep);
                    c_hits++;
                    if (pp.hit_id) {
                        const unsigned hidx = wave_append1(pp.n_hits);
                        if ((int64_t)hidx < pp.hits_cap) {
                            int32_t *hi = pp.hit_id + 5 * (size_t)hidx;
                            float *hf = pp.hit_f + 5 * (size_t)hidx;
                            hi[0] = parn; hi[1] = pann; hi[2] = M_id; hi[3] = L_id; hi[4] = ty;
                            hf[0] = ed; hf[1] = (float)tof; hf[2] = x2; hf[3] = y2; hf[4] = z2;
                        } else {
                            c_droph++;
                        }
                    }
                    if (ty != 3) {
                        if (MODE == PANEL_HARVEST && !got) {
                            fx = x2; fy = y2; fz = z2;
                            got = true;
                        }
                        d_adder(L, nev, site, tof, ed, x2, y2, z2, c_drop);
                    }
                }
            }
            if (stop) break;
            if (rot) d_rotate(vx2, vy2, vz2, cth, uphi);
        }
        bool alive[GPET_MAX_PH_EVENTS];
        int rsite[GPET_MAX_PH_EVENTS];
        d_readout(L, nev, pp.readout_depth, pp.readout_policy, pann, det.crystalN, alive, rsite);
        int nout = 0;
#pragma unroll
        for (int j = 0; j < GPET_MAX_PH_EVENTS; ++j) nout += alive[j] ? 1 : 0;
        /* wave-aggregated append of nout (0..4) events per lane */
        unsigned long long lt = lanemask_lt();
        int off = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < GPET_MAX_PH_EVENTS; ++j) {
            unsigned long long b = __ballot(nout > j);
            off += (int)__popcll(b & lt);
            tot += (int)__popcll(b);
        }
        /* leader = first active lane (lanes that handed their photon to
         * the surrogate have left this iteration); the wave's events take
         * LDS slots [lb, lb + nl) and, past the stage, global slots from gb */
        unsigned lb = 0, nl = 0, gb = 0;
        if (tot) {
            const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
            if ((int)__lane_id() == leader) {
                lb = atomicAdd(&s_evn, (unsigned)tot);
                nl = lb < EV_STAGE ? min((unsigned)tot, EV_STAGE - lb) : 0u;
                if (nl < (unsigned)tot) gb = atomicAdd(pp.n_events, (unsigned)tot - nl);
            }
            lb = __shfl(lb, leader);
            nl = __shfl(nl, leader);
            gb = __shfl(gb, leader);
        }
        int o = 0;
#pragma unroll
        for (int j = 0; j < GPET_MAX_PH_EVENTS; ++j) {
            if (alive[j]) {
                gpet_event ev;
                const int pm = L.cs[j] / det.crystalN;   /* panel * moduleN + module */
                ev.parn = parn; ev.pann = pann; ev.modn = pm - pann * det.moduleN; ev.cryn = L.cs[j] - pm * det.crystalN;
                ev.siten = rsite[j]; ev.eventid = eid; ev.t = L.t[j];
                ev.E = L.E[j]; ev.x = L.x[j]; ev.y = L.y[j]; ev.z = L.z[j];
                const unsigned e = (unsigned)(off + o);
                if (e < nl) {
                    s_ev[lb + e] = ev;
                } else {
                    const int64_t dst = (int64_t)gb + (e - nl);
                    if (dst < pp.events_cap) pp.events[dst] = ev;
                }
                o++;
            }
        }
        if (MODE == PANEL_HARVEST) {
            const float E0 = q.E[i], vx0 = q.vx[i], vy0 = q.vy[i], vz0 = q.vz[i], y0 = q.y[i], z0 = q.z[i];
            const bool hv = surr_valid(E0, vx0);
            unsigned long long m = __ballot(hv);
            unsigned long long base = 0;
            if (m) {
                int leader = __ffsll((unsigned long long)m) - 1;
                if ((int)__lane_id() == leader) base = atomicAdd(pp.harvest_count, (unsigned long long)__popcll(m));
                base = __shfl(base, leader);
            }
            if (hv) {
                const unsigned long long dst = base + __popcll(m & lanemask_lt());
                if ((int64_t)dst < pp.harvest_cap) {
                    float *rec = pp.harvest + dst * SURR_REC;
                    surr_features(P, E0, vx0, vy0, vz0, y0, z0, rec);
                    int j0 = -1;
#pragma unroll
                    for (int j = GPET_MAX_PH_EVENTS - 1; j >= 0; --j)
                        if (alive[j]) j0 = j;
                    float dep = 0.0f, tdy = 0.0f, tdz = 0.0f, fe = 0.0f;
                    if (j0 >= 0 && got) {
#pragma unroll
                        for (int j = 0; j < GPET_MAX_PH_EVENTS; ++j)
                            if (j == j0) fe = L.E[j] / E0;
                        surr_targets(P, vx0, vy0, vz0, y0, z0, fx, fy, fz, dep, tdy, tdz);
                    }
                    rec[6] = (j0 >= 0 && got) ? 1.0f : 0.0f;
                    rec[7] = fe;
                    rec[8] = dep;
                    rec[9] = tdy;
                    rec[10] = tdz;
                    rec[11] = __int_as_float(parn);   /* sort key: records come out in photon order */
                }
            }
        }
        c_steps += (unsigned long long)nsteps;
        if (q.trace_slot) {
            int ts = q.trace_slot[i];
            if (ts >= 0) {
                gpet_trace *tr = pp.trace + ts;
                tr->n_steps_panel = nsteps;
                tr->n_int_panel = nint;
                tr->n_events = nout;
                tr->draws = r.n;
            }
        }
    }
    /* write out the staged events: one reservation for the block, then a
     * coalesced copy (12 dwords per event) */
    __syncthreads();
    const unsigned ns = min(s_evn, EV_STAGE);
    if (threadIdx.x == 0) s_evbase = ns ? atomicAdd(pp.n_events, ns) : 0u;
    __syncthreads();
    {
        const int64_t base = (int64_t)s_evbase;
        const int64_t nfit = max<int64_t>(0, min<int64_t>((int64_t)ns, pp.events_cap - base));
        const unsigned *src = reinterpret_cast<const unsigned *>(s_ev);
        unsigned *dstw = reinterpret_cast<unsigned *>(pp.events + base);
        const int nw = (int)nfit * (int)(sizeof(gpet_event) / 4);
        for (int k = threadIdx.x; k < nw; k += blockDim.x) dstw[k] = src[k];
    }
    {
        unsigned long long *const dst[5] = {
            pp.cnt + CNT_HITS * CNT_STRIPES, pp.cnt + CNT_DROPPED_EVENTS * CNT_STRIPES,
            pp.cnt + CNT_DROPPED_HITS * CNT_STRIPES, pp.cnt + CNT_MAJORANT * CNT_STRIPES,
            pp.cnt + CNT_PANEL_STEPS * CNT_STRIPES};
        const unsigned long long v[5] = {c_hits, c_drop, c_droph, c_maj, c_steps};
        flush_counters<5>(dst, v);
    }
}

} // namespace

const int gpet_wpe_variants[GPET_N_WPE] = {0, 5, 6};

namespace {

int grid_of(int64_t n, const LaunchCfg &cfg, int n_cu)
{
    int64_t nb = (n + cfg.block - 1) / cfg.block;
    int64_t cap = cfg.blocks_per_cu > 0 ? (int64_t)n_cu * cfg.blocks_per_cu : 1048576;
    return (int)std::max<int64_t>(1, std::min(nb, cap));
}

template <int MODE, bool BATCH, bool TQ>
void launch_phantom_wpe(const RunParams &rp, const DevTables &tab, const DevPhantom &ph, const DevDetector &det,
                        const DevQueue &q, int grid, int block, int wpe, hipStream_t s)
{
    const size_t lds = gpet_panel_lds_bytes(det);
    if (wpe == 5)
        hipLaunchKernelGGL((k_phantom<MODE, BATCH, 5, TQ>), dim3(grid), dim3(block), lds, s, rp, tab, ph, det, q);
    else if (wpe == 6)
        hipLaunchKernelGGL((k_phantom<MODE, BATCH, 6, TQ>), dim3(grid), dim3(block), lds, s, rp, tab, ph, det, q);
    else
        hipLaunchKernelGGL((k_phantom<MODE, BATCH, 0, TQ>), dim3(grid), dim3(block), lds, s, rp, tab, ph, det, q);
}

__global__ void k_zero(unsigned *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

} // namespace

int gpet_launch_zero(unsigned *p, int n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return (int)hipGetLastError();
}

int gpet_phantom_grid(int64_t items, const LaunchCfg &cfg, int n_cu) { return grid_of(items, cfg, n_cu); }

int gpet_launch_phantom(const RunParams &rp, const DevTables &tab, const DevPhantom &ph, const DevDetector &det,
                        const DevQueue &q, bool batch, const LaunchCfg &cfg, int n_cu, hipStream_t s)
{
    if (rp.nh <= 0) return 0;
    if (cfg.block <= 0 || cfg.block > 256 || cfg.block % 64) return (int)hipErrorInvalidValue;
    const int grid = grid_of(batch ? rp.n_vblocks * cfg.block : rp.nh, cfg, n_cu);
    /* photon regrouping: source runs only (the phase-space modes finish
     * their photons in the phantom kernel) */
    const bool tq = q.tq && rp.mode == RUN_DECAY && (batch || ph.present);
    if (rp.mode == RUN_DECAY) {
        const size_t lds = gpet_panel_lds_bytes(det);
        if (tq) {
            if (batch) launch_phantom_wpe<RUN_DECAY, true, true>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
            else launch_phantom_wpe<RUN_DECAY, false, true>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
        } else {
            if (batch) launch_phantom_wpe<RUN_DECAY, true, false>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
            else launch_phantom_wpe<RUN_DECAY, false, false>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
        }
        if (rp.prange && q.lw) {
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return (int)e;
            if (batch && tq)
                hipLaunchKernelGGL((k_longwalk<true, true>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det, q);
            else if (batch)
                hipLaunchKernelGGL((k_longwalk<true, false>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det,
                                   q);
            else if (tq)
                hipLaunchKernelGGL((k_longwalk<false, true>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det,
                                   q);
            else
                hipLaunchKernelGGL((k_longwalk<false, false>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det,
                                   q);
        }
    } else if (rp.mode == RUN_PSF_POSITRON) {
        launch_phantom_wpe<RUN_PSF_POSITRON, false, false>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
    } else {
        launch_phantom_wpe<RUN_PSF_PHOTON, false, false>(rp, tab, ph, det, q, grid, cfg.block, cfg.wpe, s);
    }
    if (tq) {
        /* transit rounds: alternate between the two buffers; the last
         * launched round (or the only one) finishes every photon */
        int nr = 0;
        while (nr < 3 && cfg.rounds[nr] > 0) ++nr;
        if (nr == 0) nr = 1;
        const size_t lds = gpet_panel_lds_bytes(det);
        const bool b = batch;
        for (int i = 0; i < nr; ++i) {
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return (int)e;
            const bool last = i == nr - 1;
            if (!last && !q.tq2) return (int)hipErrorInvalidValue;
            const TransitItem *in = (i & 1) ? q.tq2 : q.tq;
            const unsigned *inc = (i & 1) ? q.tq2_count : q.tq_count;
            TransitItem *out = last ? nullptr : ((i & 1) ? q.tq : q.tq2);
            unsigned *outc = last ? nullptr : ((i & 1) ? q.tq_count : q.tq2_count);
            const int steps = cfg.rounds[i] > 0 ? cfg.rounds[i] : D_STEP_CAP;
            if (b)
                hipLaunchKernelGGL((k_transit<true>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det, q, in,
                                   inc, out, outc, steps);
            else
                hipLaunchKernelGGL((k_transit<false>), dim3(grid), dim3(cfg.block), lds, s, rp, tab, ph, det, q, in,
                                   inc, out, outc, steps);
        }
    }
    return (int)hipGetLastError();
}

int gpet_launch_panel(const PanelParams &pp, const DevTables &tab, const DevDetector &det, const DevQueue &q,
                      int64_t max_items, const LaunchCfg &cfg, int n_cu, hipStream_t s, int mode)
{
    if (max_items <= 0) return 0;
    if (cfg.block <= 0 || cfg.block > 256 || cfg.block % 64) return (int)hipErrorInvalidValue;
    const int grid = grid_of(max_items, cfg, n_cu);
    const size_t lds = sizeof(DevPanel) * (size_t)det.npanels + sizeof(unsigned) * (size_t)(q.nregions + 1);
    if (mode == PANEL_HARVEST)
        hipLaunchKernelGGL((k_panel<0, PANEL_HARVEST>), dim3(grid), dim3(cfg.block), lds, s, pp, tab, det, q);
    else if (mode == PANEL_SURROGATE)
        hipLaunchKernelGGL((k_panel<0, PANEL_SURROGATE>), dim3(grid), dim3(cfg.block), lds, s, pp, tab, det, q);
    else if (cfg.wpe == 5)
        hipLaunchKernelGGL((k_panel<5, PANEL_PLAIN>), dim3(grid), dim3(cfg.block), lds, s, pp, tab, det, q);
    else if (cfg.wpe == 6)
        hipLaunchKernelGGL((k_panel<6, PANEL_PLAIN>), dim3(grid), dim3(cfg.block), lds, s, pp, tab, det, q);
    else
        hipLaunchKernelGGL((k_panel<0, PANEL_PLAIN>), dim3(grid), dim3(cfg.block), lds, s, pp, tab, det, q);
    return (int)hipGetLastError();
}
