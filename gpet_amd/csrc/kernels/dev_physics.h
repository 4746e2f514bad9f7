/*
 * dev_physics.h - device-side physics of the transport kernels.
 *
 * Each helper restates one reference routine (cited) in explicit FP32 with a
 * fixed operation order; with -ffp-contract=off and correctly rounded
 * division/sqrt the results are bit-identical to the host oracle built from
 * the same rules.
 */
#ifndef GPET_DEV_PHYSICS_H
#define GPET_DEV_PHYSICS_H

#include <hip/hip_runtime.h>
#include "../common/gp_math.h"
#include "../host/internal.h"

#define D_C_LIGHT 29979.2458       /* cm/us, gPET_kernals.cu:264 */
#define D_INV_C (1.0 / 29979.2458)
#define D_MC2F 510999.1f           /* constants.h:34 */
#define D_IMC2F 1.95695061e-06f    /* constants.h:35 */
#define D_SZERO 1.0e-4f
#define D_ZEROF 1.0e-20f
#define D_STEP_CAP 1000000

/* Linear interpolation on a uniform grid (replaces 1-D texture filtering,
 * gPET_kernals.cu:31-64). */
__device__ __forceinline__ float d_lerp(const float *__restrict__ T, int n, float e0, float idl, float E)
{
    float x = (E - e0) * idl;
    x = fminf(fmaxf(x, 0.0f), (float)(n - 1));
    int j = (int)x;
    if (j > n - 2) j = n - 2;
    float f = x - (float)j;
    float a = T[j], b = T[j + 1];
    return a + f * (b - a);
}

__device__ __forceinline__ float d_lam(const DevWck &w, float E)
{
    return d_lerp(w.lam, GPET_NWCK, w.e0, w.idl, E);
}

/* d_lam with the first n_lds nodes of the majorant copied to LDS (s_lam):
 * the same index and arithmetic, the nodes read from LDS when both lie
 * there. */
__device__ __forceinline__ float d_lam_lds(const float *s_lam, int n_lds, const DevWck &w, float E)
{
    float x = (E - w.e0) * w.idl;
    x = fminf(fmaxf(x, 0.0f), (float)(GPET_NWCK - 1));
    int j = (int)x;
    if (j > GPET_NWCK - 2) j = GPET_NWCK - 2;
    float f = x - (float)j;
    const bool in = j + 1 < n_lds;
    float a = in ? s_lam[j] : w.lam[j], b = in ? s_lam[j + 1] : w.lam[j + 1];
    return a + f * (b - a);
}

/* mu/rho of (total, Compton, Rayleigh) for one material at E: two 16-byte
 * nodes, three lerps (same arithmetic as three separate table reads). */
__device__ __forceinline__ float3 d_mu3(const DevTables &t, int mat, float E)
{
    float x = (E - t.e0) * t.idl;
    x = fminf(fmaxf(x, 0.0f), (float)(t.n1d - 1));
    int j = (int)x;
    if (j > t.n1d - 2) j = t.n1d - 2;
    float f = x - (float)j;
    const float4 *p = t.mu4 + (size_t)mat * t.n1d + j;
    float4 a = p[0], b = p[1];
    return make_float3(a.x + f * (b.x - a.x), a.y + f * (b.y - a.y), a.z + f * (b.z - a.z));
}

/* Bilinear inverse-CDF surface read (gPET_kernals.cu:66-88,128-146). */
__device__ __forceinline__ float d_surf(const DevTables &t, int k, int mat, float E, float u)
{
    const int ne = t.ne, ncp = t.ncp;
    float x = E * t.ide[k];
    x = fminf(fmaxf(x, 0.0f), (float)(ne - 1));
    float y = u * t.idcp[k];
    y = fminf(fmaxf(y, 0.0f), (float)(ncp - 1));
    int ix = (int)x;
    if (ix > ne - 2) ix = ne - 2;
    int iy = (int)y;
    if (iy > ncp - 2) iy = ncp - 2;
    float fx = x - (float)ix, fy = y - (float)iy;
    const float *S = t.surf[k] + (size_t)mat * ncp * ne;
    float v00 = S[iy * ne + ix], v01 = S[iy * ne + ix + 1];
    float v10 = S[(iy + 1) * ne + ix], v11 = S[(iy + 1) * ne + ix + 1];
    float a = v00 + fx * (v01 - v00);
    float b = v10 + fx * (v11 - v10);
    float v = a + fy * (b - a);
    return fminf(fmaxf(v, -1.0f), 1.0f);
}

/* getAbsVox (gPET_kernals.cu:19-29) with voxel 0 inside (Q1 fixed). */
__device__ __forceinline__ bool d_vox(const DevPhantom &p, float x, float y, float z, int &ix, int &iy, int &iz)
{
    float fx = (x - p.offx) * p.idx;
    float fy = (y - p.offy) * p.idy;
    float fz = (z - p.offz) * p.idz;
    if (!(fx >= 0.0f && fx < (float)p.nx && fy >= 0.0f && fy < (float)p.ny && fz >= 0.0f && fz < (float)p.nz))
        return false;
    ix = (int)fx;
    iy = (int)fy;
    iz = (int)fz;
    return true;
}

/* rotate (gPET_kernals.cu:172-254), azimuth given as u with phi = 2 pi u. */
__device__ __forceinline__ void d_rotate(float &u, float &v, float &w, float costh, float uphi)
{
    float rho2 = u * u + v * v;
    float norm = rho2 + w * w;
    if (fabsf(norm - 1.0f) > D_SZERO) {
        norm = 1.0f / sqrtf(norm);
        u = u * norm;
        v = v * norm;
        w = w * norm;
        rho2 = u * u + v * v;
    }
    float sinphi, cosphi;
    gp_sincos2pi(uphi, &sinphi, &cosphi);
    float temp = costh * costh;
    if (rho2 > D_ZEROF) {
        float sthrho = (temp < 1.0f) ? sqrtf((1.0f - temp) / rho2) : 0.0f;
        float urho = u * sthrho;
        float vrho = v * sthrho;
        float nu = u * costh - vrho * sinphi + (w * urho) * cosphi;
        float nv = v * costh + urho * sinphi + (w * vrho) * cosphi;
        float nw = w * costh - (rho2 * sthrho) * cosphi;
        u = nu;
        v = nv;
        w = nw;
    } else {
        float sinth = (temp < 1.0f) ? sqrtf(1.0f - temp) : 0.0f;
        v = sinth * sinphi;
        if (w > 0.0f) {
            u = sinth * cosphi;
            w = costh;
        } else {
            u = -sinth * cosphi;
            w = -costh;
        }
    }
}

/* rotate with sin(theta) given (pair noncollinearity: avoids the FP32
 * collapse of sqrt(1 - cos^2) for small angles). */
__device__ __forceinline__ void d_rotate_cs(float &u, float &v, float &w, float costh, float sinth, float uphi)
{
    float rho2 = u * u + v * v;
    float norm = rho2 + w * w;
    if (fabsf(norm - 1.0f) > D_SZERO) {
        norm = 1.0f / sqrtf(norm);
        u = u * norm;
        v = v * norm;
        w = w * norm;
        rho2 = u * u + v * v;
    }
    float sinphi, cosphi;
    gp_sincos2pi(uphi, &sinphi, &cosphi);
    if (rho2 > D_ZEROF) {
        float sthrho = sinth / sqrtf(rho2);
        float urho = u * sthrho;
        float vrho = v * sthrho;
        float nu = u * costh - vrho * sinphi + (w * urho) * cosphi;
        float nv = v * costh + urho * sinphi + (w * vrho) * cosphi;
        float nw = w * costh - (rho2 * sthrho) * cosphi;
        u = nu;
        v = nv;
        w = nw;
    } else {
        v = sinth * sinphi;
        if (w > 0.0f) {
            u = sinth * cosphi;
            w = costh;
        } else {
            u = -sinth * cosphi;
            w = -costh;
        }
    }
}

/* Klein-Nishina rejection sampling (gPET_kernals.cu:90-126). */
__device__ __forceinline__ void d_kn(float E, gp_rng &r, float &efrac, float &costhe)
{
    float e0 = E * D_IMC2F;
    float twoe = 2.0f * e0;
    float kmin2 = 1.0f / ((1.0f + twoe) * (1.0f + twoe));
    float loge = gp_logf(1.0f + twoe);
    float a1 = loge + (twoe * (1.0f + e0)) * kmin2;
    efrac = 1.0f;
    for (int it = 0; it < D_STEP_CAP; ++it) {
        float u1 = gp_rng_uniform(&r);
        float u2 = gp_rng_uniform(&r);
        if (u1 * a1 < loge)
            efrac = gp_expf(-u2 * loge);
        else
            efrac = sqrtf(kmin2 + u2 * (1.0f - kmin2));
        float mess = ((e0 * e0) * efrac) * (1.0f + efrac * efrac);
        float u3 = gp_rng_uniform(&r);
        if (u3 * mess <= mess - (1.0f - efrac) * ((1.0f + twoe) * efrac - 1.0f)) break;
    }
    costhe = 1.0f - (1.0f - efrac) / (efrac * e0);
}

/* crystalSearch (gPET_kernals.cu:1236-1279): 0 crystal, 1 gap/excluded. */
__device__ __forceinline__ int d_crystal(const DevDetector &d, const DevPanel &P, float x, float y, float z,
                                         int &M_id, int &L_id)
{
    for (int q = 0; q < d.nquad; ++q) {
        const float *s = &d.quad[q * 10];
        float v = s[0] * x * x + s[1] * y * y + s[2] * z * z + s[3] * x * y + s[4] * x * z + s[5] * y * z +
                  s[6] * x + s[7] * y + s[8] * z + s[9];
        if (v < 0.0f) return 1;
    }
    float yy = P.leny / 2.0f + y;
    float zz = P.lenz / 2.0f + z;
    float mpy = P.mody + P.mspy, mpz = P.modz + P.mspz;
    int my = floorf(yy / mpy) > 0.0f ? (int)(yy / mpy) : 0;
    int mz = floorf(zz / mpz) > 0.0f ? (int)(zz / mpz) : 0;
    M_id = mz * d.moduleNy + my;
    yy = yy - (float)my * mpy;
    zz = zz - (float)mz * mpz;
    if (yy > P.mody || zz > P.modz) return 1;
    float cpy = P.cry + P.cspy, cpz = P.crz + P.cspz;
    int ly = floorf(yy / cpy) > 0.0f ? (int)(yy / cpy) : 0;
    int lz = floorf(zz / cpz) > 0.0f ? (int)(zz / cpz) : 0;
    L_id = lz * d.crystalNy + ly;
    yy = yy - (float)ly * cpy;
    zz = zz - (float)lz * cpz;
    if (yy > P.cry || zz > P.crz) return 1;
    return 0;
}

#endif
