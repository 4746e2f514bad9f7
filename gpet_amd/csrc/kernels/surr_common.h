/*
 * surr_common.h - what the panel kernel (harvest, gate) and the surrogate
 * kernels share: sizes of the network's inputs and outputs, the validity
 * range, the input features and the mapping between a photon's first
 * deposit and the network's targets.
 */
#ifndef GPET_SURR_COMMON_H
#define GPET_SURR_COMMON_H

#include "dev_physics.h"

#define SURR_NX 6          /* features: E/511 keV, |vx|, vy, vz, y/(leny/2), z/(lenz/2) */
#define SURR_LDX 8         /* input row: 6 features, the bias 1, a zero             */
#define SURR_NY 5          /* targets: event, fE, depth/lenx, dy, dz (pitches)       */
#define SURR_REC 12        /* harvested record: features, targets, padding          */
#define SURR_PP_MIN 0.999f /* a deposited-energy fraction at or above this is a
                              full absorption (photopeak)                          */
/* Head outputs (190; the head is stored with 192 rows):
 *   [0]                 event logit
 *   [SURR_CLS + c]      logits of the 61 energy classes: c = 0 full
 *                       absorption, c = 1..48 the fraction below the Compton
 *                       edge fE_edge(E) in 48 equal bins of fE / fE_edge,
 *                       c = 49..60 the fraction between the edge and
 *                       SURR_PP_MIN in 12 equal bins
 *   [SURR_WPP + k]      logits of the 16 position components for a full
 *                       absorption, [SURR_WCT + k] for a continuum fraction
 *   [SURR_MU + 3k + d]  mean of component k, target d (depth, dy, dz)
 *   [SURR_LS + 3k + d]  its log sigma, clamped to the range below */
#define SURR_NLO 48
#define SURR_NHI 12
#define SURR_NC (1 + SURR_NLO + SURR_NHI)
#define SURR_K 16          /* position mixture components                          */
#define SURR_D 3           /* position mixture dimensions                          */
#define SURR_CLS 1
#define SURR_WPP (SURR_CLS + SURR_NC)
#define SURR_WCT (SURR_WPP + SURR_K)
#define SURR_MU (SURR_WCT + SURR_K)
#define SURR_LS (SURR_MU + SURR_D * SURR_K)
#define SURR_NH (SURR_LS + SURR_D * SURR_K)
#define SURR_LOGSIG_MIN (-5.0f)
#define SURR_LOGSIG_MAX (2.0f)
#define SURR_E_MIN 100.0e3f /* validity: entry energy (eV) ...                     */
#define SURR_E_MAX 520.0e3f
#define SURR_COS_MIN 0.5f   /* ... and incidence up to 60 degrees                  */

__host__ __device__ __forceinline__ bool surr_valid(float E, float vx)
{
    return E >= SURR_E_MIN && E <= SURR_E_MAX && fabsf(vx) >= SURR_COS_MIN;
}

__host__ __device__ __forceinline__ void surr_features(const DevPanel &P, float E, float vx, float vy, float vz,
                                                       float y, float z, float f[SURR_NX])
{
    f[0] = E * (1.0f / 511.0e3f);
    f[1] = fabsf(vx);
    f[2] = vy;
    f[3] = vz;
    f[4] = y / (0.5f * P.leny);
    f[5] = z / (0.5f * P.lenz);
}

/* Compton edge of an entry of e511 = E / 511 keV as a fraction of E: the
 * largest energy one Compton scatter deposits, 2a / (1 + 2a). */
__host__ __device__ __forceinline__ float surr_fe_edge(float e511)
{
    const float a2 = 2.0f * e511;
    return a2 / (1.0f + a2);
}

/* Energy class of a deposited fraction fe for an entry of e511. */
__host__ __device__ __forceinline__ int surr_class(float e511, float fe)
{
    if (fe >= SURR_PP_MIN) return 0;
    const float ed = surr_fe_edge(e511);
    if (fe < ed) {
        const int b = (int)(fe / ed * (float)SURR_NLO);
        return 1 + (b < 0 ? 0 : (b > SURR_NLO - 1 ? SURR_NLO - 1 : b));
    }
    const int b = (int)((fe - ed) / (SURR_PP_MIN - ed) * (float)SURR_NHI);
    return 1 + SURR_NLO + (b < 0 ? 0 : (b > SURR_NHI - 1 ? SURR_NHI - 1 : b));
}

/* A fraction in energy class c at position r in [0, 1) of its bin. */
__host__ __device__ __forceinline__ float surr_class_fe(float e511, int c, float r)
{
    if (c == 0) return 1.0f;
    const float ed = surr_fe_edge(e511);
    if (c <= SURR_NLO) return ((float)(c - 1) + r) * (1.0f / (float)SURR_NLO) * ed;
    return ed + ((float)(c - 1 - SURR_NLO) + r) * (1.0f / (float)SURR_NHI) * (SURR_PP_MIN - ed);
}

static_assert(SURR_NH == 190 && SURR_NH <= 192, "head layout");

/* Targets of a first deposit at local (x, y, z) for an entry at (0, y0, z0)
 * with local direction v: depth as a fraction of the panel depth, lateral
 * offsets from the straight entry ray at that depth in crystal pitches. */
__host__ __device__ __forceinline__ void surr_targets(const DevPanel &P, float vx, float vy, float vz, float y0,
                                                      float z0, float x, float y, float z, float &depth, float &dy,
                                                      float &dz)
{
    const float s = x / vx;
    depth = (x * P.dirx) / P.lenx;
    dy = (y - (y0 + s * vy)) / (P.cry + P.cspy);
    dz = (z - (z0 + s * vz)) / (P.crz + P.cspz);
}

/* Inverse of surr_targets; false if the point falls outside the panel. */
__host__ __device__ __forceinline__ bool surr_position(const DevPanel &P, float vx, float vy, float vz, float y0,
                                                       float z0, float depth, float dy, float dz, float &x, float &y,
                                                       float &z)
{
    const float d = fminf(fmaxf(depth, 0.0f), 1.0f);
    x = (d * P.lenx) * P.dirx;
    const float s = x / vx;
    y = (y0 + s * vy) + dy * (P.cry + P.cspy);
    z = (z0 + s * vz) + dz * (P.crz + P.cspz);
    return fabsf(y) <= 0.5f * P.leny && fabsf(z) <= 0.5f * P.lenz;
}

/* Site index of a photon's single output event after adder (crystal level)
 * and readout at (depth, policy) (gPET_kernals.cu:756-813). */
__host__ __device__ __forceinline__ int surr_site(int depth, int policy, int pann, int M_id, int L_id, int moduleN,
                                                  int crystalN)
{
    if (depth == 3) return pann * moduleN * crystalN + M_id * crystalN + L_id;
    if (policy == 1) depth = 2;
    if (depth == 0) return 0;
    if (depth == 1) return pann;
    return pann * moduleN + M_id;
}

#endif
