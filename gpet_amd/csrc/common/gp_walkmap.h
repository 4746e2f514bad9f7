/*
 * gp_walkmap.h - uniform-region map of a phantom for the positron-range walk
 * (host code, shared by the HIP library's phantom upload and the CPU oracle
 * so that both walk the same regions).
 *
 * The walk (setPositronRange, gPET_kernals.cu:380-415) crosses voxels one by
 * one, charging each segment at the density of the voxel it enters.  Inside
 * a box of voxels that all hold the same density every crossing but the one
 * leaving the box charges that density, so the walk may advance from inside
 * the box to the last crossing before the box's exit layer in one step.  The
 * map gives, for every 4 x 4 x 4 brick of the padded voxel grid (phantom
 * voxel (i, j, k) is padded voxel (i+1, j+1, k+1); the padding layer is
 * outside the phantom), such a box:
 *
 *   word 0..2: lo | hi << 16, the box's first and last padded voxel along
 *              x, y, z (whole bricks, including this one);
 *   word 3:    the float bits of the box's density, or of -1.0f when the
 *              brick is not uniform (no box).
 *
 * A brick is uniform when its 64 padded voxels all lie in the phantom and
 * hold the same density (bit for bit).  Its box starts as the largest cube
 * of bricks of Chebyshev radius r <= GP_WALKMAP_RMAX around it whose bricks
 * are all uniform with the same density (r rounds of erosion over the 26
 * neighbours), and then grows face by face, in turn -x, +x, -y, +y, -z, +z,
 * by one brick layer at a time while the grown box still holds only such
 * bricks (checked on a summed-volume table of the bricks of that density),
 * by at most GP_WALKMAP_GROW layers per face: a slab of air beside a
 * cylinder becomes a slab-shaped box, which a walk crossing it at a grazing
 * angle leaves in a few jumps instead of many.  Densities held by fewer
 * than 64 uniform bricks keep the cubes.
 */
#ifndef GP_WALKMAP_H
#define GP_WALKMAP_H

#include <stdint.h>
#include <string.h>

#include <map>
#include <vector>

#define GP_WALKMAP_RMAX 15
#define GP_WALKMAP_GROW 24

/* Bricks of the padded grid along an axis of n voxels (n + 2 padded voxels). */
inline int gp_walkmap_bricks(int n) { return (n + 2 + 3) / 4; }

/* dens: nx * ny * nz densities, x fastest.  out: 4 words per padded brick,
 * bricks x fastest.  Returns false (and leaves out empty) when no brick is
 * uniform: the walk then needs no map. */
inline bool gp_build_walkmap(int nx, int ny, int nz, const float *dens, std::vector<uint32_t> &out)
{
    const int bx = gp_walkmap_bricks(nx), by = gp_walkmap_bricks(ny), bz = gp_walkmap_bricks(nz);
    const size_t nb = (size_t)bx * by * bz;
    std::vector<uint32_t> rho(nb, 0u);          /* density bits of uniform bricks */
    std::vector<int8_t> rad(nb, -1);            /* Chebyshev radius, -1 = none     */
    bool any = false;
    for (int kz = 0; kz < bz; ++kz)
        for (int ky = 0; ky < by; ++ky)
            for (int kx = 0; kx < bx; ++kx) {
                /* padded voxels 4k..4k+3 are phantom voxels 4k-1..4k+2 */
                const int x0 = 4 * kx - 1, y0 = 4 * ky - 1, z0 = 4 * kz - 1;
                if (x0 < 0 || y0 < 0 || z0 < 0 || x0 + 3 >= nx || y0 + 3 >= ny || z0 + 3 >= nz) continue;
                uint32_t v0;
                memcpy(&v0, &dens[((size_t)z0 * ny + y0) * nx + x0], 4);
                bool same = true;
                for (int z = z0; z < z0 + 4 && same; ++z)
                    for (int y = y0; y < y0 + 4 && same; ++y)
                        for (int x = x0; x < x0 + 4 && same; ++x) {
                            uint32_t v;
                            memcpy(&v, &dens[((size_t)z * ny + y) * nx + x], 4);
                            same = v == v0;
                        }
                if (!same) continue;
                float f;
                memcpy(&f, &v0, 4);
                if (!(f >= 0.0f)) continue;
                const size_t b = ((size_t)kz * by + ky) * bx + kx;
                rho[b] = v0;
                rad[b] = 0;
                any = true;
            }
    out.clear();
    if (!any) return false;
    /* erosion: radius r + 1 when the brick and its 26 neighbours all have
     * radius >= r and the same density (bricks at the grid's edge are never
     * uniform, so neighbours exist) */
    std::vector<int8_t> next(rad);
    for (int r = 0; r < GP_WALKMAP_RMAX; ++r) {
        bool grew = false;
        for (int kz = 1; kz + 1 < bz; ++kz)
            for (int ky = 1; ky + 1 < by; ++ky)
                for (int kx = 1; kx + 1 < bx; ++kx) {
                    const size_t b = ((size_t)kz * by + ky) * bx + kx;
                    if (rad[b] != r) continue;
                    bool ok = true;
                    for (int dz = -1; dz <= 1 && ok; ++dz)
                        for (int dy = -1; dy <= 1 && ok; ++dy)
                            for (int dx = -1; dx <= 1 && ok; ++dx) {
                                const size_t n = ((size_t)(kz + dz) * by + (ky + dy)) * bx + (kx + dx);
                                ok = rad[n] >= r && rho[n] == rho[b];
                            }
                    if (ok) {
                        next[b] = (int8_t)(r + 1);
                        grew = true;
                    }
                }
        rad = next;
        if (!grew) break;
    }
    /* box growth on summed-volume tables, one per density held by at least
     * 64 uniform bricks */
    std::vector<int> lo3(3 * nb), hi3(3 * nb);
    for (int kz = 0; kz < bz; ++kz)
        for (int ky = 0; ky < by; ++ky)
            for (int kx = 0; kx < bx; ++kx) {
                const size_t b = ((size_t)kz * by + ky) * bx + kx;
                const int r = rad[b] < 0 ? 0 : rad[b];
                const int k[3] = {kx, ky, kz};
                for (int a = 0; a < 3; ++a) {
                    lo3[3 * b + a] = k[a] - r;
                    hi3[3 * b + a] = k[a] + r;
                }
            }
    std::map<uint32_t, size_t> count;
    for (size_t b = 0; b < nb; ++b)
        if (rad[b] >= 0) count[rho[b]]++;
    const size_t sx = (size_t)bx + 1, sy = (size_t)by + 1, sz = (size_t)bz + 1;
    std::vector<int32_t> sat(sx * sy * sz);
    for (const auto &kv : count) {
        if (kv.second < 64) continue;
        const uint32_t v = kv.first;
        /* sat[z][y][x] = number of v-bricks with indices below (x, y, z) */
        for (size_t z = 0; z < sz; ++z)
            for (size_t y = 0; y < sy; ++y)
                for (size_t x = 0; x < sx; ++x) {
                    int32_t c = 0;
                    if (x && y && z) {
                        const size_t b = ((z - 1) * by + (y - 1)) * bx + (x - 1);
                        c = (rad[b] >= 0 && rho[b] == v) ? 1 : 0;
                        c += sat[(z * sy + y) * sx + x - 1] + sat[(z * sy + y - 1) * sx + x] +
                             sat[((z - 1) * sy + y) * sx + x] - sat[(z * sy + y - 1) * sx + x - 1] -
                             sat[((z - 1) * sy + y) * sx + x - 1] - sat[((z - 1) * sy + y - 1) * sx + x] +
                             sat[((z - 1) * sy + y - 1) * sx + x - 1];
                    }
                    sat[(z * sy + y) * sx + x] = c;
                }
        auto S = [&](int x, int y, int z) { return sat[((size_t)z * sy + y) * sx + x]; };
        /* v-bricks in the box [l, h] (inclusive brick indices) */
        auto inside = [&](const int *l, const int *h) {
            const int x0 = l[0], y0 = l[1], z0 = l[2], x1 = h[0] + 1, y1 = h[1] + 1, z1 = h[2] + 1;
            return (int64_t)S(x1, y1, z1) - S(x0, y1, z1) - S(x1, y0, z1) - S(x1, y1, z0) + S(x0, y0, z1) +
                   S(x0, y1, z0) + S(x1, y0, z0) - S(x0, y0, z0);
        };
        const int n3[3] = {bx, by, bz};
        for (size_t b = 0; b < nb; ++b) {
            if (rad[b] < 0 || rho[b] != v) continue;
            int *l = &lo3[3 * b], *h = &hi3[3 * b];
            int grown[6] = {0, 0, 0, 0, 0, 0};
            for (bool more = true; more;) {
                more = false;
                for (int f = 0; f < 6; ++f) {
                    const int a = f >> 1;
                    if (grown[f] >= GP_WALKMAP_GROW) continue;
                    int tl[3] = {l[0], l[1], l[2]}, th[3] = {h[0], h[1], h[2]};
                    if (f & 1) th[a] += 1; else tl[a] -= 1;
                    if (tl[a] < 0 || th[a] >= n3[a]) continue;
                    const int64_t vol = (int64_t)(th[0] - tl[0] + 1) * (th[1] - tl[1] + 1) * (th[2] - tl[2] + 1);
                    if (inside(tl, th) != vol) continue;
                    l[a] = tl[a];
                    h[a] = th[a];
                    grown[f]++;
                    more = true;
                }
            }
        }
    }
    out.assign(4 * nb, 0u);
    const float none = -1.0f;
    uint32_t none_bits;
    memcpy(&none_bits, &none, 4);
    for (int kz = 0; kz < bz; ++kz)
        for (int ky = 0; ky < by; ++ky)
            for (int kx = 0; kx < bx; ++kx) {
                const size_t b = ((size_t)kz * by + ky) * bx + kx;
                uint32_t *w = &out[4 * b];
                const int r = rad[b];
                if (r < 0) {
                    w[3] = none_bits;
                    continue;
                }
                for (int a = 0; a < 3; ++a) {
                    const uint32_t lo = (uint32_t)(4 * lo3[3 * b + a]), hi = (uint32_t)(4 * hi3[3 * b + a] + 3);
                    w[a] = lo | (hi << 16);
                }
                w[3] = rho[b];
            }
    return true;
}

#endif
