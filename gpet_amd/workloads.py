"""Synthetic workloads: phantoms, scanners, sources and digitizer settings.

The reference ships none of the inputs these need (the cylinder phantom and
the PSF are missing blobs, and no scanner matches the benchmark geometry),
so they are generated here from their descriptions, in the reference's file
formats where one exists (config8.geo text, source.txt rows).

Material indices are rows of data/input4gPET.matter: 0 DryAir, 1 Water,
2 TissueICRU, 3 CorticalBone, 7 LSO.
"""
from __future__ import annotations

import os
import tempfile

import numpy as np

from . import capi
from .engine import read_geometry

AIR, WATER, TISSUE, BONE, LSO = 0, 1, 2, 3, 7


def water_cube(n: int = 64, voxel: float = 0.05):
    """Uniform water cube centred on the origin (int32 mat, float32 density)."""
    mat = np.full((n, n, n), WATER, np.int32)
    dens = np.ones((n, n, n), np.float32)
    size = n * voxel
    return mat, dens, (n, n, n), (-size / 2,) * 3, (size,) * 3


def cylinder_phantom(n: int = 200, size: float = 1.0, radius: float = 0.4, rod_radius: float = 0.1,
                     air_density: float = 0.0012):
    """Water cylinder with a cortical-bone rod in air; axis along z, arrays
    indexed [z, y, x] so that x is fastest in memory."""
    d = size / n
    c = (np.arange(n) + 0.5) * d - size / 2
    yy, xx = np.meshgrid(c, c, indexing="ij")
    r2 = xx * xx + yy * yy
    sl_mat = np.full((n, n), AIR, np.int32)
    sl_den = np.full((n, n), air_density, np.float32)
    sl_mat[r2 <= radius * radius] = WATER
    sl_den[r2 <= radius * radius] = 1.0
    sl_mat[r2 <= rod_radius * rod_radius] = BONE
    sl_den[r2 <= rod_radius * rod_radius] = 1.85
    mat = np.broadcast_to(sl_mat, (n, n, n)).copy()
    dens = np.broadcast_to(sl_den, (n, n, n)).copy()
    return mat, dens, (n, n, n), (-size / 2,) * 3, (size,) * 3


def geometry_text(n_panels: int = 8, radius: float = 8.0, ncry: int = 20, pitch: float = 0.15,
                  gap: float = 0.0005, depth: float = 1.0, crystal_mat: int = LSO, crystal_den: float = 7.4,
                  rings: int = 1, ring_spacing: float = 0.0) -> str:
    """A config8.geo-format scanner: n_panels flat panels of ncry x ncry
    crystals (one module each) on a ring of inner radius `radius`.  The
    crystal size is pitch - gap with a gap > 0 so the reference's count
    formula floor(M/(L+s))+1 gives ncry (initialize.cu:1074-1086)."""
    cry = pitch - gap
    mod = ncry * pitch - gap
    ang = 360.0 / n_panels
    lines = [
        "panel repetition times", str(n_panels),
        "panel rotational axis in the global coordinate", "0 0 1",
        "panel rotational angle (in degree, plus: counterclockwise, minus: clockwise)", "%.10g" % ang,
        "material density (in the order of crystal material, air)",
        "%d %.6g" % (crystal_mat, crystal_den), "0 0.001025",
        "%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%",
        "starting panel index", "0",
        "dimension(cm)", "%.6g %.6g %.6g" % (depth, mod, mod),
        "module size", "%.6g %.6g %.6g" % (depth, mod, mod),
        "module space", "%.6g %.6g %.6g" % (depth, 0.01, 0.01),
        "crystal size (cm)", "%.6g %.6g %.6g" % (depth, cry, cry),
        "crystal space(cm)", "%.6g %.6g %.6g" % (0.0, gap, gap),
        "panel expansion direction in the local coordinate (x y z)", "-1 1 1",
        "offset of the panel surface center facing the phantom (cm)", "0 %.6g 0" % (-radius),
        "unit vector along the local x direction in the global coordinate", "0 1 0",
        "unit vector along the local y direction in the global coordinate", "1 0 0",
        "unit vector along the local z direction in the global coordinate", "0 0 -1",
        "%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%%",
    ]
    if rings > 1:
        lines += ["number of rings and ring spacing along the rotation axis (cm)", "%d %.6g" % (rings, ring_spacing)]
    return "\n".join(lines) + "\n"


def octagon_scanner(**kw):
    """Panels and materials of the 8-panel LSO octagon, parsed from the
    config8.geo text by the library's own reader."""
    txt = geometry_text(**kw)
    with tempfile.NamedTemporaryFile("w", suffix=".geo", delete=False) as fh:
        fh.write(txt)
        path = fh.name
    try:
        return read_geometry(path)
    finally:
        os.unlink(path)


# Body workload: 256 x 256 x 512 voxels inside a 4-ring scanner of 12 panels
# per ring.  The ring's inscribed diameter is 40 cm; the voxel grid must lie
# inside the panels (photons are carried to the grid boundary and only then
# projected onto the panels), and a square grid of half width h clears flat
# panels at 30-degree steps when 1.366 h < 20 cm, so the voxels are 0.1 cm
# (25.6 x 25.6 x 51.2 cm grid).
BODY_DIMS = (256, 256, 512)
BODY_VOXEL = 0.1
BODY_AXES = (11.5, 8.5, 24.0)      # water ellipsoid semi-axes (cm)
LUNG_AXES = (3.5, 4.5, 7.0)
LUNG_CENTRES = ((-5.0, -0.5, 6.0), (5.0, -0.5, 6.0))
LUNG_DENSITY = 0.3                 # TissueICRU at lung density
SPINE_CENTRE = (0.0, 5.5)          # cylinder along z
SPINE_RADIUS = 1.2
BONE_DENSITY = 1.85
N_HOT = 20
HOT_RATIO = 8.0                    # sphere : background activity concentration


def _inside_ellipsoid(x, y, z, c, a):
    return ((x - c[0]) / a[0]) ** 2 + ((y - c[1]) / a[1]) ** 2 + ((z - c[2]) / a[2]) ** 2 <= 1.0


def hot_spheres(seed: int = 5):
    """Centres and radii (cm) of the hot spheres: radii 0.5-1.5 cm, placed by
    rejection inside the soft tissue of the body (clear of lungs and spine)
    with a deterministic generator."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < N_HOT:
        r = float(rng.uniform(0.5, 1.5))
        c = rng.uniform(-1.0, 1.0, 3) * np.array(BODY_AXES) * 0.8
        # sphere fully inside the body: shrink the axes by r
        if not _inside_ellipsoid(c[0], c[1], c[2], (0, 0, 0), tuple(a - r for a in BODY_AXES)):
            continue
        if any(_inside_ellipsoid(c[0], c[1], c[2], lc, tuple(a + r for a in LUNG_AXES)) for lc in LUNG_CENTRES):
            continue
        if np.hypot(c[0] - SPINE_CENTRE[0], c[1] - SPINE_CENTRE[1]) < SPINE_RADIUS + r:
            continue
        if any(np.linalg.norm(c - np.array(o[0])) < r + o[1] + 0.2 for o in out):
            continue
        out.append((tuple(float(v) for v in c), r))
    return out


def body_phantom(dims=BODY_DIMS, voxel: float = BODY_VOXEL, air_density: float = 0.0012):
    """Ellipsoidal water body with two lung ellipsoids, a cortical-bone spine
    cylinder and (for transport) water hot spheres, in air, centred on the
    origin; arrays indexed [z, y, x].  The grid covers dims * voxel cm; a
    coarse grid of the same extent (e.g. 32 x 32 x 64 at 0.8 cm) gives the
    same body at low resolution."""
    nx, ny, nz = dims
    cx = (np.arange(nx, dtype=np.float32) + 0.5) * voxel - nx * voxel / 2
    cy = (np.arange(ny, dtype=np.float32) + 0.5) * voxel - ny * voxel / 2
    cz = (np.arange(nz, dtype=np.float32) + 0.5) * voxel - nz * voxel / 2
    mat = np.full((nz, ny, nx), AIR, np.int32)
    dens = np.full((nz, ny, nx), air_density, np.float32)
    yy, xx = np.meshgrid(cy, cx, indexing="ij")
    for k in range(nz):
        z = cz[k]
        m = mat[k]
        d = dens[k]
        body = _inside_ellipsoid(xx, yy, z, (0, 0, 0), BODY_AXES)
        m[body] = WATER
        d[body] = 1.0
        for lc in LUNG_CENTRES:
            lung = body & _inside_ellipsoid(xx, yy, z, lc, LUNG_AXES)
            m[lung] = TISSUE
            d[lung] = LUNG_DENSITY
        spine = body & ((xx - SPINE_CENTRE[0]) ** 2 + (yy - SPINE_CENTRE[1]) ** 2 <= SPINE_RADIUS ** 2)
        m[spine] = BONE
        d[spine] = BONE_DENSITY
    size = (nx * voxel, ny * voxel, nz * voxel)
    return mat, dens, dims, tuple(-v / 2 for v in size), size


def body_scanner(rings: int = 4, **kw):
    """4 rings x 12 flat LSO panels of 20 x 20 crystals (0.53 cm pitch,
    2 cm deep) with a 40 cm inscribed ring diameter."""
    args = dict(n_panels=12, radius=20.0, ncry=20, pitch=0.53, gap=0.005, depth=2.0, rings=rings,
                ring_spacing=10.7)
    args.update(kw)
    return octagon_scanner(**args)


def body_sources(seed: int = 5, concentration: float = 1.0e7):
    """Background F-18 in the body ellipsoid (shape 3) plus the hot spheres
    at HOT_RATIO : 1 (each sphere adds HOT_RATIO - 1 times the background
    concentration on top of the background that already covers it);
    natom in atoms per cm^3 * volume."""
    a, b, c = BODY_AXES
    srcs = [source(3, c3=a, c4=b, c5=c, natom=concentration * 4.0 / 3.0 * np.pi * a * b * c)]
    for centre, r in hot_spheres(seed):
        srcs.append(source(2, centre=centre, c3=r,
                           natom=(HOT_RATIO - 1.0) * concentration * 4.0 / 3.0 * np.pi * r ** 3))
    return srcs


# Batched workload: many independent phantoms of uneven size on the octagon
# scanner.  Every grid spans the same 10 cm cube (inside the 8 cm panels:
# 1.307 x 5 cm < 8 cm), so voxel size is 10/n cm.
BATCH_EXTENT = 10.0
BATCH_BODY_RADIUS = 4.5
LABEL_DENSITY = {AIR: 0.0012, WATER: 1.0, BONE: 1.85, TISSUE: 0.3}   # TissueICRU as lung


def blob_phantom(n: int, rng: np.random.Generator, extent: float = BATCH_EXTENT):
    """n^3 labels from a smoothed Gaussian random field: a coarse 8^3 white
    noise field smoothed and upsampled linearly, cut at its 15th/85th
    percentiles into lung / water / bone inside a cylinder of radius 4.5 cm
    (air outside); densities are the label density times U(0.95, 1.05)."""
    from scipy import ndimage

    coarse = ndimage.gaussian_filter(rng.normal(size=(8, 8, 8)), 1.0, mode="wrap").astype(np.float32)
    # separable linear upsampling 8 -> n: f = A (x) A (x) A applied to coarse
    u = np.clip((np.arange(n) + 0.5) * (8.0 / n) - 0.5, 0.0, 7.0)
    i0 = np.minimum(u.astype(int), 6)
    w = (u - i0).astype(np.float32)
    A = np.zeros((n, 8), np.float32)
    A[np.arange(n), i0] = 1.0 - w
    A[np.arange(n), i0 + 1] += w
    f = np.tensordot(A, coarse, axes=(1, 0))              # [n, 8, 8]
    f = np.tensordot(f, A, axes=(2, 1))                   # [n, 8, n]
    f = np.einsum("jb,ibk->ijk", A, f, optimize=True)     # [n, n, n]
    f = np.ascontiguousarray(f, dtype=np.float32)
    lo, hi = np.quantile(f, [0.15, 0.85])
    mat = np.full((n, n, n), WATER, np.int32)
    mat[f < lo] = TISSUE
    mat[f > hi] = BONE
    d = extent / n
    c = (np.arange(n, dtype=np.float32) + 0.5) * d - extent / 2
    yy, xx = np.meshgrid(c, c, indexing="ij")
    outside = (xx * xx + yy * yy) > BATCH_BODY_RADIUS ** 2
    mat[:, outside] = AIR
    lut = np.zeros(max(LABEL_DENSITY) + 1, np.float32)
    for lab, rho in LABEL_DENSITY.items():
        lut[lab] = rho
    dens = rng.random((n, n, n), dtype=np.float32)
    dens *= np.float32(0.1)
    dens += np.float32(0.95)
    dens *= lut[mat]
    return mat, dens, (n, n, n), (-extent / 2,) * 3, (extent,) * 3


def batch_workload(n_phantoms: int = 64, seed: int = 3, size_range=(32, 256), decay_range=(1e5, 1e7)):
    """(phantoms, jobs, sources, isotopes): phantom k has n_k ~ U{lo..hi}
    voxels per axis and one job of F-18 uniform in a cylinder (r 4 cm,
    h 8 cm) over 10 s with a log-uniform decay count."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(size_range[0], size_range[1] + 1, n_phantoms)
    counts = np.floor(10.0 ** rng.uniform(np.log10(decay_range[0]), np.log10(decay_range[1]), n_phantoms))
    phantoms = [blob_phantom(int(n), rng) for n in sizes]
    sources, jobs = [], []
    for k in range(n_phantoms):
        sources.append(source(1, c3=4.0, c4=8.0, natom=1.0e12))
        jobs.append(dict(phantom=k, first_source=k, n_sources=1, t0=0.0, t1=10.0, n_decays=int(counts[k])))
    return phantoms, jobs, sources, [isotope_f18()]


def setup_batch(eng, tables, phantoms, dig=None):
    """Tables, the phantom slots, the octagon scanner and the digitizer."""
    eng.set_tables(tables)
    for k, ph in enumerate(phantoms):
        eng.set_phantom(*ph, slot=k)
    panels, mats = octagon_scanner()
    eng.set_detector(panels, mats, np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32))
    eng.set_digitizer(dig or digitizer())
    return panels, mats


def isotope_f18() -> capi.Isotope:
    """Row 0 of data/isotopes.txt (F-18)."""
    iso = capi.Isotope()
    iso.half_life_s = 6586.26
    iso.beta_ratio = 0.97
    for k, v in enumerate([1.144, 0.5209, 0, 0, 10.2088, -30.4551, 28.4376, -7.9828]):
        iso.coef[k] = v
    return iso


def isotope_pure_pairs() -> capi.Isotope:
    """A positron emitter that always emits (ratio 1) with a half-life long
    enough for a flat emission rate; used for 'N back-to-back pairs'."""
    iso = isotope_f18()
    iso.half_life_s = 1.0e9
    iso.beta_ratio = 1.0
    return iso


def source(shape: int, centre=(0.0, 0.0, 0.0), c3=0.0, c4=0.0, c5=0.0, natom=1.0e12, isotope=0) -> capi.Source:
    s = capi.Source()
    s.natom = natom
    s.isotope = isotope
    s.shape = shape
    for k, v in enumerate([centre[0], centre[1], centre[2], c3, c4, c5]):
        s.coef[k] = v
    return s


def digitizer(readout_depth=2, readout_policy=1, e_threshold=50e3, blur_policy=1, e_ref=511e3, r_ref=0.12,
              slope=0.0, space_blur=0.0, dead_level=3, dead_type=0, dead_time_us=0.0, win_min=350e3,
              win_max=650e3, time_blur_fwhm_us=300e-6, coinc_window_us=4e-3, coinc_policy=0,
              record_rayleigh=0, noise_rate=0.0, noise_e_mean=0.0, noise_e_sigma=0.0,
              delay_us=0.0) -> capi.Digitizer:
    """Digitizer of the benchmark scanner: 12% FWHM at 511 keV, 350-650 keV,
    300 ps time blur, 4 ns coincidence window, no deadtime.  noise_rate
    (singles/s over the scanner) adds background singles; delay_us > 0 adds
    the delayed-window randoms estimate."""
    d = capi.Digitizer()
    d.readout_depth = readout_depth
    d.readout_policy = readout_policy
    d.e_threshold = e_threshold
    d.e_upholder = 2.0e6
    d.blur_policy = blur_policy
    d.e_ref = e_ref
    d.r_ref = r_ref
    d.slope = slope
    d.space_blur = space_blur
    d.dead_level = dead_level
    d.dead_type = dead_type
    d.dead_time_us = dead_time_us
    d.win_min = win_min
    d.win_max = win_max
    d.time_blur_fwhm_us = time_blur_fwhm_us
    d.coinc_window_us = coinc_window_us
    d.coinc_policy = coinc_policy
    d.record_rayleigh = record_rayleigh
    d.noise_rate = noise_rate
    d.noise_e_mean = noise_e_mean
    d.noise_e_sigma = noise_e_sigma
    d.delay_us = delay_us
    return d


def setup_engine(eng, tables, config: int = 2, phantom_n=None, scanner_kw=None, dig=None):
    """Load tables, phantom, scanner and digitizer of configuration 1 (water
    cube), 2 (cylinder with bone rod) or 5 (body in the 48-panel scanner;
    phantom_n, if given, is the x/y voxel count of a coarser grid of the same
    extent, z twice that)."""
    eng.set_tables(tables)
    if config == 1:
        ph = water_cube(phantom_n or 64)
    elif config == 5:
        if phantom_n:
            ph = body_phantom((phantom_n, phantom_n, 2 * phantom_n), BODY_VOXEL * BODY_DIMS[0] / phantom_n)
        else:
            ph = body_phantom()
    else:
        ph = cylinder_phantom(phantom_n or 200)
    eng.set_phantom(*ph)
    if config == 5:
        panels, mats = body_scanner(**(scanner_kw or {}))
    else:
        panels, mats = octagon_scanner(**(scanner_kw or {}))
    eng.set_detector(panels, mats, np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32))
    eng.set_digitizer(dig or digitizer())
    return panels, mats


def workload_sources(config: int):
    """(sources, isotopes, t0, t1) of configuration 1, 2 or 5."""
    if config == 1:
        return [source(2, natom=1.0e15)], [isotope_pure_pairs()], 0.0, 1.0
    if config == 5:
        return body_sources(), [isotope_f18()], 0.0, 60.0
    # F-18 uniform in the r = 0.4 cm water cylinder (full 1 cm height)
    return [source(1, c3=0.4, c4=1.0, natom=1.0e11)], [isotope_f18()], 0.0, 10.0
