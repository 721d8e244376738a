"""Benchmark and validation inputs as pmcxcl-style configuration dicts.

``bench`` mirrors the reference package's built-in dicts (pmcxcl/pmcxcl/bench.py:22-46:
cube60, cube60b, cube60planar).  The ``*_workload`` builders produce the
benchmark workloads measured by bench.py and exercised by the tests; every
volume is generated from fixed seeds by this module.
"""

import copy

import numpy as np

CUBE60_PROP = [[0, 0, 1, 1], [0.005, 1, 0.01, 1.37], [0.002, 5, 0.9, 1]]
CUBE60_DETS = [[29, 19, 0, 1], [29, 39, 0, 1], [19, 29, 0, 1], [39, 29, 0, 1]]

bench = {}
bench["cube60"] = {
    "nphoton": 1000000,
    "vol": np.ones([60, 60, 60], dtype="uint8"),
    "tstart": 0,
    "tend": 5e-9,
    "tstep": 5e-9,
    "srcpos": [29, 29, 0],
    "srcdir": [0, 0, 1],
    "prop": CUBE60_PROP,
    "isreflect": 0,
    "seed": 1648335518,
    "session": "cube60",
    "detpos": CUBE60_DETS,
    "issrcfrom0": 1,
}
bench["cube60b"] = copy.deepcopy(bench["cube60"])
bench["cube60b"]["isreflect"] = 1
bench["cube60planar"] = copy.deepcopy(bench["cube60b"])
bench["cube60planar"]["srctype"] = "planar"
bench["cube60planar"]["srcpos"] = [10, 10, -10]
bench["cube60planar"]["srcparam1"] = [40, 0, 0, 0]
bench["cube60planar"]["srcparam2"] = [0, 40, 0, 0]


def sphere_volume(dim=60, center=(30.0, 30.0, 30.0), radius=15.0, inside=2, outside=1):
    """Label volume with a ball, voxel-centre test |(i+.5,j+.5,k+.5)-O|<=R as the
    reference Sphere rasteriser (src/mcx_shapes.c:249-263), Fortran-indexed [x,y,z]."""
    c = np.arange(dim, dtype=np.float32) + np.float32(0.5)
    dx = c[:, None, None] - np.float32(center[0])
    dy = c[None, :, None] - np.float32(center[1])
    dz = c[None, None, :] - np.float32(center[2])
    r2 = np.float32(radius) * np.float32(radius)
    ball = (dx * dx + dy * dy + dz * dz) <= r2
    vol = np.full((dim, dim, dim), outside, dtype=np.uint8)
    vol[ball] = inside
    return vol


def homogeneous_cube_workload(nphoton=1000000, seed=1648335518, **over):
    """60^3 uint8 label-1 cube, 1 mm voxels, mua=0.005 mus=1 g=0.01 n=1.37 in n=1
    background with boundary reflection, pencil beam at (29,29,0) along +z, one
    0-5 ns gate, float32 fluence-rate output, no detectors."""
    cfg = {
        "nphoton": int(nphoton),
        "vol": np.ones([60, 60, 60], dtype=np.uint8),
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 5e-9,
        "srcpos": [29, 29, 0],
        "srcdir": [0, 0, 1],
        "prop": CUBE60_PROP,
        "isreflect": 1,
        "seed": int(seed),
        "issrcfrom0": 1,
        "issavedet": 0,
        "session": "cube60b",
    }
    cfg.update(over)
    return cfg


def sphere_cube_workload(nphoton=100000000, seed=1648335518, **over):
    """60^3 cube with a radius-15 ball of label 2 (mua=0.002 mus=5 g=0.9 n=1.0) at
    (30,30,30); 10 time gates over 0-5 ns; four r=1 mm disk detectors on the z=0 face
    at +-10 voxels; partial path lengths and exit positions saved (D,P,X)."""
    cfg = homogeneous_cube_workload(nphoton=nphoton, seed=seed)
    cfg.update({
        "vol": sphere_volume(),
        "tstep": 5e-10,
        "detpos": CUBE60_DETS,
        "issavedet": 1,
        "savedetflag": "DPX",
        "maxdetphoton": 4000000,
        "session": "cubesph60b",
    })
    cfg.update(over)
    return cfg


LAYER_PROP = [[0, 0, 1, 1], [0.019, 7.8, 0.89, 1.37], [0.019, 7.8, 0.89, 1.37], [0.0004, 0.009, 0.89, 1.37],
              [0.02, 9.0, 0.89, 1.37]]


def _smooth_surface(rng, n, amplitude, sigma=12.0):
    """a seeded random surface: white noise low-passed by a Gaussian of width
    sigma (voxels) in Fourier space, scaled to +-amplitude"""
    noise = rng.standard_normal((n, n))
    k = np.fft.fftfreq(n)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    filt = np.exp(-2.0 * (np.pi * sigma) ** 2 * (kx * kx + ky * ky))
    s = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    s -= s.mean()
    return s * (amplitude / max(np.abs(s).max(), 1e-12))


def layered_volume(dim=200, depths=(20, 40, 70), amplitude=8.0, seed=20251202):
    """dim^3 uint8 volume of 4 layers (labels 1-4 from the top, z = 0) whose
    interfaces lie at the given mean depths (voxels) displaced by seeded smooth
    random surfaces of +-amplitude voxels"""
    rng = np.random.default_rng(seed)
    z = np.arange(dim, dtype=np.float64)[None, None, :] + 0.5
    vol = np.ones((dim, dim, dim), dtype=np.uint8)
    for lab, d in enumerate(depths, start=2):
        surf = d + _smooth_surface(rng, dim, amplitude)
        vol[z >= surf[:, :, None]] = lab
    return vol


def layered_workload(nphoton=100000000, dim=200, seed=1648335518, **over):
    """200^3 4-layer head-like volume (interfaces from seeded Gaussian-filtered
    noise, 8-voxel amplitude), layer mua/mus/g/n (0.019,7.8,0.89,1.37),
    (0.019,7.8,0.89,1.37), (0.0004,0.009,0.89,1.37), (0.02,9.0,0.89,1.37) per mm,
    one pencil source at the centre of the top face and 8 detectors of radius
    2 mm on a 20-voxel ring around it, 0-5 ns in one gate, partial path lengths
    saved"""
    c = dim / 2.0
    dets = [[c + 20 * np.cos(a), c + 20 * np.sin(a), 0, 2.0] for a in np.arange(8) * (np.pi / 4)]
    cfg = {
        "nphoton": int(nphoton),
        "vol": layered_volume(dim),
        "prop": LAYER_PROP,
        "tstart": 0,
        "tend": 5e-9,
        "tstep": 5e-9,
        "srcpos": [c, c, 0],
        "srcdir": [0, 0, 1],
        "isreflect": 1,
        "seed": seed,
        "issrcfrom0": 1,
        "detpos": dets,
        "issavedet": 1,
        "savedetflag": "DP",
        "maxdetphoton": 5000000,
    }
    cfg.update(over)
    return cfg


def _random_rotation(rng):
    """a uniformly random 3x3 rotation (QR of a Gaussian matrix with the sign fixed)"""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoid_volume(rng, dim=64):
    """uint8 [dim]^3 label volume: label 1 with one randomly placed, sized and
    oriented ellipsoid of label 2 (centre in the middle half of the cube,
    semi-axes 4-16 voxels), Fortran-indexed [x, y, z]"""
    centre = rng.uniform(0.3 * dim, 0.7 * dim, 3)
    axes = rng.uniform(4.0, 16.0, 3)
    rot = _random_rotation(rng)
    c = np.arange(dim, dtype=np.float32) + np.float32(0.5)
    x, y, z = np.meshgrid(c - centre[0], c - centre[1], c - centre[2], indexing="ij")
    pts = np.stack([x, y, z], -1) @ rot.astype(np.float32)   # body-frame coordinates
    inside = ((pts / axes.astype(np.float32)) ** 2).sum(-1) <= 1.0
    vol = np.ones((dim, dim, dim), dtype=np.uint8)
    vol[inside] = 2
    return vol


def batched_workload(ninst=256, dim=64, seed=20251203, photons=(1e4, 1e6), srctypes=("pencil", "isotropic", "disk")):
    """Many independent simulations: ``ninst`` [dim]^3 volumes, each with a random
    ellipsoid inclusion; per instance the background and the inclusion draw
    mua in U[0.001, 0.05]/mm, mus in U[0.5, 20]/mm, g in U[0.7, 0.95] (n = 1.37,
    boundary reflection on), the photon count is log-uniform in ``photons`` and
    the source is a pencil beam or a disk (radius 2-8 voxels) entering the z=0
    face at its centre, or an isotropic point at a random interior voxel.  One
    0-5 ns gate, fluence-rate output.  Returns a list of pmcxcl dicts."""
    rng = np.random.default_rng(seed)
    cfgs = []
    for i in range(ninst):
        vol = ellipsoid_volume(rng, dim)
        prop = [[0, 0, 1, 1]]
        for _ in range(2):
            prop.append([float(rng.uniform(0.001, 0.05)), float(rng.uniform(0.5, 20.0)),
                         float(rng.uniform(0.7, 0.95)), 1.37])
        kind = srctypes[int(rng.integers(len(srctypes)))]
        cfg = {
            "nphoton": int(round(10 ** rng.uniform(np.log10(photons[0]), np.log10(photons[1])))),
            "vol": vol,
            "prop": prop,
            "tstart": 0.0,
            "tend": 5e-9,
            "tstep": 5e-9,
            "srctype": kind,
            "srcpos": [dim / 2.0, dim / 2.0, 0.0],
            "srcdir": [0, 0, 1],
            "isreflect": 1,
            "issrcfrom0": 1,
            "issavedet": 0,
            "seed": int(seed + 7919 * i),
            "session": "batch%03d" % i,
        }
        if kind == "isotropic":
            cfg["srcpos"] = [float(v) for v in rng.uniform(0.25 * dim, 0.75 * dim, 3)]
        elif kind == "disk":
            cfg["srcparam1"] = [float(rng.uniform(2.0, 8.0)), 0, 0, 0]
        cfgs.append(cfg)
    return cfgs


def _smooth_field(rng, shape, sigma=10.0):
    """seeded white noise low-passed by a Gaussian of width sigma (voxels) in
    Fourier space, scaled to [-1, 1]"""
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    kk = sum(np.meshgrid(*[ki ** 2 for ki in k], indexing="ij"))
    smooth = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-2.0 * (np.pi * sigma) ** 2 * kk)))
    return (smooth / np.abs(smooth).max()).astype(np.float32)


def inverse_workload(dim=128, photons_per_source=10000000, nside=4, top=30, seed=20251204, **over):
    """Multi-source differentiable step: a [dim]^3 two-layer background (top
    ``top`` voxels mua 0.015/mus 8, below mua 0.03/mus 12 per mm, g 0.9, n 1.37)
    with smooth random per-voxel perturbations of +-30% on mua and on mus,
    nside x nside pencil sources and nside x nside r=2 mm detectors interleaved
    on the z=0 face, photons_per_source photons per source in one launch.
    Returns (cfg, mua, mus) with mua/mus float32 [dim]^3 in 1/mm."""
    rng = np.random.default_rng(seed)
    vol = np.ones((dim, dim, dim), dtype=np.uint8)
    vol[:, :, top:] = 2
    prop = [[0, 0, 1, 1], [0.015, 8.0, 0.9, 1.37], [0.03, 12.0, 0.9, 1.37]]
    base = np.asarray(prop, np.float32)
    mua = base[vol, 0] * (1 + 0.3 * _smooth_field(rng, vol.shape))
    mus = base[vol, 1] * (1 + 0.3 * _smooth_field(rng, vol.shape))
    step = dim / (nside + 0.5)
    src = [[step * (i + 0.5), step * (j + 0.5), 0] for j in range(nside) for i in range(nside)]
    det = [[step * (i + 1.0), step * (j + 1.0), 0, 2] for j in range(nside) for i in range(nside)]
    cfg = {
        "nphoton": int(photons_per_source) * len(src),
        "vol": vol,
        "prop": prop,
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 5e-9,
        "srcpos": src,
        "srcdir": [[0, 0, 1, 0]] * len(src),
        "detpos": det,
        "isreflect": 1,
        "issrcfrom0": 1,
        "savedetflag": "DP",
        "maxdetphoton": 20000000,
        "seed": int(seed),
        "session": "inverse",
    }
    cfg.update(over)
    return cfg, mua.astype(np.float32), mus.astype(np.float32)
