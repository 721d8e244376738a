"""Pins of the oracle's photon-launch and scattering samplers against the
distributions the reference's algorithms define (launch: src/mcx_core.cl:
1632-2210; scattering angles: :2446-2500).

The samples come straight from the oracle's own launch and scattering code
(oracle_launch_samples, oracle_scatter_samples), so every device parity test of
these branches inherits the pin.  Each distribution is compared with a closed
form: a CDF by the Kolmogorov-Smirnov test (seeded samples, so the p-values are
fixed numbers; the bar is p > 1e-3), a geometric identity, or an exact value.
A plausible restatement mistake fails one of them: a radius drawn linearly
instead of by its square root, a Gaussian width off by a factor, an annulus
offset left out, a swapped axis, a missing renormalisation, interpolation over
nphase instead of nphase-1 intervals, a roulette survivor not reweighted."""

import math

import numpy as np
import pytest
from scipy import special, stats

import oracle
from mcxcl_amd import workloads

N = 100000
PMIN = 1e-3


def _cube(**over):
    over.setdefault("nphoton", 1000)
    return workloads.homogeneous_cube_workload(**over)


def _ks(x, cdf):
    p = stats.kstest(np.asarray(x, np.float64), cdf).pvalue
    assert p > PMIN, p
    return p


def _uniform(lo, hi):
    return lambda x: np.clip((x - lo) / (hi - lo), 0.0, 1.0)


def _disk_cdf(rmax, rmin=0.0):
    return lambda r: np.clip((r * r - rmin * rmin) / (rmax * rmax - rmin * rmin), 0.0, 1.0)


def _samples(n=N, **over):
    s = oracle.launch_samples(_cube(**over), n)
    assert len(s["w"]) == n
    return s


# ---------------- position samplers ----------------

def test_disk_uniform_over_annulus_along_axis():
    """disk (param1 = {R, r}): r = sqrt(U (R^2 - r^2) + r^2) (src/mcx_core.cl:1894),
    so the radius has CDF (rho^2 - r^2) / (R^2 - r^2) and the azimuth is uniform;
    along +z the points stay in the source plane"""
    s = _samples(srctype="disk", srcpos=[30, 30, 0], srcparam1=[8, 3, 0, 0])
    d = s["p"].astype(np.float64) - [30, 30, 0]
    assert np.all(d[:, 2] == 0)
    rho = np.hypot(d[:, 0], d[:, 1])
    _ks(rho, _disk_cdf(8, 3))
    _ks(np.arctan2(d[:, 1], d[:, 0]), _uniform(-math.pi, math.pi))
    assert np.all(s["v"] == np.float32([0, 0, 1]))


def test_disk_on_tilted_axis_lies_in_the_normal_plane():
    """off-axis direction: the offset is rotated into the plane normal to the
    direction (src/mcx_core.cl:1915-1925), keeping the radius distribution"""
    d0 = np.array([0, 0.6, 0.8])
    s = _samples(srctype="disk", srcpos=[30, 30, 30], srcdir=list(d0), srcparam1=[8, 0, 0, 0])
    d = s["p"].astype(np.float64) - [30, 30, 30]
    assert np.abs(d @ d0).max() < 2e-5
    _ks(np.linalg.norm(d, axis=1), _disk_cdf(8))
    e1 = np.array([1.0, 0, 0])
    e2 = np.cross(d0, e1)
    _ks(np.arctan2(d @ e2, d @ e1), _uniform(-math.pi, math.pi))


def test_ring_angle_range():
    """ring with an angle range (param1 = {R, r, phi1, phi2}): the azimuth is
    uniform over [min, min + |phi1 - phi2|] (src/mcx_core.cl:1868-1870)"""
    s = _samples(srctype="ring", srcpos=[30, 30, 0], srcparam1=[8, 4, 0.5, 2.5])
    d = s["p"].astype(np.float64) - [30, 30, 0]
    _ks(np.arctan2(d[:, 1], d[:, 0]), _uniform(0.5, 2.5))
    _ks(np.hypot(d[:, 0], d[:, 1]), _disk_cdf(8, 4))


def test_gaussian_waist():
    """gaussian (param1.x = waist w): r = w sqrt(-log(U)/2), i.e. the 1/e^2
    intensity radius: P(rho <= x) = 1 - exp(-2 x^2 / w^2) (src/mcx_core.cl:1899-1901)"""
    s = _samples(srctype="gaussian", srcpos=[30, 30, 0], srcparam1=[5, 0, 0, 0])
    d = s["p"].astype(np.float64) - [30, 30, 0]
    _ks(np.hypot(d[:, 0], d[:, 1]), lambda x: 1.0 - np.exp(-2.0 * x * x / 25.0))
    assert np.all(s["v"] == np.float32([0, 0, 1]))


def test_focused_gaussian_beam():
    """gaussian with a focus (srcdir.w = f, param1.y = wavelength): the waist
    grows to w sqrt(1 + f^2/zR^2) with zR = pi w^2 / lambda (src/mcx_core.cl:
    1902-1904), and every ray points at the focal point srcpos + f dir
    (src/mcx_core.cl:2140-2152)"""
    w, lam, f = 2.0, 0.5, 20.0
    s = _samples(srctype="gaussian", srcpos=[30, 30, 0], srcdir=[0, 0, 1, f], srcparam1=[w, lam, 0, 0])
    p = s["p"].astype(np.float64)
    zr = math.pi * w * w / lam
    wz2 = w * w * (1.0 + f * f / (zr * zr))
    _ks(np.hypot(p[:, 0] - 30, p[:, 1] - 30), lambda x: 1.0 - np.exp(-2.0 * x * x / wz2))
    to_focus = np.array([30, 30, f]) - p
    to_focus /= np.linalg.norm(to_focus, axis=1)[:, None]
    assert np.abs(np.cross(to_focus, s["v"].astype(np.float64))).max() < 1e-5
    assert np.all((to_focus * s["v"]).sum(1) > 0)


def test_planar_uniform_over_parallelogram():
    """planar (param1, param2 = edge vectors): p = pos + U1 a + U2 b"""
    s = _samples(srctype="planar", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 0], srcparam2=[0, 40, 0, 0])
    u = (s["p"][:, 0].astype(np.float64) - 10) / 40
    v = (s["p"][:, 1].astype(np.float64) - 10) / 40
    assert np.all(s["p"][:, 2] == 0)
    _ks(u, _uniform(0, 1))
    _ks(v, _uniform(0, 1))
    assert abs(np.corrcoef(u, v)[0, 1]) < 4.0 / math.sqrt(N)


def test_pencil_array_grid():
    """pencilarray (param1.w x param2.w beams): each of the 4 x 3 grid points
    pos + i a/(nx-1) + j b/(ny-1) is equally likely (src/mcx_core.cl:1701-1703)"""
    s = _samples(srctype="pencilarray", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 4], srcparam2=[0, 40, 0, 3])
    i = np.rint((s["p"][:, 0].astype(np.float64) - 10) / (40.0 / 3)).astype(int)
    j = np.rint((s["p"][:, 1].astype(np.float64) - 10) / 20.0).astype(int)
    np.testing.assert_allclose(s["p"][:, 0], 10 + i * np.float32(40.0 / 3), atol=1e-4)
    np.testing.assert_allclose(s["p"][:, 1], 10 + j * 20.0, atol=1e-4)
    counts = np.bincount(i * 3 + j, minlength=12)
    assert counts.size == 12 and counts.min() > 0
    assert stats.chisquare(counts).pvalue > PMIN


def test_fourier_pattern_weight():
    """fourier (param1.w = kx + phase, param2.w = ky + 1 - modulation): the
    launch weight is (cos(2 pi (kx u + ky v + phase)) (1 - m) + 1) / 2 at the
    launch point's parallelogram coordinates (u, v) (src/mcx_core.cl:1694-1696);
    with integer frequencies it averages 1/2"""
    s = _samples(srctype="fourier", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 2.25], srcparam2=[0, 40, 0, 1.5])
    u = (s["p"][:, 0].astype(np.float64) - 10) / 40
    v = (s["p"][:, 1].astype(np.float64) - 10) / 40
    expect = (np.cos(2 * math.pi * (2 * u + 1 * v + 0.25)) * 0.5 + 1.0) * 0.5
    np.testing.assert_allclose(s["w"], expect, atol=2e-5)
    assert abs(s["w"].mean() - 0.5) < 4 * s["w"].std() / math.sqrt(N)


def test_line_source():
    """line (param1 = segment): the start point is uniform along the segment and
    the direction is isotropic in the plane normal to it (src/mcx_core.cl:
    2023-2041)"""
    s = _samples(srctype="line", srcpos=[20, 29.5, 29.5], srcparam1=[20, 0, 0, 0])
    assert np.all(s["p"][:, 1:] == np.float32(29.5))
    _ks(s["p"][:, 0].astype(np.float64) - 20, _uniform(0, 20))
    v = s["v"].astype(np.float64)
    assert np.abs(v[:, 0]).max() < 1e-6
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    _ks(np.arctan2(v[:, 2], v[:, 1]), _uniform(-math.pi, math.pi))


def test_line_source_fan():
    """line with a fan angle (param2.x = a): the direction is srcdir projected
    normal to the line, turned about the line by a uniform angle in [-a, a]"""
    s = _samples(srctype="line", srcpos=[20, 29.5, 29.5], srcdir=[0, 1, 0], srcparam1=[20, 0, 0, 0],
                 srcparam2=[0.5, 0, 0, 0])
    v = s["v"].astype(np.float64)
    assert np.abs(v[:, 0]).max() < 1e-6
    _ks(np.arctan2(v[:, 2], v[:, 1]), _uniform(-0.5, 0.5))


def test_slit_source():
    """slit: uniform along the segment, direction srcdir; with a spread (param2 =
    {sx, sy}) the direction is srcdir + sx N1 e_c + sy N2 e_s renormalised, N1, N2
    independent standard normals (Box-Muller on -log U), e_s along the slit and
    e_c = slit x srcdir (src/mcx_core.cl:2052-2071)"""
    s = _samples(srctype="slit", srcpos=[20, 29.5, 0], srcparam1=[20, 0, 0, 0])
    _ks(s["p"][:, 0].astype(np.float64) - 20, _uniform(0, 20))
    assert np.all(s["v"] == np.float32([0, 0, 1]))
    s = _samples(srctype="slit", srcpos=[20, 29.5, 0], srcparam1=[20, 0, 0, 0], srcparam2=[0.1, 0.2, 0, 0])
    v = s["v"].astype(np.float64)
    # e_s = +x, e_c = (20,0,0) x (0,0,1) / 20 = -y
    n_c = -v[:, 1] / v[:, 2] / 0.1
    n_s = v[:, 0] / v[:, 2] / 0.2
    _ks(n_c, stats.norm.cdf)
    _ks(n_s, stats.norm.cdf)
    assert abs(np.corrcoef(n_c, n_s)[0, 1]) < 4.0 / math.sqrt(N)


# ---------------- direction samplers ----------------

def test_isotropic_source():
    """isotropic: cos(theta) = 2U - 1 about srcdir, uniform azimuth"""
    s = _samples(srctype="isotropic", srcpos=[30, 30, 30])
    v = s["v"].astype(np.float64)
    _ks(v[:, 2], _uniform(-1, 1))
    _ks(np.arctan2(v[:, 1], v[:, 0]), _uniform(-math.pi, math.pi))


def test_arcsine_source():
    """arcsine: the polar angle itself is uniform on [0, pi]"""
    s = _samples(srctype="arcsine", srcpos=[30, 30, 30])
    _ks(np.arccos(np.clip(s["v"][:, 2].astype(np.float64), -1, 1)), _uniform(0, math.pi))


@pytest.mark.parametrize("mode", [0, 1])
def test_cone_source(mode):
    """cone (param1.x = half angle a): param1.y = 0 rejects acos(2U-1) > a, so
    cos(theta) is uniform on [cos a, 1]; param1.y > 0 rejects 2 pi U > a, so theta
    is uniform on [0, a] (src/mcx_core.cl:1957-1959)"""
    a = 0.3
    s = _samples(srctype="cone", srcpos=[30, 30, 30], srcparam1=[a, mode, 0, 0])
    c = s["v"][:, 2].astype(np.float64)
    if mode == 0:
        _ks(c, _uniform(math.cos(a), 1))
    else:
        _ks(np.arccos(np.clip(c, -1, 1)), _uniform(0, a))


def test_zgaussian_source():
    """zgaussian (param1.x = s): theta = s sqrt(-2 log U1) (1 - 2 U2)
    (src/mcx_core.cl:1990), so |theta| = s R V with R Rayleigh and V uniform on
    [0, 1]: P(|theta| <= x) = 1 - exp(-a) + sqrt(pi a) erfc(sqrt a), a = x^2/(2 s^2)"""
    sg = 0.3
    s = _samples(srctype="zgaussian", srcpos=[30, 30, 30], srcparam1=[sg, 0, 0, 0])
    th = np.arccos(np.clip(s["v"][:, 2].astype(np.float64), -1, 1))

    def cdf(x):
        a = x * x / (2 * sg * sg)
        return 1.0 - np.exp(-a) + np.sqrt(math.pi * a) * special.erfc(np.sqrt(a))

    _ks(th, cdf)


# ---------------- scattering ----------------

def _hg_cdf(g):
    return lambda mu: np.clip((1 - g * g) / (2 * g) * (1.0 / np.sqrt(1 + g * g - 2 * g * mu) - 1.0 / (1 + g)), 0, 1)


@pytest.mark.parametrize("g", [0.9, 0.5, -0.4])
def test_henyey_greenstein_cdf(g):
    """HG inversion (src/mcx_core.cl:2484-2497) against the HG CDF; the rotation
    turns the direction by exactly that angle"""
    cfg = _cube(prop=[[0, 0, 1, 1], [0.005, 1.0, g, 1.37]])
    s = oracle.scatter_samples(cfg, 1, 0, N)
    _ks(s["ctheta"], _hg_cdf(g))
    np.testing.assert_allclose(s["cosdev"], s["ctheta"], atol=2e-6)
    np.testing.assert_allclose(s["stheta"] ** 2 + s["ctheta"] ** 2, 1.0, atol=3e-6)


def test_similarity_rule_switches_to_isotropic_after_gscatter():
    """gscatter = k: scattering events after the k-th use g = 0 (src/mcx_core.cl:
    2484): the k-th is still HG, the (k+1)-th is isotropic"""
    cfg = _cube(prop=[[0, 0, 1, 1], [0.005, 1.0, 0.9, 1.37]], gscatter=3)
    _ks(oracle.scatter_samples(cfg, 1, 3, N)["ctheta"], _hg_cdf(0.9))
    _ks(oracle.scatter_samples(cfg, 1, 4, N)["ctheta"], _uniform(-1, 1))


def test_similarity_rule_flight_lengths():
    """flights after gscatter events are drawn with mus (1 - g) (src/mcx_core.cl:
    2680-2682): with gscatter = 0 every flight is, so in a non-absorbing medium
    the scattering events per unit path length are mus (1 - g) = 2.5/mm, against
    mus = 10/mm without the rule.  The path length is the pathlength output
    (w0 L with w0 = 1 when mua = 0); only each photon's escaping flight ends
    without an event, a bias below 1%."""
    base = dict(nphoton=20000, vol=np.ones((20, 20, 20), np.uint8), srcpos=[10, 10, 0], isreflect=0,
                prop=[[0, 0, 1, 1], [0.0, 10.0, 0.75, 1.0]], tend=1e-8, tstep=1e-8, outputtype="length",
                autopilot=0, nthread=256, nblocksize=64, rawfield=1)
    for gsc, rate in ((0, 2.5), (None, 10.0)):
        cfg = _cube(**base) if gsc is None else _cube(gscatter=gsc, **base)
        out = oracle.run(cfg, nthread=256)
        assert out["stat"]["energyabs"] == 0
        measured = out["counters"]["nscat"] / out["flux64"].sum()
        assert abs(measured / rate - 1) < 0.01, (gsc, measured)


def test_sampled_phase_function_cdf():
    """invcdf table c_0 = -1 < ... < c_{n-1} = 1: cos(theta) is the table linearly
    interpolated at U (n - 1) (src/mcx_core.cl:2459-2475), so its CDF is
    piecewise linear through (c_k, k / (n - 1))"""
    inner = np.array([-0.6, -0.2, 0.1, 0.35, 0.55, 0.7, 0.8, 0.88, 0.94, 0.98], np.float32)
    cfg = _cube(invcdf=inner)
    table = np.concatenate([[-1.0], inner.astype(np.float64), [1.0]])
    s = oracle.scatter_samples(cfg, 1, 0, N)
    _ks(s["ctheta"], lambda mu: np.interp(mu, table, np.linspace(0, 1, table.size)))
    np.testing.assert_allclose(s["cosdev"], s["ctheta"], atol=2e-6)


# ---------------- roulette ----------------

def test_russian_roulette_column():
    """Russian roulette (src/mcx_core.cl:3031-3061) in a non-scattering column
    where every photon has the same weight history w_k = exp(-0.1 k) at the
    entry of voxel k: once the weight drops below minenergy = 0.05 (voxel 30,
    and again at voxel 54 for survivors) a photon survives with probability
    1/10 and continues with 10 w.  The energy deposits per voxel are then
    exactly N (w_k - w_k+1) before voxel 30, 10 n1 (w_k - w_k+1) up to voxel 53
    and 100 n2 (w_k - w_k+1) after, with n1 ~ Bin(N, 1/10) survivors and
    n2 ~ Bin(n1, 1/10).  The reference does not reset the voxel-entry weight w0
    when it multiplies w by 10, so the voxel where a photon survives receives
    w_k (1 - 10 e^-0.1) < 0: a negative deposit, kept here for parity (the
    absorbed energy of a run with roulette is biased low by 0.9 w per roulette
    event)."""
    n = 20000
    cfg = _cube(nphoton=n, vol=np.ones((1, 1, 60), np.uint8), srcpos=[0.5, 0.5, 0], isreflect=0,
                prop=[[0, 0, 1, 1], [0.1, 1e-7, 0.0, 1.0]], tend=1e-9, tstep=1e-9, outputtype="energy",
                minenergy=0.05, autopilot=0, nthread=256, nblocksize=64, rawfield=1)
    out = oracle.run(cfg, nthread=256)
    assert out["counters"]["nscat"] == 0
    w = np.exp(-0.1 * np.arange(61))
    ratio = out["flux64"].ravel() / (w[:-1] - w[1:])
    np.testing.assert_allclose(ratio[:30], n, rtol=1e-5)
    n1 = int(round(ratio[31] / 10))
    np.testing.assert_allclose(ratio[31:53], 10 * n1, rtol=1e-5)
    n2 = int(round(ratio[55] / 100))
    np.testing.assert_allclose(ratio[55:], 100 * n2, rtol=1e-5)
    assert abs(n1 - 0.1 * n) < 4 * math.sqrt(n * 0.09)
    assert abs(n2 - 0.1 * n1) < 4 * math.sqrt(n1 * 0.09)
    # the survival voxels: deposit w_k - 10 w_k e^-0.1 per survivor
    for k, ns in ((30, n1), (53, n2)):
        np.testing.assert_allclose(ratio[k], ns * (1 - 10 * math.exp(-0.1)) / (1 - math.exp(-0.1)) *
                                   (1 if k == 30 else 10), rtol=1e-4)


# ---------------- pattern, fourierx, hyperboloid and angle-CDF launches ----------------

JBO = np.float32(0.9998)   # JUST_BELOW_ONE of src/mcx_core.cl


def _pixel(u, n):
    """pixel index of a uniform coordinate u in [0, 1) over n pixels, as the
    reference computes it: (int)(u * JUST_BELOW_ONE * n)"""
    return np.floor(np.float32(u) * JBO * np.float32(n)).astype(int)


def test_pattern_weights_follow_the_pixel():
    """pattern (param1.w x param2.w pixels over the parallelogram param1, param2):
    the launch weight is the pattern value of the pixel under the launch point
    (src/mcx_core.cl:1678-1693); pixels are hit uniformly"""
    pat = np.arange(1, 13, dtype=np.float32).reshape(4, 3, order="F")
    s = _samples(srctype="pattern", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 4], srcparam2=[0, 40, 0, 3],
                 srcpattern=pat)
    u = (s["p"][:, 0].astype(np.float64) - 10) / 40
    v = (s["p"][:, 1].astype(np.float64) - 10) / 40
    ix, iy = _pixel(u, 4), _pixel(v, 3)
    ok = (np.abs(u * 4 - np.rint(u * 4)) > 1e-4) & (np.abs(v * 3 - np.rint(v * 3)) > 1e-4)   # off pixel edges
    assert ok.mean() > 0.99
    np.testing.assert_array_equal(s["w"][ok], pat[ix[ok], iy[ok]])
    counts = np.bincount(ix * 3 + iy, minlength=12)
    assert stats.chisquare(counts).pvalue > PMIN


def test_pattern3d_box_and_weights():
    """pattern3d (a param1.x x param1.y x param1.z box of voxels from srcpos): the
    launch point is uniform in the box and carries the pattern value of its
    voxel (src/mcx_core.cl:1661-1667, :1708-1727)"""
    pat = np.random.default_rng(5).uniform(0.5, 2.0, (4, 3, 2)).astype(np.float32)
    s = _samples(srctype="pattern3d", srcpos=[10, 12, 14], srcparam1=[4, 3, 2, 0], srcpattern=pat)
    d = s["p"].astype(np.float64) - [10, 12, 14]
    for k, n in enumerate((4, 3, 2)):
        _ks(d[:, k] / n, _uniform(0, 1))
    ix, iy, iz = _pixel(d[:, 0] / 4, 4), _pixel(d[:, 1] / 3, 3), _pixel(d[:, 2] / 2, 2)
    ok = np.all(np.abs(d - np.rint(d)) > 1e-4, axis=1)
    np.testing.assert_array_equal(s["w"][ok], pat[ix[ok], iy[ok], iz[ok]])


@pytest.mark.parametrize("kind", ["fourierx", "fourierx2d"])
def test_fourierx_patterns(kind):
    """fourierx / fourierx2d: the parallelogram is param1 and param1.w dir x
    param1/|param1| (length param1.w; src/mcx_core.cl:1776-1787); the weight is
    (cos(2 pi (kx u + ky v + phase)) (1 - m) + 1) / 2 (fourierx, param2 = {kx, ky,
    phase, m}) or (sin(2 pi (kx u + px)) sin(2 pi (ky v + py)) + 1) / 2
    (fourierx2d, param2 = {kx, ky, px, py}) at the launch point (u, v)"""
    p2 = [2.0, 1.0, 0.25, 0.5] if kind == "fourierx" else [2.0, 3.0, 0.1, 0.3]
    s = _samples(srctype=kind, srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 20], srcparam2=p2)
    u = (s["p"][:, 0].astype(np.float64) - 10) / 40
    v = (s["p"][:, 1].astype(np.float64) - 10) / 20     # second edge: 20 along dir x param1 = +y
    assert np.all(s["p"][:, 2] == 0)
    _ks(u, _uniform(0, 1))
    _ks(v, _uniform(0, 1))
    if kind == "fourierx":
        expect = (np.cos(2 * math.pi * (p2[0] * u + p2[1] * v + p2[2])) * (1 - p2[3]) + 1) / 2
    else:
        expect = (np.sin(2 * math.pi * (p2[0] * u + p2[2])) * np.sin(2 * math.pi * (p2[1] * v + p2[3])) + 1) / 2
    np.testing.assert_allclose(s["w"], expect, atol=5e-5)
    assert abs(s["w"].mean() - 0.5) < 4 * s["w"].std() / math.sqrt(N)


@pytest.mark.parametrize("axis", [(0, 0, 1), (0, 0.6, 0.8)])
def test_hyperboloid_gaussian_rays(axis):
    """hyperboloid (param1 = {w, f, z0}): every ray passes the beam axis at its
    closest distance r at f along the axis from srcpos, with r the waist sample
    r = w sqrt(-log(U)/2): P(r <= x) = 1 - exp(-2 x^2 / w^2)
    (src/mcx_core.cl:1820-1852)"""
    w, f, z0 = 3.0, 10.0, 20.0
    c = np.array([30.0, 30.0, 20.0])
    s = _samples(srctype="hyperboloid", srcpos=list(c), srcdir=list(axis), srcparam1=[w, f, z0, 0])
    a = np.array(axis, np.float64)
    p, v = s["p"].astype(np.float64), s["v"].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=2e-6)
    # closest approach of each ray to the axis line c + t a
    w0 = p - c
    b = v @ a
    d_ = (w0 * v).sum(1)
    e = w0 @ a
    t_ray = (b * e - d_) / (1 - b * b)
    t_axis = (e - b * d_) / (1 - b * b)
    gap = np.linalg.norm(w0 + t_ray[:, None] * v - t_axis[:, None] * a, axis=1)
    # (the closest point is ill-conditioned for rays almost parallel to the axis)
    wide = gap > 0.5
    assert wide.mean() > 0.9
    np.testing.assert_allclose(t_axis[wide], f, atol=1e-2)
    _ks(gap, lambda x: 1.0 - np.exp(-2.0 * x * x / (w * w)))


@pytest.mark.parametrize("step", [False, True])
def test_angle_cdf_launch(step):
    """angleinvcdf (theta / pi at nangle equal CDF steps, src/mcx_core.cl:
    2098-2125): interpolated linearly when srcdir.w <= 0, so theta/pi has the
    piecewise-linear CDF through (table_k, k / (nangle - 1)); taken as steps when
    srcdir.w > 0, so each entry is equally likely; the azimuth is uniform"""
    table = np.array([0.0, 0.05, 0.08, 0.12, 0.2, 0.3, 0.45, 0.6], np.float32)
    s = _samples(srctype="pencil", srcpos=[30, 30, 30], srcdir=[0, 0, 1, 1.0 if step else 0.0], angleinvcdf=table)
    v = s["v"].astype(np.float64)
    th = np.arccos(np.clip(v[:, 2], -1, 1)) / math.pi
    if step:
        k = np.argmin(np.abs(th[:, None] - table[None, :].astype(np.float64)), axis=1)
        np.testing.assert_allclose(th, table[k], atol=2e-4)
        assert stats.chisquare(np.bincount(k, minlength=table.size)).pvalue > PMIN
    else:
        _ks(th, lambda x: np.interp(x, table.astype(np.float64), np.linspace(0, 1, table.size)))
    sel = th > 0.02
    _ks(np.arctan2(v[sel, 1], v[sel, 0]), _uniform(-math.pi, math.pi))
