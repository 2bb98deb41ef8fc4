"""Synthetic inputs of the benchmark configurations (fixed seeds).

All images are CAR with CRVAL2 = 0 (plate carree), so pixel <-> sky is the
closed form lon = crval1 + cdelt1 (p1 - crpix1), lat = cdelt2 (p2 - crpix2).
The King-PSF and effective-area IRF grids are built in the reference's
padded layout (irf/irf.prepare_grid.cpp:26-102: bin centres plus -1/+1 in
cos theta and 0/10 in log10 E, edge-duplicated parameters).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .capi import Wcs


@dataclass
class Config:
    name: str
    n_src: int
    npix: int
    cdelt: float
    n_energy: int = 37
    emin: float = 100.0
    emax: float = 1e6
    rule: str = "deg7"          # "deg7" (reference production rule) or "gl8"
    seed: int = 0
    out_f32: bool = True
    src_mode: str = "centre"    # "centre" or "uniform"
    extra: dict = field(default_factory=dict)

    @property
    def evals(self) -> int:
        return self.n_src * self.n_energy * self.npix * self.npix


CONFIGS = {
    # one source at the centre of a 100x100 0.1 deg image, FP64 output
    "c0": Config("c0", 1, 100, 0.1, rule="deg7", seed=0, out_f32=False, src_mode="centre"),
    # 100 sources uniform over a 200x200 0.05 deg image, 8x8 Gauss-Legendre, FP32
    "c1": Config("c1", 100, 200, 0.05, rule="gl8", seed=1, out_f32=True, src_mode="uniform"),
    # the 2,000-source 400x400 0.05 deg, 37-energy workload shape (the size of
    # the larger configurations) through the plain cubature, FP32: 1.18e10
    # evaluations, a 47 GB map
    "c3p": Config("c3p", 2000, 400, 0.05, rule="gl8", seed=3, out_f32=True, src_mode="uniform"),
    # the surrogate workload: the same 2,000 sources on 400x400 at 0.05 deg,
    # 37 energies, the 8x8 Gauss-Legendre rule, FP32 map
    "c3": Config("c3", 2000, 400, 0.05, rule="gl8", seed=3, out_f32=True, src_mode="uniform"),
    # the large sky: 10,000 sources with fluxes log-uniform over 4 decades,
    # uniform over a 1000x1000 image at 0.02 deg (20 x 20 deg), 100 energies,
    # 8x8 Gauss-Legendre rule; output the flux-weighted summed model
    # [100, 1000, 1000] FP32 (the per-source maps would be 4 TB)
    "c4": Config("c4", 10000, 1000, 0.02, n_energy=100, rule="gl8", seed=4, out_f32=True,
                 src_mode="uniform", extra={"flux_decades": 4}),
    # the large sky's sources and physics on a 160x160 window at 0.02 deg
    # (12 sources), for oracle comparisons of the summed model
    "c4w": Config("c4w", 12, 160, 0.02, n_energy=100, rule="gl8", seed=4, out_f32=True,
                  src_mode="uniform", extra={"flux_decades": 4}),
}


def wcs_for(cfg: Config) -> Wcs:
    n = cfg.npix
    return Wcs.make((n, n), (n / 2 + 0.5, n / 2 + 0.5), (180.0, 0.0), (-cfg.cdelt, cfg.cdelt),
                    0.0, "CAR")


def pix2sph_car(wcs: Wcs, p1, p2):
    assert wcs.crval2 == 0.0 and wcs.crota2 == 0.0 and wcs.proj == 0
    lon = wcs.crval1 + wcs.cdelt1 * (np.asarray(p1) - wcs.crpix1)
    lat = wcs.cdelt2 * (np.asarray(p2) - wcs.crpix2)
    return np.mod(lon, 360.0), lat


def gauss_legendre_rule(n: int) -> np.ndarray:
    """n x n tensor Gauss-Legendre rule on [-1, 1]^2 as (u, v, w) rows, sum w = 4."""
    x, w = np.polynomial.legendre.leggauss(n)
    u, v = np.meshgrid(x, x, indexing="ij")
    ww = np.outer(w, w)
    return np.stack([u.ravel(), v.ravel(), ww.ravel()], axis=1)


def rule_for(cfg: Config):
    """None selects the library's built-in degree-7 rule."""
    return None if cfg.rule == "deg7" else gauss_legendre_rule(int(cfg.rule[2:]))


def energies(cfg: Config) -> np.ndarray:
    k = np.arange(cfg.n_energy)
    return cfg.emin * (cfg.emax / cfg.emin) ** (k / (cfg.n_energy - 1))


def sources(cfg: Config, wcs: Wcs) -> np.ndarray:
    rng = np.random.default_rng(cfg.seed + 1000)
    if cfg.src_mode == "centre":
        p = np.full((cfg.n_src, 2), cfg.npix / 2 + 0.5)
    else:
        p = rng.uniform(0.5, cfg.npix + 0.5, (cfg.n_src, 2))
    lon, lat = pix2sph_car(wcs, p[:, 0], p[:, 1])
    return np.stack([lon, lat], axis=1)


def fluxes(cfg: Config) -> np.ndarray:
    """Per-source flux, log-uniform over cfg.extra["flux_decades"] decades
    (1 for every source without it)."""
    dec = cfg.extra.get("flux_decades", 0)
    if not dec:
        return np.ones(cfg.n_src)
    rng = np.random.default_rng(cfg.seed + 2000)
    return 10.0 ** rng.uniform(-dec, 0.0, cfg.n_src)


def _padded_grid(e_edges, c_edges, params_raw):
    """prepare_grid layout: [Mc0 + 2, Me0 + 2, Ng] with padded axes."""
    e_lo, e_hi = e_edges[:-1], e_edges[1:]
    c_lo, c_hi = c_edges[:-1], c_edges[1:]
    cos = np.concatenate([[-1.0], 0.5 * (c_lo + c_hi), [1.0]])
    loge = np.concatenate([[0.0], 0.5 * np.log10(e_lo * e_hi), [10.0]])
    mc0, me0 = len(c_lo), len(e_lo)
    t_ = np.clip(np.arange(mc0 + 2) - 1, 0, mc0 - 1)
    e_ = np.clip(np.arange(me0 + 2) - 1, 0, me0 - 1)
    params = params_raw[np.ix_(t_, e_)]
    return cos, loge, params, float(c_lo[0])


@dataclass
class Inputs:
    cfg: Config
    wcs: Wcs
    src_sph: np.ndarray
    logE: np.ndarray
    costh: np.ndarray
    king_raw: dict           # side -> (params [Mc,Me,6] before normalisation, cos, loge)
    psf_scale: tuple
    aeff_grid: dict          # side -> (cos, loge, params [Mc, Me], minc)
    effic: tuple             # (p0, p1) float32[6]
    cube: np.ndarray         # [12 nside^2, C] livetime
    wcube: np.ndarray
    nside: int
    rule: np.ndarray | None


def make_inputs(cfg: Config) -> Inputs:
    rng = np.random.default_rng(cfg.seed)
    wcs = wcs_for(cfg)
    src = sources(cfg, wcs)
    logE = np.log10(energies(cfg))
    # King PSF IRF: 23 log energy bins x 8 cos theta bins
    e_edges = 10.0 ** (0.75 + 0.25 * np.arange(24))
    c_edges = np.linspace(0.2, 1.0, 9)
    king = {}
    for side in ("front", "back"):
        raw = np.zeros((8, 23, 6))
        raw[..., 0] = 1.0
        raw[..., 1] = rng.uniform(0.05, 0.5, (8, 23))
        raw[..., 2] = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), (8, 23)))
        raw[..., 3] = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), (8, 23)))
        raw[..., 4] = rng.uniform(1.5, 4.0, (8, 23))
        raw[..., 5] = rng.uniform(1.5, 4.0, (8, 23))
        cos, loge, params, _ = _padded_grid(e_edges, c_edges, raw)
        king[side] = (params, cos, loge)
    # Aeff: smooth log-parabola in E times linear in cos theta
    ae_edges = 10.0 ** (0.75 + (6.5 - 0.75) * np.arange(41) / 40)
    ac_edges = np.linspace(0.2, 1.0, 17)
    ec = np.sqrt(ae_edges[:-1] * ae_edges[1:])
    cc = 0.5 * (ac_edges[:-1] + ac_edges[1:])
    shape_e = 0.9 * np.exp(-0.08 * np.log(ec / 3e3) ** 2)
    aeff = {}
    for side, scale in (("front", 1.0), ("back", 0.7)):
        raw = scale * np.outer(0.2 + 0.8 * cc, shape_e)
        cos, loge, params, minc = _padded_grid(ae_edges, ac_edges, raw)
        aeff[side] = (cos, loge, params, minc)
    p0 = np.array([-1.2, 3.9, -0.5, 2.3, -0.07, 3.5], np.float32)
    p1 = np.array([1.9, -4.2, 0.47, 2.3, 0.066, 3.5], np.float32)
    nside = 32
    cube = rng.uniform(0.0, 1e8, (12 * nside * nside, 40))
    wcube = 0.3 * cube
    costh = 1.0 - ((np.arange(40) + 0.5) / 40) ** 2      # cosmin 0, sqrt binning
    return Inputs(cfg, wcs, src, logE, costh, king, (0.0, 1.0, 0.0), aeff, (p0, p1), cube,
                  wcube, nside, rule_for(cfg))


# ------------------------------------------------------------- several images
@dataclass
class BatchConfig:
    """Sources over several CAR images (plate carree, CRVAL2 = 0, centred at
    different longitudes), one shared instrument and livetime cube."""
    name: str
    rois: tuple              # ((npix, n_src), ...)
    cdelt: float = 0.05
    n_energy: int = 37
    rule: str = "gl8"
    seed: int = 2
    lon0: tuple = (40.0, 110.0, 200.0, 300.0)

    @property
    def n_src(self) -> int:
        return sum(n for _, n in self.rois)

    def dense_evals(self) -> int:
        return sum(n * p * p for p, n in self.rois) * self.n_energy


BATCH_CONFIGS = {
    # 2,000 sources across four images of 64^2, 128^2, 256^2 and 400^2 pixels
    # at 0.05 deg, in proportion to their areas
    "c2": BatchConfig("c2", ((64, 33), (128, 133), (256, 532), (400, 1302))),
}


@dataclass
class BatchInputs:
    cfg: BatchConfig
    wcs: list                # [n_images] Wcs
    src_begin: np.ndarray    # [n_images + 1] int64
    phys: Inputs             # src_sph = every source, ordered by image


def make_batch_inputs(cfg: BatchConfig) -> BatchInputs:
    rng = np.random.default_rng(cfg.seed + 1000)
    wcs, srcs, begin = [], [], [0]
    for (n, ns), lon0 in zip(cfg.rois, cfg.lon0):
        w = Wcs.make((n, n), (n / 2 + 0.5, n / 2 + 0.5), (lon0, 0.0), (-cfg.cdelt, cfg.cdelt),
                     0.0, "CAR")
        p = rng.uniform(0.5, n + 0.5, (ns, 2))
        lon, lat = pix2sph_car(w, p[:, 0], p[:, 1])
        srcs.append(np.stack([lon, lat], axis=1))
        wcs.append(w)
        begin.append(begin[-1] + ns)
    # the instrument and livetime of a single-image configuration with the
    # same seed and energies; its sources replaced by every image's
    base = make_inputs(Config(cfg.name, 1, 16, cfg.cdelt, n_energy=cfg.n_energy, rule=cfg.rule,
                              seed=cfg.seed, src_mode="uniform"))
    phys = Inputs(**{**base.__dict__, "src_sph": np.concatenate(srcs)})
    return BatchInputs(cfg, wcs, np.array(begin, np.int64), phys)


# ------------------------------------------------------------- diffuse sources
@dataclass
class DiffuseConfig:
    """A diffuse-source model map: an all-sky galactic CAR cube of the
    galactic diffuse MapCube's shape (2880 x 1441 at 0.125 deg, 28 energy
    planes) under a celestial CAR model map, its grid resampled to
    target_pix_size."""
    name: str
    npix: int = 400
    cdelt: float = 0.05
    crval: tuple = (180.0, 0.0)
    cube_shape: tuple = (2880, 1441)
    cube_cdelt: float = 0.125
    planes: int = 28
    n_energy: int = 37
    target: float = 0.05
    seed: int = 7

    @property
    def values(self) -> int:
        return self.n_energy * self.npix * self.npix


DIFFUSE_CONFIGS = {
    # the model map of the 2,000-source workload (400 x 400 at 0.05 deg,
    # 37 energies) under the galactic diffuse cube
    "diffuse": DiffuseConfig("diffuse"),
    # the same on a 100 x 100 image at 0.2 deg over a 1 deg cube with 8
    # planes (oracle-sized)
    "diffuse_small": DiffuseConfig("diffuse_small", npix=100, cdelt=0.2, cube_shape=(360, 181),
                                   cube_cdelt=1.0, planes=8, n_energy=6, target=0.5),
}


def diffuse_inputs(cfg: DiffuseConfig):
    """(roi Wcs, cube Wcs, cube float32 [planes, N2, N1], upsf [401, E],
    exposure [E]): a cube bright along the galactic plane with a power-law
    spectrum and pixel-scale structure, a King-like PSF narrowing with
    energy."""
    rng = np.random.default_rng(cfg.seed)
    n = cfg.npix
    roi = Wcs.make((n, n), (n / 2 + 0.5, n / 2 + 0.5), cfg.crval, (-cfg.cdelt, cfg.cdelt), 0.0,
                   "CAR")
    N1, N2 = cfg.cube_shape
    cube_wcs = Wcs.make((N1, N2), ((N1 + 1) / 2, (N2 + 1) / 2), (0.0, 0.0),
                        (-cfg.cube_cdelt, cfg.cube_cdelt), 0.0, "CAR", galactic=True)
    lon = cube_wcs.cdelt1 * (np.arange(1, N1 + 1) - cube_wcs.crpix1)
    lat = cube_wcs.cdelt2 * (np.arange(1, N2 + 1) - cube_wcs.crpix2)
    prof_b = np.exp(-np.abs(lat) / 3.0) + 0.2 * np.exp(-np.abs(lat) / 25.0)
    prof_l = 1.0 + 0.5 * np.cos(np.radians(lon))
    spec = (10.0 ** np.linspace(2, 6, cfg.planes) / 100.0) ** -2.7
    tex = (1.0 + 0.1 * rng.standard_normal((N2, N1))).astype(np.float32)
    cube = (spec[:, None, None] * (prof_b[:, None] * prof_l[None, :] * tex)[None]).astype(
        np.float32)
    sep = np.zeros(401)
    sep[1:] = 1e-4 * np.exp(np.arange(400) * np.log(7e5) / 399.0)
    energy = 10.0 ** np.linspace(2, 6, cfg.n_energy)
    sigma = np.maximum(3.0 * (energy / 100.0) ** -0.8, 0.05)
    g = 2.5
    upsf = (1.0 + (sep[:, None] / sigma[None, :]) ** 2 / (2 * g)) ** -g
    expo = 3e10 * (0.5 + 0.5 * np.log10(energy) / 6.0)
    return roi, cube_wcs, cube, upsf, expo
