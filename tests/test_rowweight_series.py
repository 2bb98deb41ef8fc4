"""The FP32 row-weight pass's separation series and row-weight identities,
checked in float64 against the definitions they replace.

The reference weights a rule point at separation s between grid nodes
sep[i-1] <= s < sep[i] by t = (s - sep[i-1]) / (sep[i] - sep[i-1]) to node
i and 1 - t to node i-1 (convolve_pixel_psf_logsep.hpp:46-54), with the
point's separation s = 2 asin(|p - src| / 2) from the bilinear corner blend
(points_weights.hpp:37-51).  The GPU pass does not form s per point: it
evaluates sigma = (s - s1) / (s1 - s0) as a polynomial in x = y / y_c - 1
(xts_modelmap.hip, poly_pixel) and sums w sigma, w |sigma| and
w |sigma - e^step|, from which poly_rows recovers the row weights.  These
tests restate poly_pixel's coefficients and the remainder bound behind
kPolyReach in float64 and compare them with s computed by arcsin, and
check poly_rows against per-point hat weights.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "xtsrcmaps_amd", "csrc", "xts_modelmap.hip")
STEP = math.log(7e5) / 399.0            # log spacing of grid nodes 1 .. 400
R2D = 180.0 / math.pi


def _source_constants():
    text = open(SRC).read()
    def arr(name):
        m = re.search(name + r"\[[^\]]*\]\s*=\s*\{([^}]*)\}", text)
        return [float(v.rstrip("f")) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    return {"gq": arr("kGQuartic"), "reach": arr("kPolyReach")}


C = _source_constants()


def sep_node(k):
    """Grid node k >= 1 in degrees (psf/psf.hxx:24-31)."""
    return 1e-4 * math.exp((k - 1) * STEP)


def pixel(sc_deg, size_deg, az, tilt):
    """A square pixel of side size_deg centred sc_deg from a source at the
    pole, rotated by tilt: returns the corner offsets D[4][3] (LL LR UR UL)."""
    src = np.array([0.0, 0.0, 1.0])
    th = math.radians(sc_deg)
    c = np.array([math.sin(th) * math.cos(az), math.sin(th) * math.sin(az), math.cos(th)])
    e1 = np.cross(src, c)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(c, e1)
    h = math.radians(size_deg) / 2
    ct, st = math.cos(tilt), math.sin(tilt)
    corners = []
    for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
        da, db = a * ct - b * st, a * st + b * ct
        q = c + h * (da * e1 + db * e2)
        corners.append(q / np.linalg.norm(q))
    return np.array(corners) - src


def geometry(D):
    """pixel_geom: A, B, C, E of p(u, v) / 2 and the bounds x0, hr."""
    A = 0.125 * (D[0] + D[1] + D[2] + D[3])
    B = 0.125 * ((D[1] - D[0]) + (D[2] - D[3]))
    Cv = 0.125 * ((D[3] - D[0]) + (D[2] - D[1]))
    E = 0.125 * ((D[0] - D[1]) + (D[2] - D[3]))
    x0 = np.linalg.norm(A)
    hr = max(np.linalg.norm(0.5 * D[k] - A) for k in range(4))
    hr = hr * 1.0001 + 1e-5 * x0
    return A, B, Cv, E, x0, hr


def asin_series():
    """kAsinSeries as written in the source (fractions)."""
    text = open(SRC).read()
    m = re.search(r"kAsinSeries\[[^\]]*\]\s*=\s*\{([^}]*)\}", text)
    vals = []
    for v in m.group(1).replace("\n", " ").split(","):
        num, _, den = v.strip().partition("/")
        vals.append(float(num) / (float(den) if den else 1.0))
    return vals


def poly_coeffs(yc, s1, D):
    """poly_pixel's sigma(x) coefficients c_0 .. c_D (float64), expanded
    about y = yc, for row r0+1 at separation s1 (deg)."""
    g1, g2, g3, g4 = C["gq"]
    G = sum(a * yc ** n for n, a in enumerate(asin_series()))
    em = math.sqrt(yc) * G * (2 * R2D / s1) - 1.0
    inv_d0 = 1.0 / (1.0 - math.exp(-STEP))
    K = (1.0 + em) * inv_d0
    ga = yc * (g1 + yc * (2 * g2 + yc * (3 * g3 + yc * 4 * g4)))
    gb = yc * yc * (g2 + yc * (3 * g3 + yc * 6 * g4))
    gc = yc ** 3 * (g3 + yc * 4 * g4)
    ex = [1.0, ga, gb + 0.5 * ga * ga, gc + ga * gb + ga ** 3 / 6]
    bs = [1.0, 0.5, -0.125, 0.0625, -0.0390625, 0.02734375, -0.0205078125]
    c = [em * inv_d0]
    for k in range(1, D + 1):
        c.append(K * sum(bs[k - j] * ex[j] for j in range(0, min(k, 3) + 1)))
    return c


def gl_points(n):
    x, w = np.polynomial.legendre.leggauss(n)
    u, v = np.meshgrid(x, x, indexing="ij")
    ww = np.outer(w, w)
    return u.ravel(), v.ravel(), ww.ravel() / 4


def reach(x0, hr):
    rho = hr / x0
    return rho * (2 + rho)


def degree_for(r):
    for d, lim in enumerate(C["reach"], start=2):
        if r <= lim:
            return d
    return None


CASES = [(sc, size) for sc in (0.8, 1.2, 2.0, 3.5, 6.0, 12.0, 25.0) for size in (0.05, 0.1, 0.2)]


@pytest.mark.parametrize("sc,size", CASES)
def test_sigma_series_matches_arcsin(sc, size):
    """sigma(x) from the series agrees with (s - s1) / (s1 - s0) computed by
    arcsin at every 8x8 Gauss-Legendre point of the pixel to 1.5e-7 (row
    spacings) wherever kPolyReach admits the pixel."""
    rng = np.random.default_rng(int(sc * 100 + size * 1000))
    for _ in range(4):
        D4 = pixel(sc, size, rng.uniform(0, 2 * np.pi), rng.uniform(0, np.pi / 2))
        A, B, Cv, E, x0, hr = geometry(D4)
        d = degree_for(reach(x0, hr))
        if d is None:
            continue
        u, v, _ = gl_points(8)
        p = A[None] + u[:, None] * B[None] + v[:, None] * (Cv[None] + u[:, None] * E[None])
        y = np.sum(p * p, axis=1)
        s = 2 * R2D * np.arcsin(np.sqrt(y))
        yc = float(A @ A)
        assert np.all(np.abs(y / yc - 1) <= reach(x0, hr))
        # the pixel's first row as the kernel picks it: ilo - 1 for the
        # interval of the lowest separation
        slo = 2 * R2D * math.asin(x0 - hr)
        r0 = int(math.floor(math.log(slo / 1e-4) / STEP)) + 1
        s0, s1 = sep_node(r0), sep_node(r0 + 1)
        c = poly_coeffs(yc, s1, d)
        x = y / yc - 1
        sig = np.polyval(c[::-1], x)
        exact = (s - s1) / (s1 - s0)
        assert np.max(np.abs(sig - exact)) < 1.5e-7, (sc, size, d)


def test_asin_series_table():
    """kAsinSeries holds a_n = C(2n, n) / (4^n (2n + 1)) and reaches 2e-13 at
    the largest separation the series path takes."""
    a = asin_series()
    ref = [math.comb(2 * n, n) / (4 ** n * (2 * n + 1)) for n in range(len(a))]
    np.testing.assert_allclose(a, ref, rtol=1e-15)
    y = math.sin(math.radians(14.25)) ** 2
    G = sum(c * y ** n for n, c in enumerate(a))
    assert abs(G - math.asin(math.sqrt(y)) / math.sqrt(y)) < 2e-13


def test_fast_pixels_within_reach():
    """A pixel the pass keeps in registers spans at most three row
    intervals (shi / slo < e^(3 step)), so its x range lies within the
    reach of the degree-6 series."""
    rho = (math.exp(3 * STEP) - 1) / (math.exp(3 * STEP) + 1)   # (1+rho)/(1-rho) = e^(3 step)
    assert rho * (2 + rho) < C["reach"][-1]


def test_reach_table_is_the_remainder_bound():
    """kPolyReach[D - 2] is where 1.1 K |binom(1/2, D + 1)| r^(D+1) / (1 - r)
    reaches 1e-7, K = e^(2 step) / (1 - e^-step)."""
    K = math.exp(2 * STEP) / (1 - math.exp(-STEP))
    for d, lim in enumerate(C["reach"], start=2):
        b = abs(math.comb(2 * (d + 1), d + 1) / (4 ** (d + 1) * (2 * (d + 1) - 1)))
        bound = 1.1 * K * b * lim ** (d + 1) / (1 - lim)
        assert 0.95e-7 < bound <= 1.0e-7, (d, bound)


def hat_rows(sig, w, G):
    """Per-point linear interpolation in s over rows r0 .. r0+3, in units
    of s1 - s0 (row r0 at sigma = -1, r0+1 at 0, r0+2 at G, r0+3 at G + G^2)."""
    nodes = np.array([-1.0, 0.0, G, G + G * G])
    W = np.zeros(4)
    for sg, wq in zip(sig, w):
        k = int(np.searchsorted(nodes, sg, side="right")) - 1
        k = min(max(k, 0), 2)
        t = (sg - nodes[k]) / (nodes[k + 1] - nodes[k])
        W[k] += wq * (1 - t)
        W[k + 1] += wq * t
    return W


def poly_rows(wtot, A, B1, B2, G, nm):
    P1 = 0.5 * (B1 + A)
    W0 = 0.5 * (B1 - A)
    if nm < 2:
        W2, W3 = P1 / G, 0.0
    else:
        P2 = 0.5 * (B2 + A - G * wtot)
        W2, W3 = P1 / G - P2 * (1 / G + 1 / G ** 2), P2 / G ** 2
    return np.array([W0, wtot - W0 - W2 - W3, W2, W3])


@pytest.mark.parametrize("nm", [1, 2])
def test_poly_rows_recover_hat_weights(nm):
    G = math.exp(STEP)
    rng = np.random.default_rng(5 + nm)
    hi = G if nm == 1 else G + G * G
    for _ in range(20):
        n = int(rng.integers(1, 80))
        sig = rng.uniform(-1, hi, n)
        if rng.random() < 0.3:   # all on one side of a node
            sig = rng.uniform(0, G, n) if nm == 2 else rng.uniform(-1, 0, n)
        w = rng.uniform(0.01, 1, n)
        A, B1, B2 = np.sum(w * sig), np.sum(w * np.abs(sig)), np.sum(w * np.abs(sig - G))
        got = poly_rows(w.sum(), A, B1, B2, G, nm)
        np.testing.assert_allclose(got, hat_rows(sig, w, G), rtol=0, atol=1e-12)
