"""Closed-form pins of the oracle's material model (orc_bsdf_eval /
orc_bsdf_sample, oracle/oracle.cpp): the Lambertian lobe of
LambertianBxDF.cpp:27-55, the normalised modified Phong lobe (no reference
counterpart; used by the glossy materials of the batched and cache
workloads), and their combination by BSDF::SampleF's rules
(BSDF.cpp:141-211).

Each check is a property the mathematics fixes, chosen so that a plausible
slip fails it: a wrong normalisation constant ((n+1) for (n+2) or the
reverse) fails the pdf integral or the albedo; a wrong sampling exponent
fails E[cos a] and the chi-square test; a mirror vector reflected about the
wrong axis fails the chi-square test (it is taken about the mirror
direction); a mixture pdf that disagrees with the sampling density biases
the estimator of the directional albedo.
"""
import numpy as np
import pytest

EXPONENTS = (10.0, 60.0, 1000.0)


def _frame(w):
    w = w / np.linalg.norm(w)
    a = np.array([1.0, 0.0, 0.0]) if abs(w[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(a, w)
    u /= np.linalg.norm(u)
    return u, np.cross(w, u), w


def _dirs_about(axis, c, phi):
    """Unit directions at angle acos(c) from axis, azimuth phi (float64)."""
    u, v, w = _frame(np.asarray(axis, np.float64))
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    return (s * np.cos(phi))[:, None] * u + (s * np.sin(phi))[:, None] * v + c[:, None] * w


def _phong_only_pdf(orc, n, wo, wi):
    # the evaluated pdf of the Lambert + Phong mixture is (|z|/pi + phong) / 2
    _, pdf = orc.bsdf_eval((0.5, 0.5, 0.5), (0.5, 0.5, 0.5), n, 1, wo, wi)
    return 2.0 * pdf.astype(np.float64) - np.abs(wi[:, 2]) / np.pi


@pytest.mark.parametrize("n", EXPONENTS)
def test_phong_pdf_integrates_to_one(orc, n):
    # over the sphere, about the mirror direction r: the lobe is zero where
    # cos a <= 0, so the integral is 2 pi * int_0^1 pdf(c) dc; Gauss-Legendre
    # with 600 nodes integrates c^n exactly up to degree 1199
    wo = np.array([0.3, -0.2, 0.9327379])
    wo /= np.linalg.norm(wo)
    r = np.array([-wo[0], -wo[1], wo[2]])
    x, wts = np.polynomial.legendre.leggauss(600)
    c = 0.5 * (x + 1.0)
    wi = _dirs_about(r, c, np.full_like(c, 0.7))
    pdf = _phong_only_pdf(orc, n, np.tile(wo, (len(c), 1)), wi)
    total = 2.0 * np.pi * 0.5 * np.sum(wts * pdf)
    assert abs(total - 1.0) < 2e-4, total


@pytest.mark.parametrize("n", EXPONENTS)
def test_phong_albedo_at_normal_incidence_is_ks(orc, n):
    # wo = z: r = z and the whole lobe lies above the horizon, so
    # int f cos dw = ks (n+2)/(2 pi) * 2 pi int_0^1 c^(n+1) dc = ks
    ks = (0.25, 0.5, 0.75)
    x, wts = np.polynomial.legendre.leggauss(600)
    c = 0.5 * (x + 1.0)
    wi = _dirs_about((0.0, 0.0, 1.0), c, np.full_like(c, 1.3))
    f, _ = orc.bsdf_eval((0.0, 0.0, 0.0), ks, n, 1, np.tile((0.0, 0.0, 1.0), (len(c), 1)), wi)
    albedo = 2.0 * np.pi * 0.5 * np.sum(wts[:, None] * f.astype(np.float64) * c[:, None], axis=0)
    np.testing.assert_allclose(albedo, ks, rtol=2e-4)


def _phong_samples(orc, n, wo, count, seed):
    rng = np.random.default_rng(seed)
    uvw = rng.random((count, 3), dtype=np.float32)
    uvw[:, 2] = 0.5 + 0.5 * uvw[:, 2]          # w in [0.5, 1): the Phong lobe (lobe index 1 of 2)
    uvw[:, 2] = np.minimum(uvw[:, 2], np.float32(0.99999))
    wi, f, pdf = orc.bsdf_sample((0.5, 0.5, 0.5), (0.5, 0.5, 0.5), n, 1, np.tile(wo, (count, 1)), uvw)
    return wi.astype(np.float64), f, pdf


@pytest.mark.parametrize("n", EXPONENTS)
def test_phong_sampling_mean_cosine(orc, n):
    # cos a ~ density (n+1) c^n on [0, 1]: E[cos a] = (n+1)/(n+2),
    # Var = (n+1)/(n+3) - E^2
    wo = np.array([0.5, 0.1, 0.8602325])
    wo /= np.linalg.norm(wo)
    r = np.array([-wo[0], -wo[1], wo[2]])
    wi, _, _ = _phong_samples(orc, n, wo.astype(np.float32), 200000, 11)
    c = wi @ r
    mean, var = (n + 1) / (n + 2), (n + 1) / (n + 3) - ((n + 1) / (n + 2)) ** 2
    se = np.sqrt(var / len(c))
    assert abs(c.mean() - mean) < 5 * se + 1e-6, (c.mean(), mean, se)


@pytest.mark.parametrize("n", EXPONENTS)
def test_phong_sampling_chi_square(orc, n):
    # 20 equal-probability bins of cos a (CDF c^(n+1)) x 8 azimuth bins
    # about the mirror direction: 160 cells, 159 degrees of freedom
    from scipy import stats
    wo = np.array([-0.4, 0.6, 0.6928203])
    wo /= np.linalg.norm(wo)
    r = np.array([-wo[0], -wo[1], wo[2]])
    count = 160000
    wi, _, _ = _phong_samples(orc, n, wo.astype(np.float32), count, 12)
    u, v, w = _frame(r)
    c = np.clip(wi @ w, 0.0, 1.0)
    phi = np.mod(np.arctan2(wi @ v, wi @ u), 2 * np.pi)
    cb = np.minimum((c ** (n + 1) * 20).astype(int), 19)
    pb = np.minimum((phi / (2 * np.pi) * 8).astype(int), 7)
    hist = np.bincount(cb * 8 + pb, minlength=160)
    chi2 = float(np.sum((hist - count / 160) ** 2 / (count / 160)))
    p = stats.chi2.sf(chi2, 159)
    assert p > 1e-4, (chi2, p)


def test_sampled_pdf_equals_evaluated_pdf(orc):
    # the pdf SampleF reports equals the mixture pdf evaluated at its direction
    wo = np.array([0.2, 0.3, 0.9327379], np.float32)
    rng = np.random.default_rng(3)
    uvw = rng.random((4000, 3), dtype=np.float32)
    wi, f, pdf = orc.bsdf_sample((0.4, 0.3, 0.2), (0.3, 0.3, 0.3), 60.0, 1, np.tile(wo, (4000, 1)), uvw)
    f2, pdf2 = orc.bsdf_eval((0.4, 0.3, 0.2), (0.3, 0.3, 0.3), 60.0, 1, np.tile(wo, (4000, 1)), wi)
    ok = pdf > 0
    np.testing.assert_allclose(pdf[ok], pdf2[ok], rtol=1e-5)
    np.testing.assert_allclose(f[ok], f2[ok], rtol=1e-5)


@pytest.mark.parametrize("kind,n", [(0, 10.0), (1, 10.0), (1, 60.0)])
def test_importance_sampled_albedo_matches_quadrature(orc, kind, n):
    # E[f |cos| / pdf] over SampleF's samples equals int f |cos| dw by
    # quadrature, at an oblique wo where part of the Phong lobe falls below
    # the horizon (f = 0 there): this holds only if the reported pdf is the
    # density the directions are drawn from
    kd, ks = (0.5, 0.4, 0.3), (0.3, 0.3, 0.3)
    wo = np.array([0.6, 0.0, 0.8], np.float32)
    # quadrature over the upper hemisphere in (cos theta, phi)
    x, wts = np.polynomial.legendre.leggauss(200)
    ct = 0.5 * (x + 1.0)
    phis = (np.arange(720) + 0.5) * (2 * np.pi / 720)
    C, P = np.meshgrid(ct, phis, indexing="ij")
    st = np.sqrt(1 - C ** 2)
    wi = np.stack([st * np.cos(P), st * np.sin(P), C], -1).reshape(-1, 3)
    f, _ = orc.bsdf_eval(kd, ks, n, kind, np.tile(wo, (len(wi), 1)), wi)
    wgt = (0.5 * wts)[:, None] * np.full_like(P, 2 * np.pi / 720)
    quad = np.sum((f.astype(np.float64) * wi[:, 2:3]) * wgt.reshape(-1, 1), axis=0)
    if kind == 0:
        np.testing.assert_allclose(quad, kd, rtol=1e-4)   # Lambertian: albedo = kd
    rng = np.random.default_rng(7)
    m = 400000
    uvw = rng.random((m, 3), dtype=np.float32)
    wi_s, f_s, pdf_s = orc.bsdf_sample(kd, ks, n, kind, np.tile(wo, (m, 1)), uvw)
    f_s, wi_s, pdf_s = f_s.astype(np.float64), wi_s.astype(np.float64), pdf_s.astype(np.float64)
    est = np.where(pdf_s[:, None] > 0, f_s * np.abs(wi_s[:, 2:3]) / np.maximum(pdf_s[:, None], 1e-30), 0.0)
    mean, se = est.mean(axis=0), est.std(axis=0) / np.sqrt(m)
    assert np.all(np.abs(mean - quad) < 5 * se + 1e-4), (mean, quad, se)
