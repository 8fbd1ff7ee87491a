"""Pins of the oracle's Microfacet BSDF (the metric workload's material) by
facts the mathematics fixes, independently of the oracle's own formulas.

Reference: LibTensorRay/Renderer/Microfacet.cpp:37-165 (EvalInner,
SampleOnly, PdfInner), BSDF.cpp:232-302 (GGX_D, Smith_G, GGX_VNDF_Pdf,
GGX_VNDF_Sample).  What each test fixes:

* the GGX normal distribution with alpha = r^2 integrates to one in its
  projected form, int D cos(theta_h) d(omega) = 1 (Walter et al. 2007);
* Schlick's Fresnel in the spherical-Gaussian form has F(vh = 0) = 1,
  F(vh = 1) = F0 + (1 - F0) 2^-12.38633, and stays within 0.009 of Schlick's
  (1 - vh)^5 (Karis 2013, "Real Shading in Unreal Engine 4");
* Smith-Schlick G1 with k = (r + 1)^2 / 8: G1(1) = 1, G1(0) = 0, and
  G1(1/2) = 2/3 at r = 1;
* the full eval equals the textbook Cook-Torrance product computed here in
  float64 (Walter's tan form of D, a different algebraic route from the
  reference's alpha / (cos^2 (alpha^2 - 1) + 1)), is reciprocal, vanishes
  below either horizon, and its directional albedo is at most one;
* the VNDF sampler: Heitz's distribution of visible normals D_v(h) =
  G1(v) max(0, v.h) D(h) / v_z (alpha = r; Heitz 2018, JCGT 7(4)) is pinned
  by int D_v d(omega) = 1 and by a chi-square test of Heitz's own listing;
  viewed along the normal the oracle's sampler passes the same chi-square
  test; for tilted views it equals Heitz's listing sample by sample except
  for one step the reference writes differently -- it takes the hemisphere
  height from the disk point before the lerp (BSDF.cpp:292-296) -- so the
  reference, and the oracle with it, does not sample the VNDF there;
* the reference's GGX_VNDF_Pdf is D(h) Smith_G(v, h) |v.h| / |v_z| with
  the reference's Smith_G, whose closed-form values are pinned at designed
  points.  That Smith_G takes tan^2 of the half vector and squares
  alpha^2 tan^2 (BSDF.cpp:243-256), so the pdf the reference divides by is
  not the density its sampler draws from; the oracle reproduces this, and
  the test records the size of the difference;
* BSDF sampling returns the pdf that BSDF::Pdf gives for the sampled
  direction (the reflection Jacobian 1 / (4 wi.wh), Walter 2007 eq. 14).
"""
import numpy as np
import pytest
from scipy import stats

import oracle

MICROFACET = 1
ROUGH = (0.1, 0.3, 0.7)
# chi-square acceptance per test: the file runs about a dozen such tests,
# so p > 0.001 each keeps the family-wise false-alarm rate near 1 %
P_MIN = 1e-3


def _dirs(theta, phi):
    theta = np.asarray(theta, np.float64)
    phi = np.asarray(phi, np.float64)
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


# ---- textbook forms in float64 ----
def ggx_walter(cos_h, alpha):
    """D(h) = 1 / (pi alpha^2 cos^4 (1 + tan^2 / alpha^2)^2) (Walter 2007, eq. 33)."""
    c2 = cos_h * cos_h
    t2 = (1.0 - c2) / c2
    return 1.0 / (np.pi * alpha * alpha * c2 * c2 * (1.0 + t2 / (alpha * alpha)) ** 2)


def smith_g1_ggx(cos_v, alpha):
    """Exact Smith G1 of GGX: 2 / (1 + sqrt(1 + alpha^2 tan^2 theta_v))."""
    t2 = (1.0 - cos_v * cos_v) / (cos_v * cos_v)
    return 2.0 / (1.0 + np.sqrt(1.0 + alpha * alpha * t2))


def vndf_heitz(v, h, alpha):
    """Distribution of visible normals D_v(h) = G1(v) max(0, v.h) D(h) / v_z."""
    vh = np.maximum((v * h).sum(-1), 0.0)
    return smith_g1_ggx(v[..., 2], alpha) * vh * ggx_walter(h[..., 2], alpha) / v[..., 2]


def cook_torrance(wo, wi, r, f0, kd=0.0):
    """kd / pi + D F G / (4 nv nl + 1e-6) with D = GGX (alpha = r^2), F =
    Schlick-SG, G = Smith-Schlick (k = (r + 1)^2 / 8); zero below either horizon."""
    nv, nl = wo[..., 2], wi[..., 2]
    h = wo + wi
    h = h / np.linalg.norm(h, axis=-1, keepdims=True)
    vh = (wo * h).sum(-1)
    D = ggx_walter(h[..., 2], r * r)
    F = f0 + (1.0 - f0) * 2.0 ** ((-5.55473 * vh - 6.8316) * vh)
    k = (r + 1.0) ** 2 / 8.0
    G = nv / (nv * (1.0 - k) + k) * nl / (nl * (1.0 - k) + k)
    f = kd / np.pi + D * F * G / (4.0 * np.abs(nv) * np.abs(nl) + 1e-6)
    return np.where((nv > 0) & (nl > 0), f, 0.0)


@pytest.mark.parametrize("r", ROUGH)
def test_ggx_projected_area_is_one(r):
    """int D cos(theta_h) d(omega) = 1 for the oracle's D, alpha = r^2.  With
    x = tan^2 theta / alpha^2 and y = x / (1 + x), the exact integrand is 1
    on [0, 1), so the Gauss-Legendre sum of the oracle's values is 1 up to
    float32 rounding of cos^2 near the normal."""
    a = r * r
    y, w = np.polynomial.legendre.leggauss(4000)
    y = 0.5 * (y + 1.0)
    w = 0.5 * w
    x = y / (1.0 - y)
    cos_h = 1.0 / np.sqrt(1.0 + a * a * x)
    D, _, _ = oracle.microfacet_parts(cos_h, np.zeros_like(cos_h), np.ones_like(cos_h), r)
    integrand = D.astype(np.float64) * np.pi * a * a / (1.0 + a * a * x) ** 2 / (1.0 - y) ** 2
    total = (integrand * w).sum()
    assert abs(total - 1.0) < (3e-3 if r < 0.2 else 2e-4), total
    # the textbook form gives exactly one in the same quadrature
    ref = (ggx_walter(cos_h, a) * np.pi * a * a / (1.0 + a * a * x) ** 2 / (1.0 - y) ** 2 * w).sum()
    assert abs(ref - 1.0) < 1e-9


def test_schlick_sg_fresnel_end_points_and_accuracy():
    vh = np.linspace(0.0, 1.0, 2001)
    _, e, _ = oracle.microfacet_parts(np.ones_like(vh), vh, np.ones_like(vh), 0.3)
    assert e[0] == 1.0  # F(0) = F0 + (1 - F0) = 1 for any F0
    assert abs(e[-1] - 2.0 ** -12.38633) < 1e-9  # F(1) = F0 + (1 - F0) 2^-12.38633
    assert np.abs(e - (1.0 - vh) ** 5).max() < 0.009
    assert np.all(np.diff(e) <= 0)


@pytest.mark.parametrize("r", ROUGH + (1.0,))
def test_smith_schlick_g1_closed_forms(r):
    c = np.array([1.0, 0.0, 0.5, 0.25], np.float32)
    _, _, g = oracle.microfacet_parts(np.ones(4), np.zeros(4), c, r)
    k = (r + 1.0) ** 2 / 8.0
    assert g[0] == 1.0 and g[1] == 0.0
    assert np.allclose(g[2:], c[2:] / (c[2:] * (1.0 - k) + k), rtol=1e-6)
    if r == 1.0:
        assert abs(g[2] - 2.0 / 3.0) < 1e-7


@pytest.mark.parametrize("r", ROUGH)
def test_microfacet_eval_matches_textbook_cook_torrance(r):
    rng = np.random.default_rng(11)
    n = 4000
    wo = _dirs(np.arccos(rng.uniform(0.02, 1.0, n)), rng.uniform(0, 2 * np.pi, n))
    wi = _dirs(np.arccos(rng.uniform(-0.3, 1.0, n)), rng.uniform(0, 2 * np.pi, n))
    f0 = np.array([0.04, 0.5, 1.0])
    kd = np.array([0.0, 0.2, 0.7])
    got, _ = oracle.bsdf_local(MICROFACET, r, kd, f0, wo, wi)
    wo32 = wo.astype(np.float32).astype(np.float64)
    wi32 = wi.astype(np.float32).astype(np.float64)
    ref = np.stack([cook_torrance(wo32, wi32, r, f0[c], kd[c]) for c in range(3)], -1)
    below = (wo32[:, 2] <= 0) | (wi32[:, 2] <= 0)
    assert np.all(got[below] == 0.0)
    tol = 2e-3 if r < 0.2 else 2e-4  # float32 cancellation in cos^2 (alpha^2 - 1) + 1 near the peak
    err = np.abs(got[~below] - ref[~below]) / np.maximum(ref[~below], 1e-3)
    assert err.max() < tol, err.max()


@pytest.mark.parametrize("r", ROUGH)
def test_microfacet_eval_is_reciprocal(r):
    rng = np.random.default_rng(12)
    n = 2000
    wo = _dirs(np.arccos(rng.uniform(0.05, 1.0, n)), rng.uniform(0, 2 * np.pi, n))
    wi = _dirs(np.arccos(rng.uniform(0.05, 1.0, n)), rng.uniform(0, 2 * np.pi, n))
    f0 = [0.04, 0.3, 0.9]
    a, _ = oracle.bsdf_local(MICROFACET, r, [0.1] * 3, f0, wo, wi)
    b, _ = oracle.bsdf_local(MICROFACET, r, [0.1] * 3, f0, wi, wo)
    assert np.allclose(a, b, rtol=2e-5 if r > 0.2 else 2e-3, atol=1e-6)


@pytest.mark.parametrize("r", (0.3, 0.7, 1.0))
def test_microfacet_directional_albedo_at_most_one(r):
    """int f cos(theta_i) d(omega_i) <= 1 with F0 = 1 (the largest Fresnel) and
    kd = 0, from grazing to normal incidence (Gauss-Legendre in cos, uniform
    in phi; r >= 0.3 so that the quadrature resolves the lobe)."""
    x, w = np.polynomial.legendre.leggauss(300)
    mu = 0.5 * (x + 1.0)
    wmu = 0.5 * w
    nphi = 600
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    M, P = np.meshgrid(mu, phi, indexing="ij")
    wi = _dirs(np.arccos(M), P).reshape(-1, 3)
    for mu_o in (0.05, 0.2, 0.5, 0.8, 1.0):
        wo = np.broadcast_to(_dirs(np.arccos(mu_o), 0.0), wi.shape)
        f, _ = oracle.bsdf_local(MICROFACET, r, [0.0] * 3, [1.0] * 3, wo, wi)
        alb = (f[:, 0].reshape(M.shape) * M * wmu[:, None]).sum() * 2 * np.pi / nphi
        assert 0.0 < alb <= 1.0, (mu_o, alb)


def _vndf_grid_probs(v, alpha, ny=16, nphi=16, sub=64):
    """Probability of each (y, phi) bin under Heitz's D_v, y = x / (1 + x),
    x = tan^2 theta_h / alpha^2.  In y the density has an integrable
    (1 - y)^-1/2 singularity at grazing half vectors, so each bin is
    integrated in t = sqrt(1 - y) (dy = 2 t dt), where it is smooth: a
    sub x sub midpoint rule per bin."""
    edges = np.sqrt(1.0 - np.arange(ny + 1) / ny)  # t at the y edges, decreasing
    frac = (np.arange(sub) + 0.5) / sub
    T = (edges[:-1, None] + (edges[1:] - edges[:-1])[:, None] * frac[None, :]).ravel()
    dT = np.repeat(edges[:-1] - edges[1:], sub) / sub
    ps = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2 * np.pi
    Tm, Pp = np.meshgrid(T, ps, indexing="ij")
    Y = 1.0 - Tm * Tm
    X = Y / (1.0 - Y)
    theta = np.arctan(alpha * np.sqrt(X))
    h = _dirs(theta, Pp)
    # d(omega) = alpha^2 / (2 (1 + alpha^2 x)^1.5) dx dphi, dx = dy / (1 - y)^2, dy = 2 t dt
    jac = alpha * alpha / (2.0 * (1.0 + alpha * alpha * X) ** 1.5) / (1.0 - Y) ** 2 * 2.0 * Tm
    dens = vndf_heitz(np.broadcast_to(v, h.shape), h, alpha) * jac
    cell = dT[:, None] * (2 * np.pi / (nphi * sub))
    return (dens * cell).reshape(ny, sub, nphi, sub).sum(axis=(1, 3))


def _disk_concentric64(u):
    """Shirley-Chiu concentric map of [0,1)^2 onto the unit disk (float64)."""
    x = 2.0 * u[:, 0] - 1.0
    y = 2.0 * u[:, 1] - 1.0
    q13 = np.abs(x) < np.abs(y)
    r = np.where(q13, y, x)
    rp = np.where(q13, x, y)
    with np.errstate(divide="ignore", invalid="ignore"):
        p_ = 0.25 * np.pi * rp / r
    phi = np.where(q13, 0.5 * np.pi - p_, p_)
    phi = np.where((x == 0) & (y == 0), 0.0, phi)
    return r * np.cos(phi), r * np.sin(phi)


def heitz_vndf_vector_form(v, alpha, t1, t2, z_from_unlerped=False):
    """Heitz 2018 (JCGT 7(4), listing in section 4) in float64, given a disk
    point (t1, t2): orthonormal basis around the stretched view, the lerp of
    t2 toward the visible half disk, the hemisphere point, unstretch of the
    normal.  z_from_unlerped = True takes the hemisphere height from the disk
    point before the lerp, the one step in which BSDF.cpp:292-296 differs
    (it computes z from p after assigning the lerped y to a new variable)."""
    Vh = np.array([alpha * v[0], alpha * v[1], v[2]])
    Vh /= np.linalg.norm(Vh)
    lensq = Vh[0] ** 2 + Vh[1] ** 2
    T1 = np.array([-Vh[1], Vh[0], 0.0]) / np.sqrt(lensq) if lensq > 0 else np.array([1.0, 0.0, 0.0])
    T2 = np.cross(Vh, T1)
    s = 0.5 * (1.0 + Vh[2])
    t2l = (1.0 - s) * np.sqrt(np.maximum(1.0 - t1 * t1, 0.0)) + s * t2
    zz = np.sqrt(np.maximum(1.0 - t1 * t1 - (t2 if z_from_unlerped else t2l) ** 2, 0.0))
    Nh = t1[:, None] * T1 + t2l[:, None] * T2 + zz[:, None] * Vh
    if z_from_unlerped:
        # that height can put the point below the tangent plane; the slope
        # form then returns the opposite normal (normalize(-slope, 1) keeps
        # z > 0), where Heitz's listing (which never gets there) clamps z
        Nh = Nh * np.where(Nh[:, 2] < 0.0, -1.0, 1.0)[:, None]
    Ne = np.stack([alpha * Nh[:, 0], alpha * Nh[:, 1], np.maximum(Nh[:, 2], 0.0)], -1)
    return Ne / np.linalg.norm(Ne, axis=1, keepdims=True)


def _chi_square_p(wh, v, alpha):
    probs = _vndf_grid_probs(v, alpha)
    assert abs(probs.sum() - 1.0) < 2e-4  # int D_v d(omega) = 1 pins the density
    n = len(wh)
    c2 = np.clip(wh[:, 2], 1e-12, 1.0) ** 2
    x = (1.0 - c2) / c2 / (alpha * alpha)
    y = x / (1.0 + x)
    phi = np.mod(np.arctan2(wh[:, 1], wh[:, 0]), 2 * np.pi)
    iy = np.minimum((y * 16).astype(int), 15)
    ip = np.minimum((phi / (2 * np.pi) * 16).astype(int), 15)
    obs = np.bincount(iy * 16 + ip, minlength=256).astype(np.float64)
    exp = probs.ravel() * n
    keep = exp >= 5.0
    # bins with fewer than 5 expected samples are pooled into one
    o = np.append(obs[keep], obs[~keep].sum())
    e = np.append(exp[keep], exp[~keep].sum())
    if e[-1] < 5.0:
        o, e = o[:-1], e[:-1]
    chi2 = ((o - e) ** 2 / e).sum()
    return stats.chi2.sf(chi2, len(o) - 1)


@pytest.mark.parametrize("r", ROUGH)
def test_heitz_vector_form_samples_the_vndf(r):
    """The float64 restatement of Heitz's listing draws from D_v (chi-square
    on 10^6 samples): the yardstick of the next two tests is itself pinned."""
    for theta_v in (0.2, 0.9, 1.4):
        v = _dirs(theta_v, 0.7)
        u = np.random.default_rng(2024).random((1_000_000, 2))
        t1, t2 = _disk_concentric64(u)
        wh = heitz_vndf_vector_form(v, r, t1, t2)
        assert np.all(wh @ v > -1e-12)
        assert _chi_square_p(wh, v, r) > P_MIN


@pytest.mark.parametrize("r", ROUGH)
def test_vndf_sampler_draws_the_vndf_at_normal_view(r):
    """Viewed along the normal the lerp is the identity (s = 1) and the
    reference's sampler is exactly Heitz's: chi-square of 10^6 oracle samples
    against D_v with alpha = r."""
    v = np.array([0.0, 0.0, 1.0])
    u = np.random.default_rng(2024).random((1_000_000, 2), dtype=np.float32)
    wh, _ = oracle.vndf(np.broadcast_to(v, (len(u), 3)), r, u)
    assert _chi_square_p(wh.astype(np.float64), v, r) > P_MIN


@pytest.mark.parametrize("r", ROUGH + (1.0,))
@pytest.mark.parametrize("theta_v", (0.2, 0.9, 1.4))
def test_vndf_sampler_is_heitz_with_the_references_hemisphere_height(r, theta_v):
    """Sample by sample, the oracle's GGX_VNDF_Sample equals Heitz's vector
    form fed the same concentric disk point, with the hemisphere height taken
    from the point before the lerp (BSDF.cpp:292-296): this pins the
    stretch, the lerp, the slope conversion, the rotation and the unstretch
    of the slope-form code, which is written differently.  That height makes
    the sampler leave the VNDF for tilted views (it even returns back-facing
    half vectors); the fraction is recorded."""
    v = _dirs(theta_v, 0.7)
    n = 200_000
    u = np.random.default_rng(9).random((n, 2), dtype=np.float32)
    wh, _ = oracle.vndf(np.broadcast_to(v, (n, 3)), r, u)
    t1, t2 = _disk_concentric64(u.astype(np.float64))
    # the slope form returns the T1 component as +x (Heitz's normal has -x
    # there): a mirror across the view plane, which leaves the distribution
    # unchanged; feeding -t1 compares sample by sample
    ref = heitz_vndf_vector_form(v, r, -t1, t2, z_from_unlerped=True)
    assert np.abs(wh - ref).max() < 2e-4
    correct = heitz_vndf_vector_form(v, r, -t1, t2)
    back = float((wh.astype(np.float64) @ v < 0).mean())
    print(f"r = {r}, theta_v = {theta_v}: back-facing {back:.4f}, max |h_ref - h_vndf| = "
          f"{np.abs(ref - correct).max():.3f}")


def test_reference_smith_g_closed_forms():
    """BSDF::Smith_G(v, wh, a) = 2 / (1 + sqrt(1 + (a^2 tan^2 theta_h)^2)),
    1 at tan theta_h = 0, 0 when (v.wh) v_z <= 0: at a^2 tan^2 theta_h =
    sqrt(3), sqrt(8), sqrt(15) it is 2/3, 1/2, 2/5."""
    a = 0.5
    v = np.array([[0.0, 0.0, 1.0]] * 5)
    targets = np.array([np.sqrt(3.0), np.sqrt(8.0), np.sqrt(15.0)])
    th = np.arctan(np.sqrt(targets) / a)
    wh = np.concatenate([_dirs(th, 0.3), [[0.0, 0.0, 1.0]], [_dirs(0.6, 0.0)]])
    v[4] = _dirs(1.4, np.pi)  # v above the horizon, back-facing to wh: (v.wh) v_z < 0
    g = oracle.smith_g(v, wh, a)
    assert np.allclose(g[:3], [2.0 / 3.0, 0.5, 0.4], rtol=2e-6)
    assert g[3] == 1.0 and g[4] == 0.0


@pytest.mark.parametrize("r", ROUGH)
def test_vndf_pdf_is_reference_composition_and_its_mismatch(r):
    """GGX_VNDF_Pdf = GGX_D(h, a) Smith_G(v, h, a) |v.h| / |v_z| (composition
    of the pinned factors), and the ratio to Heitz's density is
    Smith_G / G1(v).  Its mean under the sampler is the integral of the
    reference's pdf over the sphere of half vectors: recorded, not one."""
    v = _dirs(0.9, 0.7)
    n = 200_000
    u = np.random.default_rng(5).random((n, 2), dtype=np.float32)
    wh, pdf = oracle.vndf(np.broadcast_to(v, (n, 3)), r, u)
    D = oracle.ggx_d(wh[:, 2], r).astype(np.float64)
    G = oracle.smith_g(np.broadcast_to(v, (n, 3)), wh, r).astype(np.float64)
    v32 = v.astype(np.float32).astype(np.float64)
    comp = D * G * np.abs(wh.astype(np.float64) @ v32) / abs(v32[2])
    assert np.allclose(pdf, comp, rtol=2e-5, atol=1e-30)
    heitz = vndf_heitz(np.broadcast_to(v32, wh.shape), wh.astype(np.float64), r)
    ok = heitz > 0
    ratio = pdf[ok] / heitz[ok]
    assert np.allclose(ratio, G[ok] / smith_g1_ggx(v32[2], r), rtol=1e-3)
    integral = ratio.mean()
    print(f"r = {r}: integral of GGX_VNDF_Pdf over the half vectors = {integral:.4f}")
    assert 0.0 < integral < 2.0


@pytest.mark.parametrize("r", ROUGH)
def test_sampled_pdf_equals_bsdf_pdf(r):
    rng = np.random.default_rng(3)
    n = 20000
    wo = _dirs(np.arccos(rng.uniform(0.05, 1.0, n)), rng.uniform(0, 2 * np.pi, n))
    u = rng.random((n, 2), dtype=np.float32)
    wi, pdf = oracle.bsdf_sample_local(MICROFACET, r, wo, u)
    _, pdf2 = oracle.bsdf_local(MICROFACET, r, [0.0] * 3, [0.04] * 3, wo, wi)
    up = wi[:, 2] > 0
    assert np.all(pdf[~up] == 0.0)
    assert np.allclose(pdf[up], pdf2[up], rtol=1e-3, atol=1e-6)
    # the reflection Jacobian: pdf(wi) = pdf_h(wh) / (4 wi.wh)
    wh = wo.astype(np.float32) + wi
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    _, ph = oracle.vndf(wo, r, u)
    good = up & (pdf > 0)
    j = ph[good] / (4.0 * (wi[good] * wh[good]).sum(1))
    assert np.allclose(pdf[good], np.abs(j), rtol=1e-3)
