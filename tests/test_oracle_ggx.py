"""Pins the oracle's GGX reflection lobe (the Glossy shader: oracle.cpp bsdf_eval / bsdf_sample,
type 2) by mathematics that does not restate its formulas. The lobe has no reference counterpart
(the reference ships LambertShader and NullShader only, LambertShader.cpp:27-73); it follows the
ShaderInterface contract of those shaders: evaluate(in, out, n) -> (f, cos, pdf) and
sample(u1, u2, out, n) -> in with its evaluation (ShaderInterface.h:34-95).

What pins each part:
* D: its projected integral over the hemisphere is 1 (the microfacet normalisation), by quadrature.
* G1: Smith's visible-area identity -- the microfacets seen from v cover the macrosurface's
  projected area: G1(v) * integral of max(0, v.h) D(h) dh = cos(theta_v) (Heitz 2014, eq. 2+7).
* pdf: integrates over the upper hemisphere to the sampler's acceptance rate; a chi-square test of
  sampled directions against it on bins centred on the mirror direction.
* f: reciprocity; the sampled weight f cos / pdf equals G1(i) G1(o) (o.h) / (cos_o cos_h)
  (Walter et al. 2007, eq. 41 with F = 1), which fixes f once D, G1 and the pdf are pinned;
  directional albedo <= 1 and close to 1 for a nearly smooth surface at normal incidence.
* Planted mistakes, built into copies of the oracle on the CPU, must each fail one of these: the
  1/(4 o.h) Jacobian dropped, G1's arguments swapped, D without its 1/pi, G1's cos squared lost,
  the 4 of f's denominator dropped, the sampler's alpha^2 replaced by alpha.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as orc

ALPHAS = (0.05, 0.3, 0.6)
N_UP = np.array([0.0, 0.0, 1.0])


def _dir(theta, phi=0.0):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _lobe_grid(alpha, n_theta):
    """Quadrature nodes for a function of theta_h peaked at 0 with width alpha: theta = atan(alpha
    tan s) with s midpoints on (0, pi/2); returns theta and d theta."""
    s = (np.arange(n_theta) + 0.5) * (0.5 * math.pi / n_theta)
    theta = np.arctan(alpha * np.tan(s))
    dtheta = alpha / np.cos(s) ** 2 / (1.0 + (alpha * np.tan(s)) ** 2) * (0.5 * math.pi / n_theta)
    return theta, dtheta


# ---- checks (each raises AssertionError when the lobe breaks its law) -----------------------

def check_d_normalised(b, alpha):
    theta, dth = _lobe_grid(alpha, 20000)
    d = b.ggx_d(np.cos(theta)).astype(np.float64)
    integral = 2.0 * math.pi * np.sum(d * np.cos(theta) * np.sin(theta) * dth)
    assert abs(integral - 1.0) < 2e-3, f"alpha {alpha}: projected D integrates to {integral}"


def check_smith_visible_area(b, alpha):
    theta, dth = _lobe_grid(alpha, 6000)
    nphi = 2048
    phi = (np.arange(nphi) + 0.5) * (2.0 * math.pi / nphi)
    d = b.ggx_d(np.cos(theta)).astype(np.float64)
    for tv in (0.0, 0.5, 1.0, 1.3, 1.5):
        v = _dir(tv)
        vh = (np.sin(theta)[:, None] * np.cos(phi)[None, :] * v[0] + np.cos(theta)[:, None] * v[2])
        proj = np.maximum(vh, 0.0).sum(axis=1) * (2.0 * math.pi / nphi)
        area = np.sum(proj * d * np.sin(theta) * dth)
        g1 = float(b.ggx_g1(np.array([math.cos(tv)]))[0])
        assert abs(g1 * area - math.cos(tv)) < 3e-3 * max(math.cos(tv), 0.05), \
            f"alpha {alpha} theta_v {tv}: G1 * visible area {g1 * area} != cos {math.cos(tv)}"


def _mirror_frame(wo):
    r = np.array([-wo[0], -wo[1], wo[2]])
    a = np.array([1.0, 0.0, 0.0]) if abs(r[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t = np.cross(r, a)
    t /= np.linalg.norm(t)
    return r, t, np.cross(r, t)


def _lobe_dirs(wo, alpha, n_psi, n_chi):
    """Directions around the mirror direction of wo, spaced to the lobe width: psi (angle from the
    mirror direction) = atan(4 alpha tan s), chi the azimuth around it; with d omega."""
    r, t, s_ = _mirror_frame(wo)
    s = (np.arange(n_psi) + 0.5) * (0.5 * math.pi / n_psi)
    k = 4.0 * alpha
    psi = np.arctan(k * np.tan(s)) * 2.0  # covers [0, pi)
    dpsi = 2.0 * k / np.cos(s) ** 2 / (1.0 + (k * np.tan(s)) ** 2) * (0.5 * math.pi / n_psi)
    chi = (np.arange(n_chi) + 0.5) * (2.0 * math.pi / n_chi)
    sp, cp = np.sin(psi)[:, None], np.cos(psi)[:, None]
    w = (cp[..., None] * r + sp[..., None] * (np.cos(chi)[None, :, None] * t + np.sin(chi)[None, :, None] * s_))
    dom = (np.sin(psi) * dpsi)[:, None] * np.full(n_chi, 2.0 * math.pi / n_chi)[None, :]
    return w.reshape(-1, 3), dom.ravel(), psi, chi


def _pdf_integral(b, wo, alpha):
    w, dom, _, _ = _lobe_dirs(wo, alpha, 3000, 720)
    _, _, pdf = b.eval(w.astype(np.float32), wo.astype(np.float32), N_UP.astype(np.float32))
    return float(np.sum(pdf.astype(np.float64) * dom))


def _sample(b, wo, n, seed):
    rng = np.random.default_rng(seed)
    u = rng.random((n, 2), dtype=np.float32)
    return b.sample(u, wo.astype(np.float32), N_UP.astype(np.float32))


def check_pdf_integrates_to_acceptance(b, alpha, theta_o):
    wo = _dir(theta_o, 0.3)
    integral = _pdf_integral(b, wo, alpha)
    n = 400000
    wi, f, c, pdf, ok = _sample(b, wo, n, 11)
    acc = ok.mean()
    se = math.sqrt(acc * (1 - acc) / n) + 1e-4
    assert integral <= 1.0 + 3e-3, f"alpha {alpha} theta_o {theta_o}: pdf integrates to {integral} > 1"
    assert abs(integral - acc) < 5 * se + 3e-3, \
        f"alpha {alpha} theta_o {theta_o}: pdf integrates to {integral}, acceptance {acc}"


def check_chi2(b, alpha, theta_o, seed=3):
    """Sampled directions against the evaluated pdf on bins around the mirror direction."""
    from scipy import stats

    wo = _dir(theta_o, 0.7)
    n = 400000
    wi, f, c, pdf, ok = _sample(b, wo, n, seed)
    r, t, s_ = _mirror_frame(wo)
    wi = wi[ok].astype(np.float64)
    psi_s = np.arccos(np.clip(wi @ r, -1.0, 1.0))
    chi_s = np.mod(np.arctan2(wi @ s_, wi @ t), 2.0 * math.pi)
    # bin edges: psi on the lobe scale, 16 azimuth sectors
    psi_edges = np.concatenate([[0.0], np.minimum(alpha * np.array([0.25, 0.5, 0.8, 1.2, 1.7, 2.4, 3.5, 5.0, 8.0]),
                                                  math.pi * 0.99), [math.pi]])
    psi_edges = np.unique(psi_edges)
    nchi = 16
    obs, _, _ = np.histogram2d(psi_s, chi_s, bins=[psi_edges, np.linspace(0, 2 * math.pi, nchi + 1)])
    # expected counts from the pdf, by quadrature inside each bin
    w, dom, psi, chi = _lobe_dirs(wo, alpha, 4000, 960)
    _, _, pdf_g = b.eval(w.astype(np.float32), wo.astype(np.float32), N_UP.astype(np.float32))
    mass = (pdf_g.astype(np.float64) * dom).reshape(psi.size, chi.size)
    pi = np.clip(np.searchsorted(psi_edges, psi, side="right") - 1, 0, len(psi_edges) - 2)
    ci = np.clip((chi / (2 * math.pi / nchi)).astype(int), 0, nchi - 1)
    exp = np.zeros_like(obs)
    np.add.at(exp, (pi[:, None].repeat(chi.size, 1), ci[None, :].repeat(psi.size, 0)), mass)
    exp *= n
    keep = exp > 20
    o, e = obs[keep], exp[keep]
    chi2 = float(np.sum((o - e) ** 2 / e))
    dof = int(keep.sum()) - 1
    p = stats.chi2.sf(chi2, dof)
    lost = obs[~keep].sum()
    assert lost < 0.002 * n + 50, f"alpha {alpha}: {lost} samples fell in bins the pdf gives no mass"
    assert p > 1e-4, f"alpha {alpha} theta_o {theta_o}: chi2 {chi2:.1f} on {dof} dof, p = {p:.2e}"


def check_weight_identity(b, alpha, theta_o):
    wo = _dir(theta_o, 1.1)
    wi, f, c, pdf, ok = _sample(b, wo, 20000, 5)
    wi, f, c, pdf = wi[ok].astype(np.float64), f[ok, 0].astype(np.float64), c[ok], pdf[ok]
    h = wi + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    g_i = b.ggx_g1(wi[:, 2]).astype(np.float64)
    g_o = float(b.ggx_g1(np.array([wo[2]]))[0])
    expect = g_i * g_o * (h @ wo) / (wo[2] * h[:, 2])
    weight = f * c / pdf
    err = np.abs(weight - expect) / np.maximum(expect, 1e-6)
    assert np.quantile(err, 0.999) < 2e-3, f"alpha {alpha}: weight f cos / pdf off by {err.max()}"


def check_reciprocity(b, alpha):
    rng = np.random.default_rng(17)
    n = 20000
    mu = rng.random((2, n)) * 0.98 + 0.02
    ph = rng.random((2, n)) * 2 * math.pi
    d = np.stack([np.sqrt(1 - mu ** 2) * np.cos(ph), np.sqrt(1 - mu ** 2) * np.sin(ph), mu], -1)
    f1, _, _ = b.eval(d[0].astype(np.float32), d[1].astype(np.float32), N_UP.astype(np.float32))
    f2, _, _ = b.eval(d[1].astype(np.float32), d[0].astype(np.float32), N_UP.astype(np.float32))
    err = np.abs(f1 - f2) / np.maximum(np.abs(f1), 1e-6)
    assert err.max() < 1e-4, f"alpha {alpha}: f(i, o) != f(o, i) by {err.max()}"


def albedo(b, alpha, theta_o):
    wo = _dir(theta_o, 0.2)
    w, dom, _, _ = _lobe_dirs(wo, alpha, 3000, 720)
    f, c, _ = b.eval(w.astype(np.float32), wo.astype(np.float32), N_UP.astype(np.float32))
    return float(np.sum(f[:, 0].astype(np.float64) * c * dom))


def check_albedo(b, alpha):
    for to in (0.0, 0.6, 1.2, 1.45):
        a = albedo(b, alpha, to)
        assert 0.0 < a <= 1.0 + 2e-3, f"alpha {alpha} theta_o {to}: albedo {a}"
    if alpha <= 0.05:
        a = albedo(b, alpha, 0.0)
        assert a > 0.98, f"a nearly smooth lobe at normal incidence reflects only {a}"


# ---- the oracle's lobe passes every check ----------------------------------------------------

@pytest.fixture(params=ALPHAS, ids=lambda a: f"alpha{a}")
def lobe(request):
    return orc.Bsdf(2, roughness=request.param), request.param


def test_ggx_d_normalised(lobe):
    check_d_normalised(*lobe)


def test_ggx_g1_smith_visible_area(lobe):
    check_smith_visible_area(*lobe)


@pytest.mark.parametrize("theta_o", [0.0, 0.8, 1.35])
def test_ggx_pdf_integrates_to_acceptance(lobe, theta_o):
    check_pdf_integrates_to_acceptance(*lobe, theta_o)


@pytest.mark.parametrize("theta_o", [0.1, 0.8, 1.35])
def test_ggx_sampling_chi2(lobe, theta_o):
    check_chi2(*lobe, theta_o)


@pytest.mark.parametrize("theta_o", [0.0, 0.7, 1.4])
def test_ggx_sample_weight_identity(lobe, theta_o):
    check_weight_identity(*lobe, theta_o)


def test_ggx_reciprocity(lobe):
    check_reciprocity(*lobe)


def test_ggx_albedo(lobe):
    check_albedo(*lobe)


def test_ggx_mirror_limit():
    """As roughness goes to 0 the lobe becomes a mirror: sampled directions reflect wo about n
    (GGX's tail is heavy: P(tan theta_h > x alpha) = 1 / (1 + x^2), so 90% lie within ~6 alpha)."""
    b = orc.Bsdf(2, roughness=1e-3)
    wo = _dir(0.9, 0.4)
    wi, f, c, pdf, ok = _sample(b, wo, 4000, 2)
    mirror = np.array([-wo[0], -wo[1], wo[2]])
    dev = np.arccos(np.clip(wi.astype(np.float64) @ mirror, -1, 1))
    assert ok.all() and np.quantile(dev, 0.9) < 8e-3


def test_ggx_below_horizon_is_black():
    b = orc.Bsdf(2, roughness=0.3)
    f, c, pdf = b.eval(np.array([[0.3, 0.0, -0.95]], np.float32), _dir(0.4).astype(np.float32),
                       N_UP.astype(np.float32))
    assert f.max() == 0.0 and pdf[0] == 0.0
    f, c, pdf = b.eval(_dir(0.4).astype(np.float32), np.array([[0.3, 0.0, -0.95]], np.float32),
                       N_UP.astype(np.float32))
    assert f.max() == 0.0 and pdf[0] == 0.0


# ---- planted mistakes must fail ---------------------------------------------------------------

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "oracle.cpp")

PLANTS = {
    "jacobian_dropped": ("e.pdf = oh > 0.f ? D * ch / (4.f * oh) : 0.f;", "e.pdf = oh > 0.f ? D * ch : 0.f;",
                         ["pdf", "chi2"]),
    "g1_arguments_swapped": ("float G = ggx_g1(a2, ci) * ggx_g1(a2, co);",
                             "float G = ggx_g1(ci, a2) * ggx_g1(co, a2);", ["weight"]),
    "d_without_pi": ("return a2 / (kPi * k * k);", "return a2 / (k * k);", ["d"]),
    "g1_cos_not_squared": ("std::sqrt(a2 + (1.f - a2) * c * c)", "std::sqrt(a2 + (1.f - a2) * c)", ["g1"]),
    "f_without_4": ("(4.f * ci * co)", "(ci * co)", ["weight", "albedo"]),
    "sampler_alpha_not_squared": ("float tan2 = a2 * u2 / (1.f - u2);", "float tan2 = a * u2 / (1.f - u2);",
                                  ["chi2"]),
}


def _run_check(name, b, alpha):
    if name == "pdf":
        check_pdf_integrates_to_acceptance(b, alpha, 0.8)
    elif name == "chi2":
        check_chi2(b, alpha, 0.8)
    elif name == "weight":
        check_weight_identity(b, alpha, 0.7)
    elif name == "d":
        check_d_normalised(b, alpha)
    elif name == "g1":
        check_smith_visible_area(b, alpha)
    elif name == "albedo":
        check_albedo(b, alpha)


@pytest.fixture(scope="module")
def planted_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("planted_oracle")


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails(plant, planted_dir):
    old, new, checks = PLANTS[plant]
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"plant site for {plant} not found once in oracle.cpp"
    path = planted_dir / f"{plant}.cpp"
    path.write_text(src.replace(old, new))
    so = planted_dir / f"lib{plant}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    "-I", os.path.dirname(ORACLE_SRC), "-o", str(so), str(path)], check=True)
    L = orc.load(str(so))
    for chk in checks:
        for alpha in (0.3,):
            b = orc.Bsdf(2, roughness=alpha, L=L)
            with pytest.raises(AssertionError):
                _run_check(chk, b, alpha)
