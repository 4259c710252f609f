"""Pins of the oracle's microfacet BSDFs against the published models they implement (no GPU).

Each check evaluates the oracle at configurations where the reference's code reduces to a
textbook formula, written here from the publication, not from the oracle:

  GGX distribution (Walter et al. 2007, eq. 33), anisotropic form with
  alpha = roughness^2 and aspect = sqrt(1 - 0.9 anisotropy) (Burley 2012, "Physically based
  shading at Disney", addendum B; reference materials.cu:1220-1240)
      MicrofacetReflection at normal incidence: pdf(n -> l) = D(h) / 4 (visible-normal pdf
      with G1(n) = 1 and l.h = n.h), so the pdf's shape is D's; D(n) = 1 / (pi ax ay)
  conductor Fresnel (pbrt-v2 FrCond, Pharr & Humphreys 2010, sec. 8.2.1; reference
  materials.cu:13-46)
      normal incidence: ((eta - 1)^2 + k^2) / ((eta + 1)^2 + k^2); all angles through the
      MicrofacetReflection check below, with directions down to 3 degrees above the horizon
  MicrofacetReflection (Walter et al. 2007, eq. 20 with the separable Smith G1 of eq. 34)
      f = F(v.h) D(h) G1(v) G1(l) / (4 |v.n| |l.n|)
  UE4 material (Karis 2013, "Real shading in Unreal Engine 4": alpha = roughness^2,
  F0 = lerp(0.08 specular, base, metallic) with specular 0.5, diffuse = base (1 - metallic);
  reference materials.cu:1959-1978)
      specular lobe: Schlick's Fresnel (Schlick 1994) with exponent 5, GGX D and the
      height-correlated Smith G (Heitz 2014, eq. 99)
      diffuse lobe: Frostbite's renormalised Disney diffuse (Lagarde & de Rousiers 2014,
      listing 1): energyBias = 0.5 r, energyFactor = lerp(1, 1 / 1.51, r),
      F_D90 = energyBias + 2 (l.h)^2 r, fd = Schlick(1, F_D90, n.l) Schlick(1, F_D90, n.v)
      energyFactor / pi
      the one-sample MIS pdf at normal incidence: (cos / pi iB + D / 4 iS) / (iB + iS) with
      iB = importance(diffuse) energyFactor and iS = importance(F0)
  SampledSpectrum::importance (reference spectrum_types.h:250-264): 0.9 of the selected
  wavelength's value plus 0.1 spread evenly over the other three; checked through the UE4
  selection weights in the spectral build, where a mean or the selected value alone would
  give another pdf.

RGB parameters pass through the RGB -> XYZ -> RGB matrices in float, hence the 3e-4
tolerances; the formulas themselves are evaluated in float64.
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from vlr_amd import scenes

RTOL = 3e-4


def q_of(wo, wi, wl_select=0.0):
    n = len(wo)
    q = np.zeros(n, ob.query_dtype())
    q["dirLocal"] = wo
    q["evalDir"] = wi
    q["geomNormalLocal"] = (0.0, 0.0, 1.0)
    q["wlOffset"] = 0.5
    q["wlSelect"] = wl_select
    q["uComponent"] = 0.5
    q["uDir"] = 0.5
    return q


def hemi(rng, n, min_z=0.05):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = np.abs(d[:, 2])
    return d[d[:, 2] > min_z]


def ue4(base, roughness, metallic):
    return scenes.MaterialDesc(scenes.UE4, {"base color": (scenes.REC709,) + tuple(base)},
                               {"roughness": roughness, "metallic": metallic})


def conductor(eta, k, roughness, anisotropy=0.0):
    return scenes.MaterialDesc(scenes.MICROFACET_REFLECTION, {"eta": (scenes.REC709,) + tuple(eta),
                                                              "k": (scenes.REC709,) + tuple(k)},
                               {"roughness": roughness, "anisotropy": anisotropy, "rotation": 0.0})


# ---- textbook formulas (float64) ----
def ggx_d(h, ax, ay):
    t = (h[:, 0] / ax) ** 2 + (h[:, 1] / ay) ** 2 + h[:, 2] ** 2
    return 1.0 / (math.pi * ax * ay * t * t)


def smith_lambda(v, a):
    tan2 = (v[:, 0] ** 2 + v[:, 1] ** 2) / v[:, 2] ** 2
    return (-1.0 + np.sqrt(1.0 + a * a * tan2)) / 2.0


def half(v, l):
    h = v + l
    return h / np.linalg.norm(h, axis=1, keepdims=True)


def fr_conductor(cos, eta, k):
    cos = np.abs(cos)[:, None]
    tmp_f = eta * eta + k * k
    tmp = tmp_f * cos * cos
    rparl2 = (tmp - 2 * eta * cos + 1) / (tmp + 2 * eta * cos + 1)
    rperp2 = (tmp_f - 2 * eta * cos + cos * cos) / (tmp_f + 2 * eta * cos + cos * cos)
    return (rparl2 + rperp2) / 2


def schlick(f0, f90, u):
    return f0 + (f90 - f0) * (1.0 - u) ** 5


# ---- GGX distribution, alpha = roughness^2 ----
@pytest.mark.parametrize("roughness,anisotropy", [(0.2, 0.0), (0.5, 0.0), (0.9, 0.0), (0.4, 0.6), (0.7, 1.0)])
def test_ggx_distribution_shape_at_alpha_roughness_squared(roughness, anisotropy):
    rng = np.random.default_rng(101)
    wi = hemi(rng, 4000, 0.02)
    wo = np.tile([0.0, 0.0, 1.0], (len(wi), 1))
    r = ob.bsdf_queries(conductor((1.2, 1.2, 1.2), (3.0, 3.0, 3.0), roughness, anisotropy), False, q_of(wo, wi))
    aspect = math.sqrt(1.0 - 0.9 * anisotropy)
    a = roughness ** 2
    ax, ay = max(0.001, a / aspect), max(0.001, a * aspect)
    h = half(wo, wi)
    np.testing.assert_allclose(r["evalPDF"], ggx_d(h, ax, ay) / 4.0, rtol=2e-5)
    # alpha = roughness would put the normal-incidence peak D(n) = 1 / (pi ax ay) elsewhere
    peak = ob.bsdf_queries(conductor((1.2,) * 3, (3.0,) * 3, roughness, anisotropy), False,
                           q_of(np.array([[0.0, 0.0, 1.0]]), np.array([[0.0, 0.0, 1.0]])))["evalPDF"][0]
    assert abs(peak - 1.0 / (4 * math.pi * ax * ay)) < 2e-5 * peak
    wrong = 1.0 / (4 * math.pi * (roughness / aspect) * (roughness * aspect))
    assert abs(peak - wrong) > 0.1 * peak


# ---- conductor Fresnel ----
@pytest.mark.parametrize("eta,k", [((0.2, 0.9, 1.1), (3.9, 2.4, 2.2)), ((1.5, 1.5, 1.5), (0.0, 0.0, 0.0)),
                                   ((0.05, 0.4, 1.6), (1.0, 2.0, 6.0))])
def test_conductor_fresnel_normal_incidence(eta, k):
    roughness = 0.5
    n = np.array([[0.0, 0.0, 1.0]])
    r = ob.bsdf_queries(conductor(eta, k, roughness), False, q_of(n, n))
    eta_, k_ = np.array(eta), np.array(k)
    fn = ((eta_ - 1) ** 2 + k_ ** 2) / ((eta_ + 1) ** 2 + k_ ** 2)
    alpha = roughness ** 2
    # f(n, n) = F(1) D(n) G1 G1 / 4 with G1(n) = 1 and D(n) = 1 / (pi alpha^2)
    np.testing.assert_allclose(r["evalValue"][0, :3], fn / (4 * math.pi * alpha * alpha), rtol=RTOL)


def test_microfacet_reflection_is_walter_with_pbrt_conductor_fresnel():
    rng = np.random.default_rng(102)
    eta, k, roughness = np.array([0.2, 0.9, 1.1]), np.array([3.9, 2.4, 2.2]), 0.45
    wo, wi = hemi(rng, 6000), hemi(rng, 6000)
    n = min(len(wo), len(wi))
    wo, wi = wo[:n], wi[:n]
    r = ob.bsdf_queries(conductor(eta, k, roughness), False, q_of(wo, wi))
    a = roughness ** 2
    h = half(wo, wi)
    g1v = 1.0 / (1.0 + smith_lambda(wo, a))
    g1l = 1.0 / (1.0 + smith_lambda(wi, a))
    f = fr_conductor(np.sum(wo * h, axis=1), eta, k) * (ggx_d(h, a, a) * g1v * g1l / (4 * wo[:, 2] * wi[:, 2]))[:, None]
    np.testing.assert_allclose(r["evalValue"][:, :3], f, rtol=RTOL, atol=1e-7)


# ---- UE4 ----
@pytest.mark.parametrize("roughness", [0.05, 0.3, 0.6, 1.0])
@pytest.mark.parametrize("metallic", [0.0, 0.3, 1.0])
def test_ue4_value_at_normal_incidence(roughness, metallic):
    base = np.array([0.8, 0.45, 0.2])
    n = np.array([[0.0, 0.0, 1.0]])
    r = ob.bsdf_queries(ue4(base, roughness, metallic), False, q_of(n, n))
    alpha = roughness ** 2
    f0 = 0.04 * (1 - metallic) + base * metallic
    energy = 1.0 + (1.0 / 1.51 - 1.0) * roughness
    # F(l.h = 1) = F0; D(n) = 1 / (pi alpha^2); G = 1; the Disney Fresnel factors are 1 at n
    expect = base * (1 - metallic) * energy / math.pi + f0 / (4 * math.pi * alpha * alpha)
    np.testing.assert_allclose(r["evalValue"][0, :3], expect, rtol=RTOL)


@pytest.mark.parametrize("roughness", [0.25, 0.7])
@pytest.mark.parametrize("metallic", [0.0, 1.0, 0.4])
def test_ue4_value_is_karis_specular_plus_frostbite_diffuse(roughness, metallic):
    rng = np.random.default_rng(103)
    base = np.array([0.7, 0.3, 0.55])
    wo, wi = hemi(rng, 6000), hemi(rng, 6000)
    n = min(len(wo), len(wi))
    wo, wi = wo[:n], wi[:n]
    r = ob.bsdf_queries(ue4(base, roughness, metallic), False, q_of(wo, wi))
    a = roughness ** 2
    h = half(wo, wi)
    ldoth = np.sum(wi * h, axis=1)
    f0 = 0.04 * (1 - metallic) + base * metallic
    g2 = 1.0 / (1.0 + smith_lambda(wo, a) + smith_lambda(wi, a))
    spec = schlick(f0[None, :], 1.0, ldoth[:, None]) * (ggx_d(h, a, a) * g2 / (4 * wo[:, 2] * wi[:, 2]))[:, None]
    fd90 = 0.5 * roughness + 2.0 * ldoth ** 2 * roughness
    energy = 1.0 + (1.0 / 1.51 - 1.0) * roughness
    fd = schlick(1.0, fd90, wi[:, 2]) * schlick(1.0, fd90, wo[:, 2]) * energy / math.pi
    diff = (base * (1 - metallic))[None, :] * fd[:, None]
    np.testing.assert_allclose(r["evalValue"][:, :3], spec + diff, rtol=RTOL, atol=1e-7)


@pytest.mark.parametrize("gray,roughness,metallic", [(0.6, 0.5, 0.0), (0.3, 0.9, 0.5), (0.8, 0.2, 0.2)])
def test_ue4_selection_weights_at_normal_incidence(gray, roughness, metallic):
    rng = np.random.default_rng(104)
    wi = hemi(rng, 3000, 0.02)
    wo = np.tile([0.0, 0.0, 1.0], (len(wi), 1))
    r = ob.bsdf_queries(ue4((gray,) * 3, roughness, metallic), False, q_of(wo, wi))
    a = roughness ** 2
    energy = 1.0 + (1.0 / 1.51 - 1.0) * roughness
    i_b = gray * (1 - metallic) * energy
    i_s = 0.04 * (1 - metallic) + gray * metallic
    h = half(wo, wi)
    expect = (wi[:, 2] / math.pi * i_b + ggx_d(h, a, a) / 4 * i_s) / (i_b + i_s)
    np.testing.assert_allclose(r["evalPDF"], expect, rtol=RTOL)


@pytest.mark.parametrize("sel", [0, 1, 2, 3])
def test_spectral_importance_weights_the_selected_wavelength(sel):
    """UE4, metallic 0, roughness 0.5, dark strongly coloured base, spectral: at wo = wi = n
    the value gives each wavelength's diffuse reflectance c[i]; the pdf there is
    (iB / pi + iS / (4 pi alpha^2)) / (iB + iS) with iB = importance(c) energyFactor and iS the
    importance of the flat F0 = 0.04."""
    n = np.array([[0.0, 0.0, 1.0]])
    roughness = 0.5
    mat = ue4((0.12, 0.006, 0.003), roughness, 0.0)
    r = ob.bsdf_queries(mat, True, q_of(n, n, wl_select=(sel + 0.5) / 4))
    f = r["evalValue"][0].astype(np.float64)
    f0, alpha = 0.04, roughness ** 2
    energy = 1.0 + (1.0 / 1.51 - 1.0) * roughness
    spec_at_n = 1.0 / (4 * math.pi * alpha * alpha)
    c = (f - f0 * spec_at_n) * math.pi / energy
    assert c.max() > 3 * c.min()           # the weighting matters for this spectrum

    def pdf_for(imp):
        i_b = imp * energy
        return (i_b / math.pi + f0 * spec_at_n) / (i_b + f0)

    imp = 0.9 * c[sel] + 0.1 / 3 * (c.sum() - c[sel])
    got = r["evalPDF"][0]
    assert abs(got - pdf_for(imp)) < 1e-3 * got
    # neither the plain mean nor the selected value alone gives this pdf
    for other in (c.mean(), c[sel]):
        assert abs(got - pdf_for(other)) > 5e-3 * got
