"""Closed-form pins of the oracle's Disney BSDF evaluation (bxdf.h:93-174).

The Disney model (Burley 2012, "Physically-Based Shading at Disney", with
the reference's constants: specular scale 0.08, Fd90 = 0.5 + 2 cos^2(theta_d)
roughness, the Hanrahan-Krueger subsurface factor 1.25, clearcoat 0.25 with
F0 = 0.04 and the GTR1 distribution, GTR2/GTR2-aniso and Smith GGX masking)
reduces to hand-derived numbers at three configurations, which this file
evaluates from the published formulas at those angles, term by term:

- normal incidence and exit (L = V = n): every Schlick factor at cos 1 is
  F0 (0 for the diffuse, retro-reflection and sheen terms), both Smith
  factors are 1, D_GTR2aniso = 1/(pi ax ay), D_GTR1 = (g^2-1)/(pi ln(g^2) g^2);
- V = n, L at 60 degrees: the half vector is at 30 degrees, cos theta_d =
  sqrt(3)/2, FL = 1/32, and the masking of L is 1/(1/2 + sqrt(a^2 + 1/4 -
  a^2/4));
- straight transmission (V = n, L = -n): D_GTR2 = 1/(pi a^2), the
  generalized half vector is n, and the refraction Jacobian factor is
  eta^2 / (1 - eta)^2.

A one-constant slip in bxdf.h:105-170 (the 0.08, the factor 2 of Fd90, the
order of a lerp, the 0.25 or 0.04 of the clearcoat, 1.25, the sheen tint)
moves at least one of these values.  A further test isolates the sheen
lobe at random directions as a difference of two evaluations.

Material parameters are float32 in the record and the reference's literals
are float (0.08f, 0.9f, ...), so the expectations use those float values in
float64 arithmetic.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from srt_amd import scenes


def F(x):
    """A float literal or float parameter of the reference, read in float64."""
    return float(np.float32(x))


INV_PI = 1.0 / math.pi


def _lum(c):
    return F(0.2126) * c[0] + F(0.7152) * c[1] + F(0.0722) * c[2]


def _params(mat):
    keys = ["ior", "metallic", "subsurface", "roughness", "specular", "specular_tint", "anisotropic", "sheen",
            "sheen_tint", "clearcoat", "clearcoat_gloss", "spec_trans"]
    return {k: float(mat[k]) for k in keys}


def _schlick(c, eta=1.0):
    r0 = ((eta - 1.0) / (eta + 1.0)) ** 2
    m = min(max(1.0 - c, 0.0), 1.0)
    return r0 + (1.0 - r0) * m ** 5


def _smith(n_v, a):
    return 2.0 * n_v / (n_v + math.sqrt(a * a + n_v * n_v - a * a * n_v * n_v))


def _gtr1(c, g):
    if g >= 1.0:
        return INV_PI
    g2 = g * g
    return (g2 - 1.0) / (math.pi * math.log(g2) * (1.0 + (g2 - 1.0) * c * c))


def _colors(p, color):
    color = np.asarray(color, np.float64)
    lum = _lum(color)
    ctint = color / lum if lum > 0 else np.ones(3)
    tint = 1.0 + p["specular_tint"] * (ctint - 1.0)
    base = p["specular"] * F(0.08) * tint
    spec_color = base + p["metallic"] * (color - base)
    sheen_color = 1.0 + p["sheen_tint"] * (ctint - 1.0)
    return color, lum, spec_color, sheen_color


def _aniso(p):
    aspect = math.sqrt(1.0 - p["anisotropic"] * F(0.9))
    a2 = p["roughness"] ** 2
    return max(F(0.001), a2 / aspect), max(F(0.001), a2 * aspect)


def expected_normal(p, color):
    """f(L = n, V = n): diffuse with FL = FV = FH = 0, specular and clearcoat
    with both Smith factors 1."""
    color, lum, spec_color, _ = _colors(p, color)
    m, st = p["metallic"], p["spec_trans"]
    eta = 1.0 / p["ior"]
    r0 = ((eta - 1.0) / (eta + 1.0)) ** 2
    f = np.zeros(3)
    if lum * (1 - m) * (1 - st) > 0:
        # Fd = 1, ss = 1.25 (Fss (1/2 - 1/2) + 1/2) = 0.625, sheen F = 0
        f += color * INV_PI * (1.0 - 0.375 * p["subsurface"]) * (1 - m) * (1 - st)
    fm = r0 * (1.0 - m)                      # fresnel_mix at cos 1: F0(eta), or 0 for the metal part
    Fs = spec_color + fm * (1.0 - spec_color)
    ax, ay = _aniso(p)
    f += Fs * (INV_PI / (ax * ay)) / 4.0
    cc = p["clearcoat"]
    if cc * (1 - m) > 0:
        r0c = ((1.0 / F(1.5) - 1.0) / (1.0 / F(1.5) + 1.0)) ** 2
        Fc = F(0.04) + r0c * (1.0 - F(0.04))
        f += F(0.25) * cc * Fc * _gtr1(1.0, p["clearcoat_gloss"]) / 4.0
    return f


def expected_60(p, color):
    """f(L at 60 degrees from n in the xz plane, V = n): H at 30 degrees."""
    color, lum, spec_color, sheen_color = _colors(p, color)
    m, st, r = p["metallic"], p["spec_trans"], p["roughness"]
    eta = 1.0 / p["ior"]
    c = math.sqrt(3.0) / 2.0                 # cos theta_d = L.H = V.H = H.z
    f = np.zeros(3)
    if lum * (1 - m) * (1 - st) > 0:
        FL = 0.5 ** 5                        # (1 - L.z)^5, L.z = 1/2
        Fd90 = 0.5 + 2.0 * 0.75 * r
        Fd = 1.0 + (Fd90 - 1.0) * FL         # FV = 0
        Fss90 = 0.75 * r
        Fss = 1.0 + (Fss90 - 1.0) * FL
        ss = F(1.25) * (Fss * (1.0 / 1.5 - 0.5) + 0.5)
        FH = (1.0 - c) ** 5
        diff = Fd + p["subsurface"] * (ss - Fd)
        f += (INV_PI * diff * color + FH * p["sheen"] * sheen_color) * (1 - m) * (1 - st)
    fm = _schlick(c, eta) + m * (_schlick(c) - _schlick(c, eta))
    Fs = spec_color + fm * (1.0 - spec_color)
    ax, ay = _aniso(p)
    D = INV_PI / (ax * ay * ((0.5 / ax) ** 2 + 0.75) ** 2)
    G = 1.0 / (0.5 + math.sqrt(r * r + 0.25 - 0.25 * r * r))     # Smith of L.z = 1/2; of V.z = 1 it is 1
    f += Fs * D * G / (4.0 * 0.5)
    cc = p["clearcoat"]
    if cc * (1 - m) > 0:
        Fc = F(0.04) + _schlick(c, 1.0 / F(1.5)) * (1.0 - F(0.04))
        Gc = 1.0 / (0.5 + math.sqrt(0.0625 + 0.25 - 0.0625 * 0.25))
        f += F(0.25) * cc * Fc * _gtr1(c, p["clearcoat_gloss"]) * Gc / (4.0 * 0.5)
    return f


def expected_straight_through(p, color, inner):
    """f(L = -n, V = n): only the refraction lobe; H = n, V.H = 1, L.H = -1."""
    color = np.asarray(color, np.float64)
    m, st, r = p["metallic"], p["spec_trans"], p["roughness"]
    eta = p["ior"] if inner else 1.0 / p["ior"]
    r0 = ((eta - 1.0) / (eta + 1.0)) ** 2
    D = 1.0 / (math.pi * r * r)
    G = eta * eta / (1.0 - eta) ** 2
    return np.sqrt(color) * (1 - m) * st * (1.0 - r0) * D * G


REFLECT = [
    dict(roughness=1.0),
    dict(roughness=0.5),
    dict(roughness=0.3, metallic=1.0),
    dict(roughness=0.6, metallic=0.4, specular=0.8, specular_tint=0.6),
    dict(roughness=0.2, clearcoat=1.0, clearcoat_gloss=0.3),
    dict(roughness=0.7, clearcoat=0.5, clearcoat_gloss=1.0),
    dict(roughness=0.7, sheen=1.0, sheen_tint=0.5, subsurface=0.6),
    dict(roughness=0.9, sheen=0.8, sheen_tint=0.0, subsurface=1.0),
    dict(roughness=0.4, anisotropic=0.8, specular_tint=0.7),
    dict(roughness=0.45, spec_trans=0.5, ior=1.33),
    dict(roughness=0.8, metallic=0.2, spec_trans=0.3, clearcoat=0.7, clearcoat_gloss=0.1, sheen=0.4,
         sheen_tint=1.0, subsurface=0.3, anisotropic=0.5, specular_tint=0.2, specular=0.3, ior=1.7),
]
COLORS = [(0.8, 0.3, 0.1), (0.05, 0.6, 0.9), (1.0, 1.0, 1.0)]


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("color", COLORS)
@pytest.mark.parametrize("params", REFLECT)
def test_disney_eval_normal_incidence(params, color, fp64):
    m = scenes.material(**params)
    f = O.bxdf(m, "eval", wi=[[0, 0, 1]], wo=[[0, 0, 1]], color=color, fp64=fp64)[0]
    exp = expected_normal(_params(m), color)
    np.testing.assert_allclose(f, exp, rtol=1e-12 if fp64 else 3e-6)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("color", COLORS)
@pytest.mark.parametrize("params", REFLECT)
def test_disney_eval_sixty_degrees(params, color, fp64):
    m = scenes.material(**params)
    s = math.sqrt(3.0) / 2.0
    wi = [[s, 0.0, 0.5]] if fp64 else [[np.float32(s), 0.0, 0.5]]
    f = O.bxdf(m, "eval", wi=wi, wo=[[0, 0, 1]], color=color, fp64=fp64)[0]
    exp = expected_60(_params(m), color)
    np.testing.assert_allclose(f, exp, rtol=1e-12 if fp64 else 3e-6)


@pytest.mark.parametrize("inner", [False, True])
@pytest.mark.parametrize("params", [dict(roughness=0.3, spec_trans=1.0), dict(roughness=0.6, spec_trans=0.5,
                                                                                metallic=0.25, ior=1.33),
                                    dict(roughness=0.9, spec_trans=0.8, ior=2.0, clearcoat=1.0)])
def test_disney_eval_straight_transmission(params, inner):
    m = scenes.material(**params)
    color = (0.64, 0.25, 0.81)
    f = O.bxdf(m, "eval", wi=[[0, 0, -1]], wo=[[0, 0, 1]], color=color, inner=inner)[0]
    exp = expected_straight_through(_params(m), color, inner)
    np.testing.assert_allclose(f, exp, rtol=1e-12)
    # ior 1.5 from outside: G = (4/9) / (1/9) = 4 exactly; F0 = 0.04
    if params.get("ior", 1.5) == 1.5 and not inner:
        np.testing.assert_allclose(f, np.sqrt(color) * 0.96 * 4.0 / (math.pi * F(0.3) ** 2), rtol=1e-7)


def test_disney_reference_numbers():
    """Two hand-evaluated values: default Disney (roughness 0.5, specular 0.5)
    on white at normal incidence is 1/pi + (0.04 + 0.04 * 0.96) / (4 pi 0.25^2)
    = 0.318310 + 0.099822; the default clearcoat gloss 0.001 gives
    D_GTR1(1) = (1e-6 - 1) / (pi ln(1e-6) 1e-6)."""
    m = scenes.material(roughness=0.5)
    f = O.bxdf(m, "eval", wi=[[0, 0, 1]], wo=[[0, 0, 1]], color=(1, 1, 1))[0]
    spec = (F(0.04) * 1.0 + 0.04 * (1 - F(0.04))) / (4 * math.pi * 0.0625)
    np.testing.assert_allclose(f, 1 / math.pi + spec, rtol=1e-7)
    assert abs(f[0] - 0.4181318) < 2e-6
    m = scenes.material(roughness=1.0, clearcoat=1.0)
    g = F(0.001)
    d1 = (g * g - 1) / (math.pi * math.log(g * g) * g * g)
    f2 = O.bxdf(m, "eval", wi=[[0, 0, 1]], wo=[[0, 0, 1]], color=(1, 1, 1))[0]
    f1 = O.bxdf(scenes.material(roughness=1.0), "eval", wi=[[0, 0, 1]], wo=[[0, 0, 1]], color=(1, 1, 1))[0]
    r0c = 0.04
    np.testing.assert_allclose(f2 - f1, 0.25 * (0.04 + r0c * 0.96) * d1 / 4.0, rtol=1e-6)


@pytest.mark.parametrize("tint", [0.0, 0.5, 1.0])
def test_disney_sheen_lobe_isolated(tint):
    """f(sheen = s) - f(sheen = 0) = s (1 - L.H)^5 sheen_color (1 - metallic)(1 - specTrans)
    at random directions above the surface (only the diffuse term holds sheen)."""
    rng = np.random.default_rng(7)
    n = 500
    wo = rng.normal(size=(n, 3)); wo[:, 2] = np.abs(wo[:, 2]) + 0.05
    wi = rng.normal(size=(n, 3)); wi[:, 2] = np.abs(wi[:, 2]) + 0.05
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    color = (0.7, 0.2, 0.4)
    base = dict(roughness=0.6, metallic=0.3, spec_trans=0.2, sheen_tint=tint)
    m1 = scenes.material(sheen=0.75, **base)
    m0 = scenes.material(sheen=0.0, **base)
    d = O.bxdf(m1, "eval", wi=wi, wo=wo, color=color) - O.bxdf(m0, "eval", wi=wi, wo=wo, color=color)
    h = wi + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    lh = np.sum(wi * h, axis=1)
    p = _params(m1)
    _, _, _, sheen_color = _colors(p, color)
    exp = (p["sheen"] * (1 - lh) ** 5)[:, None] * sheen_color[None, :] * (1 - p["metallic"]) * (1 - p["spec_trans"])
    np.testing.assert_allclose(d, exp, rtol=1e-9, atol=1e-15)


def test_closed_forms_catch_constant_slips():
    """The normal-incidence and 60-degree expectations separate the reference's
    constants from plausible slips: a 0.04 specular scale instead of 0.08, Fd90
    = 0.5 + L_H^2 r, and clearcoat 0.5 instead of 0.25 each move the value by
    far more than the test tolerance."""
    m = scenes.material(roughness=0.8, subsurface=0.0, clearcoat=0.6, clearcoat_gloss=0.2)
    p = _params(m)
    color = (0.8, 0.3, 0.1)
    f60 = O.bxdf(m, "eval", wi=[[math.sqrt(3) / 2, 0, 0.5]], wo=[[0, 0, 1]], color=color)[0]
    np.testing.assert_allclose(f60, expected_60(p, color), rtol=1e-12)
    p_spec = dict(p, specular=p["specular"] / 2)          # what a 0.04 scale would give
    assert np.max(np.abs(expected_60(p_spec, color) / f60 - 1)) > 1e-3
    # Fd90 = 0.5 + L_H^2 r (a missing factor 2) changes the diffuse term by color_diffuse_delta
    assert np.max(np.abs(color_diffuse_delta(p, color) / f60)) > 1e-4
    p_cc = dict(p, clearcoat=2 * p["clearcoat"])          # 0.5 instead of 0.25
    assert np.max(np.abs(expected_60(p_cc, color) / f60 - 1)) > 1e-3


def color_diffuse_delta(p, color):
    """Change of the 60-degree diffuse term when Fd90 = 0.5 + L_H^2 r instead of 0.5 + 2 L_H^2 r."""
    color = np.asarray(color, np.float64)
    r, m, st = p["roughness"], p["metallic"], p["spec_trans"]
    FL = 0.5 ** 5
    right = 1.0 + (0.5 + 1.5 * r - 1.0) * FL
    wrong = 1.0 + (0.5 + 0.75 * r - 1.0) * FL
    return color * INV_PI * (right - wrong) * (1 - m) * (1 - st)
