"""Mathematical pins of the oracle's BSDF code (oracle/cpu_ref.cpp: Diffuse_BRDF, Metal_BRDF with GGX
VNDF sampling, conductor and dielectric Fresnel; kennyalive/YAR src/ref/bsdf.cpp:60-200,
scattering.cpp:17-269, sampling.cpp:43,311-359).  No GPU: closed forms, quadrature and Monte Carlo
identities that a dropped term, a wrong sign or a swapped argument would break."""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_lib

DIFFUSE, METAL = 1, 2
F3 = C.c_float * 3


def _lib():
    L = oracle_lib.lib()
    L.oref_bsdf_eval_local.argtypes = [C.c_int32, F3, F3, F3, C.c_float, C.c_float, C.c_float, F3, F3, F3,
                                       C.POINTER(C.c_float)]
    L.oref_bsdf_sample_local.argtypes = [C.c_int32, F3, F3, F3, C.c_float, C.c_float, C.c_float, F3,
                                         C.c_float * 2, F3, F3, C.POINTER(C.c_float)]
    L.oref_bsdf_sample_local.restype = C.c_int32
    L.oref_dielectric_fresnel.argtypes = [C.c_float, C.c_float]
    L.oref_dielectric_fresnel.restype = C.c_float
    L.oref_conductor_fresnel.argtypes = [C.c_float, C.c_float, F3, F3, F3]
    return L


class Bsdf:
    def __init__(self, type_, refl=(0.5, 0.5, 0.5), eta=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.1), ax=0.3, ay=0.3):
        self.L = _lib()
        self.args = (type_, F3(*refl), F3(*eta), F3(*k), 1.0, ax, ay)

    def eval(self, wo, wi):
        f = F3()
        pdf = C.c_float()
        self.L.oref_bsdf_eval_local(*self.args, F3(*wo), F3(*wi), f, C.byref(pdf))
        return np.array(f[:]), pdf.value

    def sample(self, wo, u):
        wi, f, pdf = F3(), F3(), C.c_float()
        ok = self.L.oref_bsdf_sample_local(*self.args, F3(*wo), (C.c_float * 2)(*u), wi, f, C.byref(pdf))
        return bool(ok), np.array(wi[:]), np.array(f[:]), pdf.value


def _hemisphere(n_theta=160, n_phi=320):
    th = (np.arange(n_theta) + 0.5) * (np.pi / 2) / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return d, dw


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def test_dielectric_fresnel_closed_forms():
    L = _lib()
    assert L.oref_dielectric_fresnel(1.0, 1.5) == pytest.approx(0.04, abs=1e-6)        # ((n-1)/(n+1))^2
    assert L.oref_dielectric_fresnel(1.0, 1.0) == pytest.approx(0.0, abs=1e-7)
    # leaving glass past the critical angle: total internal reflection
    assert L.oref_dielectric_fresnel(0.3, 1.0 / 1.5) == 1.0
    for ci in np.linspace(0.05, 1.0, 12):
        eta = 1.5
        st = np.sqrt(1 - ci * ci) / eta
        ct = np.sqrt(1 - st * st)
        rs = (ci - eta * ct) / (ci + eta * ct)
        rp = (eta * ci - ct) / (eta * ci + ct)
        assert L.oref_dielectric_fresnel(float(ci), eta) == pytest.approx(0.5 * (rs * rs + rp * rp), abs=2e-6)


def test_conductor_fresnel_closed_forms():
    L = _lib()
    eta, k = (0.2, 0.9, 1.1), (3.9, 2.4, 2.1)
    out = F3()
    L.oref_conductor_fresnel(1.0, 1.0, F3(*eta), F3(*k), out)
    ref = [((n - 1) ** 2 + kk ** 2) / ((n + 1) ** 2 + kk ** 2) for n, kk in zip(eta, k)]
    assert np.allclose(out[:], ref, rtol=1e-5)
    # with k = 0 the conductor formula is the dielectric one
    for ci in np.linspace(0.1, 1.0, 10):
        L.oref_conductor_fresnel(float(ci), 1.0, F3(1.5, 1.3, 2.0), F3(0, 0, 0), out)
        for c, n in zip(out[:], (1.5, 1.3, 2.0)):
            assert c == pytest.approx(L.oref_dielectric_fresnel(float(ci), n), abs=1e-5)


def test_diffuse_white_furnace_and_sampling():
    b = Bsdf(DIFFUSE, refl=(0.8, 0.5, 0.2))
    wo = _unit([0.3, -0.2, 0.9])
    d, dw = _hemisphere()
    fs = np.array([b.eval(wo, w)[0] for w in d[::7]])
    cos = d[::7, 2]
    est = (fs * (cos * dw[::7] * 7)[:, None]).sum(axis=0)
    assert np.allclose(est, [0.8, 0.5, 0.2], rtol=5e-3)                   # integral of f cos = albedo
    rng = np.random.default_rng(1)
    for u in rng.uniform(size=(200, 2)):
        ok, wi, f, pdf = b.sample(wo, u)
        assert ok and wi[2] > 0 and abs(np.linalg.norm(wi) - 1) < 1e-5
        assert np.allclose(f * wi[2] / pdf, [0.8, 0.5, 0.2], rtol=1e-4)  # cosine sampling: weight = albedo


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.1, 0.4), (0.6, 0.2)])
def test_ggx_sampling_consistent_with_pdf_and_eval(ax, ay):
    b = Bsdf(METAL, ax=ax, ay=ay)
    wo = _unit([0.4, 0.1, 0.7])
    rng = np.random.default_rng(2)
    n = 40000
    acc = np.zeros(3)
    valid = 0
    for u in rng.uniform(size=(n, 2)):
        ok, wi, f, pdf = b.sample(wo, u)
        if not ok:
            continue
        valid += 1
        f2, pdf2 = b.eval(wo, wi)
        assert np.allclose(f, f2, rtol=1e-6) and pdf == pytest.approx(pdf2, rel=1e-5)
        acc += f * wi[2] / pdf
    mc = acc / n
    d, dw = _hemisphere(200, 400)
    ev = [b.eval(wo, w) for w in d]
    fq = np.array([e[0] for e in ev])
    pq = np.array([e[1] for e in ev])
    quad = (fq * (d[:, 2] * dw)[:, None]).sum(axis=0)
    mass = (pq * dw).sum()
    print(f"alpha ({ax}, {ay}): MC {mc}, quadrature {quad}, pdf mass above horizon {mass:.4f}, "
          f"valid fraction {valid / n:.4f}")
    # importance-sampled integral of f cos equals its quadrature; the pdf's mass over the upper
    # hemisphere equals the fraction of samples that reflect above the horizon
    assert np.allclose(mc, quad, rtol=2e-2)
    assert mass == pytest.approx(valid / n, abs=1e-2)
    assert mass <= 1.0 + 5e-3


def test_ggx_reciprocity_and_energy_bound():
    b = Bsdf(METAL, eta=(1.5, 1.5, 1.5), k=(0.0, 0.0, 0.0), ax=0.25, ay=0.25)
    rng = np.random.default_rng(3)
    for _ in range(50):
        wo = _unit(np.append(rng.normal(size=2) * 0.6, 1.0))
        wi = _unit(np.append(rng.normal(size=2) * 0.6, 1.0))
        f1, _ = b.eval(wo, wi)
        f2, _ = b.eval(wi, wo)
        assert np.allclose(f1, f2, rtol=1e-5)
    # a dielectric-Fresnel microfacet reflector never reflects more than it receives
    d, dw = _hemisphere(120, 240)
    wo = _unit([0.2, 0.0, 0.98])
    fq = np.array([b.eval(wo, w)[0] for w in d])
    alb = (fq * (d[:, 2] * dw)[:, None]).sum(axis=0)
    assert np.all(alb > 0.0) and np.all(alb < 1.0)
