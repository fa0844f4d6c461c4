"""The real spherical-harmonic evaluator (rbfx_sph_harm) that generates the multi-field
workloads, against the closed forms the reference itself uses (src/utils/spherical_harmonics.h:
39-170, Mathematica CForm output: Y_3^2, Y_10^5, Y_20^20, valid off the unit sphere too) and
against scipy's complex Y_l^m (an independent implementation) for every l <= 20."""
import numpy as np
from scipy.special import sph_harm_y

import rbffd_gpu_amd as R

PI = np.pi


def points():
    # directions at several radii (the closed forms depend on the direction only)
    xyz = R.nodes_fibonacci(600, 1e-3, 5)
    return xyz * np.linspace(0.5, 2.0, xyz.shape[0])[:, None]


def test_matches_reference_closed_forms():
    xyz = points()
    x, y, z = xyz.T
    r2 = x * x + y * y + z * z
    rho2 = x * x + y * y
    at = np.arctan2(y, x)  # Mathematica ArcTan[Xx, Yy]
    y32 = (np.sqrt(105 / PI) * (x - y) * (x + y) * z) / (4.0 * r2 ** 1.5)
    y105 = (-3 * np.sqrt(1001 / (2.0 * PI)) * rho2 ** 2.5 * z * np.sqrt(1 / r2) *
            (15 * rho2 ** 2 - 140 * rho2 * z * z + 168 * z ** 4) * np.cos(5 * at)) / (128.0 * r2 ** 4.5)
    y2020 = (3 * np.sqrt(156991880045 / (2.0 * PI)) * rho2 ** 10 * np.cos(20 * at)) / (524288.0 * r2 ** 10)
    for (l, m), ref in (((3, 2), y32), ((10, 5), y105), ((20, 20), y2020)):
        got = R.sph_harm(l, m, xyz)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (l, m)
    # the sum class Sph32105 is the sum of the two
    assert np.allclose(R.sph_harm(3, 2, xyz) + R.sph_harm(10, 5, xyz), y32 + y105, atol=1e-13)


def test_matches_scipy_for_all_degrees_up_to_20():
    xyz = points()
    r = np.linalg.norm(xyz, axis=1)
    theta = np.arccos(np.clip(xyz[:, 2] / r, -1, 1))
    phi = np.arctan2(xyz[:, 1], xyz[:, 0])
    for l in range(21):
        for m in range(-l, l + 1):
            c = sph_harm_y(l, abs(m), theta, phi)  # Condon-Shortley phase included
            ref = c.real if m == 0 else np.sqrt(2) * (c.real if m > 0 else c.imag)
            assert np.abs(R.sph_harm(l, m, xyz) - ref).max() < 1e-12, (l, m)


def test_orthonormal_on_the_sphere():
    """Gauss-Legendre x uniform quadrature in (cos theta, phi), exact for these degrees."""
    ct, wt = np.polynomial.legendre.leggauss(24)
    ph = np.arange(48) * 2 * PI / 48
    C, P = np.meshgrid(ct, ph, indexing="ij")
    s = np.sqrt(1 - C * C)
    xyz = np.column_stack([(s * np.cos(P)).ravel(), (s * np.sin(P)).ravel(), C.ravel()])
    w = (wt[:, None] * np.full(48, 2 * PI / 48)[None, :]).ravel()
    lm = [(l, m) for l in range(0, 8) for m in range(-l, l + 1)]
    Y = np.stack([R.sph_harm(l, m, xyz) for l, m in lm])
    G = (Y * w) @ Y.T
    assert np.abs(G - np.eye(len(lm))).max() < 1e-12


def test_fields_are_seeded():
    xyz = R.nodes_fibonacci(300, 1e-3, 0)
    a, spec = R.make_fields(xyz, 4, seed=3)
    b, _ = R.make_fields(xyz, 4, seed=3)
    assert np.array_equal(a, b) and len(spec) == 4 and all(l <= 20 for f in spec for l, _, _ in f)
    f0 = sum(c * R.sph_harm(l, m, xyz) for l, m, c in spec[0])
    assert np.allclose(a[:, 0], f0, rtol=0, atol=1e-13)
