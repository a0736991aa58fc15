"""Closed-form and invariant checks that pin the oracle where no golden file does.

Each check is chosen so that a plausible restatement mistake fails it:
  * FD coefficients: Taylor moment conditions (wrong sign/ordering/value fails)
  * sparse operators: exactness on polynomials row by row, which fixes the
    local (reduced) order and the inward shift at the outermost points
  * free-surface image operators: antisymmetric image rule at y = 0
  * Ricker, CPML and ABS profiles: their formulas evaluated independently in FP64
  * averaging: arithmetic (density) and harmonic (shear modulus) means with
    the clamps of ModelparameterSeismic.cpp:421-433
  * linearity in the source amplitude and FP32-vs-FP64 agreement of a CPML run
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10, 12])
def test_fd_coefficients_moments(order):
    c = [Fraction(x) for x in po.fd_coef(order)]
    offs = [Fraction(2 * (j - order // 2) + 1, 2) for j in range(order)]  # half-point offsets
    for m in range(order):
        s = sum(cj * o ** m for cj, o in zip(c, offs))
        scale = sum(abs(cj * o ** m) for cj, o in zip(c, offs))
        want = 1 if m == 1 else 0
        # the coefficients are the correctly rounded doubles of the exact rationals
        assert abs(float(s) - want) <= 4e-16 * float(scale) * order, (order, m, float(s))


def _params(NX=14, NY=11, order=8, **kw):
    return po.make_params(NX=NX, NY=NY, NZ=1, NT=1, order=order, DH=1.0, DT=1.0, **kw)


def _expected_order(c, L, N, forward):
    """Local order of the sparse operator at position c (Derivatives.cpp:166-179, 743-756)."""
    if forward:
        n = min(N, 2 * c + 2, 2 * (L - 1 - c)) if L - 1 - c > 0 else 2
        shift = -1 if L - 1 - c == 0 else 0
    else:
        n = min(N, 2 * c, 2 * (L - c)) if c > 0 else 2
        shift = 1 if c == 0 else 0
    return n, shift


@pytest.mark.parametrize("order", [2, 4, 8, 12])
@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_sparse_operator_rows(order, which):
    NX, NY = 15, 13
    p = _params(NX, NY, order)
    D = po.operator(p, which)
    forward = which in (0, 2)
    axis = 0 if which in (0, 1) else 1
    L = NX if axis == 0 else NY
    for y in range(NY):
        for x in range(NX):
            i = x + y * NX
            c = x if axis == 0 else y
            pos = np.array([(xx if axis == 0 else yy) for yy in range(NY) for xx in range(NX)], float)
            n_exp, shift = _expected_order(c, L, order, forward)
            row = D[i]
            nz = np.nonzero(row)[0]
            assert len(nz) == n_exp, (which, x, y, len(nz), n_exp)
            # derivative location: half a cell forward/backward of the (shifted) centre
            at = c + shift + (0.5 if forward else -0.5)
            assert abs(row @ np.ones_like(pos)) < 1e-12
            assert abs(row @ pos - 1.0) < 1e-12
            assert abs(row @ pos ** 2 - 2 * at) < 1e-10
            # exact for the local order
            if n_exp >= 4:
                assert abs(row @ pos ** 3 - 3 * at ** 2) < 1e-8
            else:
                assert abs(row @ pos ** 3 - 3 * at ** 2 - 0.25) < 1e-8
            # all points stay inside the grid and on the row's line
            for k in nz:
                assert (k % NX == x) if axis == 1 else (k // NX == y)


def test_stencil_operators_zero_padded():
    p = _params(9, 7, 4, stencil=1)
    Dxf = po.operator(p, 0)
    Dxb = po.operator(p, 1)
    # backward stencil operator is minus the transpose of the forward one (FDTD2D.cpp:171-174)
    assert np.allclose(Dxb, -Dxf.T)
    # interior rows equal the sparse ones, edge rows are truncated (not reduced)
    sp = po.operator(_params(9, 7, 4), 0)
    assert np.allclose(Dxf[4], sp[4])
    assert np.count_nonzero(Dxf[0]) == 3 and np.count_nonzero(sp[0]) == 2


@pytest.mark.parametrize("order", [4, 8, 12])
def test_free_surface_image_operators(order):
    NX, NY = 5, 16
    p = _params(NX, NY, order, free_surface=1, physics=1)
    dyf = po.operator(p, 2)
    dyffs = po.operator(p, 6)
    dybfs = po.operator(p, 7)
    c = po.fd_coef(order)
    H = order // 2
    for y in range(NY):
        i = y * NX
        for j in range(order):
            # DyfFreeSurface: weight of row y, tap j = c_j - c_image (Derivatives.cpp:408-417)
            Y = y + j - H + 1
            img = order - 2 - 2 * y - j
            w = c[j] - (c[img] if img >= 0 else 0.0)
            if 0 <= Y < NY:
                assert abs(dyffs[i, Y * NX] - w) < 1e-12
            Yb = y + j - H
            imgb = order - 1 - 2 * y - j
            wb = c[j] - (c[imgb] if imgb >= 0 else 0.0)
            if 0 <= Yb < NY:
                assert abs(dybfs[i, Yb * NX] - wb) < 1e-12
    # far from the surface and the bottom the image operator is the plain one
    mid = (NY // 2) * NX
    assert np.allclose(dyffs[mid], dyf[mid])
    # at y = 0 the antisymmetric image removes the stress derivative's surface tap
    assert abs(dyffs[0, 0]) < 1e-12


def test_ricker_closed_form():
    NT, DT, fc, amp = 400, 1e-3, 12.0, 2.5
    s = po.ricker(NT, DT, fc, amp, 0.0, "f64")
    t = np.arange(NT) * DT
    tau = np.pi * fc * (t - 1.5 / fc)
    ref = amp * (1 - 2 * tau ** 2) * np.exp(-tau ** 2)
    assert np.max(np.abs(s - ref)) < 1e-12
    s32 = po.ricker(NT, DT, fc, amp, 0.0, "f32")
    assert np.max(np.abs(s32 - ref)) < 1e-5 * amp
    # peak at t = 1.5/fc (sample 125 here) equals amp
    k = int(round(1.5 / fc / DT))
    assert abs(s[k] - amp * (1 - 2 * (np.pi * fc * (k * DT - 1.5 / fc)) ** 2)) < 1e-3


def test_cpml_profile_closed_form():
    W, npow, fc, vmax, DT, DH = 20, 4.0, 50.0, 2000.0, 2e-4, 1.0
    for half in (0, 1):
        a, b = po.cpml_profile(W, npow, fc, vmax, DT, DH, half, "f64")
        d0 = -(npow + 1) * vmax * np.log(8e-4) / (2 * W * DH)
        pos = (W - np.arange(W) - 0.5 * half) / W
        d = d0 * pos ** npow
        al = np.pi * fc * (1 - pos)
        bb = np.exp(-(d + al) * DT)
        aa = d * (bb - 1) / (d + al)
        assert np.allclose(b, bb, rtol=1e-12) and np.allclose(a, aa, rtol=1e-12)
        assert np.all(a <= 0) and np.all((b > 0) & (b < 1))


def test_abs_profile_closed_form():
    p = po.make_params(NX=30, NY=30, NZ=1, width=9, damping_coeff=8.0, free_surface=0)
    d = po.abs_layout(p, "f64").reshape(30, 30)
    a = np.sqrt(-np.log(1 - 0.08) / 81)
    assert abs(d[0, 15] - np.exp(-(a * 9) ** 2)) < 1e-12   # outermost row: coeff[0]
    assert abs(d[8, 15] - np.exp(-(a * 1) ** 2)) < 1e-12   # innermost frame row: coeff[W-1]
    assert d[15, 15] == 1.0
    # with a free surface the top rows are damped only near the side edges
    p.free_surface = 1
    d = po.abs_layout(p, "f64").reshape(30, 30)
    assert d[0, 15] == 1.0 and abs(d[0, 2] - np.exp(-(a * 7) ** 2)) < 1e-12


def test_averaging():
    NX, NY = 6, 5
    p = po.make_params(NX=NX, NY=NY, NZ=1)
    rng = np.random.default_rng(0)
    rho = rng.uniform(1000, 3000, NX * NY)
    vs = rng.uniform(0, 2000, NX * NY)
    vs[7] = 0.0  # fluid cell: mu < 1 -> clamped to 1
    vp = 2 * vs + 1500
    irx = po.elastic_coefficients(p, vp, vs, rho, 0, "f64").reshape(NY, NX)
    iry = po.elastic_coefficients(p, vp, vs, rho, 1, "f64").reshape(NY, NX)
    muxy = po.elastic_coefficients(p, vp, vs, rho, 3, "f64").reshape(NY, NX)
    mu = po.elastic_coefficients(p, vp, vs, rho, 7, "f64").reshape(NY, NX)
    R = rho.reshape(NY, NX)
    MU = np.maximum(R * vs.reshape(NY, NX) ** 2, 1.0)
    assert np.allclose(mu, MU)
    assert np.allclose(irx[:, :-1], 2 / (R[:, :-1] + R[:, 1:])) and np.allclose(irx[:, -1], 1 / R[:, -1])
    assert np.allclose(iry[:-1], 2 / (R[:-1] + R[1:])) and np.allclose(iry[-1], 1 / R[-1])
    h = 4 / (1 / MU[:-1, :-1] + 1 / MU[:-1, 1:] + 1 / MU[1:, :-1] + 1 / MU[1:, 1:])
    h[h < 4] = 0
    assert np.allclose(muxy[:-1, :-1], h)
    assert muxy[1, 1] == 0.0 and muxy[1, 0] == 0.0  # neighbours of the fluid cell vanish (< 4 Pa)
    assert np.allclose(muxy[-1, :-1], 2 / (1 / MU[-1, :-1] + 1 / MU[-1, 1:]))
    assert np.allclose(muxy[-1, -1], MU[-1, -1])


def _small_elastic(precision, amp=1.0, NT=300):
    N = 48
    p = po.make_params(dim=2, physics=1, NX=N, NY=N, NT=NT, order=8, damping=2, width=10, DH=1.0, DT=1e-4,
                       npower=4.0, fc_cpml=30.0, vmax_cpml=3000.0, nthreads=4)
    n = N * N
    vp = np.full(n, 2500.0)
    vs = vp / 1.8
    rho = np.full(n, 2000.0)
    sig = po.ricker(NT, 1e-4, 60.0, amp, 0.0, precision)[None]
    rec = [24 + 14 * N, 30 + 14 * N]
    return po.run(p, vp, vs, rho, [24 + 24 * N], [3], sig, rec, [2, 3], precision)[0]


def test_linearity_and_precision():
    a = _small_elastic("f64", 1.0)
    b = _small_elastic("f64", 3.0)
    assert np.allclose(b, 3 * a, rtol=1e-10, atol=1e-14 * np.abs(a).max())
    s32 = _small_elastic("f32", 1.0)
    assert np.linalg.norm(s32 - a) / np.linalg.norm(a) < 1e-5
