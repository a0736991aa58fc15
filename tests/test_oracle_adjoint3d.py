"""The oracle's FP64 3D elastic adjoint-state gradient against central differences.

The reference has no adjoint, so the gradient is pinned by the mathematics:
the discrete adjoint of the oracle's 3D time loop must equal the derivative of
the oracle's own FP64 misfit (oracle_run, ForwardSolver3Delastic.cpp:120-416).
The grid is not a cube (16 x 14 x 12), so a term read along the wrong axis
cannot cancel.  One case has CPML on all six faces.  The other has no boundary
and a source next to the far corner, so the edge fallbacks of the averaging
rules (Modelparameter.cpp:336-526) carry real gradient at the cells checked.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

NX, NY, NZ, NT = 16, 14, 12, 100


def _idx(x, y, z):
    return x + z * NX + y * NX * NZ


def _setup(damping, stencil, corner):
    n = NX * NY * NZ
    p = po.make_params(dim=3, physics=1, NX=NX, NY=NY, NZ=NZ, NT=NT, order=4, stencil=stencil, damping=damping,
                       width=3, DH=1.0, DT=1e-4, npower=4.0, fc_cpml=400.0, vmax_cpml=3000.0, nthreads=4)
    rng = np.random.default_rng(11)
    vp = 2500.0 * (1 + 0.05 * rng.standard_normal(n))
    rho = 2100.0 * (1 + 0.05 * rng.standard_normal(n))
    vs = vp / 1.8 * (1 + 0.03 * rng.standard_normal(n))
    w = po.ricker(NT, 1e-4, 400.0, 1.0, precision="f64")
    if corner:
        src = np.array([_idx(NX - 3, NY - 2, NZ - 3), _idx(NX - 2, NY - 4, NZ - 2)])
    else:
        src = np.array([_idx(6, 5, 5), _idx(9, 7, 6)])
    sty = np.array([1, 3], np.int32)
    sig = np.stack([w, 0.5 * w])
    cells = [_idx(11, 8, 7), _idx(5, 9, 4), _idx(10, 4, 8), _idx(7, 7, 7), _idx(NX - 1, NY - 1, NZ - 1),
             _idx(3, 10, 8)]
    rec = np.repeat(np.array(cells), 4)
    rty = np.tile(np.array([1, 2, 3, 4], np.int32), len(cells))
    m0 = [np.full(n, 2500.0), np.full(n, 2500.0 / 1.8), np.full(n, 2100.0)]
    obs, _ = po.run(p, *m0, src, sty, sig, rec, rty, "f64")
    return p, [vp, vs, rho], (src, sty, sig, rec, rty), obs


def _misfit(p, m, sr, obs):
    s, _ = po.run(p, m[0], m[1], m[2], *sr, "f64")
    return 0.5 * np.sum((s - obs) ** 2)


@pytest.mark.parametrize("damping,stencil,corner", [(2, 0, False), (0, 1, True)])
def test_adjoint3d_gradient_equals_central_differences(damping, stencil, corner):
    p, m, sr, obs = _setup(damping, stencil, corner)
    seis, J, gcoef, gmod = po.adjoint_elastic3d(p, *m, *sr, obs)
    assert J > 0 and np.isclose(J, _misfit(p, m, sr, obs), rtol=1e-12)
    assert np.all(np.isfinite(gcoef)) and np.all(np.isfinite(gmod))
    if corner:  # far faces, edges and the corner: the averaging rules' fallbacks
        X, Y, Z = NX - 1, NY - 1, NZ - 1
        cells = [_idx(X, 9, 8), _idx(12, Y, 8), _idx(12, 9, Z), _idx(X, Y, 8), _idx(X, 9, Z), _idx(12, Y, Z),
                 _idx(X, Y, Z), _idx(13, 11, 9)]
    else:
        rng = np.random.default_rng(5)
        cells = [_idx(int(x), int(y), int(z)) for x, y, z in
                 zip(rng.integers(0, NX, 5), rng.integers(0, NY, 5), rng.integers(0, NZ, 5))]
        cells += [_idx(1, 6, 6), _idx(8, 1, 6), _idx(8, 6, 1)]  # inside the CPML frame
    for pi in range(3):
        scale = np.abs(gmod[pi]).max()
        assert scale > 0
        for i in cells:
            mm = [a.copy() for a in m]
            h = 1e-4 * mm[pi][i]
            mm[pi][i] += h
            jp = _misfit(p, mm, sr, obs)
            mm[pi][i] -= 2 * h
            jm = _misfit(p, mm, sr, obs)
            fd = (jp - jm) / (2 * h)
            # central differences with a relative step of 1e-4 agree to ~1e-6
            # here (9e-6 at the far corner); the absolute slack is below 4e-4 of
            # the smallest gradient checked
            assert abs(fd - gmod[pi][i]) <= 2e-5 * abs(fd) + 1e-10 * scale, (pi, i, fd, gmod[pi][i])


def test_adjoint3d_rejects_free_surface_and_abs():
    p, m, sr, obs = _setup(2, 0, False)
    p.free_surface = 1
    with pytest.raises(RuntimeError):
        po.adjoint_elastic3d(p, *m, *sr, obs)
    p.free_surface = 0
    p.damping = 1
    with pytest.raises(RuntimeError):
        po.adjoint_elastic3d(p, *m, *sr, obs)
