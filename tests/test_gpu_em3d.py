"""3D EM (ForwardSolver3Demem.cpp:117-235) and 3D visco-EM (ForwardSolver3Dviscoemem.cpp:170-305)
on the GPU, through the C ABI.

No reference fixture exists; the oracle is pinned by its line-source reductions to the 2D TM / TE
loops (test_oracle_invariants.py).  The GPU runs blocky epsilon_r / sigma / mu_r models on a
44 x 36 x 30 grid (no two axes equal, ragged against the x and z tiles), EZ / EX / EY / HZ sources
and receivers of all four components, against the oracle's float build (relative L2 <= 1e-5 of
the electric seismogram and of the magnetic one) and FP64 run: CPML at orders 2, 4, 8, 12; ABS at
every z-tile height; one to three relaxation mechanisms; batched == single-shot; graph segments;
the final wavefields.
"""
import numpy as np
import pytest

from wavesim import cases

from _cases import oracle_shot, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
FIELDS = ("hx", "hy", "hz", "ex", "ey", "ez")


def _check(case, fields=False, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    out = []
    nm = len(case.visco["freqs"]) if case.visco else 0
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=fields)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        out.append(g)
        assert np.abs(o32).max() > 0
        # the electric seismogram (EZ, EX, EY traces) and the magnetic one (HZ): their amplitudes
        # differ by the impedance, so each is compared on its own
        for label, m in (("E", sh.rec_type != 4), ("H", sh.rec_type == 4)):
            e32, e64 = rel_l2(g[m], o32[m]), rel_l2(g[m], o64[m])
            print(f"{case.name} order {case.settings['spatial_fd_order']} mechanisms {nm} {kw} shot {k} {label}: "
                  f"relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
            assert e32 <= TOL, (k, label, e32)
            assert e64 <= 1e-4, (k, label, e64)
        if fields:
            for group in (FIELDS[:3], FIELDS[3:]):
                g_f = np.concatenate([sol.wavefield(name, shot=k) for name in group])
                o_f = np.concatenate([f32[FIELDS.index(name)] for name in group])
                ef = rel_l2(g_f, o_f)
                print(f"  final {'/'.join(group)}: relL2 vs f32 oracle {ef:.2e}", flush=True)
                assert ef <= TOL, (group, ef)
    return out


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_em3d_cpml_orders(order):
    _check(cases.emem_3d(order=order, nshots=4 if order == 8 else 2), fields=order == 8)


@pytest.mark.parametrize("tile_y", [4, 8, 16])
def test_em3d_abs_tiles(tile_y):
    _check(cases.emem_3d(damping=1, W=8, nshots=4), fields=tile_y == 8, tile_y=tile_y)


@pytest.mark.parametrize("freqs", [(3e8,), (1e8, 6e8), (5e7, 3e8, 1.2e9)])
def test_visco_em3d_mechanisms(freqs):
    _check(cases.viscoemem_3d(freqs=freqs, nshots=4), fields=len(freqs) == 2)


def test_visco_em3d_abs():
    _check(cases.viscoemem_3d(freqs=(2e8, 8e8), damping=1, W=8, nshots=2), tile_y=16)


@pytest.mark.parametrize("kind", ["emem_3d", "viscoemem_3d"])
def test_em3d_batched_equals_single_and_graphs(kind):
    c = getattr(cases, kind)(nshots=3, NT=200)
    batched = cases.make_solver(c)
    batched.run(0, 200)
    graphed = cases.make_solver(c, steps_per_graph=50)
    graphed.run(0, 200)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, shots=[sh])
        one.run(0, 200)
        assert np.array_equal(one.seismograms(0), batched.seismograms(k)), k
        assert np.array_equal(graphed.seismograms(k), batched.seismograms(k)), k
