"""wv_check_finite_noalloc (no device allocation; scratch in a static device array)
against wv_check_finite, through the C ABI, on two calls in a row where the second
handle is smaller than the first: the status must agree, the count of non-finite
values must equal (reference-sequence path, no padding) or bound from above (padded
planes of the hand-written kernels) the host's count over the read-back wavefields
and seismograms, and a finite state after a non-finite one must report zero (the
scratch is cleared at every call)."""
import numpy as np
import pytest

from wavesim import cases
from wavesim._abi import WV_ENONFINITE

pytestmark = pytest.mark.gpu


def _host_nonfinite(sol, case):
    names = sol.FIELDS[(sol.dimension, sol.type)]
    bad = 0
    for k in range(len(case.shots)):
        for f in range(len(names)):
            bad += int((~np.isfinite(sol.wavefield(f, shot=k))).sum())
        bad += int((~np.isfinite(sol.seismograms(k))).sum())
    return bad


def _status(sol):
    from wavesim import WaveError

    try:
        sol.check_finite()
        return 0
    except WaveError:
        return WV_ENONFINITE


def _unstable(case):
    case.settings["DT"] *= 4.0
    return case


@pytest.mark.parametrize("big,small,kw,exact", [
    (lambda: _unstable(cases.elastic_c2(N=96, NT=400, W=12)), lambda: cases.elastic_c2(N=64, NT=300, W=10),
     dict(path="strip", tile_y=16), False),
    (lambda: _unstable(cases.tmem_c3(NX=96, NY=64, NT=400, nshots=3, W=10)),
     lambda: cases.tmem_c3(NX=64, NY=48, NT=300, nshots=2, W=8), dict(), False),
    (lambda: _unstable(cases.acoustic_c1(N=80, NT=400)), lambda: cases.acoustic_c1(N=60, NT=300),
     dict(path="plain"), True),
    (lambda: _unstable(cases.elastic_c5(N=40, NT=200, nshots=2, W=8, nrec_side=8, spacing=2)),
     lambda: cases.elastic_c5(N=32, NT=150, nshots=1, W=6, nrec_side=6, spacing=2), dict(), False),
])
def test_noalloc_matches_check_finite_two_calls(big, small, kw, exact):
    c1 = big()
    s1 = cases.make_solver(c1, **kw)
    s1.run(0, c1.settings["NT"])
    rc, n = s1.check_finite_noalloc()
    host = _host_nonfinite(s1, c1)
    print(f"{c1.name} (unstable DT): noalloc status {rc}, {n} non-finite values; host count {host}", flush=True)
    assert _status(s1) == WV_ENONFINITE
    assert rc == WV_ENONFINITE
    assert host > 0 and (n == host if exact else n >= host)
    rc2, n2 = s1.check_finite_noalloc()  # same state again: same count
    assert (rc2, n2) == (rc, n)
    del s1

    c2 = small()
    s2 = cases.make_solver(c2, **kw)
    s2.run(0, c2.settings["NT"])
    rc, n = s2.check_finite_noalloc()
    print(f"{c2.name} (stable, smaller): noalloc status {rc}, {n} non-finite values", flush=True)
    assert _status(s2) == 0
    assert (rc, n) == (0, 0)
    assert _host_nonfinite(s2, c2) == 0
