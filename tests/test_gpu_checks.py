"""The per-shot non-finite check (Simulation.cpp:519) on the GPU.

An unstable DT (well above the CFL bound, which wv_check_parameters rejects)
drives the fields to Inf/NaN -- ordinary floating-point overflow, no fault --
and wv_check_finite must then report WV_ENONFINITE with the reference's
message; a stable run must pass.  Covered on the 2D two-launch and row-strip
kernels, the reference-sequence path and the 3D kernels.
"""
import numpy as np
import pytest

from wavesim import WaveError, cases
from wavesim.solver import check_parameters

pytestmark = pytest.mark.gpu


def _unstable(case, factor=4.0):
    case.settings["DT"] = case.settings["DT"] * factor
    return case


@pytest.mark.parametrize("make,kw", [
    (lambda: cases.elastic_c2(N=96, NT=400, W=12), dict(path="strip", tile_y=16)),
    (lambda: cases.elastic_c2(N=96, NT=400, W=12), dict(path="phased", tile_y=8)),
    (lambda: cases.acoustic_c1(N=80, NT=400), dict(path="plain")),
    (lambda: cases.tmem_c3(NX=96, NY=64, NT=400, nshots=3, W=10), dict()),
    (lambda: cases.elastic_c5(N=40, NT=200, nshots=2, W=8, nrec_side=8, spacing=2), dict()),
])
def test_nonfinite_detected_and_stable_run_passes(make, kw):
    good = make()
    sol = cases.make_solver(good, **kw)
    sol.run(0, good.settings["NT"])
    sol.check_finite()  # stable: no error

    bad = _unstable(make())
    with pytest.raises(WaveError, match="Courant-Friedrichs-Lewy"):
        fc = {"acoustic": 50.0, "elastic": 30.0, "tmem": 2e8}[bad.type]
        check_parameters(bad.dimension, bad.type, bad.settings, bad.m0, bad.m1, bad.m2, fc_max=fc)
    sol = cases.make_solver(bad, **kw)
    sol.run(0, bad.settings["NT"])
    seis = sol.seismograms(0)
    assert not np.isfinite(seis).all() or not np.isfinite(sol.wavefield(0)).all()
    with pytest.raises(WaveError, match=r"status -5: Infinite or NaN"):
        sol.check_finite()
