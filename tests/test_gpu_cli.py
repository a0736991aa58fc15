"""The C++ host driver (wavesim_run) end to end on the reference's CI cases:
configuration file -> model files -> acquisition files -> GPU time loop ->
seismogram .mtx files, compared with the reference's golden seismograms with
the criterion of Test_CompareSeismogram.cpp:84 (misfit <= 5e-7)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CI = os.path.join(ROOT, "tests", "golden", "ci")
EXE = os.path.join(ROOT, "wave-simulation_amd", "bin", "wavesim_run")


def _write_vec(path, v):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{len(v)} 1\n")
        np.savetxt(f, np.asarray(v, np.float64), fmt="%.9g")


@pytest.mark.parametrize("name,comp,extra", [("2D.acoustic", "p", []), ("2D.elastic", "vy", ["useStencilMatrix=1"]),
                                             ("3D.elastic", "vy", [])])
def test_cli_reproduces_golden_seismogram(tmp_path, name, comp, extra):
    cfg_path = os.path.join(CI, f"configuration_ci.{name}.txt")
    cfg = po.read_config(cfg_path)
    NX, NY, NZ = (po.cfg_num(cfg, k, int) for k in ("NX", "NY", "NZ"))
    if name.startswith("2D"):
        NZ = 1
    model = tmp_path / "model"
    model_args = []
    if po.cfg_num(cfg, "ModelRead", int) == 1:
        # the CI models come from the reference's TwoLayer tool (TwoLayer.cpp:25-86)
        vp, vs, rho = po.two_layer_model(NX, NY, NZ)
        for ext, v in (("vp", vp), ("vs", vs), ("density", rho)):
            _write_vec(f"{model}.{ext}.mtx", v)
        model_args = [f"ModelFilename={model}", "fileFormat=1"]
    src = os.path.join(CI, os.path.basename(cfg["sourcefilename"].split()[0]))
    rec = os.path.join(CI, os.path.basename(cfg["receiverfilename"].split()[0]))
    out = tmp_path / "seis"
    args = [EXE, cfg_path, f"SourceFilename={src}", f"ReceiverFilename={rec}", f"SeismogramFilename={out}"]
    args += model_args + extra
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    shot = po.read_sources(src + ".txt")[0]["shot"]
    seis = po.read_mtx_array(f"{out}.shot_{shot}.{comp}.mtx")
    ref = po.read_mtx_array(os.path.join(CI, f"seismogram.{name}.ref.{comp}.mtx"))
    assert seis.shape == ref.shape
    m = po.ci_misfit(seis, ref)
    print(f"{name}: misfit {m:.3e}")
    assert m <= 5e-7


def test_cli_reports_errors(tmp_path):
    r = subprocess.run([EXE, str(tmp_path / "missing.txt")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not found" in r.stderr
