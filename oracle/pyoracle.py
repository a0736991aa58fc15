"""ctypes driver for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads liboracle_f32.so / liboracle_f64.so built from
wave_oracle.c by the Makefile next to it.

Besides the ctypes binding it holds the small pieces of the reference's input
handling the tests need to drive the oracle from the reference's own files:
  * KEY=VALUE configuration parsing   (src/Configuration/Configuration.cpp:45-87)
  * source / receiver txt files       (src/Acquisition/AcquisitionSettings.hpp:55-100, 238-270)
  * MatrixMarket array read/write     (src/IO/IO.hpp:96-121; column-major values)
  * the TwoLayer CI model             (src/Tools/CreateModel/TwoLayer.cpp:25-86)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PHYSICS = {"acoustic": 0, "elastic": 1, "tmem": 2, "sh": 3, "viscoelastic": 4, "visco": 4, "viscosh": 5,
           "emem": 6, "viscotmem": 7, "viscoemem": 8}


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("physics", C.c_int),
        ("NX", C.c_int),
        ("NY", C.c_int),
        ("NZ", C.c_int),
        ("NT", C.c_int),
        ("order", C.c_int),
        ("stencil", C.c_int),
        ("free_surface", C.c_int),
        ("damping", C.c_int),
        ("width", C.c_int),
        ("nthreads", C.c_int),
        ("DH", C.c_double),
        ("DT", C.c_double),
        ("damping_coeff", C.c_double),
        ("npower", C.c_double),
        ("fc_cpml", C.c_double),
        ("vmax_cpml", C.c_double),
    ]


_LIBS = {}


def build():
    """Compile the oracle libraries in place (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(precision="f32"):
    if precision not in _LIBS:
        path = os.path.join(HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.oracle_run.restype = C.c_int
        L.oracle_operator.restype = C.c_long
        _LIBS[precision] = L
    return _LIBS[precision]


def dtype_of(precision):
    return np.float32 if precision == "f32" else np.float64


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(**kw) -> Params:
    p = Params()
    defaults = dict(dim=2, physics=1, NX=0, NY=0, NZ=1, NT=0, order=8, stencil=0, free_surface=0, damping=0, width=0,
                    nthreads=0, DH=1.0, DT=1e-4, damping_coeff=8.0, npower=4.0, fc_cpml=5.0, vmax_cpml=3500.0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def run(params: Params, m0, m1, m2, src_idx, src_type, src_sig, rec_idx, rec_type, precision="f32", want_fields=False,
        t_end=-1, visco=None):
    """Run one shot; returns (seis [nrec, NT], fields or None).  Viscoelastic, visco-SH and visco-EM
    (physics 4, 5, 7, 8) need visco = dict(freqs=[relaxation frequencies, one per mechanism],
    tauP=..., tauS=... per cell); visco-EM: tauP = tau of the dielectric permittivity, tauS = tau of
    the electric conductivity [s]."""
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    n = params.NX * params.NY * params.NZ
    m0 = np.ascontiguousarray(m0, dtype=dt).ravel()
    m1 = np.ascontiguousarray(m1 if m1 is not None else np.zeros(n), dtype=dt).ravel()
    m2 = np.ascontiguousarray(m2, dtype=dt).ravel()
    src_idx = np.ascontiguousarray(src_idx, dtype=np.int64)
    src_type = np.ascontiguousarray(src_type, dtype=np.int32)
    src_sig = np.ascontiguousarray(src_sig, dtype=dt)
    rec_idx = np.ascontiguousarray(rec_idx, dtype=np.int64)
    rec_type = np.ascontiguousarray(rec_type, dtype=np.int32)
    seis = np.zeros((len(rec_idx), params.NT), dtype=dt)
    nf = {0: 3 if params.dim == 2 else 4, 1: 5 if params.dim == 2 else 9, 2: 3, 3: 3,
          4: 5 if params.dim == 2 else 9, 5: 3, 6: 3 if params.dim == 2 else 6, 7: 3,
          8: 3 if params.dim == 2 else 6}[params.physics]
    fields = np.zeros((nf, n), dtype=dt) if want_fields else None
    if params.physics in (4, 5, 7, 8):
        freqs = np.ascontiguousarray(visco["freqs"], dtype=np.float64)
        tp = np.ascontiguousarray(np.broadcast_to(visco["tauP"], (n,)), dtype=dt)
        ts = np.ascontiguousarray(np.broadcast_to(visco["tauS"], (n,)), dtype=dt)
        if L.oracle_set_visco(C.c_int(len(freqs)), _ptr(freqs, C.c_double), _ptr(tp, real), _ptr(ts, real)) != 0:
            raise RuntimeError("oracle_set_visco: 1..4 mechanisms")
    rc = L.oracle_run(C.byref(params), _ptr(m0, real), _ptr(m1, real), _ptr(m2, real), C.c_int(len(src_idx)),
                      _ptr(src_idx, C.c_long), _ptr(src_type, C.c_int), _ptr(src_sig, real), C.c_int(len(rec_idx)),
                      _ptr(rec_idx, C.c_long), _ptr(rec_type, C.c_int), _ptr(seis, real),
                      _ptr(fields, real) if want_fields else None, C.c_int(t_end))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed with {rc}")
    return seis, fields


def run_fused_elastic2d(params: Params, vp, vs, rho, src_idx, src_type, src_sig, rec_idx, rec_type, precision="f32",
                        t_end=-1):
    """Fused-loop CPU mode of the 2D elastic time loop (two passes per step); returns seis [nrec, NT].

    Same per-cell arithmetic as run(); the CPU baseline's best-effort mode."""
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    a = [np.ascontiguousarray(m, dtype=dt).ravel() for m in (vp, vs, rho)]
    src_idx = np.ascontiguousarray(src_idx, dtype=np.int64)
    src_type = np.ascontiguousarray(src_type, dtype=np.int32)
    src_sig = np.ascontiguousarray(src_sig, dtype=dt)
    rec_idx = np.ascontiguousarray(rec_idx, dtype=np.int64)
    rec_type = np.ascontiguousarray(rec_type, dtype=np.int32)
    seis = np.zeros((len(rec_idx), params.NT), dtype=dt)
    rc = L.oracle_run_fused_elastic2d(C.byref(params), _ptr(a[0], real), _ptr(a[1], real), _ptr(a[2], real),
                                      C.c_int(len(src_idx)), _ptr(src_idx, C.c_long), _ptr(src_type, C.c_int),
                                      _ptr(src_sig, real), C.c_int(len(rec_idx)), _ptr(rec_idx, C.c_long),
                                      _ptr(rec_type, C.c_int), _ptr(seis, real), C.c_int(t_end))
    if rc != 0:
        raise RuntimeError(f"oracle_run_fused_elastic2d failed with {rc}")
    return seis


def adjoint_elastic2d(params: Params, vp, vs, rho, src_idx, src_type, src_sig, rec_idx, rec_type, obs,
                      precision="f64"):
    """2D elastic misfit J = 1/2 ||seis - obs||^2 and its adjoint-state gradient.

    Returns (seis [nrec, NT], J, gcoef [5, n] w.r.t. (1/rho_x, 1/rho_y, M, mu,
    mu_xy), gmodel [3, n] w.r.t. (vp, vs, rho)); gradients are FP64.
    """
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    n = params.NX * params.NY
    a = [np.ascontiguousarray(m, dtype=dt).ravel() for m in (vp, vs, rho)]
    src_idx = np.ascontiguousarray(src_idx, dtype=np.int64)
    src_type = np.ascontiguousarray(src_type, dtype=np.int32)
    src_sig = np.ascontiguousarray(src_sig, dtype=dt)
    rec_idx = np.ascontiguousarray(rec_idx, dtype=np.int64)
    rec_type = np.ascontiguousarray(rec_type, dtype=np.int32)
    obs = np.ascontiguousarray(obs, dtype=dt).reshape(len(rec_idx), params.NT)
    seis = np.zeros((len(rec_idx), params.NT), dtype=dt)
    J = C.c_double()
    gcoef = np.zeros((5, n), dtype=np.float64)
    gmodel = np.zeros((3, n), dtype=np.float64)
    rc = L.oracle_adjoint_elastic2d(C.byref(params), _ptr(a[0], real), _ptr(a[1], real), _ptr(a[2], real),
                                    C.c_int(len(src_idx)), _ptr(src_idx, C.c_long), _ptr(src_type, C.c_int),
                                    _ptr(src_sig, real), C.c_int(len(rec_idx)), _ptr(rec_idx, C.c_long),
                                    _ptr(rec_type, C.c_int), _ptr(obs, real), _ptr(seis, real), C.byref(J),
                                    _ptr(gcoef, C.c_double), _ptr(gmodel, C.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint_elastic2d failed with {rc}")
    return seis, J.value, gcoef, gmodel


def adjoint_elastic3d(params: Params, vp, vs, rho, src_idx, src_type, src_sig, rec_idx, rec_type, obs,
                      precision="f64"):
    """3D elastic misfit J = 1/2 ||seis - obs||^2 and its adjoint-state gradient.

    Returns (seis [nrec, NT], J, gcoef [8, n] w.r.t. (1/rho_x, 1/rho_y,
    1/rho_z, M, mu, mu_xy, mu_xz, mu_yz), gmodel [3, n] w.r.t. (vp, vs,
    rho)); gradients are FP64.  Receiver types 1 pressure, 2 vx, 3 vy, 4 vz.
    """
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    n = params.NX * params.NY * params.NZ
    a = [np.ascontiguousarray(m, dtype=dt).ravel() for m in (vp, vs, rho)]
    src_idx = np.ascontiguousarray(src_idx, dtype=np.int64)
    src_type = np.ascontiguousarray(src_type, dtype=np.int32)
    src_sig = np.ascontiguousarray(src_sig, dtype=dt)
    rec_idx = np.ascontiguousarray(rec_idx, dtype=np.int64)
    rec_type = np.ascontiguousarray(rec_type, dtype=np.int32)
    obs = np.ascontiguousarray(obs, dtype=dt).reshape(len(rec_idx), params.NT)
    seis = np.zeros((len(rec_idx), params.NT), dtype=dt)
    J = C.c_double()
    gcoef = np.zeros((8, n), dtype=np.float64)
    gmodel = np.zeros((3, n), dtype=np.float64)
    rc = L.oracle_adjoint_elastic3d(C.byref(params), _ptr(a[0], real), _ptr(a[1], real), _ptr(a[2], real),
                                    C.c_int(len(src_idx)), _ptr(src_idx, C.c_long), _ptr(src_type, C.c_int),
                                    _ptr(src_sig, real), C.c_int(len(rec_idx)), _ptr(rec_idx, C.c_long),
                                    _ptr(rec_type, C.c_int), _ptr(obs, real), _ptr(seis, real), C.byref(J),
                                    _ptr(gcoef, C.c_double), _ptr(gmodel, C.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint_elastic3d failed with {rc}")
    return seis, J.value, gcoef, gmodel


def plant_cpml_swap(psi_index, precision="f64"):
    """Tests only: make the oracle's elastic loops (2D and 3D) use the full-grid CPML coefficients
    where the half-grid ones belong (and the reverse) for one memory variable, given by its index
    in the loop (2D: 0..7, 3D: 0..17); -1 restores the correct placement."""
    lib(precision).oracle_plant_cpml_swap(C.c_int(psi_index))


def ricker(NT, DT, fc, amp, tshift=0.0, precision="f32"):
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    out = np.zeros(NT, dtype=dt)
    L.oracle_ricker(C.c_int(NT), C.c_double(DT), C.c_double(fc), C.c_double(amp), C.c_double(tshift), _ptr(out, real))
    return out


def fd_coef(order):
    out = np.zeros(order, dtype=np.float64)
    if lib("f64").oracle_fd_coef(C.c_int(order), _ptr(out, C.c_double)) != 0:
        raise ValueError(order)
    return out


def operator(params: Params, which, precision="f64"):
    """Assembled CSR operator: which = 0 Dxf 1 Dxb 2 Dyf 3 Dyb 4 Dzf 5 Dzb 6 DyfFS 7 DybFS -> dense ndarray."""
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    nnz = L.oracle_operator(C.byref(params), C.c_int(which), None, None, None, C.c_long(0))
    if nnz < 0:
        return None
    n = params.NX * params.NY * params.NZ
    rp = np.zeros(n + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int64)
    val = np.zeros(nnz, dtype=dt)
    L.oracle_operator(C.byref(params), C.c_int(which), _ptr(rp, C.c_long), _ptr(col, C.c_long), _ptr(val, real),
                      C.c_long(nnz))
    D = np.zeros((n, n), dtype=np.float64)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            D[i, col[k]] += val[k]
    return D


def elastic_coefficients(params: Params, vp, vs, rho, kind, precision="f32"):
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    n = params.NX * params.NY * params.NZ
    a = [np.ascontiguousarray(x, dtype=dt).ravel() for x in (vp, vs, rho)]
    out = np.zeros(n, dtype=dt)
    L.oracle_elastic_coefficients(C.byref(params), _ptr(a[0], real), _ptr(a[1], real), _ptr(a[2], real), C.c_int(kind),
                                  _ptr(out, real))
    return out


def cpml_profile(W, npower, fc, vmax, DT, DH, shift, precision="f32"):
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    a = np.zeros(W, dtype=dt)
    b = np.zeros(W, dtype=dt)
    L.oracle_cpml_profile(C.c_int(W), C.c_double(npower), C.c_double(fc), C.c_double(vmax), C.c_double(DT),
                          C.c_double(DH), C.c_int(shift), _ptr(a, real), _ptr(b, real))
    return a, b


def abs_layout(params: Params, precision="f32"):
    L = lib(precision)
    dt = dtype_of(precision)
    real = C.c_float if precision == "f32" else C.c_double
    out = np.zeros(params.NX * params.NY * params.NZ, dtype=dt)
    L.oracle_abs_layout(C.byref(params), _ptr(out, real))
    return out


# --------------------------------------------------------------------------
# reference input formats
# --------------------------------------------------------------------------
def read_config(path):
    """KEY=VALUE file: '#' starts a comment, keys are lower-cased, later keys win."""
    cfg = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0]
            if "=" not in line:
                continue
            k, v = line.split("=", 1)
            k = k.strip().lower()
            v = v.strip()
            if not k:
                continue
            cfg[k] = v.split()[0] if v.split() else ""
    return cfg


def cfg_num(cfg, key, typ=float, default=None):
    key = key.lower()
    if key not in cfg:
        if default is None:
            raise KeyError(key)
        return default
    v = cfg[key]
    if typ is int:
        import re
        m = re.match(r"[+-]?\d+", v)
        return int(m.group(0))
    return typ(v.rstrip(";"))


def read_sources(path):
    rows = []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.lstrip().startswith("#"):
                continue
            t = line.split()
            rows.append(dict(shot=int(t[0]), x=int(t[1]), y=int(t[2]), z=int(t[3]), type=int(t[4]), wtype=int(t[5]),
                             shape=int(t[6]), fc=float(t[7]), amp=float(t[8]), tshift=float(t[9])))
    return rows


def read_receivers(path):
    rows = []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.lstrip().startswith("#"):
                continue
            t = line.split()
            rows.append(dict(x=int(t[0]), y=int(t[1]), z=int(t[2]), type=int(t[3])))
    return rows


def read_mtx_array(path):
    """MatrixMarket 'array real general': header rows cols, values column-major."""
    with open(path) as f:
        lines = [l for l in f if not l.startswith("%")]
    r, c = (int(x) for x in lines[0].split()[:2])
    vals = np.array([float(x) for x in lines[1:1 + r * c]], dtype=np.float64)
    return vals.reshape(c, r).T.copy()


def two_layer_model(NX, NY, NZ):
    """TwoLayer.cpp:25-86: vp/vs/rho 3500/2000/2000 above depth 40, 4550/2600/2600 below."""
    shape = (NY, NZ, NX)
    vp = np.full(shape, 3500.0)
    vs = np.full(shape, 2000.0)
    rho = np.full(shape, 2000.0)
    vp[40:] = 4550.0
    vs[40:] = 2600.0
    rho[40:] = 2600.0
    return vp.ravel(), vs.ravel(), rho.ravel()


def ci_case(cfg_path):
    """Assemble oracle inputs for one of the reference's par/ci cases."""
    d = os.path.dirname(cfg_path)
    cfg = read_config(cfg_path)
    dim = 3 if cfg["dimension"].lower() == "3d" else 2
    physics = PHYSICS[cfg["equationtype"].lower()]
    NX, NY, NZ = (cfg_num(cfg, k, int) for k in ("NX", "NY", "NZ"))
    DT = cfg_num(cfg, "DT")
    T = cfg_num(cfg, "T")
    NT = int(np.float32(T) / np.float32(DT) + np.float32(0.5))
    stencil = cfg_num(cfg, "useStencilMatrix", int, 0)
    p = make_params(dim=dim, physics=physics, NX=NX, NY=NY, NZ=NZ, NT=NT, order=cfg_num(cfg, "spatialFDorder", int),
                    stencil=stencil, free_surface=cfg_num(cfg, "FreeSurface", int),
                    damping=cfg_num(cfg, "DampingBoundary", int), width=cfg_num(cfg, "BoundaryWidth", int),
                    DH=cfg_num(cfg, "DH"), DT=DT, damping_coeff=cfg_num(cfg, "DampingCoeff"),
                    npower=cfg_num(cfg, "NPower", default=4.0),
                    fc_cpml=cfg_num(cfg, "CenterFrequencyCPML", default=5.0),
                    vmax_cpml=cfg_num(cfg, "VMaxCPML", default=3500.0))
    if cfg_num(cfg, "ModelRead", int) == 1:
        vp, vs, rho = two_layer_model(NX, NY, NZ)
    else:
        n = NX * NY * NZ
        vp = np.full(n, cfg_num(cfg, "velocityP"))
        vs = np.full(n, cfg_num(cfg, "velocityS"))
        rho = np.full(n, cfg_num(cfg, "rho"))
    base = os.path.join(d, os.path.basename(cfg["sourcefilename"]) + ".txt")
    srcs = read_sources(base)
    recs = read_receivers(os.path.join(d, os.path.basename(cfg["receiverfilename"]) + ".txt"))

    def lin(r):
        return r["x"] + r["z"] * NX + r["y"] * NX * NZ

    ref_name = os.path.basename(cfg["seismogramfilename"]).split(".ci")[0] + ".ref"
    visco = None
    if physics in (4, 5):
        # Viscoelastic.cpp:160-200: numRelaxationMechanisms, relaxationFrequency(2..4); tauP/tauS
        # from the model files (TwoLayer.cpp:25-26, 83-86: 0.1 in both layers) or the configuration
        nl = cfg_num(cfg, "numRelaxationMechanisms", int)
        keys = ["relaxationFrequency", "relaxationFrequency2", "relaxationFrequency3", "relaxationFrequency4"]
        freqs = [cfg_num(cfg, keys[l]) for l in range(nl)]
        if cfg_num(cfg, "ModelRead", int) == 1:
            tp = ts = np.full(NX * NY * NZ, 0.1)
        else:
            tp = np.full(NX * NY * NZ, cfg_num(cfg, "tauP"))
            ts = np.full(NX * NY * NZ, cfg_num(cfg, "tauS"))
        visco = dict(freqs=freqs, tauP=tp, tauS=ts)
    return dict(cfg=cfg, params=p, vp=vp, vs=vs, rho=rho, sources=srcs, receivers=recs, visco=visco,
                src_idx=[lin(s) for s in srcs], src_type=[s["type"] for s in srcs],
                rec_idx=[lin(r) for r in recs], rec_type=[r["type"] for r in recs], ref_base=os.path.join(d, ref_name))


def ci_misfit(test, ref):
    """Test_CompareSeismogram.cpp:62-84 criterion: sum ||test-ref||_2 / (max * nsamples * ntraces)."""
    test = np.asarray(test, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    mx = max(np.abs(ref).max(), np.abs(test).max())
    return np.linalg.norm(test - ref) / (mx * ref.shape[1] * ref.shape[0])
