"""Python face of the MI355X forward solver.

Mirrors the reference's ForwardSolver interface (src/ForwardSolver/ForwardSolver.hpp:26-70
and ForwardSolverFactory.cpp:4-66): a factory ``create(dimension, type)``,
``init_forward_solver`` (= initForwardSolver), ``prepare_for_modelling``
(= Modelparameter/ForwardSolver prepareForModelling), ``run(t)`` for one step,
``run(t_start, t_end)`` for a segment, ``reset_cpml`` / ``reset_wavefields``.
All work happens in libwavehip.so on the GPU; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ForwardSolver:
    """One solver handle: a grid, a model, nshots acquisitions, device wavefields."""

    FIELDS = {
        ("2D", "acoustic"): ("vx", "vy", "p"),
        ("2D", "elastic"): ("vx", "vy", "sxx", "syy", "sxy"),
        ("2D", "viscoelastic"): ("vx", "vy", "sxx", "syy", "sxy"),
        ("2D", "tmem"): ("hx", "hy", "ez"),
        ("2D", "emem"): ("hz", "ex", "ey"),
        ("2D", "viscotmem"): ("hx", "hy", "ez"),
        ("2D", "viscoemem"): ("hz", "ex", "ey"),
        ("2D", "sh"): ("vz", "sxz", "syz"),
        ("2D", "viscosh"): ("vz", "sxz", "syz"),
        ("3D", "elastic"): ("vx", "vy", "vz", "sxx", "syy", "szz", "syz", "sxz", "sxy"),
        ("3D", "viscoelastic"): ("vx", "vy", "vz", "sxx", "syy", "szz", "syz", "sxz", "sxy"),
        ("3D", "acoustic"): ("vx", "vy", "vz", "p"),
        ("3D", "emem"): ("hx", "hy", "hz", "ex", "ey", "ez"),
        ("3D", "viscoemem"): ("hx", "hy", "hz", "ex", "ey", "ez"),
    }

    def __init__(self, dimension: str, type_: str):
        L = _abi.load()
        self._h = C.c_void_p()
        _abi.check(L.wv_create(dimension.encode(), type_.encode(), C.byref(self._h)))
        self.dimension = dimension.upper()
        self.type = type_.lower()
        self.cfg = None
        self._nrec = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _abi.load().wv_destroy(h)
            self._h = C.c_void_p()

    # -- reference-named lifecycle -------------------------------------------------
    def init_forward_solver(self, NX, NY, NT, DH, DT, spatial_fd_order=8, NZ=1, ops_mode="sparse", free_surface=0,
                            damping_boundary=0, boundary_width=0, damping_coeff=8.0, npower=4.0, fc_cpml=5.0,
                            vmax_cpml=3500.0, nshots=1, path="fused", steps_per_graph=0, tile_y=0):
        cfg = _abi.Config()
        cfg.NX, cfg.NY, cfg.NZ, cfg.NT = NX, NY, NZ, NT
        cfg.DH, cfg.DT = DH, DT
        cfg.spatial_fd_order = spatial_fd_order
        cfg.ops_mode = {"sparse": _abi.WV_OPS_SPARSE, "stencil": _abi.WV_OPS_STENCIL}[ops_mode]
        cfg.free_surface = free_surface
        cfg.damping_boundary = damping_boundary
        cfg.boundary_width = boundary_width
        cfg.damping_coeff, cfg.npower, cfg.fc_cpml, cfg.vmax_cpml = damping_coeff, npower, fc_cpml, vmax_cpml
        cfg.nshots = nshots
        cfg.path = {"fused": _abi.WV_PATH_FUSED, "plain": _abi.WV_PATH_PLAIN, "phased": _abi.WV_PATH_PHASED,
                    "strip": _abi.WV_PATH_STRIP, "persistent": _abi.WV_PATH_PERSISTENT}[path]
        cfg.steps_per_graph = steps_per_graph
        cfg.tile_y = tile_y
        _abi.check(_abi.load().wv_init(self._h, C.byref(cfg)))
        self.cfg = cfg
        return self

    def prepare_for_modelling(self, m0, m1, m2):
        n = self.cfg.NX * self.cfg.NY * self.cfg.NZ
        arrs = []
        for m in (m0, m1, m2):
            arrs.append(None if m is None else np.ascontiguousarray(m, dtype=np.float32).reshape(-1)[:n].copy())
        for a in arrs:
            if a is not None and a.size != n:
                raise ValueError("model vector size mismatch")
        ptr = [(_fp(a) if a is not None else None) for a in arrs]
        _abi.check(_abi.load().wv_set_model(self._h, *ptr))
        return self

    def set_relaxation(self, freqs, tauP, tauS):
        """Viscoelastic: relaxation frequencies [Hz] (1..4 mechanisms) and tauP, tauS per cell (or
        scalars), before prepare_for_modelling (Modelparameter::Viscoelastic, Viscoelastic.cpp:160-260).
        Visco-EM: tauP = tau of the dielectric permittivity, tauS = tau of the electric conductivity
        [s] (Modelparameter::getTauDielectricPermittivity / getTauElectricConductivity); the model
        vectors are then the static relative permittivity, static conductivity and mu_r."""
        n = self.cfg.NX * self.cfg.NY * self.cfg.NZ
        f = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
        tp = np.ascontiguousarray(np.broadcast_to(np.asarray(tauP, np.float32), (n,)))
        ts = np.ascontiguousarray(np.broadcast_to(np.asarray(tauS, np.float32), (n,)))
        _abi.check(_abi.load().wv_set_relaxation(self._h, len(f), f.ctypes.data_as(C.POINTER(C.c_double)), _fp(tp),
                                                 _fp(ts)))
        return self

    def set_acquisition(self, shot, src_index, src_type, src_signal, rec_index, rec_type):
        si = np.ascontiguousarray(src_index, dtype=np.int64)
        st = np.ascontiguousarray(src_type, dtype=np.int32)
        sg = np.ascontiguousarray(src_signal, dtype=np.float32).reshape(len(si), self.cfg.NT)
        ri = np.ascontiguousarray(rec_index, dtype=np.int64)
        rt = np.ascontiguousarray(rec_type, dtype=np.int32)
        _abi.check(_abi.load().wv_set_acquisition(
            self._h, shot, len(si), si.ctypes.data_as(C.POINTER(C.c_longlong)), st.ctypes.data_as(C.POINTER(C.c_int)),
            _fp(sg), len(ri), ri.ctypes.data_as(C.POINTER(C.c_longlong)), rt.ctypes.data_as(C.POINTER(C.c_int))))
        self._nrec[shot] = len(ri)
        return self

    def reset(self):
        """Wavefields::resetWavefields + ForwardSolver::resetCPML (+ clears seismograms)."""
        _abi.check(_abi.load().wv_reset(self._h))

    reset_wavefields = reset
    reset_cpml = reset

    def run(self, t_start, t_end=None):
        """run(t): one time step as the reference; run(t0, t1): steps [t0, t1)."""
        if t_end is None:
            t_end = t_start + 1
        _abi.check(_abi.load().wv_run(self._h, int(t_start), int(t_end)))

    TUNE_CHOICES = ("phased-8", "phased-16", "phased-32", "fused-8", "fused-16", "fused-24", "3d-z4", "3d-z8",
                    "3d-z16", "strip-8", "strip-16", "strip-32", "persistent-16")

    def tune(self, nsteps=100):
        """Time every kernel variant on this problem, keep the fastest (resets the wavefields)."""
        ch = C.c_int()
        _abi.check(_abi.load().wv_tune(self._h, int(nsteps), C.byref(ch)))
        self.kernel_choice = self.TUNE_CHOICES[ch.value] if ch.value >= 0 else "plain"
        return self.kernel_choice

    def synchronize(self):
        _abi.check(_abi.load().wv_synchronize(self._h))

    def check_finite(self):
        """Simulation.cpp:519: raise WaveError if any wavefield value or seismogram sample is Inf/NaN."""
        _abi.check(_abi.load().wv_check_finite(self._h))

    def check_finite_noalloc(self):
        """The same test through wv_check_finite_noalloc (no device allocation); returns
        (status, number of non-finite values found) without raising."""
        n = C.c_longlong(-1)
        rc = _abi.load().wv_check_finite_noalloc(self._h, C.byref(n))
        return rc, n.value

    def last_run_ms(self):
        ms = C.c_double()
        _abi.check(_abi.load().wv_last_run_ms(self._h, C.byref(ms)))
        return ms.value

    # -- adjoint (2D elastic) -------------------------------------------------------
    def set_adjoint_store(self, max_gib):
        """Device-memory budget for recording every step instead of checkpointing (< 0 auto, 0 never)."""
        _abi.check(_abi.load().wv_set_adjoint_store(self._h, float(max_gib)))

    def set_checkpointing(self, every=50, fp16=True):
        """Store checkpoints every `every` steps in the following forward runs (resets the wavefields)."""
        _abi.check(_abi.load().wv_set_checkpointing(self._h, int(every), 1 if fp16 else 0))
        return self

    def set_adjoint_source(self, shot, adj):
        a = np.ascontiguousarray(adj, dtype=np.float32).reshape(self._nrec.get(shot, 0), self.cfg.NT)
        _abi.check(_abi.load().wv_set_adjoint_source(self._h, shot, _fp(a)))

    def run_adjoint(self):
        _abi.check(_abi.load().wv_run_adjoint(self._h))

    def adjoint_plan(self):
        """3D elastic: (sub-segment steps, sub-segments per checkpoint segment, checkpoints on the host)."""
        k2, ns, oh = C.c_int(), C.c_int(), C.c_int()
        _abi.check(_abi.load().wv_adjoint_plan(self._h, C.byref(k2), C.byref(ns), C.byref(oh)))
        return k2.value, ns.value, bool(oh.value)

    GRADIENTS = ("vp", "vs", "rho", "inv_rho_x", "inv_rho_y", "M", "mu", "mu_xy")
    GRADIENTS3D = ("vp", "vs", "rho", "inv_rho_x", "inv_rho_y", "inv_rho_z", "M", "mu", "mu_xy", "mu_xz", "mu_yz")

    def gradient(self, param="vp"):
        names = self.GRADIENTS3D if self.dimension == "3D" else self.GRADIENTS
        k = names.index(param) if isinstance(param, str) else int(param)
        n = self.cfg.NX * self.cfg.NY * (self.cfg.NZ if self.dimension == "3D" else 1)
        out = np.zeros(n, dtype=np.float64)
        _abi.check(_abi.load().wv_get_gradient(self._h, k, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- results ------------------------------------------------------------------
    def seismograms(self, shot=0):
        out = np.zeros((self._nrec.get(shot, 0), self.cfg.NT), dtype=np.float32)
        if out.size:
            _abi.check(_abi.load().wv_get_seismograms(self._h, shot, _fp(out)))
        return out

    def wavefield(self, name_or_index, shot=0):
        idx = name_or_index
        if isinstance(idx, str):
            idx = self.FIELDS[(self.dimension, self.type)].index(idx)
        out = np.zeros(self.cfg.NX * self.cfg.NY * self.cfg.NZ, dtype=np.float32)
        _abi.check(_abi.load().wv_get_wavefield(self._h, shot, idx, _fp(out)))
        return out

    def set_wavefield(self, name_or_index, values, shot=0):
        idx = name_or_index
        if isinstance(idx, str):
            idx = self.FIELDS[(self.dimension, self.type)].index(idx)
        v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        _abi.check(_abi.load().wv_set_wavefield(self._h, shot, idx, _fp(v)))


def check_parameters(dimension, type_, settings, m0, m1, m2, fc_max):
    """CheckParameter::checkNumericalArtefactsAndInstabilities (CheckParameter.hpp:166-246).

    `settings` holds the keys of ``init_forward_solver`` (NX, NY, NZ, DH, DT,
    spatial_fd_order).  Raises WaveError when the CFL criterion (with the
    reference's h = 1 for every order) or, for elastic, vp/vs >= sqrt(2) fails;
    returns (dispersion_ok, message) for the dispersion criterion, which the
    reference only reports.
    """
    L = _abi.load()
    cfg = _abi.Config()
    cfg.NX, cfg.NY, cfg.NZ = settings["NX"], settings["NY"], settings.get("NZ", 1)
    cfg.DH, cfg.DT = settings["DH"], settings["DT"]
    cfg.spatial_fd_order = settings.get("spatial_fd_order", 8)
    n = cfg.NX * cfg.NY * (cfg.NZ if dimension.upper() == "3D" else 1)
    arrs = [None if m is None else np.ascontiguousarray(m, dtype=np.float32).reshape(-1)[:n].copy()
            for m in (m0, m1, m2)]
    ptr = [(_fp(a) if a is not None else None) for a in arrs]
    ok = C.c_int(1)
    _abi.check(L.wv_check_parameters(dimension.upper().encode(), type_.encode(), C.byref(cfg), *ptr,
                                     float(fc_max), C.byref(ok)))
    return bool(ok.value), ("" if ok.value else L.wv_last_error().decode())


def create(dimension: str, type_: str) -> ForwardSolver:
    """ForwardSolver::Factory::Create(dimension, type)."""
    return ForwardSolver(dimension, type_)
