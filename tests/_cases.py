"""Helpers shared by the parity tests: run a wavesim Case through the oracle."""
import os

import numpy as np

from oracle import pyoracle as po

PHYS = {"acoustic": 0, "elastic": 1, "tmem": 2, "sh": 3, "viscoelastic": 4, "viscosh": 5, "emem": 6, "viscotmem": 7,
        "viscoemem": 8}


def oracle_threads():
    """OpenMP threads of the oracle: the host's cores, at most 16 (the share a GPU box gives one GPU)."""
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1))


def oracle_params(case, precision_threads=None, ops_mode="sparse"):
    s = case.settings
    return po.make_params(dim=2 if case.dimension == "2D" else 3, physics=PHYS[case.type], NX=s["NX"], NY=s["NY"],
                          NZ=s.get("NZ", 1), NT=s["NT"], order=s["spatial_fd_order"],
                          stencil=1 if ops_mode == "stencil" else 0, free_surface=s.get("free_surface", 0),
                          damping=s.get("damping_boundary", 0), width=s.get("boundary_width", 0), DH=s["DH"],
                          DT=s["DT"], damping_coeff=s.get("damping_coeff", 8.0), npower=s.get("npower", 4.0),
                          fc_cpml=s.get("fc_cpml", 5.0), vmax_cpml=s.get("vmax_cpml", 3500.0),
                          nthreads=precision_threads or oracle_threads())


def oracle_shot(case, shot, precision="f32", ops_mode="sparse", want_fields=False):
    p = oracle_params(case, ops_mode=ops_mode)
    m1 = case.m1 if case.m1 is not None else np.zeros_like(case.m0)
    dt = np.float32 if precision == "f32" else np.float64
    return po.run(p, case.m0.astype(dt), m1.astype(dt), case.m2.astype(dt), shot.src_index, shot.src_type,
                  shot.src_signal.astype(dt), shot.rec_index, shot.rec_type, precision, want_fields=want_fields,
                  visco=case.visco)


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))
