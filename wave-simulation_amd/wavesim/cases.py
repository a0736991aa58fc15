"""Deterministic synthetic workloads (numpy PCG64) for the benchmark and tests.

Each function returns a Case: the solver settings (reference configuration
keys), model vectors in grid order (x fastest, then depth) and the per-shot
acquisition.  Sizes default to the benchmark configurations; tests pass
smaller grids / fewer steps of the same recipe.

  acoustic_c1   2D acoustic, order 4, 200x200, dx 1 m, dt 2e-4 s, 2000 steps, 20-cell CPML,
                homogeneous vp 2000 rho 2000, 50 Hz Ricker P source at the centre,
                100 P receivers on a row 10 cells below the top
  elastic_c2    2D elastic P-SV, order 8, 1000x1000, dx 1 m, dt 1e-4 s, 4000 steps, 30-cell CPML,
                10 layers vp 1500-4000, vs = vp/1.8, rho 1800-2600, x(1 + 0.03 g) with g a unit-std
                Gaussian random field (20 m correlation), 30 Hz Ricker VY force, 500 VX + 500 VY receivers
  tmem_c3       2D TM-mode EM, order 8, 800x400, dx 0.02 m, dt 3e-11 s, 3000 steps, 20-cell CPML,
                eps_r ~ U(1,25) and sigma ~ U(0,0.01) per 0.5 m block, 64 shots of 200 MHz Ricker EZ
                sources with 20..300 EZ receivers each
  elastic_c4    2D elastic, order 8, 600x600, 3000 steps, layered + 80 m-radius vp/vs anomaly (+10 %),
                16 shots, VX/VY receivers every 2 m
  elastic_c5    3D elastic, order 8, 320^3, dx 2 m, dt 2e-4 s, 2500 steps, 24-cell CPML, 6 depth layers
                vp 1800-3800, vs = vp/1.8, rho 1800-2500, x(1 + 0.03 g) (10-cell correlation), 8 shots of
                15 Hz Ricker VY forces at depth 26, 40x40 receiver grids (spacing 4) of vx, vy, vz
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .acquisition import ricker


@dataclass
class Shot:
    src_index: np.ndarray
    src_type: np.ndarray
    src_signal: np.ndarray
    rec_index: np.ndarray
    rec_type: np.ndarray


@dataclass
class Case:
    name: str
    dimension: str
    type: str
    settings: dict
    m0: np.ndarray
    m1: np.ndarray | None
    m2: np.ndarray
    shots: list = field(default_factory=list)
    visco: dict | None = None  # viscoelastic: freqs (per mechanism), tauP, tauS (per cell)

    @property
    def cells(self):
        s = self.settings
        return s["NX"] * s["NY"] * s.get("NZ", 1)


def gaussian_field(shape, sigma, seed, clip=4.0):
    """Unit-std Gaussian random field: white noise smoothed by a periodic Gaussian kernel of std `sigma` cells."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    grids = np.meshgrid(*k, indexing="ij")
    r2 = sum(g * g for g in grids)
    spec = np.exp(-2.0 * (np.pi * sigma) ** 2 * r2)
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * spec))
    g = (g - g.mean()) / g.std()
    return np.clip(g, -clip, clip)


def layered(NY, NX, nlayers, lo, hi):
    rows = np.arange(NY) * nlayers // NY
    vals = lo + (hi - lo) * rows / (nlayers - 1)
    return np.repeat(vals[:, None], NX, axis=1)


def acoustic_c1(N=200, NT=2000, DT=2e-4):
    NX = NY = N
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=DT, spatial_fd_order=4, damping_boundary=2, boundary_width=20,
             npower=4.0, fc_cpml=50.0, vmax_cpml=2000.0)
    n = NX * NY
    vp = np.full(n, 2000.0, np.float32)
    rho = np.full(n, 2000.0, np.float32)
    c = Case("acoustic_c1", "2D", "acoustic", s, vp, None, rho)
    cx = NX // 2
    rx = np.arange(NX // 4, NX // 4 + min(100, NX // 2))
    c.shots.append(Shot(np.array([cx + (NY // 2) * NX]), np.array([1]), ricker(NT, DT, 50.0, 1.0)[None, :],
                        rx + 10 * NX, np.ones(len(rx), np.int32)))
    return c


def elastic_model(NX, NY, seed_v, seed_r, sigma=20.0, pert=0.03):
    vp = layered(NY, NX, 10, 1500.0, 4000.0)
    rho = layered(NY, NX, 10, 1800.0, 2600.0)
    g = gaussian_field((NY, NX), sigma, seed_v)
    gr = gaussian_field((NY, NX), sigma, seed_r)
    vp = vp * (1.0 + pert * g)
    vs = vp / 1.8
    rho = rho * (1.0 + pert * gr)
    return (vp.astype(np.float32).ravel(), vs.astype(np.float32).ravel(), rho.astype(np.float32).ravel())


def elastic_c2(N=1000, NT=4000, DT=1e-4, W=30):
    NX = NY = N
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=DT, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0)
    vp, vs, rho = elastic_model(NX, NY, 2, 102)
    c = Case("elastic_c2", "2D", "elastic", s, vp, vs, rho)
    x = np.arange(NX // 4, NX // 4 + NX // 2)
    yr = max(W + 10, NY // 25)
    rec = np.concatenate([x + yr * NX, x + yr * NX])
    rty = np.concatenate([np.full(len(x), 2), np.full(len(x), 3)]).astype(np.int32)
    c.shots.append(Shot(np.array([NX // 2 + (NY // 2) * NX]), np.array([3]), ricker(NT, DT, 30.0, 1.0)[None, :],
                        rec, rty))
    return c


def sh_2d(NX=180, NY=120, NT=600, DT=1e-4, nshots=3, order=8, W=16, damping=2, free_surface=0, seed=7):
    """2D SH (Love waves, ForwardSolver2Dsh.cpp): the C2-style layered model with Gaussian
    perturbations, VZ Ricker sources at seeded positions, VZ receivers on a row and a column
    (uneven counts per shot)."""
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=30.0, vmax_cpml=2500.0, free_surface=free_surface,
             damping_coeff=8.0)
    vp, vs, rho = elastic_model(NX, NY, seed, seed + 100, sigma=8.0)
    c = Case("sh_2d", "2D", "sh", s, vp, vs, rho)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = ricker(NT, DT, 60.0, 1.0)
    lo = W + 4 if damping else 4
    for k in range(nshots):
        sx, sy = int(rng.integers(lo, NX - lo)), int(rng.integers(lo if not free_surface else 1, NY - lo))
        nr = int(rng.integers(10, 40))
        rx = np.sort(rng.choice(np.arange(lo, NX - lo), nr, replace=False))
        ry = 0 if free_surface else lo
        rec = np.concatenate([rx + ry * NX, (NX // 2) + np.arange(lo, NY - lo, 7) * NX])
        c.shots.append(Shot(np.array([sx + sy * NX]), np.array([4]), w[None, :], rec, np.full(len(rec), 4, np.int32)))
    return c


def tmem_c3(NX=800, NY=400, NT=3000, DT=3e-11, nshots=64, seed=3, block=25, W=20):
    s = dict(NX=NX, NY=NY, NT=NT, DH=0.02, DT=DT, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=2e8, vmax_cpml=3e8)
    rng = np.random.Generator(np.random.PCG64(seed))
    nby, nbx = -(-NY // block), -(-NX // block)
    eps_b = rng.uniform(1.0, 25.0, (nby, nbx))
    sig_b = rng.uniform(0.0, 0.01, (nby, nbx))
    eps = np.kron(eps_b, np.ones((block, block)))[:NY, :NX]
    sig = np.kron(sig_b, np.ones((block, block)))[:NY, :NX]
    mur = np.ones((NY, NX))
    c = Case("tmem_c3", "2D", "tmem", s, eps.astype(np.float32).ravel(), sig.astype(np.float32).ravel(),
             mur.astype(np.float32).ravel())
    lo, hi = W + 5, NX - W - 5
    y = W + 2
    w = ricker(NT, DT, 2e8, 1.0)
    for _ in range(nshots):
        sx = int(rng.integers(lo, hi))
        nrec = int(rng.integers(20, 301))
        nrec = min(nrec, hi - lo)
        rx = np.sort(rng.choice(np.arange(lo, hi), nrec, replace=False))
        c.shots.append(Shot(np.array([sx + y * NX]), np.array([1]), w[None, :], rx + y * NX,
                            np.ones(nrec, np.int32)))
    return c


def emem_2d(NX=170, NY=97, NT=500, DT=3e-11, nshots=3, order=8, W=12, damping=2, seed=13, block=10):
    """2D TE-EM (ForwardSolver2Demem.cpp): blocky epsilon_r (1-16), sigma (0-0.01 S/m) and mu_r
    (1-3) so every average differs from the node values; per shot EX, EY and HZ sources (one or
    two) at seeded positions and a receiver row of all three components, uneven counts."""
    s = dict(NX=NX, NY=NY, NT=NT, DH=0.02, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=2e8, vmax_cpml=3e8, damping_coeff=8.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    nby, nbx = -(-NY // block), -(-NX // block)
    blocks = [rng.uniform(lo, hi, (nby, nbx)) for lo, hi in ((1.0, 16.0), (0.0, 0.01), (1.0, 3.0))]
    eps, sig, mur = (np.kron(b, np.ones((block, block)))[:NY, :NX].astype(np.float32).ravel() for b in blocks)
    c = Case("emem_2d", "2D", "emem", s, eps, sig, mur)
    lo = W + 4 if damping else 4
    w = ricker(NT, DT, 2e8, 1.0)
    kinds = ([3], [2, 4], [4])
    for k in range(nshots):
        ty = kinds[k % 3]
        pos = [int(rng.integers(lo, NX - lo)) + int(rng.integers(lo, NY - lo)) * NX for _ in ty]
        sigs = np.stack([w * (1.0 if j == 0 else 0.6) for j in range(len(ty))])
        nr = int(rng.integers(10, 40))
        rx = np.sort(rng.choice(np.arange(lo, NX - lo), nr, replace=False))
        ry = lo + 2 * k
        rec = np.concatenate([rx + ry * NX, (NX // 2) + np.arange(lo, NY - lo, 9) * NX])
        rty = (np.arange(len(rec)) % 3 + 2).astype(np.int32)
        c.shots.append(Shot(np.array(pos), np.array(ty), sigs, rec, rty))
    return c


def _em_visco(c, freqs, tau_eps, tau_sig, seed):
    """Relaxation of a 2D EM case: tau of the permittivity (relative) and of the conductivity [s]
    per 10-cell block around tau_eps, tau_sig; L mechanisms at `freqs` [Hz]."""
    s = c.settings
    rng = np.random.Generator(np.random.PCG64(seed + 500))
    nby, nbx = -(-s["NY"] // 10), -(-s["NX"] // 10)
    te = np.kron(rng.uniform(0.5, 1.5, (nby, nbx)) * tau_eps, np.ones((10, 10)))[:s["NY"], :s["NX"]]
    ts = np.kron(rng.uniform(0.0, 2.0, (nby, nbx)) * tau_sig, np.ones((10, 10)))[:s["NY"], :s["NX"]]
    c.visco = dict(freqs=list(freqs), tauP=te.astype(np.float32).ravel(), tauS=ts.astype(np.float32).ravel())
    return c


def viscotmem_2d(NX=170, NY=97, NT=500, DT=3e-11, nshots=3, order=8, W=12, damping=2, seed=17, freqs=(3e8,),
                 tau_eps=0.3, tau_sig=2e-10):
    """2D visco TM-EM (ForwardSolver2Dviscotmem.cpp): the emem_2d model (blocky static epsilon_r,
    sigma, mu_r), tau of the permittivity and of the conductivity per block, L mechanisms at
    `freqs`; per shot one EZ source and a row and a column of EZ receivers (uneven counts)."""
    c = emem_2d(NX=NX, NY=NY, NT=NT, DT=DT, nshots=nshots, order=order, W=W, damping=damping, seed=seed)
    c.name, c.type = "viscotmem_2d", "viscotmem"
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    lo = W + 4 if damping else 4
    w = ricker(NT, DT, 2e8, 1.0)
    c.shots = []
    for k in range(nshots):
        src = int(rng.integers(lo, NX - lo)) + int(rng.integers(lo, NY - lo)) * NX
        nr = int(rng.integers(10, 40))
        rx = np.sort(rng.choice(np.arange(lo, NX - lo), nr, replace=False))
        rec = np.concatenate([rx + (lo + 2 * k) * NX, (NX // 2) + np.arange(lo, NY - lo, 9) * NX])
        c.shots.append(Shot(np.array([src]), np.array([1]), w[None, :], rec, np.ones(len(rec), np.int32)))
    return _em_visco(c, freqs, tau_eps, tau_sig, seed)


def viscoemem_2d(NX=170, NY=97, NT=500, DT=3e-11, nshots=3, order=8, W=12, damping=2, seed=19, freqs=(3e8,),
                 tau_eps=0.3, tau_sig=2e-10):
    """2D visco TE-EM (ForwardSolver2Dviscoemem.cpp): the emem_2d model and acquisition with tau of
    the permittivity and of the conductivity per block, L mechanisms at `freqs`."""
    c = emem_2d(NX=NX, NY=NY, NT=NT, DT=DT, nshots=nshots, order=order, W=W, damping=damping, seed=seed)
    c.name, c.type = "viscoemem_2d", "viscoemem"
    return _em_visco(c, freqs, tau_eps, tau_sig, seed)


def emem_3d(NX=44, NY=36, NZ=30, NT=300, DT=3e-11, nshots=2, order=8, W=8, damping=2, seed=23, block=6):
    """3D EM (ForwardSolver3Demem.cpp): blocky epsilon_r (1-16), sigma (0-0.01 S/m) and mu_r (1-3)
    in 6-cell blocks so every average differs from the node values; per shot EZ, EX, EY and HZ
    sources (one or two) at seeded positions and a receiver line of all four components."""
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=0.02, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=2e8, vmax_cpml=3e8, damping_coeff=8.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = [-(-n // block) for n in (NY, NZ, NX)]
    ones = np.ones((block, block, block))
    blocks = [rng.uniform(lo, hi, nb) for lo, hi in ((1.0, 16.0), (0.0, 0.01), (1.0, 3.0))]
    eps, sig, mur = (np.kron(b, ones)[:NY, :NZ, :NX].astype(np.float32).ravel() for b in blocks)
    c = Case("emem_3d", "3D", "emem", s, eps, sig, mur)
    lo = W + 3 if damping else 3
    w = ricker(NT, DT, 2e8, 1.0)
    kinds = ([1], [2, 4], [3], [4, 1])
    for k in range(nshots):
        ty = kinds[k % 4]
        pos = [int(rng.integers(lo, NX - lo)) + int(rng.integers(lo, NZ - lo)) * NX +
               int(rng.integers(lo, NY - lo)) * NX * NZ for _ in ty]
        sigs = np.stack([w * (1.0 if j == 0 else 0.6) for j in range(len(ty))])
        rx = np.arange(lo, NX - lo, 2)
        cells = rx + (NZ // 2 + k) * NX + (lo + 2) * NX * NZ
        rec = np.concatenate([cells, cells[::2], cells[1::2], cells[::3]])
        rty = np.concatenate([np.full(len(cells), 1), np.full(len(cells[::2]), 2), np.full(len(cells[1::2]), 3),
                              np.full(len(cells[::3]), 4)]).astype(np.int32)
        c.shots.append(Shot(np.array(pos), np.array(ty), sigs, rec, rty))
    return c


def viscoemem_3d(freqs=(3e8,), tau_eps=0.3, tau_sig=2e-10, seed=29, **kw):
    """3D visco-EM (ForwardSolver3Dviscoemem.cpp): the emem_3d model and acquisition with tau of the
    permittivity and of the conductivity per 6-cell block, L mechanisms at `freqs`."""
    c = emem_3d(seed=seed, **kw)
    c.name, c.type = "viscoemem_3d", "viscoemem"
    s = c.settings
    rng = np.random.Generator(np.random.PCG64(seed + 500))
    nb = [-(-n // 6) for n in (s["NY"], s["NZ"], s["NX"])]
    ones = np.ones((6, 6, 6))
    te = np.kron(rng.uniform(0.5, 1.5, nb) * tau_eps, ones)[:s["NY"], :s["NZ"], :s["NX"]]
    ts = np.kron(rng.uniform(0.0, 2.0, nb) * tau_sig, ones)[:s["NY"], :s["NZ"], :s["NX"]]
    c.visco = dict(freqs=list(freqs), tauP=te.astype(np.float32).ravel(), tauS=ts.astype(np.float32).ravel())
    return c


def elastic_c4(N=600, NT=3000, DT=1e-4, nshots=16, seed=4, W=30, anomaly=True):
    NX = NY = N
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=DT, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0)
    vp = layered(NY, NX, 10, 1500.0, 4000.0)
    rho = layered(NY, NX, 10, 1800.0, 2600.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    cy, cx = rng.uniform(N / 4, 3 * N / 4, 2)
    if anomaly:
        yy, xx = np.mgrid[0:NY, 0:NX]
        disc = (xx - cx) ** 2 + (yy - cy) ** 2 <= (80.0 * N / 600) ** 2
        vp = np.where(disc, vp * 1.10, vp)
    vs = vp / 1.8
    c = Case("elastic_c4", "2D", "elastic", s, vp.astype(np.float32).ravel(), vs.astype(np.float32).ravel(),
             rho.astype(np.float32).ravel())
    y = W + 10
    xr = np.arange(W + 10, NX - W - 10, 2)
    rec = np.concatenate([xr + y * NX, xr + y * NX])
    rty = np.concatenate([np.full(len(xr), 2), np.full(len(xr), 3)]).astype(np.int32)
    w = ricker(NT, DT, 30.0, 1.0)
    for _ in range(nshots):
        sx = int(rng.integers(W + 10, NX - W - 10))
        c.shots.append(Shot(np.array([sx + y * NX]), np.array([3]), w[None, :], rec, rty))
    return c


def visco_2d(NX=170, NY=97, NT=500, DT=1e-4, nshots=3, order=8, W=14, damping=2, free_surface=0, freqs=(30.0,),
             tau=0.15, seed=11):
    """2D viscoelastic P-SV (ForwardSolver2Dviscoelastic.cpp): the C2-style layered model with
    Gaussian perturbations, tauP / tauS varying by layer around `tau`, VX / VY / P Ricker sources at
    seeded positions, VX, VY and P receivers on a row and a column (uneven counts per shot)."""
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0, free_surface=free_surface,
             damping_coeff=8.0)
    vp, vs, rho = elastic_model(NX, NY, seed, seed + 100, sigma=8.0)
    rows = np.repeat((np.arange(NY) * 4 // NY)[:, None], NX, axis=1).ravel()
    tp = (tau * (1.0 + 0.25 * rows)).astype(np.float32)
    ts = (tau * (1.2 - 0.15 * rows)).astype(np.float32)
    c = Case("visco_2d", "2D", "viscoelastic", s, vp, vs, rho,
             visco=dict(freqs=list(freqs), tauP=tp, tauS=ts))
    rng = np.random.Generator(np.random.PCG64(seed))
    w = ricker(NT, DT, 60.0, 1.0)
    lo = W + 4 if damping else 4
    for k in range(nshots):
        sx, sy = int(rng.integers(lo, NX - lo)), int(rng.integers(lo if not free_surface else 1, NY - lo))
        nr = int(rng.integers(10, 40))
        rx = np.sort(rng.choice(np.arange(lo, NX - lo), nr, replace=False))
        ry = 0 if free_surface else lo
        rec = np.concatenate([rx + ry * NX, (NX // 2) + np.arange(lo, NY - lo, 7) * NX])
        rty = np.array([1 + (j % 3) for j in range(len(rec))], np.int32)
        c.shots.append(Shot(np.array([sx + sy * NX]), np.array([1 + k % 3]), w[None, :], rec, rty))
    return c


def viscosh_2d(NX=180, NY=120, NT=600, DT=1e-4, nshots=3, order=8, W=16, damping=2, free_surface=0, freqs=(30.0,),
               tau=0.15, seed=7):
    """2D visco-SH (ForwardSolver2Dviscosh.cpp): the sh_2d acquisition and model with tauS varying
    by layer around `tau` and L relaxation mechanisms at `freqs`."""
    c = sh_2d(NX=NX, NY=NY, NT=NT, DT=DT, nshots=nshots, order=order, W=W, damping=damping,
              free_surface=free_surface, seed=seed)
    rows = np.repeat((np.arange(NY) * 4 // NY)[:, None], NX, axis=1).ravel()
    ts = (tau * (1.2 - 0.15 * rows)).astype(np.float32)
    c.name, c.type = "viscosh_2d", "viscosh"
    c.visco = dict(freqs=list(freqs), tauP=np.zeros_like(ts), tauS=ts)
    return c


def elastic_c5(N=320, NT=2500, DT=2e-4, nshots=8, seed=5, W=24, nrec_side=40, spacing=4, src_depth=26):
    NX = NY = NZ = N
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=2.0, DT=DT, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=15.0, vmax_cpml=4400.0)
    # model arrays in the reference's index order: x fastest, then z, then depth y
    rows = np.arange(NY) * 6 // NY
    vp_y = 1800.0 + 2000.0 * rows / 5
    rho_y = 1800.0 + 700.0 * rows / 5
    g = gaussian_field((NY, NZ, NX), 10.0, seed)
    gr = gaussian_field((NY, NZ, NX), 10.0, seed + 100)
    vp = vp_y[:, None, None] * (1.0 + 0.03 * g)
    rho = rho_y[:, None, None] * (1.0 + 0.03 * gr)
    vs = vp / 1.8
    c = Case("elastic_c5", "3D", "elastic", s, vp.astype(np.float32).ravel(), vs.astype(np.float32).ravel(),
             rho.astype(np.float32).ravel())
    rng = np.random.Generator(np.random.PCG64(seed))
    lo = W + max(2, min(16, (N - 2 * W) // 4))
    hi = N - lo
    w = ricker(NT, DT, 15.0, 1.0)
    y = min(src_depth, NY - W - 2)
    for _ in range(nshots):
        sx, sz = (int(v) for v in rng.integers(lo, hi, 2))
        off = (np.arange(nrec_side) - nrec_side // 2) * spacing
        rx = sx + off
        rz = sz + off
        rx = rx[(rx >= W) & (rx < NX - W)]
        rz = rz[(rz >= W) & (rz < NZ - W)]
        gx, gz = np.meshgrid(rx, rz)
        cells = (gx + gz * NX + (W + 2) * NX * NZ).ravel()
        rec = np.concatenate([cells, cells, cells])
        rty = np.repeat(np.array([2, 3, 4], np.int32), len(cells))
        c.shots.append(Shot(np.array([sx + sz * NX + y * NX * NZ]), np.array([3]), w[None, :], rec, rty))
    return c


def acoustic_3d(NX=72, NY=60, NZ=52, NT=260, DT=2e-4, nshots=2, order=8, W=10, damping=2, free_surface=0,
                seed=9):
    """3D acoustic (ForwardSolver3Dacoustic.cpp): layered vp/rho with Gaussian perturbations, P and
    VX/VY/VZ sources at seeded positions, receivers of every component on a line at depth."""
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=2.0, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=80.0, vmax_cpml=3000.0, free_surface=free_surface,
             damping_coeff=8.0)
    rows = np.arange(NY) * 5 // NY
    g = gaussian_field((NY, NZ, NX), 4.0, seed)
    vp = (1800.0 + 250.0 * rows)[:, None, None] * (1.0 + 0.03 * g)
    rho = (1900.0 + 100.0 * rows)[:, None, None] * np.ones((1, NZ, NX))
    c = Case("acoustic_3d", "3D", "acoustic", s, vp.astype(np.float32).ravel(), None, rho.astype(np.float32).ravel())
    rng = np.random.Generator(np.random.PCG64(seed))
    w = ricker(NT, DT, 80.0, 1.0)
    lo = W + 3 if damping else 3
    for k in range(nshots):
        sx, sy, sz = (int(rng.integers(lo, n - lo)) for n in (NX, NY, NZ))
        ty = 1 if k % 2 == 0 else 2 + k % 3
        rx = np.arange(lo, NX - lo, 3)
        yr = 0 if free_surface else lo
        cells = rx + (NZ // 2) * NX + yr * NX * NZ
        rec = np.concatenate([cells, cells[::2], cells[1::2], cells[::3]])
        rty = np.concatenate([np.full(len(cells), 1), np.full(len(cells[::2]), 2), np.full(len(cells[1::2]), 3),
                              np.full(len(cells[::3]), 4)]).astype(np.int32)
        c.shots.append(Shot(np.array([sx + sz * NX + sy * NX * NZ]), np.array([ty]), w[None, :], rec, rty))
    return c


def visco_3d(NX=44, NY=36, NZ=30, NT=260, DT=2e-4, nshots=2, order=8, W=8, damping=2, free_surface=0,
             freqs=(30.0,), tau=0.15, seed=13):
    """3D viscoelastic (ForwardSolver3Dviscoelastic.cpp): layered vp/vs/rho with Gaussian
    perturbations, tauP / tauS varying by layer around `tau`, L relaxation mechanisms at `freqs`,
    P / VX / VY / VZ Ricker sources at seeded positions, receivers of every component on a line."""
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=2.0, DT=DT, spatial_fd_order=order, damping_boundary=damping,
             boundary_width=W, npower=4.0, fc_cpml=40.0, vmax_cpml=3600.0, free_surface=free_surface,
             damping_coeff=8.0)
    rows = np.arange(NY) * 4 // NY
    g = gaussian_field((NY, NZ, NX), 4.0, seed)
    vp = (1800.0 + 400.0 * rows)[:, None, None] * (1.0 + 0.03 * g)
    rho = (1900.0 + 150.0 * rows)[:, None, None] * np.ones((1, NZ, NX))
    vs = vp / 1.8
    lay = np.repeat(rows, NZ * NX)
    tp = (tau * (1.0 + 0.25 * lay)).astype(np.float32)
    ts = (tau * (1.2 - 0.15 * lay)).astype(np.float32)
    c = Case("visco_3d", "3D", "viscoelastic", s, vp.astype(np.float32).ravel(), vs.astype(np.float32).ravel(),
             rho.astype(np.float32).ravel(), visco=dict(freqs=list(freqs), tauP=tp, tauS=ts))
    rng = np.random.Generator(np.random.PCG64(seed))
    w = ricker(NT, DT, 60.0, 1.0)
    lo = W + 3 if damping else 3
    for k in range(nshots):
        sx, sy, sz = (int(rng.integers(lo, n - lo)) for n in (NX, NY, NZ))
        if free_surface:
            sy = int(rng.integers(1, NY - lo))
        ty = 1 + k % 4
        rx = np.arange(lo, NX - lo, 3)
        yr = 0 if free_surface else lo
        cells = rx + (NZ // 2) * NX + yr * NX * NZ
        rec = np.concatenate([cells, cells[::2], cells[1::2], cells[::3]])
        rty = np.concatenate([np.full(len(cells), 1), np.full(len(cells[::2]), 2), np.full(len(cells[1::2]), 3),
                              np.full(len(cells[::3]), 4)]).astype(np.int32)
        c.shots.append(Shot(np.array([sx + sz * NX + sy * NX * NZ]), np.array([ty]), w[None, :], rec, rty))
    return c


def make_solver(case, nshots=None, path="fused", ops_mode="sparse", steps_per_graph=0, shots=None, tile_y=0):
    """Create and fully prepare a solver for `case` (the given shots, default all)."""
    from .solver import create

    shots = case.shots if shots is None else shots
    n = len(shots) if nshots is None else nshots
    s = dict(case.settings)
    sol = create(case.dimension, case.type)
    sol.init_forward_solver(nshots=n, path=path, ops_mode=ops_mode, steps_per_graph=steps_per_graph, tile_y=tile_y,
                            **s)
    if case.visco is not None:
        sol.set_relaxation(case.visco["freqs"], case.visco["tauP"], case.visco["tauS"])
    sol.prepare_for_modelling(case.m0, case.m1, case.m2)
    for k, sh in enumerate(shots[:n]):
        sol.set_acquisition(k, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type)
    sol.reset()
    return sol
