/*
 * wave_abi.h -- C ABI of the MI355X forward-solver time loop (libwavehip.so).
 *
 * This is the drop-in boundary for WAVE-Simulation's ForwardSolver /
 * ForwardSolverEM hot path.  Each entry point replaces one method of the
 * reference interface (paths relative to the WAVE-Simulation repository):
 *
 *   wv_create          ForwardSolver::Factory<ValueType>::Create(dimension, type)
 *                      src/ForwardSolver/ForwardSolverFactory.cpp:4-66
 *   wv_init            ForwardSolver::initForwardSolver(config, derivatives, wavefield, model, coords, ctx, DT)
 *                      src/ForwardSolver/ForwardSolver.hpp:52 (+ Derivatives init FDTD2D.cpp:49-59,
 *                      Wavefields init Wavefields2Delastic.cpp:30-38, CPML/ABS init ForwardSolver.cpp:34-54)
 *   wv_set_model       Modelparameter::prepareForModelling + ForwardSolver::prepareForModelling(model, DT)
 *                      src/Modelparameter/Elastic.cpp:28-47, ForwardSolver.hpp:42, ForwardSolver2Dtmem.cpp:77-84
 *   wv_set_acquisition Sources::init / Receivers::init for one shot
 *                      src/Acquisition/Sources.cpp:16-46, Receivers.cpp:14-33
 *   wv_reset           Wavefields::resetWavefields + ForwardSolver::resetCPML
 *                      src/Wavefields/Wavefields2Delastic.cpp:104-113, ForwardSolver.hpp:40
 *   wv_run             ForwardSolver::run(receiver, sources, model, wavefield, derivatives, t)
 *                      for t in [tStart, tEnd)   src/ForwardSolver/ForwardSolver.hpp:38, Simulation.cpp:444-476
 *   wv_get_seismograms receivers.getSeismogramHandler() data, src/Acquisition/Seismogram.cpp:82-156
 *   wv_get_wavefield   Wavefields::getRef*() (read back one component)
 *   wv_last_error      the message of the exception the reference would have thrown
 *
 * Conventions: status 0 = ok, negative = error (WV_E*); the message is then
 * available from wv_last_error() (thread-local).  Host pointers only; the
 * handle owns all device memory and one HIP stream (not thread-safe; separate
 * handles may be used concurrently).  Grid index i = x + z*NX + y*NX*NZ
 * (x fastest, y = depth slowest), as Coordinates.cpp:687.
 */
#ifndef WAVE_ABI_H
#define WAVE_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define WV_OK 0
#define WV_EINVAL (-1)   /* invalid argument or unsupported combination */
#define WV_ESTATE (-2)   /* call out of order (e.g. run before init)   */
#define WV_EDEVICE (-3)  /* HIP runtime error                           */
#define WV_ENOMEM (-4)   /* device allocation failed                    */
#define WV_ENONFINITE (-5) /* Inf or NaN in the wavefields or seismograms (wv_check_finite) */

/* operator edge mode */
#define WV_OPS_SPARSE 0  /* sparse matrices with edge order reduction (Derivatives.cpp:129-1234) */
#define WV_OPS_STENCIL 1 /* LAMA StencilMatrix: points outside the grid dropped (FDTD2D.cpp:156-185) */

/* execution path */
#define WV_PATH_FUSED 0  /* hand-written kernels, all shots per launch; one launch per time step
                            where a single-pass kernel exists (2D elastic), else WV_PATH_PHASED */
#define WV_PATH_PLAIN 1  /* reference operator sequence: CSR SpMV + vector passes, shot by shot */
#define WV_PATH_PHASED 2 /* hand-written kernels, velocity-type and stress-type update as two launches */
#define WV_PATH_STRIP 3  /* 2D elastic and TM-EM: two launches of row-strip kernels (each wavefront
                            owns tile_y/4 whole rows of 64 columns; y derivatives from registers) */
#define WV_PATH_PERSISTENT 4 /* 2D TM-EM without ABS: one cooperative launch per run call runs every
                                time step of every shot, grid barriers between the H and E updates
                                (no hipGraph; tile_y 8, 16, 32) */

typedef struct wv_solver wv_solver;

typedef struct {
    int NX, NY, NZ;        /* grid (NZ = 1 in 2D)                                  */
    int NT;                /* number of time steps (seismogram length)             */
    double DH, DT;         /* spacing [m], time step [s]                          */
    int spatial_fd_order;  /* 2,4,6,8,10,12                                       */
    int ops_mode;          /* WV_OPS_SPARSE (reference default) or WV_OPS_STENCIL  */
    int free_surface;      /* FreeSurface key (0 or 1)                            */
    int damping_boundary;  /* DampingBoundary key: 0 none, 1 ABS, 2 CPML          */
    int boundary_width;    /* BoundaryWidth                                        */
    double damping_coeff;  /* DampingCoeff (ABS)                                  */
    double npower;         /* NPower (CPML)                                       */
    double fc_cpml;        /* CenterFrequencyCPML                                  */
    double vmax_cpml;      /* VMaxCPML                                             */
    int nshots;            /* shots simulated together (>= 1)                      */
    int path;              /* WV_PATH_FUSED or WV_PATH_PLAIN                      */
    int steps_per_graph;   /* time steps per captured hipGraph segment (0 = none; rounded up to even) */
    int tile_y;            /* rows per tile (0 = default 16): single-pass elastic kernel 8, 16, 24,
                              negative selects direct global->LDS loads for the tile;
                              two-launch kernels 8, 16, 32; row-strip kernels 8, 16, 32;
                              3D elastic: rows of the x-z tile 4, 8, 16 (default 8) */
} wv_config;

/* dimension "2D"/"3D", type (case-insensitive): 2D "acoustic", "elastic", "viscoelastic", "sh",
 * "viscosh", "tmem", "viscotmem", "emem" / "viscoemem" (TE mode: fields hz, ex, ey; EX / EY / HZ
 * sources and receivers, types 2 / 3 / 4); 3D "acoustic", "elastic", "viscoelastic", "emem" /
 * "viscoemem" (fields hx, hy, hz, ex, ey, ez; EZ / EX / EY / HZ sources and receivers, types
 * 1..4), as ForwardSolverFactory.cpp:4-66 creates them.  The combinations the reference factory
 * answers with an empty pointer (3D "tmem", "sh", "viscosh", "viscotmem") return WV_EINVAL
 * ("not provided"), unknown dimensions and types WV_EINVAL. */
int wv_create(const char *dimension, const char *type, wv_solver **out);
int wv_destroy(wv_solver *s);

int wv_init(wv_solver *s, const wv_config *cfg);

/* model vectors [NX*NY*NZ] in grid index order:
 *   acoustic: m0 = velocityP, m1 ignored (may be NULL), m2 = density
 *   elastic : m0 = velocityP, m1 = velocityS, m2 = density
 *   viscoelastic: as elastic, after wv_set_relaxation
 *   sh      : m0 ignored (may be NULL), m1 = velocityS, m2 = density
 *   tmem, emem: m0 = epsilon_r, m1 = sigma [S/m], m2 = mu_r
 *   viscotmem, viscoemem: m0 = static epsilon_r, m1 = static sigma [S/m], m2 = mu_r (what
 *             Modelparameter::getDielectricPermittivity / getElectricConductivity hold, over
 *             eps0), after wv_set_relaxation                                           */
int wv_set_model(wv_solver *s, const float *m0, const float *m1, const float *m2);

/* Relaxation mechanisms of the viscoelastic solvers ("viscoelastic" 2D / 3D, "viscosh"; tauP
 * unused and may be NULL for visco-SH, ViscoSH.cpp:450-575) and of the visco-EM solvers
 * ("viscotmem", "viscoemem": tauP = tau of the dielectric permittivity (relative), tauS = tau of
 * the electric conductivity [s], as Modelparameter::getTauDielectricPermittivity /
 * getTauElectricConductivity hold them; ForwardSolver2Dviscotmem.cpp:77-100,
 * ForwardSolverEM.cpp:70-157 derive Ca, Cb, Cc, Cd from them): nmech
 * (1..4) relaxation frequencies [Hz] and tauP, tauS per cell [NX*NY] (grid order), as
 * Modelparameter::Viscoelastic holds them (numRelaxationMechanisms, relaxationFrequency..4,
 * the .tauP / .tauS model files; Viscoelastic.cpp:160-260, 585-596) and
 * FD2Dviscoelastic / FD3Dviscoelastic::initForwardSolver derive their constants from them
 * (ForwardSolver2Dviscoelastic.cpp:40-60, ForwardSolver3Dviscoelastic.cpp:50-66).  Call after
 * wv_init and before wv_set_model: the
 * moduli velocityP / velocityS given to wv_set_model are taken at the reference frequency
 * CenterFrequencyCPML (wv_config.fc_cpml) and scaled to the relaxed ones
 * (Viscoelastic.cpp:650-700).  WV_EINVAL for other solvers or bad arguments. */
int wv_set_relaxation(wv_solver *s, int nmech, const double *freqs, const float *tauP, const float *tauS);

/* Acquisition of one shot.  type: 1 = P (EZ for tmem), 2 = VX, 3 = VY, 4 = VZ.
 * src_signal is [nsrc][NT] (already scaled, as the reference's source seismogram).
 * Receivers keep the order given; seismograms come back in that order.      */
int wv_set_acquisition(wv_solver *s, int shot, int nsrc, const long long *src_index, const int *src_type,
                       const float *src_signal, int nrec, const long long *rec_index, const int *rec_type);

/* Synthetic source signal of wavelet shape `shape` into out[0, NT), evaluated in float
 * as the reference's float build: Sources::generateSyntheticSignal
 * (src/Acquisition/Sources.cpp:157-220) with the shapes of src/Acquisition/SourceSignal/:
 * 1 Ricker, 2 SinW, 3 SinThree, 4 FGaussian, 5 Spike, 6 IntgSinThree, 7 Ricker_GprMax,
 * 8 Berlage, 9 Sin; fc [Hz], amp, tshift [s] as in the source file's columns.  Host
 * code only (no GPU, no handle).  WV_EINVAL for an unknown shape, NT/DT/fc <= 0, or a
 * shape with a window (2, 3, 5, 6) shifted beyond the signal, with the reference's messages. */
int wv_source_signal(int shape, int NT, double DT, double fc, double amp, double tshift, float *out);

int wv_reset(wv_solver *s);

/* advance all shots over time steps [t_start, t_end); asynchronous on the handle's stream */
int wv_run(wv_solver *s, int t_start, int t_end);
int wv_synchronize(wv_solver *s);

/* seismograms of one shot, [nrec][NT] in receiver order (synchronizes) */
int wv_get_seismograms(wv_solver *s, int shot, float *out);

/* one wavefield component of one shot, [NX*NY*NZ] (synchronizes).
 * field order: acoustic vx vy p ; elastic 2D vx vy sxx syy sxy ; tmem hx hy ez ;
 *              elastic 3D vx vy vz sxx syy szz syz sxz sxy                    */
int wv_get_wavefield(wv_solver *s, int shot, int field, float *out);
int wv_set_wavefield(wv_solver *s, int shot, int field, const float *in);

/* Autotuner: times `nsteps` steps of every kernel variant that can run this
 * configuration (the two-launch kernels with 8-, 16- and 32-row tiles; for 2D
 * elastic on WV_PATH_FUSED also the single-pass kernel with 8-, 16- and 24-row
 * tiles and the row-strip kernels with 8-, 16- and 32-row blocks) on the handle's
 * own model and acquisition, keeps the fastest and
 * resets the wavefields and seismograms.  *choice receives the index of the
 * variant kept (0..2 = two-launch with 8/16/32 rows, 3..5 = single pass with
 * 8/16/24 rows, 6..8 = 3D elastic with 4/8/16-row x-z tiles, 9..11 = 2D
 * elastic row-strip kernels with 8/16/32-row blocks, -1 = the handle has a
 * single variant: WV_PATH_PLAIN).
 * No counterpart in the reference. */
int wv_tune(wv_solver *s, int nsteps, int *choice);

/* elapsed GPU time [ms] of the last wv_run (hipEvents on the handle's stream) */
int wv_last_run_ms(wv_solver *s, double *ms);

/* Per-kernel timing.  With wv_set_kernel_timing(s, 1), wv_run launches every
 * kernel eagerly with start/stop hipEvents attached to the dispatch itself
 * (hipExtLaunchKernelGGL) on the handle's stream and accumulates the
 * elapsed time per kernel kind; wv_kernel_stats then reports, for kind
 * 0 = velocity-type phase, 1 = stress-type phase, 2 = end-of-segment source/receiver flush
 * (and for the 2D elastic adjoint pass, wv_run_adjoint: 3 = recomputed velocity phase,
 * 4 = recomputed stress phase, 5 = adjoint stress step, 6 = adjoint velocity step):
 * the kernel name, its average duration [ms] per launch, the number of
 * launches, and the algorithmic HBM bytes one launch must move (every field
 * and coefficient of the batch read once, updated fields written once, CPML
 * memory variables read and written on frame cells).  Returns WV_EINVAL for an
 * unknown kind.  No counterpart in the reference (instrumentation only). */
int wv_set_kernel_timing(wv_solver *s, int on);
int wv_kernel_stats(wv_solver *s, int kind, const char **name, double *avg_ms, long long *launches,
                    double *bytes_per_launch);

const char *wv_last_error(void);

/* CheckParameter::checkNumericalArtefactsAndInstabilities
 * (src/CheckParameter/CheckParameter.hpp:166-246) for a regular grid; host only, needs no
 * GPU and no handle.  Model vectors as wv_set_model (m1 = velocityS for elastic / sh);
 * fc_max = the largest centre frequency of the shot's sources.
 *  - elastic: WV_EINVAL when vp/vs < sqrt(2) at a cell (the reference asserts);
 *  - checkStabilityCriterion (:60-112): WV_EINVAL when DT > DH / (h sqrt(D) vmax), with
 *    vmax = max vp (vs for sh, 1/sqrt(eps mu) for EM) and h = 1 for every order (the
 *    reference's h = 7/6, 149/120, ... are integer divisions);
 *  - checkNumericalDispersion (:121-162): *dispersion_ok = 0 when DH > vmin / (2 fc_max N),
 *    N = 12, 8, 7, 6, 5, 4 for orders 2..12, vmin = the smallest positive vp (acoustic),
 *    vs (elastic, sh) or EM velocity; a warning, as in the reference: the status stays
 *    WV_OK and wv_last_error() holds the message.                                        */
int wv_check_parameters(const char *dimension, const char *type, const wv_config *cfg, const float *m0,
                        const float *m1, const float *m2, double fc_max, int *dispersion_ok);

/* The per-shot test of Simulation.cpp:519 (WavefieldsSeismic::isFinite,
 * SeismogramHandler::isFinite): WV_ENONFINITE, with the reference's message, when any
 * wavefield value of any shot or any seismogram sample is Inf or NaN (the reference
 * samples one wavefield cell and the last sample of each trace; this tests all of
 * them).  Synchronizes the handle's stream. */
int wv_check_finite(wv_solver *s);

/* The same test as wv_check_finite, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): it calls
 * no hipMalloc / hipFree and no allocating library routine.  No counterpart in the
 * reference.  Its scratch memory -- one 64-bit count of non-finite values per state buffer
 * of the handle, at most WV_NOALLOC_MAX_BUFFERS buffers (128 B) -- is a static device
 * array of the library, one copy per GPU, held until the process ends; it is cleared by
 * a fill kernel at every call.  Calls that overlap on one GPU (from different threads or
 * handles) share that array and race.  *nonfinite (may be NULL) receives the number of
 * Inf/NaN values found in the wavefields of every shot (padding included) and in the
 * seismograms.  Returns WV_ENONFINITE as wv_check_finite does, WV_EINVAL when the handle
 * has more state buffers than the scratch holds.  Synchronizes the handle's stream. */
#define WV_NOALLOC_MAX_BUFFERS 16
int wv_check_finite_noalloc(wv_solver *s, long long *nonfinite);

/* ---- adjoint-state gradient (2D elastic; CPML or no absorbing frame, no free
 * surface).  No counterpart in the reference: WAVE-Simulation has no adjoint
 * or gradient code (only inert flags, Modelparameter.hpp:94-95, 289); these
 * entry points serve the WAVE-Inversion-style caller that needs dJ/dm.      */

/* Checkpoint every `every` steps during the following wv_run calls (0 = off);
 * fp16 != 0 stores the wavefields in FP16 with one scale per field and shot,
 * otherwise FP32; CPML memory variables are kept in FP32.  Switches a handle
 * on the single-pass kernel to the two-launch kernels, uses graph segments of
 * `every` steps and resets the wavefields. */
int wv_set_checkpointing(wv_solver *s, int every, int fp16);

/* Device-memory budget [GiB] for recording, during the checkpointed forward
 * run, the five derivative sums of every step (FP16, one scale per row of 64
 * cells, field and shot: about 10 B per cell, shot and step) instead of
 * checkpointing and recomputing; used when the handle runs the row-strip
 * kernels and the record fits.  < 0 (default): 3/4 of the free device memory
 * at wv_set_checkpointing; 0: always checkpoint and recompute.  Takes effect
 * at the next wv_set_checkpointing.  No counterpart in the reference. */
int wv_set_adjoint_store(wv_solver *s, double max_gib);

/* adjoint source of one shot: dJ/d(seismogram), [nrec][NT] in receiver order */
int wv_set_adjoint_source(wv_solver *s, int shot, const float *adj);

/* Time-reversed adjoint pass over [0, NT) after a checkpointed wv_run(0, NT):
 * recomputes each checkpoint segment (or reads the per-step record) and
 * correlates with the adjoint fields;
 * the gradient is summed over all shots of the handle in FP64.  Asynchronous
 * on the handle's stream; wv_last_run_ms reports its duration. */
int wv_run_adjoint(wv_solver *s);

/* gradient [NY][NX] (x fastest): param 0 vp, 1 vs, 2 rho (chain rule through
 * the averaging rules of wv_set_model), 3..7 w.r.t. the coefficients
 * 1/rho_x, 1/rho_y, M, mu, mu_xy.  3D elastic: [NY][NZ][NX] in grid index order,
 * params 3..10 = 1/rho_x, 1/rho_y, 1/rho_z, M, mu, mu_xy, mu_xz, mu_yz. */
int wv_get_gradient(wv_solver *s, int param, double *out);

/* 3D elastic: the memory plan wv_set_checkpointing chose from the budget of
 * wv_set_adjoint_store (default 3/4 of the free device memory).  The forward run keeps
 * checkpoints every `every` steps, on the device or (*on_host = 1) in pinned host memory;
 * the adjoint pass recomputes each segment in sub-segments of *k2 steps (*nsub of them,
 * FP32 state saved at the start of each when *nsub > 1), recording the derivative sums
 * of every step of one sub-segment.  No counterpart in the reference. */
int wv_adjoint_plan(wv_solver *s, int *k2, int *nsub, int *on_host);

#ifdef __cplusplus
}
#endif
#endif
