/*
 * phix_amd.h -- C ABI of the MI355X (gfx950) phase-field hot path.
 *
 * Every entry point takes plain pointers and sizes. Device pointers address
 * ghost-padded stored arrays in the PhiX layout:
 *
 *   storedDims[a] = n[a] + 2*ghost          (for all three axes, also unused ones)
 *   index(i,j,k)  = (i+g) + sx*((j+g) + sy*(k+g)),  sx = storedDims[0], sy = storedDims[1]
 *
 * (include/field/ScalarField.h:14-26, 90-96 of the reference). Physical cells are
 * i in [0,n). For a 2D mesh the physical plane is k = 0 (stored plane g).
 *
 * Conventions
 *   - return 0 (PHX_OK) on success, a negative phx_status on failure;
 *     phx_last_error() returns the message of the last failure on this thread.
 *   - the caller owns every buffer; nothing here allocates device memory except
 *     where a function says so (graph handles, the tuner cache).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Work is stream-ordered: functions return once work is enqueued, unless
 *     they say otherwise.
 *
 * Reference interfaces replaced (file:line in WangHanwei1996/PhiX):
 *   BC kernels           src/boundary/Boundary.cu:95-195, 204-351
 *   DSL term kernels     src/equation/Equation.cu:40-124,386-404, include/equation/TermPW.inl:18-90
 *   Solver kernels       src/solver/Solver.cu:30-71, src/solver/VectorSolver.cu:27-59
 *   Euler multi-step     src/solver/Solver.cu:317-336 (CH: applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu:84-106)
 *   RK4 single equation  src/solver/Solver.cu:192-259 with RHS  M*lap(pw(c,f') - kappa*lap(c), bcs)
 *   RK4 vector equation  src/solver/VectorSolver.cu:149-228
 *   mean / free energy   src/field/ScalarField.cu:260-280, develop/Spinodal Decomposition/computeFreeEnergy.py:76-98
 */
#ifndef PHIX_AMD_H
#define PHIX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_ABI_VERSION 1

typedef enum {
    PHX_OK              = 0,
    PHX_ERR_INVALID     = -1,   /* bad argument, shape or layout            */
    PHX_ERR_UNSUPPORTED = -2,   /* valid request outside what a kernel does  */
    PHX_ERR_HIP         = -3,   /* a HIP runtime call failed                 */
    PHX_ERR_NOMEM       = -4
} phx_status;

typedef enum { PHX_F64 = 0, PHX_F32 = 1 } phx_dtype;

/* Boundary kinds (PeriodicBC / NoFluxBC / FixedBC of the reference). FIXED is a
 * constant ghost fill (include/boundary/FixedBC.h:11-16, src/boundary/Boundary.cu:166-195). */
typedef enum {
    PHX_BC_NONE     = 0,
    PHX_BC_PERIODIC = 1,
    PHX_BC_NOFLUX   = 2,
    PHX_BC_FIXED    = 3
} phx_bc_kind;

typedef enum { PHX_SIDE_LOW = 0, PHX_SIDE_HIGH = 1, PHX_SIDE_BOTH = 2 } phx_side;

/* Grid, layout and per-face boundary table. bc[a][0] is the low side of axis a,
 * bc[a][1] the high side. Axes >= dim must have n = 1 and are never stenciled. */
typedef struct phx_grid {
    int    dim;
    int    n[3];
    int    ghost;
    double d[3];
    int    bc[3][2];
    double bc_value[3][2];
} phx_grid;

/* Cahn-Hilliard with the double-well derivative of the reference apps
 *   f'(c) = 2*rho*(c-ca)*(c-cb)*(2c-ca-cb)      (Cahn-Hillard_double-well.cu:87-90)
 *   mu    = f'(c) - kappa*lap(c),   dc/dt = M*lap(mu)
 * f = c^2(1-c)^2 is rho=1, ca=0, cb=1. */
typedef struct phx_ch_params {
    double rho, ca, cb;
    double kappa, M;
    double dt;
} phx_ch_params;

/* Multi-order-parameter Allen-Cahn (grain growth):
 *   d(eta_i)/dt = L*kappa*lap(eta_i) + L*(eta_i - eta_i^3) - 2*L*gamma*eta_i*sum_{j!=i} eta_j^2 */
typedef struct phx_ac_params {
    double L, kappa, gamma;
    double dt;
} phx_ac_params;

typedef void* phx_stream;

/* ---------------------------------------------------------------- misc ---- */
int         phx_abi_version(void);
const char* phx_last_error(void);
/* Number of stored cells of one field in this layout (0 if the grid is invalid). */
size_t      phx_stored_size(const phx_grid* grid);
/* Blocks the host until all work on `stream` has finished. */
int         phx_stream_sync(phx_stream stream);

/* ------------------------------------------------------ boundaries ---- */
/* One BoundaryCondition object: fills the ghost layers of `axis` on `side`
 * over the full stored extent of the other two axes, exactly as
 * PeriodicBC/NoFluxBC/FixedBC::applyOnGPU (src/boundary/Boundary.cu:224-239,
 * 276-296, 331-351). `kind` is a phx_bc_kind, `side` a phx_side (PERIODIC
 * ignores it). Only grid->n, grid->ghost are read. */
int phx_bc_apply(const phx_grid* grid, int dtype, void* data,
                 int axis, int kind, int side, double value, phx_stream stream);

/* Applies grid->bc of the active axes in the order X, Y, Z, in place -- the
 * order of a {bc_x, bc_y, bc_z} list handed to the reference Solver
 * (Solver.cu:149-151). A periodic axis is one PeriodicBC object, any other
 * axis one object per face; faces set to PHX_BC_NONE are left alone. When
 * every active face has a condition and ghost <= n on every active axis, all
 * faces are filled in one launch with the same values (each ghost cell takes
 * what the face fills leave, corners included); otherwise one launch per
 * face. */
int phx_bc_fill(const phx_grid* grid, int dtype, void* data, phx_stream stream);

/* ------------------------------------------ generic DSL term kernels ---- */
/* FP64 per-term kernels behind the generic (unfused) path of the C++ API.
 * They accumulate into physical cells of rhs only. */
/* rhs += coeff*lap(src)            (Equation.cu:40-80) */
int phx_lap_accumulate(const phx_grid* grid, double* rhs, const double* src,
                       double coeff, phx_stream stream);
/* rhs += coeff*d(src)/d(x_axis)    (Equation.cu:89-124) */
int phx_grad_accumulate(const phx_grid* grid, double* rhs, const double* src,
                        int axis, double coeff, phx_stream stream);
/* rhs += coeff*s1*s2               (Equation.cu:386-404) */
int phx_mul_accumulate(const phx_grid* grid, double* rhs, const double* s1,
                       const double* s2, double coeff, phx_stream stream);
/* rhs += coeff*f'(src) with the double-well f' of phx_ch_params (TermPW.inl:18-37
 * instantiated with the functor of Cahn-Hillard_double-well.cu:87-90). */
int phx_pw_doublewell_accumulate(const phx_grid* grid, double* rhs, const double* src,
                                 double rho, double ca, double cb, double coeff,
                                 phx_stream stream);
/* Whole-array kernels over n stored values (Solver.cu:30-71). */
int phx_fill_f64(double* dst, double value, size_t n, phx_stream stream);
int phx_copy_f64(double* dst, const double* src, size_t n, phx_stream stream);
int phx_axpy_f64(double* dst, const double* src, double coeff, size_t n, phx_stream stream);
int phx_rk4_tmp_f64(double* tmp, const double* phi, const double* k, double coeff,
                    size_t n, phx_stream stream);
int phx_rk4_update_f64(double* phi, const double* k1, const double* k2, const double* k3,
                       const double* k4, double dt6, size_t n, phx_stream stream);
/* Precision conversion of n stored values (round to nearest / exact widening).
 * Used by the C++ Solver's FP32-bulk option; no reference counterpart. */
int phx_convert_f64_to_f32(float* dst, const double* src, size_t n, phx_stream stream);
int phx_convert_f32_to_f64(double* dst, const float* src, size_t n, phx_stream stream);

/* ------------------------------------------- fused Cahn-Hilliard ---- */
/* nsteps of the reference CH multi-step Euler solver
 *   {c, bcs, mu = f'(c) - kappa*lap(c), STEADY}, {mu, bcs, dc/dt = M*lap(mu), TRANSIENT}
 * (Solver.cu:317-336). On return c holds c_{n+nsteps} on physical cells and the
 * BC fill of c_{n+nsteps-1} on its ghosts; mu holds mu(c_{n+nsteps-1}) with BC-filled
 * ghosts -- the state the reference leaves. Every face of grid->bc must be set. */
int phx_ch_euler_steps(const phx_grid* grid, int dtype, void* c, void* mu,
                       const phx_ch_params* p, int nsteps, phx_stream stream);
/* No reference counterpart. Every Euler step but the last runs as one launch
 * (the exact-form stage kernel: mu on a ring around each tile, then
 * c + dt * M lap(mu); in 3D the z-register kernel streaming z-chunks of
 * 128x16 columns) instead of two, with the same results bit for
 * bit; mu serves as the second state buffer until the last step. on = 0
 * keeps two launches per step (process-wide). Returns PHX_OK. */
int phx_ch_euler_set_fused(int on);
/* No reference counterpart. On 2D grids with n >= 8 on both axes (any BC
 * kinds), the one-launch steps run in blocks of four per launch (the 2D whole-step
 * kernel's Euler form: c on the tile and an 8-cell ring in LDS, mu and the
 * state updated four times there; the same bits). tile: 0 (default) 32x16
 * cells per 512-thread workgroup, 1: 32x32/512, 2: 64x16/256, 3: 32x16/256;
 * -1 turns the blocks off (process-wide). PHX_ERR_INVALID for other values. */
int phx_ch_euler_set_quad(int tile);

/* Constants of the Fe-B glass-formation model as the reference's GFA_2D
 * settings name them ("constants" block, GFA.cu:104-115), plus the step dt.
 * Units: SI (J, mol, m, s, K). */
typedef struct phx_gfa_params {
    double dt;
    double M_eta, eps_sq, beta_sq;
    double w_phi, w_eta, w_ex;
    double T, T_g, alpha;
    double V_m_L, V_m_S, D_Am;
} phx_gfa_params;

/* Workspace for phx_gfa_euler_steps, in bytes (three stored FP64 fields). */
size_t phx_gfa_workspace_bytes(const phx_grid* grid);
/* nsteps Euler steps of the GFA_2D system on FP64 fields c, phi, eta (stored
 * layout of `grid`, 2D, periodic or zero-flux faces): the reference's six
 * multi-steps per step (GFA.cu:143-201 with Solver.cu:317-338: mu; D; D; c;
 * phi; eta) as one fused launch -- mu and D are evaluated in LDS on each tile
 * and a one-cell ring through the BC index maps, then c, phi (reading the new
 * c) and eta (reading the new phi) are updated at each cell. The temperature-
 * only quantities (M_phi, D_L, D_S, f_S, df_Am, the SGTE energies) are formed
 * once on the host as GFA.cu:117-121 does. On return the physical cells hold
 * the state after nsteps and the x/y ghost cells the BC fill of the state
 * before the last step, as the reference leaves them; z ghost planes are not
 * touched. work: phx_gfa_workspace_bytes(grid). Returns PHX_OK or
 * PHX_ERR_INVALID (dim != 2, a FIXED face, ghost > n, a null pointer, a
 * non-finite constant). */
int phx_gfa_euler_steps(const phx_grid* grid, double* c, double* phi, double* eta, void* work,
                        const phx_gfa_params* p, int nsteps, phx_stream stream);

/* Workspace for phx_ch_rk4_steps, in bytes (three stored fields). */
size_t phx_ch_rk4_workspace_bytes(const phx_grid* grid, int dtype);
/* nsteps of classic RK4 on dc/dt = M*lap(f'(c) - kappa*lap(c)) with the BCs of
 * grid->bc applied to every stage state and to mu (Solver.cu:192-259 with the
 * composite lap(expr,bcs) of Equation.cu:537-651). Each stage is one fused
 * kernel: mu and lap(mu) live in LDS, the BCs are index maps of the halo load.
 * On return c holds c_{n+nsteps} on physical cells and the BC fill of
 * c_{n+nsteps-1} on its ghosts, as the reference leaves it. */
int phx_ch_rk4_steps(const phx_grid* grid, int dtype, void* c, void* work,
                     const phx_ch_params* p, int nsteps, phx_stream stream);

/* nbatch independent fields of one 2D grid (periodic, no-flux or fixed
 * faces, n >= 8 on both axes), stored back to back in c (each
 * phx_ch_rk4_workspace_bytes(grid)/3 bytes apart, i.e. its full stored
 * size), advanced by nsteps RK4 steps together: the whole-step kernel of the
 * grid's 2D launch runs all fields per launch (the batch on the launch's z
 * axis). Each field ends exactly as phx_ch_rk4_steps would leave it, bit for
 * bit. work: nbatch * phx_ch_rk4_workspace_bytes(grid) bytes. No reference
 * counterpart (the reference advances one field per solver). */
int phx_ch_rk4_steps_batch(const phx_grid* grid, int dtype, int nbatch, void* c, void* work,
                           const phx_ch_params* p, int nsteps, phx_stream stream);
/* Times the whole-step tiles on `steps` steps of the caller's batch (it is
 * advanced: pass a scratch copy) and keeps the fastest for
 * phx_ch_rk4_steps_batch with this shape, dtype and nbatch (no reference
 * counterpart); every tile returns the same bits. */
int phx_ch_rk4_batch_autotune(const phx_grid* grid, int dtype, int nbatch, void* c, void* work,
                              const phx_ch_params* p, int steps, double* best_ms_per_step,
                              phx_stream stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). phx_ch_rk4_steps with the RK4 workspace held by the library:
 * a static device array of PHX_CH_RK4_NOALLOC_BYTES per GPU (64 MiB: e.g.
 * 2048^2 FP32, 1024^2 FP64, 128^3 FP64 with one ghost layer), taken from the
 * library's load until the process ends. Grids whose workspace
 * (phx_ch_rk4_workspace_bytes) exceeds it return PHX_ERR_UNSUPPORTED. Calls no
 * allocating routine; results equal phx_ch_rk4_steps's for the same launch.
 * Every call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (other streams, other host threads) race and must be serialised by the
 * caller. Launches made through it are never timed by phx_ch_timing. */
#define PHX_CH_RK4_NOALLOC_BYTES (64u << 20)
int phx_ch_rk4_steps_noalloc(const phx_grid* grid, int dtype, void* c,
                             const phx_ch_params* p, int nsteps, phx_stream stream);

/* Launch shape of the RK4 kernels, per (dim, n, dtype). Defaults: 3D per
 * stage class (phx_ch_rk4_set_stage_launch); 2D with n >= 8 the whole-step
 * tiles 7 (FP32 on grids of 256^2 and more) or 5, else 0. In 3D the planes are split into cdiv(n_z, zc) balanced chunks of
 * floor or ceil(n_z / chunks) planes, one workgroup column each. The
 * autotuner times every tile shape x z-chunk in {4,8,16,32,64} plus the
 * chunk counts that fill 1, 2 or 3 whole waves of resident workgroups (from
 * the occupancy calculator) on `steps` steps of the given state (advancing
 * it: pass a scratch copy), keeps the fastest and reports it. Host-side
 * settings; no device allocation. */
int phx_ch_rk4_set_zchunk(const phx_grid* grid, int dtype, int zc);
/* Tile shape of the 3D stage kernel: 0: 64x16, 1: 64x8, 2: 128x8, 3: 32x16,
 * 4: 64x4 (x by y columns per 256-thread workgroup). Tiles 5: 64x8, 6: 32x16,
 * 7: 32x8, 8: 64x4 select the two-stage kernel (stages 1+2 and 3+4 each in one
 * launch: 7 field passes per step instead of 16/17; same results bit for bit;
 * needs n >= 4 on every axis, else PHX_ERR_UNSUPPORTED). Tiles 9: 64x16/256,
 * 10: 128x8/256, 11: 128x16/512, 12: 256x8/512, 13: 256x16/1024, 14: 256x4/256,
 * 15: 256x8/1024, 16: 128x16/1024, 17: 128x8/512 (columns / workgroup
 * threads) select the one-stage kernel that keeps each
 * column's z-neighbours in registers and only one plane of the stage state
 * and of mu in LDS, so wide tiles (long contiguous rows in HBM) keep their
 * occupancy; same results bit for bit; any n >= 2. In 2D: 0 is the
 * one-stage kernel (64x16), 1: 64x16, 2: 64x32, 3: 128x16, 4: 32x32 (256
 * threads), 5: 32x32, 6: 64x32 (512 threads), 7: 128x32 (1024 threads; FP64
 * runs 6 instead), 8: 32x64 (512 threads) the whole-step kernel (a full RK4
 * step per launch, 2 field passes; same results bit for bit; needs n >= 8 on
 * both axes). */
int phx_ch_rk4_set_tile(const phx_grid* grid, int dtype, int tile);
int phx_ch_rk4_get_launch(const phx_grid* grid, int dtype, int* zc, int* tile);
/* XCD-aware workgroup order of the 3D stage kernel (on by default): each of
 * the 8 XCDs gets a contiguous range of tiles so halo re-reads hit its L2. */
int phx_ch_set_xcd_remap(int on);
/* Arithmetic form of FP64 RK4 (no reference counterpart; process-wide, host
 * setting). 0, the default: explicit FMAs with M folded into lap(mu)'s
 * coefficients and, in 3D, the step finished from the stage states
 * u2, u3, u4 (13 field passes per step instead of 17; FP64 12, with stage 2
 * also writing b = u2 + 2 u3 - c_n for stage 4); within a few ulp of
 * the reference's expression per step. 1: the reference's operation order
 * without contraction and its k1+2k2+2k3+k4 accumulator
 * (src/solver/Solver.cu:61-71, src/equation/Equation.cu:164-180), equal to
 * the FP64 oracle bit for bit. Every kernel of one form returns the same bits.
 * phx_ch_get_fp64_exact returns the current form. */
int phx_ch_set_fp64_exact(int on);
int phx_ch_get_fp64_exact(void);
/* The autotuner also refines each stage class of a winning 3D one-stage
 * launch (see phx_ch_rk4_set_stage_launch). */
int phx_ch_rk4_autotune(const phx_grid* grid, int dtype, void* c, void* work,
                        const phx_ch_params* p, int steps, int* best_zc,
                        double* best_ms_per_step, phx_stream stream);
/* Launch of one stage class of the 3D one-stage schedule (no reference
 * counterpart): cls 0 = stage 1, 1 = stages 2 and 3, 2 = stage 4; tile a
 * one-stage tile (0-4 or 9-17), z-chunk zc. Every one-stage kernel returns
 * the same bits, so classes may mix launches. phx_ch_rk4_set_tile/_set_zchunk
 * set all three classes; phx_ch_rk4_get_launch reports class 1, whose tile
 * also chooses the schedule (a two-stage tile there runs every stage
 * two-stage). Defaults in 3D FP32: classes 0 and 1 tile 11 at 24 z-chunks,
 * class 2 tile 12 at 16; FP64: class 0 tile 11 at 8 z-chunks, classes 1 and
 * 2 tile 17 at 4 and 8 z-chunks. */
int phx_ch_rk4_set_stage_launch(const phx_grid* grid, int dtype, int cls, int tile, int zc);
int phx_ch_rk4_get_stage_launch(const phx_grid* grid, int dtype, int cls, int* zc, int* tile);

/* Benchmark timing (no reference counterpart): while on, each launch of the
 * fused RK4 kernels is bracketed by HIP events on its stream. phx_ch_timing(1)
 * also clears the record. phx_ch_timing_read sums the durations of the
 * recorded launches of one kind (CH: 0 one-stage kernel, 1 two-stage (1,2),
 * 2 two-stage (3,4), 3 2D whole step; Allen-Cahn: 4 one-stage kernel,
 * 5 two-stage (1,2), 6 two-stage (3,4); ensemble: 7 one-stage kernel, 8
 * two-stage (1,2), 9 two-stage (3,4); GFA: 10 the fused step kernel),
 * waiting for them. Launches captured
 * into a graph (phx_ensemble_graph_steps) must not run while it is on. One
 * thread at a time. */
int phx_ch_timing(int on);
int phx_ch_timing_read(int kind, int* launches, double* total_ms);

/* ------------------------------------------ FP64 diagnostics ---- */
/* Workspace in bytes for phx_mass_energy. */
size_t phx_reduce_workspace_bytes(const phx_grid* grid);
/* out_dev (device, 2 doubles) <- { sum of c over physical cells,
 *   total free energy sum[rho(c-ca)^2(cb-c)^2 + kappa/2 |grad c|^2] * dV }
 * with central-difference gradients (c[+1]-c[-1])/(2d) whose off-grid
 * neighbours follow grid->bc (computeFreeEnergy.py:76-98 uses a periodic roll).
 * Accumulation is FP64 with Neumaier-compensated sums and a fixed reduction tree,
 * whatever the storage dtype, so the result is deterministic. */
int phx_mass_energy(const phx_grid* grid, int dtype, const void* c,
                    const phx_ch_params* p, void* work, double* out_dev,
                    phx_stream stream);
/* phx_mass_energy of nbatch same-shape stored fields laid back to back in c
 * (no reference counterpart; the parareal judge's energies of all slice
 * boundaries in one launch pair): out_dev holds 2*nbatch doubles, field b's
 * pair at 2b, each the bits of phx_mass_energy on that field; work is
 * nbatch * phx_reduce_workspace_bytes(grid) bytes. */
int phx_mass_energy_batch(const phx_grid* grid, int dtype, int nbatch, const void* c,
                          const phx_ch_params* p, void* work, double* out_dev,
                          phx_stream stream);

/* ------------------------------------- fused multi-field Allen-Cahn ---- */
size_t phx_ac_rk4_workspace_bytes(const phx_grid* grid, int dtype, int nfields);
/* nsteps of RK4 for nfields coupled order parameters (VectorSolver.cu:186-228
 * with a VectorEquation whose component i is
 *   L*kappa*lap(eta_i) + pw(eta_i, e - e^3, L) + sum_{j!=i} pw(eta_i, eta_j, a*b*b, -2*L*gamma)).
 * eta is an array of nfields host-held device pointers (one stored field each),
 * 1 <= nfields <= 8, dim 1 or 2. work is phx_ac_rk4_workspace_bytes (3*nfields fields). */
int phx_ac_rk4_steps(const phx_grid* grid, int dtype, int nfields, void* const* eta,
                     void* work, const phx_ac_params* p, int nsteps, phx_stream stream);
/* No reference counterpart. 2D grids run two RK4 stages per launch by default
 * (stages 1+2, then 3+4: 7 field passes per step instead of 16), with results
 * equal bit for bit to the one-stage kernel's. on = 0 selects the one-stage
 * kernel (process-wide; 1D grids always use it). Returns PHX_OK. */
int phx_ac_set_two_stage(int on);

/* ------------------------------------ batched uneven-size ensemble ---- */
/* No reference counterpart (PhiX runs one problem per process). An ensemble
 * is nruns independent 2D periodic Cahn-Hilliard runs (f = c^2(1-c)^2,
 * per-run kappa and mobility M), side lengths sides[r], advanced together by
 * one persistent-kernel launch per RK4 stage. FP32 state: the runs' physical
 * n_r x n_r fields back to back in one arena of phx_ensemble_cells() floats,
 * run r at element phx_ensemble_offset(r), x fastest, no ghost cells.
 * work: 3 * phx_ensemble_cells() floats. create allocates the device tables;
 * the other calls allocate nothing (graph instantiation aside). */
typedef struct phx_ensemble phx_ensemble;
int       phx_ensemble_create(int nruns, const int* sides, const double* kappa,
                              const double* mobility /* NULL: all 1 */, phx_ensemble** out);
int       phx_ensemble_destroy(phx_ensemble* ens);
long long phx_ensemble_cells(const phx_ensemble* ens);
long long phx_ensemble_offset(const phx_ensemble* ens, int run);
/* nsteps RK4 steps (2 or 4 launches each, see phx_ensemble_set_two_stage) on
 * `stream`. */
int phx_ensemble_rk4_steps(phx_ensemble* ens, float* c, float* work, double dt, int nsteps,
                           phx_stream stream);
/* Captures steps_per_graph RK4 steps into a hipGraph (once per (c, work,
 * steps, tiling, dt)) and replays it nreplays times on `stream`. */
int phx_ensemble_graph_steps(phx_ensemble* ens, float* c, float* work, double dt,
                             int steps_per_graph, int nreplays, phx_stream stream);
/* out_dev[r] (device, nruns doubles) <- FP64 compensated sum of run r. */
int phx_ensemble_mass(const phx_ensemble* ens, const float* c, double* out_dev, phx_stream stream);
/* Tile shape index (0..4: 64x16, 32x16, 64x8, 128x8, 32x32) and persistent
 * grid size; grid <= 0 picks 8 workgroups per CU. Every shape returns the
 * same bits. */
int phx_ensemble_set_tiling(phx_ensemble* ens, int shape, int grid);
int phx_ensemble_get_tiling(const phx_ensemble* ens, int* shape, int* grid, int* tx, int* ty);
/* on = 1 (the default): two RK4 stages per launch (stages 1+2, then 3+4; 7
 * field passes per step instead of 16); on = 0: one stage per launch. Both
 * kernels are persistent (grid from phx_ensemble_set_tiling) and return the
 * same bits. get returns 0/1 (-1 for NULL). */
int phx_ensemble_set_two_stage(phx_ensemble* ens, int on);
int phx_ensemble_get_two_stage(const phx_ensemble* ens);
/* Times both kernels at every tile shape and {2,4,8,16} workgroups per CU,
 * on `steps` steps of the given state
 * (advancing it: pass a scratch copy), keeps the fastest and returns its ms
 * per step. Blocks the host. */
int phx_ensemble_autotune(phx_ensemble* ens, float* c, float* work, double dt, int steps,
                          double* best_ms_per_step, phx_stream stream);

/* ------------------------------------------------ learned surrogate ---- */
/* No reference counterpart. A 6-layer 3x3 circular-padding conv net mapping
 * c(t) -> c(t + 100 dt) on B periodic H x W FP32 fields (H, W multiples of 16):
 *   relu(conv 1->32), 4 x relu(conv 32->32), conv 32->1 = d,
 *   c' = c + d - mean(d)  (exact per-sample mean projection).
 * BF16 weights and activations, FP32 accumulation; the 32->32 layers run on
 * v_mfma_f32_16x16x32_bf16. params: phx_sur_param_count() FP32 master weights,
 * [out][in][kh][kw] per layer followed by its bias, layers in order.
 * ws: phx_sur_workspace_bytes(B, H, W) bytes; forward keeps the activations
 * there for the following backward. */
int    phx_sur_param_count(void);
size_t phx_sur_workspace_bytes(int B, int H, int W);
/* flags: PHX_SUR_REUSE_WEIGHTS skips re-packing the BF16 weights (valid when
 * params are unchanged since the last call on this ws). PHX_SUR_FUSED
 * evaluates the six layers of each 16x16 output tile in one launch with the
 * activations in LDS (inference: the activations are not kept, so no
 * phx_sur_backward may follow). Every layer runs on the MFMA there, the
 * first and last with their FP32 operands split into BF16 high and low
 * parts (~16 significant bits): the output follows the layer-by-layer
 * kernels to within the BF16 rounding of the activations, not bit for bit. */
#define PHX_SUR_REUSE_WEIGHTS 1
#define PHX_SUR_FUSED 2
int    phx_sur_forward(const float* params, void* ws, const float* c_in, float* c_out,
                       int B, int H, int W, int flags, phx_stream stream);
/* jumps >= 1 inference jumps c_in -> c_out (no reference counterpart): the
 * bits of `jumps` phx_sur_forward calls with PHX_SUR_FUSED, with each
 * jump's mean projection formed by the next jump's launch as it loads its
 * input (one launch per jump, plus one projection at the end). flags:
 * PHX_SUR_REUSE_WEIGHTS as above. c_out may be c_in. */
int    phx_sur_rollout(const float* params, void* ws, const float* c_in, float* c_out,
                       int B, int H, int W, int jumps, int flags, phx_stream stream);
/* As phx_sur_rollout, keeping every `every`-th state (no reference
 * counterpart; the surrogate sweeps of a parallel-in-time solver):
 * states[k] (B*H*W floats each, k = 0 .. jumps/every - 1) <- the state after
 * every*(k+1) jumps, projected to the mean and, when corr is not NULL, plus
 * corr[k] (FP32 add); the jumps after a kept state continue from it. The
 * same bits as jumps/every phx_sur_rollout calls of `every` jumps, each
 * followed by the addition of corr[k]. Needs every >= 1 dividing jumps;
 * states must not overlap c_in or corr. */
int    phx_sur_rollout_states(const float* params, void* ws, const float* c_in, float* states,
                              const float* corr, int B, int H, int W, int jumps, int every,
                              int flags, phx_stream stream);
/* Loss = mean((c_out - c_target)^2); grad (param layout) <- dLoss/dparams;
 * loss_dev (device double, may be NULL) <- sum of squared errors (divide by
 * B*H*W for the loss). Needs the preceding phx_sur_forward on the same ws. */
int    phx_sur_backward(const float* params, void* ws, const float* c_in, const float* c_out,
                        const float* c_target, float* grad, double* loss_dev,
                        int B, int H, int W, phx_stream stream);
/* Adam (Kingma & Ba 2015) on the master weights, bias-corrected at `step` >= 1. */
int    phx_sur_adam(float* params, const float* grad, float* m, float* v, int step,
                    double lr, double beta1, double beta2, double eps, phx_stream stream);
/* One 32->32 layer, as the network runs it (for parity tests).
 * x, y, dy, dx, dz, mask: BF16 NHWC [B][H][W][32]; w: FP32 [32][32][3][3];
 * ws: phx_conv3x3_ws_bytes() bytes of scratch.
 *   fwd  : y  = relu(conv(x, w) + bias)
 *   dgrad: dx = conv^T(dy, w) * (mask > 0)
 *   wgrad: dw = d/dw sum(dz * conv(x, w)),  db = sum over pixels of dz  */
size_t phx_conv3x3_ws_bytes(void);
int    phx_conv3x3_bf16_fwd(const void* x, const float* w, const float* bias, void* y, void* ws,
                            int B, int H, int W, phx_stream stream);
int    phx_conv3x3_bf16_dgrad(const void* dy, const float* w, const void* mask, void* dx, void* ws,
                              int B, int H, int W, phx_stream stream);
int    phx_conv3x3_bf16_wgrad(const void* x, const void* dz, float* dw, float* db, void* ws,
                              int B, int H, int W, phx_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* PHIX_AMD_H */
