/* hes.h -- C-ABI of libhes.so, the gfx950 finite-volume time step of the
 * Barton-2009 (two-phase Savenkov-Alekseev) hyperelastic solver.
 *
 * Conventions
 *   - Every function returns an int status: HES_OK (0) or a negative HES_ERR_*;
 *     hes_last_error() gives the message of the calling thread's last failure.
 *     HES_ERR_DOMAIN corresponds to the Julia DomainError the reference raises
 *     for sqrt of a negative det(rho A) (HyperelasticityMPh.jl:114).
 *   - Arrays are FP64 in the reference's Julia layout: a set of n states is the
 *     column-major Matrix (ncomp, n), i.e. component fastest.  ncomp is 14 for
 *     one phase (the phase block without the volume fraction:
 *     [rho, rho u1, rho u2, rho u3, rho E, rho A11, rho A21, rho A31, rho A12,
 *     ..., rho A33]) and 30 for two phases
 *     (HyperelasticityMPh.jl:80-84: [a, a rho, a rho u(3), a rho E, a rho vec(A)]
 *     per phase).  3x3 tensors are column-major 9-vectors (Julia vec()).
 *   - Materials are `nphase` consecutive blocks of 8 doubles:
 *     (rho0, c0, cv, t0, b0, alpha, beta, gamma), the keyword arguments of the
 *     Barton2009 constructor (EquationsOfState.jl:90); b0^2 and
 *     k0 = c0^2 - 4/3 b0^2 are derived inside (EquationsOfState.jl:111-112).
 *   - Host pointers are never retained.  A solver handle is not thread-safe;
 *     the point-wise functions are re-entrant.
 */
#ifndef HES_H
#define HES_H

#ifdef __cplusplus
extern "C" {
#endif

#define HES_ABI_VERSION 1

#define HES_OK 0
#define HES_ERR_ARG (-1)
#define HES_ERR_HIP (-2)
#define HES_ERR_DOMAIN (-3)

#define HES_FLUX_RUSANOV 0 /* lxf with lambda = max(lambda_l, lambda_r), main.jl:46-53 */
#define HES_FLUX_HLL 1     /* hll, main.jl:56-57 */
#define HES_FLUX_LXF 2     /* Lax-Friedrichs: lxf with lambda = dx / dt on every face,
                              main.jl:30-41 and the variant of main.jl:50 */

typedef struct hes_solver hes_solver;

const char* hes_last_error(void);
int hes_version(void);
int hes_device_info(char* arch_name, int len, int* compute_units, double* mem_gb);

/* ---- Solver: replaces the time loop of main.jl:202-241 together with the
 *      per-cell update_cell (main.jl:43-60), the CFL reduction (main.jl:204-212)
 *      and NumFluxes.lxf / hll called per face.  The state lives on the GPU as
 *      structure-of-arrays FP64; cells 1 and n are frozen (main.jl:219-220). */

/* Allocates device state for `ncells` cells.  dx = X / nx (main.jl:147). */
int hes_solver_create(hes_solver** out, int nphase, int flux, long long ncells,
                      const double* materials, double cfl, double dx);
int hes_solver_destroy(hes_solver* h);

/* Upload Q0 (ncomp x ncells, Julia layout) and set the time to t0; computes the
 * sound speeds and the CFL maximum of the state (main.jl:179, 204-212). */
int hes_solver_set_state(hes_solver* h, const double* Q_host, double t0);
int hes_solver_set_state_device(hes_solver* h, const double* Q_device, double t0);
/* Piecewise-constant state built on the device: `nstates` states (ncomp x
 * nstates) separated by nstates-1 sorted interfaces in (0,1); cell i (0-based,
 * centre (i+1/2)/ncells) takes state #{j : centre > interface_j}, which for one
 * interface at 1/2 is initial_condition (main.jl:99-106). */
int hes_solver_set_piecewise(hes_solver* h, int nstates, const double* states,
                             const double* interfaces, double t0);
/* Download the current state (Julia layout). */
int hes_solver_get_state(hes_solver* h, double* Q_host);
/* Per-phase sound speed max_k sqrt|eig_k(Omega)| of every cell, (nphase, ncells)
 * (single-phase Rusanov from 2^21 cells, whose steps keep none: formed for
 * the current state on the call). */
int hes_solver_get_speeds(hes_solver* h, double* c_host);
/* Time and number of steps taken since the state was set. */
int hes_solver_time(hes_solver* h, double* t, long long* steps);
/* dt = cfl dx / max_i lambda_i of the current state (main.jl:212). */
int hes_solver_next_dt(hes_solver* h, double* dt);

/* Advance while steps < max_steps and t < t_end (main.jl:202 `while t < T`;
 * pass INFINITY for a fixed step count).  use_graph replays the launches of
 * each 60-step chunk from a captured hipGraph.  Outputs may be NULL:
 * steps_done, t_now, elapsed_ms (HIP events on the solver stream around the
 * whole run) and dt_out (max_steps doubles: the dt of every executed step). */
int hes_solver_run(hes_solver* h, long long max_steps, double t_end, int use_graph,
                   long long* steps_done, double* t_now, float* elapsed_ms, double* dt_out);

/* Optional source, two phases only (the reference has no source terms):
 * interphase velocity relaxation with time scale tau, applied to the interior
 * cells after each flux update.  The relative velocity decays exactly by
 * exp(-dt / tau); mixture momentum, phase masses and total energy are kept (the
 * dissipated kinetic energy goes to the two phases in equal halves).
 * tau = INFINITY switches it off (the default). */
int hes_solver_set_relaxation(hes_solver* h, double tau);

/* Optional stiff source, two phases only (the reference has none): pressure
 * relaxation with time scale tau, solved implicitly per interior cell after
 * each flux update (and after the velocity relaxation):
 *     a1' - a1 = (dt / tau) (p1' - p2') / Z,   Z = sum_k rho_k c0_k^2 / a_k,
 * p_k = -tr(sigma_k) / 3 of the end state.  Each phase keeps its mass,
 * momentum and distortion shape (a rho A scales by (a'/a)^(1/3)), and
 * a_k rho_k E_k changes by -p_I (a_k' - a_k) with p_I the mean pressure of the
 * start state: mixture mass, momentum and energy are conserved, a1 + a2 is
 * kept, and as tau -> 0 each step ends in pressure equilibrium.
 * tau = INFINITY switches it off (the default). */
int hes_solver_set_pressure_relaxation(hes_solver* h, double tau);

/* Time `reps` back-to-back steps (the step kernels and nothing else) with HIP
 * events on the solver stream; the state advances by reps steps.  The steps
 * are captured into one graph before the first event (as hes_solver_run
 * replays them), so the timed region holds only the graph's launch. */
int hes_solver_time_steps(hes_solver* h, long long reps, float* elapsed_ms);

/* Measured HBM bandwidth of a plain device copy of `bytes` bytes (read plus
 * written bytes per second, GB/s, best of `reps`): the second roofline
 * denominator beside the 8 TB/s specification.  No counterpart in the reference. */
int hes_copy_bandwidth(long long bytes, int reps, double* gbs);

/* ---- Batched solver: many independent 1D problems of uneven length (the
 *      reference runs one problem per process, main.jl:130-152) advanced with
 *      one launch per step.  Problem p has ncells[p] >= 3 cells, its own
 *      materials (nphase blocks of 8), CFL number, dx and hence its own dt
 *      (main.jl:204-214 per problem) and frozen end cells.  The packed state
 *      is ncomp x sum(ncells), problems one after another (Julia layout). */
typedef struct hes_batch hes_batch;

/* cfl, dx: per problem, or NULL for 0.6 and 1/ncells[p] (main.jl:144-147). */
int hes_batch_create(hes_batch** out, int nphase, int flux, int nproblems, const long long* ncells,
                     const double* materials, const double* cfl, const double* dx);
int hes_batch_destroy(hes_batch* h);
int hes_batch_info(hes_batch* h, long long* total_cells, int* ntiles, int* block, int* xcd,
                   int* tiles_per_wave);
/* t0: per problem start times, or NULL for 0. */
int hes_batch_set_state(hes_batch* h, const double* Q_host, const double* t0);
/* Per-problem piecewise-constant state built on the device: problem p has
 * nstates[p] conservative states (consecutive in `states`, ncomp each) and
 * nstates[p]-1 sorted interfaces in (0,1) (consecutive in `interfaces`). */
int hes_batch_set_piecewise(hes_batch* h, const int* nstates, const double* states,
                            const double* interfaces, const double* t0);
int hes_batch_get_state(hes_batch* h, double* Q_host);
/* Time and executed steps of every problem. */
int hes_batch_get_times(hes_batch* h, double* t, long long* steps);
/* Per-problem end time (main.jl:202 `while t < T`), NULL for none. */
int hes_batch_set_t_end(hes_batch* h, const double* t_end);
/* Launch shape: block of 64 or 256 threads; xcd != 0 deals consecutive blocks
 * to one XCD; tiles_per_wave (1..64) adjacent tiles of 62 cells per wave in
 * turn (a grid of ceil(tiles / tiles_per_wave) waves, each lane taking that
 * many cells). */
int hes_batch_set_launch(hes_batch* h, int block, int xcd, int tiles_per_wave);
/* Advance every problem by `steps` steps.  mode 0: one launch per step;
 * 1: the whole step loop captured in one hipGraph and launched once;
 * 2: one launch per problem per step (the per-problem loop). */
int hes_batch_run(hes_batch* h, long long steps, int mode, float* elapsed_ms);
/* Time every launch shape (block 64/256 x XCD order on/off x 1, 2 or 4 tiles
 * per wave) for `steps` steps on the current state (which is restored
 * afterwards) and keep the fastest. */
int hes_batch_autotune(hes_batch* h, int steps, int* block, int* xcd, int* tiles_per_wave,
                       float* ms_per_step);

/* ---- Adjoint of a fixed-length single-phase Rusanov run (the reference has
 *      no adjoint; its forward model is main.jl:199-225 with NumFluxes.jl:25-60
 *      and HyperelasticityMPh.jl:63-175).  Loss of the final state:
 *          L = sum_i (rho_i - rho*_i)^2 + (u1_i - u1*_i)^2
 *      (rho, u1 as cons2prim gives them, HyperelasticityMPh.jl:106-121).
 *      Gradients with respect to the primitive initial pieces (u(3), vec(A)(9),
 *      S per piece; rho = rho0 / det A) and the material constants b0, c0.
 *      Every `every`-th state is stored; the states in between are recomputed
 *      by the backward pass.  Only alpha = 1, beta = 3, gamma = 2 materials. */
typedef struct hes_adjoint hes_adjoint;

/* material: the 8 Barton2009 constants; nsteps: steps of the run (no end time);
 * ncells below 2^28 (the reverse step addresses a row's bytes with 32-bit
 * offsets), else HES_ERR_ARG. */
int hes_adjoint_create(hes_adjoint** out, long long ncells, const double* material, double cfl,
                       double dx, long long nsteps, long long every);
int hes_adjoint_destroy(hes_adjoint* h);
/* nstates pieces, prims[13 * k + (0:3 u, 3:12 vec(A), 12 S)], nstates-1 sorted
 * interfaces (cell i takes piece #{j : (i+1/2)/ncells > s_j}, as
 * hes_solver_set_piecewise). */
int hes_adjoint_set_initial(hes_adjoint* h, int nstates, const double* prims,
                            const double* interfaces);
/* Target profiles rho*, u1* (ncells each). */
int hes_adjoint_set_target(hes_adjoint* h, const double* rho, const double* u1);
/* Forward run from the initial pieces: stores the checkpoints, returns L. */
int hes_adjoint_forward(hes_adjoint* h, double* loss, float* elapsed_ms);
/* Reverse pass after hes_adjoint_forward: grad_states[13 * nstates] (same
 * layout as prims), grad_material[2] = (dL/db0, dL/dc0). */
int hes_adjoint_backward(hes_adjoint* h, double* grad_states, double* grad_material,
                         float* elapsed_ms);
/* on != 0 (the default): reverse steps skip tiles of cells whose cotangent is
 * zero on the tile and its two halo cells (the result is the same bit for bit:
 * such a tile's contribution is exactly zero); 0 evaluates every tile. */
int hes_adjoint_set_sparse(hes_adjoint* h, int on);
/* use_graph != 0 (the default): each pass is one captured hipGraph, replayed;
 * 0: the kernels of a pass are launched one by one. */
int hes_adjoint_set_launch(hes_adjoint* h, int use_graph);
/* Final state of the forward run (14 x ncells, Julia layout). */
int hes_adjoint_get_state(hes_adjoint* h, double* Q_host);
/* rho and u1 of every cell of the final state, computed exactly as the misfit
 * computes them (targets made from a run's profiles then give a zero misfit
 * wherever the two runs agree bit for bit). */
int hes_adjoint_get_profiles(hes_adjoint* h, double* rho, double* u1);
/* Device memory held by the handle and the number of stored checkpoints. */
int hes_adjoint_info(hes_adjoint* h, long long* device_bytes, long long* checkpoints);

/* ---- Adjoint of a batch of independent fixed-length single-phase Rusanov
 *      runs (the reference has neither an adjoint nor a batch; each problem
 *      is the forward model of hes_adjoint_*, main.jl:199-225 with
 *      NumFluxes.jl:25-60, with its own material, dx and CFL dt).  Problem p
 *      owns packed cells [off_p, off_p+1), off_0 = 0, off_p+1 = off_p +
 *      ncells[p]; every problem takes the same number of steps.  Its loss is
 *          L_p = sum_{i in p} (rho_i - rho*_i)^2 + (u1_i - u1*_i)^2
 *      and the gradients are those of L_p with respect to its own pieces and
 *      its own b0, c0.  Only alpha = 1, beta = 3, gamma = 2 materials. */
typedef struct hes_badjoint hes_badjoint;

/* materials: 8 Barton2009 constants per problem (8 x nproblems); dx: one per
 * problem, or NULL for 1 / ncells[p]; nsteps: steps of every problem; fewer
 * than 2^28 cells in all, else HES_ERR_ARG. */
int hes_badjoint_create(hes_badjoint** out, int nproblems, const long long* ncells,
                        const double* materials, double cfl, const double* dx, long long nsteps,
                        long long every);
int hes_badjoint_destroy(hes_badjoint* h);
/* nstates[p] pieces per problem; prims: 13 per piece, problems in order
 * (layout of hes_adjoint_set_initial); interfaces: nstates[p] - 1 sorted
 * positions in (0, 1) per problem, problems in order. */
int hes_badjoint_set_initial(hes_badjoint* h, const int* nstates, const double* prims,
                             const double* interfaces);
/* Target profiles rho*, u1* over the packed cells. */
int hes_badjoint_set_target(hes_badjoint* h, const double* rho, const double* u1);
/* Forward runs: stores the checkpoints; losses[p] = L_p (may be NULL), total = sum_p L_p. */
int hes_badjoint_forward(hes_badjoint* h, double* losses, double* total, float* elapsed_ms);
/* Reverse pass: grad_states (13 per piece, layout of prims), grad_material[2p],
 * [2p + 1] = (dL_p/db0_p, dL_p/dc0_p). */
int hes_badjoint_backward(hes_badjoint* h, double* grad_states, double* grad_material,
                          float* elapsed_ms);
/* As hes_adjoint_set_sparse / hes_adjoint_set_launch. */
int hes_badjoint_set_sparse(hes_badjoint* h, int on);
int hes_badjoint_set_launch(hes_badjoint* h, int use_graph);
/* Final packed state (14 x total cells, Julia layout) and its rho, u1 profiles
 * as the misfit computes them. */
int hes_badjoint_get_state(hes_badjoint* h, double* Q_host);
int hes_badjoint_get_profiles(hes_badjoint* h, double* rho, double* u1);
int hes_badjoint_info(hes_badjoint* h, long long* total_cells, long long* device_bytes,
                      long long* checkpoints);

/* ---- Point-wise functions over n states (host arrays, synchronous). ---- */

/* HyperelasticityMPh.cons2prim_mph (HyperelasticityMPh.jl:99-133):
 * P = [a, rho, u(3), S, vec(A)] per phase (without a for one phase). */
int hes_cons2prim(int nphase, const double* materials, const double* Q, double* P, long long n);
/* HyperelasticityMPh.prim2cons_mph (HyperelasticityMPh.jl:63-90). */
int hes_prim2cons(int nphase, const double* materials, const double* P, double* Q, long long n);
/* HyperelasticityMPh.flux_mph (HyperelasticityMPh.jl:140-175). */
int hes_flux(int nphase, const double* materials, const double* Q, double* F, long long n);
/* HyperelasticityMPh.get_eigvals(eos, Q, nvec) (HyperelasticityMPh.jl:252-266), one
 * direction nvec (3 doubles; NULL means [1,0,0]) for all states: per phase
 * [u.n + c_k (k=1..3), u.n - c_k], c_k = sqrt|eig_k| in descending eig order. */
int hes_get_eigvals(int nphase, const double* materials, const double* Q, const double* nvec,
                    double* eig, long long n);
/* maximum(abs.(get_eigvals(eos, Q, [1,0,0]))) per state (main.jl:209-210). */
int hes_max_speed(int nphase, const double* materials, const double* Q, double* lambda,
                  long long n);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  hes_max_speed
 * on device arrays: Q_dev (ncomp x n, Julia layout, device memory) gives
 * lambda_dev[i] (device, n doubles; may be NULL) as hes_max_speed does, and
 * *lmax (host) receives max_i lambda_i, the CFL maximum of main.jl:204-212
 * (0 for n = 0).  It allocates nothing: the per-block partial maxima and
 * status words live in a fixed device array of 4 MB that each GPU the library
 * runs on holds until the process ends, so n is limited to
 * HES_NOALLOC_MAX_STATES (larger n: HES_ERR_ARG).  Synchronous, on the null
 * stream of the current device.  Calls that overlap on one GPU (from several
 * host threads) share that array and race. */
#define HES_NOALLOC_MAX_STATES (1LL << 26)
int hes_max_speed_noalloc(int nphase, const double* materials, const double* Q_dev,
                          double* lambda_dev, double* lmax, long long n);
/* HyperelasticityMPh.noncons_flux (HyperelasticityMPh.jl:178-250), two phases:
 * dense 30x30 column-major per state (n x 900 doubles). */
int hes_noncons_flux(const double* materials, const double* Q, double* B, long long n);

/* NumFluxes.lxf (NumFluxes.jl:25-60; aux = lambda per face) and NumFluxes.hll
 * (NumFluxes.jl:70-132; aux = [minimum(eigvals[1]), maximum(eigvals[2])] per
 * face) for n faces (QL, QR: ncomp x n).  Outputs are the (cons, dm, dp)
 * triple the Julia functions return. */
int hes_face_flux(int nphase, int flux, const double* materials, const double* QL,
                  const double* QR, const double* aux, double* cons, double* dm, double* dp,
                  long long n);

/* One phase block of HyperelasticityMPh (15 components [a, a rho, a rho u(3),
 * a rho E, a rho vec(A)], one material): the per-phase methods
 * cons2prim(eos, Q) (HyperelasticityMPh.jl:106-133; P = [a, rho, u(3), S,
 * vec(A)]), flux(eos, Q) (HyperelasticityMPh.jl:146-175) and
 * get_eigvals(eos, Q, n) (HyperelasticityMPh.jl:258-266; 6 per state as
 * hes_get_eigvals, nvec NULL for [1,0,0]). */
int hes_phase_cons2prim(const double* material, const double* Q, double* P, long long n);
int hes_phase_flux(const double* material, const double* Q, double* F, long long n);
int hes_phase_get_eigvals(const double* material, const double* Q, const double* nvec,
                          double* eig, long long n);

/* Hyperelasticity, the single-phase module (Hyperelasticity.jl:25-114), in its
 * own 13-component layout Q = [rho u (3), rho F (9, row-major: rho F[i,j] at
 * 3 i + j, 1-based), rho E], one material.  The Julia module cannot run as
 * written (it calls density, energy and stress without the equation of
 * state); these take the material first and keep the formulas, with
 * den = sqrt(det(rho F) / rho0) as HyperelasticityMPh forms it.
 *   hes_sp_cons2prim (Hyperelasticity.jl:25-36): P = [den, vel (3), vec(F) (9,
 *     column-major: the memory of Julia's 3x3 F), e_int], 14 per state;
 *   hes_sp_prim2cons (Hyperelasticity.jl:70-96): P = [vel (3), vec(F) (9,
 *     column-major), S], 13 per state, den = rho0 / det F;
 *   hes_sp_flux (Hyperelasticity.jl:99-114): 13 per state. */
int hes_sp_cons2prim(const double* material, const double* Q, double* P, long long n);
int hes_sp_prim2cons(const double* material, const double* P, double* Q, long long n);
int hes_sp_flux(const double* material, const double* Q, double* F, long long n);

/* EquationsOfState (single material, 8 doubles) */
/* energy(eos, S, G) (EquationsOfState.jl:118-137); G: 9 x n */
int hes_energy(const double* material, const double* S, const double* G, double* e, long long n);
/* entropy(eos, e_int, G) (EquationsOfState.jl:139-156) */
int hes_entropy(const double* material, const double* e_int, const double* G, double* S,
                long long n);
/* ForwardDiff.derivative(S -> energy(eos, S, G), S) (HyperelasticityMPh.jl:212) */
int hes_temperature(const double* material, const double* S, const double* G, double* T,
                    long long n);
/* stress(eos, ent, F) (EquationsOfState.jl:179-190); sig: 9 x n column-major */
int hes_stress(const double* material, const double* ent, const double* F, double* sig,
               long long n);
/* stress(eos, den, e_int, F) (EquationsOfState.jl:162-176) */
int hes_stress_den(const double* material, const double* den, const double* e_int,
                   const double* F, double* sig, long long n);
/* acoustic(eos, ent, F, n) (EquationsOfState.jl:223-246); nvec: 3 x n */
int hes_acoustic(const double* material, const double* ent, const double* F, const double* nvec,
                 double* ac, long long n);
/* acoustic(eos, P, n) (EquationsOfState.jl:192-221); P: 15 x n phase primitives */
int hes_acoustic_prim(const double* material, const double* P, const double* nvec, double* ac,
                      long long n);

/* Strains (Strains.jl:26-52) */
int hes_finger(const double* A, double* G, long long n);
int hes_invariants(const double* G, double* inv, long long n);

#ifdef __cplusplus
}
#endif
#endif /* HES_H */
