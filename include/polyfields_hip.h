/*
 * polyfields_hip.h -- C ABI of the MI355X backend for PolyFields.jl's SCFT hot path.
 *
 * One FieldSystem lives on one GPU; the library owns its device grids, the host owns
 * every buffer it passes in.  Grids cross the boundary as float64 arrays of n = nx*ny*nz
 * entries in the reference's memory order (Julia Array{Float64,3}: x fastest), so a
 * Julia `ccall` passes `pointer(A)` without a transpose.  Every function returns 0 on
 * success and a negative code on failure; `pf_last_error()` returns the message (thread
 * local).  Codes: -1 bad argument, -2 invalid model or state, -3 HIP runtime error,
 * -4 singular matrix, -5 divergence.  A system and the updaters bound to it are not
 * re-entrant: one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to PolyFields.jl/src):
 *   pf_system_create ........ FieldSystem(dims, cell; ensemble, compress, mde)  system/system.jl:51-74
 *                             with UnitCell(dim, crystal, params...)             system/cell.jl:31-54
 *   pf_system_add_monomer ... add_monomer!(sys, Monomer(id, vol, charge))       system/system.jl:113-128
 *   pf_system_add_multiblock  add_species!(sys, Multiblock(...), phi)           system/system.jl:137-161,
 *                                                                               species/multiblock.jl:32-77
 *   pf_system_add_point ..... add_species!(sys, Point(mon), phi)                species/point.jl:19-22
 *   pf_system_set_flory ..... set_interaction!(FloryInteraction, a, b, chi)     interactions/flory.jl:25-37
 *   pf_system_set_edwards ... set_interaction!(EdwardsInteraction, a, u)        interactions/edwards.jl:25-32
 *   pf_system_set_field ..... sys.fields[mid] .= w (fieldinit! dict form)       fileio.jl:25-54
 *   pf_system_fieldinit_seeded fieldinit!(sys; seed, scale)                     fileio.jl:14-23
 *   pf_system_get_grid ...... sys.fields / density / residuals / potentials     system/system.jl:35-38
 *   pf_system_get_species_density  species.density[mid]                         species/multiblock.jl:82
 *   pf_system_get_propagator  species.q[s] / species.qc[s]                      species/multiblock.jl:80-81
 *   pf_system_species_scalars species.Q, sys.phi_species, sys.mu_species         system/system.jl:44-45
 *   pf_system_cell .......... cell.params / update!(cell)                       system/cell.jl:96-109
 *   pf_system_add_cell ...... add_cell!(sys, cell)                              system/system.jl:102-106
 *   pf_validate ............. validate(sys)                                     system/system.jl:192-237
 *   pf_density .............. density!(sys)                                     system/system.jl:266-289
 *   pf_residuals ............ residuals!(sys)                                   system/system.jl:317-340
 *   pf_field_error .......... field_error(sys)                                  system/system.jl:470-486
 *   pf_free_energy .......... free_energy(sys)                                  system/system.jl:402-433
 *   pf_free_energy_bulk ..... free_energy_bulk(sys)                             system/system.jl:440-458
 *   pf_scfstress ............ scfstress(sys)                                    system/system.jl:384-395
 *   pf_updater_anderson ..... AndersonUpdater(; nhist, iter_euler, lam_euler)   updaters/anderson.jl:36-48
 *   pf_updater_euler ........ EulerUpdater(; method, lam)                       updaters/euler.jl:20-28
 *   pf_updater_momentum ..... MomentumUpdater(; method, lam, gamma, beta1, beta2) updaters/momentum.jl:30-41
 *   pf_updater_setup/step ... setup!(u, sys) / step!(u)                         updaters/base.jl:7-9
 *   pf_domain_create ........ DomainUpdater(; nper, nskip, lam, pmax, rlxn)     updaters/domain.jl:20-24
 *   pf_domain_setup/step .... setup!(d, sys) / step!(d; nsteps, tol)            updaters/domain.jl:32-92
 *   pf_domain_stress ........ domain.stress / stress_error(d)                   updaters/domain.jl:30,44-48
 *   pf_scft ................. scft!(sys, updater; domain, opts) -> SCFTResults  simulate.jl:60-148
 *   pf_scft_save ............ the same, with savefields every nsave steps      simulate.jl:124-128
 *   pf_cell_geometry ........ make_basis / basis! / dbasis! / ddbasis!          system/basis.jl:11-116,
 *                                                                               system/cell.jl:113-258
 *   pf_chain_grid ........... chain_grid(N_block, Ns)                           species/base.jl:27-56
 *
 * Reverse mode (no counterpart in the reference, which has no derivative code; these
 * differentiate its forward path and replace what a ChainRulesCore.rrule of
 * `free_energy(sys)` would bind):
 *   pf_fhelm_grad ........... d free_energy / d(w, cell, chi, u, phi, N_blk)    system/system.jl:402-433
 *                             through density!(sys)                             system/system.jl:266-289
 *   pf_density_vjp .......... vector-Jacobian product of density!(sys)         species/multiblock.jl:125-231
 *   pf_species_set_blocks ... continuous block lengths on a fixed contour grid  species/base.jl:27-56
 */
#ifndef POLYFIELDS_HIP_H
#define POLYFIELDS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_PRECISION_FP64 0 /* every grid and transform in float64 ("plain") */
#define PF_PRECISION_FP32 1 /* float32 propagator and transforms, float64 fields and sums */

#define PF_CANONICAL 0
#define PF_GRAND 1

#define PF_GRID_FIELD 0
#define PF_GRID_DENSITY 1
#define PF_GRID_RESIDUAL 2
#define PF_GRID_POTENTIAL 3

typedef struct pf_system pf_system;
typedef struct pf_updater pf_updater;
typedef struct pf_domain pf_domain;

typedef struct {
    int nsteps;    /* SCFTOptions.nsteps (default 100)  */
    int nsout;     /* print progress every nsout steps (0: quiet) */
    int nsave;     /* pf_scft_save calls its save function every nsave steps (0: never) */
    double ftol;   /* default 1e-5 */
    double stol;   /* default 1e-4 */
} pf_scft_options;

typedef struct {
    double time;
    int nsteps;
    int converged;
    double ferr;
    double serr;
    double fhelm;
} pf_scft_results;

const char* pf_last_error(void);
int pf_version(void);
int pf_device_count(int* count);

/* Host-only helpers (no GPU needed). R9/G9 are row-major 3x3 (column j = vector j);
   dGG has 9*nparams entries indexed [(i*3+j)*nparams + k]. */
int pf_cell_geometry(int dim, const char* crystal, const double* params, int nparams,
                     double* R9, double* G9, double* volume, double* dGG);
int pf_chain_grid(int nblk, const double* Nref_block, int Ns, int* begin /* nblk+1, 1-based */,
                  double* ds_block, double* Nref);

/* System */
int pf_system_create(pf_system** out, int nx, int ny, int nz, int dim, const char* crystal,
                     const double* params, int nparams, int ensemble, double inv_zeta,
                     int precision, int device);
int pf_system_destroy(pf_system* sys);
int pf_system_add_monomer(pf_system* sys, int id, double vol, double charge);
int pf_system_add_multiblock(pf_system* sys, int nblk, const int* block_mids, const int* N_block,
                             const double* b_block, int Ns, double phi_or_mu, int* species_index,
                             int* Ns_used);
int pf_system_add_point(pf_system* sys, int mid, double phi_or_mu, int* species_index);
int pf_system_set_flory(pf_system* sys, int mid_a, int mid_b, double chi);
int pf_system_set_edwards(pf_system* sys, int mid, double u);
int pf_system_set_field(pf_system* sys, int mid, const double* src);
int pf_system_get_grid(pf_system* sys, int which, int mid, double* dst);
/* fieldinit!(sys; seed, scale) (fileio.jl:14-23): Random.seed!(seed), then
   sys.fields[mid] .= scale * randn(dims) for each monomer in the iteration order of the
   reference's Dict{Int} of fields (system/system.jl:35,122; monomers in the order they
   were added), then residuals!.  The draws restate Julia 1.5's MersenneTwister (dSFMT-19937)
   and randn (256-layer ziggurat), so the fields equal the reference's for the same seed. */
int pf_system_fieldinit_seeded(pf_system* sys, uint64_t seed, double scale);
/* The same stream without a system: Random.seed!(seed) then randn(counts[k]) for
   k = 0..narrays-1, written one after another to out (sum of counts values).  Host code;
   no GPU call.  No counterpart as a function in the reference: it exposes the draws of
   fileio.jl:16-19 for checking. */
int pf_julia_randn(uint64_t seed, const int64_t* counts, int narrays, double* out);
/* Seeding from densities (fieldinit! with field_type = :rho, fileio.jl:25-54): write
   sys.density[mid], then pf_uniform_fields sets w = potentials(rho), mean-subtracted in
   the canonical ensemble (uniform_fields!, system/system.jl:348-357). */
int pf_system_set_density(pf_system* sys, int mid, const double* src);
int pf_uniform_fields(pf_system* sys);
int pf_system_get_species_density(pf_system* sys, int species, int mid, double* dst);
int pf_system_get_propagator(pf_system* sys, int species, int s, int backward, double* dst);
int pf_system_species_scalars(pf_system* sys, int species, double* Q, double* phi, double* mu);
/* set != 0: write params (length nparams) and update!(cell); else read them. */
int pf_system_cell(pf_system* sys, double* params, int nparams, int set);
/* add_cell!(sys, UnitCell(dim, crystal, params...)): replace the system's cell, crystal
   system included (the dimension must match the grid's); as fieldinit!(sys, fpath;
   load_cell = true) does with the cell of a field file. */
int pf_system_add_cell(pf_system* sys, int dim, const char* crystal, const double* params,
                       int nparams);
int pf_system_synchronize(pf_system* sys);
/* Switch the propagator precision (PF_PRECISION_FP64 / _FP32).  Fields, densities and
   every sum are float64 in both modes, so the state carries over; the q stacks are
   reallocated.  Used to polish an FP32-converged state in FP64. */
int pf_system_set_precision(pf_system* sys, int precision);
/* Propagator scheme: 0 RK2 (default), 1 RQM4 (mde.jl:83-104; 1-D, 2-D and 3-D grids).
   Replaces the `mde` keyword of FieldSystem (system/system.jl:51-56). */
int pf_system_set_mde(pf_system* sys, int method);

int pf_validate(pf_system* sys);
int pf_density(pf_system* sys);
int pf_residuals(pf_system* sys);
int pf_field_error(pf_system* sys, double* ferr);
int pf_free_energy(pf_system* sys, double* fhelm);
int pf_free_energy_bulk(pf_system* sys, double* fhelm);
int pf_scfstress(pf_system* sys, double* stress, int nparams);
/* pf_scfstress without device allocation.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  Every device buffer of the stress pass (the
   row table, the per-row and k-space partial sums, the transforms of q_s and q+_s) is
   carved from a fixed 32 MiB array in the library's code object on the system's GPU
   (held until the process ends), so nothing is allocated or freed.  Requirements: the
   system has run density! since its species last changed (the propagator tables exist),
   its stacks are full (no checkpointing), and one stress row's transforms fit: grids up to
   128^3 in FP32 and below 128^3 in FP64; beyond that it fails with -1 and a message giving
   the bytes needed.  The array is shared by every call on that GPU, so calls that overlap
   on one GPU race; the contents left by one call are never read by the next. */
int pf_scfstress_noalloc(pf_system* sys, double* stress, int nparams);

/* Free-energy functionals differentiated by pf_fhelm_grad. */
#define PF_FHELM_REFERENCE 0 /* free_energy exactly as system.jl:402-433 (mu = phi) */
#define PF_FHELM_TEXTBOOK 1  /* the same with mu = log(phi / Q) (system.jl:248-259) */

/* Free energy and its gradient at the current fields (canonical ensemble, RK2).  Runs
   density!(sys) and a reverse sweep of every chain; the fields are used as given (not
   mean-subtracted).  Outputs (each may be NULL): dw [nmon][n] (slot = ascending monomer
   id, x fastest), dcell [nparams], dchi [nmon][nmon] (Flory pairs; zero elsewhere),
   du [nmon] (Edwards parameters), dphi [nspecies], dN [ndN] = every block of every chain
   species in species order (derivative with the contour grid held fixed, as
   pf_species_set_blocks changes it). */
int pf_fhelm_grad(pf_system* sys, int functional, double* fhelm, double* dw, double* dcell,
                  int nparams, double* dchi, double* du, double* dphi, double* dN, int ndN);
/* Vector-Jacobian product of density!(sys): cotangents rho_bar [nmon][n] of the system
   densities and Q_bar [nspecies] of the species partition functions -> dw [nmon][n],
   dcell [nparams], dN [ndN] (each output may be NULL). */
int pf_density_vjp(pf_system* sys, const double* rho_bar, const double* Q_bar, double* dw,
                   double* dcell, int nparams, double* dN, int ndN);
/* Checkpointed propagator sweeps (K > 0: a checkpoint every K slices; -1: K = ceil(sqrt(Ns+1));
   0: full stacks, the default).  Each chain then keeps q and q+ at the checkpoints plus one
   segment of K+1 slices per direction, about 2 (Ns/K + K) slices instead of 2 (Ns+1);
   density!, residuals!, free energies, pf_scfstress (so also the DomainUpdater),
   pf_fhelm_grad and pf_density_vjp recompute segments as they need them, with RK2 or RQM4
   steps (the gradient is RK2 only).  The propagator read-back needs full stacks and
   raises.  No counterpart in the reference, which keeps every slice
   (species/multiblock.jl:109-110). */
int pf_system_set_checkpoint(pf_system* sys, int K);
/* Set the block lengths of chain species isp to continuous values N_block[nblk]; every
   block keeps its number of contour steps and ds_b = v_b N_b / steps_b. */
int pf_species_set_blocks(pf_system* sys, int isp, const double* N_block, int nblk);

/* Field updaters.  method: Euler 0 PEC, 1 PECE; Momentum 0 MOM, 1 RMS, 2 ADAM, 3 NAG.
   Anderson accepts any nhist >= 1, as the reference does (its default is 50): depths up
   to 16 reduce the Gram matrix in one pass over the history, deeper ones in 8 x 8 tiles;
   the history holds nhist + 1 field and residual grids in device memory. */
int pf_updater_anderson(pf_updater** out, int nhist, int iter_euler, double lam_euler);
int pf_updater_euler(pf_updater** out, int method, double lam);
int pf_updater_momentum(pf_updater** out, int method, double lam, double gamma, double beta1,
                        double beta2);
int pf_updater_setup(pf_updater* u, pf_system* sys);
int pf_updater_step(pf_updater* u);
int pf_updater_destroy(pf_updater* u);

/* Variable cell */
int pf_domain_create(pf_domain** out, int nper, int nskip, double lam, double pmax, double rlxn);
int pf_domain_setup(pf_domain* d, pf_system* sys);
/* Forward-difference step of the stress Jacobian; the reference fixes it at 1e-8
   (updaters/domain.jl:101), which is the default.  Larger steps make the Jacobian
   insensitive to last-bit rounding (the reference's step amplifies it ~1e8 times). */
int pf_domain_set_fd_step(pf_domain* d, double dx);
int pf_domain_step(pf_domain* d, int nsteps, double tol);
int pf_domain_stress(pf_domain* d, int refresh, double* stress, int nparams, double* serr);
int pf_domain_destroy(pf_domain* d);

/* Driver.  domain may be NULL (fixed cell). */
int pf_scft(pf_system* sys, pf_updater* u, pf_domain* d, const pf_scft_options* opts,
            pf_scft_results* res);
/* pf_scft with field saving (src/simulate.jl:124-128): save(user, step) is called where
   the reference calls savefields(sys, opts.savepath) -- after step `step` of the loop
   when nsave > 0 and step % nsave == 0 and the run has not converged -- and once more
   with step = -1 after the final residuals! (the reference's own final write names an
   option that SCFTOptions lacks, src/simulate.jl:140-142).  The callback may read the
   system's grids; it must not change them.  NULL save: pf_scft. */
typedef void (*pf_save_fn)(void* user, int step);
int pf_scft_save(pf_system* sys, pf_updater* u, pf_domain* d, const pf_scft_options* opts,
                 pf_scft_results* res, pf_save_fn save, void* user);

/* Batches.  Systems of one precision on one device advance together: every propagator
   launch covers the lanes of all systems (grouped by grid size and propagator method),
   every system-level kernel covers all systems.  A checkpointed system
   (pf_system_set_checkpoint) runs its own segment-wise density! inside the batch's
   launch sequence, which then runs eagerly instead of from a graph.  The systems stay
   owned by the caller and must outlive the batch; no other call may use them while a
   batch call runs.  pf_batch_scft runs
   scft!(sys, AndersonUpdater(; nhist, iter_euler, lam_euler)) with a fixed cell for every
   system in lockstep, dropping each from the active set once ferr < ftol; `results`
   receives one entry per system (time is the whole batch's wall time).  A system whose
   field error turns NaN (scft! would raise "Diverging field updater") is dropped and
   reported with ferr = NaN and converged = 0 (so is one whose Anderson matrix turns
   singular); the others continue.  Systems with mask[i] == 0 are left untouched and
   reported with ferr = NaN.
   No counterpart in the reference: it extends scft! to many independent systems. */
typedef struct pf_batch pf_batch;
int pf_batch_create(pf_batch** out, pf_system* const* systems, int nsys);
int pf_batch_destroy(pf_batch* b);
int pf_batch_density(pf_batch* b);
int pf_batch_set_precision(pf_batch* b, int precision);
int pf_batch_residuals(pf_batch* b);
/* The propagator sweeps of density! alone (src/species/multiblock.jl:136-172): the
   exponential tables from the current fields and the q / q+ sweeps of every chain of
   every system, leaving the stacks filled (readable with pf_system_get_propagator); no
   Q, Simpson, densities or pooling.  The chain of propagate! calls as one launch
   sequence, replayed from a graph like pf_batch_density. */
int pf_batch_sweeps(pf_batch* b);
/* Graph replay of the density!/residuals! launch sequence (default on): the launches of
   every lane group, Q, Simpson, pooling and residuals are captured once per active set
   and replayed; off runs them eagerly. */
int pf_batch_set_graph(pf_batch* b, int on);

/* Autotuner (no counterpart in the reference).  Times every candidate kernel schedule of
   the grid on the caller's own workload, `reps` density! evaluations each, and keeps the
   fastest for this grid size and precision for the rest of the process (the four
   fastest are re-timed in interleaved rounds).  Candidates:
   3-D xy-slab form: 0 generic, 1 compile-time 256 threads, 2 compile-time 512 threads,
     3 shuffle-exchange (FP32 64 x 64 planes), 4 compile-time 1024 threads (FP32 128^2);
   z-pencil form: 0 generic, 1 compile-time LDS, 2 register-resident, 3 two threads per
     column (FP32, 64-point lines), 4 LDS with one ky row per workgroup (64-point lines) or
     512 threads (128-point lines), 5 LDS with one ky row and 512 threads (64-point lines)
     or 1024 threads (FP32 128-point lines), 6 LDS with half a ky row per workgroup (64-
     and 128-point lines);
   in a batch, the streams per 3-D lane group (1 or 2);
   1-D/2-D planar kernel: 0 generic, 1 compile-time (2-D: 256 threads), 2 (2-D: 512),
     3 (2-D: 1024).
   Schedules with different transform forms round differently (about 1 ulp per step).
   sched receives {slab, pencil, planar, split} (-1: not a candidate for this grid). */
int pf_system_autotune(pf_system* sys, int reps, int* sched, double* best_ms);
/* Per distinct grid size of the batch: sched[8k .. 8k+7] = {bytes, nx, ny, nz, slab,
   pencil, planar, split} for k < min(cap, *count). */
int pf_batch_autotune(pf_batch* b, int reps, int* sched, int cap, int* count);
/* Forget every tuned schedule (built-in defaults again). */
int pf_schedule_reset(void);
/* Set the schedule of one grid size and precision (elem_bytes 4 or 8) by hand:
   sched = {slab, pencil, planar, split} with the form numbers above, -1 for the default
   of an entry (a form the grid does not have falls back to another compile-time form, or
   the generic one).  Systems already built pick it up at their next launch; batches
   rebuild on their next membership or cell change. */
int pf_schedule_set(int elem_bytes, int nx, int ny, int nz, const int* sched);
int pf_batch_scft(pf_batch* b, int nhist, int iter_euler, double lam_euler, int nsteps,
                  double ftol, const int* mask /* nsys flags or NULL: all */,
                  pf_scft_results* results);
/* pf_batch_scft with a variable cell: scft!(sys, AndersonUpdater(; nhist, iter_euler,
   lam_euler); domain = DomainUpdater(...)) for every system in lockstep
   (src/simulate.jl:60-148, src/updaters/domain.jl:50-120).  Each system gets its own
   DomainUpdater with the settings of `domain` (nper, nskip, lam, pmax, rlxn and the
   finite-difference step; `domain` itself is not set up or changed).  Every nskip steps
   each active system's stress is taken from its last residuals! and, above stol, its
   cell is relaxed by the Newton iterations, system by system; a system stops once
   ferr < ftol and serr <= stol (src/simulate.jl:37-49).  results[i].serr is the stress
   error the system's updater last computed, as scft! reports it.  A singular cell
   Jacobian drops the system like a singular Anderson matrix.  No counterpart in the
   reference beyond scft! itself. */
int pf_batch_scft_domain(pf_batch* b, int nhist, int iter_euler, double lam_euler,
                         const pf_domain* domain, int nsteps, double ftol, double stol,
                         const int* mask, pf_scft_results* results);
/* The batched loop with scft!(sys, MomentumUpdater(; method, lam, gamma, beta1, beta2);
   domain) for every system (src/updaters/momentum.jl:47-165): method 0 :MOM, 1 :RMS,
   2 :ADAM, 3 :NAG (residuals! of every active system at its look-ahead fields, inside
   the step).  Systems may be in the grand ensemble (Anderson's loops refuse them, as
   updaters/anderson.jl:61-63 does).  domain NULL: a fixed cell (stol unused); else as
   pf_batch_scft_domain.
   One kernel applies every active system's moment update and field step.  No
   counterpart in the reference beyond scft! itself. */
int pf_batch_scft_momentum(pf_batch* b, int method, double lam, double gamma, double beta1,
                           double beta2, const pf_domain* domain, int nsteps, double ftol,
                           double stol, const int* mask, pf_scft_results* results);
/* The batched loop with scft!(sys, EulerUpdater(; method, lam); domain) for every system
   (src/updaters/euler.jl:36-66): method 0 :PEC, 1 :PECE.  Every step runs the predictor
   of all active systems in one kernel, one batched residuals! at the predicted fields,
   the corrector, and for PECE a second batched residuals!.  Systems and domain as in
   pf_batch_scft_momentum.  No counterpart in the reference beyond scft! itself. */
int pf_batch_scft_euler(pf_batch* b, int method, double lam, const pf_domain* domain,
                        int nsteps, double ftol, double stol, const int* mask,
                        pf_scft_results* results);
int pf_batch_timing(pf_batch* b, int on);
int pf_batch_timing_read(pf_batch* b, int kind, double* total_ms, int* launches);
/* Mean number of propagator lanes covered by the sampled launches of one kind (3-D
   groups of 4 or more lanes are launched as two halves on two streams). */
int pf_batch_timing_lanes(pf_batch* b, int kind, double* lanes);

/* Measurement: HIP-event timing of the propagator kernels on the system's stream.  kind:
   0 xy-slab pass, 1 z-pencil pass, 2 whole-plane (1-D/2-D) sweep.  One launch in 8 of each
   kind runs 4 times back to back between two events; *launches counts every timed run.
   The repeats act on their own output, so densities computed with timing on are not
   meaningful (batches alike). */
int pf_system_timing(pf_system* sys, int on);
int pf_system_timing_read(pf_system* sys, int kind, double* total_ms, int* launches);

#ifdef __cplusplus
}
#endif

#endif /* POLYFIELDS_HIP_H */
