/*
 * mqe.h -- C ABI of the MI355X walker engine for MPS projector QMC.
 *
 * The engine owns one GPU-resident walker population and replaces the
 * per-walker CPU propagation of the reference's MPSQMC2 driver.  Every call
 * returns 0 on success or a negative MQE_E* code, never throws, and leaves a
 * message retrievable with mqe_last_error().  Host arrays are borrowed for the
 * duration of the call only.  One host thread per engine; calls on one engine
 * are ordered on its stream.
 *
 * Layout conventions (identical to the reference):
 *   MPS site tensor k:  A[row + dimL*(col + dimR*s)], dimL = vd[k], dimR = vd[k+1]
 *                       (MPStensor.cpp:91-95, :168-173); a flat MPS is the site
 *                       tensors concatenated in site order.
 *   couplings:          listed lexicographically over (first < second)
 *                       (MPSQMC2.cpp:286-299).
 *   Trotter tables:     per coupling c, S[c*4 + k], U[c*16 + i + d*(j + d*k)],
 *                       VT[c*16 + k + d*d*(i + d*j)]  (TrotterHeisenberg.h:83-90).
 *   grid:               grid[i + dim*g]  (GridGenerator.h:48-49).
 */
#ifndef MQE_H
#define MQE_H

#ifdef __cplusplus
extern "C" {
#endif

#define MQE_OK 0
#define MQE_EINVAL (-1)     /* bad argument / shape */
#define MQE_EDEVICE (-2)    /* HIP runtime error */
#define MQE_ECAPACITY (-3)  /* population exceeds the walker capacity */
#define MQE_ENUMERIC (-4)   /* zero PDF norm or non-finite weight on some walker */
#define MQE_ESTATE (-5)     /* call out of order (e.g. propagate before load) */

#define MQE_PREC_FP64 0     /* all contractions FP64 (parity mode) */
#define MQE_PREC_FP32 1     /* contractions FP32, QR/overlap/weights/energy FP64 */

typedef struct mqe_engine mqe_engine;

typedef struct {
    int L;                 /* chain length */
    int d;                 /* local dimension; the engine supports d = 2 */
    int precision;         /* MQE_PREC_FP64 or MQE_PREC_FP32 */
    int n_desired;         /* totalNDesiredWalkers (MPSQMC2.h:54) */
    int capacity;          /* walker slots; the reference uses max(1000, 3N) (MPSQMC2.cpp:84) */
    int max_bond;          /* largest walker bond dimension any walker will hold */
    double dtau;           /* imaginary time step */
    double field;          /* h of -h sum_i S^z_i (HeisenbergMPO.h:98-103) */
    unsigned long long seed; /* device RNG key (used when no host uniforms are given) */
} mqe_params;

typedef struct {
    int L;
    int d;
    const int *vd;         /* L+1 bond dimensions, vd[0] = vd[L] = 1 */
    const double *data;    /* concatenated site tensors */
} mqe_mps;

/* Replaces MPSQMC2::MPSQMC2(...) set-up of the trial and the propagator tables
 * (MPSQMC2.cpp:68-103, :219-332).  The trial is used as given (it must already
 * be left-normalized as SetupTrial does, MPSQMC2.cpp:239).  J[c] is the coupling
 * of pair (first[c], second[c]). */
int mqe_create(const mqe_params *p, const mqe_mps *trial, int n_couplings, const int *first,
               const int *second, const double *J, const double *tab_S, const double *tab_U,
               const double *tab_VT, const double *grid, int n_grid, int grid_dim, mqe_engine **out);

void mqe_destroy(mqe_engine *e);

/* Replaces MPSQMC2::SetupWalkers (MPSQMC2.cpp:381-441): n walkers, walker i has
 * bond dims vd[i*(L+1) ...] and its tensors at data[offset_i ...] (walkers
 * concatenated), weight w[i], overlap O[i], energy history eh[i]. */
int mqe_load_walkers(mqe_engine *e, int n, const int *vd, const double *data, const double *w,
                     const double *O, const double *eh);

/* Replaces the EnergyFunctionAndHistory(1, &E, false) call at the start of
 * MPSQMC2::Walk (MPSQMC2.cpp:447): local energies with the stored overlaps, no
 * history update.  Also brings the walkers to the engine's canonical form. */
int mqe_prepare(mqe_engine *e, double *E_start);

/* Replaces MPSQMC2::PropagateSeparately (MPSQMC2.cpp:518-611) and computes the
 * per-walker local energies of the propagated walkers (Walker.cpp:72-77).
 * uniforms: n_walkers*n_couplings draws, walker-major, or NULL for the device
 * RNG: the k-th call with NULL draws the numbers of step k of mqe_run (Philox
 * keyed by seed, step word k, population position and coupling), so a walk of
 * mqe_propagate / mqe_energy / mqe_population_control calls with device draws
 * repeats mqe_run(1, ...) exactly.  *sum_w receives the sum of the weights. */
int mqe_propagate(mqe_engine *e, const double *uniforms, double *sum_w);

/* Replaces MPSQMC2::EnergyFunctionAndHistory (MPSQMC2.cpp:688-747): E_P and
 * Omega from the local energies of the last mqe_propagate. */
int mqe_energy(mqe_engine *e, int step, int do_fluct, double *E_P, double *omega);

/* Replaces MPSQMC2::SeparatePopulationControl (MPSQMC2.cpp:613-686).  The
 * k-th walker (in population order) whose scaled weight lies outside
 * [0.25, 1.5] consumes uniforms[k]; uniforms may be NULL for the device RNG
 * (the k-th such call draws the numbers of step k of mqe_run).  *n_new receives
 * the new population size, *n_used the draws consumed.  MQE_ECAPACITY: the
 * copies would exceed the capacity; the population is left unchanged and the
 * engine refuses further steps (the reference writes past its arrays,
 * MPSQMC2.cpp:650-662). */
int mqe_population_control(mqe_engine *e, double scaling, const double *uniforms, int n_uniforms,
                           int *n_new, int *n_used);

/* One full Walk step (MPSQMC2.cpp:465-511) on the device with the device RNG:
 * propagate, scaling = n_desired / sum_w, E_T, energy with history, population
 * control.  No host synchronisation; the per-step record (step, N, E_P, E_T,
 * Omega, sum_w) is kept on the device and read with mqe_read_trace. */
int mqe_step(mqe_engine *e, int step);

/* Run steps first_step .. first_step+n_steps-1 with mqe_step.  use_graph != 0:
 * the n_steps steps are captured once in a HIP graph (kept until a run with
 * another n_steps) and replayed; step numbers live in a device control block
 * that the graph advances, so a replay gives exactly the eager result.  Graph
 * replay needs the reference proposal without sample collection (MQE_EINVAL
 * otherwise) and is not timed by mqe_set_timing. */
int mqe_run(mqe_engine *e, int first_step, int n_steps, int use_graph);

/* mqe_run(e, first_step, n_steps, 1): the steps replayed from a captured HIP graph. */
int mqe_step_graph(mqe_engine *e, int first_step, int n_steps);

/* Copy the trace rows of steps first_step .. first_step+n_steps-1 (6 doubles each). */
int mqe_read_trace(mqe_engine *e, int first_step, int n_steps, double *out);

/* MPSstate::CompressState(targets[i]) (MPSstate.cpp:379-404) for every walker i
 * of the population on the device, followed by Walker::setOverlap
 * (Walker.cpp:54-66), as SetupWalkers does for DW < DT (MPSQMC2.cpp:402-409):
 * a left-to-right and a right-to-left sweep of FP64 two-site SVDs (one-sided
 * Jacobi) keeping min(targets[i], number of Schmidt values >= 1e-13) states per
 * bond (TwoSiteObject.cpp:16,119-184), the state normalised, its phase chosen so
 * that the overlap with the trial is positive, and that overlap stored.
 * targets[i] < 1 leaves walker i (and its overlap) unchanged.  Call
 * mqe_prepare afterwards. */
int mqe_compress_walkers(mqe_engine *e, const int *targets);

/* ---- tile shapes of the general path ----------------------------------------
 * No counterpart in the reference (its contractions are BLAS calls on one
 * walker at a time).  The general path's FP32 products run on the matrix cores
 * with one of three output tile shapes per wave: 16 x 16 (v_mfma_f32_16x16x4),
 * 32 x 32 or 32 x 64 (two 32 x 32 tiles sharing their A panel).  The shape is
 * chosen per walker by its bond bucket (largest bond 1-16, 17-32, ..., > 112)
 * from an 8-entry table: tiles[bucket] in {16, 32, 64}.  mqe_create takes the
 * table tuned for the engine's key (L, couplings, trial and walker bonds,
 * precision) from the cache file ($MQE_TUNE_CACHE, else
 * ~/.cache/mqe_tiles.txt) when it has one.  Results do not depend on the
 * table beyond FP32 rounding; FP64 engines and the tile path ignore it. */
int mqe_get_tiles(mqe_engine *e, int *tiles);
int mqe_set_tiles(mqe_engine *e, const int *tiles);
/* Time `reps` propagations of the current population with each candidate
 * shape of each bond bucket present (coordinate descent, largest bonds first;
 * the population is restored after every trial), keep the fastest table,
 * append it to the cache file and return it; *ms_default / *ms_tuned: walker
 * kernel time of the starting and of the chosen table (either may be NULL).
 * Needs mqe_prepare; a no-op (returning the table) on FP64 engines and the
 * tile path. */
int mqe_autotune(mqe_engine *e, int reps, int *tiles_out, double *ms_default, double *ms_tuned);

/* 1 when the engine runs the FP32 MFMA tile path (nearest-neighbour chain,
 * all bonds <= 32, h = 0), 0 for the general path. */
int mqe_kernel_path(mqe_engine *e);

/* Re-orthonormalise every walker (FP64 Householder LQ sweep) once in `steps`
 * propagations (default 64), staggered over the walker slots so that every
 * step does the same share; in between the energy sweep carries right norm
 * environments instead, which leaves every computed quantity unchanged and
 * changes only the gauge of the stored tensors.  1 = every step. */
int mqe_set_canonical_interval(mqe_engine *e, int steps);

int mqe_sync(mqe_engine *e);
int mqe_num_walkers(mqe_engine *e, int *n);
int mqe_num_couplings(mqe_engine *e, int *n);

/* Walker scalars in population order (any pointer may be NULL). */
int mqe_read_scalars(mqe_engine *e, double *w, double *O, double *eh, double *EL);

/* Sampled grid point of every (walker, coupling) of the last propagation,
 * walker-major; idx holds (its population, see mqe_last_kernel_ms) x nC ints. */
int mqe_read_indices(mqe_engine *e, int *idx);

/* Walker i (population order): bond dims and tensors (gauge is the engine's,
 * the state is the same as the reference's up to gauge). size = flat length. */
int mqe_read_walker(mqe_engine *e, int i, int *vd, double *data, long *size);

/* out[i] = <other | walker_i> by transfer matrices (MPSstate.cpp:248-284),
 * batched over the population; other may have any bond dimensions. */
int mqe_overlaps(mqe_engine *e, const mqe_mps *other, double *out);

/* mqe_overlaps without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its workspace (the
 * other MPS and one transfer-matrix workspace per walker of a chunk) is a
 * fixed 64 MiB device array per GPU, held until the process ends; walkers run
 * in chunks of as many workspaces as fit.  MQE_ECAPACITY when the other MPS and
 * one walker's workspace (4 Dmax^2 elements, Dmax the largest bond of the
 * walkers and of `other`) exceed it.  Every call on one GPU uses the same
 * array: calls that overlap on one GPU (two host threads, two engines) race. */
int mqe_overlaps_noalloc(mqe_engine *e, const mqe_mps *other, double *out);

/* Device time of the walker kernel of the last mqe_propagate in milliseconds
 * (HIP events on the engine's stream), and the population of the last
 * propagation (mqe_propagate or a step of mqe_run). */
int mqe_last_kernel_ms(mqe_engine *e, double *ms, int *walkers);

/* Per-launch timing of the walker kernel inside mqe_step / mqe_run: enable
 * (resets the record), then read the summed device time (HIP events on the
 * engine's stream) and the number of timed launches. */
int mqe_set_timing(mqe_engine *e, int on);
int mqe_kernel_time(mqe_engine *e, double *total_ms, int *launches);

/* Snapshot of the walker population (slots, tensors and environments as
 * stored, weights, overlaps, energy histories, local energies, bond energies
 * and recorded labels of the learned proposal, the propagation and device-draw
 * counters) and a caller step number, written to a file; mqe_load_state
 * restores it into an engine created with the same parameters and trial.  A
 * restored engine continues the walk exactly: the device random numbers are
 * keyed by seed, step and population position.  (The reference has no resume;
 * MPSQMC2.cpp:749-758 only writes the energy trace.) */
int mqe_save_state(mqe_engine *e, const char *path, long step);
int mqe_load_state(mqe_engine *e, const char *path, long *step);

const char *mqe_last_error(void);

/* ---- learned proposal -------------------------------------------------------
 * The reference has no learned part (its proposal is the exact PDF,
 * MPSQMC2.cpp:552-588); these calls have no counterpart there.
 *
 * Features of walker b at coupling c (MQE_NFEAT floats, from the start of the
 * step): log O, log w, E_L, e_{c-1}, e_c, e_{c+1}, c/(nC-1), with e_k the
 * local energy of bond k from the previous energy sweep (0 beyond the ends),
 * then the 16 labels of the same coupling in the walker's previous
 * propagation with collection or the learned proposal on (0 before that).
 * Labels: the 16 normalised two-site elements combos[a+4b]/O of the coupling
 * (MPSQMC2.cpp:540-549), so that ratio(x) = sum_ab label[a+4b] x_a x_b.
 *
 * Learned proposal: with predicted labels p, q(x) = (1-mix) qhat(x)/sum qhat +
 * mix/N_grid, qhat(x) = max(0, sum_ab p[a+4b] x_a x_b) (uniform when all qhat
 * are 0); x is drawn from q by the reference's CDF scan and the weight is
 * multiplied by pdf(x)/q(x), pdf(x) = max(0, ratio(x)) 4/N_grid (the
 * reference's unnormalised PDF), so E[weight] is the reference's w * PDFnorm
 * (unbiased).  If pdf(x) = 0 the weight becomes 0 and the walker continues
 * with the reference's own sample.  Tile path only (nearest-neighbour chains,
 * bonds <= 32, FP32). */
#define MQE_NFEAT 23
#define MQE_NLABEL 16

typedef struct mqe_mlp mqe_mlp;

/* MLP n_in -> hidden -> hidden -> n_out with ReLU, FP32, He-normal weights
 * drawn from std::mt19937_64(seed).  n_in <= 32, hidden <= 128, n_out <= 32. */
int mqe_mlp_create(int n_in, int hidden, int n_out, unsigned long long seed, mqe_mlp **out);
void mqe_mlp_destroy(mqe_mlp *m);
long mqe_mlp_num_params(const mqe_mlp *m);
/* parameters: W1 [hidden x n_in], b1, W2 [hidden x hidden], b2, W3 [n_out x hidden], b3 (column-major) */
int mqe_mlp_get_params(mqe_mlp *m, float *p);
int mqe_mlp_set_params(mqe_mlp *m, const float *p);   /* also resets the Adam state */
/* input/output standardisation: x' = (x - x_mean)/x_std, y = y' y_std + y_mean */
int mqe_mlp_get_scaling(mqe_mlp *m, float *x_mean, float *x_std, float *y_mean, float *y_std);
int mqe_mlp_set_scaling(mqe_mlp *m, const float *x_mean, const float *x_std, const float *y_mean,
                        const float *y_std);
/* Y[n x n_out] = net(X[n x n_in]) (row-major host arrays) */
int mqe_mlp_forward(mqe_mlp *m, int n, const float *X, float *Y);
/* mean squared error of the standardised outputs over n rows and its gradient
 * with respect to the parameters (num_params floats) */
int mqe_mlp_loss_grad(mqe_mlp *m, int n, const float *X, const float *Y, double *loss, float *grad);
/* Adam (lr, 0.9, 0.999, 1e-8) over `epochs` seeded shuffles in minibatches of
 * `batch` rows; standardize != 0 first sets the scaling from the data.
 * *loss: the last epoch's mean loss. */
int mqe_mlp_fit(mqe_mlp *m, int n, const float *X, const float *Y, int epochs, int batch, double lr,
                int standardize, double *loss);

/* Collect (feature, label) rows during the following propagations, up to
 * max_samples rows (0 stops collecting and frees the buffer). */
int mqe_collect_samples(mqe_engine *e, long max_samples);
int mqe_num_samples(mqe_engine *e, long *n);
int mqe_read_samples(mqe_engine *e, long first, long n, float *X, float *Y);
/* mqe_mlp_fit on the collected rows, which stay on the device */
int mqe_train_proposal(mqe_engine *e, mqe_mlp *m, int epochs, int batch, double lr, double *loss);
/* Sample the following propagations from the network's proposal (m = NULL:
 * the reference's proposal).  The network must have n_in = MQE_NFEAT and
 * n_out = MQE_NLABEL and stay alive while it is in use. */
int mqe_set_proposal(mqe_engine *e, mqe_mlp *m, double mix);
/* Fixed predictions instead of a network: n_walkers * nC * 16 floats in walker
 * order (NULL: the reference's proposal); walkers beyond n_walkers get the
 * uniform proposal. */
int mqe_set_proposal_table(mqe_engine *e, const float *pred, int n_walkers, double mix);
/* Local energies of the bonds (couplings) from the last energy sweep, n * nC
 * doubles in walker order; they sum to each walker's E_L. */
int mqe_read_bond_energies(mqe_engine *e, double *out);

/* ---- set-up helpers (host code; no GPU needed) ----------------------------- */

/* Bond dimensions of MPSstate(L, D, d) (MPSstate.cpp:20-28). */
void mqe_bond_dims(int L, int D, int d, int *vd);

/* Operator-Schmidt tables of exp(-dtau J S_i.S_j) in the reference's storage
 * (TrotterHeisenberg.cpp:102-285).  d = 2 uses the closed form in the
 * Pauli basis {I, Sx, iSy, Sz}; returns MQE_EINVAL for other d. */
int mqe_trotter_tables(int d, double J, double dtau, double *S, double *U, double *VT);

/* Marsaglia grid on the unit sphere in dim dimensions with p points per
 * coordinate (GridGenerator.cpp:167-225).  grid must hold dim*p^dim doubles;
 * *n receives the number of points. */
int mqe_marsaglia_grid(int dim, int p, double *grid, int *n);

/* MPSstate::LeftNormalize (MPSstate.cpp:216-229) on a flat MPS, in place;
 * *norm receives the final 1x1 R factor. */
int mqe_left_normalize(int L, int d, const int *vd, double *data, double *norm);

/* MPSstate::RightNormalize (MPSstate.cpp:231-246), in place: LQ per site from
 * the right end, Lmat pushed into the left neighbour; *norm receives the final
 * 1x1 factor. */
int mqe_right_normalize(int L, int d, const int *vd, double *data, double *norm);

/* MPSstate::ApplyMPO(HeisenbergMPO, Psi0) (MPSstate.cpp:413-471): H|psi> for
 * H = sum_c J[c] S_first[c].S_second[c] - field sum_i S^z_i as an MPS of bond
 * dims vd_out (bond k: channel w and state index r combined as w*vd[k] + r).
 * The channels are those of a minimal MPO (2 + 3 x the sites whose couplings
 * cross the bond), fewer than the reference's HeisenbergMPO; the vector is the
 * same.  out may be NULL to query vd_out and *size (the flat length). d = 2. */
int mqe_apply_mpo(int L, int d, int n_couplings, const int *first, const int *second, const double *J,
                  double field, const int *vd, const double *data, int *vd_out, double *out, long *size);

/* MPSstate::CompressState(Dt) (MPSstate.cpp:379-404): two SVD sweeps with the
 * 1e-13 absolute Schmidt cutoff.  vd_out receives the new bond dims and out
 * the tensors (capacity: flat size of the input). */
int mqe_compress(int L, int d, const int *vd, const double *data, int Dt, int *vd_out, double *out);

/* Trial wave function by two-site DMRG (DMRG.cpp:50-100 as called from
 * MPSQMC2::SetupTrial, MPSQMC2.cpp:225-236): ground state of
 * H = sum_c J[c] S_first[c].S_second[c] - field sum_i S^z_i in the bond profile
 * mqe_bond_dims(L, D, d), from a random start drawn from std::mt19937(seed),
 * sweeping until the energy changes by less than tol (or max_sweeps).  The
 * result is left-normalized.  vd_out: L+1 ints; data_out: the flat size of
 * that profile; *energy: the variational energy.  d must be 2. */
int mqe_dmrg(int L, int d, int n_couplings, const int *first, const int *second, const double *J, double field,
             int D, int max_sweeps, double tol, unsigned int seed, int *vd_out, double *data_out, double *energy);

/* ---- host random stream (the build's RNG contract) ------------------------- */
/* Random::rand (Random.cpp:49-59) with a fixed seed: one std::mt19937 stream,
 * u = x * 2^-32.  peek fills out without consuming; discard consumes k draws. */
typedef struct mqe_rng mqe_rng;
int mqe_rng_create(unsigned int seed, mqe_rng **out);
void mqe_rng_destroy(mqe_rng *r);
int mqe_rng_uniforms(mqe_rng *r, long n, double *out);
int mqe_rng_peek(mqe_rng *r, long n, double *out);
int mqe_rng_discard(mqe_rng *r, long k);

#ifdef __cplusplus
}
#endif

#endif /* MQE_H */
