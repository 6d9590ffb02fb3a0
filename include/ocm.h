/*
 * ocm.h -- C ABI of the MI355X MD step for OpenCafeMol force fields.
 *
 * This ABI replaces the OpenMM objects the reference builds and steps:
 *   - OpenMM::System / OpenMM::Force, produced by SystemGenerator::generate
 *     (src/SystemGenerator.cpp:34-111) and ForceFieldGeneratorBase::generate
 *     (src/forcefield/ForceFieldGeneratorBase.hpp:9-16), become the plain
 *     descriptors ocm_system_desc / ocm_force_desc below;
 *   - OpenMM::LangevinIntegrator, produced by LangevinIntegratorGenerator
 *     (src/IntegratorGenerator.cpp:5-14), becomes ocm_integrator_desc;
 *   - OpenMM::Context (src/Simulator.cpp:17), its setPositions (:41),
 *     setVelocities[ToTemperature] (src/Simulator.hpp:49-65), Integrator::step
 *     (src/Simulator.cpp:60) and getState (src/Observer.cpp:34,107,113,339-344)
 *     become ocm_context_* below.
 *
 * Units are OpenMM's: nm, ps, amu, kJ/mol, elementary charge, K.
 * Every function returns an int status (OCM_OK on success); no C++ exception
 * crosses this boundary.  ocm_last_error() returns a thread-local message for
 * the last failing call on the calling thread.
 *
 * Ownership: descriptor arrays are copied at create time; the caller may free
 * them afterwards.  Output arrays are caller-allocated.  One context may be
 * driven by one host thread at a time.
 */
#ifndef OCM_H
#define OCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define OCM_OK               0
#define OCM_ERR_INVALID      1  /* malformed descriptor or argument            */
#define OCM_ERR_OOM          2  /* device allocation failed / HBM budget hit   */
#define OCM_ERR_HIP          3  /* HIP runtime error                           */
#define OCM_ERR_NUMERIC      4  /* NaN or fixed-point force overflow detected  */
#define OCM_ERR_NO_DEVICE    5  /* no usable gfx950 device                     */
#define OCM_ERR_STATE        6  /* call not valid in the current state         */
#define OCM_ERR_TIMEOUT      7  /* nothing ready within the given time          */

/* ----------------------------------------------------------- force kinds */
/* Bonded kinds: `indices` holds n_terms * arity bead indices and `params`
 * holds n_terms * nparams doubles, in the order given.                     */
enum ocm_force_kind {
    /* E = 1/2 k (r - r0)^2 ; params {r0, k}
     * HarmonicBondForceFieldGenerator.cpp:3-14 (OpenMM HarmonicBondForce). */
    OCM_HARMONIC_BOND = 1,
    /* E = k exp(-(r - v0)^2 / (2 sigma^2)) ; params {k, v0, sigma}
     * GaussianBondForceFieldGenerator.cpp:5-20.                             */
    OCM_GAUSSIAN_BOND = 2,
    /* E = k (5 (r0/r)^12 - 6 (r0/r)^10) ; params {k, r0}
     * GoContactForceFieldGenerator.cpp:5-17.                                */
    OCM_GO_CONTACT = 3,
    /* E = 1/2 k (theta - theta0)^2, theta at the middle bead ;
     * params {theta0, k}.  HarmonicAngleForceFieldGenerator.cpp:3-15.      */
    OCM_HARMONIC_ANGLE = 4,
    /* E = k (spline(theta) - Emin + step(tmin-theta)(30(tmin-theta)+y0)
     *        + step(theta-tmax)(30(theta-tmax)+y9)) ; params {k} ;
     * globals {y0..y9, tmin, tmax, Emin}; spline is the natural cubic
     * spline through (tmin + i (tmax-tmin)/9, y_i), zero outside [tmin,tmax].
     * FlexibleLocalAngleForceFieldGenerator.cpp:3-45.                       */
    OCM_FLEXIBLE_LOCAL_ANGLE = 5,
    /* E = k exp(-d^2/(2 sigma^2)), d = (phi - phi0) wrapped to [-pi, pi) ;
     * params {k, phi0, sigma}.  GaussianDihedralForceFieldGenerator.cpp:3-27 */
    OCM_GAUSSIAN_DIHEDRAL = 6,
    /* E = k (c - Emin + sum_n a_n cos(n phi) + b_n sin(n phi)), n = 1..3 ;
     * params {k} ; globals {c, a1, b1, a2, b2, a3, b3, Emin}.
     * FlexibleLocalDihedralForceFieldGenerator.cpp:5-63.                    */
    OCM_FLEXIBLE_LOCAL_DIHEDRAL = 7,
    /* Go 12-10 contact with a linear cap: with c = capping_ratio and
     * x = r / r0, E = k (a x + b) for x <= c, else k (5 x^-12 - 6 x^-10),
     * a = -60 (c^-13 - c^-11), b = 65 c^-12 - 66 c^-10 (value and slope
     * continuous at x = c) ; params {k, r0, capping_ratio}.
     * CappedGoContactForceFieldGenerator.cpp:3-30.                         */
    OCM_CAPPED_GO_CONTACT = 8,
    /* E = k (1 - cos(n (phi - t0))), phi the torsion angle ; params {k, t0, n}.
     * CosineDihedralForceFieldGenerator.cpp:5-43 (its TOML reader adds pi
     * to v0, ReadTOMLForceFieldGenerator.cpp:600-611).                       */
    OCM_COSINE_DIHEDRAL = 9,

    /* Nonbonded kinds: `particle_params` holds n_particles * npp doubles.
     * Energy is zero at r >= cutoff (OpenMM CutoffNonPeriodic/Periodic).   */
    /* E = eps (s_i + s_j)^12 (r^-12 - rc^-12) ; npp 1 {radius} ;
     * globals {eps}.  ExcludedVolumeForceFieldGenerator.cpp:23-92.          */
    OCM_EXCLUDED_VOLUME = 16,
    /* E = q_i q_j K (exp(-r/lambda)/r - corr) ; npp 1 {charge} ;
     * globals {K = 1/(4 pi eps0 epsk), lambda, corr}.
     * DebyeHuckelForceFieldGenerator.cpp:10-108.                            */
    OCM_DEBYE_HUCKEL = 17,
    /* E = step(r*-r) e (4 (s^12 - s^6) + 1), s = sig/r, r* = 2^(1/6) sig,
     * sig = (sig_i+sig_j)/2, e = (e_i+e_j)/2 ; npp 2 {epsilon, sigma}.
     * LennardJonesRepulsiveForceFieldGenerator.cpp:34-112.                  */
    OCM_LJ_REPULSIVE = 18,
    /* E = e (step(r-r*) 4 (s^12 - s^6) - step(r*-r)) * S(r), S the OpenMM
     * switching function from switch_distance to cutoff ; npp 2 {epsilon,
     * sigma}.  LennardJonesAttractiveForceFieldGenerator.cpp:38-121.        */
    OCM_LJ_ATTRACTIVE = 19,
    /* Weeks-Chandler-Andersen: E = step(2^(1/6) sig - r) (4 e (s^12 - s^6) + e),
     * s = sig/r, sig = (sig_i+sig_j)/2, e = sqrt(e_i e_j) ; npp 2 {sigma,
     * epsilon}.  WeeksChandlerAndersenForceFieldGenerator.cpp:32-114.       */
    OCM_WCA = 20,
    /* WCA with one sigma and epsilon for every pair of its categories:
     * E = 4 e (s^12 - s^6 + 1/4), cutoff 2^(1/6) sig ; npp 0 ;
     * globals {sigma, epsilon}.
     * UniformWeeksChandlerAndersenForceFieldGenerator.cpp:160-185.          */
    OCM_UNIFORM_WCA = 21,
    /* E = e (step(r - r*) 4 (s^12 - s^6) - step(r* - r)) S(r), r* = 2^(1/6)
     * sig, one sigma and epsilon, S the switch from switch_distance ; npp 0 ;
     * globals {sigma, epsilon}.
     * UniformLennardJonesAttractiveForceFieldGenerator.cpp:143-194.         */
    OCM_UNIFORM_LJ_ATTRACTIVE = 22,
    /* iSoLF attraction: b = 2^(1/6) sig, E = -e (step(b - r) + step(r - b)
     * step(b + w - r) cos^2(pi (r - b) / (2 w))), sig and w the arithmetic
     * means, e = sqrt(e_i e_j) ; npp 3 {sigma, epsilon, omega}.
     * iSoLFAttractiveForceFieldGenerator.cpp:41-145.                       */
    OCM_ISOLF_ATTRACTIVE = 23
};

/* One force (one reference ForceFieldGenerator::generate() result). */
typedef struct ocm_force_desc {
    int32_t        kind;            /* enum ocm_force_kind                    */
    int32_t        group;           /* force group id, 0..31                  */
    int32_t        use_periodic;    /* minimum-image displacements            */
    int32_t        n_terms;         /* bonded: number of terms                */
    const int32_t* indices;         /* bonded: n_terms * arity                */
    const double*  params;          /* bonded: n_terms * nparams              */
    int32_t        n_globals;       /* count of `globals`                     */
    int32_t        n_categories;    /* nonbonded: categories in use (>=1) or 0 */
    const double*  globals;         /* kind-specific force-level constants    */
    /* ---- nonbonded only ---- */
    const double*  particle_params; /* n_particles * npp                      */
    const int32_t* category;        /* n_particles; -1 = not a participant;
                                       NULL = every bead is category 0        */
    const uint8_t* category_allow;  /* n_categories^2 symmetric 0/1 matrix;
                                       NULL = every category pair interacts   */
    double         cutoff;          /* nm                                     */
    double         switch_distance; /* nm; <= 0 disables switching            */
    int64_t        n_exclusions;
    const int32_t* exclusions;      /* n_exclusions * 2 bead indices          */
} ocm_force_desc;

typedef struct ocm_system_desc {
    int32_t               n_particles;
    int32_t               n_forces;
    const double*         masses;       /* amu; 0 = immobile bead            */
    int32_t               use_periodic_box;
    int32_t               reserved;
    double                box[3];       /* nm, cuboid with a corner at origin */
    const ocm_force_desc* forces;
} ocm_system_desc;

/* Langevin thermostat (LangevinIntegratorGenerator, IntegratorGenerator.hpp:19-42).
 * Integrated with the BAOAB splitting; noise from counter-based
 * Philox4x32-10 keyed by (seed, bead index) and counted by step.          */
typedef struct ocm_integrator_desc {
    double   temperature;  /* K     */
    double   friction;     /* 1/ps  */
    double   dt;           /* ps    */
    uint64_t seed;
} ocm_integrator_desc;

typedef struct ocm_options {
    int32_t device;            /* HIP device ordinal                         */
    int32_t rebuild_interval;  /* steps between neighbor-list rebuilds (>=1) */
    double  skin;              /* nm, neighbor-list buffer beyond cutoff     */
    int32_t team_size;         /* lanes per pair-list row: 4,8,16,32,64; 0 = auto */
    int32_t use_graph;         /* capture step chunks as hipGraphs           */
    int32_t chunk_steps;       /* steps between host checks of device flags;
                                  over 2^17 pair-group members at most
                                  max(2 * rebuild_interval, 32), since a
                                  stale list redoes the whole chunk         */
    int32_t prune_interval;    /* steps between prunes of the inner list
                                  (cutoff + 0.1 nm, capped by the skin);
                                  <= 1: every pair pass reads the full list,
                                  as does a batch whose list is estimated
                                  under 16,384 entries                      */
    int32_t resident;          /* batches whose every system has <= 1024
                                  beads (<= 8 nonbonded lists, <= 64
                                  parameter sets): one workgroup per system
                                  steps whole chunks on chip, with the same
                                  bits as the standard path.  -1 (default):
                                  when the batch qualifies; 0: never;
                                  1: the same as -1                         */
    int32_t entries_per_lane;  /* with team_size set: list entries each lane
                                  of the pair kernel takes per round, 4
                                  (0 = 4), or 2 with a team of 8, 16 or 32;
                                  ignored when the shape is autotuned       */
} ocm_options;

/* getState selection bits */
#define OCM_STATE_POSITIONS   1
#define OCM_STATE_VELOCITIES  2
#define OCM_STATE_FORCES      4
#define OCM_STATE_ENERGY      8

typedef struct ocm_state {
    double* positions;      /* n*3 or NULL                                     */
    double* velocities;     /* n*3 or NULL                                     */
    double* forces;         /* n*3 or NULL (kJ/mol/nm)                         */
    double  energy_by_group[32]; /* kJ/mol, filled when ENERGY requested       */
    double  potential_energy;    /* sum over the requested groups              */
    double  kinetic_energy;
    int64_t step;           /* steps taken since creation                      */
} ocm_state;

typedef struct ocm_context ocm_context;

/* Fills `opt` with defaults (skin 0.7 nm, rebuild every 40 steps, inner
 * list pruned every 5 steps, ...).
 * Skin and interval change speed only: forces are fixed-point sums over the
 * pairs inside the cutoff, so any valid list gives the same bits.  When a bead
 * moves more than skin/2 between rebuilds the chunk is redone with half the
 * interval; the interval doubles back after a run of clean chunks.          */
void ocm_options_default(ocm_options* opt);

/* OpenMM::Context(system, integrator, platform), src/Simulator.cpp:17.
 * Uploads the system to the device; positions start at the origin until
 * ocm_set_positions is called.                                             */
int ocm_context_create(const ocm_system_desc* system,
                       const ocm_integrator_desc* integrator,
                       const ocm_options* options, ocm_context** out);
int ocm_context_destroy(ocm_context* ctx);

/* Context::setPositions, src/Simulator.cpp:41. xyz is n*3 nm. */
int ocm_set_positions(ocm_context* ctx, const double* xyz, int32_t n);
/* Context::setVelocities, src/Simulator.hpp:62-65. xyz is n*3 nm/ps. */
int ocm_set_velocities(ocm_context* ctx, const double* xyz, int32_t n);
/* Context::setVelocitiesToTemperature, src/Simulator.hpp:49-60: Maxwell-
 * Boltzmann velocities from Philox stream 1 keyed by (seed, bead).       */
int ocm_set_velocities_to_temperature(ocm_context* ctx, double temperature,
                                      uint64_t seed);
/* Integrator::step(n), src/Simulator.cpp:60.
 * Steps run in chunks (chunk_steps) that are captured as hipGraphs.  The
 * neighbor-list ages carry over from call to call: the list is rebuilt every
 * rebuild_interval steps of the trajectory, not at the start of each call. */
int ocm_step(ocm_context* ctx, int64_t nsteps);
/* Builds the neighbor lists if needed and captures the step graphs that
 * ocm_step(ctx, nsteps) will launch from the current state, without stepping,
 * so that call spends no time on capture or instantiation (a rollback to a
 * shorter list interval may still capture during it).  No counterpart in the
 * reference, where OpenMM compiles its kernels when the Context is created
 * (src/Simulator.cpp:17).                                                  */
int ocm_prepare_steps(ocm_context* ctx, int64_t nsteps);
/* Context::getState(types, false, groups), src/Observer.cpp:336-344.
 * Forces and energies are restricted to the groups in `groups`.          */
int ocm_get_state(ocm_context* ctx, int32_t what, uint32_t groups,
                  ocm_state* out);

/* LocalEnergyMinimizer::minimize(context, tolerance, max_iterations),
 * src/Simulator.cpp:42-46.  L-BFGS (10 correction pairs) over the mobile
 * beads, with device forces and FP64 device energies, a backtracking line
 * search, and no bead moving more than 0.05 nm per iteration (so starts with
 * overlapping beads relax safely).  Stops when the RMS force component of
 * every system is <= tolerance (kJ/mol/nm), as OpenMM's does, or after
 * max_iterations (0 = 100000).  OpenMM's implementation differs in its
 * line search and history, so the point reached is a nearby minimum, not the
 * same bits.                                                                */
int ocm_minimize(ocm_context* ctx, double tolerance, int32_t max_iterations);

/* Neighbor pairs (bead index pairs i<j, r < cutoff) that the device list of
 * nonbonded force `force_index` interacts through, at the current positions.
 * The list is rebuilt there (its rows widened if they no longer fit).  Writes
 * the first `cap` pairs in (i, j) order to out_pairs (2 ints each); *count
 * gets the total.  No counterpart in the reference, whose lists live inside
 * OpenMM; the diagnostic that pins the list against brute force.           */
int ocm_get_neighbor_pairs(ocm_context* ctx, int32_t force_index,
                           int32_t* out_pairs, int64_t cap, int64_t* count);

/* ocm_get_neighbor_pairs without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The pairs pass through a fixed device scratch array of
 * OCM_NOALLOC_MAX_PAIRS pairs that each GPU holds until the process ends.
 * Differences from ocm_get_neighbor_pairs:
 *  - rows are not widened: if a row no longer fits, OCM_ERR_OOM (a step or
 *    ocm_get_neighbor_pairs widens them);
 *  - cap = 0 only counts; otherwise cap must be >= the pair count, and more
 *    than OCM_NOALLOC_MAX_PAIRS pairs is an error (OCM_ERR_INVALID, with
 *    *count set in both cases);
 *  - every call on one GPU shares the scratch array: calls that overlap on
 *    one GPU (from several threads) race.                                   */
#define OCM_NOALLOC_MAX_PAIRS (1 << 21)
int ocm_get_neighbor_pairs_noalloc(ocm_context* ctx, int32_t force_index,
                                   int32_t* out_pairs, int64_t cap, int64_t* count);

/* Per-kernel timing.  enable = 1: HIP events on the context's stream around
 * every launch (steps run eagerly), read with kernel 0 = pair forces,
 * 1 = bonded forces, 2 = integrator, 3 = list build, 4 = record pack,
 * 6 = inner-list prune.
 * enable = 2: steps keep running as captured graphs and every pair-force
 * launch of a step records its first block start and last block end on the
 * device wall clock; read with kernel 5.  enable = 0: off.                  */
int ocm_set_timing(ocm_context* ctx, int32_t enable);
int ocm_get_timing(ocm_context* ctx, int32_t kernel, double* total_ms,
                   int64_t* launches);

/* Work figures of the current neighbor lists: listed (directed) entries and
 * member beads summed over all pair groups.                                */
int ocm_get_list_stats(ocm_context* ctx, int64_t* list_entries,
                       int64_t* members, int64_t* rebuilds);
/* Inner (pruned) list: padded entries now held, and the prune interval in
 * use (lowered after a bead moved more than margin/2 between prunes, raised
 * back after clean chunks; 0 when the inner list is off).  No counterpart in
 * the reference, whose list lives inside OpenMM; a diagnostic.            */
int ocm_get_inner_list_stats(ocm_context* ctx, int64_t* inner_entries,
                             int32_t* prune_interval);
/* Exact entry counts of the current lists, summed over all pair groups:
 * counts[0] outer entries holding a partner (every listed pair sits in both
 * of its rows, so unique listed pairs are counts[0] / 2), counts[1] outer
 * entries including the -1 padding, counts[2] and counts[3] the same for the
 * inner list (0 when it is off).  Copies the lists to the host; a diagnostic
 * with no counterpart in the reference.                                    */
int ocm_get_list_counts(ocm_context* ctx, int64_t* counts);
/* The pair kernel's autotuned launch shape: lanes per member row, list
 * entries per lane per round, workgroups of its fixed grid, and the prune's
 * lanes per row.  A diagnostic with no counterpart in the reference.       */
int ocm_get_launch_shape(ocm_context* ctx, int32_t* team, int32_t* entries_per_lane,
                         int32_t* grid, int32_t* prune_team);
/* Resident stepping: *threads receives the workgroup size each system steps
 * in (1024), or 0 when the batch steps through the standard path.  A
 * diagnostic with no counterpart in the reference.                        */
int ocm_get_resident(ocm_context* ctx, int32_t* threads);

/* ------------------------------------------------------------------ batch */
/* Many independent systems packed into one set of device arrays; every
 * stage is one launch over all of them.  A context is a batch of one.     */
int ocm_batch_create(const ocm_system_desc* systems,
                     const ocm_integrator_desc* integrators, int32_t n_systems,
                     const ocm_options* options, ocm_context** out);
int ocm_batch_set_positions(ocm_context* ctx, int32_t system,
                            const double* xyz, int32_t n);
int ocm_batch_set_velocities(ocm_context* ctx, int32_t system,
                             const double* xyz, int32_t n);
int ocm_batch_get_state(ocm_context* ctx, int32_t system, int32_t what,
                        uint32_t groups, ocm_state* out);
/* Per-system error status after the last ocm_step (OCM_OK or OCM_ERR_NUMERIC). */
int ocm_batch_system_status(ocm_context* ctx, int32_t system);
int32_t ocm_batch_size(ocm_context* ctx);

/* Checkpoint and resume.  A checkpoint holds one system's positions,
 * velocities and steps taken; with the integrator's seed these fix the
 * Philox noise of every later step, so a run resumed from a checkpoint gives
 * the same bits as one that never stopped.  The file records the seed and
 * bead count, and reading it into a system with another seed or size is
 * OCM_ERR_INVALID.  No counterpart in the reference, whose runs cannot be
 * resumed; the state is what its Context holds between Integrator::step
 * calls (src/Simulator.cpp:49-76).                                          */
int ocm_batch_write_checkpoint(ocm_context* ctx, int32_t system, const char* path);
int ocm_batch_read_checkpoint(ocm_context* ctx, int32_t system, const char* path);
int ocm_write_checkpoint(ocm_context* ctx, const char* path);   /* system 0 */
int ocm_read_checkpoint(ocm_context* ctx, const char* path);

/* ----------------------------------------------------------------- server */
/* A request loop over batches: jobs (one system each) are admitted as job
 * slots and the device-memory budget allow, packed into one batch with the
 * jobs already running and advanced together; the batch is repacked whenever
 * a job enters or leaves.  Frames are copied out asynchronously and queued.
 * A job's trajectory and result are the same bits as when it runs alone.
 * No counterpart in the reference, which runs one system per process
 * (src/OpenAICG2Plus.cpp:12-93); each job is that run's MD loop
 * (src/Simulator.cpp:49-76) without its file output.
 * Thread safety: submit, next_frame, wait and get_stats may be called from
 * any host threads; one worker thread owns the device.                      */
typedef struct ocm_server ocm_server;

typedef struct ocm_server_options {
    ocm_options engine;          /* options of each packed batch             */
    int64_t hbm_budget_bytes;    /* device bytes for admitted jobs, by estimate;
                                    0 = 80% of the free memory at creation    */
    int32_t max_active_jobs;     /* job slots; 0 = 256                        */
    int32_t max_queued_frames;   /* uncollected frames before the worker waits;
                                    0 = 4096.  Callers must drain frames.     */
    int64_t quantum_steps;       /* longest stretch between admission checks;
                                    0 = 2000                                  */
} ocm_server_options;

typedef struct ocm_server_stats {
    int64_t submitted, completed, failed, active, queued;
    int64_t repacks;             /* batches built                             */
    int64_t frames;              /* frames snapshotted                        */
    double  latency_p50_ms;      /* submit -> completion, completed jobs      */
    double  latency_p99_ms;
    double  wait_p50_ms;         /* submit -> admission                       */
    double  wait_p99_ms;
    double  simulated_ns;        /* sum over ended jobs of steps * dt         */
    double  busy_ms;             /* worker time spent stepping or repacking,
                                    summed over the two lane workers         */
    double  repack_ms;           /* of which repacking                        */
    double  makespan_ms;         /* first submission -> last completion       */
    int64_t device_bytes;        /* current batch allocation                  */
    int64_t budget_bytes;
} ocm_server_stats;

typedef struct ocm_job_info {
    int32_t status;              /* OCM_OK, OCM_ERR_NUMERIC, ...              */
    int32_t n_particles;
    int64_t steps_done;
    double  submit_ms, start_ms, end_ms;   /* since server creation          */
} ocm_job_info;

void ocm_server_options_default(ocm_server_options* opt);
int ocm_server_create(const ocm_server_options* options, ocm_server** out);
/* Stops the worker (running jobs are abandoned) and frees everything.      */
int ocm_server_destroy(ocm_server* srv);
/* Queues a job: n particles at `positions` (n*3 nm); `velocities` n*3 nm/ps,
 * or NULL for Maxwell-Boltzmann velocities at the integrator temperature
 * from its seed.  Runs nsteps steps and queues a frame (positions) every
 * frame_interval steps (0 = none).  OCM_ERR_OOM if the job alone exceeds
 * the budget.                                                              */
int ocm_server_submit(ocm_server* srv, const ocm_system_desc* system,
                      const ocm_integrator_desc* integrator,
                      const double* positions, const double* velocities,
                      int32_t n, int64_t nsteps, int64_t frame_interval,
                      int64_t* job_id);
/* Pops the oldest frame into xyz (cap particles).  timeout_ms < 0 waits
 * without limit.  OCM_ERR_TIMEOUT if none arrived; OCM_ERR_INVALID if cap is
 * too small (*n gets the size and the frame stays queued).                 */
int ocm_server_next_frame(ocm_server* srv, int32_t timeout_ms, int64_t* job,
                          int64_t* step, double* xyz, int32_t cap, int32_t* n);
/* Waits for a job to end and all of its frames to be queued, then copies its
 * final positions and velocities (either may be NULL; cap particles) and
 * forgets the job.  OCM_ERR_TIMEOUT if it is still running.               */
int ocm_server_wait(ocm_server* srv, int64_t job, int32_t timeout_ms,
                    double* positions, double* velocities, int32_t cap,
                    ocm_job_info* info);
int ocm_server_get_stats(ocm_server* srv, ocm_server_stats* out);

const char* ocm_last_error(void);
const char* ocm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OCM_H */
