/*
 * mcg.h -- C ABI of the MI355X phonon Monte Carlo trajectory engine.
 *
 * The engine replaces the per-bundle loop of the reference's field solve,
 * FieldProblem::solve (montecarlo/problem.cpp:370-445), and everything it
 * calls per scatter-step: free flight (Subdomain::advect, subdomain.cpp:161-192),
 * relaxation-time scattering (Material::scatter, material.cpp:226-231),
 * boundary reflection (the Boundary::scatter family, boundary.cpp:283-522)
 * and grid tallying (Field::accumulate, field.cpp:92-220).
 *
 * Conventions
 *  - Plain C: fixed-width integers, doubles, plain pointers and sizes.
 *  - Every entry point returns an int status (MCG_OK = 0); on failure the
 *    message is available from mcg_last_error() (thread-local).  No C++
 *    exception crosses this boundary and nothing aborts on bad input.
 *  - The caller owns every host buffer passed in or out.  A scene owns the
 *    device memory it uploads.  One host thread drives a scene at a time.
 *  - Matrices are 3x3, column-major (the reference stores them as Eigen
 *    Matrix3d, which is column-major): m[0..2] is column 0.
 *  - Lengths in metres, times in seconds, temperatures in kelvin.  Wall and
 *    volume temperatures are deviational (relative to the material's T).
 */
#ifndef MCG_H
#define MCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCG_ABI_VERSION 2

/* ---------------------------------------------------------------- status */
enum mcg_status {
    MCG_OK = 0,
    MCG_EINVAL = 1,    /* invalid argument or inconsistent scene          */
    MCG_ENOMEM = 2,    /* host or device allocation failed                */
    MCG_EDEVICE = 3,   /* HIP runtime error or no usable GPU              */
    MCG_ERANGE = 4,    /* output buffer too small / size beyond a limit   */
    MCG_EAGAIN = 5,    /* serving: result not ready yet                   */
    MCG_ENOTFOUND = 6  /* serving: unknown request id                     */
};

/* ------------------------------------------------------------- material
 * Inputs of Material::Material(disp, relax, T) (material.cpp:82-162),
 * already parsed from the reference's two text formats:
 *  disp:  nw rows x (2 + 2*np) columns, row-major: omega, domega, then for
 *         each branch p the pair (vel_p, dos_p)        (material.cpp:90-107)
 *  relax: np rows x 8 columns, row-major: two mechanisms x (c0, c1, c2, c3),
 *         1/tau = sum_j c0 * omega^c1 * T^c2 * exp(-c3/T), a mechanism with
 *         c0 == 0 is skipped                            (material.cpp:113-134)
 */
typedef struct mcg_material_spec {
    int64_t nw;
    int64_t np;
    const double* disp;
    const double* relax;
    double temp;
} mcg_material_spec;

/* Tables Material::Material derives (material.cpp:116-161).  Arrays are
 * nw x np row-major and owned by the caller (any pointer may be NULL). */
typedef struct mcg_material_tables {
    double* tau;
    double* vel;
    double* energy_pdf;
    double* flux_pdf;
    double* scat_pdf;
    double energy_sum;   /* Material::energySum  (heat capacity C)      */
    double flux_sum;     /* Material::fluxSum    (sum of v*C)           */
    double scat_sum;     /* Material::scatSum                            */
    double cond;         /* Material::cond       (bulk k)                */
    int64_t nmodes;      /* cells with energy_pdf > 0 (the draw support) */
} mcg_material_tables;

/* ------------------------------------------------------------- geometry */
enum mcg_bdry_kind {
    MCG_SPEC = 0,      /* SpecBoundary     boundary.cpp:283-287             */
    MCG_DIFF = 1,      /* DiffBoundary     boundary.cpp:308-312             */
    MCG_ISOT = 2,      /* IsotBoundary<S>  boundary.cpp:455-460 (emits)     */
    MCG_PERI = 3,      /* PeriBoundary<S>  boundary.cpp:516-522 (emits)     */
    MCG_INTER = 4,     /* InterBoundary    boundary.cpp:349-359             */
    MCG_PARTIAL = 5    /* extension: specular with probability p, else
                          diffuse (the p = 1 / p = 0 ends are Spec / Diff)  */
};

enum mcg_cell_kind {
    MCG_PARALLELEPIPED = 0,   /* subdomain.h:128-198, faces Bac Lef Bot Fro Rig Top */
    MCG_TRIPRISM = 1,         /* subdomain.h:206-277, faces Bac Lef Bot Dia Top     */
    MCG_TETRAHEDRON = 2,      /* subdomain.h:285-352, faces Bac Lef Bot Dia         */
    MCG_PRISM = 3,            /* subdomain.h:364-492, faces Bot Top Sid0..Sid(N-1)  */
    MCG_PYRAMID = 4           /* subdomain.h:504-630, faces Bot Sid0..Sid(N-1)      */
};

/* Sides of the N-gon base of a prism or pyramid: 4 <= N <= MCG_MAX_SIDES
 * (Polygon<N> is instantiated for N = 4..9, boundary.cpp:253-258). */
#define MCG_MAX_SIDES 9
#define MCG_MAX_FACES (MCG_MAX_SIDES + 2)

/* One face of a cell: the boundary type chosen by the template argument in
 * the reference (e.g. Parallelepiped<IsotBoundary<Parallelogram>, Spec,
 * Spec>) and the constructor's temperature T(i). */
typedef struct mcg_face_spec {
    int32_t kind;          /* mcg_bdry_kind                                   */
    int32_t reserved;
    double temp;           /* deviational wall temperature, Isot / Peri only  */
    double specularity;    /* MCG_PARTIAL only, in [0, 1]                     */
} mcg_face_spec;

/* A convex cell: the constructor arguments (o, mat, div, gradT, T) of the
 * reference's cell templates.  The parallelepiped, triangular prism and
 * tetrahedron take three columns of mat and div per axis.  A prism or
 * pyramid with an N-gon base takes N columns: column 0 is the prism's axis
 * (the pyramid's apex), columns 1..N-1 the base's vertices after the origin,
 * all relative to the origin; as in the reference (subdomain.h:388-391) its
 * tally is one cell, or none when div[0] < 0 (div[1], div[2] are unused). */
typedef struct mcg_subdomain_spec {
    int32_t cell;              /* mcg_cell_kind                                 */
    int32_t nface;             /* 6, 5, 4; N + 2 (prism) or N + 1 (pyramid)     */
    double origin[3];
    double mat[3 * MCG_MAX_SIDES]; /* column-major: columns are the cell edges  */
    int64_t div[3];            /* tally divisions (subdomain.cpp:41-71)         */
    double grad_t[3];          /* imposed gradient: volumetric emitter          */
    mcg_face_spec face[MCG_MAX_FACES]; /* in the reference's container order    */
} mcg_subdomain_spec;

/* makePair for periodic faces (boundary.cpp:524-550: translation and
 * rotation) or for interfaces (boundary.cpp:361-369; transl/rot unused). */
typedef struct mcg_pair_spec {
    int32_t sdom_a, face_a;
    int32_t sdom_b, face_b;
    double transl[3];
    double rot[9];             /* column-major; all zeros means identity */
} mcg_pair_spec;

/* Extension for the porous-film family: a cylindrical pore whose axis is
 * parallel to z, cut out of one subdomain.  The pore wall is a boundary of
 * kind MCG_DIFF, MCG_SPEC or MCG_PARTIAL. */
typedef struct mcg_pore_spec {
    int32_t sdom;
    int32_t kind;
    double center[2];          /* axis position (x, y) */
    double radius;
    double specularity;
} mcg_pore_spec;

typedef struct mcg_domain_spec {
    int32_t nsdom;
    int32_t npair;
    int32_t npore;
    int32_t reserved;
    const mcg_subdomain_spec* sdom;
    const mcg_pair_spec* pair;
    const mcg_pore_spec* pore;
} mcg_domain_spec;

/* -------------------------------------------------------------- problem */
enum mcg_tally_kind {
    MCG_TEMP = 0,      /* TempProblem     problem.cpp:468-481  rows 1          */
    MCG_FLUX = 1,      /* FluxProblem     problem.cpp:501-514  rows 3          */
    MCG_MULTI = 2,     /* MultiProblem    problem.cpp:534-551  rows 4          */
    MCG_CUMTEMP = 3,   /* CumTempProblem  problem.cpp:557-599  rows size+1     */
    MCG_CUMFLUX = 4    /* CumFluxProblem  problem.cpp:605-648  rows 3(size+1)  */
};

enum mcg_precision {
    MCG_FP64 = 0,          /* flight, positions and tallies in FP64            */
    MCG_FP32_FLIGHT = 1    /* positions, directions, flight times, velocities
                              in FP32; tallies and sampling tables in FP64     */
};

/* FieldProblem(mat, dom, nemit, maxscat, maxloop) (problem.cpp:315-342)
 * plus the engine's own run controls. */
typedef struct mcg_problem_spec {
    int32_t tally;         /* mcg_tally_kind                                   */
    int32_t precision;     /* mcg_precision                                    */
    int64_t nemit;         /* requested bundles (rounded per emitter)          */
    int64_t maxscat;       /* scatter cap (intrinsic + diffuse events)         */
    int64_t maxloop;       /* step cap; 0 means 100 * maxscat                  */
    int64_t size;          /* CumTemp / CumFlux only                           */
    uint64_t seed;         /* Philox key of the solve                          */
    int64_t first_bundle;  /* bundles [first, first + count) are run; the      */
    int64_t bundle_count;  /* normalisation always uses the full nemit.
                              count = 0 means all of them                       */
} mcg_problem_spec;

/* What the constructor of FieldProblem derives, plus the output shape. */
typedef struct mcg_problem_info {
    int64_t nemit;         /* sum of the rounded per-emitter counts            */
    int64_t rows;          /* tally rows                                       */
    int64_t cols;          /* tally columns (cells over all subdomains)        */
    int64_t maxloop;       /* effective step cap                               */
    int64_t nemitter;      /* emitters in registry order (domain.cpp:88-102)   */
    double power;          /* power_ (problem.cpp:341)                         */
    double weight_sum;     /* sum of emitter weights                           */
} mcg_problem_info;

/* Optional per-bundle outputs (deterministic-mode parity data). Each array
 * has bundle_count entries, indexed by bundle id - first_bundle. */
typedef struct mcg_counts {
    uint32_t* nscat;       /* Phonon::nscat at termination                     */
    uint32_t* nbdry;       /* boundary events (advect returned a boundary)     */
    uint32_t* nstep;       /* executed loop iterations (advect calls)          */
    uint32_t* fate;        /* 0 absorbed, 1 escaped, 2 maxscat, 3 maxloop      */
} mcg_counts;

enum mcg_fate { MCG_ABSORBED = 0, MCG_ESCAPED = 1, MCG_MAXSCAT = 2, MCG_MAXLOOP = 3 };

typedef struct mcg_stats {
    int64_t bundles;       /* bundles run                                      */
    int64_t steps;         /* executed advect calls (the scatter-step count)   */
    int64_t escapes;       /* Progress::incrEsc count (problem.cpp:408-426)    */
    int64_t absorbed;
    int64_t truncated;     /* stopped by maxscat or maxloop                    */
    double kernel_ms;      /* device time of the trajectory kernels            */
    double wall_ms;        /* host wall time of the whole call                 */
} mcg_stats;

/* --------------------------------------------------------------- library */
int mcg_abi_version(void);
const char* mcg_last_error(void);

/* Host-only (no GPU needed). */
int mcg_material_compute(const mcg_material_spec* mat, mcg_material_tables* out);
int mcg_problem_plan(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                     const mcg_problem_spec* prob, mcg_problem_info* info,
                     double* emit_weight, int64_t* emit_count, int64_t nemitter_cap);
/* Per-column tally cell volumes of a domain (the CellVolF of problem.cpp:302-306
 * over Field's columns, field.cpp:53-77): out receives min(cap, ncols)
 * doubles, *ncols the column count.  A cell a pore cuts reports its material
 * volume; a cell wholly inside a pore reports 0 (its normalised tallies are
 * 0). */
int mcg_cell_volumes(const mcg_domain_spec* dom, double* out, int64_t cap, int64_t* ncols);
int mcg_alias_build(const double* pdf, int64_t n, uint32_t* thresh, uint32_t* alias);
void mcg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Scene: material tables and geometry resident on the current GPU. */
typedef struct mcg_scene mcg_scene;
int mcg_scene_create(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                     mcg_scene** out);
int mcg_scene_destroy(mcg_scene* scene);

/* Launch shape of the scene's single-solve kernel (block size, bundle ids
 * per wavefront grab = chunk >= 64, workgroups per CU); NULL or zeros
 * restore the defaults.  Results do not depend on it. */
typedef struct mcg_tune mcg_tune;
int mcg_scene_tune(mcg_scene* scene, const mcg_tune* tune);

/* mcg_problem_plan for a problem on an existing scene (no table rebuild). */
int mcg_scene_plan(const mcg_scene* scene, const mcg_problem_spec* prob, mcg_problem_info* info);

/* FieldProblem::solve (problem.cpp:370-445) for the whole bundle set:
 * out receives rows x cols doubles, column-major, equal to the reference's
 * return value power_ * postProc(data) / cellVol.  counts may be NULL.
 * The scene keeps its device workspace between solves: one host thread
 * solves (or traces, or autotunes) a given scene at a time. */
int mcg_solve(mcg_scene* scene, const mcg_problem_spec* prob,
              double* out, int64_t out_len, mcg_counts* counts, mcg_stats* stats);

/* The solve that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is
 * mcg_solve without per-bundle counts, on a scene already created: every
 * per-solve device buffer (tallies, the range-add array, counters, emitter
 * counts, the folded output) lives in one static device array of
 * MCG_STATIC_SCRATCH_BYTES on each GPU, held until the process ends, and no
 * hipMalloc, hipFree or hipHostMalloc is called.  A problem whose buffers do
 * not fit (rows x cols of about 2^18 doubles at most) is refused with
 * MCG_ERANGE.  Every call on a GPU shares that GPU's array and clears what it
 * uses: calls that overlap on one GPU (two host threads, or any use while a
 * call is in progress) race and give wrong results.  stats.kernel_ms is the
 * host-measured time from launch to completion. */
#define MCG_STATIC_SCRATCH_BYTES (8u * 1024u * 1024u)
int mcg_solve_static(mcg_scene* scene, const mcg_problem_spec* prob, double* out, int64_t out_len,
                     mcg_stats* stats);

/* Trajectory of one bundle of a field solve, step by step: the GPU
 * counterpart of TrajProblem::solve (problem.cpp:226-299), which prints the
 * hops of one bundle.  Record i (MCG_TRACE_WIDTH doubles) describes loop
 * iteration i: step, hit (plane index, -1 intrinsic, -2 escape, -3 pore),
 * distance moved, position after the flight (x, y, z), direction, free path
 * left, mode and scatter count after the event, subdomain, fate (-1 while
 * alive), position after the event (x, y, z; a periodic face maps it,
 * boundary.cpp:517-522).  nrec receives the step count (capped at cap; a
 * last record with fate -1 means the trajectory went on past cap). */
#define MCG_TRACE_WIDTH 17
int mcg_trace(mcg_scene* scene, const mcg_problem_spec* prob, int64_t bundle, double* rec,
              int64_t cap, int64_t* nrec);

/* TrajProblem(mat, dom, [prop,] pos, [dir,] maxscat, maxloop)::solve with a
 * given start (problem.cpp:226-299, the pos_ branch; constructors
 * problem.cpp:163-198): one trajectory from pos.  prop_cell < 0 draws the
 * mode from the scattering distribution (drawScatProp) as the reference
 * does, otherwise it is the (w, p) cell w * np + p, which must have a
 * positive energy pdf; dir NULL draws an isotropic direction (drawIso), a
 * given one is normalised.  The start cell is the first subdomain that
 * contains pos (Domain::locate, domain.cpp:58-66); "Position not inside
 * domain" otherwise.  The draws use event 0 of bundle prob->first_bundle of
 * prob's seed; maxscat, maxloop and precision come from prob (nemit and
 * the tally kind only need to be valid: the trajectory draws nothing from
 * the emitters, so the domain needs no source).  Records and nrec as
 * mcg_trace; the fate of the last record tells how it ended. */
int mcg_trace_from(mcg_scene* scene, const mcg_problem_spec* prob, int64_t prop_cell,
                   const double pos[3], const double* dir, double* rec, int64_t cap, int64_t* nrec);

/* --------------------------------------------------------------- batching
 * Many independent solves in one persistent launch.  Each solve has its
 * own scene, problem and output; results equal those of mcg_solve. */
enum mcg_engine {
    MCG_ENGINE_HISTORY = 0,    /* persistent kernel, a bundle per lane for its
                                  whole life, state in registers (default)    */
    MCG_ENGINE_EVENT = 1       /* structure-of-arrays state in HBM, one step
                                  per bundle per round, bundles regrouped by
                                  next event and cell kind between rounds     */
};

struct mcg_tune {
    int32_t block;             /* threads per block: a multiple of 64 up to 1024 (0 = auto) */
    int32_t per_thread;        /* event engine: bundles per thread per round
                                  launch, 1, 2, 4 or 8 (0 = 1); the history
                                  engine carries one bundle per lane          */
    int32_t chunk;             /* history: bundle ids per wavefront grab         */
    int32_t blocks_per_cu;     /* occupancy target (0 = auto)                   */
    int32_t lds_tables;        /* sampling tables in LDS: 1 yes, -1 no, 0 auto  */
    int32_t engine;            /* mcg_engine (single solves, mcg_solve)         */
    int32_t compact_every;     /* event engine: rounds between regroupings, the
                                  others keep bundles in place (0 = 1)         */
    int32_t defer_walks;       /* history engine, scenes with 2-D / 3-D tally
                                  grids: a step's crossing walk deferred to the
                                  wavefront's walk phase, 1 yes, -1 no (walked
                                  inside the step), 0 auto (yes when some grid
                                  has at least 32 cells).  FP64 defers only in
                                  single solves.  Results differ by rounding
                                  only.                                       */
    int64_t live;              /* event engine: bundles in flight (0 = auto)    */
};

int mcg_batch_solve(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                    double* const* outs, const int64_t* out_lens,
                    mcg_stats* stats, const mcg_tune* tune);

/* A batch captured once into a HIP graph and replayed: the scenes' device
 * buffers, the tally and counter buffers and the launch are prepared by
 * mcg_batch_create (all problems in one precision); each mcg_batch_run zeroes
 * the tallies and runs the batch kernel by one hipGraphLaunch, optionally
 * with new seeds (seeds[i] replaces problem i's seed; NULL keeps them), and
 * returns what mcg_batch_solve returns for the same problems.  The scenes
 * must outlive the batch; one host thread runs a batch at a time. */
typedef struct mcg_batch mcg_batch;
int mcg_batch_create(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                     const mcg_tune* tune, mcg_batch** out);
int mcg_batch_run(mcg_batch* batch, const uint64_t* seeds, double* const* outs,
                  const int64_t* out_lens, mcg_stats* stats);
int mcg_batch_destroy(mcg_batch* batch);

/* ---------------------------------------------------------------- serving
 * Continuous batching.  A request is one field solve (its own material,
 * domain and problem).  mcg_server_step admits queued requests in arrival
 * order while their device memory fits the budget, runs one persistent batch
 * launch over fair shares of every active request's remaining bundles
 * (step_bundles in total, default 2^24, at least 2^16 per request), and
 * retires the requests whose bundles are all done.  Tallies and counters of an admitted request stay on the GPU between
 * steps.  Results equal mcg_solve's for the same problem, alone or mixed with
 * other requests.  Steps are pipelined: mcg_server_step enqueues the next
 * step before it waits for the previous one, so a request retires one call
 * after its last bundles were handed out.  One host thread drives a server.
 * mcg_server_submit builds the request's domain, checks that a lower bound
 * of its device bytes fits the budget, copies the material's arrays and the
 * problem, and returns; the server's own host threads build the material
 * tables, the problem and the request's packed device image meanwhile.
 * mcg_server_step waits for a request's build when it reaches the head of
 * the queue and admits it with one device allocation and one copy (blocks
 * of retired requests are freed lazily, when no step is in flight, or at
 * once when that allocation fails).  A request whose material or problem is
 * invalid, that could never fit the budget, or whose allocation or upload
 * failed, is listed by mcg_server_completed and its error is returned by
 * mcg_server_poll. */
typedef struct mcg_server mcg_server;

typedef struct mcg_server_stats {
    int64_t queued, active;
    uint64_t bytes_used, bytes_budget;
    double kernel_ms;       /* device time of all steps so far          */
    int64_t steps;          /* scatter-steps of retired requests        */
} mcg_server_stats;

int mcg_server_create(uint64_t hbm_budget_bytes, int64_t step_bundles, const mcg_tune* tune,
                      mcg_server** out);
int mcg_server_submit(mcg_server* server, const mcg_material_spec* mat,
                      const mcg_domain_spec* dom, const mcg_problem_spec* prob,
                      int64_t* request_id);
int mcg_server_step(mcg_server* server, int32_t* active, int32_t* queued, int32_t* completed);
/* Ids of the requests the last mcg_server_step retired (n = how many). */
int mcg_server_completed(const mcg_server* server, int64_t* ids, int64_t cap, int64_t* n);
/* MCG_OK and the result (then the request is forgotten), MCG_EAGAIN while it
 * runs, MCG_ENOTFOUND for an unknown id, or the error of a request that
 * failed (then it is forgotten too).  stats.wall_ms = submit to done. */
int mcg_server_poll(mcg_server* server, int64_t request_id, double* out, int64_t out_len,
                    mcg_stats* stats);
int mcg_server_cancel(mcg_server* server, int64_t request_id);
int mcg_server_info(const mcg_server* server, mcg_server_stats* out);
int mcg_server_destroy(mcg_server* server);

/* ------------------------------------------------------------- autotuner
 * Times candidate shapes on a bounded sample of the problem's bundles and
 * returns the fastest; mcg_autotune also installs it on the scene
 * (mcg_scene_tune).  mcg_autotune's candidates are both engines: the history
 * kernel's block (64-1024), table placement (LDS or not) and bundle ids per
 * wavefront grab, and the event engine's bundles in flight, bundles per
 * thread (1, 4) and rounds between regroupings (1, 4); on scenes with 2-D /
 * 3-D tally grids, then the best history shape with its walks deferred and
 * in the step (defer_walks 1, -1).  mcg_autotune_batch tunes the batch
 * kernel's block, table placement and grab (and, for FP32 batches with
 * grids, defer_walks).  Only
 * scheduling changes, never results.  sample_bundles <= 0 means
 * min(nemit, 2^20); sample_fraction <= 0 means 0.1. */
int mcg_autotune(mcg_scene* scene, const mcg_problem_spec* prob, int64_t sample_bundles,
                 mcg_tune* best, double* best_steps_per_s);
int mcg_autotune_batch(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                       double sample_fraction, mcg_tune* best, double* best_steps_per_s);

#ifdef __cplusplus
}
#endif

#endif /* MCG_H */
