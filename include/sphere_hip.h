/*
 * sphere_hip.h -- C ABI of the MI355X (gfx950) DEM + Darcy library.
 *
 * The reference program has no library boundary: python/sphere.py
 * writes input/<sid>.bin and runs "cd ..; ./sphere [flags] input/<sid>.bin"
 * (python/sphere.py:4482-4540), and the CUDA binary reads/writes the
 * versioned binary format (src/file_io.cpp:30-731) and runs the time loop
 * DEM::startTime (src/device.cu:833-2834).  This header is the drop-in
 * replacement of that process boundary: every entry point below replaces
 * one step of that contract and is cited at its declaration.
 *
 * Conventions
 *   - extern "C", plain pointers and sizes, no exceptions across the ABI.
 *   - int status: 0 = OK, negative = error (sph_last_error() has the text).
 *   - One context is single threaded and not reentrant; every call runs on
 *     the context's own HIP stream and returns when its work is complete.
 *   - Caller-owned buffers are copied in or out and never retained.
 */
#ifndef SPHERE_HIP_H_
#define SPHERE_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_VERSION 2.15      /* src/version.h:5 */
#define SPH_NC 32             /* contact slots per particle, src/constants.h:21 */
#define SPH_NC_MAX 64         /* largest slot count sph_set_contact_slots accepts */
#define SPH_MAXWALLS 6        /* src/constants.h:12 */

/* Status codes */
#define SPH_OK 0
#define SPH_ERR_ARG -1        /* bad argument / inconsistent sizes */
#define SPH_ERR_FILE -2       /* file missing, short or wrong version */
#define SPH_ERR_VALUE -3      /* failed value check (src/sphere.cpp:145-345) */
#define SPH_ERR_HIP -4        /* HIP runtime error or no device */
#define SPH_ERR_GRID -5       /* a particle left the sorting grid */
#define SPH_ERR_NAN -6        /* NaN Darcy residual (src/darcy.cpp:276-285) */
#define SPH_ERR_OVERFLOW -7   /* more than SPH_NC contacts on one particle */
#define SPH_ERR_STATE -8      /* call out of order (no state loaded, ...) */

/*
 * In-memory image of one v2.15 binary file (field order of
 * src/file_io.cpp:420-731 and python/sphere.py:1355-1553).  All arrays are
 * caller owned.  Per-particle vectors are packed rows; fluid cell arrays
 * are interior cells in file order (x fastest, then y, then z).
 */
typedef struct sph_state {
    /* header, src/file_io.cpp:47-104 */
    int32_t nd;
    uint32_t np;
    double time_dt, time_current, time_total, time_file_dt;
    uint32_t time_step_count;
    double origo[3], L[3];
    uint32_t num[3];
    int32_t periodic, adaptive;

    /* particles, src/file_io.cpp:106-164 */
    double *x;       /* np*4: x, y, z, radius      */
    double *xyzsum;  /* np*3                        */
    double *vel;     /* np*4: vx, vy, vz, fixvel    */
    double *force;   /* np*3 */
    double *angpos;  /* np*3 */
    double *angvel;  /* np*3 */
    double *torque;  /* np*3 */
    double *es_dot, *es, *ev_dot, *ev, *p;  /* np each */

    /* params, src/file_io.cpp:166-186 (k_r carries Poisson's ratio for the
     * Hertz-Mindlin model, see contactmodel 3) */
    double g[3];
    double k_n, k_t, k_r, E, gamma_n, gamma_t, gamma_r;
    double mu_s, mu_d, mu_r, gamma_wn, gamma_wt, mu_ws, mu_wd, rho;
    uint32_t contactmodel;
    double kappa, db, V_b;

    /* walls, src/file_io.cpp:188-220 */
    uint32_t nw;
    int32_t wmode[SPH_MAXWALLS];
    double w_nx[SPH_MAXWALLS * 4];    /* normal x,y,z and position        */
    double w_mvfd[SPH_MAXWALLS * 4];  /* mass, velocity, force, sigma0    */
    double sigma0_A, sigma0_f, tau_x;

    /* bonds, src/file_io.cpp:222-249 */
    double lambda_bar;
    uint32_t nb0;
    double sigma_b, tau_b;
    uint32_t *bonds;          /* nb0*2 (broken bonds hold nb0 in column 0) */
    double *bonds_delta_n;    /* nb0   */
    double *bonds_delta_t;    /* nb0*3 */
    double *bonds_omega_n;    /* nb0   */
    double *bonds_omega_t;    /* nb0*3 */

    /* Darcy fluid block, present iff fluid != 0 (src/file_io.cpp:329-398) */
    int32_t fluid;            /* not in the file: the --fluid flag */
    int32_t cfd_solver;       /* must be 1 (Darcy) */
    double mu;
    double *v_f;              /* cells*3 */
    double *p_f, *phi, *dphi; /* cells each; dphi as the native solver stores it */
    double rho_f, p_mod_A, p_mod_f, p_mod_phi;
    int32_t bc_xn, bc_xp, bc_yn, bc_yp, bc_bot, bc_top;
    int32_t free_slip_bot, free_slip_top;
    double bc_bot_flux, bc_top_flux;
    int32_t *p_constant;      /* cells */
    double tolerance;
    uint32_t maxiter, ndem;
    double c_phi;
    double *f_p;              /* np*3 */
    double beta_f, k_c;

    int32_t *color;           /* np */
} sph_state;

/* Options of sph_run: the CLI flags of src/main.cpp:45-163 that affect a run. */
typedef struct sph_run_opts {
    const char *input_path;   /* input/<sid>.bin (relative to cwd)          */
    const char *output_dir;   /* "output" in the reference layout           */
    int32_t fluid;            /* --fluid                                    */
    int32_t verbose;          /* 0 = -q                                     */
    int32_t dry;              /* --dry: read + check only                   */
    int32_t device;           /* -d N (negative: device 0)                  */
} sph_run_opts;

/* Darcy solver statistics since the last sph_load_* call. */
typedef struct sph_darcy_stats {
    uint64_t fluid_steps;     /* Jacobi solves performed                     */
    uint64_t sweeps;          /* total Jacobi sweeps                          */
    uint64_t nonconverged;    /* solves that hit maxiter (warning, not error) */
    uint32_t last_nijac;      /* nijac at the last convergence check          */
    uint32_t pad;
    double last_max_residual; /* max |norm| at the last check                 */
    double solve_ms;          /* device time spent in Jacobi loops            */
    uint64_t corrections;     /* learned corrections applied (sph_set_darcy_net) */
} sph_darcy_stats;

/* Kernel identifiers for sph_kernel_time. */
#define SPH_K_HASH 0
#define SPH_K_SORT 1
#define SPH_K_REORDER 2
#define SPH_K_CONTACT 3       /* contact forces, walls and sums (interact) */
#define SPH_K_BONDS 4
#define SPH_K_INTEGRATE 5
#define SPH_K_WALLS 6
#define SPH_K_DARCY_SETUP 7
#define SPH_K_DARCY_SWEEP 8
#define SPH_K_NLIST 9         /* neighbour-list rebuild (no reference counterpart) */
#define SPH_K_TOPOLOGY 10     /* contact search and slot bookkeeping (topology) */
#define SPH_K_COUNT 11

/* Neighbour-list statistics (sph_get_neighbor_stats). */
typedef struct {
    uint32_t enabled;   /* 1 when the contact search uses the lists */
    uint32_t range;     /* cells searched on each side at a rebuild */
    uint32_t cap;       /* list entries per particle */
    uint32_t overflow;  /* 1 once a list overflowed (rebuilt every step since) */
    double skin;        /* listed surface gap [m] */
    uint64_t rebuilds;  /* rebuilds since the last load */
    uint64_t entries;   /* list entries over all particles at the last rebuild */
    uint32_t longest;   /* longest list (cap + 1 if one overflowed) */
    uint32_t full_steps; /* steps whose search tested every list entry (rebuilds,
                            periodic wraps); the others test only entries that
                            can have come into contact */
} sph_neighbor_stats;

/* Context.  Replaces DEM::DEM(..., initCuda=1, device) (src/sphere.cpp:16-95)
 * and initializeGPU (src/device.cu:65-155); device < 0 picks device 0. */
typedef struct sph_ctx sph_ctx;
sph_ctx *sph_create(int device);
/* Replaces DEM::~DEM + freeing of device memory (src/device.cu:602-664). */
void sph_destroy(sph_ctx *ctx);
/* Text of the last error on this context (or of sph_create failure if ctx is NULL). */
const char *sph_last_error(const sph_ctx *ctx);

/* Contact slots per particle for the next sph_load_*: 32 (the reference's
 * compile-time NC, src/constants.h:21; the default) or 64.  Slots are filled
 * from the top, so the slot order of the contacts -- and the force summation
 * order -- is the same for both sizes while a particle has at most 32 contacts;
 * 64 serves packings whose largest grains touch more than 32 others, where the
 * reference writes past the slot row (src/contactsearch.cuh:298-299). */
int sph_set_contact_slots(sph_ctx *ctx, int nc);

/* Neighbour lists for contact models 2 and 3.  No counterpart in the
 * reference, which searches the 27 cells around every particle each step
 * (src/contactsearch.cuh:358-369); the lists give the same contacts, slot
 * assignment and forces.  skin_fraction is the listed surface gap as a
 * fraction of the largest radius (default 0.1; 0 disables the lists, and so
 * does a grid that leaves no room for it); cap is the number of list entries
 * per particle (1..128, default 128; a particle with more neighbours than
 * that is searched over the cells, and the lists are then rebuilt every
 * step).  Takes effect at the next load.  In list mode the cell sort runs
 * only on rebuild steps, so sph_read_sort returns the arrays of the last
 * rebuild. */
int sph_set_neighbor_lists(sph_ctx *ctx, double skin_fraction, int cap);
int sph_get_neighbor_stats(sph_ctx *ctx, sph_neighbor_stats *out);

/* How list mode rebuilds its lists.  No counterpart in the reference.
 * fused = 1 (default): the cell sort and reorder run as one kernel with grid
 * barriers, followed by the list build; both are launched every step and
 * return at once unless the rebuild flag is set.  fused = 0: the sort
 * passes, reorder and list build are separate launches that each return at
 * once between rebuilds.  Runs with fluid rebuild every step and always use
 * the separate launches.  Both give the same arrays.  Takes effect at the
 * next load. */
int sph_set_rebuild_kernel(sph_ctx *ctx, int fused);

/* Bound on the rebuild kernel's grid-barrier waits, in sleeps of a waiting
 * block (0: the default, 2^22).  No counterpart in the reference.  A wait
 * that runs past it ends the kernel on every block, the list and contact
 * kernels after it skip their work, and the step returns SPH_ERR_HIP; a
 * small bound forces that path from the host, for tests of the failure
 * handling.  Takes effect at once. */
int sph_set_barrier_spin_limit(sph_ctx *ctx, uint32_t sleeps);

/* Load a state: replaces DEM::readbin + checkValues + transfer to the device
 * (src/file_io.cpp:30-417, src/sphere.cpp:145-345, src/device.cu:667-752).
 * Resets the time-loop counters (iter = 0, as every reference invocation). */
int sph_load_state(sph_ctx *ctx, const sph_state *s);
int sph_load_file(sph_ctx *ctx, const char *path, int fluid);

/* Header of the loaded state (scalars only; array pointers set to NULL), so
 * that a caller can size buffers for sph_read_state. */
int sph_get_header(sph_ctx *ctx, sph_state *hdr);
/* Copy the device state into caller buffers (any NULL array is skipped):
 * replaces transferFromGlobalDeviceMemory (src/device.cu:754-829). */
int sph_read_state(sph_ctx *ctx, sph_state *out);
/* Write the current state as a v2.15 file: replaces DEM::writebin
 * (src/file_io.cpp:420-731). */
int sph_write_file(sph_ctx *ctx, const char *path);

/* Advance n DEM time steps (one pass of the loop body of
 * src/device.cu:1038-2545 each, including the Darcy solve when fluid). */
int sph_step(sph_ctx *ctx, int nsteps);
/* Full run: replaces main() + DEM::startTime (src/main.cpp:28-225,
 * src/device.cu:833-2834): outputs <dir>/<sid>.outputNNNNN.bin, status.dat,
 * <sid>-conv.log exactly as the reference names and formats them. */
int sph_run(const sph_run_opts *opts);

/* Diagnostics for bit-exact parity (the reference keeps these on the
 * device: src/device.cu:1066-1174).  Any NULL pointer is skipped.
 *   cellid, index: np each (the sorted key/value pairs of the last step)
 *   cell_start, cell_end: num[0]*num[1]*num[2] each (0xffffffff = empty)  */
int sph_read_sort(sph_ctx *ctx, uint32_t *cellid, uint32_t *index,
                  uint32_t *cell_start, uint32_t *cell_end);
/* Per-contact history (src/device.cu:436-439 layout, orig index * nc + slot,
 * nc = the slot count of the context):
 *   contacts: np*nc (empty slot = np), distmod: np*nc*3, delta_t: np*nc*3 */
int sph_read_contacts(sph_ctx *ctx, uint32_t *contacts, double *distmod,
                      double *delta_t);
/* Restore contact history (a restart side car; absent from the v2.15 file). */
int sph_write_contacts(sph_ctx *ctx, const uint32_t *contacts,
                       const double *distmod, const double *delta_t);
/* Restart side car (no counterpart in the reference, whose restarts drop
 * the contact history and restart TY3 with zero previous accelerations,
 * src/file_io.cpp:420-731, src/integration.cuh:74-81): the occupied contact
 * slots with their tangential springs, the previous accelerations of grains
 * and walls, the step counter, the velocity-step state and the Darcy top
 * pressure base.  sph_run writes <sid>.contacts.NNNNN.bin beside every output
 * file and, when <input>.contacts.bin lies beside its input file and belongs
 * to it (same time, output number and sizes), restores it, so that a run
 * continued from an output file equals the uninterrupted run bit for bit.
 * sph_read_restart needs the state it was written with to be loaded. */
int sph_write_restart(sph_ctx *ctx, const char *path);
int sph_read_restart(sph_ctx *ctx, const char *path);

/* Number of occupied contact slots summed over all particles. */
int sph_count_contacts(sph_ctx *ctx, uint64_t *ncontacts);

int sph_get_darcy_stats(sph_ctx *ctx, sph_darcy_stats *out);

/* Per-kernel device timing with HIP events on the context stream. */
int sph_set_profiling(sph_ctx *ctx, int on);
int sph_kernel_time(sph_ctx *ctx, int kernel_id, double *total_ms,
                    uint64_t *launches);

/* ------------------------------------------------------------------------
 * Learned correction of the Darcy solve.  No counterpart in the reference:
 * it changes only how many Jacobi sweeps a solve takes.  The sweeps, their
 * convergence test (max relative change <= tolerance every 10 sweeps) and
 * maxiter stay those of the reference (src/device.cu:2184-2386); after the
 * test of sweeps 0, 10, 20, ... has not ended the solve, a 3-D U-Net (BF16
 * on the matrix cores, FP32 accumulation) maps the fixed-point residual of
 * the current iterate to an estimate of its error, which is added to the
 * iterate.  Plain Jacobi stays the fallback and decides convergence.
 * ---------------------------------------------------------------------- */
typedef struct sph_net sph_net;

/* Number of FP32 parameters of the network (per layer w[27][cout][cin],
 * then b[cout]). */
uint64_t sph_net_param_count(void);
/* Network for an nx*ny*nz grid (each a multiple of 8) on a device;
 * periodic: bit 0 x, bit 1 y wrap around (else zero padding). */
sph_net *sph_net_create(int device, uint32_t nx, uint32_t ny, uint32_t nz,
                        uint32_t periodic);
void sph_net_destroy(sph_net *net);
const char *sph_net_last_error(void);
/* Device pointers; stream is a hipStream_t (NULL: the null stream).
 * forward: input[V][3] -> output[V], V = nx*ny*nz, x fastest.
 * backward: of the last forward; adds dLoss/dparams to grad_params given
 * grad_output = dLoss/doutput. */
int sph_net_forward(sph_net *net, const float *params, const float *input,
                    float *output, void *stream);
int sph_net_backward(sph_net *net, const float *params, const float *grad_output,
                     float *grad_params, void *stream);

/* One of the network's 3x3x3 convolutions on its own (for tests):
 * in[V][cin] and out[V][cout] FP32, w[27][cout][cin], b[cout] or NULL;
 * operands rounded to BF16, FP32 sums; device pointers. */
int sph_net_conv3d(const float *in, int cin, const float *w, const float *b, int cout,
                   uint32_t nx, uint32_t ny, uint32_t nz, uint32_t periodic, float *out,
                   void *stream);

/* The convolution of sph_net_conv3d without any device allocation: no
 * counterpart in the reference.  It is for callers that cannot allocate --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The BF16 staging of the input and weights goes to a
 * library-owned scratch array of 16 MiB per GPU (a static device array each
 * GPU holds until the process ends), so V * pad(cin) * 2 bytes (rounded up to
 * 256) plus 27 * pad(cout) * pad(cin) * 2 bytes must fit in 16 MiB, with
 * pad(c) = 4 for c <= 4, else c rounded up to 16; larger calls return
 * SPH_ERR_ARG.  The work is enqueued on `stream` and the call returns
 * without waiting for it.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (different streams, or host threads) race.
 * The first call on a device resolves the scratch address; make it outside
 * graph capture. */
int sph_net_conv3d_noalloc(const float *in, int cin, const float *w, const float *b, int cout,
                           uint32_t nx, uint32_t ny, uint32_t nz, uint32_t periodic, float *out,
                           void *stream);

/* Use the network in the context's Darcy solves: params on the host
 * (count = sph_net_param_count()), at most max_corrections per solve;
 * params == NULL or max_corrections == 0 switches it off. */
int sph_set_darcy_net(sph_ctx *ctx, const float *params, uint64_t count,
                      int max_corrections);

/* Training data from the last Darcy solve of the context (device
 * pointers; fields with the ghost layer, (nx+2)(ny+2)(nz+2) doubles).  They
 * run on the context's stream and return when done; the caller finishes
 * its own pending work on the buffers first. */
#define SPH_DARCY_P_START 0   /* the iterate the solve started from (p_old) */
#define SPH_DARCY_P_FINAL 1   /* the converged pressure                     */
int sph_darcy_export(sph_ctx *ctx, int which, double *dst);
/* nsweeps plain Jacobi sweeps of the last solve's system applied to p. */
int sph_darcy_sweeps(sph_ctx *ctx, double *p, int nsweeps);
/* Network input [V][3] for iterate p, and the residual scale s (host):
 * the network's target for p is (p_fixed - p) / s on the cells. */
int sph_darcy_net_input(sph_ctx *ctx, const double *p, float *input, double *scale);

/* ------------------------------------------------------------------------
 * Ensembles of independent boxes (no counterpart in the reference, which
 * runs one simulation per process and gets its parallel runs from job
 * scripts, python/sphere.py:4549-4692,7638-7720).  The boxes are contexts
 * loaded as usual (any sizes and parameters, one device); sph_ensemble_step
 * advances every box by nsteps steps and returns when all are done.
 *
 * Boxes on the neighbour-list path without fluid (contact models 2 and 3,
 * no adaptive grid, no shear-stress wall) are stepped together: each stage
 * of a step is one kernel launch for all of them, with per-box task
 * records in device memory, and pairs of steps are captured in one HIP
 * graph and replayed.  Darcy-coupled boxes on the list path (one DEM step
 * per fluid step, no modulated top pressure, the same maxiter; boxes with
 * a learned correction need grids in multiples of 8 cells and the same
 * number of corrections per solve) are stepped together too, one launch
 * per stage of the fluid step, with their Jacobi sweeps batched until every
 * box's device-side test has ended its solve and the networks' forward
 * passes as multi-box launches.  Other boxes are stepped box by box
 * on their own streams.  Either way each box's results equal those of
 * sph_step on that box alone, bit for bit.  An output event or the velocity
 * step of t = 10 s / 15 s inside the steps is refused (SPH_ERR_VALUE).
 * ---------------------------------------------------------------------- */
typedef struct sph_ensemble sph_ensemble;

sph_ensemble *sph_ensemble_create(sph_ctx **boxes, int n);
void sph_ensemble_destroy(sph_ensemble *ens);
const char *sph_ensemble_last_error(const sph_ensemble *ens);
int sph_ensemble_step(sph_ensemble *ens, int nsteps);
/* steps taken so far by every box */
int sph_ensemble_stats(const sph_ensemble *ens, uint64_t *steps);

typedef struct {
    uint32_t boxes;              /* boxes in the ensemble                      */
    uint32_t multi_boxes;        /* of them stepped by the multi-box launches  */
    uint32_t launches_per_step;  /* kernel launches per step for those boxes   */
    uint32_t graph;              /* 1: a captured two-step graph is in use     */
    uint64_t multi_steps;        /* steps taken by the multi-box launches      */
    uint64_t graph_replays;      /* replays of the captured graph (2 steps)    */
    uint32_t fluid_boxes;        /* Darcy-coupled boxes in multi-box launches  */
    uint32_t pad;
    uint64_t fluid_steps;        /* steps taken by them                        */
} sph_ensemble_info_t;
int sph_ensemble_info(const sph_ensemble *ens, sph_ensemble_info_t *out);
/* on = 0: enqueue the multi-box launches step by step instead of replaying
 * a captured graph (same results). */
int sph_ensemble_set_graph(sph_ensemble *ens, int on);
/* Autotuner of each box's neighbour-list skin (fraction of its largest
 * radius): for each of the nfrac fractions, every box runs trial_steps
 * steps with it, timed with HIP events on its stream; each box keeps its
 * fastest.  The lists are exact for any skin, so the trial steps are
 * ordinary steps of the run (every box advances by nfrac * trial_steps
 * steps) and the results do not depend on the choice.  chosen[box]
 * (0 for a box without lists) and ms_per_step[box * nfrac + f] (-1 where
 * the fraction does not fit the box's grid) may be NULL. */
int sph_ensemble_autotune(sph_ensemble *ens, int trial_steps, const double *fractions,
                          int nfrac, double *chosen, double *ms_per_step);

/* ------------------------------------------------------------------------
 * Navier-Stokes fluid (cfd_solver = 0): the reference's fluid part of
 * each time step, src/device.cu:1213-1898, on a staggered grid
 * (cell-centred p, face velocities, one ghost layer of cells and of
 * faces): porosities (1 without grains, src/navierstokes.cuh:1319-1328),
 * viscous and advective terms, predicted face velocities (Chorin projection
 * with beta = 0, or a pressure-dependent one with beta > 0), the Jacobi
 * iteration of the pressure change epsilon with its normalized-residual
 * test every 10 iterations, optional smoothing, and the pressure and
 * velocity updates (src/navierstokes.cuh:261-3514).  On its own a fluid
 * has no grains (or grains held in place, sph_ns_set_grains); attached to
 * a context (sph_ns_attach) it moves the context's grains and sees its top
 * wall.  Boundaries: bc 0 fixed pressure, 1 no flow with free slip, 2 no
 * flow without slip (python/sphere.py:3527-3598); x and y are periodic.
 * ---------------------------------------------------------------------- */
typedef struct sph_ns_params {
    uint32_t nx, ny, nz;          /* cells, each >= 3 */
    double L[3];                  /* domain size (grid.L) */
    double dt;                    /* the DEM time step (time.dt) */
    double t0;                    /* time.current at the first step */
    double mu, rho_f;             /* dynamic viscosity, fluid density */
    double g[3];                  /* gravity (params.g) */
    double beta, theta, gamma;    /* projection, relaxation, smoothing */
    double tolerance, c_v;        /* Jacobi test, velocity scaling */
    uint32_t maxiter, ndem;       /* Jacobi limit, DEM steps per fluid step */
    int32_t bc_bot, bc_top;       /* 0, 1 or 2, as above */
    double p_mod_A, p_mod_f, p_mod_phi; /* top pressure modulation */
    double c_phi;                 /* porosity scaling (grains only) */
} sph_ns_params;

typedef struct sph_ns_stats {
    uint64_t steps;               /* time steps taken */
    uint64_t solves;              /* Jacobi solves (steps with iter % ndem == 0) */
    uint64_t iterations;          /* Jacobi iterations in all */
    uint64_t nonconverged;        /* solves that reached maxiter */
    uint32_t last_nijac;          /* iterations of the last solve */
    double last_max_norm;         /* last tested max normalized residual */
} sph_ns_stats;

typedef struct sph_ns sph_ns;

/* A fluid on `device` from host arrays of the interior cells (index
 * x + nx*(y + ny*z)): v[3*cell] the cell-centred velocity (ns.v), p[cell]
 * the pressure (ns.p).  The first step forms the face velocities from v
 * (src/device.cu:1242-1259).  NULL on failure (sph_ns_last_error()). */
sph_ns *sph_ns_create(int device, const sph_ns_params *params, const double *v,
                      const double *p);
void sph_ns_destroy(sph_ns *ns);
const char *sph_ns_last_error(void);
/* nsteps time steps (DEM iterations with np = 0).  SPH_ERR_NAN: a NaN
 * normalized residual (src/navierstokes.cpp:275-297). */
int sph_ns_step(sph_ns *ns, int nsteps);
/* The interior cells' v, p, phi, dphi (any may be NULL), as the reference
 * transfers them for output (src/navierstokes.cuh:177-200,
 * src/device.cu:2595).  The host copy
 * of p that the reference's step reads for its boundary values
 * (ns.p[idx(0,0,nz-1)], src/device.cu:1357-1360) is refreshed here, as
 * there. */
int sph_ns_read(sph_ns *ns, double *v, double *p, double *phi, double *dphi);
/* Whole ghosted arrays for checks: which 0-2 the face velocities v_x, v_y,
 * v_z ((nx+3)(ny+3)(nz+3) each), 3 epsilon ((nx+2)(ny+2)(nz+2)). */
int sph_ns_read_field(sph_ns *ns, int which, double *dst);
int sph_ns_get_stats(const sph_ns *ns, sph_ns_stats *out);
/* Grains held in place in the fluid (np = 0 clears them): rows of 4,
 * x = (x, y, z, radius) and vel = (vx, vy, vz, unused), the grid's origin
 * and periodic mode (0 none, 1 x and y, 2 x; src/contactsearch.cuh:12-84).
 * Each later step forms the porosities from them
 * (findPorositiesVelocitiesDiametersSpherical, src/navierstokes.cuh:
 * 1136-1318), the drag, pressure and viscous force on each grain
 * (findInteractionForce, :3141-3282) and the drag's reaction on the fluid
 * (applyInteractionForceToFluid, :3284-3345), which the velocity
 * prediction adds when mu > 0 (src/device.cu:1213-1351).  The grains do not
 * move.  SPH_ERR_GRID: a grain outside the grid. */
int sph_ns_set_grains(sph_ns *ns, uint32_t np, const double *x, const double *vel,
                      const double *origo, int32_t periodic);
/* The last step's forces on the grains, np rows of 3 each (any may be NULL):
 * drag, pressure, viscous and their sum (ns.f_d, f_p, f_v, f_sum). */
int sph_ns_read_grain_forces(sph_ns *ns, double *f_d, double *f_p, double *f_v, double *f_sum);
/* Couples the fluid to a loaded context without a Darcy fluid whose grid
 * (num, L) is the fluid's: each sph_step of the context then runs the
 * fluid's iteration with the context's grains between their contact and
 * bond forces and the integration, and adds its force on each grain
 * (src/device.cu:1213-1351, findInteractionForce's dev_force += f_pf); the
 * fluid must not have stepped before.  sph_ns_read and
 * sph_ns_read_grain_forces read it between steps; sph_ns_step and
 * sph_ns_set_grains are refused while attached.  ns = NULL detaches.
 * Detach (or destroy the context) before destroying the fluid. */
int sph_ns_attach(sph_ns *ns, sph_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* SPHERE_HIP_H_ */
