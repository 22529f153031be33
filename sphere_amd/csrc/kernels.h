// kernels.h -- host-side launchers of the gfx950 kernels (dem_kernels.hip,
// darcy_kernels.hip).  Every launcher enqueues on the given stream and
// returns immediately.
#pragma once

#include "common.h"
#include "multi.h"

// Step counter and simulation time on the device (read by integrate and the
// wall integration, advanced by k_clock_advance after them), so that a step
// replayed from a captured graph sees its own values.
struct StepClock {
    long long iter;
    double time;
};

struct DemBuffers {
    // particle state, original order (src/datatypes.h:16-32)
    // rows without a fourth component are double3 (24 B)
    double4 *x, *vel;
    double3 *angvel, *force, *torque, *angpos, *acc, *angacc, *xyzsum;
    double *es_dot, *es, *ev_dot, *ev, *p;
    // sort (src/datatypes.h:52-60)
    uint32_t *keys, *vals, *keys_tmp, *vals_tmp, *hist;
    StepClock *clock;
    uint32_t *cell_start, *cell_end;
    double4 *x_s, *vel_s;
    double3 *angvel_s;
    // contact history, original index * nc + slot
    uint32_t *contacts;   // np*nc: partner index | image code << 28 (ct_word), empty = np
    uint64_t *cmask;      // np, bit i set iff contacts[.. + i] != np
    double *delta_t;      // np*nc*3
    // walls
    double4 *w_nx, *w_mvfd;
    double *w_acc;
    double *walls_force_pp;   // nw*np
    double *wall_partial;     // nblocks
    double *tau_pp, *tau_partial, *tau_eff;  // wmode 3
    // bonds
    uint2 *bonds;
    uint8_t *bond_broken;
    double4 *bond_delta, *bond_omega;  // (t.x, t.y, t.z, n)
    double4 *bond_f, *bond_t;
    uint8_t *bond_applied;
    uint32_t *bp_list, *bp_off, *bp_ent;  // bonded particles, CSR offsets, entries
    uint32_t *pp_off;                     // np+1: CSR offsets into bp_ent per particle
    uint32_t nbp;
    uint32_t *errflags;
    uint32_t *rb_bar;    // 2 words: the rebuild kernel's grid barrier (arrivals, generation)
    uint32_t *rb_part;   // one word per block of the rebuild kernel
    // neighbour lists
    uint32_t *cellc;     // np, packed cell coordinates of this step
    double4 *x_ref;      // np, positions at the last rebuild
    uint32_t *nl;        // cap*np, column-major by original index
    uint32_t *nl_cnt;    // np, entries per particle (NL_OVF: overflow)
    uint32_t *nl_pos;    // np, position in the cell sort of the last rebuild
    uint8_t *nl_slot;    // cap*np like nl: slot of each list entry in the row (0xff: none)
    uint8_t *slot_cand;  // np*nc like contacts: list entry of each slot (0xff: none)
    uint32_t *nl_flags;  // NLF_WORDS words, see common.h
    float *nl_gap;       // cap*np like nl: surface gap at the rebuild, rounded down
    float *nl_gmin;      // np: lower bound of nl_gap over the entries not in the row
    float *nl_disp;      // np: displacement since the rebuild, rounded up
};

// Bond contributions gathered inside integrate (off == nullptr: none).
struct BondGather {
    const uint32_t *off, *ent;  // per-particle CSR of (bond << 1 | side)
    const double4 *f, *t;
    const uint8_t *applied;
};

// The next step's cell hash, formed inside integrate (keys == nullptr: none).
struct NextHash {
    uint32_t *keys, *vals, *cellc;
    const double4 *x_ref;
    float *disp;
    uint32_t *flags;
    uint32_t cur;   // rebuild-flag word of the next step
    double disp2;
};

void launch_hash(hipStream_t s, const StepArgs &a, const DemBuffers &b);
void launch_hash_nl(hipStream_t s, const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                    uint32_t cur, uint32_t force);
// Stable LSD radix sort of (keys, vals) by key (bits = key width).  With a
// non-null cond every kernel returns at once if *cond == 0 on the device.
void launch_sort(hipStream_t s, uint32_t n, int bits, DemBuffers &b, const uint32_t *cond);
void launch_reorder(hipStream_t s, const StepArgs &a, const DemBuffers &b, uint32_t ncells,
                    const uint32_t *cond);
// The list rebuild: the cell sort and reorder as one launch of `blocks`
// resident blocks, then the list build; both return at once unless the
// rebuild flag at cond is set.
struct RebuildArgs {
    const uint32_t *cond;
    uint32_t *keys[2], *vals[2];
    uint32_t *hist, *part;
    uint32_t n, ntiles, passes, ncells;
    uint32_t *cell_start, *cell_end;
    const double4 *x, *vel;
    const double3 *angvel;
    double4 *x_s, *vel_s;
    double3 *angvel_s;
    uint32_t *bar, *err;
    uint32_t spin_limit;  // sleeps a barrier waiter takes before it gives up
};
int rebuild_blocks_per_cu();
void launch_rebuild(hipStream_t s, const StepArgs &a, const NLArgs &n, DemBuffers &b,
                    uint32_t ncells, int bits, uint32_t blocks, const uint32_t *cond,
                    uint32_t spin_limit);
void launch_nl_build(hipStream_t s, const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                     const uint32_t *cond);
// out[0] += sum of list lengths, out[1] = max(out[1], longest list)
void launch_nl_stats(hipStream_t s, const DemBuffers &b, uint32_t np, uint32_t cap,
                     unsigned long long *out);
// Contact search (models 2, 3): new overlaps take slots (topology).  near:
// list mode on a grid searched one cell deep; cur: the step's flag word.
void launch_topology(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool nlist,
                     bool near, uint32_t cur);
// Contact forces, walls and per-particle sums (interact).
void launch_contact(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool nlist);
// gather: add the bond forces to force/torque here (else integrate does it)
void launch_bonds(hipStream_t s, const StepArgs &a, const DemBuffers &b, bool gather);
void launch_shear_stress(hipStream_t s, const StepArgs &a, const DemBuffers &b);
void launch_clock_advance(hipStream_t s, StepClock *clock, double dt);
void launch_integrate(hipStream_t s, const StepArgs &a, const DemBuffers &b,
                      int change_velocity_state, double velocity_factor, double tau_x,
                      bool gather_bonds, const NextHash &nh);
void launch_integrate_walls(hipStream_t s, const StepArgs &a, const DemBuffers &b);
void launch_fill_u32(hipStream_t s, uint32_t *p, size_t n, uint32_t v);
void launch_fill_f64(hipStream_t s, double *p, size_t n, double v);
void launch_count_contacts(hipStream_t s, const uint64_t *cmask, uint32_t np,
                           unsigned long long *out);
void launch_rebuild_mask(hipStream_t s, const DemBuffers &b, uint32_t np, uint32_t nc);

// ------------------------------------------------------------------
// Multi-box launches (ensembles of independent boxes, no counterpart in
// the reference): one launch runs one stage of a step for every box.  The
// launch's blocks are divided among the boxes: box k owns the blocks
// [start[k], start[k] + count[k]).  The starts are multiples of 8, so that a
// box's blocks meet the eight XCDs in the order of a launch of its own;
// the blocks between one box's last block and the next start return at
// once.  Each box's arguments are one task record in device memory, the
// arguments of the single-box kernel, which runs the same body.
// ------------------------------------------------------------------
struct RebuildTask {
    StepArgs a;
    RebuildArgs r;
};

struct NlBuildTask {
    StepArgs a;
    NLArgs n;
    const uint32_t *cond, *vals, *cell_start, *cell_end;
    const double4 *x_s;
    uint32_t *nl, *nl_cnt, *nl_pos;
    double4 *x_ref;
    uint32_t *flags;
    const uint32_t *contacts;
    const uint64_t *cmask;
    uint8_t *nl_slot, *slot_cand;
    float *nl_gap, *nl_gmin;
    const uint32_t *err;
};

struct TopoTask {
    StepArgs a;
    const uint32_t *vals, *cell_start, *cell_end;
    const double4 *x, *x_s;
    uint32_t *contacts;
    uint64_t *cmask;
    uint32_t *err;
    const uint32_t *nl;
    uint8_t *nl_slot, *slot_cand;
    const uint32_t *nl_cnt, *nl_pos, *cellc;
    const float *nl_gap, *nl_gmin, *nl_disp;
    uint32_t *flags;
    uint32_t cur;
};

struct ForcesTask {
    StepArgs a;
    const uint32_t *order;
    const double4 *x, *vel;
    const double3 *angvel;
    double3 *force, *torque;
    double *es_dot, *ev_dot, *es, *ev, *p;
    const double4 *w_nx, *w_mvfd;
    double *walls_force_pp;
    uint32_t *contacts;
    uint64_t *cmask;
    double *delta_t;
    uint8_t *nl_slot, *slot_cand;
    const float *nl_gap;
    float *nl_gmin;
    const uint32_t *err;
};

struct BondTask {
    StepArgs a;
    const uint2 *bonds;
    uint8_t *broken;
    double4 *bdelta, *bomega;
    const double4 *x, *vel;
    const double3 *angvel;
    double4 *bf, *bt;
    uint8_t *applied;
};

struct IntegrateTask {
    StepArgs a;
    const StepClock *clock;
    int change_velocity_state;
    double velocity_factor, tau_x;
    double4 *x, *vel;
    double3 *angvel, *force, *torque, *angpos, *acc, *angacc, *xyzsum;
    const double4 *w_mvfd;
    const double *tau_eff, *walls_force_pp;
    double *wall_partial;
    BondGather bg;
    NextHash nh;
};

struct WallsTask {
    StepArgs a;
    const StepClock *clock;
    double4 *w_nx, *w_mvfd;
    double *w_acc;
    const double *partial;
    uint32_t nparts;
};

struct ClockTask {
    StepClock *clock;
    double dt;
};

struct HashTask {
    StepArgs a;
    NLArgs n;
    const double4 *x;
    uint32_t *keys, *vals, *cellc;
    const double4 *x_ref;
    float *disp;
    uint32_t *flags;
    uint32_t cur, force;
    uint32_t *err;
};

struct BondGatherTask {
    uint32_t nbp;
    const uint32_t *plist, *off, *ent;
    const double4 *bf, *bt;
    const uint8_t *applied;
    double3 *force, *torque;
};

// Task records of one box for one step, as the single-box launchers form
// them (rebuild: the sort passes' key buffers as they are before the step).
RebuildTask rebuild_task(const StepArgs &a, const DemBuffers &b, uint32_t ncells, int bits,
                         const uint32_t *cond, uint32_t spin_limit);
NlBuildTask nl_build_task(const StepArgs &a, const NLArgs &n, const DemBuffers &b,
                          const uint32_t *cond);
TopoTask topology_task(const StepArgs &a, const DemBuffers &b, uint32_t cur);
ForcesTask forces_task(const StepArgs &a, const DemBuffers &b, bool nlist);
BondTask bond_task(const StepArgs &a, const DemBuffers &b);
IntegrateTask integrate_task(const StepArgs &a, const DemBuffers &b, int change_velocity_state,
                             double velocity_factor, double tau_x, bool gather_bonds,
                             const NextHash &nh);
WallsTask walls_task(const StepArgs &a, const DemBuffers &b);
HashTask hash_task(const StepArgs &a, const NLArgs &n, const DemBuffers &b, uint32_t cur,
                   uint32_t force);
BondGatherTask bond_gather_task(const DemBuffers &b);
// Blocks a box needs in each multi-box stage.
uint32_t topology_blocks(uint32_t np);
uint32_t forces_blocks(uint32_t np);
uint32_t nl_build_blocks(uint32_t np);
// Multi-box launches of `blocks` blocks (the last start plus its count,
// rounded up) over device task arrays; near/hertz pick the kernel variant
// (all boxes of one launch share it).
void launch_rebuild_multi(hipStream_t s, const RebuildTask *t, MultiMap m, uint32_t blocks);
void launch_nl_build_multi(hipStream_t s, const NlBuildTask *t, MultiMap m, uint32_t blocks);
void launch_topology_multi(hipStream_t s, const TopoTask *t, MultiMap m, uint32_t blocks,
                           bool near);
void launch_forces_multi(hipStream_t s, const ForcesTask *t, MultiMap m, uint32_t blocks,
                         bool hertz);
void launch_bonds_multi(hipStream_t s, const BondTask *t, MultiMap m, uint32_t blocks);
void launch_integrate_multi(hipStream_t s, const IntegrateTask *t, MultiMap m, uint32_t blocks);
void launch_hash_multi(hipStream_t s, const HashTask *t, MultiMap m, uint32_t blocks);
void launch_bond_gather_multi(hipStream_t s, const BondGatherTask *t, MultiMap m, uint32_t blocks);
// one block per task
void launch_walls_multi(hipStream_t s, const WallsTask *t, uint32_t n);
void launch_clock_multi(hipStream_t s, const ClockTask *t, uint32_t n);

// A max over all the blocks of a launch (the bit pattern of a non-negative
// double, or NaN above +inf) is kept in RED_LANES words RED_STRIDE words
// apart: block b folds its max into word b % RED_LANES with one atomic max,
// so that the atomics of a launch meet 64 addresses on separate 128-B lines
// rather than one; readers take the max of the 64 words.
#define RED_LANES 64
#define RED_STRIDE 16
#define RED_WORDS (RED_LANES * RED_STRIDE)
// words of DarcyBuffers::maxres: two alternating check slots (check slot k
// at k * RED_WORDS); DARCY_SCALE: the max |residual| slot of a correction;
// DARCY_STATE: solve state (ended, nijac of the last check, its max |norm|
// bits, corrections applied); DARCY_ZERO: always 0 (the state of sweeps run
// outside a solve)
#define DARCY_SCALE (2 * RED_WORDS)
#define DARCY_STATE (3 * RED_WORDS)
#define DARCY_ZERO (DARCY_STATE + 4)
#define DARCY_MAXRES_WORDS (DARCY_STATE + 8)

struct DarcyBuffers {
    double *p, *p_old, *p_new, *dp_expl, *phi, *dphi, *k;
    double *v, *vp_avg, *grad_p;  // 3 per cell
    int *p_constant;
    double4 *f_p;
    unsigned long long *maxres;  // DARCY_MAXRES_WORDS words, see above
    // constants of one solve's Jacobi sweeps: face conductances on the x, y
    // and z faces, then per interior cell (x fastest) the factor A of
    // div(k grad p), the forcing F and the base p_old + dp_expl/2 (at most
    // 9 nx*ny*nz doubles); and per cell whether dp is held at zero
    double *jc;
    uint8_t *jfrozen;
};

void launch_darcy_ghosts(hipStream_t s, const DarcyArgs &d, double *f, int ncomp);
void launch_darcy_ghosts_flux(hipStream_t s, const DarcyArgs &d, double *f, const double *k);
void launch_darcy_pressure_gradient(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b);
void launch_darcy_pressure_force(hipStream_t s, const StepArgs &a, const DarcyArgs &d,
                                 const DarcyBuffers &b, const DemBuffers &db);
void launch_darcy_porosities(hipStream_t s, const StepArgs &a, const DarcyArgs &d,
                             const DarcyBuffers &b, const DemBuffers &db, long iter);
void launch_darcy_porosities_to_edges(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b);
void launch_darcy_top_pressure(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b,
                               double new_pressure);
void launch_darcy_permeabilities(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b);
void launch_darcy_copy_interior(hipStream_t s, const DarcyArgs &d, const double *src, double *dst);
void launch_darcy_zero_interior(hipStream_t s, const DarcyArgs &d, double *f);
void launch_darcy_prepare(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b);
void launch_darcy_sweep(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b,
                        const double *p_in, double *p_out, unsigned long long *check,
                        unsigned long long *clear, const unsigned long long *prev,
                        double tolerance, uint32_t prev_nijac, bool free_running = false,
                        bool zero_scale = false);
void launch_darcy_residual_max(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, const double *p,
                               bool free_running);
// the network input of iterate p: FP32 rows of 3 (input) and/or the
// network's BF16 rows padded to 4 (in4); either may be null
void launch_darcy_net_input(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, const double *p,
                            float *input, bool free_running, uint16_t *in4 = nullptr);
void launch_darcy_net_apply(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b, double *p,
                            const float *e);
void launch_darcy_decide(hipStream_t s, const DarcyBuffers &b, const unsigned long long *slot,
                         double tolerance, uint32_t nijac);
void launch_darcy_velocities(hipStream_t s, const DarcyArgs &d, const DarcyBuffers &b);

// ------------------------------------------------------------------
// Darcy stages of the multi-box launches (ensembles of fluid boxes): one
// task record per box; the solve's two iterate buffers are b.p (where
// every solve starts and, after dm_normalize, ends) and b.p_new.
// ------------------------------------------------------------------
struct DarcyTask {
    StepArgs a;
    DarcyArgs d;
    DarcyBuffers b;
    const double4 *x;
    double3 *force;
    const uint32_t *cell_start, *cell_end;
    const double4 *x_s, *vel_s;
    uint32_t *err;
    const StepClock *clock;
    double tolerance;
    uint32_t maxiter;
    uint16_t *net_in4;    // the box's network input rows (BF16), null without a network
    const float *net_out; // the network's output for the box
};
// fields of DarcyTask::b that the ghost stage updates
#define DM_P 0
#define DM_GRAD_P 1
#define DM_PHI 2
#define DM_K 3
uint32_t darcy_cell_blocks(const DarcyArgs &d);          // 256-thread blocks over the cells
uint32_t darcy_sweep_blocks(const DarcyArgs &d);         // the sweep's blocks
uint32_t darcy_all_cell_blocks(const DarcyArgs &d);      // 256-thread blocks over cells and ghosts
void launch_dm_ghosts(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, int field);
void launch_dm_ghosts_flux(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_pressure_gradient(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_pressure_force(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_porosities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_por_edges(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_permeabilities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_solve_start(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_prepare(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
// clears each box's solve state and check slots (one block per box)
void launch_dm_reset(hipStream_t s, const DarcyTask *t, uint32_t n);
// sweep nijac of every box's solve (from b.p into b.p_new for even nijac,
// back for odd); correct: the correction that follows clears its scale slot
void launch_dm_sweep(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac,
                     bool correct);
void launch_dm_decide(hipStream_t s, const DarcyTask *t, uint32_t n, uint32_t nijac);
// the learned correction of the iterate sweep nijac wrote (boxes with a
// network): residual maximum, the network's input, and, after the forward
// pass, p += s e
void launch_dm_net_input(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac);
void launch_dm_net_apply(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks, uint32_t nijac);
// the solve state words of every box into out[4 k .. 4 k + 3]
void launch_dm_states(hipStream_t s, const DarcyTask *t, uint32_t n, unsigned long long *out);
// the converged iterate of a box whose solve ended in b.p_new copied into b.p
void launch_dm_normalize(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_finish(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
void launch_dm_velocities(hipStream_t s, const DarcyTask *t, MultiMap m, uint32_t blocks);
