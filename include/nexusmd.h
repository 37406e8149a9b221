/*
 * nexusmd.h -- C-ABI of the MI355X force-and-integrate engine for NexusMD.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes.  No C++
 * types, no exceptions and no torch types cross this boundary.  The library
 * (libnexusmd_hip.so) is built for gfx950 only.
 *
 * Reference interfaces replaced (NgFEP/NexusMD, paths from the repo root,
 * "core/" = "NexusMD/core/"):
 *   - BaseLine::Engine (core/baseline/headers/Engine.h:25-93) and the loop in
 *     Engine::RunSimulation (core/baseline/src/Engine.cpp:162-188);
 *   - Forces::AddHBond/AddHAngle/AddPTorsion/AddNonBondElectroPME
 *     (core/baseline/headers/Forces/Forces.h:21-24);
 *   - VerletIntegration::Advance (core/baseline/headers/VerletIntegration.h:17);
 *   - KineticEnergy::calculateKineticEnergy (core/baseline/headers/KineticEnergy.h:12);
 *   - obtainNeighborList (core/baseline/headers/Forces/NeighborList.h:62-69).
 * Units are the reference's: nm, ps, amu, kJ/mol, e.
 *
 * Threading: one context per host thread; calls on one context serialise on
 * its HIP stream.  A server handle is likewise single-threaded.
 */
#ifndef NEXUSMD_H
#define NEXUSMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int nx_status;
enum {
    NX_OK = 0,
    NX_ERR_ARG = 1,       /* invalid argument (index out of range, bad size, NULL) */
    NX_ERR_HIP = 2,       /* HIP runtime error */
    NX_ERR_CAPACITY = 3,  /* a device list overflowed its capacity */
    NX_ERR_STATE = 4,     /* call out of order (e.g. no topology yet) */
    NX_ERR_NOMEM = 5,     /* device allocation failed */
    NX_ERR_NUMERIC = 6    /* non-finite energy or position detected */
};

enum { NX_PRECISION_FP64 = 0, NX_PRECISION_MIXED = 1 };

/* Enable bits, one per TaskDispatcher::Enable*Force (core/dispatcher/headers/TaskDispatcher.h:22-25) */
enum { NX_FORCE_BOND = 1, NX_FORCE_ANGLE = 2, NX_FORCE_TORSION = 4, NX_FORCE_NONBONDED = 8, NX_FORCE_ALL = 15 };

/* Parameter records; same fields and meaning as BondParams / AngleParams /
 * PTorsionParams (core/common/headers/SystemXMLParser.h:10-27). */
typedef struct { int p1, p2; double d, k; } nx_bond;
typedef struct { int p1, p2, p3; double a, k; } nx_angle;
typedef struct { int p1, p2, p3, p4; double k, phase; int periodicity; } nx_torsion;

/* Nonbonded options.  cutoff and ewald_tolerance follow NonbondedParams
 * (SystemXMLParser.h:34-39); alpha = sqrt(-ln(2 tol)) / cutoff
 * (core/baseline/src/Forces/NonbondedForce.cpp:297-300). */
typedef struct {
    double cutoff;
    double ewald_tolerance;
    int grid[3];            /* PME grid; {0,0,0} = reference rule (next pow2 >= 10*ceil(L)) */
    int pme_order;          /* must be 5 (NonbondedForce.cpp:474) */
    int use_switch;         /* switching function (NonbondedForce.cpp:553-557, applied per pair) */
    double switch_distance;
    double skin;            /* Verlet buffer in nm; 0 = rebuild every step as the reference does */
} nx_nonbonded_opts;

typedef struct {
    int n_atoms;
    const double* mass;                               /* [n_atoms] amu */
    int n_bonds;    const nx_bond* bonds;
    int n_angles;   const nx_angle* angles;
    int n_torsions; const nx_torsion* torsions;
    const double* charge;                             /* [n_atoms] e */
    const double* sigma;                              /* [n_atoms] nm */
    const double* epsilon;                            /* [n_atoms] kJ/mol */
    const int* excl_offsets;                          /* CSR [n_atoms+1] or NULL: derive 1-2/1-3/1-4 from bonds */
    const int* excl_indices;                          /*  (Exclusions.cpp:7-44, bondCutoff = 3) */
    nx_nonbonded_opts nb;
} nx_topology;

/* Per-term energies, kJ/mol.  reference_pe is what the reference's
 * TotalEnergy report prints: with the nonbonded force enabled, the
 * reciprocal energy is assigned over the bonded sum (NonbondedForce.cpp:387). */
typedef struct {
    double bond, angle, torsion;
    double coulomb_direct, lj, recip, self, exclusion;
    double potential;       /* physical total of the enabled terms */
    double reference_pe;
    double kinetic;         /* half-step-shifted KE (KineticEnergy.cpp:21-41) */
} nx_energies;

typedef struct nx_ctx nx_ctx;

/* Context lifetime.  Replaces Engine::Engine (Engine.cpp:19-50). */
nx_status nx_create(int device, int precision, nx_ctx** out);
nx_status nx_destroy(nx_ctx* ctx);
/* Text of the context's last error; with NULL, the calling thread's last error
 * of a context-free call (nx_host_alloc, nx_host_free, nx_fft3d_r2c_noalloc). */
const char* nx_last_error(const nx_ctx* ctx);
const char* nx_status_string(nx_status s);

/* Topology upload (Engine::InitializeSimulationParameters, Engine.cpp:53-73,
 * plus Exclusions::createExclusions, Engine.cpp:171). */
nx_status nx_set_topology(nx_ctx* ctx, const nx_topology* top);

/* Box from lower/upper bounds (BoxInfo, core/common/headers/PeriodicBoundaryCondition.h:11-19). */
nx_status nx_set_box(nx_ctx* ctx, const double lb[3], const double ub[3]);
/* The reference's rule: bounding box of the positions widened by 1% per side
 * (PeriodicBoundaryCondition::extractBoxBoundaries, PeriodicBoundaryCondition.cpp:11-88). */
nx_status nx_box_from_positions(int n, const double* pos, double lb[3], double ub[3]);

/* State in/out, AoS double[n][3] as Coords3D.  vel may be NULL (zeros, as
 * Initializer::InitializeForcesAndVelocities, Initializer.cpp:18-24). */
nx_status nx_set_state(nx_ctx* ctx, const double* pos, const double* vel);
nx_status nx_get_state(nx_ctx* ctx, double* pos, double* vel, double* forces);

/* Neighbour-list state for checkpoints (no counterpart in the reference, which
 * rebuilds its list every step and keeps no state to restart from): the
 * positions the current device list was built from (AoS double[n][3]) and
 * whether a rebuild is due.  A context given a checkpoint's positions,
 * velocities and list state builds its list from the same positions as the
 * checkpointed run did, so the mixed path continues bit for bit.
 * nx_set_list_state must follow nx_set_state (which resets the list state);
 * the context is laid out again on the next call.  build_pos may be NULL in
 * nx_get_list_state. */
nx_status nx_get_list_state(nx_ctx* ctx, double* build_pos, int* rebuild_due);
nx_status nx_set_list_state(nx_ctx* ctx, const double* build_pos, int rebuild_due);

/* Forces and energies at the current positions (Engine::CalculateForces,
 * Engine.cpp:75-95).  forces may be NULL; e may be NULL. */
nx_status nx_compute_forces(nx_ctx* ctx, int force_mask, double* forces, nx_energies* e);

/* nsteps x [forces -> leapfrog] on the device (the loop of
 * Engine::RunSimulation without reports; VerletIntegration::Advance,
 * VerletIntegration.cpp:49-89). */
nx_status nx_step(nx_ctx* ctx, int nsteps, double dt, int force_mask);

/* One step whose report matches Engine::Report (Engine.cpp:117-158): forces
 * and energies at x(t), kinetic energy with the half-step shift, and x(t),
 * v(t-dt/2), F(t) copied out before the leapfrog update.  Outputs may be NULL. */
nx_status nx_step_report(nx_ctx* ctx, double dt, int force_mask, double* pos, double* vel, double* forces,
                         nx_energies* e);

/* nx_step_report without waiting (no counterpart in the reference, whose
 * reports stop its loop, Engine.cpp:117-158): enqueues the same report step
 * -- forces and energies at x(t), the half-step-shifted kinetic energy, x(t),
 * v(t-dt/2) and F(t) copied out, then the leapfrog update -- and returns at
 * once with a ticket.  The copies run on a stream of their own after the
 * report's kernels, so the update and later steps do not wait for them.  The
 * output buffers (any may be NULL) must stay valid and untouched until
 * nx_report_wait(ticket) returns; pinned host memory (nx_host_alloc) makes the
 * copies asynchronous, pageable memory makes this call wait for them.  At most
 * 4 reports may be in flight per context (NX_ERR_STATE beyond).  A re-layout
 * of the context (nx_set_topology, nx_set_box, nx_set_list_state) drops the
 * reports in flight. */
nx_status nx_step_report_async(nx_ctx* ctx, double dt, int force_mask, double* pos, double* vel, double* forces,
                               int64_t* ticket);
/* Waits until a ticket's outputs are in place; fills its energies (e may be NULL). */
nx_status nx_report_wait(nx_ctx* ctx, int64_t ticket, nx_energies* e);
/* Pinned (page-locked) host memory for the outputs of nx_step_report_async. */
nx_status nx_host_alloc(size_t bytes, void** out);
nx_status nx_host_free(void* p);

/* The explicit pair set of the current positions: (i, j), i > j, r^2 < rc^2,
 * not excluded -- the set obtainNeighborList emits (NeighborList.cpp:129-154).
 * Writes up to cap pairs; *count receives the full count. */
nx_status nx_neighbor_pairs(nx_ctx* ctx, int* pairs, int64_t cap, int64_t* count);

/* Kernel timing of the last nx_step/nx_compute_forces call, in milliseconds
 * per launch (average over launches in the call), from HIP events on the
 * stream each kernel runs on.  A step runs the bonded and exclusion terms and
 * PME on a second stream beside the pair kernel.  Modes:
 *   NX_TIMING_KERNELS     every kernel family launched directly, one after
 *                         the other on one stream, with events around it:
 *                         "nlist", "pack", "bonded", "atom_forces", "pair",
 *                         "pme_spread", "fft", "pme_gather", "integrate"
 *   NX_TIMING_GRAPH_PAIR  nx_step replays every step as graphs around the pair
 *                         kernel (list before it, terms and PME beside it,
 *                         update after it) with events on the main stream
 *                         around the pair kernel and after every step: "pair"
 *                         (the pair kernel as it runs beside PME) and "step";
 *                         around the graphs before the fork and after the
 *                         join: "list" (the list kernels) and "update"; and
 *                         on the second stream around the side graph: "side"
 *                         (terms and PME as they run beside the pair kernel)
 * Environment, read by nx_create: NEXUSMD_GRAPHS=0 makes every step run as
 * direct launches (NX_TIMING_GRAPH_PAIR then times as NX_TIMING_KERNELS);
 * NEXUSMD_OVERLAP=0 keeps every kernel on one stream; NEXUSMD_LITE=0 keeps
 * the list kernels in every step graph of a one-system context.
 * No counterpart in the reference (it has no GPU path to time). */
enum { NX_TIMING_OFF = 0, NX_TIMING_KERNELS = 1, NX_TIMING_GRAPH_PAIR = 2 };
nx_status nx_enable_timing(nx_ctx* ctx, int mode);
nx_status nx_kernel_time(nx_ctx* ctx, const char* name, double* ms_per_launch, int64_t* launches);

/* Launch shapes (no counterpart in the reference; its FFTW plan choice,
 * NonbondedForce.cpp:264-290, is the nearest analogue).  0 selects the
 * built-in rule.  No shape changes what any kernel computes for an atom, a
 * pair or a grid line, so every choice gives the same bits in the mixed path:
 *   pair_per_cu    persistent pair-kernel workgroups per CU (capped by occupancy)
 *   spread_block   PME spread tile side, in neighbour-list columns
 *   fft_xblocks    plane FFT: number of workgroups the x pass is split into
 *   fft_lines      line FFT: lines per workgroup */
typedef struct { int pair_per_cu, spread_block, fft_xblocks, fft_lines; } nx_tuning;
/* Shapes for the next layout of the context (taken at its next call). */
nx_status nx_set_tuning(nx_ctx* ctx, const nx_tuning* t);
nx_status nx_get_tuning(nx_ctx* ctx, nx_tuning* t);
/* Autotuner: times steps_per_trial steps (>= 4) for each candidate of each
 * shape in turn (coordinate descent), from the context's current state,
 * which is restored after every trial; keeps the fastest, applies it and
 * caches it under the system's key (precision, CU count, atoms, box, grid,
 * cutoff, buffer) for every later context of the process with that key --
 * and across processes when NEXUSMD_TUNE_CACHE names a file.  Reports the
 * chosen shapes and ms per step of the built-in shapes and of the choice. */
nx_status nx_autotune(nx_ctx* ctx, int steps_per_trial, nx_tuning* chosen, double* ms_default, double* ms_best);

/* Counters of the last call: pairs evaluated by the pair kernel (cluster
 * tiles x 64), pair-list entries, neighbour-list rebuilds.  Over the
 * context's life: steps run as the step graph without the list kernels
 * (one-system contexts; NEXUSMD_LITE=0 turns it off), and steps run again
 * because a step without the list kernels found a build due and froze. */
typedef struct {
    int64_t n_slots, n_clusters, n_super, list_entries, tiles, rebuilds, steps;
    int grid[3];
    double alpha;
    int64_t steps_without_list, steps_repeated;
} nx_stats;
nx_status nx_get_stats(nx_ctx* ctx, nx_stats* s);

/* Standalone 3D FFT through the engine's Stockham kernels, for verification:
 * real input [nx][ny][nz] (row-major) -> half spectrum [nx][ny][nz/2+1]
 * (interleaved complex, unnormalised, sign -1), and the inverse (sign +1)
 * back to real.  precision as nx_create. */
nx_status nx_fft3d_r2c(int device, int precision, int nx, int ny, int nz, const double* in, double* out);
nx_status nx_fft3d_c2r(int device, int precision, int nx, int ny, int nz, const double* in, double* out);

/* nx_fft3d_r2c without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code around graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * result and bits as nx_fft3d_r2c; every device buffer (grid, spectrum,
 * twiddles, work items) lives in a 20 MiB array that each GPU holds from the
 * library's load until the process ends.  Limit: every dimension <= 256 and
 * at most 2^20 grid points (e.g. 96^3, 64x128x128); larger grids return
 * NX_ERR_CAPACITY.  The array keeps the last call's contents; calls on one GPU
 * share it and the library holds a per-device lock through each call, so
 * calls from different threads run one after the other (anything else that
 * wrote to the array during a call would race with it).  Runs on the device's
 * null stream and returns after the result is in `out`.  The reason of a
 * failure is nx_last_error(NULL) on the calling thread. */
nx_status nx_fft3d_r2c_noalloc(int device, int precision, int nx, int ny, int nz, const double* in, double* out);

/* ---------------------------------------------------------------- serving
 * Continuous batching of many independent small systems on one GPU (no
 * counterpart in the reference).  Requests are admitted while their device
 * footprint fits under hbm_budget_bytes, advance together in batched
 * launches, and retire as they finish; new requests join at the next step
 * boundary. */
typedef struct nx_server nx_server;
typedef struct {
    int64_t submitted, admitted, finished, failed, active;
    double p50_step_ms, p99_step_ms;      /* per (request, step) latency: device time between the ends of
                                             successive batched steps (events on the batch stream); the
                                             first step after retirement, admission or a re-layout also
                                             counts the host time those took */
    double aggregate_ns_per_day;          /* sum of simulated time / wall time */
    double wall_s;
    int64_t batched_steps;                /* batched device steps executed */
    int64_t bytes_in_use, bytes_budget;
} nx_server_stats;

nx_status nx_server_create(int device, int precision, int64_t hbm_budget_bytes, int max_active, nx_server** out);
nx_status nx_server_destroy(nx_server* s);
const char* nx_server_last_error(const nx_server* s);
/* Copies topology and state; box from lb/ub.  *req receives the request id. */
nx_status nx_submit(nx_server* s, const nx_topology* top, const double lb[3], const double ub[3],
                    const double* pos, const double* vel, int nsteps, double dt, int64_t* req);
/* Runs batched steps until every active request has advanced `steps` more
 * steps or finished (steps <= 0: until all admitted work is done). */
nx_status nx_server_run(nx_server* s, int steps);
/* status: 0 queued, 1 running, 2 done, 3 failed */
nx_status nx_poll(nx_server* s, int64_t req, int* status, int* steps_done);
nx_status nx_fetch(nx_server* s, int64_t req, double* pos, double* vel, nx_energies* e);
nx_status nx_server_get_stats(nx_server* s, nx_server_stats* st);

#ifdef __cplusplus
}
#endif
#endif
