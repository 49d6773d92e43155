/*
 * orrery_lbm.h -- C ABI of the MI355X lattice Boltzmann library (liborrery_lbm.so).
 *
 * Drop-in boundary for the solver level of Orrery's aero LBM path: every entry point mirrors one
 * member of orrery::aero::lbm::Lbm3DSolver / Lbm2DSolver or of the factory make_gpu_solver,
 * which the reference binds to an IComputeBackend. Citations are reference-relative paths.
 *
 * Conventions
 *  - Every call returns LBM_OK (0) or an error code; lbm_last_error() gives the message of the
 *    last failing call on the calling thread. No C++ exception crosses this boundary.
 *  - Host arrays are owned by the caller and are copied (or filled) before the call returns,
 *    as IComputeBackend::upload/readback do (src/aero/compute/include/orrery/aero/compute/
 *    compute_backend.h:89-93).
 *  - Lattice layout: node (x, y, z) has index x + nx (y + ny z); populations are reported
 *    direction-major, f[i * cells + cell], with the weights added back (the layout of
 *    initial_populations, src/aero/lbm3d/include/orrery/aero/lbm3d/lbm3d.h:248-252).
 *  - One handle binds one HIP device and one stream; one thread may call into it at a time.
 */
#ifndef ORRERY_LBM_H
#define ORRERY_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
enum {
    LBM_OK = 0,
    LBM_EINVAL = 1,  /* refused configuration or argument (validate(), check_binary32, ...) */
    LBM_ENOMEM = 2,  /* refused before allocating: device memory (check_memory) */
    LBM_EDEVICE = 3, /* a HIP runtime call failed */
    LBM_ESTATE = 4   /* handle in a state that cannot serve the call */
};

/* ---- configuration (mirrors lbm3d.h:83-103, lbm2d.h:43-78) -------------------------------- */
enum { LBM_X_PERIODIC = 0, LBM_X_INFLOW_OUTFLOW = 1 };           /* XBoundary, lbm2d.h:43 */
enum { LBM_OUTLET_CONVECTIVE = 0, LBM_OUTLET_EXTRAPOLATION = 1 }; /* Outlet, lbm2d.h:50 */
enum { LBM_WALL_PERIODIC = 0, LBM_WALL_SLIP = 1, LBM_WALL_NO_SLIP = 2 }; /* WallKind, lbm2d.h:51 */
enum { LBM_SOLID_HALFWAY = 0, LBM_SOLID_INTERPOLATED = 1 };       /* SolidWall, lbm3d.h:161 */

/* Population storage of the device lattice. The collision always runs in binary32 except for
 * LBM_STORE_F64, the FP64 instantiation used as the on-device oracle. The 16-bit formats store
 * the populations shifted by their lattice weights (f_i - w_i), as the reference's binary32
 * buffers do (src/aero/lbm3d/lbm3d_gpu.cpp:1-7). */
enum { LBM_STORE_F32 = 0, LBM_STORE_F16S = 1, LBM_STORE_BF16S = 2, LBM_STORE_F64 = 3 };

typedef struct lbm_wall {
    int32_t kind;      /* LBM_WALL_* */
    double velocity_x; /* tangential wall speed along x, no-slip walls only (Wall, lbm2d.h:53-56) */
} lbm_wall;

typedef struct lbm3d_config { /* Lbm3DConfig, lbm3d.h:83-103 (+ Sponge :61-66) */
    uint64_t nx, ny, nz;
    double tau;
    int32_t x_boundary; /* LBM_X_* */
    int32_t outlet;     /* LBM_OUTLET_* */
    lbm_wall y_low, y_high, z_low, z_high;
    double inlet_ux, inlet_uy, inlet_uz;
    double body_force_x, body_force_y, body_force_z;
    int32_t smagorinsky_enabled;
    double smagorinsky_constant;
    double sponge_strength;
    uint64_t sponge_inlet, sponge_outlet, sponge_sides;
    int32_t solid_wall; /* LBM_SOLID_* */
} lbm3d_config;

typedef struct lbm2d_config { /* Lbm2DConfig, lbm2d.h:65-78 */
    uint64_t nx, ny;
    double tau;
    int32_t x_boundary;
    int32_t outlet;
    lbm_wall bottom, top;
    double inlet_ux, inlet_uy;
    double body_force_x, body_force_y;
    int32_t smagorinsky_enabled;
    double smagorinsky_constant;
} lbm2d_config;

typedef struct lbm3d lbm3d; /* opaque solver handle (one device lattice) */
typedef struct lbm2d lbm2d;

/* ---- library ------------------------------------------------------------------------------ */
/* Message of the last failing call on this thread ("" when none). */
const char* lbm_last_error(void);
/* Library version string. */
const char* lbm_version(void);
/* Free device memory and the largest single allocation, from hipMemGetInfo; replaces
 * IComputeBackend::memory() (compute_backend.h:94-129). A fake limit set by
 * lbm_set_memory_limit() caps `available` (the NullComputeBackend::set_memory analogue,
 * src/aero/compute/include/orrery/aero/compute/null_compute_backend.h:22-86). */
int lbm_device_memory(size_t* available, size_t* max_alloc);
/* 0 clears the fake limit. */
int lbm_set_memory_limit(size_t bytes);
/* Name, architecture and compute-unit count of the current device (the autotuner's cache key). */
int lbm_device_name(char* name, size_t capacity);

/* ---- D3Q19 -------------------------------------------------------------------------------- */
/* make_gpu_solver (src/aero/lbm3d/lbm3d_gpu.cpp:487-513): validate() -> 32-bit index bound ->
 * binary32 range -> boundary tables -> device-memory check -> allocate -> initialise with the
 * uniform inlet state. Refuses before allocating, with the reference's reasons. `distance`
 * (cells doubles) may be NULL unless solid_wall is interpolated. */
int lbm3d_create(const lbm3d_config* config, const uint8_t* solid, const double* distance,
                 int32_t storage, lbm3d** out);
/* ~GpuLbm3D: frees the device memory. NULL is accepted. */
int lbm3d_destroy(lbm3d* h);
/* Lbm3DSolver::initialize (lbm3d_gpu.cpp:298-331): equilibrium of the given state, solids at
 * rest; resets the step count and the inlet velocity. `moments` holds 4 planes of `cells`
 * doubles (rho, ux, uy, uz), or NULL for uniform_inlet_state (lbm3d_common.cpp:262-270). */
int lbm3d_initialize(lbm3d* h, const double* moments);
/* initialize() with density 1 and the inlet velocity (rest when x is periodic) plus a seeded
 * uniform perturbation of +-amplitude on each velocity component of every fluid node, drawn
 * from lbm_perturbation(seed, cell, component). The state is built on the device. */
int lbm3d_initialize_perturbed(lbm3d* h, uint64_t seed, double amplitude);
/* Lbm3DSolver::step (lbm3d_gpu.cpp:333-355): advances exactly `count` steps. */
int lbm3d_step(lbm3d* h, uint64_t count);
/* Lbm3DSolver::steps. */
uint64_t lbm3d_steps(const lbm3d* h);
/* set_inlet_velocity / inlet_velocity (lbm3d_gpu.cpp:268-289, check lbm3d_common.cpp:163-174). */
int lbm3d_set_inlet(lbm3d* h, const double velocity[3]);
int lbm3d_get_inlet(const lbm3d* h, double velocity[3]);
/* Lbm3DSolver::force (lbm3d_cpu.cpp:192-203, lbm3d_common.cpp:213-241): momentum exchange
 * over the boundary links, moment about the lattice origin at the wall points; summed in
 * FP64 on the device in a fixed order (deterministic). */
int lbm3d_force(lbm3d* h, double force[3], double moment[3]);
/* Lbm3DSolver::mass (lbm3d_cpu.cpp:205-215): sum of the populations over fluid nodes, FP64. */
int lbm3d_mass(lbm3d* h, double* mass);
/* Lbm3DSolver::macroscopic (lbm3d_cpu.cpp:154-186): density and velocity (Guo half-force
 * corrected) per node; solids report 1 and 0. Any pointer may be NULL. */
int lbm3d_macroscopic(lbm3d* h, double* rho, double* ux, double* uy, double* uz);
/* macroscopic() at n listed nodes only, computed on the device: 4 n doubles cross the bus
 * instead of 4 cells (the adapter's wall-pressure samples and inlet-plane density). Any output
 * may be NULL; values are those lbm3d_macroscopic reports at the same nodes. */
int lbm3d_sample(lbm3d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy, double* uz);
/* lbm3d_sample without allocations. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid hipMalloc). Its node list,
 * boundary slots and outputs live in a fixed device array of the library (2.5 MiB per device,
 * held until the process ends) instead of a per-call buffer; at most 65536 nodes per call
 * (LBM_EINVAL beyond, before any device work). Same values as lbm3d_sample bit for bit. Calls
 * on one device share its array: they are serialised within the process, and return once the
 * outputs are on the host. */
int lbm3d_sample_fixed(lbm3d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy,
                       double* uz);
/* max_speed(macroscopic()) (lbm3d_common.cpp:200-211), reduced on the device. */
int lbm3d_max_speed(lbm3d* h, double* speed);
/* Lbm3DSolver::populations (lbm3d_gpu.cpp:394-405): 19 * cells doubles, logical layout. */
int lbm3d_populations(lbm3d* h, double* f);
/* Lbm3DSolver::links (boundary_links, lbm3d_common.cpp:176-198): writes up to `capacity`
 * links and sets *count to the total; any output pointer may be NULL. */
int lbm3d_links(const lbm3d* h, uint32_t* cell, uint32_t* direction, double* q, size_t capacity,
                size_t* count);
/* Device bytes held by the handle (populations, node classes, boundary tables, history). */
int lbm3d_device_bytes(const lbm3d* h, size_t* bytes);
/* Blocks until the handle's stream has drained (the reference's finish()). */
int lbm3d_synchronize(lbm3d* h);
/* Launch shape of the bulk kernel: block size (64..256, multiple of 64) and cells per thread
 * (1, 2 or 4 consecutive cells, each direction of the run moved as one access). 0 keeps the
 * current value. Used by the autotuner. */
int lbm3d_set_launch(lbm3d* h, int32_t block, int32_t cells_per_thread);
/* Cells per thread of the bulk kernel for the A step (odd step count) and the B step (even)
 * separately (1, 2 or 4; 0 keeps the current value); lbm3d_set_launch sets both. No result
 * changes. */
int lbm3d_set_cells_per_thread(lbm3d* h, int32_t a_step, int32_t b_step);
int lbm3d_get_cells_per_thread(const lbm3d* h, int32_t* a_step, int32_t* b_step);
/* How a step is launched: LBM_SCHEDULE_SERIAL (default) runs the bulk kernel then the
 * boundary kernel on the handle's stream; LBM_SCHEDULE_STREAMS runs the two concurrently on
 * two streams (the second is created when this schedule is first selected); LBM_SCHEDULE_FUSED
 * runs bulk and boundary nodes in one kernel in cell order.
 * The launch shape of lbm3d_set_launch applies to the bulk kernel of the first two. All three
 * give the same results bit for bit. */
enum { LBM_SCHEDULE_FUSED = 0, LBM_SCHEDULE_SERIAL = 1, LBM_SCHEDULE_STREAMS = 2 };
int lbm3d_set_schedule(lbm3d* h, int32_t schedule);
/* Whether lbm3d_step replays captured HIP graphs (on by default; no reference counterpart). A
 * call of at least 8 steps under the serial or fused schedule runs them as replays of one graph
 * of up to 64 steps, captured again when the inlet, schedule or launch shape changes; the
 * results are those of the same steps launched one by one, bit for bit. ORRERY_LBM_GRAPHS=0 in the
 * environment makes off the default for handles created afterwards. */
int lbm3d_set_graphs(lbm3d* h, int32_t enabled);
/* The schedule and bulk launch shape lbm3d_step uses. */
int lbm3d_get_launch(const lbm3d* h, int32_t* schedule, int32_t* block, int32_t* cells_per_thread);
/* Cap on the workgroups of the bulk kernel that share a compute unit, for the A step (odd step
 * counts) and the B step (even), 1 to 32, 0 for none (the default): each workgroup reserves LDS
 * it does not use, so that no more than that many fit (with 256-thread workgroups, the cap is
 * waves per SIMD). Fewer resident waves stream the 19 population directions with fewer DRAM
 * conflicts; results do not change. Refused when the device's LDS cannot express the cap. No
 * reference counterpart; the autotuner sets it. */
int lbm3d_set_occupancy(lbm3d* h, int32_t groups_a, int32_t groups_b);
int lbm3d_get_occupancy(const lbm3d* h, int32_t* groups_a, int32_t* groups_b);
/* One candidate of the autotuner: a schedule with a bulk launch shape, and its measured time. */
typedef struct lbm_launch_choice {
    int32_t schedule;          /* LBM_SCHEDULE_* */
    int32_t block;             /* threads per block of the bulk kernel */
    int32_t cells_per_thread;  /* consecutive cells per thread of the bulk kernel */
    int32_t groups_a, groups_b; /* occupancy caps of lbm3d_set_occupancy (0: none) */
    double ms_per_step;
} lbm_launch_choice;
/* Autotuner (no reference counterpart): times `steps` steps (rounded up to even) under every
 * bulk launch shape (block 64/128/192/256 x 1/2/4 cells per thread) of the serial schedule, and
 * the fused schedule unless serial_only; under the fastest serial shape it then tries caps of
 * 7, 6, 5, 4 and 3 workgroups per compute unit on the A step, then on the B step, then on the A step
 * again (lbm3d_set_occupancy); the four fastest, and the default launch if it is not among
 * them, are timed once more, and the fastest of that final round is kept on the handle. The lattice is
 * copied on the device first and restored afterwards, with its step count: its later results
 * are those it would have had without tuning, bit for bit (all schedules and shapes compute
 * the same values). Needs device memory for that copy (LBM_ENOMEM otherwise). `tried` (capacity
 * entries, may be NULL) receives every candidate in the order timed; *count their number. */
int lbm3d_autotune(lbm3d* h, uint32_t steps, int32_t serial_only, lbm_launch_choice* tried, size_t capacity,
                   size_t* count, lbm_launch_choice* best);
/* Bytes one step of this lattice moves by the algorithmic count: 2 * 19 * bytes per stored
 * population, times nx * ny * nz (the reference's MLUPS convention). */
int lbm3d_step_bytes(const lbm3d* h, double* bytes);
/* Node counts of the lattice: bulk (plain streaming), boundary, solid, and boundary links, as the
 * reference classifies them (classify_nodes, lbm3d_internal.h:219-248). */
int lbm3d_counts(const lbm3d* h, size_t* bulk, size_t* boundary, size_t* solid, size_t* links);
/* No counterpart in the reference (its kernels do not split the boundary nodes): how this
 * library's step kernels divide the fluid nodes. `bulk` nodes (with the nodes of the velocity
 * inlet face off the y and z faces and away from solids) are streamed in place by the bulk
 * kernel; `side` nodes (the boundary nodes on slip or periodic y and z faces, away from the x
 * faces and from solids) by the bulk kernel in A steps and the side kernel in B steps; `listed`
 * nodes by the boundary kernel, through source codes and the stash. bulk + side + listed is the
 * number of fluid nodes. */
int lbm3d_stream_counts(const lbm3d* h, size_t* bulk, size_t* side, size_t* listed);
/* Advances `count` steps and times them with HIP events on the handle's stream (the
 * reference's bench protocol, tests/bench/aero/bench_lbm_gpu.cpp:54-64). With per_kernel = 0,
 * one event pair brackets all steps (*step_ms); with per_kernel = 1, event pairs also bracket
 * the bulk kernel and the rest of each step (side and boundary kernels), giving their summed
 * durations in *bulk_ms and *boundary_ms. */
int lbm3d_time_steps(lbm3d* h, uint64_t count, int32_t per_kernel, double* step_ms, double* bulk_ms,
                     double* boundary_ms);
/* Advances `count` steps with HIP events around each kernel on the handle's stream: ms[0], ms[1]
 * the bulk kernel's mean duration in A and B steps, ms[2], ms[3] the side kernel's (it runs in B
 * steps only: ms[2] = 0), ms[4], ms[5] the boundary kernel's. */
int lbm3d_time_kernels(lbm3d* h, uint64_t count, double ms[6]);

/* ---- D2Q9 (same shape; lbm2d.h:173-211, lbm2d_gpu.cpp:68-378) ---------------------------- */
int lbm2d_create(const lbm2d_config* config, const uint8_t* solid, int32_t storage, lbm2d** out);
int lbm2d_destroy(lbm2d* h);
int lbm2d_initialize(lbm2d* h, const double* moments /* 3 planes: rho, ux, uy, or NULL */);
int lbm2d_initialize_perturbed(lbm2d* h, uint64_t seed, double amplitude);
int lbm2d_step(lbm2d* h, uint64_t count);
uint64_t lbm2d_steps(const lbm2d* h);
/* As lbm3d_set_graphs, for lbm2d_step. */
int lbm2d_set_graphs(lbm2d* h, int32_t enabled);
/* LBM_SCHEDULE_FUSED (default): bulk cells and boundary nodes of a step in one launch;
 * LBM_SCHEDULE_SERIAL: the bulk kernel, then the boundary kernel. Same results bit for bit. */
int lbm2d_set_schedule(lbm2d* h, int32_t schedule);
int lbm2d_force(lbm2d* h, double* fx, double* fy, double* moment);
int lbm2d_mass(lbm2d* h, double* mass);
int lbm2d_macroscopic(lbm2d* h, double* rho, double* ux, double* uy);
/* Moments of n listed nodes (cell indices), the values lbm2d_macroscopic reports there; any
 * output may be NULL. Used for wall-pressure samples without reading the whole field back. */
int lbm2d_sample(lbm2d* h, const uint32_t* cells, size_t n, double* rho, double* ux, double* uy);
/* max_speed (lbm2d_common.cpp:168-176): the largest hypot(ux, uy) over the fluid nodes. */
int lbm2d_max_speed(lbm2d* h, double* speed);
int lbm2d_populations(lbm2d* h, double* f);
int lbm2d_links(const lbm2d* h, uint32_t* cell, uint32_t* direction, size_t capacity,
                size_t* count);
int lbm2d_synchronize(lbm2d* h);

/* ---- batched lattices (no reference counterpart: Orrery runs one lattice per solver) --------- */
/* Many resident D3Q19 lattices of uneven size advanced together: each launch sweeps every
 * lattice of one storage format, and `sample_every` steps plus a force sample per lattice are
 * captured once into a hipGraph and replayed. Lattices are admitted against an HBM budget and
 * keep their own step count, inlet and history; a lattice's results do not depend on the
 * other lattices in the batch. */
typedef struct lbm_batch lbm_batch;
/* budget_bytes 0 means no budget; sample_every >= 1 steps between force samples. */
int lbm_batch_create(size_t budget_bytes, uint32_t sample_every, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
/* Device bytes lbm_batch_add / lbm3d_create would allocate for this lattice. */
int lbm_batch_required_bytes(const lbm3d_config* config, const uint8_t* solid, const double* distance,
                             int32_t storage, size_t* bytes);
/* Admits a lattice (LBM_ENOMEM before allocating when it exceeds the free budget), started at
 * the uniform inlet state, perturbed by lbm_perturbation(seed, ...) * amplitude when
 * amplitude != 0. */
int lbm_batch_add(lbm_batch* b, const lbm3d_config* config, const uint8_t* solid, const double* distance,
                  int32_t storage, uint64_t seed, double amplitude, int32_t* id);
int lbm_batch_remove(lbm_batch* b, int32_t id);
/* Advances every resident lattice by `steps` (a multiple of sample_every), sampling the force
 * of each every sample_every steps; returns when the device is done. */
int lbm_batch_run(lbm_batch* b, uint64_t steps);
int lbm_batch_count(const lbm_batch* b, size_t* count);
/* The lattice's handle (owned by the batch): lbm3d_force/_macroscopic/_populations/_steps
 * work on it between runs. The calls that would step, re-initialise, retarget or free it
 * (lbm3d_step, _initialize*, _set_inlet, _time_steps, _autotune, _destroy) refuse it with
 * LBM_EINVAL: the batch keeps its A-A parity on the device and steps it alone. */
int lbm_batch_lattice(lbm_batch* b, int32_t id, lbm3d** h);
/* Force samples of a lattice: time (step count) and 6 doubles each (force xyz, moment xyz). */
int lbm_batch_history(lbm_batch* b, int32_t id, double* time, double* force, size_t capacity,
                      size_t* count);
int lbm_batch_bytes(const lbm_batch* b, size_t* used, size_t* budget);
/* Launch shape of the bulk kernel for the lattices of one storage format (autotuner). */
int lbm_batch_set_launch(lbm_batch* b, int32_t storage, int32_t block, int32_t cells_per_thread);
/* Occupancy cap of the bulk launch of one storage group (lbm3d_set_occupancy, one cap for both
 * step parities, since one launch advances lattices on either): 0 (no cap) to 32 workgroups per
 * compute unit, refused if the device's LDS cannot express it. Results do not change. The
 * autotuner leaves batches uncapped (a cap ran a batch of uneven lattices 2% slower). */
int lbm_batch_set_occupancy(lbm_batch* b, int32_t storage, int32_t groups_per_cu);
/* Autotunes the bulk launch shape of each storage format present in the batch on a scratch
 * copy of its largest lattice (lbm3d_autotune, serial schedule; the batch's lattices are not
 * advanced) and applies it with lbm_batch_set_launch, uncapped (lbm_batch_set_occupancy 0).
 * best[storage] (4 entries, indexed by LBM_STORE_*) receives the choice as applied (groups 0),
 * block 0 for a storage the batch does not hold. */
int lbm_batch_autotune(lbm_batch* b, uint32_t steps, lbm_launch_choice* best);
/* set_inlet_velocity of one lattice of the batch. */
int lbm_batch_set_inlet(lbm_batch* b, int32_t id, const double velocity[3]);
/* Per-kernel times of a batch run (HIP events on the batch's stream), indexed by LBM_STORE_*
 * where per storage group. */
typedef struct lbm_batch_timing {
    double step_ms;          /* wall time of the run / steps (force samples included) */
    double bulk_ms[4];       /* bulk and side kernels of the group, mean per step (0: no such group) */
    double boundary_ms[4];   /* boundary kernel of the group, mean per step */
    double force_ms;         /* force kernels of every group, mean per sample */
    uint64_t bulk_nodes[4];  /* bulk and side nodes the group's bulk launch updates */
    uint64_t boundary_nodes[4]; /* listed boundary nodes the group's boundary launch updates */
} lbm_batch_timing;
/* lbm_batch_run with every kernel launched on its own (no graph replay) between HIP events:
 * the same kernels in the same order, so the lattices and their histories advance exactly as
 * lbm_batch_run would advance them. */
int lbm_batch_run_timed(lbm_batch* b, uint64_t steps, lbm_batch_timing* timing);

/* ---- host-only setup (no device needed) --------------------------------------------------- */
/* validate() (src/aero/lbm3d/lbm3d_common.cpp:159-161): LBM_OK, or LBM_EINVAL with the
 * reference's reason in lbm_last_error(). */
int lbm3d_validate(const lbm3d_config* config, const uint8_t* solid, const double* distance);
int lbm2d_validate(const lbm2d_config* config, const uint8_t* solid);
/* Node classes (0 solid, 1 bulk, 2 boundary; classify_nodes, lbm3d_internal.h:219-248). */
int lbm3d_classify(const lbm3d_config* config, const uint8_t* solid, uint8_t* classes);
/* boundary_links (lbm3d_common.cpp:176-198, lbm2d_common.cpp:79-97) of a validated
 * configuration, in the reference's order, as lbm3d_links/lbm2d_links report them. */
int lbm3d_boundary_links(const lbm3d_config* config, const uint8_t* solid, const double* distance,
                         uint32_t* cell, uint32_t* direction, double* q, size_t capacity, size_t* count);
int lbm2d_boundary_links(const lbm2d_config* config, const uint8_t* solid, uint32_t* cell,
                         uint32_t* direction, size_t capacity, size_t* count);

/* ---- force-coefficient history (src/aero/lbm/include/orrery/aero/lbm/force_history.h) ---- */
typedef struct lbm_oscillation { /* Oscillation, force_history.h:48-54 */
    double period;
    double frequency;
    size_t cycles;
    double amplitude;
    double period_spread;
} lbm_oscillation;
/* force_coefficients (force_history.cpp:9-13): C = F / (U^2 L / 2), lattice units. */
void lbm_force_coefficients(double force_x, double force_y, double speed, double length, double* drag,
                            double* lift);
/* mean_since (force_history.cpp:36-47). */
double lbm_mean_since(const double* time, const double* values, size_t n, double from);
/* dominant_oscillation (force_history.cpp:49-100): 1 and *out set when at least two complete
 * periods of hysteresis mean crossings are found, else 0. */
int lbm_dominant_oscillation(const double* time, const double* values, size_t n, double from,
                             lbm_oscillation* out);
/* steady_since (force_history.cpp:149-157) with SettledCriteria::steady_tolerance. */
int lbm_steady_since(const double* time, const double* drag, const double* lift, size_t n, double from,
                     double steady_tolerance);
/* settled_oscillation (force_history.cpp:159-193). */
int lbm_settled_oscillation(const double* time, const double* drag, const double* lift, size_t n,
                            double from, double max_period_spread, double stationarity,
                            lbm_oscillation* out);
/* forces_settled_3d (src/aero/lbm3d/lbm_aero_solver3d.cpp:60-83): 1 when drag, lift and side
 * force are each steady over the samples at time >= from (at least two samples, width within
 * steady_tolerance x |mean drag|) or settled periodic shedding (lbm_settled_oscillation of the
 * drag with that component), else 0. */
int lbm_forces_settled(const double* time, const double* drag, const double* lift, const double* side, size_t n,
                       double from, double steady_tolerance, double max_period_spread, double stationarity);

/* ---- geometry: meshes, voxelisation, signed distance (src/geom) --------------------------- */
/* Host-only, deterministic. Winding numbers follow MeshTree (winding_number.cpp: Barill et al.
 * 2018 dipole far field, beta 2, leaves of 8 triangles); masks and distances are bit-identical
 * to the reference's for the same mesh and grid. Errors: lbm_geom_last_error(). */
typedef struct lbm_mesh lbm_mesh;
const char* lbm_geom_last_error(void);
/* TriangleMesh from vertices [nv][3] and triangles [nt][3] (triangle_mesh.h). */
int lbm_mesh_create(const double* vertices, size_t nv, const uint32_t* triangles, size_t nt, lbm_mesh** out);
int lbm_mesh_destroy(lbm_mesh* m);
int lbm_mesh_size(const lbm_mesh* m, size_t* nv, size_t* nt);
int lbm_mesh_data(const lbm_mesh* m, double* vertices, uint32_t* triangles);
/* RevolutionProfile::mesh (primitives.cpp:312-379) of sphere_profile (kind 0: a radius),
 * cylinder_profile (1: a radius, b height), cone_profile (2: a radius, b height) or
 * capsule_profile (3: a max diameter, b nose radius, c shoulder radius, d afterbody half
 * angle, e top diameter), with RevolutionResolution{around, arc_step}. */
int lbm_mesh_revolution(int32_t kind, double a, double b, double c, double d, double e, int32_t around,
                        double arc_step, lbm_mesh** out);
/* Wedge::mesh (primitives.cpp:590-604). */
int lbm_mesh_wedge(double length, double half_angle, double span, lbm_mesh** out);
/* box_mesh (primitives.cpp:557-571), centred on the origin. */
int lbm_mesh_box(double lx, double ly, double lz, lbm_mesh** out);
/* naca_wing (naca.cpp:184-257) of Naca4{max_camber, camber_position, thickness}. */
int lbm_mesh_naca_wing(double max_camber, double camber_position, double thickness, double chord, double span,
                       int32_t n, int32_t span_divisions, lbm_mesh** out);
/* Naca4::section (src/geom/naca.cpp:155-172): 2 n points (x, z) per unit chord into points[4 n]. */
int lbm_naca_section(double max_camber, double camber_position, double thickness, int32_t n, double* points);
/* polygon_mask (src/aero/lbm/solid_mask.cpp:45-73): node (x, y) solid when inside the polygon of
 * n points (x, z) in lattice units (even-odd rule); solid [nx ny]. */
int lbm_polygon_mask(uint64_t nx, uint64_t ny, const double* points, size_t n, uint8_t* solid);
/* v <- R (v - pivot) + translation, R row-major. */
int lbm_mesh_transform(lbm_mesh* m, const double rotation[9], const double pivot[3], const double translation[3]);
int lbm_mesh_bounds(const lbm_mesh* m, double lo[3], double hi[3]);
/* voxelize (voxel_grid.cpp:81-102): cell (i,j,k) at index i + nx (j + ny k), centre
 * origin + (i + 1/2, j + 1/2, k + 1/2) spacing; solid where winding >= threshold. */
int lbm_voxelize(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny, uint64_t nz,
                 double threshold, uint8_t* solid, double* winding);
/* signed_distance_band (voxel_grid.cpp:104-138): distance to the surface, exact within
 * band_cells * spacing, negative in solid cells; in_band optional. */
int lbm_signed_distance_band(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny,
                             uint64_t nz, const uint8_t* solid, double band_cells, double* distance,
                             uint8_t* in_band);
/* MeshTree::winding_number at n points [n][3]. */
int lbm_winding_numbers(const lbm_mesh* m, const double* points, size_t n, double* out);
/* lbm_voxelize on the device (no reference counterpart: the reference voxelises on the host):
 * one thread per cell walks the same tree. The mask equals lbm_voxelize's bit for bit: cells
 * whose device winding number lies within 1e-9 of the threshold are decided again by the host
 * evaluation (*rechecked, optional, counts them); elsewhere device and host numbers differ by
 * a few ulps of atan2 at most (`winding`, optional, holds the numbers that decided). */
int lbm_voxelize_device(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny,
                        uint64_t nz, double threshold, uint8_t* solid, double* winding, size_t* rechecked);
/* lbm_signed_distance_band on the device: bit for bit the host's (additions, products,
 * divisions, square roots and comparisons in the host's order, all correctly rounded). */
int lbm_signed_distance_band_device(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx,
                                    uint64_t ny, uint64_t nz, const uint8_t* solid, double band_cells,
                                    double* distance, uint8_t* in_band);

/* ---- the aerodynamic adapter (LbmAeroSolver3D, src/aero/lbm3d/lbm_aero_solver3d.h:45-166) ---- */
/* A closed body mesh in SI units (body axes, metres) goes in; coefficients, SI force and moment,
 * the force history, wall-pressure samples and validity warnings come out. The lattice is
 * wind-aligned (x along the free stream, y the side force, z the lift); the body is voxelised
 * with the geometry above and placed `upstream` reference lengths behind a velocity inlet and
 * `downstream` ahead of a convective outlet, in a square channel `width` lengths wide.
 * Errors: lbm_aero3d_last_error(). */
enum { LBM_LES_AUTOMATIC = 0, LBM_LES_ON = 1, LBM_LES_OFF = 2 }; /* LesMode, lbm_aero_solver.h:32-36 */

typedef struct lbm_aero3d_options { /* LbmSolver3DOptions, lbm_aero_solver3d.h:45-81 */
    double cells_per_length, lattice_velocity; /* LatticeResolution{16, 0.05} */
    double upstream, downstream, width;        /* reference lengths: 5, 10, 6 */
    int32_t sides;                             /* LBM_WALL_SLIP or LBM_WALL_PERIODIC */
    int32_t wall;                              /* LBM_SOLID_INTERPOLATED (default) or _HALFWAY */
    int32_t les;                               /* LBM_LES_* */
    double smagorinsky_constant;               /* 0.1 */
    double duration, average_window, start_ramp; /* convective times: 30, 8, 3 */
    double sponge_strength, sponge_inlet, sponge_outlet, sponge_sides; /* 0.1; 0, 2, 0 lengths */
    uint64_t steps_per_call, sample_every;     /* 200, 10 */
    uint64_t max_cells;                        /* 0: the device memory, else 2^24 */
    int32_t storage;                           /* LBM_STORE_* of the lattice (no reference counterpart) */
} lbm_aero3d_options;

typedef struct lbm_flow_conditions { /* FlowConditions, src/aero/include/orrery/aero/flow_conditions.h */
    double speed, density, viscosity, temperature, angle_of_attack, sideslip, gamma, gas_constant;
} lbm_flow_conditions;

typedef struct lbm_reference_values { /* ReferenceValues, flow_conditions.h */
    double area, length, moment_centre[3];
} lbm_reference_values;

typedef struct lbm_aero3d_result { /* the scalars of AeroResult, src/aero/include/orrery/aero/aero_result.h */
    double cd, cl, cy, cm;
    double force[3], moment[3]; /* N and N m, body axes, moment about the moment centre */
    uint64_t iterations;
    int32_t converged;
    double residual;            /* last relative change of cd */
    size_t samples, wall_samples, warnings;
} lbm_aero3d_result;

typedef struct lbm_aero3d_info { /* what setup() decided */
    lbm3d_config config;
    double origin[3];  /* wind-frame position (m) of lattice node (0, 0, 0) */
    double centre[3];  /* moment centre, lattice coordinates */
    uint64_t first[3], count[3]; /* voxelised box inside the lattice */
    double axes[9];    /* wind axes (drag, side, lift) in body coordinates, row by row */
    double dx, dt, reynolds, lattice_velocity, cells_per_length;
    uint64_t total_steps;
    size_t window_samples;
    double ramp_steps, blockage, surface_area;
    size_t solid_nodes;
} lbm_aero3d_info;

typedef struct lbm_aero3d lbm_aero3d;
const char* lbm_aero3d_last_error(void);
void lbm_aero3d_default_options(lbm_aero3d_options* options);
void lbm_default_flow_conditions(lbm_flow_conditions* conditions);
int lbm_aero3d_create(const lbm_aero3d_options* options, lbm_aero3d** out);
int lbm_aero3d_destroy(lbm_aero3d* a);
/* setup (lbm_aero_solver3d.cpp:85-339) without the device lattice: the lattice configuration,
 * mask and warnings only, on the host. */
int lbm_aero3d_plan(lbm_aero3d* a, const lbm_mesh* mesh, const lbm_reference_values* reference,
                    const lbm_flow_conditions* conditions);
/* setup: plan, then the device lattice (lbm3d_create) and the gradual start. */
int lbm_aero3d_setup(lbm_aero3d* a, const lbm_mesh* mesh, const lbm_reference_values* reference,
                     const lbm_flow_conditions* conditions);
/* step (lbm_aero_solver3d.cpp:342-411): *finished = 1 once the run is complete. */
int lbm_aero3d_step(lbm_aero3d* a, int32_t* finished);
int lbm_aero3d_get_result(const lbm_aero3d* a, lbm_aero3d_result* out);
/* Force history: sample time (lattice steps), drag, lift and side-force coefficients. */
int lbm_aero3d_history(const lbm_aero3d* a, double* time, double* drag, double* lift, double* side, size_t capacity,
                       size_t* count);
int lbm_aero3d_residuals(const lbm_aero3d* a, double* residuals, size_t capacity, size_t* count);
/* Coefficients (drag, lift, side) of the last sample. */
int lbm_aero3d_instantaneous(const lbm_aero3d* a, double coefficients[3]);
/* Wall samples (SurfaceField): points and normals [n][3], body axes; areas; cp. */
int lbm_aero3d_surface(const lbm_aero3d* a, double* points, double* normals, double* areas, double* cp,
                       size_t capacity, size_t* count);
int lbm_aero3d_warning_count(const lbm_aero3d* a, size_t* count);
int lbm_aero3d_warning(const lbm_aero3d* a, size_t index, char* code, size_t code_capacity, char* message,
                       size_t message_capacity);
int lbm_aero3d_get_info(const lbm_aero3d* a, lbm_aero3d_info* out);
/* The lattice mask of the plan: solid [cells]; distance [cells] for the interpolated wall. */
int lbm_aero3d_mask(const lbm_aero3d* a, uint8_t* solid, double* distance);
/* The device lattice (borrowed; valid until the next setup or destroy). */
int lbm_aero3d_lattice(const lbm_aero3d* a, lbm3d** out);
/* The SI velocity field [cells][3] on the lattice, on request (zero angle of attack and
 * sideslip only). */
int lbm_aero3d_velocity(lbm_aero3d* a, double* values);

/* ---- the 2D aerodynamic adapter (LbmAeroSolver2D, src/aero/lbm/lbm_aero_solver.h:38-106) ---- */
/* A closed section (x, z) in SI units (metres, per unit span) goes in; coefficients, SI force and
 * pitching moment, the force history, the Strouhal number, wall-pressure samples and validity
 * warnings come out. The section is rotated by the angle of attack so the free stream runs along
 * lattice x, rasterised by lbm_polygon_mask and placed `upstream` reference lengths behind a
 * velocity inlet and `downstream` ahead of a convective outlet, in a free-slip channel `height`
 * lengths high. Results use lbm_aero3d_result (cy = 0; moment about body y).
 * Errors: lbm_aero2d_last_error(). */
typedef struct lbm_aero2d_options { /* LbmSolverOptions, lbm_aero_solver.h:38-52 */
    double cells_per_length, lattice_velocity; /* LatticeResolution{40, 0.05} */
    double upstream, downstream, height;       /* reference lengths: 8, 16, 16 */
    int32_t les;                               /* LBM_LES_* */
    double smagorinsky_constant;               /* 0.1 */
    double duration, average_window;           /* convective times: 60, 6 */
    uint64_t steps_per_call, sample_every;     /* 500, 10 */
    int32_t storage;                           /* LBM_STORE_* of the lattice (no reference counterpart) */
} lbm_aero2d_options;

typedef struct lbm_aero2d_info { /* what setup() decided */
    lbm2d_config config;
    double origin[2];  /* wind-frame position (x, z in m) of lattice node (0, 0) */
    double centre[2];  /* moment centre, lattice coordinates */
    double dx, dt, reynolds, lattice_velocity, cells_per_length;
    uint64_t total_steps;
    size_t window_samples;
    double blockage;   /* section height / channel height */
    size_t solid_nodes;
} lbm_aero2d_info;

typedef struct lbm_aero2d lbm_aero2d;
const char* lbm_aero2d_last_error(void);
void lbm_aero2d_default_options(lbm_aero2d_options* options);
int lbm_aero2d_create(const lbm_aero2d_options* options, lbm_aero2d** out);
int lbm_aero2d_destroy(lbm_aero2d* a);
/* setup (lbm_aero_solver.cpp:39-198) without the device lattice: configuration, mask and
 * warnings only, on the host. section: n points (x, z), interleaved. */
int lbm_aero2d_plan(lbm_aero2d* a, const double* section, size_t n, const lbm_reference_values* reference,
                    const lbm_flow_conditions* conditions);
/* setup: plan, then the device lattice (lbm2d_create). */
int lbm_aero2d_setup(lbm_aero2d* a, const double* section, size_t n, const lbm_reference_values* reference,
                     const lbm_flow_conditions* conditions);
/* step (lbm_aero_solver.cpp:200-235): *finished = 1 once the run is complete. */
int lbm_aero2d_step(lbm_aero2d* a, int32_t* finished);
int lbm_aero2d_get_result(const lbm_aero2d* a, lbm_aero3d_result* out);
/* CoefficientHistory: sample time (lattice steps), drag and lift coefficients. */
int lbm_aero2d_history(const lbm_aero2d* a, double* time, double* drag, double* lift, size_t capacity,
                       size_t* count);
int lbm_aero2d_residuals(const lbm_aero2d* a, double* residuals, size_t capacity, size_t* count);
/* Coefficients (drag, lift) of the last sample. */
int lbm_aero2d_instantaneous(const lbm_aero2d* a, double coefficients[2]);
/* strouhal() (lbm_aero_solver.cpp:249-263): *found = 0 when the lift shows no oscillation. */
int lbm_aero2d_strouhal(const lbm_aero2d* a, double* strouhal, int32_t* found);
/* Wall samples (SurfaceField): points and normals [n][3] in body axes (y = 0); areas (dx, per
 * metre of span); cp. */
int lbm_aero2d_surface(const lbm_aero2d* a, double* points, double* normals, double* areas, double* cp,
                       size_t capacity, size_t* count);
int lbm_aero2d_warning_count(const lbm_aero2d* a, size_t* count);
int lbm_aero2d_warning(const lbm_aero2d* a, size_t index, char* code, size_t code_capacity, char* message,
                       size_t message_capacity);
int lbm_aero2d_get_info(const lbm_aero2d* a, lbm_aero2d_info* out);
/* The lattice mask of the plan: solid [cells]. */
int lbm_aero2d_mask(const lbm_aero2d* a, uint8_t* solid);
/* The device lattice (borrowed; valid until the next setup or destroy). */
int lbm_aero2d_lattice(const lbm_aero2d* a, lbm2d** out);
/* The SI velocity field [cells][3] = (u, 0, w) on the lattice, on request (zero angle of attack
 * only). */
int lbm_aero2d_velocity(lbm_aero2d* a, double* values);

/* ---- shared helpers ----------------------------------------------------------------------- */
/* Seeded perturbation in [-1, 1): SplitMix64 of (seed, cell, component), 53-bit mantissa.
 * The same counter-based draw on the host and the device, so every solver starts alike. */
double lbm_perturbation(uint64_t seed, uint64_t cell, uint32_t component);

#ifdef __cplusplus
}
#endif

#endif /* ORRERY_LBM_H */
