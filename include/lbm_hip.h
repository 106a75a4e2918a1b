/* lbm_hip.h — C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann time step.
 *
 * One fused step per launch: SoA pull streaming with halfway bounce-back, BGK collision and
 * Guo body force, on populations stored shifted (g = f - w). In float and double the result is
 * bit-identical to the reference's sequential v0 solver.
 *
 * The reference has no FFI: its boundary is the duck-typed C++ Solver concept consumed by
 * run_driver (src/common/driver.hpp:5-16,255-300) and by the test templates. Each entry point
 * below names the member of SolverV3 (src/v3_cuda_opt/solver_v3.hpp:27-57, solver_v3.cu) or of
 * the v4 StripEngine (src/v4_mpi_cuda/engine.hpp:36-61) that it replaces. The C++ class
 * lbm::SolverHip in lbm_hip_solver.hpp restores SolverV3's exact method set on top of it.
 *
 * Conventions
 *   - Every function returns an int status: LBM_OK, or a negative code; the message of the last
 *     failure on the calling thread is lbm_last_error(). Nothing throws or aborts across the ABI
 *     (the reference throws std::invalid_argument for bad arguments, solver_v3.cu:169-170, and
 *     aborts on CUDA errors, cuda_check.hpp:17-25).
 *   - Host arrays are caller-owned. "real" arrays are float for LBM_FLOAT and LBM_FP16 and
 *     double for LBM_DOUBLE. Grids are row-major, cell = y * nx + x, N = nx * ny.
 *   - Device memory is owned by the handle. All work runs on one HIP stream per handle
 *     (default priority); lbm_step is asynchronous, every read-back synchronises.
 *   - Handles are not thread-safe; distinct handles are independent (no process-global
 *     parameters: everything a kernel needs is passed by value as a kernel argument).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_EINVAL (-1) /* invalid argument (reference: std::invalid_argument) */
#define LBM_EHIP (-2)   /* HIP runtime or launch error (reference: CUDA_CHECK abort) */
#define LBM_ESTATE (-3) /* call out of order (e.g. set_inflow_outflow after init) */
#define LBM_ENOMEM (-4) /* device allocation failed */

#define LBM_FLOAT 0  /* fp32 state and arithmetic (reference real = float) */
#define LBM_DOUBLE 1 /* fp64 state and arithmetic (reference LBM_REAL_DOUBLE) */
#define LBM_FP16 2   /* fp16 storage of scaled g, fp32 arithmetic (no reference counterpart) */

typedef struct lbm_solver lbm_solver;

/* Message of the last failed call on this thread ("" if none). */
const char* lbm_last_error(void);

/* Library version string and the device it runs on (hipDeviceProp name, gcnArchName). */
int lbm_device_info(char* buf, size_t len);

/* ---- plain solver (replaces SolverV3) ------------------------------------------------------ */

/* SolverV3::SolverV3(nx, ny, tau, force_x, force_y, Mask) — solver_v3.cu:166-186.
 * solid: N bytes, 1 = solid. tau must be > 0.5. Values are converted to the solver's real. */
int lbm_create(int nx, int ny, double tau, double force_x, double force_y,
               const uint8_t* solid, int precision, lbm_solver** out);
/* SolverV3::~SolverV3 — solver_v3.cu:188-192. NULL is accepted. */
int lbm_destroy(lbm_solver* s);

/* SolverV3::set_inflow_outflow(u_in) — solver_v3.cu:228-232. Open x: inlet at x = 0 with the
 * equilibrium of (1, u_in, 0), zero-gradient outlet at x = nx - 1. Call before init. */
int lbm_set_inflow_outflow(lbm_solver* s, double u_in);
/* SolverV3::init_equilibrium(rho0, ux0, uy0) — solver_v3.cu:194-226 (pull preimage). */
int lbm_init_equilibrium(lbm_solver* s, double rho0, double ux0, double uy0);

/* Setters (the reference fixes these at construction, solver_v3.cu:166-167). They take effect
 * at the next step; populations are kept. set_force_field: 2N host reals, planar (fx plane then
 * fy plane); NULL returns to the uniform force of lbm_set_force. */
int lbm_set_mask(lbm_solver* s, const uint8_t* solid);
int lbm_set_force(lbm_solver* s, double force_x, double force_y);
int lbm_set_force_field(lbm_solver* s, const void* fxy);
int lbm_set_tau(lbm_solver* s, double tau);

/* Launch shape: tile of block_x columns (a multiple of 64, <= 1024) by rows_per_thread rows,
 * one thread per column walking the rows (the reference's LBM_BLOCK "bx x by", solver_v3.cu:
 * 145-153, is read from the environment at creation with the same syntax). */
int lbm_set_block(lbm_solver* s, int block_x, int rows_per_thread);
int lbm_get_block(const lbm_solver* s, int* block_x, int* rows_per_thread);
/* No reference counterpart: code-generation options of the step (same bits either way;
 * lbm_autotune tries them): streaming_stores = non-temporal stores (periodic, uniform-force
 * kernel); single_resource = one buffer resource for the nine planes when they span < 4 GiB
 * (fewer scalar registers); tile_order = the order in which workgroups take tiles: 0 the
 * hardware's round-robin over the 8 XCDs, 1 a contiguous band of tiles per XCD, 2 whole tile
 * rows per XCD (row r on XCD r % 8: the cache lines that neighbouring tiles of a row share stay
 * in one L2; needs a multiple of 8 tile rows, else round-robin). Defaults: fp32 1, 1, 0;
 * fp64 1, 1, 1; FP16 0, 1, 2. */
int lbm_set_kernel_options(lbm_solver* s, int streaming_stores, int single_resource,
                           int tile_order);
int lbm_get_kernel_options(const lbm_solver* s, int* streaming_stores, int* single_resource,
                           int* tile_order);

/* No reference counterpart: temporal blocking. With steps_per_launch = 2 (the default)
 * lbm_run and lbm_timed_run advance pairs of steps in one launch of a two-step kernel wherever
 * it applies (float and double: periodic or open x, uniform force or a per-cell field; FP16
 * storage: periodic x, uniform force): every population is read from HBM and written once per
 * two steps. The results are bit-identical to single
 * steps (same operations per cell). block_x (64..512, multiple of 64; 0 keeps the current
 * value) and rows set its tile: rows > 0 a fixed height, -1 (the default) the tallest of 64,
 * 48, 32, 24, 16 rows whose tile count fills the GPU's resident workgroup slots evenly, 0 keeps
 * the current setting; tile_order (0, 1, 2 as in lbm_set_kernel_options, -1 keeps; default 1,
 * FP16 storage 2)
 * orders its tiles over the XCDs. lbm_step always advances one step. lbm_get_steps_per_launch
 * reports the height in use and in_use whether the current configuration takes the two-step
 * path. */
int lbm_set_steps_per_launch(lbm_solver* s, int steps_per_launch, int block_x, int rows,
                             int tile_order);
int lbm_get_steps_per_launch(const lbm_solver* s, int* steps_per_launch, int* block_x, int* rows,
                             int* tile_order, int* in_use);

/* SolverV3::step() — solver_v3.cu:251-278: one asynchronous step. */
int lbm_step(lbm_solver* s);
/* SolverV3::run(steps) — solver_v3.cu:280-283: steps, then synchronise. */
int lbm_run(lbm_solver* s, int steps);
/* SolverV3::timed_run(steps) — solver_v3.cu:285-298: seconds between two HIP events recorded
 * on the solver's stream around the steps. */
int lbm_timed_run(lbm_solver* s, int steps, double* seconds);

/* SolverV3::macroscopic — solver_v3.cu:322-343: rho, ux, uy (N reals each) from the
 * post-streaming state, computed on the device with v0's operations; solid cells get 0. */
int lbm_macroscopic(lbm_solver* s, void* rho, void* ux, void* uy);
/* No counterpart in the reference: lbm_macroscopic for callers that cannot allocate device
 * memory (no hipMalloc once set up: tight field-output loops, harnesses that forbid it). Same
 * outputs, bit for bit; the fields are staged in a fixed 12 MiB device array of the library
 * (one per GPU, held until the process ends) instead of a buffer the handle allocates, so
 * 3 * nx * ny reals must fit in 12 MiB (1,048,576 cells in float or FP16 storage, 524,288 in
 * double), else LBM_EINVAL. Calls from one process are serialised; every call on a GPU reuses
 * its array. */
int lbm_macroscopic_noalloc(lbm_solver* s, void* rho, void* ux, void* uy);
/* lbm::vorticity(ux, uy, solid, nx, ny, periodic_x) — src/common/io.cpp:54-74, applied to the
 * solver's macroscopic velocity: N reals (float for FP16 storage), 0 in solid cells, computed
 * on the device with the reference's operations. periodic_x: 1 periodic, 0 one-sided
 * differences at the open ends (the cylinder driver's choice, driver.hpp:229,245), -1 the
 * solver's own x boundary. */
int lbm_vorticity(lbm_solver* s, void* out, int periodic_x);
/* SolverV3::total_mass — solver_v3.cu:345-355: fluid cells + sum of g, FP64 on the device
 * (deterministic tree order, not v0's sequential order). */
int lbm_total_mass(lbm_solver* s, double* mass);
/* No reference counterpart: FP64 population momentum sum_c sum_i c_i g_i over fluid cells. */
int lbm_momentum(lbm_solver* s, double* px, double* py);
/* SolverV3::populations — solver_v3.cu:300-320: 9N reals, post-streaming, v0 AoS layout
 * (cell * 9 + i), solid cells 0. */
int lbm_populations(lbm_solver* s, void* out);
/* SolverV3::populations_at — solver_v3.cu:234-249: the 9 post-streaming reals of one cell. */
int lbm_populations_at(lbm_solver* s, int x, int y, void* out9);
/* Upload an arbitrary post-streaming state (v0 AoS layout); stored as its pull preimage. */
int lbm_set_populations(lbm_solver* s, const void* in);
/* Raw stored state: the post-collision pull state R, SoA (plane i at i * N), in the storage
 * type (for LBM_FP16: the scaled halves as uint16). */
int lbm_get_state(lbm_solver* s, void* out);
int lbm_set_state(lbm_solver* s, const void* in);

/* SolverV3::threads — 0 on the GPU. */
int lbm_threads(const lbm_solver* s);
/* SolverV3::bench_info — solver_v3.cu:357-370: extra JSON fields (no braces) for the BENCH line:
 * gpu (the device name, or "AMD <arch>" where the runtime reports none), arch, hip_runtime,
 * hip_driver, block (one-step launch shape), variant ("soa_pull_two_step" when run() advances
 * two steps per launch, else "soa_pull_buffer"), storage, stores, planes, tile_order,
 * plane_stride, steps_per_launch, two_step_tile, two_step_order. */
int lbm_bench_info(const lbm_solver* s, char* buf, size_t len);
int lbm_dims(const lbm_solver* s, int* nx, int* ny, int* precision);

/* FP16 storage scale (power of two): stored half = g * scale. Default 2^12. */
int lbm_set_fp16_scale(lbm_solver* s, double scale);

/* Autotuner: times `steps` steps for each candidate (block_x, rows_per_thread) and code option
 * of the one-step kernel and keeps the fastest on the handle; where run() takes the two-step
 * kernel it then times its tiles (256 or 512 columns; automatic, 32 or 64 rows; tile orders
 * 0..2) and keeps the fastest, and best_mlups is the rate of that path. The state is saved and
 * restored around the trials. */
int lbm_autotune(lbm_solver* s, int steps, int* best_block_x, int* best_rows, double* best_mlups);

/* No reference counterpart (the reference samples a velocity probe, driver.hpp:186-251):
 * momentum-exchange force on the solids bordering the fluid cells of the box [x0, x0 + w) x
 * [y0, y0 + h), FP64, deterministic; drag = *fx, lift = *fy (lattice units per step). */
int lbm_momentum_exchange(lbm_solver* s, int x0, int y0, int w, int h, double* fx, double* fy);

/* Reference poiseuille_force(H, tau, u_max) in the solver precision (geometry.hpp:24-26). */
double lbm_poiseuille_force(int precision, int H, double tau, double u_max);

/* ---- strip engine: one GPU's share of the v4 strip decomposition -------------------------- */
/* Mirrors StripEngine (src/v4_mpi_cuda/engine.hpp:36-61, engine_gpu.cu:27-131). A strip holds
 * ny_local owned rows plus ghost row 0 (below) and ghost row ny_local + 1 (above); no wrap in y;
 * SoA with N = nx * (ny_local + 2). Halo: planes {2,5,6} of the top owned row go up, planes
 * {4,7,8} of the bottom owned row go down; 3 * nx reals per message. */
typedef struct lbm_strip lbm_strip;
int lbm_strip_create(int nx, int ny_local, int precision, lbm_strip** out);
/* No reference counterpart: a strip with ghost_rows = 1 (as lbm_strip_create) or 2 ghost rows
 * on each side; state [9][ny_local + 2 * ghost_rows][nx], owned rows are local rows ghost_rows
 * .. ghost_rows + ny_local - 1 (ny_local >= ghost_rows). With two, one halo exchange serves two
 * fused steps (lbm_strip_compute_rows2) and a message is 12 * nx reals: up, the 9 planes of the
 * top owned row then planes {2,5,6} of the row below it (they land in ghost rows 1 and 0 of
 * the strip above); down, the 9 planes of the bottom owned row then planes {4,7,8} of the row
 * above it (ghost rows ny_local + 2 and ny_local + 3 of the strip below). Single steps
 * (lbm_strip_compute_rows) work on either layout. */
int lbm_strip_create_ghost(int nx, int ny_local, int precision, int ghost_rows, lbm_strip** out);
int lbm_strip_destroy(lbm_strip* s);
/* Reals per halo message: 3 * nx with one ghost row, 12 * nx with two. */
int lbm_strip_halo_values(const lbm_strip* s, size_t* values);
/* StripParams (strip_update.hpp:23-30): tau, force, open x with inlet velocity u_in. */
int lbm_strip_set_params(lbm_strip* s, double tau, double force_x, double force_y, int open_x,
                         double u_in);
int lbm_strip_upload_solid(lbm_strip* s, const uint8_t* solid_strip);
int lbm_strip_upload_state(lbm_strip* s, const void* g);
int lbm_strip_download_state(lbm_strip* s, void* g);
/* compute_rows(first, last, stream): local rows inclusive, stream 0 = boundary, 1 = interior. */
int lbm_strip_compute_rows(lbm_strip* s, int first, int last, int stream);
/* No reference counterpart: two fused steps of owned local rows [first, last] in one launch of
 * the two-step kernel (two ghost rows, periodic x, uniform force); it reads rows first - 2 ..
 * last + 2 of the current state and writes rows first .. last of the next, bit-identical to
 * two single steps. Disjoint row ranges may run on the two streams at once. */
int lbm_strip_compute_rows2(lbm_strip* s, int first, int last, int stream);
/* No reference counterpart: lbm_strip_compute_rows2 that also delivers the halo. As the
 * launch stores the two outermost owned rows of each side of the new state, it stores their
 * message entries straight into the next-state ghost rows of `up` (the strip above: its rows 0
 * and 1) and `down` (the strip below: its rows ny_local + 2 and + 3), device to device, with
 * no pack/unpack (the neighbours may be the strip itself). All strips need two ghost rows,
 * the same nx and precision; a receiver's ghost rows are complete once the sender's streams
 * are synchronised (its lbm_strip_finish_step). */
int lbm_strip_compute_rows2_to(lbm_strip* s, int first, int last, int stream, lbm_strip* up,
                               lbm_strip* down);
/* Tile of the two-step launches: block_x columns (64..512 float, ..256 double, multiple of 64;
 * 0 keeps; default 256), rows (> 0 fixed, -1 automatic, 0 keeps), tile_order (0..2 as in
 * lbm_set_kernel_options, -1 keeps; default 2). Same bits for every tile. */
int lbm_strip_set_two_step_tile(lbm_strip* s, int block_x, int rows, int tile_order);
/* pack_halo: send_up / send_down (lbm_strip_halo_values reals each) from the new state; waits
 * for stream 0. */
int lbm_strip_pack_halo(lbm_strip* s, void* send_up, void* send_down);
/* unpack_halo: recv_from_down -> the ghost rows below (row 0; rows 0 and 1 with two),
 * recv_from_up -> the ghost rows above (row ny_local + 1; ny_local + 2 and + 3 with two). */
int lbm_strip_unpack_halo(lbm_strip* s, const void* recv_from_down, const void* recv_from_up);
/* finish_step: wait for both streams, swap the buffers. */
int lbm_strip_finish_step(lbm_strip* s);
int lbm_strip_set_block(lbm_strip* s, int block_x, int rows_per_thread);
/* Seconds of `steps` rounds of compute_rows over the owned rows + ghost-row copy from a device
 * halo (no host staging), HIP events on stream 0: the one-GPU share at its real shape. With two
 * ghost rows and periodic x, pairs of steps take one compute_rows2_to launch with the strip as
 * its own neighbours (no copy). */
int lbm_strip_timed_steps(lbm_strip* s, int steps, double* seconds);

/* ---- batched persistent kernel over many independent domains ------------------------------- */
/* No reference counterpart. Each domain is a plain solver of its own shape and parameters,
 * bit-identical to lbm_create + lbm_run on the same inputs. */
typedef struct lbm_domain_desc {
  int nx, ny;
  double tau, force_x, force_y;
  const uint8_t* solid; /* nx * ny bytes */
  double rho0, ux0, uy0; /* initial equilibrium */
} lbm_domain_desc;

typedef struct lbm_batch lbm_batch;
int lbm_batch_create(int n_domains, const lbm_domain_desc* domains, int precision,
                     lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
/* Steps per captured HIP graph (default 100); the graph is rebuilt on change. */
int lbm_batch_set_graph_steps(lbm_batch* b, int steps_per_graph);
/* Tile width (multiple of 64; in order 2 the threads per workgroup), rows per tile and
 * workgroups of the persistent launch. */
int lbm_batch_set_shape(lbm_batch* b, int tile_x, int tile_rows, int workgroups);
/* How the workgroups walk the work: 0 tiles w, w + G, ... (G workgroups); 1 a contiguous chunk
 * of tiles per workgroup, workgroups banded per XCD; 2 a wave work-list (every wavefront takes
 * 64-column segments of one domain over tile_rows rows: narrow domains leave few lanes idle).
 * streaming_stores: non-temporal stores. Every order computes the same bits. */
int lbm_batch_set_order(lbm_batch* b, int order, int streaming_stores);
int lbm_batch_get_order(const lbm_batch* b, int* order, int* streaming_stores);
/* Temporal blocking over the batch (every domain is periodic with a uniform force): with
 * steps_per_launch = 2 (the default) lbm_batch_run / lbm_batch_timed_run and their graphs
 * advance pairs of steps in one launch of the two-step kernel. Its work: each domain's columns
 * in sub-tiles of tile_x columns (a multiple of 64 up to the kernel's 512 threads, double 256;
 * default: that maximum) plus one of the remainder, its rows in bands of about tile_rows rows
 * (-1, the default: the height the built-in schedule model picks for these domains; 0 keeps);
 * the sub-tiles are packed into work items of one workgroup each. `workgroups`: 0 (the
 * default) one workgroup per item; G > 0 a persistent grid of G; -1 keeps. An odd count ends
 * with one single step. Same bits as single steps. */
int lbm_batch_set_steps_per_launch(lbm_batch* b, int steps_per_launch, int tile_x, int tile_rows,
                                   int workgroups);
/* tile_rows: the band height in use (modelled or set). */
int lbm_batch_get_steps_per_launch(const lbm_batch* b, int* steps_per_launch, int* tile_x,
                                   int* tile_rows, int* workgroups);
/* Threads per workgroup of the two-step kernel (a multiple of 64; at most 512, double 256, the
 * defaults): every work item packs sub-tiles into that many lanes, and the sub-tile width is
 * capped to it. */
int lbm_batch_set_two_step_threads(lbm_batch* b, int threads);
/* The two-step work list: threads per workgroup, work items, band rows in use, whether they are
 * the modelled choice, and the fill (cell rows over lane rows of the items, 0..1). Any output
 * may be null. */
int lbm_batch_two_step_layout(const lbm_batch* b, int* threads, int* items, int* band_rows,
                              int* modelled, double* fill);
int lbm_batch_run(lbm_batch* b, int steps);
int lbm_batch_timed_run(lbm_batch* b, int steps, double* seconds);
int lbm_batch_populations(lbm_batch* b, int domain, void* out);
int lbm_batch_total_cells(const lbm_batch* b, long long* cells);
/* Times every (order, stores, tile shape, workgroups) candidate for `steps` steps on a copy of
 * the state and keeps the fastest (lbm_batch_get_order reports its order); with two steps per
 * launch on, then every two-step tile (best_mlups is then the rate of that path). */
int lbm_batch_autotune(lbm_batch* b, int steps, int* tile_x, int* tile_rows, int* workgroups,
                       double* best_mlups);

/* ---- discrete adjoint ---------------------------------------------------------------------- */
/* No reference counterpart. Float or double forward of T steps with a per-cell force field and
 * tau, rest initial state; objective J = mean u_x (with the half-force term) over a probe
 * window, averaged over the last `avg_steps` states. Reverse sweep by the hand-written adjoint
 * step with checkpointing in at most `snapshots` stored states (slot 0 holds the initial
 * state). */
typedef struct lbm_adjoint lbm_adjoint;
int lbm_adjoint_create(int nx, int ny, double tau, const uint8_t* solid, int steps,
                       int probe_x0, int probe_y0, int probe_w, int probe_h, int avg_steps,
                       int snapshots, int precision, lbm_adjoint** out);
int lbm_adjoint_destroy(lbm_adjoint* a);
int lbm_adjoint_set_force_field(lbm_adjoint* a, const void* fxy);
int lbm_adjoint_set_tau(lbm_adjoint* a, double tau);
/* Forward only: J. */
int lbm_adjoint_forward(lbm_adjoint* a, double* J, double* seconds);
/* J and its gradient: dJ/dF (2N reals, planar) and dJ/dtau. `forward_steps`: forward steps of
 * the whole gradient sweep, the first pass included. */
int lbm_adjoint_gradient(lbm_adjoint* a, double* J, void* dJ_dforce, double* dJ_dtau,
                         double* seconds, long long* forward_steps);

/* Checkpointing schedules (every one gives the same gradient bits):
 *   LBM_SCHEDULE_BINOMIAL   revolve: the fewest forward steps for the snapshot budget; the
 *                           forward steps run one after another.
 *   LBM_SCHEDULE_PIPELINED  multi-level checkpoints; each reverse step shares its launch with
 *                           one recomputation step per level (about T launches for the reverse
 *                           sweep), within the snapshot budget.
 *   LBM_SCHEDULE_AUTO       every state stored when the budget holds them all (no
 *                           recomputation), else pipelined when it fits the budget, else
 *                           binomial (the default). */
#define LBM_SCHEDULE_AUTO 0
#define LBM_SCHEDULE_BINOMIAL 1
#define LBM_SCHEDULE_PIPELINED 2
int lbm_adjoint_set_schedule(lbm_adjoint* a, int kind);
/* The schedule in use: kind (BINOMIAL or PIPELINED), stored states, levels (0 for binomial),
 * forward steps and launches of the gradient sweep. */
int lbm_adjoint_schedule_info(lbm_adjoint* a, int* kind, int* slots, int* levels,
                              long long* forward_steps, long long* launches);

/* Host only (no device work): the schedule for `states` = T + 1 states within `slots` stored
 * states (built and replay-checked); any output may be null. */
int lbm_checkpoint_schedule(int states, int slots, int kind, int* slots_used, int* levels,
                            long long* forward_steps, long long* launches);
/* Host only: forward steps of the binomial schedule (first pass included). */
int lbm_checkpoint_forward_steps(int states, int slots, long long* steps);

/* ---- HBM copy ceiling (the counterpart of bench/gpu_bandwidth.cu) ---------------------------- */
/* Best-of-reps copy bandwidth (GB/s, 2 * bytes / time) on two arrays of `bytes` each:
 *   copy_gbs      best grid-stride copy kernel (scalar, 16-byte and 16-byte non-temporal
 *                 accesses, grids of 8..2048 workgroups per CU);
 *   memcpy_gbs    hipMemcpyDtoD;
 *   reference_gbs the reference's own configuration (scalar, block 256, grid = CUs x 32,
 *                 bench/gpu_bandwidth.cu:17-22,66-67);
 *   nine_gbs      nine arrays in, nine out, one element of each per thread (the stream
 *                 structure of the step's nine population planes).
 * Any output pointer may be null. */
int lbm_copy_ceiling(size_t bytes, int reps, double* copy_gbs, double* memcpy_gbs,
                     double* reference_gbs, double* nine_gbs);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
