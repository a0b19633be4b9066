/*
 * pbf.h - C ABI of the MI355X position-based-fluids solver.
 *
 * This is the drop-in boundary for the reference's per-step path.  The
 * reference exposes that path as the C++ class `Simulation`
 * (src/Simulation.hpp:30-197) built on an OpenCL kernel map
 * (src/ocl/clsetup.cpp:344-358, src/main.cpp:170-175).  Every entry point
 * below names the reference member it replaces.  No exceptions cross this
 * boundary: every call returns an int status (PBF_OK == 0) and the text of
 * the last failure on the calling thread is available from pbf_last_error().
 *
 * Threading: one HIP stream per handle; a handle is not thread-safe, callers
 * serialise calls on the same handle.  Distinct handles are independent.
 *
 * Array conventions: particle arrays are float4 records (4 floats per
 * particle) in ORIGINAL particle order (the order of the input particle
 * array), exactly as Simulation::dumpData returns them
 * (src/Simulation.cpp:578-591): pos = (x, y, z, |v|), vel = (vx, vy, vz, m).
 */
#ifndef PBF_H
#define PBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PBF_OK 0
#define PBF_ERR_INVALID 1   /* bad argument (null, size, range)              */
#define PBF_ERR_HIP 2       /* HIP runtime failure (no device, OOM, launch)  */
#define PBF_ERR_STATE 3     /* call not valid for this handle                */
#define PBF_ERR_NUMERIC 4   /* a position or velocity became NaN or infinite */

/* ---- physics modes ----------------------------------------------------- */
/* PBF_PHYSICS_REF: the equations of the reference kernels
 *   (src/kernels/compute_scaling.cl:12-140, compute_delta.cl:13-152,
 *    apply_vorticity_and_viscosity.cl:12-104), with the kernel constants
 *   rounded to 6 significant digits as src/main.cpp:149-168 formats them.
 * PBF_PHYSICS_PAPER: Macklin & Mueller 2013 (eqs. 1, 2, 8, 11, 13-17):
 *   relaxation epsilon, s_corr, XSPH and vorticity confinement, exact
 *   constants. */
#define PBF_PHYSICS_REF 0
#define PBF_PHYSICS_PAPER 1

/* ---- precision --------------------------------------------------------- */
/* PBF_PRECISION_MIXED (PAPER only): the lambda passes of iterations
 * 1 .. iterations - fp32_iterations - 1 and the XSPH and vorticity passes
 * gather 16-bit neighbour records: positions as 16-bit fixed point in units
 * of h / 4096 kept modulo 2^16, |omega|, density and velocity as FP16
 * (8-byte records; 16 bytes with the velocity).  Kernels are evaluated and
 * accumulated in FP32.  Iteration 0's lambda pass is the FP32 neighbour
 * search, whose lists every pass of the step uses; every delta pass gathers
 * FP32 records and updates FP32 positions, and the last fp32_iterations
 * iterations are full FP32. */
#define PBF_PRECISION_FP32 0
#define PBF_PRECISION_MIXED 1

/* Simulation parameters.  The first block maps 1:1 onto the reference's
 * ConfigParameters (src/Parameters.hpp:12-31) as read from a .par file
 * (src/io/ConfigReader.cpp:19-94); the file-name and output fields stay in
 * the C++ facade.  The second block holds what the reference hard-codes in
 * its kernels and host (epsilon compute_scaling.cl:16, gravity
 * predict_positions.cl:8, XSPH c apply_vorticity_and_viscosity.cl:103,
 * iteration count Simulation.cpp:503).  pbf_params_init fills defaults. */
typedef struct pbf_params {
  float time_step_length; /* timestep_length                                */
  float time_end;         /* time_end (used by the Runner facade only)      */
  float x_min, x_max, y_min, y_max, z_min, z_max; /* box                    */
  float x_n, y_n, z_n;    /* cell counts per axis (floats, as reference)    */
  float rest_density;     /* restdensity                                    */

  int32_t physics;        /* PBF_PHYSICS_*                                  */
  int32_t iterations;     /* constraint iterations per step (reference: 4)  */
  float relaxation;       /* epsilon in the lambda denominator              */
  float scorr_k;          /* s_corr k          (paper only)                 */
  float scorr_n;          /* s_corr exponent n (paper only)                 */
  float scorr_dq;         /* s_corr delta-q as a fraction of h (paper only) */
  float xsph_c;           /* XSPH viscosity c                               */
  float vorticity_eps;    /* vorticity confinement epsilon (paper only)     */
  float gravity_y;        /* gravity along y (reference: -9.81)             */
  int32_t precision;      /* PBF_PRECISION_*                                */
  int32_t fp32_iterations;/* mixed precision: trailing full-FP32 iterations */
  float sphere[4];        /* obstacle sphere (cx, cy, cz, r); r <= 0: none  */
  int32_t block_size;     /* gather work-group size 64/128/256; 0 = tuned   */
  int32_t launch_grid;    /* tile kernels: 0 = one work-group per tile of   */
                          /* block_size slots; g > 0 (a multiple of 8) =    */
                          /* persistent, g work-groups walking all tiles;   */
                          /* -1 = persistent, as many work-groups as the    */
                          /* device holds at once.                          */
  int32_t scan_cells;     /* cells per work-group of the cell-count scan:   */
                          /* 1024, 2048 or 4096; 0 = 1024 or tuned.         */
                          /* block_size, launch_grid and scan_cells all 0:  */
                          /* all three tuned                                */
  int32_t reserved[5];
} pbf_params;

/* One input particle, layout of the reference's Particle
 * (src/Particle.hpp:9-14): mass, position, velocity, unused acceleration. */
typedef struct pbf_particle {
  float m;
  float x[3];
  float v[3];
  float a[3];
} pbf_particle;

typedef struct pbf_sim pbf_sim;

/* Fill *p with the defaults of a physics mode (box and cell counts are left
 * zero).  REF: epsilon 1e4, c 0.01, g -9.81, 4 iterations, no s_corr, no
 * vorticity.  PAPER: epsilon 600, s_corr k 0.001 n 4 dq 0.2, c 0.01,
 * vorticity 1e-4, g -9.81, 4 iterations.  No reference counterpart. */
int pbf_params_init(pbf_params *p, int32_t physics);

/* Replaces Simulation::Simulation + init() + initCells()
 * (src/Simulation.hpp:41-46, src/Simulation.cpp:34-88,104-230): validates
 * parameters, allocates device state on the current HIP device, uploads the
 * particles.  n may be 0 (every later call is then a no-op). */
int pbf_create(const pbf_params *params, const pbf_particle *particles,
               uint32_t n, pbf_sim **out);

/* Many independent scenes in one handle; every kernel of a step covers all
 * scenes in one launch.  scene s has counts[s] particles at particles[s].
 * No reference counterpart (the reference simulates one scene). */
int pbf_batch_create(uint32_t nscenes, const pbf_params *params,
                     const pbf_particle *const *particles,
                     const uint32_t *counts, pbf_sim **out);

/* Replaces Simulation::step (src/Simulation.cpp:455-576), nsteps times.
 * Asynchronous with respect to the host; pbf_synchronize / pbf_dump wait. */
int pbf_step(pbf_sim *sim, uint32_t nsteps);

/* Block the host until all work queued on the handle's stream is done.
 * Replaces the mQueue.finish() at src/Simulation.cpp:573. */
int pbf_synchronize(pbf_sim *sim);

/* Replaces Simulation::dumpData (src/Simulation.cpp:578-591): fills
 * caller-owned float4[n] arrays (either may be NULL) in original order. */
int pbf_dump(pbf_sim *sim, float *pos4, float *vel4);

/* Batched handle: dump one scene. */
int pbf_batch_dump(pbf_sim *sim, uint32_t scene, float *pos4, float *vel4);

/* Replaces the positions buffer the reference shares with GL: the
 * CL buffer created on the viewer's VBO (src/Simulation.cpp:149-158),
 * acquired and released around every step (:464-467, :570-574), into which
 * updatePositions writes each particle's (x, |v|) at its index
 * (src/kernels/update_positions.cl:9-12) and which the viewer draws
 * (src/visual/visual.cpp:256-268, 319-323).  Writes the current (x, |v|) of
 * every particle of `scene` (0 for a single-scene handle) into dst, float4 by
 * particle index in input order -- the layout pbf_dump's pos4 has -- and
 * returns when the write is done.  dst is device memory of the handle's GPU
 * holding at least that scene's particle count of float4: a hipMalloc
 * buffer, or the pointer hipGraphicsResourceGetMappedPointer gives for a GL
 * buffer registered with hipGraphicsGLRegisterBuffer and mapped with
 * hipGraphicsMapResources (unmap it before drawing).  Waits first for all
 * work queued on the GPU, on any stream (the map or a fill of dst included),
 * as the reference finishes GL before acquiring the buffer (:464).  A
 * pointer HIP does not know as device memory of that GPU is rejected
 * (PBF_ERR_INVALID). */
int pbf_positions_to_device(pbf_sim *sim, uint32_t scene, void *dst);

/* Replaces Simulation::setWaveGenerator (src/Simulation.hpp:105-108):
 * moving x-min wall offset used by the collision clamp
 * (src/kernels/compute_delta.cl:130-131).  Applies to every scene. */
int pbf_set_wave_generator(pbf_sim *sim, float value);

/* Replaces getSizesMin/getSizesMax (src/Simulation.hpp:93-101). */
int pbf_get_box(const pbf_sim *sim, float min4[4], float max4[4]);

/* Replaces getNumberParticles (src/Simulation.hpp:61-63); batched handles
 * return the total over scenes. */
uint32_t pbf_num_particles(const pbf_sim *sim);
uint32_t pbf_num_scenes(const pbf_sim *sim);

/* Replaces ~Simulation (src/Simulation.cpp:90-101).  NULL is accepted. */
int pbf_destroy(pbf_sim *sim);

/* Thread-local text of the last failure ("" if none).  No reference
 * counterpart (the reference throws cl::Error / std::runtime_error,
 * src/main.cpp:180-186). */
const char *pbf_last_error(void);

/* ---- scenario loading ---------------------------------------------------- */

/* Replaces ConfigReader::read (src/io/ConfigReader.cpp:19-94): reads a .par
 * file ("key value" lines; the reference keys plus the physics keys
 * physics ref|paper, iterations, relaxation, scorr_k, scorr_n, scorr_dq,
 * xsph_c, vorticity_eps, gravity_y, precision fp32|mixed, fp32_iterations,
 * sphere cx cy cz r, block_size, launch_grid, scan_cells).  Missing keys
 * keep the REF defaults of pbf_params_init (zero for the box).
 * part_input_file is copied into part_file (NUL-terminated, at most cap
 * bytes) when part_file != NULL. */
int pbf_read_par(const char *path, pbf_params *out, char *part_file,
                 size_t cap);

/* Replaces PartReader::read (src/io/PartReader.cpp:21-97): reads a .in
 * particle file; *n receives the number of particles read and at most cap
 * of them are copied to out (out may be NULL to query the count). */
int pbf_read_particles(const char *path, pbf_particle *out, uint32_t cap,
                       uint32_t *n);

/* Advance caller-owned device state nsteps steps, allocating nothing: no
 * reference counterpart; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  pos4 / vel4 are float4[n] arrays in DEVICE memory of the
 * current device, in original particle order: pos = (x, y, z, any), vel =
 * (vx, vy, vz, m) in; the state after nsteps steps out, laid out as
 * pbf_dump returns it (pos.w = |v|).  One scene; the same kernels and
 * results as pbf_create + pbf_step + pbf_dump with the same parameters
 * (work-group size block_size, or 128 if 0; no autotuning).  Every
 * temporary lives in the library's static device scratch (16 MiB per GPU,
 * held from library load to process end): the 32,768-particle dam break
 * fits; sizes that need more are refused with PBF_ERR_INVALID.  The work
 * is queued on `stream` (a hipStream_t; NULL: the null stream) and not
 * waited for; it reports no numeric error (use a handle for that).  All
 * calls on one GPU share its scratch: calls that overlap on one GPU (other
 * streams, other threads) race. */
int pbf_step_noalloc(const pbf_params *params, float *pos4, float *vel4,
                     uint32_t n, uint32_t nsteps, void *stream);

/* ---- parity hooks and measurement (no reference counterpart) ----------- */

/* Run the neighbour-search stage of a step on the current state without
 * advancing it: predict (predict_positions.cl:5-9), cell assignment
 * (update_cells.cl:9-15), the sort, and one density pass.  Outputs, each
 * optional (NULL to skip), single-scene handles only:
 *   cell_of[n]      cell id of each particle (original order)
 *   sorted_ids[n]   original particle id at each sorted slot
 *   cell_start[C+1] first sorted slot of each cell, cell_start[C] = n
 *   nbr_count[n]    neighbours j != i with 0 < r^2 < h^2 in the 27-cell
 *                   stencil (original order)
 *   rho[n], lambda[n]  density and lambda of the first constraint pass
 *   pred4[n]        predicted positions x* (float4, original order)      */
int pbf_probe(pbf_sim *sim, int32_t *cell_of, int32_t *sorted_ids,
              int32_t *cell_start, int32_t *nbr_count, float *rho,
              float *lambda, float *pred4);

/* Number of cells C of a single-scene handle (length of cell_start is C+1).
 * REF handles bin in the x_n * y_n * z_n cells of the parameters; PAPER
 * handles bin in 4 x_n * y_n * z_n cells (each x cell split in 4 along x, so
 * that the neighbour search can cut its rows to the reach h). */
uint32_t pbf_num_cells(const pbf_sim *sim);

/* FP64 reductions on the device over the current state (all scenes), with
 * e_i = rho_i/rho0 - 1, rho_i the density at the current positions (self
 * excluded; a neighbour is a j != i with 0 < |x_i - x_j|^2 < h^2):
 * out[0] = max over all particles of max(e_i, 0) (largest compression),
 * out[1] = mean |e_i| over the particles with at least one neighbour,
 * out[2..4] = total momentum sum m v, out[5] = kinetic energy sum m v^2/2,
 * out[6] = number of particles with at least one neighbour,
 * out[7] = max |e_i| over the particles with at least one neighbour.
 * Isolated particles (e_i = -1) enter only out[2..5]. */
int pbf_diagnostics(pbf_sim *sim, double out[8]);

/* Run nsteps steps, timing every kernel launch with start/stop events
 * attached to its dispatch (the kernel's own duration, launch gaps
 * excluded); writes per-kernel-class totals (milliseconds, summed over the
 * nsteps steps) into ms[k] and launch counts into launches[k] for
 * k < *nclass (in: capacity, out: number of classes), names[k] -> static
 * strings.  The state advances as with pbf_step. */
int pbf_profile_steps(pbf_sim *sim, uint32_t nsteps, double *ms,
                      uint32_t *launches, const char **names,
                      uint32_t *nclass);

/* Time nsteps graph launches with HIP events on the handle's stream
 * (after the caller's warm-up); returns elapsed milliseconds in *ms. */
int pbf_time_steps(pbf_sim *sim, uint32_t nsteps, double *ms);

/* Average duration in *ms of one launch of the kernel class named cls
 * ("search", "lambda" or "delta"; REF: "lambda" or "delta"): after one
 * step, reps launches of that class's kernel are captured back to back into
 * a graph and timed with events around one graph launch.  The kernels
 * recompute step temporaries only; the simulation state (pbf_dump) is that
 * of the step.  No reference counterpart. */
int pbf_time_kernel(pbf_sim *sim, const char *cls, uint32_t reps, double *ms);

/* Measurement calibration: run a plain float4 copy of `bytes` bytes (read
 * bytes, write bytes) `reps` times on the current device and return the
 * average kernel time in *ms.  Its HBM counters give the scale factor for
 * the per-kernel FETCH_SIZE / WRITE_SIZE readings. */
int pbf_calibrate_copy(uint64_t bytes, uint32_t reps, double *ms);

/* Launch configuration of this handle: cfg[0] = work-group size of the
 * searching lambda pass (PAPER) or of the lambda pass (REF).  With
 * block_size 0, launch_grid 0 and at least 2^18 slots, every kernel class
 * gets its own size and launch form, chosen by timing 64, 128 and 256 with
 * one work-group per tile and persistent (resident work-groups) on the
 * first step (cached per shape, and in the file named by the PBF_TUNE_CACHE
 * environment variable); otherwise all use block_size (default 128) and
 * launch_grid.  cfg[1] = particles per
 * work-item (1), cfg[2] = sort passes (1: one counting pass), cfg[3] =
 * iterations that gather 16-bit records (mixed precision). */
int pbf_get_config(const pbf_sim *sim, int32_t cfg[4]);

/* Launch form of each kernel class of this handle, classes in the order of
 * pbf_profile_steps (predict, scan, scatter, reorder, lambda, delta, xsph,
 * vorticity, search): block[c] = work-group size, grid[c] = 0 for one
 * work-group per tile of block[c] slots, the work-group count of a
 * persistent launch, or -1 for a persistent launch of as many work-groups
 * as the device holds at once (the autotuner times the per-tile and the
 * resident persistent form; launch_grid forces one).  Fills min(n, 9)
 * entries; *count (if not NULL) receives 9 and *scan_cells (if not NULL)
 * the cells per work-group of the cell-count scan (1024, 2048 or 4096:
 * scan_cells, or tuned with the rest).  No reference counterpart (the
 * reference launches every kernel over the particle count rounded up to
 * 32, Simulation.cpp:138-141). */
int pbf_get_launch(const pbf_sim *sim, int32_t *block, int32_t *grid,
                   int32_t n, int32_t *count, int32_t *scan_cells);

#ifdef __cplusplus
}
#endif

#endif /* PBF_H */
