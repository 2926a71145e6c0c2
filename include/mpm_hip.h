/*
 * mpm_hip.h -- C-ABI of the MI355X MLS-MPM substep engine (libmpm_hip.so).
 *
 * The reference exposes its simulation step only as a C++ class with no C-ABI
 * (class Simulation, include/mpm.cuh:44-106, src/mpm.cu:180-394) whose physics
 * is bound at compile time by three plugin aliases (include/mpm.cuh:24-27):
 * TransferScheme = MLS_APIC_Scheme, MaterialModel = MMSnow, InterpolationKernel
 * = QuadraticInterpolationKernel. This header is the flat boundary that replaces
 * that class: every entry point names the reference method it stands in for.
 * Plain pointers and sizes only; streams are passed as `void*` (a hipStream_t,
 * NULL = the device's null stream). Every call returns 0 (MPM_OK) or an error
 * code, and mpm_last_error() holds the message of the last failure on a handle.
 *
 * Layout conventions at this boundary:
 *   - "SoA" host arrays are per-field arrays: x[n][3], v[n][3], C[n][3][3],
 *     F[n][3][3] (row-major 3x3: element (r,c) at r*3+c), Jp[n], material[n].
 *   - "AoS104" is the reference's own particle bytes (MLS_APIC_Particle,
 *     include/types.h:24-35 + include/TransferScheme.h:46-54): u8 material@0,
 *     float x[3]@4, v[3]@16, F[9]@28 column-major, C[9]@64 column-major,
 *     Jp@100; 104 bytes, 4-byte aligned.
 *   - Downloads return particles in the order they were added, whatever order
 *     the engine keeps them in on the device.
 *
 * Calls on one handle are serialised by a mutex of the handle (mpm_last_error
 * returns a pointer that the next failing call on the handle overwrites);
 * distinct handles are independent.
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 1

/* status codes */
#define MPM_OK 0
#define MPM_ERR_INVALID 1   /* bad argument (null pointer, N out of range, ...) */
#define MPM_ERR_HIP 2       /* a HIP runtime call failed */
#define MPM_ERR_STATE 3     /* call order violated (e.g. advance before upload) */
#define MPM_ERR_CAPACITY 4  /* size beyond what the handle can hold */
#define MPM_ERR_IO 5        /* file could not be read or written */
#define MPM_ERR_NOGPU 6     /* no HIP device visible */

/* material kinds.  SNOW is the reference's only bound model
 * (MMSnow, include/MaterialModel.cuh:66-115; "elastic" scenes are SNOW with
 * hardening 0 and clamps [0, 1e30], scenes/rubber_duck.toml:1-8).
 * FIXED_COROTATED and JELLY are the reference's other plugin classes
 * (MaterialModel.cuh:39-64, :118-150).  WC_LIQUID is a weakly compressible
 * J-based liquid with no reference counterpart: tau = K (J-1) J I, J kept in
 * the Jp slot, F reset to J^(1/3) I after each substep. */
#define MPM_MATERIAL_SNOW 0
#define MPM_MATERIAL_FIXED_COROTATED 1
#define MPM_MATERIAL_JELLY 2
#define MPM_MATERIAL_WC_LIQUID 3

/* boundary modes for the grid update (reference: mpm.cu:92-105) */
#define MPM_BOUNDARY_FRACTION 0 /* walls where node_index/N < b or > 1-b (reference, b=0.05) */
#define MPM_BOUNDARY_CELLS 1    /* walls on the first/last b node layers of each axis */

/* interpolation kernels (option "interpolation") */
#define MPM_INTERP_QUADRATIC 0 /* QuadraticInterpolationKernel (reference) */
#define MPM_INTERP_CUBIC 1     /* cubic B-spline, generic D^-1 */

/* compatibility with result-changing details of the reference step */
#define MPM_COMPAT_REFERENCE 0 /* reproduce them (the default for parity) */
#define MPM_COMPAT_FIXED 1     /* every particle updated, P2G clipped, J = det F */

/* Simulation parameters.  Replaces SimulationParameters
 * (include/TransferScheme.h:6-29) plus the constants the reference hard-codes
 * in src/mpm.cu:6 (gravity) and :92 (boundary). dx and 1/dx are derived as
 * float(1.0/N) and float(1.0/dx), as TransferScheme.h:26-27 does. */
typedef struct mpm_params {
  float dt;              /* substep length (options.h --dt, default 1e-4) */
  uint32_t N;            /* grid nodes per axis, 4 <= N <= 1024 (options.h -N) */
  float gravity;         /* y acceleration; the reference uses -9.81 */
  int32_t boundary_mode; /* MPM_BOUNDARY_* */
  float boundary;        /* fraction (0.05) or node-layer count */
  int32_t compat;        /* MPM_COMPAT_* */
} mpm_params;

/* One material.  Replaces the MaterialModel plugin instance
 * (MaterialModelBase + MMFixedCorotated + MMSnow fields,
 * include/MaterialModel.cuh:18-35,47-48,70-72). Build it with
 * mpm_material_init() to get the reference's float arithmetic for the
 * Lame constants. */
typedef struct mpm_material {
  int32_t kind;          /* MPM_MATERIAL_* */
  float particle_volume; /* MaterialModelBase::particleVolume */
  float particle_mass;   /* density * volume */
  float mu0;             /* E / (2 (1 + nu)) */
  float lambda0;         /* E nu / ((1 + nu)(1 - 2 nu)) */
  float hardening;       /* xi in exp(xi (1 - Jp)) */
  float clamp_lo;        /* plasticity clamp on singular values */
  float clamp_hi;
  float bulk_modulus;    /* K, WC_LIQUID only */
} mpm_material;

/* Fill *m the way the reference's constructors do (MaterialModel.cuh:27-29,
 * :50-54, :75-82; main.cu:36-42 passes the TOML doubles as float). */
int mpm_material_init(mpm_material* m, int32_t kind, double volume, double density,
                      double E, double nu, double hardening, double clamp_lo,
                      double clamp_hi, double bulk_modulus);

/* ---- lifetime (replaces Simulation::Simulation / ~Simulation, mpm.cu:180-195) ---- */
typedef struct mpm_sim* mpm_handle;

/* Create a handle on the current HIP device. */
int mpm_create(const mpm_params* params, mpm_handle* out);
/* Release all device and host memory of the handle. NULL is accepted. */
int mpm_destroy(mpm_handle h);
/* Message of the last failed call on h (h == NULL: last failure of
 * mpm_create or any call given a NULL handle). Never NULL. */
const char* mpm_last_error(mpm_handle h);
/* ABI version of the library (MPM_ABI_VERSION). */
int mpm_abi_version(void);

/* ---- materials (replaces the material vector handed to Simulation,
 *      main.cu:9,25,45 and its upload in initCuda, mpm.cu:198-201) ---- */
int mpm_set_materials(mpm_handle h, const mpm_material* materials, int32_t n);

/* ---- particles (replaces Simulation::addObject/addParticles host staging,
 *      mpm.cu:237-251,347-394, and initCuda's upload, mpm.cu:197-207,278-286) ---- */
/* Append n particles to the host staging set. x and material are required;
 * v, C, F, Jp may be NULL (defaults: v = 0, C = 0, F = I, Jp = 1, as
 * MLS_APIC_Particle's constructor sets them, TransferScheme.h:50-53). */
int mpm_add_particles_soa(mpm_handle h, int64_t n, const float* x, const float* v,
                          const float* C, const float* F, const float* Jp,
                          const uint8_t* material);
/* Append n particles given in the reference's 104-byte layout. */
int mpm_add_particles_aos104(mpm_handle h, const void* particles, int64_t n);
/* Drop the host staging set (device state is kept until the next upload). */
int mpm_clear_particles(mpm_handle h);
/* Copy the staging set to the device (allocating as needed) and reset the
 * simulated time to 0 (Simulation::initCuda, mpm.cu:197-207). */
int mpm_upload(mpm_handle h, void* stream);
/* Copy the particles staged since the last mpm_upload / mpm_append_particles
 * (or mpm_clear_particles) to the device behind the ones there, which stay
 * on the device untouched; the new particles take the next original indices
 * (downloads list them last) and the simulated time goes on. Replaces the
 * reference's host round trip when an object's lifetime begins
 * (Simulation::particlesToHost + reallocateParticles, mpm.cu:288-314). */
int mpm_append_particles(mpm_handle h, void* stream);
/* Remove the device particles whose original index (the order downloads
 * list them in) lies in [first, first + count): they leave the simulation,
 * the others stay on the device in their order, and the indices above the
 * range move down by count (downloads list the remaining particles as before,
 * without the removed ones). With `removed` non-NULL it first receives the
 * removed particles' last state in the reference's 104-byte layout, count
 * entries in index order. Blocking. Replaces the reference's host round trip
 * when an object's lifetime ends (SimObject::isActive, mpm.cuh:38-40;
 * Simulation::particlesToHost + reallocateParticles, mpm.cu:288-314). */
int mpm_remove_particles(mpm_handle h, int64_t first, int64_t count, void* removed, void* stream);
/* Number of particles on the device (Simulation::getFullParticleCount,
 * mpm.cu:253-259, for the uploaded set). */
int mpm_count(mpm_handle h, int64_t* out);

/* ---- the substep (replaces Simulation::advance, mpm.cu:323-329) ---- */
/* Run `substeps` substeps, enqueued on `stream`; returns without waiting. */
int mpm_advance(mpm_handle h, int32_t substeps, void* stream);
/* Simulated time (Simulation::t, accumulated in double as mpm.cu:328 does). */
int mpm_time(mpm_handle h, double* out);

/* ---- download (replaces Simulation::syncDevice/particlesToHost,
 *      mpm.cu:209-211,288-306) ---- blocking; NULL outputs are skipped. */
int mpm_download_soa(mpm_handle h, float* x, float* v, float* C, float* F, float* Jp,
                     uint8_t* material, void* stream);
int mpm_download_aos104(mpm_handle h, void* particles, void* stream);
/* Grid after the last grid update: N^3 nodes x (vx, vy, vz, m) float, node
 * index N*N*i + N*j + k (reference layout, mpm.cu:49,55,66). After a grid
 * update the reference leaves m = 1 on nodes that had mass (mpm.cu:86-88). */
int mpm_download_grid(mpm_handle h, float* grid, void* stream);

/* ---- surface mesh of the particles (replaces MeshBuilder::computeMesh,
 *      include/mesh_builder.h:181-211, with its distance field,
 *      fillVoxelGrid_distance, mesh_builder.h:103-149; computed on the GPU
 *      from the device particles, no host copy of them) ----
 * mpm_mesh_build: distance field on grid_side^3 voxels (voxel_dx = N dx /
 * grid_side; each voxel the distance to the nearest particle whose box of
 * voxels, (particle_radius + 1) voxels around it, reaches it, else that
 * cut-off), then the iso-surface at particle_radius voxels by marching cubes
 * (a closed surface; the case table, mpm_mesh_case_table, is this library's,
 * not libigl's). Returns the triangle count. Blocking.
 * mpm_mesh_triangles: the triangles, 9 floats each (3 vertices, normal
 * pointing out by the right-hand rule), and for each vertex the grid edge it
 * lies on (lower node index << 32 | upper node index; NULL skips them).
 * mpm_mesh_distance_field: the grid_side^3 distances of the last build
 * (index (i*side+j)*side+k; grid_side must be that build's).
 * mpm_mesh_write_obj: the mesh as OBJ text, one vertex per crossed grid edge. */
int mpm_mesh_build(mpm_handle h, int32_t grid_side, int32_t particle_radius, int64_t* triangles,
                   void* stream);
int mpm_mesh_triangles(mpm_handle h, float* xyz, uint64_t* edges);
int mpm_mesh_distance_field(mpm_handle h, int32_t grid_side, float* out);
int mpm_mesh_write_obj(mpm_handle h, const char* path);
/* The last mesh indexed: one vertex per crossed grid edge (xyz doubles, in
 * edge-key order), faces as 0-based vertex triples. Pass V = F = NULL to get
 * the counts *nv, *nf first. */
int mpm_mesh_indexed(mpm_handle h, int64_t* nv, int64_t* nf, double* V, int32_t* F);
/* The marching-cubes case table: for case c (bit q: cube corner q = (q>>2&1,
 * q>>1&1, q&1) inside), ntri[c] triangles, tri[15 c + 3 t + k] the cube edge
 * of vertex k of triangle t; edges 0-3 along x from corners 0-3, 4-7 along y
 * from corners 0,1,4,5, 8-11 along z from corners 0,2,4,6. */
int mpm_mesh_case_table(uint8_t ntri[256], int8_t tri[256 * 15]);

/* ---- debug view (replaces Renderer, include/renderer.h:7-59 and
 *      src/renderer.cpp, which draw to a window; no reference C function) ----
 * The device particles as the reference draws them (points of 2 pixels at
 * clip position (x - 0.5) * 1.8, colour (0.59, 0.54, 0.2) on 0.1 grey, no
 * depth test) into `rgb`, a host buffer of width * height RGB8 pixels, rows
 * top to bottom (the reference's window: 800 x 600). A point covers the
 * pixels whose centres lie within one pixel of its window position (right
 * and top edges included). Blocking. */
int mpm_render_frame(mpm_handle h, int32_t width, int32_t height, uint8_t* rgb, void* stream);

/* ---- host-side mesh steps and particle files (replace MeshBuilder's
 *      libigl calls, mesh_builder.h:196-246, and ParticleWriter's Partio
 *      call, particle_writer.h:13-44; neither library is used). Vertices are
 *      double xyz triples, faces 0-based int32 triples. ----
 * mpm_mesh_smooth: MeshBuilder::smoothMesh in place: one implicit step
 * (M - 0.001 L) U = M V (L the cotangent Laplacian, M the barycentric mass
 * matrix of the input), then U /= sqrt(total area of U) (mesh_builder.h:
 * 215-246).
 * mpm_mesh_decimate: igl::decimate with its default cost and placement
 * (mesh_builder.h:203-207): shortest edge first, merged at its midpoint, only
 * where the link condition holds, until at most max_faces faces remain; U
 * (capacity nv vertices) and G (capacity nf faces) receive the result.
 * mpm_obj_write: igl::writeOBJ (v lines, then 1-based f lines).
 * mpm_bgeo_write: the particles (x, v: n xyz float triples) as the reference
 * writes them (id, position, velocity, radius 0.1) in Houdini's classic
 * .bgeo format (version 5, gzip-compressed; particle_writer.h:13-44). */
int mpm_mesh_smooth(int64_t nv, double* V, int64_t nf, const int32_t* F);
int mpm_mesh_decimate(int64_t nv, const double* V, int64_t nf, const int32_t* F, int64_t max_faces,
                      double* U, int64_t* nu, int32_t* G, int64_t* ng);
int mpm_obj_write(const char* path, int64_t nv, const double* V, int64_t nf, const int32_t* F);
int mpm_bgeo_write(const char* path, int64_t n, const float* x, const float* v);

/* ---- full-state checkpoint (no reference counterpart: the reference writes
 *      only x and v, particle_writer.h:32-37). Binary, little-endian. ---- */
int mpm_save_state(mpm_handle h, const char* path);
int mpm_load_state(mpm_handle h, const char* path);

/* ---- tuning knobs (integer options; no reference counterpart: the reference
 *      fixes its launch shapes at compile time, mpm.cu:323-329) ----
 *   "p2g_mode"   1 (default): cell-sorted particles, LDS-tiled fused kernel,
 *                tile stores reduced per node block; 0: global FP32 atomics
 *                per node in particle order (the reference's scheme)
 *   "interpolation" the InterpolationKernel of the MLS-APIC TransferScheme
 *                (include/InterpolationKernel.cuh:12-74, bound at
 *                include/mpm.cuh:24-27): MPM_INTERP_QUADRATIC (default, the
 *                reference's, constant D^-1 = 4/dx^2 I) or MPM_INTERP_CUBIC (a
 *                cubic B-spline over 4 nodes per axis whose D^-1 is the generic
 *                per-particle inverse, InterpolationKernel.cuh:21-50; no
 *                reference counterpart). The tiles of p2g_mode 1 fit the
 *                quadratic stencil: the cubic kernel runs the direct kernels
 *   "fused"      1 (default): G2P of substep n fused with P2G of substep n+1;
 *                0: the reference's order P2G, grid update, G2P
 *   "sort_every" re-sort the particles by cell every k substeps (default 16)
 *   "chunk"      particles per work item, 64..192 in steps of 64 (0 = 192)
 *   "radix_bits" digit width of the radix sort: 4, 6 or 8 (default 8)
 *   "block"      home block edge in cells: 4 (0 = default 4)
 *   "derive_f"   1 (default): when every material is WC_LIQUID and the
 *                uploaded state has F = I, Jp = 1, the fused step carries F
 *                implicitly as cbrt(Jp) I (the value it would store), saving
 *                72 bytes of particle traffic per substep; downloads and the
 *                other paths see the same F. 0: F planes read and written.
 *   "xcd_map"    1 (default): the fused kernel hands each of the 8 XCDs a
 *                contiguous run of work items (neighbouring home blocks share
 *                its L2); 0: the hardware's round-robin order
 *   "reduce_mlp" -1 (default): chosen at upload from how crowded the home
 *                blocks are; 1: the grid update reads the further work items
 *                of crowded blocks 8 tiles at a time (more registers); 0: one
 *                at a time (more wavefronts resident). Same result either way.
 *   "combine"    -1 (default): chosen at upload as for reduce_mlp; 1: before
 *                the grid update a pass sums the tiles of every home block
 *                holding several work items into its first (one workgroup per
 *                block, 16 items in flight), and the grid update reads one
 *                tile per block; 0: the grid update reads every item's tile.
 *                Same result either way (both sum in item order).
 *   "timing"    k > 0: record HIP events around the kernel phases of one
 *                substep in every k (k = 1: every substep); 0 (default): off
 *   "markers"   1: roctx ranges (rocprofiler-sdk) named "mpm_advance" around
 *                each mpm_advance call and "sort", "g2p2g", "grid_update", ...
 *                around the enqueue of each phase, for `rocprofv3
 *                --marker-trace`; host-side (when the launches were issued,
 *                not when the kernels ran). 0 (default): none
 * Results do not depend on the options beyond FP32 summation order. */
int mpm_set_option(mpm_handle h, const char* name, int64_t value);
int mpm_get_option(mpm_handle h, const char* name, int64_t* value);

/* ---- per-kernel timing (no reference counterpart) ----
 * With option "timing" = k the engine records a HIP event pair on the launch
 * stream around each kernel phase of one substep in every k; the timed
 * substeps rotate through the positions of a k-substep cycle (substep i is
 * timed when i % k == (i / k) % k, counting from the last reset). The events
 * carry no system-scope fence, so they do not slow the kernels they bracket.
 * mpm_timing_read waits for the
 * recorded events and returns the accumulated milliseconds and launch count
 * of one phase: "grid_update", "g2p2g", "p2g", "g2p", "sort" or "clear". */
int mpm_timing_read(mpm_handle h, const char* phase, double* total_ms, int64_t* launches);
int mpm_timing_reset(mpm_handle h);

/* ---- autotuner (no reference counterpart) ----
 * Times `trial_substeps` substeps (at least two sort intervals) of the
 * uploaded scene under each candidate sort interval (8, 12, 16, 24, 32, 48,
 * 64, 96 substeps) and work-item size (192 and 128 particles), each from the same
 * snapshot of the current state, after one interval of warm-up; then, at the
 * fastest of those, radix digits of 6 and 4 bits against the 8-bit ones (the
 * sort bin width: a stable sort gives the same order with any). Keeps the
 * fastest as the "sort_every", "chunk" and "radix_bits" options (tiled fused
 * mode) and puts the state and simulated time back as they were. The report
 * (JSON, truncated to report_len) lists the trials. Blocking. */
int mpm_autotune(mpm_handle h, int32_t trial_substeps, char* report, int32_t report_len, void* stream);

/* ---- kernel-level entry points used by the parity tests ---- */
/* Run only the particle-to-grid scatter of the current state into a cleared
 * grid and copy the raw grid (momentum, mass: no normalisation) to host
 * (the work of particleToGrid, mpm.cu:14-74). Does not change particles. */
int mpm_debug_p2g(mpm_handle h, float* grid, void* stream);
/* Stable cell-key sort of the current particle positions with the engine's
 * radix-sort kernels: writes the keys and the permutation (original indices
 * in sorted order). The engine itself sorts by home block with the same
 * kernels and orders each work item by cell. keys/perm hold n entries. */
int mpm_debug_sort(mpm_handle h, uint32_t* keys, uint32_t* perm, void* stream);
/* 3x3 SVD of n row-major matrices on the device (the McAdams routine the
 * reference vendors, include/svd3_cuda.h:35-1043): U, V rotations, sigma
 * sorted by magnitude with the sign of det A on sigma[2]. Host pointers. */
int mpm_debug_svd(int64_t n, const float* A, float* U, float* sigma, float* V);

/* Runtime and device description for logs: HIP runtime and driver versions,
 * device name, architecture and CU count, and whether this library's device
 * code object loads on the device (no kernel is launched). No reference
 * counterpart. Returns MPM_ERR_NOGPU when no device is visible. */
int mpm_device_info(char* buf, int32_t len);
/* Runs a fixed sequence of small device operations (copies both ways, a fill
 * kernel, the SVD kernel on three 3x3 matrices, a pinned copy) and writes
 * the status of each into buf. No reference counterpart; for bring-up logs. */
int mpm_debug_selftest(char* buf, int32_t len);

/* ---- batches of independent scenes (no reference counterpart: the
 *      reference runs one Simulation per process, src/main.cu:96-101) ----
 * A batch advances several handles in lock step. Their substeps are recorded
 * once into a HIP graph -- every scene's kernels on one of a few streams
 * forked from the batch's own stream, so independent scenes run concurrently
 * -- and replayed with one graph launch per cycle of 2 sort intervals. The
 * results equal advancing each handle on its own: bit for bit while no
 * particle takes the direct path (at the grid faces, or drifted past its
 * tile between sorts), whose global atomics add in hardware order. Handles stay
 * owned by the caller and must outlive the batch; calls on a handle between
 * batch advances are allowed (a changed scene is recorded again). */
typedef struct mpm_batch_s* mpm_batch;
int mpm_batch_create(mpm_batch* out);
int mpm_batch_destroy(mpm_batch b);
const char* mpm_batch_last_error(mpm_batch b);
/* Append a handle (uploaded, with its materials set). */
int mpm_batch_add(mpm_batch b, mpm_handle h);
int mpm_batch_count(mpm_batch b, int32_t* out);
/* Advance every handle by `substeps`, ordered after earlier work on `stream`;
 * returns without waiting (later work on `stream` is ordered after it). */
int mpm_batch_advance(mpm_batch b, int32_t substeps, void* stream);
/* "graph" (1: record and replay, 0: launch kernel by kernel), "streams"
 * (1..16, default 4), "fuse" (1: when every scene is in the steady state of
 * the tiled fused mode, record each phase of all scenes as one multi-scene
 * launch; 0: per-scene kernel chains), "multi_sort" (1: the due sorts of all
 * scenes in one launch per sort kernel; 0: per-scene kernel chains, as with
 * "fuse" 0), "chains"
 * (1..4, default 4: the scenes split into that many chains of launches,
 * balanced by particles weighted by material kind, each on a stream of its
 * own), "item_share" (1..100, default 100: percent of each scene's estimated
 * work items given a workgroup of the fused pass, whose workgroups take
 * further items of their scene in turn; lower only to exercise that path);
 * read-only "graph_launches", "fused_captures" and "fused_eager_substeps".
 * Results do not depend on these settings. */
int mpm_batch_set_option(mpm_batch b, const char* name, int64_t value);
int mpm_batch_get_option(mpm_batch b, const char* name, int64_t* value);
/* Batch tuner: times candidate settings through the batch itself (its graph
 * capture and multi-scene launches included), each from the same snapshot of
 * every scene, and keeps the fastest: every scene's own mpm_autotune choice
 * (sort interval and work-item size; the batch sorts with 8-bit digits), or
 * one sort interval (16..128) and work-item size (192; 128 at intervals 32
 * and 64) for all scenes.
 * Scenes with different sort intervals share a cycle of their least common
 * multiple. The scenes' states and times are put back. The report (JSON,
 * truncated to report_len) lists the trials in microseconds per batch
 * substep. Blocking. Every scene must be on the tiled path. */
int mpm_batch_autotune(mpm_batch b, int32_t trial_substeps, char* report, int32_t report_len,
                       void* stream);

/* ---- scene input (src/main.cu:31-66, Simulation::loadMesh / addParticles
 *      src/mpm.cu:331-394) ----
 * Loads a scene TOML file: [[material]] tables (density, E, Nu, hardening,
 * plast_clamp_lower, plast_clamp_higher with the reference's defaults;
 * extension: model = "snow" | "fixed_corotated" | "jelly" | "wc_liquid" and
 * bulk_modulus) and [[object]] tables (material, mesh under <scene dir>/meshes/,
 * size, position, velocity, lifetime_begin, lifetime_end). Each object's OBJ
 * mesh is scaled to `size`, moved to `position`, and filled by rejection
 * sampling at u32(particle_count) particles per unit volume, drawing with
 * glibc's rand() sequence for `seed` (the reference never seeds: seed 1).
 * Host memory only; no GPU needed. Errors: mpm_scene_last_error(). */
typedef struct mpm_scene_s* mpm_scene;
int mpm_scene_load(const char* toml_path, uint32_t particle_count, uint32_t seed, mpm_scene* out);
int mpm_scene_destroy(mpm_scene s);
const char* mpm_scene_last_error(void);
int mpm_scene_info(mpm_scene s, int32_t* n_materials, int32_t* n_objects, int64_t* n_particles);
/* n_materials entries, particle volume 1/particle_count (main.cu:36) */
int mpm_scene_materials(mpm_scene s, mpm_material* out);
/* object i: its material, its particles [first, first+count) of the scene's
 * particle list, and its lifetime [begin, end) in simulated seconds */
int mpm_scene_object(mpm_scene s, int32_t i, int32_t* material, int64_t* first, int64_t* count,
                     double* lifetime_begin, double* lifetime_end);
/* every object's particles in object order: x[n][3], v[n][3], material[n] */
int mpm_scene_particles_soa(mpm_scene s, float* x, float* v, uint8_t* material);
/* One mesh at a time (Simulation::addObject's loadMesh + addParticles,
 * mpm.cu:237-251,331-394): a sampler owns one generator stream, so objects
 * filled one after the other continue the sequence as the reference's global
 * rand() does. fill_mesh scales the OBJ at obj_path to `size`, moves its
 * lowest corner to `position`, samples it at u32(1/particle_volume) points
 * per unit volume and keeps the points; mpm_sampler_points copies them
 * (count x 3 floats). */
typedef struct mpm_sampler_s* mpm_sampler;
int mpm_sampler_create(uint32_t seed, mpm_sampler* out);
int mpm_sampler_destroy(mpm_sampler s);
int mpm_sampler_fill_mesh(mpm_sampler s, const char* obj_path, double size, const float position[3],
                          float particle_volume, int64_t* count);
int mpm_sampler_points(mpm_sampler s, float* x);
/* the first n outputs of the generator for `seed` (tests) */
int mpm_debug_glibc_rand(uint32_t seed, int32_t n, int32_t* out);

/* ---- differentiable substep (no reference counterpart: the reference has no
 *      gradients) ----
 * One substep of the fixed-corotated model (MMFixedCorotated,
 * MaterialModel.cuh:56-61, with J = det F) under compat FIXED, with
 * per-particle Lame parameters mu[n], lambda[n], one particle mass and volume,
 * and no plasticity; and its adjoint. Stateless: every buffer is a device
 * pointer the caller owns. State is plane-major, 24 planes of n floats:
 * x0 x1 x2, v0 v1 v2, C (9, row-major), F (9, row-major); plane q of particle
 * p at state[q*n + p]. params->compat must be MPM_COMPAT_FIXED.
 * Scratch sizes (floats) come from mpm_diff_sizes: `grid` (4 N^3) is scratch
 * of one call; `tape` (8 N^3) is written by the forward and read by the
 * backward of the same substep. Grids and tapes must be 16-byte aligned. */
int mpm_diff_sizes(const mpm_params* params, int64_t* grid_floats, int64_t* tape_floats);
/* state_out = substep(state); state_out must not alias state. */
int mpm_diff_substep(const mpm_params* params, float particle_mass, float particle_volume,
                     int64_t n, const float* state, float* state_out, const float* mu,
                     const float* lambda, float* grid, float* tape, void* stream);
/* mpm_diff_substep without the `grid` workspace, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * scatter grid is a static array of the library, one per device, sized for
 * N <= 64 (4 MiB, held until the process ends); N > 64 returns
 * MPM_ERR_CAPACITY. Nothing is allocated: the caller passes the inputs and the
 * outputs (state_out, tape). Every call on a device shares that device's
 * array, so calls that overlap on one device (concurrent streams, several
 * threads) race; the array is cleared by a kernel at the start of each call. */
int mpm_diff_substep_noalloc(const mpm_params* params, float particle_mass, float particle_volume,
                             int64_t n, const float* state, float* state_out, const float* mu,
                             const float* lambda, float* tape, void* stream);
/* Given adj_out = dL/d(state_out) of the substep that read `state` (and wrote
 * `tape`), writes adj_in = dL/d(state) and ADDS dL/dmu, dL/dlambda into
 * adj_mu[n], adj_lambda[n]. adj_in must not alias adj_out. */
int mpm_diff_substep_backward(const mpm_params* params, float particle_mass,
                              float particle_volume, int64_t n, const float* state,
                              const float* mu, const float* lambda, const float* tape,
                              const float* adj_out, float* adj_in, float* adj_mu,
                              float* adj_lambda, float* grid, void* stream);

/* A run of `steps` >= 1 differentiable substeps in one call: the kernels of
 * mpm_diff_substep, two launches per substep, all enqueued here. From
 * state0, the state after substep j+1 goes to states + j*state_stride
 * (state_stride >= 24 n floats; state0 must not lie in that output) and its
 * grid tape to tapes + j*tape_stride (tape_stride a multiple of 4 >= 8 N^3 floats, or 0:
 * every substep writes the same tape, when no backward follows). grid: 4 N^3
 * floats of scratch. No reference counterpart. */
int mpm_diff_forward_steps(const mpm_params* params, float particle_mass, float particle_volume,
                           int64_t n, int32_t steps, const float* state0, float* states,
                           int64_t state_stride, const float* mu, const float* lambda, float* grid,
                           float* tapes, int64_t tape_stride, void* stream);
/* The adjoint of a run of `steps` substeps: states + j*state_stride is the
 * state before substep j (j = 0..steps-1) and tapes + j*tape_stride its tape,
 * as mpm_diff_forward_steps wrote them (tape_stride a multiple of 4 >= 8 N^3
 * floats when steps > 1); adj_out is dL/d(state after the last
 * substep). Writes adj_in = dL/d(state before substep 0) and ADDS dL/dmu,
 * dL/dlambda into adj_mu[n], adj_lambda[n]. Scratch: grid 8 N^3 floats, work
 * 24 n floats (adj_in may be work; neither may alias adj_out). Two launches
 * per substep. No reference counterpart. */
int mpm_diff_backward_steps(const mpm_params* params, float particle_mass, float particle_volume,
                            int64_t n, int32_t steps, const float* states, int64_t state_stride,
                            const float* mu, const float* lambda, const float* tapes,
                            int64_t tape_stride, const float* adj_out, float* adj_in, float* adj_mu,
                            float* adj_lambda, float* grid, float* work, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MPM_HIP_H */
