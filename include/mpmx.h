/*
 * mpmx.h — C-ABI of the MI355X-native extended Matter MPM step.
 *
 * The reference has no FFI layer: its boundary is the C++ class `Simulation`
 * (matter/src/simulation/simulation.hpp:25-281), driven by the renderer adapter
 * `MPMIntegrationSim` (raytracing/src/mpm/MPMIntegrationSim.cpp:16-116) and by
 * tests through public members + `simulate()`.  This header is the thin C layer
 * that a `Simulation` host facade (pigment_mixing_amd/csrc/host/simulation.hpp)
 * calls; every entry point cites the reference member it replaces.
 *
 * Conventions
 *   - Every function returns an mpmx_status (0 = OK).  No C++ exception crosses
 *     the ABI.  On error, mpmx_last_error(h) returns a message.
 *   - Plain pointers and sizes only.  Host arrays are borrowed for the call.
 *     Arrays are structure-of-arrays, row-major per particle:
 *       x, v      : n*3 floats   (x0,y0,z0, x1,y1,z1, ...)
 *       F, C      : n*9 floats   row-major 3x3 per particle (F[r*3+c])
 *       pig       : n*7 floats   Mixbox latent per particle
 *       scalars   : n floats
 *       active    : n bytes (0/1)
 *   - Device memory is owned by the handle.  One handle per host thread; calls on
 *     one handle are serialised on the handle's HIP stream (a learned-closure
 *     step forks its closure onto a second stream of the handle and joins it
 *     within the step).  Not reentrant, like
 *     the reference's single GL-thread use (raytracing/src/main.cpp:270-285).
 *   - FP32 throughout (matter/src/tools.hpp:19, `typedef float T`), 3-D, quadratic
 *     B-spline (tools.hpp:20-21).
 */
#ifndef MPMX_H
#define MPMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMX_ABI_VERSION 1

typedef enum {
  MPMX_OK = 0,
  MPMX_ERR_ARG = 1,       /* invalid argument / size                              */
  MPMX_ERR_HIP = 2,       /* HIP runtime error (no device, launch failure, ...)   */
  MPMX_ERR_STATE = 3,     /* call out of order (e.g. step before prepare)         */
  MPMX_ERR_DEVICE = 4,    /* device status word set (reference `exit = 1`)        */
  MPMX_ERR_CAPACITY = 5,  /* exceeds a fixed capacity                              */
  MPMX_ERR_IO = 6         /* file could not be read / written                      */
} mpmx_status;

/* Boundary conditions — reference enum class BC { NoSlip, SlipStick, SlipFree }
 * (matter/src/tools.hpp:59). */
typedef enum { MPMX_BC_NOSLIP = 0, MPMX_BC_SLIPSTICK = 1, MPMX_BC_SLIPFREE = 2 } mpmx_bc;

/* Plate types — reference enum class PlateType (matter/src/tools.hpp:55). */
typedef enum {
  MPMX_PLATE_TOP = 0, MPMX_PLATE_BOTTOM = 1, MPMX_PLATE_LEFT = 2,
  MPMX_PLATE_RIGHT = 3, MPMX_PLATE_FRONT = 4, MPMX_PLATE_BACK = 5
} mpmx_plate_type;

/* Plastic models on the path (reference enum class PlasticModel, tools.hpp:57):
 * VMVisc, von Mises/Perzyna (plasticity.cpp:163-238), the benchmark material;
 * DPVisc, Drucker-Prager/Perzyna with Pradhana's volume correction
 * (plasticity.cpp:240-340), the paint app's material (simulation_units.cpp:127-129). */
typedef enum { MPMX_PLASTIC_NONE = 0, MPMX_PLASTIC_VMVISC = 5, MPMX_PLASTIC_DPVISC = 6 } mpmx_plastic_model;

/* Return-map closure used inside the step. */
typedef enum {
  MPMX_CLOSURE_EXACT = 0,    /* von Mises/Perzyna return map (plasticity.cpp:163-238) */
  MPMX_CLOSURE_LEARNED = 1,  /* BF16 MLP on MFMA predicts (h', dgamma); accepted when
                                the predicted state's yield residual is within
                                closure_tol * q_y, else the exact map (fallback)     */
  MPMX_CLOSURE_LEARNED_REFINED = 2 /* as 1, but the predicted dgamma starts one Newton
                                step on the consistency residual and h' follows the
                                flow rule; same acceptance test and fallback          */
} mpmx_closure;

/* Knife (palette knife / spatula) shapes. */
typedef enum {
  MPMX_KNIFE_SLAB = 0,       /* exact box SDF, half extents hx,hy,hz                 */
  MPMX_KNIFE_SPATULA = 1     /* reference trapezoid SDF (object_spatula.hpp:40-82)   */
} mpmx_knife_shape;

/* Knife kinematics (replaces ObjectSpatula::move keyframes, object_spatula.hpp:164-233). */
typedef enum {
  MPMX_MOTION_STATIC = 0,
  MPMX_MOTION_LINEAR = 1,    /* centre = p0 + velocity * t                           */
  MPMX_MOTION_CIRCULAR = 2,  /* centre on a circle in the xz plane, heading tangent */
  MPMX_MOTION_LINEAR_THEN_CIRCULAR = 3, /* linear for t < t_switch, then circular    */
  MPMX_MOTION_KEYFRAMES = 4  /* 60 fps pos+quat frames, lerp/slerp (spatula_motion.bin) */
} mpmx_motion;

/* Scene/material configuration.  Field names follow Simulation's public members
 * (simulation.hpp:38-126) and the keys of sim_config.json / pigment_config.json
 * (loaded at simulation_units.cpp:32-80). */
typedef struct {
  /* material — sim_config.json: E, nu, rho, flip_ratio, perzyna_exp, perzyna_visc */
  float E, nu, rho;
  float flip_ratio;          /* [-1,0): AFLIP blend, [0,1]: PIC/FLIP (simulation.hpp:64) */
  int   plastic_model;       /* mpmx_plastic_model                                    */
  float q_max, q_min;        /* von Mises yield (Pa), simulation.hpp:94-95            */
  float xi;                  /* softening (simulation.hpp:91)                         */
  float p_min;               /* capped von Mises (simulation.hpp:96)                  */
  float perzyna_exp, perzyna_visc; /* simulation.hpp:103-104; visc is a time (s)      */
  float q_prefac;            /* 1/sqrt(2) (simulation.hpp:118)                        */
  float cfl;                 /* sim_config.json "cfl" (used by adaptive dt only)      */
  float M;                   /* tan(friction_angle) (simulation_units.cpp:60-61)      */
  float q_cohesion;          /* sim_config.json "q_cohesion"                          */
  float gravity[3];
  /* pigment_config.json */
  float D_max, D_edge0, D_edge1, start_boost_time, end_boost_time, boost_factor;
  /* grid / time */
  float dx;
  float ppc;                 /* particles per cell: particle_volume = dx^3/ppc        */
  int   grid_dims[3];        /* node counts; 0 = auto from bbox (remesh.cpp:131-153)  */
  int   origin_mode;         /* 0: bbox_min - 2dx (remesh.cpp:114-137); 1: explicit   */
  float grid_origin[3];      /* used when origin_mode == 1                            */
  float grid_reference_point[3]; /* >= 1e10 disables (simulation.hpp:52)              */
  float dt;                  /* fixed time step                                       */
  float fps;                 /* frame rate for frameFinished (simulation_units.cpp:235) */
  int   use_musl;            /* simulation.hpp:50; 0 USL, 1 MUSL (musl.cpp:6-98)    */
  int   use_particle_boundaries;           /* simulation.hpp:55-57                    */
  float particle_boundary_min[3], particle_boundary_max[3];
  int   closure;             /* mpmx_closure                                          */
  float closure_tol;         /* relative yield-residual tolerance for the fallback    */
  int   n_threads;           /* CPU threads (oracle/baseline only; ignored on GPU)    */
  int   use_pradhana;        /* DPVisc volume correction (simulation.hpp:86, default 1) */
} mpmx_config;

/* Axis-aligned plate — ObjectPlate (matter/src/objects/object_plate.hpp:13-78).
 * A plate with a velocity moves at the end of every step as ObjectPlate::move
 * does (object_plate.hpp:136-162): during the first load_factor frames
 * (frame = 1 / fps) at v / vmin_factor, later at v; pos_object advances along
 * the plate's normal axis, pos_lower / pos_upper along its tangential one.
 * vmin_factor <= 0 is taken as 1. */
typedef struct {
  float pos_object;
  int   plate_type;          /* mpmx_plate_type */
  int   bc;                  /* mpmx_bc         */
  float friction;
  float pos_lower, pos_upper;
  float v[3];                /* vx_object, vy_object, vz_object (at construction) */
  float vmin_factor;         /* object_plate.hpp:13 (default 1)                  */
  float load_factor;         /* object_plate.hpp:13 (default 0)                  */
} mpmx_plate;

/* Knife description + motion. */
typedef struct {
  int   shape;               /* mpmx_knife_shape */
  float half[3];             /* slab half extents (m); spatula: halfWidth, halfThickness, halfLength */
  float offset;              /* spatula offset (object_spatula.hpp:23)                */
  int   bc;                  /* mpmx_bc         */
  float friction;
  int   motion;              /* mpmx_motion     */
  float p0[3];               /* centre at t = 0 (linear), or circle centre (circular) */
  float tilt_axis[3];        /* body rotation applied before the heading rotation     */
  float tilt_angle;          /* radians                                               */
  float velocity[3];         /* linear motion velocity (m/s)                          */
  float radius;              /* circular motion radius (m)                            */
  float speed;               /* circular motion speed (m/s)                           */
  float phase0;              /* circular start angle (rad)                            */
  float t_switch;            /* LINEAR_THEN_CIRCULAR: switch time (s)                 */
} mpmx_knife;

/* Per-step knife state, as held by ObjectSpatula (transform, invTransform,
 * velocity, angularVelocity; object_spatula.hpp:245-270).  Matrices are 3x4
 * row-major (rotation | translation). */
typedef struct {
  float T[12];               /* local -> world                                        */
  float Tinv[12];            /* world -> local                                        */
  float Nmat[9];             /* (T.linear^-1)^T, row-major (object_spatula.hpp:114)   */
  float v[3];                /* vx_object, vy_object, vz_object                       */
  float w[3];                /* angularVelocity                                       */
} mpmx_knife_state;

/* Counters surfaced by mpmx_stats — reference "Proj: plastic_count / Np"
 * (deformation_update.cpp:49-50), the `exit` flag (simulation.hpp:30), and timings
 * (simulation.cpp:210-215). */
typedef struct {
  int64_t step;              /* completed steps                                       */
  double  time;              /* simulated time (s)                                    */
  int64_t n_particles;
  int64_t n_active;          /* active particles                                      */
  int64_t plastic_count;     /* particles projected in the last step                  */
  int64_t fallback_count;    /* learned closure: particles sent to the exact map, last step */
  int64_t active_nodes;      /* nodes with mass > 0, mean over the last mpmx_step call */
  int64_t active_blocks;     /* 4x4x4 node blocks touched, mean over the last call    */
  int64_t stray_count;       /* particles scattered outside their LDS tile, last call */
  uint32_t device_error;     /* bit mask: 1 dgamma < 0, 2 Newton, 4 unknown boundary  *
                              * condition (the reference's exit = 1 causes); internal *
                              * checks: 8 a particle moved > 1 cell in a step, 16 a   *
                              * closure-queue segment overflowed                      */
  float   ms_last_step;      /* wall ms of the last mpmx_step call / its steps        */
  int64_t learned_count;     /* learned closure: predictions accepted in the last step */
  int64_t newton_iters;      /* exact map: Newton iterations (plasticity.cpp:196-229), sum over the last call */
} mpmx_stats;

typedef struct mpmx_handle mpmx_handle;

/* ---------------- lifecycle ---------------- */

/* Defaults = Simulation's constructor values (simulation.hpp:38-126) with
 * the reference files' JSON defaults; replaces Simulation::Simulation(). */
void mpmx_config_default(mpmx_config* cfg);

/* Load sim_config.json / pigment_config.json into cfg as
 * Simulation::setupScene does (simulation_units.cpp:32-80): a key missing from
 * a readable file takes the reference's fallback value (e.g. cfl 0.2,
 * flip_ratio -0.8, E 5e5); an unreadable sim_config.json sets the same
 * fallbacks; an unreadable pigment_config.json leaves the pigment fields as
 * they are.  Either path may be NULL (skipped).  Extra keys understood beyond
 * the reference's: "yield" (q_max = q_min), "viscosity" (Pa s; perzyna_visc =
 * viscosity / yield), "q_max", "q_min", "xi", "dt", "dx", "ppc", "use_musl",
 * "gravity_y".  Returns MPMX_ERR_IO if a given file cannot be read or parsed. */
int mpmx_config_load_json(mpmx_config* cfg, const char* sim_config_path,
                          const char* pigment_config_path);

/* Create a handle on the current HIP device (Simulation::initializeBasic,
 * simulation_units.cpp:9-26). */
int mpmx_create(const mpmx_config* cfg, mpmx_handle** out);
void mpmx_destroy(mpmx_handle* h);
const char* mpmx_last_error(const mpmx_handle* h);
const char* mpmx_status_string(int status);

/* ---------------- scene ---------------- */

/* Upload particle state (host SoA, borrowed).  Any of v, F, C, pig, diff,
 * eps_pl_dev, active may be NULL (defaults: 0, I, 0, 0, 0, 0, 1), as in
 * Particles(Np) (data_structures.hpp:11-29).  particle_mass = rho*dx^3/ppc
 * (sampling_particles.hpp:28-29). */
int mpmx_upload_particles(mpmx_handle* h, int64_t n, const float* x, const float* v,
                          const float* F, const float* C, const float* pig,
                          const float* diff, const float* eps_pl_dev,
                          const uint8_t* active);

/* The volumetric plastic strains eps_pl_vol and eps_pl_vol_pradhana of a
 * scene's particles (Particles, data_structures.hpp:30-39; zero when not
 * given), for a restart from a saved state.  Call after the scene's particle
 * upload and before mpmx_prepare; scene 0 for a single-scene handle.  Either
 * pointer may be NULL (zeros). */
int mpmx_upload_plastic_state(mpmx_handle* h, int scene, const float* eps_pl_vol,
                              const float* eps_pl_vol_pradhana);

/* Plates and knives (Simulation::plates / objects, simulation.hpp:129-131;
 * setupScene adds the floor and the spatula, simulation_units.cpp:117-122).
 * At most 8 plates and 1 knife. */
int mpmx_set_objects(mpmx_handle* h, const mpmx_plate* plates, int n_plates,
                     const mpmx_knife* knives, int n_knives);

/* Keyframes for MPMX_MOTION_KEYFRAMES: n frames of {pos[3], quat_xyzw[4]} at
 * `fps` (ObjectSpatula::loadAnimation, object_spatula.hpp:235-250). */
int mpmx_set_knife_keyframes(mpmx_handle* h, int k, const float* frames7, int n_frames,
                             float fps);

/* Override the knife state used by the next step (driven externally, e.g. by an
 * app per frame).  Clears after one step; kinematics resume afterwards. */
int mpmx_set_knife_state(mpmx_handle* h, int k, const mpmx_knife_state* s);

/* Knife state that step `step` uses (host kinematics; for tests/renderer). */
int mpmx_knife_state_at(mpmx_handle* h, int k, int64_t step, mpmx_knife_state* out);

/* Simulation::updateDt (update_dt.cpp:5-71) as an option: the caller computes
 * dt = min(cfl dx / max speed, dt_max), clamped to the frame end and min_dt,
 * from mpmx_max_speed and sets it with mpmx_set_dt before the next step.
 *   mpmx_max_speed: largest |v| of the live particles (all scenes of a batch),
 *     a device reduction read back to the host (synchronises the stream).
 *   mpmx_set_dt: the step of every scene from the next step on (a batch
 *     shares one dt); the derived constants follow, the knife's clock and the
 *     plates' positions continue from the current step (re-based there with
 *     the old dt), and the captured step graph stays valid (it reads the
 *     scene constants from device memory).  Valid before and after prepare. */
int mpmx_max_speed(mpmx_handle* h, float* vmax);
int mpmx_set_dt(mpmx_handle* h, float dt);

/* Same kinematics without a handle (no device needed): knife `k`, optional
 * keyframes (n_frames x {pos[3], quat_xyzw[4]} at fps), time step dt. */
int mpmx_knife_state_eval(const mpmx_knife* k, const float* frames7, int n_frames, float fps, double dt,
                          int64_t step, mpmx_knife_state* out);

/* Fix the lattice and derived constants (Simulation::prepareSimulation,
 * simulation_units.cpp:133-215): origin rule (remesh.cpp:45-180), lambda, mu,
 * K, prefactors (simulation_units.cpp:167-183), then bin the particles and run
 * the first P2G + grid update so that mpmx_step can start. */
int mpmx_prepare(mpmx_handle* h);

/* Lattice actually used: dims[3], origin (= grid.xc/yc/zc) and node coordinate
 * tables xs[dims[0]], ys[dims[1]], zs[dims[2]] (may be NULL). */
int mpmx_get_lattice(mpmx_handle* h, int* dims, float* origin, float* xs, float* ys,
                     float* zs);

/* ---------------- stepping ---------------- */

/* Advance n steps (Simulation::step / advanceStep, simulation_units.cpp:217-233,
 * simulation.cpp:225-280).  Returns MPMX_ERR_DEVICE if the device status word
 * is set (the reference's exit flag), after completing the steps. */
int mpmx_step(mpmx_handle* h, int nsteps);

/* Frame bookkeeping (simulation_units.cpp:235-245). */
int mpmx_frame_finished(mpmx_handle* h, int* finished);

int mpmx_synchronize(mpmx_handle* h);

/* ---------------- output ---------------- */

/* Download particle state in the caller's original particle order.  Any
 * pointer may be NULL. delta_gamma is plasticity.cpp:235's rate. */
int mpmx_download_particles(mpmx_handle* h, float* x, float* v, float* F, float* C,
                            float* pig, float* diff, float* eps_pl_dev,
                            float* delta_gamma, uint8_t* active);

/* The volumetric plastic strains in the caller's original particle order:
 * eps_pl_vol and eps_pl_vol_pradhana (Particles, data_structures.hpp:30-39),
 * written by the DPVisc map (plasticity.cpp:268-275) and the capped von Mises
 * branch (plasticity.cpp:172-180).  Zero for scenes whose model does not track
 * them.  Either pointer may be NULL. */
int mpmx_download_plastic_state(mpmx_handle* h, float* eps_pl_vol, float* eps_pl_vol_pradhana);

/* Particle output as Simulation::saveParticleData (save_data.cpp:7-166):
 * binary little-endian PLY, element `vertex` with one float property per
 * field, in the caller's original particle order: x y z vx vy vz, pressure p
 * = -tr(tau)/3, deviatoric stress q = q_prefac |dev tau| and Je = det F, then
 * eps_pl_vol eps_pl_dev delta_gamma when a plastic model is set (tau is the
 * Hencky Kirchhoff stress of F).  Single-scene handles. */
int mpmx_write_particles_ply(mpmx_handle* h, const char* path);

/* Grid output as Simulation::saveGridData (save_data.cpp:168-238): every node
 * of the lattice in Simulation::ind order, x y z (node coordinates) vx vy vz
 * (post-collision velocity of the last grid update) and mass. */
int mpmx_write_grid_ply(mpmx_handle* h, const char* path);

/* Resumable checkpoint of a single-scene handle: the whole particle state
 * (including eps_pl_vol, the Pradhana accumulation and the Kirchhoff stress
 * rows), the step count and clocks, the lattice (saved as an explicit origin
 * and node counts), plates, knife and keyframes.  mpmx_checkpoint_load creates
 * a prepared handle on the current device that continues from that step.
 * The learned closure's network is not part of it (set it again after
 * loading).  No counterpart in the reference, which restarts from PLY frames
 * only. */
int mpmx_checkpoint_save(mpmx_handle* h, const char* path);
int mpmx_checkpoint_load(const char* path, mpmx_handle** out);

/* Renderer format (MPMIntegrationSim::recountParticles,
 * MPMIntegrationSim.cpp:100-116): pos4 = (x,y,z,1), lat7, diff — host arrays in
 * original particle order. */
int mpmx_pack_render(mpmx_handle* h, float* pos4, float* lat7, float* diff);

/* Same, into device memory (e.g. a mapped GL/VK buffer); no host copy. */
int mpmx_pack_render_device(mpmx_handle* h, void* d_pos4, void* d_lat7, void* d_diff);

/* Grid after the last grid update: dense node arrays (Nx*Ny*Nz, index
 * (i*Ny+j)*Nz+k as Simulation::ind, simulation.hpp:236-238).  mass is the
 * node mass m_i (kg), v the post-BC velocity, flip = v_new - v_old, pig the
 * normalised latent, shear the normalised shear intensity.  Any may be NULL. */
int mpmx_get_grid(mpmx_handle* h, float* mass, float* v, float* flip, float* pig,
                  float* shear);

int mpmx_stats_get(mpmx_handle* h, mpmx_stats* out);

/* Per-kernel timing: mean device duration (ms) of each kernel of the step,
 * measured with HIP events on the handle's stream. */
typedef struct {
  float ms_g2p;              /* G2P + F update + return map + advection (k_g2p) */
  float ms_sort;             /* block/cell prefix sums + permutation */
  float ms_p2g;              /* P2G scatter of mass, momentum, force, latent, shear */
  float ms_grid;             /* grid update + boundary collision */
  float ms_step;             /* whole step */
  int64_t steps;
  double particles_per_step; /* live particles averaged over the steps */
  double active_nodes_per_step;
  float ms_closure;          /* learned closure over the queued candidates (k_closure), 0 without it */
  double queued_per_step;    /* plastic candidates queued for the learned closure, averaged */
} mpmx_kernel_times;

/* Advance nsteps without the HIP graph, timing every kernel with events
 * (no reference counterpart: measurement for the roofline). */
int mpmx_profile_steps(mpmx_handle* h, int nsteps, mpmx_kernel_times* out);

/* ---------------- synthetic scenes ---------------- */

/* A paint blob: an axis-aligned box of cells[0]*cells[1]*cells[2] cells of size
 * dx starting at lo, with one Mixbox latent.  Replaces the VDB blob sampling of
 * the paint app (simulation_blobs.cpp:7-150) for the benchmark scenes. */
typedef struct {
  float lo[3];
  int   cells[3];
  float latent[7];
} mpmx_box;

/* Particles of the boxes: per cell per_axis^3 sub-cells, one particle each at
 * lo + (cell + (sub + u)/per_axis) * dx (float arithmetic in that order), u a
 * uniform [0,1) from a counter-based hash of (seed, particle index, axis) —
 * splitmix64(seed * 0x100000001B3 + 3*index + axis) >> 40, times 2^-24.
 * Particles are numbered box by box, cell (i,j,k) i-major, sub-cell i-major.
 * Writes x (n*3) and pig (n*7, may be NULL).  *n_out = particle count; pass
 * x = NULL to query it. */
int mpmx_scene_boxes(const mpmx_box* boxes, int n_boxes, float dx, int per_axis, uint64_t seed,
                     float* x, float* pig, int64_t capacity, int64_t* n_out);

/* ---------------- batched scenes ---------------- */

/* A handle advancing n_scenes (<= 64) independent scenes together: every
 * kernel of a step walks the blocks of all scenes, and one HIP graph holds
 * two steps of the whole batch.  Scenes differ in material, lattice, objects
 * and particles; they share dt and the closure mode.  Particles are staged per
 * scene (mpmx_upload_scene) and laid out in scene order, so downloads return
 * scene 0's particles first; mpmx_scene_range gives each scene's slice.
 * Knife index k of the knife calls is the scene index.  No counterpart in the
 * reference (one Simulation per process there). */
int mpmx_create_batch(const mpmx_config* cfgs, int n_scenes, mpmx_handle** out);
int mpmx_upload_scene(mpmx_handle* h, int scene, int64_t n, const float* x, const float* v, const float* F,
                      const float* C, const float* pig, const float* diff, const float* eps,
                      const uint8_t* active);
int mpmx_set_objects_scene(mpmx_handle* h, int scene, const mpmx_plate* plates, int n_plates,
                           const mpmx_knife* knives, int n_knives);
int mpmx_scene_range(mpmx_handle* h, int scene, int64_t* first, int64_t* count);

/* Launch-shape autotuner: times candidate workgroups-per-CU for P2G, G2P and
 * the grid update (coordinate descent, steps_per_trial ordinary steps each)
 * and keeps the fastest.  The steps it runs advance the simulation. */
typedef struct {
  int wg_p2g, wg_g2p, wg_grid;   /* chosen workgroups per CU                  */
  float ms_p2g, ms_g2p, ms_grid; /* mean kernel ms at the chosen shapes       */
  int trials;
  int64_t steps;                 /* simulation steps spent tuning             */
} mpmx_tune_report;
int mpmx_autotune(mpmx_handle* h, int steps_per_trial, mpmx_tune_report* report);
/* Set the launch shapes directly (workgroups per CU, each in 1..64), e.g. to
 * reuse an earlier autotune of the same workload.  Shapes change no result
 * (every sum has a fixed order).  No counterpart in the reference. */
int mpmx_set_launch_shape(mpmx_handle* h, int wg_p2g, int wg_g2p, int wg_grid);

/* ---------------- learned return-map closure ---------------- */

/* MLP 8 -> 64 -> 64 -> 64 -> 64 -> 4, SiLU, BF16 weights, FP32 accumulate.
 * Parameters are a flat FP32 master copy laid out layer by layer as W[out][in]
 * then b[out]; mpmx_closure_n_params() gives the count. */
int64_t mpmx_closure_n_params(void);
int mpmx_closure_set_params(mpmx_handle* h, const float* params, int64_t n);
int mpmx_closure_get_params(mpmx_handle* h, float* params, int64_t n);
/* Input/output normalisation: x_n = (x - in_mean)/in_std, y = y_n*out_std + out_mean. */
int mpmx_closure_set_norm(mpmx_handle* h, const float* in_mean, const float* in_std,
                          const float* out_mean, const float* out_std);

/* Harvest (trial, corrected) pairs from the exact solver: while enabled, each
 * step appends up to `capacity` records of 12 floats
 * (8 inputs: h1,h2,h3 (trial principal Hencky strains), ||dev h||, q_y,
 *  eta = q_y * perzyna_visc (Pa s), dt*E, nu; 4 targets: h1',h2',h3' (returned
 *  strains), dgamma (plastic increment)) for particles whose trial state is
 *  plastic (plasticity.cpp:163-238). */
int mpmx_harvest_begin(mpmx_handle* h, int64_t capacity);
int mpmx_harvest_end(mpmx_handle* h, float* out12, int64_t max_records, int64_t* n_out);

/* Training of the closure on the device (own forward/backward MFMA kernels,
 * Adam, cosine learning-rate decay from lr to lr_final).  The network learns
 * normalised (h' - h_trial, dgamma) from the 8 inputs; normalisation is the
 * per-column mean / standard deviation of the records. */
typedef struct {
  int64_t steps;       /* Adam steps                                         */
  int   batch;         /* samples per step, a positive multiple of 64        */
  float lr, lr_final;  /* learning rate, cosine-decayed to lr_final          */
  float beta1, beta2, eps;
  uint64_t seed;       /* parameter init and minibatch sampling              */
  int   init;          /* 1: seeded Xavier init; 0: continue from the handle's parameters */
  int   report_every;  /* steps between minibatch-loss samples               */
} mpmx_train_config;

typedef struct {
  float loss_first, loss_last;  /* minibatch MSE (normalised) at the first / last step */
  float loss_full;              /* MSE over all records with the final parameters      */
  int64_t steps, n_curve;       /* steps run; loss samples written to `curve`          */
  double ms;                    /* wall time of the training loop                      */
  float norm[24];               /* in_mean[8] in_std[8] out_mean[4] out_std[4]         */
} mpmx_train_report;

void mpmx_train_config_default(mpmx_train_config* tc);
/* Trains on n harvest records (12 floats each, layout as above) and installs
 * the result (parameters + normalisation) in the handle.  No counterpart in
 * the reference (the learned closure is this build's addition). */
int mpmx_closure_train(mpmx_handle* h, const float* records12, int64_t n, const mpmx_train_config* tc,
                       mpmx_train_report* report, float* curve, int64_t curve_cap);

/* The closure's accuracy in principal space: for n trial states (h1 h2 h3 of
 * the trial principal Hencky strain, eps_pl_dev) runs the learned map (the
 * handle's closure mode, tolerance and network) and the exact map
 * (plasticity.cpp:163-238) with scene 0's material.  learned4 / exact4 get
 * (h1' h2' h3' dgamma) per state; learned4 is NaN where the prediction was not
 * accepted (accepted[i] = 0).  Needs a prepared handle with a learned closure
 * mode and parameters.  No counterpart in the reference. */
int mpmx_closure_eval(mpmx_handle* h, const float* trial4, int64_t n, float* learned4, float* exact4,
                      uint8_t* accepted);

/* boundaryCollision (boundary_collision.cpp:5-316) of the device's grid
 * update, evaluated on n given grid nodes: positions X3, post-force
 * velocities v3 -> out3, against `plates` and the knife (NULL: none) at its
 * state of step `step`, with cfg's dt and dx.  The reference applies it per
 * node inside explicitEulerUpdate (explicit_euler_update.cpp:94) and has no
 * separate entry; this one exists to check the device rules node by node.
 * MPMX_ERR_DEVICE if a rule reported an invalid boundary condition. */
int mpmx_boundary_eval(const mpmx_config* cfg, const mpmx_plate* plates, int n_plates, const mpmx_knife* knife,
                       int64_t step, const float* X3, const float* v3, int64_t n, float* out3);

/* Handle-free evaluation of the network on the device, for checks against a
 * CPU implementation: norm24 = in_mean[8] in_std[8] out_mean[4] out_std[4].
 * mpmx_mlp_forward: y4 = network(x8) in output units.
 * mpmx_mlp_gradient: loss = mean over samples and outputs of the squared
 * normalised error, and its gradient with respect to the parameters. */
int mpmx_mlp_forward(const float* params, const float* norm24, const float* x8, int64_t n, float* y4);
int mpmx_mlp_gradient(const float* params, const float* norm24, const float* x8, const float* t4, int64_t n,
                      float* grad, float* loss);

/* mpmx_mlp_forward without allocation: the same result as mpmx_mlp_forward
 * (same kernels) for n <= MPMX_MLP_NOALLOC_MAX samples, computed in a fixed
 * device array of the library (about 3.2 MB, held by every device the library
 * runs on until the process ends) on the device's null stream; it calls no
 * hipMalloc / hipFree / stream creation.  n > MPMX_MLP_NOALLOC_MAX returns
 * MPMX_ERR_CAPACITY before any device work.  Calls on one device share the
 * array and the stream: calls that overlap in time on one device race.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc). */
#define MPMX_MLP_NOALLOC_MAX 65536
int mpmx_mlp_forward_noalloc(const float* params, const float* norm24, const float* x8, int64_t n, float* y4);

/* ---- scene initialisation: particle sampling of blob objects (host code).
 *
 * A blob object's level set as a dense signed-distance grid: phi[(i*dims[1] +
 * j)*dims[2] + k] at world point origin + (i, j, k) * voxel, negative inside;
 * the object is placed as X = p * scale + offset.  This is what the
 * reference's ObjectVdb (object_vdb.hpp:30-118) holds over its active voxel
 * bounding box; reading .vdb files (OpenVDB) is not part of this library. */
typedef struct mpmx_level_set {
  const float* phi;
  int32_t dims[3];
  float voxel;
  float origin[3];
  float offset[3];
  float scale;
} mpmx_level_set;

/* Poisson-disk points with minimum distance `radius` in the box [lo, hi]
 * (Bridson's algorithm, `attempts` candidates per active sample, seeded),
 * replacing thinks::PoissonDiskSampling (tph_poisson 0.3) as
 * sampling_particles_vdb.hpp:33 and :113-117 call it.  *n_out = number of
 * points; MPMX_ERR_CAPACITY when it exceeds cap (nothing written). */
int mpmx_sample_poisson_box(const float lo[3], const float hi[3], float radius, int32_t attempts, uint32_t seed,
                            float* xyz, int64_t cap, int64_t* n_out);

/* The multi-object sampler of sampling_particles_vdb.hpp:86-160: per object
 * i, box samples over its bounds (object_vdb.hpp:107-118) with seed + i, kept
 * where the trilinear level set is <= 0 (object_vdb.hpp:73-85) and tagged
 * object[p] = i (the reference's pigment index); *dx_out = cbrt(ppc / (all box
 * samples) * (sum of box volumes)), the reference's dx rule (volumes summed
 * and the cube root taken in float, as there).  The reference calls it with
 * radius 0.01, 30 attempts, seed 42, ppc 27 (simulation_blobs.cpp:149); the
 * scene's particle volume is then dx^3 / ppc (sampling_particles_vdb.hpp:148),
 * so mpmx_config.ppc must be the ppc passed here. */
int mpmx_sample_level_sets(const mpmx_level_set* objects, int32_t n_objects, float radius, int32_t attempts,
                           uint32_t seed, float ppc, float* xyz, int32_t* object, int64_t cap, int64_t* n_out,
                           float* dx_out);

/* ---- the paint app's blob scene (host code).
 *
 * One model of the blob database (simulation.hpp:272-279): its level set as a
 * dense grid (offset and scale are ignored; the models share one voxel size,
 * with origins on that voxel lattice), its volume (m^3, ascending over the
 * database) and its calibration factor. */
typedef struct mpmx_blob_model {
  mpmx_level_set shape;
  float volume;
  float calibration;
} mpmx_blob_model;

/* Where mpmx_blob_scene put blob i: the database pair (pair, pair + 1) it
 * morphs and the blend t, the target volume, the packing radius, the offset
 * applied to the morphed level set and its placed bounding box. */
typedef struct mpmx_blob_placement {
  int32_t pair;
  float t;
  float target_volume;
  float radius;
  float offset[3];
  float bbox_min[3], bbox_max[3];
} mpmx_blob_placement;

/* Simulation::blobs (simulation_blobs.cpp:7-150): per colour ratio (clamped to
 * [0.1, 0.9]) a target volume between the database's smallest and largest,
 * the bracketing pair (getMorphPair, :152-163), t by 15 bisection steps on the
 * voxel-counted volume times the blended calibration factor
 * (findTForExactVolume, :165-218), the models blended as (1 - t) A + t B over
 * the union of their extents (values outside a model's extent: 3 voxels,
 * positive); the blobs placed touching in x-z (n = 1..4 by the reference's
 * closed forms, a row beyond; packing radius = half the wider x-z extent -
 * 0.055) and centred; then sampled as mpmx_sample_level_sets does (the
 * reference calls it with radius 0.01, 30 attempts, seed 42, ppc 27).
 * object[p] = the blob (colour) index of particle p.  placement (n entries)
 * may be NULL.  Two-call protocol on cap as mpmx_sample_level_sets. */
int mpmx_blob_scene(const mpmx_blob_model* db, int32_t n_db, const float* ratios, int32_t n, float radius,
                    int32_t attempts, uint32_t seed, float ppc, float* xyz, int32_t* object, int64_t cap,
                    int64_t* n_out, float* dx_out, mpmx_blob_placement* placement);

/* ObjectSpatula::loadAnimation (object_spatula.hpp:235-250): the animation
 * file is an array of FrameData {float pos[3]; float quat[4] (x, y, z, w)}
 * (object_spatula.hpp:7-10), 28 bytes each, played at 60 fps; frames7
 * receives them as the 7-float keyframes mpmx_set_knife_keyframes takes.
 * *n_out = whole records in the file; MPMX_ERR_CAPACITY (nothing written)
 * when more than cap; MPMX_ERR_IO when the file cannot be read. */
int mpmx_load_keyframes(const char* path, float* frames7, int64_t cap, int64_t* n_out);

#ifdef __cplusplus
}
#endif

#endif /* MPMX_H */
