/*
 * libwoodem_hip -- MI355X (gfx950) implementation of the Woo DEM timestep.
 *
 * C ABI: plain pointers and sizes, no C++ or torch types.  Every entry point
 * returns WDM_OK (0) or a negative WDM_E_* code; wdm_last_error() gives the
 * message.  No exception or abort crosses the boundary.  Calls on one context
 * are not re-entrant; separate contexts may be used from separate threads.
 *
 * Each entry point names the reference interface it replaces.  Paths are
 * relative to the Woo source tree (Azeko2xo/woodem).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference throws std::runtime_error / logic_error, see
        pkg/dem/IdealElPl.cpp:76-80, InsertionSortCollider.cpp:945-949, core/Cell.cpp:96) ---- */
#define WDM_OK 0
#define WDM_E_ARG -1          /* bad argument / unknown scene */
#define WDM_E_HIP -2          /* HIP runtime failure */
#define WDM_E_OOM -3          /* device allocation failed or HBM budget exceeded */
#define WDM_E_NAN -4          /* "NaN force in contact" (IdealElPl.cpp:76-80) */
#define WDM_E_OVERSIZE -5     /* "Particle larger than half of the cell size" (InsertionSortCollider.cpp:945-949) */
#define WDM_E_DEGENERATE -6   /* "Cell is degenerate (zero volume)" (core/Cell.cpp:96) */
#define WDM_E_UNSUPPORTED -7  /* input outside the implemented path (clumps, aspherical free nodes, sheared cells, adhesive Hertz) */
#define WDM_E_STATE -8        /* call order violated (e.g. step before upload) */

/* shape codes: Sphere (pkg/dem/Sphere.hpp), Wall (pkg/dem/Wall.hpp), Facet (pkg/dem/Facet.hpp) */
#define WDM_SHAPE_NONE 0
#define WDM_SHAPE_SPHERE 1
#define WDM_SHAPE_WALL 2
#define WDM_SHAPE_FACET 3

/* contact laws: Cp2_FrictMat_FrictPhys + Law2_L6Geom_FrictPhys_IdealElPl (pkg/dem/FrictMat.cpp:13-44,
   pkg/dem/IdealElPl.cpp:12-122); Cp2_FrictMat_HertzPhys + Law2_L6Geom_HertzPhys_DMT (pkg/dem/Hertz.cpp:9-182) */
#define WDM_LAW_IDEALELPL 0
#define WDM_LAW_HERTZ_DMT 1

/* arithmetic of the contact kernel: FP64 throughout, or L6Geom/Cp2/Law2 in FP32 on FP64 state */
#define WDM_PREC_FP64 0
#define WDM_PREC_MIXED 1

/* DemData.flags bits (pkg/dem/Particle.hpp:177-182) */
#define WDM_DOF_X 1u
#define WDM_DOF_Y 2u
#define WDM_DOF_Z 4u
#define WDM_DOF_RX 8u
#define WDM_DOF_RY 16u
#define WDM_DOF_RZ 32u
#define WDM_CLUMP_CLUMPED 64u
#define WDM_CLUMP_CLUMP 128u
#define WDM_ENERGY_SKIP 256u
#define WDM_GRAVITY_SKIP 512u
#define WDM_DAMPING_SKIP 2048u

/* energy slots of EnergyTracker (core/EnergyTracker.hpp:61-79), fixed order */
#define WDM_E_KINETIC 0
#define WDM_E_ELAST 1
#define WDM_E_PLAST 2
#define WDM_E_GRAV 3
#define WDM_E_NONVISCDAMP 4
#define WDM_E_VISCN 5
#define WDM_E_VISCT 6
#define WDM_E_BROKEN 7
#define WDM_E_COUNT 8

/* Engine attributes of the timestep path; layout: doubles first, then 32-bit ints.
   Sources: Leapfrog.damping (pkg/dem/Leapfrog.hpp:50), DemField.gravity/loneMask
   (pkg/dem/Particle.hpp:316-319), InsertionSortCollider.verletDist
   (pkg/dem/InsertionSortCollider.hpp:262), Bo1_Sphere_Aabb/Cg2_Sphere_Sphere_L6Geom
   distFactor (pkg/dem/Sphere.hpp), Cp2_FrictMat_HertzPhys.{poisson,en,gamma,alpha}
   (pkg/dem/Hertz.hpp:27-32), Law2_L6Geom_FrictPhys_IdealElPl.{noSlip,noBreak,noFrict}
   (pkg/dem/IdealElPl.hpp:10-26), Law2_L6Geom_HertzPhys_DMT.noAttraction (Hertz.hpp:52),
   Scene.trackEnergy, Cg2_Any_Any_L6Geom__Base.trsfRenorm (pkg/dem/L6Geom.hpp:36),
   ContactLoop.dist00 (pkg/dem/ContactLoop.hpp:118). */
typedef struct wdm_params {
    double dt;
    double damping;
    double gravity[3];
    double verletDist;  /* < 0: |verletDist| x minimum sphere radius */
    double distFactor;
    double poisson;
    double en;
    double gamma;       /* must be 0 (adhesion is outside the implemented path) */
    double alpha;       /* must be 0 */
    int32_t law;        /* WDM_LAW_* */
    int32_t noSlip, noBreak, noFrict;
    int32_t noAttraction;
    int32_t trackEnergy;
    int32_t trsfRenorm;
    int32_t dist00;
    uint32_t loneMask;
    int32_t precision;  /* WDM_PREC_* */
} wdm_params;

typedef struct wdm_ctx wdm_ctx;

/* Context: owns device buffers and one HIP stream on `device`.  hbm_budget_bytes
   bounds what scenes may allocate (0 = no bound). */
int wdm_create(wdm_ctx** out, int device, size_t hbm_budget_bytes);
void wdm_destroy(wdm_ctx* ctx);
const char* wdm_last_error(wdm_ctx* ctx);

/* A scene slot: Scene + DemField (core/Scene.hpp:27, pkg/dem/Particle.hpp:278).
   Uninodal particles (Sphere, Wall) must use node index == particle index; Facet
   particles reference three nodes with index >= n_par. */
int wdm_scene_alloc(wdm_ctx* ctx, int32_t n_par, int32_t n_node, int32_t n_mat, int32_t* scene_id);
int wdm_scene_free(wdm_ctx* ctx, int32_t scene);
/* bytes of HBM a scene of this size needs with `contacts_per_particle` potential
   contacts (used by admission control) */
size_t wdm_scene_footprint(int32_t n_par, int32_t n_node, double contacts_per_particle);

/* Node + DemData arrays (core/Field.hpp:93-99, pkg/dem/Particle.hpp:249-262); SoA,
   caller-owned host arrays copied synchronously.  force3/torque3 may be NULL (zero). */
int wdm_upload_nodes(wdm_ctx* ctx, int32_t scene, int32_t n, const double* pos3, const double* ori4,
                     const double* vel3, const double* angVel3, const double* mass, const double* inertia3,
                     const uint32_t* flags, const double* force3, const double* torque3);
/* Particle + Shape (pkg/dem/Particle.hpp:26-130): shape code, nodes (3 per particle, unused = -1),
   Sphere.radius or Facet.halfThick, Wall.axis/sense, Particle.mask, material index */
int wdm_upload_particles(wdm_ctx* ctx, int32_t scene, int32_t n, const int32_t* shape, const int32_t* nodes3,
                         const double* radius, const int32_t* axis, const int32_t* sense,
                         const uint32_t* mask, const int32_t* mat);
/* FrictMat table (pkg/dem/FrictMat.hpp:18-28) */
int wdm_upload_materials(wdm_ctx* ctx, int32_t scene, int32_t n, const double* density, const double* young,
                         const double* ktDivKn, const double* tanPhi);
/* Clumps: ClumpData (pkg/dem/Clump.hpp:10-40) as built by ClumpData::makeClump
   (pkg/dem/Clump.cpp:185-253).  Clump c has master node master[c] (flag
   WDM_CLUMP_CLUMP) and members members[off[c]..off[c+1]) (flag WDM_CLUMP_CLUMPED),
   each with relPos (3 doubles) and relOri (w, x, y, z) in the master's frame.
   Leapfrog integrates masters (forces and torques summed from the members,
   ClumpData::forceTorqueFromMembers) and places the members (applyToMembers,
   pkg/dem/Clump.cpp:265-297); the collider does not pair particles of one clump.
   Masters with free rotational DOFs need spherical inertia.  Call after
   wdm_upload_nodes; nc = 0 removes the clumps. */
int wdm_upload_clumps(wdm_ctx* ctx, int32_t scene, int32_t nc, const int32_t* master, const int32_t* off,
                      const int32_t* members, const double* relPos3, const double* relOri4);
/* Cell (core/Cell.hpp:116-160): hSize, gradV, nextGradV row-major 3x3 */
int wdm_set_cell(wdm_ctx* ctx, int32_t scene, int32_t periodic, const double* hSize9, const double* gradV9,
                 const double* nextGradV9, int32_t homoDeform);
int wdm_set_params(wdm_ctx* ctx, int32_t scene, const wdm_params* params);

/* Engines.  One call = one Engine::run() of the reference:
   Leapfrog::run (pkg/dem/Leapfrog.cpp:103-255),
   InsertionSortCollider::run (pkg/dem/InsertionSortCollider.cpp:445-646),
   ContactLoop::run (pkg/dem/ContactLoop.cpp:93-259).
   wdm_step_begin/end are the prologue/epilogue of Scene::doOneStep
   (core/Scene.cpp:371-389); wdm_step runs whole steps for one or more scenes. */
int wdm_step_begin(wdm_ctx* ctx, int32_t scene);
int wdm_leapfrog_run(wdm_ctx* ctx, int32_t scene);
int wdm_collider_run(wdm_ctx* ctx, int32_t scene);
int wdm_contact_loop_run(wdm_ctx* ctx, int32_t scene);
int wdm_step_end(wdm_ctx* ctx, int32_t scene);
/* A scene whose kernels raise an error (NaN force, oversized particle) stops, as a Woo scene's
   run thread stops on an exception and reports it at Scene.wait() (core/Scene.cpp:90-95): in a
   wdm_step over several scenes the others take all their steps, and the call then returns the
   stopped scene's code.  The stopped scene refuses further steps (WDM_E_STATE) and can still
   be read back; free it with wdm_scene_free.  *code = its WDM_E_* code (0: running);
   wdm_last_error() then holds the message. */
int wdm_scene_error(wdm_ctx* ctx, int32_t scene, int32_t* code);
/* DynDt::run (pkg/dem/DynDt.cpp:72-97) when PeriodicEngine::isActivated (core/Engine.cpp:169-195)
   says so at the current step: the engine call after ContactLoop (see wdm_set_dyndt) */
int wdm_dyndt_run(wdm_ctx* ctx, int32_t scene);
int wdm_step(wdm_ctx* ctx, const int32_t* scenes, int32_t n_scenes, int32_t n_steps);

/* DynDt (pkg/dem/DynDt.hpp:20-26, DynDt.cpp:72-97) placed after ContactLoop in the engine
   list, with the PeriodicEngine attributes and state it runs by (core/Engine.hpp:141-153,
   PeriodicEngine::isActivated core/Engine.cpp:169-195).  The state fields (nDone,
   stepLast, stepPrev, virtLast, virtPrev) and `dt` (the value a dryRun would apply) are
   both set and read back, so a host-side engine object keeps them between calls as the
   reference's attributes do.  realPeriod (wall-clock time) is not supported: the steps it
   would fire at are not known when a window of steps is launched.  A new dt takes effect
   at the step's epilogue (core/Scene.cpp:400).  Rejected (WDM_E_UNSUPPORTED): scenes whose
   facet nodes have free DOFs and are not clumped (their stiffness comes from the facets). */
typedef struct wdm_dyndt {
    double maxRelInc;   /* DynDt.maxRelInc (1e-4); 1+maxRelInc must differ from 1 */
    double dtSafety;    /* Scene.dtSafety (.9) */
    double virtPeriod;  /* PeriodicEngine.virtPeriod (<= 0: off) */
    double virtLast, virtPrev;
    double dt;          /* DynDt.dt: the dryRun result (NaN otherwise) */
    int64_t stepPeriod; /* PeriodicEngine.stepPeriod (<= 0: off) */
    int64_t nDo, nDone, stepLast, stepPrev;
    int32_t enabled;    /* the engine is in the list and not dead */
    int32_t stepModulo, initRun, dryRun;
} wdm_dyndt;
int wdm_set_dyndt(wdm_ctx* ctx, int32_t scene, const wdm_dyndt* dyndt);
int wdm_get_dyndt(wdm_ctx* ctx, int32_t scene, wdm_dyndt* dyndt);
/* DynDt::critDt_compute (DynDt.cpp:61-70) on the contacts of the last ContactLoop:
   sqrt(min over nodes of the squared critical step); +inf when no node has stiffness */
int wdm_dyndt_crit(wdm_ctx* ctx, int32_t scene, double* crit_dt);
/* Scene.time and Scene.dt (DynDt may have changed dt) */
int wdm_get_time(wdm_ctx* ctx, int32_t scene, double* time, double* dt);

/* Launch-shape knobs of wdm_step (no counterpart in the reference; results do not depend
   on them): `window` steps between two looks of the host at the Verlet triggers, launched
   as one captured HIP graph when `graphs` is nonzero (1 = step by step); `cell_factor`
   (>= 1) scales the collider's grid cell width over the largest sphere bound.  The getter
   also returns how many windows ran, graphs were captured and graph launches were made. */
int wdm_set_tuning(wdm_ctx* ctx, int32_t window, int32_t graphs, double cell_factor);
int wdm_get_tuning(wdm_ctx* ctx, int32_t* window, int32_t* graphs, double* cell_factor, int64_t* windows,
                   int64_t* graph_captures, int64_t* graph_launches);
/* Broadphase of the collider's full runs: 0 (default) the multi-level uniform grid, 1 sort and
   sweep (bound entries radix-sorted by minimum x, swept in LDS tiles, the reference's
   per-axis principle applied at once rather than incrementally).  Both feed the same exact
   test (InsertionSortCollider.cpp:52-98, 903-957 and Collider::mayCollide), so the contact
   set does not depend on the choice; the sweep is kept as the measured alternative to the
   grid.  No counterpart in the reference. */
int wdm_set_broadphase(wdm_ctx* ctx, int32_t mode);
int wdm_get_broadphase(wdm_ctx* ctx, int32_t* mode);
/* Work-list relayout between collider full runs: every `steps` steps (default 32; 0 never)
   after a scene's last full run or relayout, in the ContactLoop work list of a scene with at
   least `min_contacts` contacts (default 2^20) the contacts that became real since the full
   run are moved right after the state rows of that run, with state rows of their own in the
   same contiguous block (otherwise they sit scattered among the pairs that stay apart).  Memory order only: the contact list, every
   value and every result stay the same (energy sums, which follow the work-list order, as
   well: every way of stepping relayouts at the same steps).  The getter also returns the
   relayouts of scene sid so far (sid < 0: none).  No counterpart in the reference. */
int wdm_set_relayout(wdm_ctx* ctx, int32_t steps, int64_t min_contacts);
int wdm_get_relayout(wdm_ctx* ctx, int32_t sid, int32_t* steps, int64_t* min_contacts, int64_t* relayouts);

/* Readback (synchronous). */
int wdm_download_nodes(wdm_ctx* ctx, int32_t scene, double* pos3, double* ori4, double* vel3, double* angVel3,
                       double* force3, double* torque3);
/* number of contacts currently held (real + potential) */
int wdm_num_contacts(wdm_ctx* ctx, int32_t scene, int64_t* n);
/* the ContactLoop's work at the last full run: contacts it processes (`active`: all but the
   dormant potential contacts, whose bounds overlap but whose surfaces stay more than
   2 verletDist apart until the next full run) and state rows handed out (`rows`); no
   counterpart in the reference, where every contact is visited */
int wdm_contact_work(wdm_ctx* ctx, int32_t scene, int64_t* active, int64_t* rows);
/* contacts sorted by (min id, max id); ids in functor order (Wall/Facet first, else lower id);
   any output pointer may be NULL */
int wdm_get_contacts(wdm_ctx* ctx, int32_t scene, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3,
                     uint8_t* isReal, double* forceLocal3, double* contPt3, double* trsf9, double* uN, int64_t* n_out);
/* phys (FrictPhys: kn, kt, tanPhi, 0, 0; HertzPhys: K, R, tanPhi, kt0, alpha_sqrtMK) and
   L6Geom.lens of the contacts in wdm_get_contacts order (NaN for non-real contacts) */
int wdm_get_contact_state(wdm_ctx* ctx, int32_t scene, int64_t cap, double* phys5, double* lens2, int64_t* n_out);

/* Checkpoint restore (no counterpart on the reference's hot path: Woo saves whole scenes
   with boost::serialization, lib/object/ObjectIO.hpp:49-90).  wdm_upload_contacts makes
   `n` real contacts (ids in functor order, any order of contacts) with their L6Geom and
   phys state the scene's contact list; the next step runs a collider full run, which keeps
   them and adds the potential contacts, so stepping on continues bit for bit as if the
   scene had not been saved (node forces/torques of the checkpoint go to wdm_upload_nodes).
   wdm_set_energy restores EnergyTracker values and created slots, wdm_set_time Scene.step
   and Scene.time. */
int wdm_upload_contacts(wdm_ctx* ctx, int32_t scene, int64_t n, const int32_t* idA, const int32_t* idB,
                        const int32_t* cellDist3, const double* trsf9, const double* forceLocal3,
                        const double* phys5, const double* lens2, const double* contPt3);
int wdm_set_energy(wdm_ctx* ctx, int32_t scene, const double* values, const int32_t* created);
int wdm_set_time(wdm_ctx* ctx, int32_t scene, int64_t step, double time);

/* Particles removed and inserted in place.  The scene keeps its particle slots: a removed
   particle (DemField::removeParticle, pkg/dem/Particle.cpp:445-493; spheres and walls, not
   clumped) leaves a hole of shape NONE whose node is fully blocked, massless and at rest, and
   its contacts are dropped; the forces of the last ContactLoop stay with its partners.
   Inserted spheres take ids as ParticleContainer::insert does (ParticleContainer.cpp:23-72):
   the holes first, in the order they were freed, then the next id past the last particle in
   use, up to n_par (allocate spare particles of shape NONE for growth); their nodes start
   with zero force and torque, as a new DemData.  Both make the next step run a collider full
   run (the particle set changed, InsertionSortCollider.cpp:475-492). */
int wdm_remove_particles(wdm_ctx* ctx, int32_t scene, int32_t n, const int32_t* ids);
int wdm_insert_spheres(wdm_ctx* ctx, int32_t scene, int32_t n, const double* pos3, const double* vel3,
                       const double* angVel3, const double* radius, const double* mass, const double* inertia,
                       const int32_t* mat, const uint32_t* mask, int32_t* ids_out);
/* holes in use order and the end of the particles in use (ParticleContainer.freeIds, size) */
int wdm_get_free_ids(wdm_ctx* ctx, int32_t scene, int32_t cap, int32_t* ids, int32_t* n, int32_t* par_end);
int wdm_set_free_ids(wdm_ctx* ctx, int32_t scene, int32_t n, const int32_t* ids, int32_t par_end);
/* Particle.mask of particles (Outlet's marking mode, pkg/dem/Outlet.cpp:95-101) */
int wdm_set_masks(wdm_ctx* ctx, int32_t scene, int32_t n, const int32_t* ids, const uint32_t* masks);
/* The particles Outlet::run (pkg/dem/Outlet.cpp:12-40) would delete (markMask 0) or mark for a
   BoxOutlet (Outlet.hpp:56: the closed box lo..hi; `inside` selects the particles in it
   rather than out of it; `mask` as Outlet.mask), uninodal particles only, ascending ids;
   n = how many (ids holds the first `cap`). */
int wdm_select_box(wdm_ctx* ctx, int32_t scene, const double* lo3, const double* hi3, int32_t inside, uint32_t mask,
                   uint32_t mark_mask, int32_t cap, int32_t* ids, int32_t* n);
/* wdm_select_box that allocates nothing (no counterpart in the reference): for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops).  Same
   arguments and result; the selection is compacted on the device in memory the library
   reserves once per GPU (16.1 MiB, held until the process ends), so scenes of at most
   4,194,304 particles are accepted (larger ones: WDM_E_ARG).  Calls that overlap on one GPU
   share that memory and race. */
int wdm_select_box_noalloc(wdm_ctx* ctx, int32_t scene, const double* lo3, const double* hi3, int32_t inside, uint32_t mask,
                           uint32_t mark_mask, int32_t cap, int32_t* ids, int32_t* n);

/* SpherePack::makeCloud (lib/sphere-pack/SpherePack.cpp:135-248) on the context's device:
   a random loose packing in the box mn..mx (periodic: the cell of that size), radii from
   rMean/rRelFuzz, num/porosity or a PSD (psdSizes diameters, psdCumm cumulative, nPsd
   points; distributeMass as the reference), placed sphere after sphere in the reference's
   order.  *rng_state is boost::minstd_rand's state (1..2^31-2), advanced by the draws (the
   reference's generator is a static, seeded once per process).  Spheres go to c3/r (cap of
   them; with num < 0 the packing stops at cap spheres or at the first sphere without a free
   spot); sheared cells (hSize) are not supported. */
int wdm_make_cloud(wdm_ctx* ctx, const double* mn3, const double* mx3, double rMean, double rRelFuzz, int32_t num,
                   int32_t periodic, double porosity, const double* psdSizes, const double* psdCumm, int32_t nPsd,
                   int32_t distributeMass, uint32_t* rng_state, int64_t cap, double* c3, double* r, int64_t* n_out,
                   double* appliedPsdScaling);

/* pack.regularOrtho / pack.regularHexa (py/pack.py:323-350) on the context's device: the
   lattice over the predicate's aabb (ortho: numpy.arange(mn+radius, mx-radius, 2 radius+gap)
   per axis; hexa: spacing a = 2 radius + gap, rows a sqrt(3)/2, layers a sqrt(6)/3, alternate
   rows and layers shifted), the points where the predicate holds with pad = radius, in the
   order of itertools.product (z fastest).  kind 0 inAlignedBox (params mn3, mx3), 1 inSphere
   (centre3, r), 2 inCylinder (c1_3, c2_3, r) (py/_packPredicates.cpp:96-205).  c3 receives
   n_out centres (cap of them; a larger lattice sets n_out and returns WDM_E_ARG); with c3
   NULL only n_out is set. */
int wdm_regular_pack(wdm_ctx* ctx, int32_t hexa, int32_t kind, const double* params, double radius, double gap, int64_t cap,
                     double* c3, int64_t* n_out);

/* EnergyTracker values and whether each slot exists */
int wdm_get_energy(wdm_ctx* ctx, int32_t scene, double* values, int32_t* created);
/* Scene.step, InsertionSortCollider.nFullRuns, resolved verletDist */
int wdm_get_stats(wdm_ctx* ctx, int32_t scene, int64_t* step, int64_t* n_full_runs, double* verlet_dist);
int wdm_get_cell(wdm_ctx* ctx, int32_t scene, double* hSize9, double* gradV9);
int wdm_sync(wdm_ctx* ctx);

/* Per-kernel timing of the last wdm_step call, measured with HIP events on the
   context's stream: fills up to cap (name, total ms, launches) entries. */
int wdm_kernel_times(wdm_ctx* ctx, int32_t cap, char* names_64, double* ms, int64_t* launches, int32_t* n_out);
int wdm_enable_timing(wdm_ctx* ctx, int32_t on);

/* Scene generation helper (no counterpart in the timestep path): pushes overlapping
   spheres apart with `iters` Jacobi relaxation sweeps on the device, deterministic. */
int wdm_relax_overlaps(wdm_ctx* ctx, int32_t scene, int32_t iters);

#ifdef __cplusplus
}
#endif
