/*
 * CPU oracle for the Woo DEM timestep -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain C++17 restatement of the reference's CPU algorithm for the
 * timestep path (Leapfrog, InsertionSortCollider, ContactLoop with
 * Cg2_{Sphere_Sphere,Wall_Sphere,Facet_Sphere}_L6Geom, Cp2_FrictMat_FrictPhys +
 * Law2_L6Geom_FrictPhys_IdealElPl, Cp2_FrictMat_HertzPhys +
 * Law2_L6Geom_HertzPhys_DMT, Cell GRADV2, EnergyTracker).  It is used only by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the
 * checker.  The product (woodem_amd/, libwoodem_hip.so) never links or calls it.
 *
 * Every function in dem_oracle.cpp cites the reference file:line it follows.
 */
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* shape codes (same numbering as the product ABI) */
enum { ORC_SHAPE_NONE = 0, ORC_SHAPE_SPHERE = 1, ORC_SHAPE_WALL = 2, ORC_SHAPE_FACET = 3 };
/* law codes */
enum { ORC_LAW_IDEALELPL = 0, ORC_LAW_HERTZ_DMT = 1 };

/* Parameters of the engines; layout: all doubles first, then all ints. */
typedef struct orc_params {
    double dt;            /* Scene.dt */
    double damping;       /* Leapfrog.damping */
    double gravity[3];    /* DemField.gravity */
    double verletDist;    /* InsertionSortCollider.verletDist (<0: relative to min sphere radius) */
    double distFactor;    /* Bo1_Sphere_Aabb.distFactor == Cg2_Sphere_Sphere_L6Geom.distFactor */
    double poisson;       /* Cp2_FrictMat_HertzPhys.poisson */
    double en;            /* Cp2_FrictMat_HertzPhys.en */
    double gamma;         /* Cp2_FrictMat_HertzPhys.gamma (must be 0) */
    double alpha;         /* Cp2_FrictMat_HertzPhys.alpha (must be 0) */
    int32_t law;          /* ORC_LAW_* */
    int32_t noSlip, noBreak, noFrict;  /* Law2_L6Geom_FrictPhys_IdealElPl flags */
    int32_t noAttraction; /* Law2_L6Geom_HertzPhys_DMT.noAttraction */
    int32_t trackEnergy;  /* Scene.trackEnergy */
    int32_t trsfRenorm;   /* Cg2_Any_Any_L6Geom__Base.trsfRenorm */
    int32_t dist00;       /* ContactLoop.dist00 */
    uint32_t loneMask;    /* DemField.loneMask */
    int32_t nThreads;     /* OpenMP threads for the oracle (0 = default) */
} orc_params;

/* energy slots, fixed order */
enum { ORC_E_KINETIC = 0, ORC_E_ELAST, ORC_E_PLAST, ORC_E_GRAV, ORC_E_NONVISCDAMP,
       ORC_E_VISCN, ORC_E_VISCT, ORC_E_BROKEN, ORC_E_N };

void* orc_create(void);
void  orc_destroy(void* h);
const char* orc_last_error(void* h);

int orc_set_nodes(void* h, int32_t n, const double* pos3, const double* ori4, const double* vel3,
                  const double* angVel3, const double* mass, const double* inertia3,
                  const uint32_t* flags, const double* force3, const double* torque3);
int orc_set_particles(void* h, int32_t n, const int32_t* shape, const int32_t* nodes3,
                      const double* radius, const int32_t* axis, const int32_t* sense,
                      const uint32_t* mask, const int32_t* mat);
int orc_set_materials(void* h, int32_t n, const double* density, const double* young,
                      const double* ktDivKn, const double* tanPhi);
int orc_set_cell(void* h, int32_t periodic, const double* hSize9, const double* gradV9,
                 const double* nextGradV9, int32_t homoDeform);
int orc_set_params(void* h, const orc_params* p);
/* clumps (ClumpData, pkg/dem/Clump.hpp): master node, member node lists (offsets
   nc+1), members' relPos (3 each) and relOri (w, x, y, z); call after orc_set_nodes */
int orc_set_clumps(void* h, int32_t nc, const int32_t* master, const int32_t* off, const int32_t* members,
                   const double* relPos3, const double* relOri4);

/* DynDt (pkg/dem/DynDt.hpp:20-26) after ContactLoop, with the PeriodicEngine attributes and
   state it runs by (core/Engine.hpp:141-153); realPeriod is not restated */
typedef struct orc_dyndt {
    double maxRelInc, dtSafety, virtPeriod, virtLast, virtPrev, dt;
    int64_t stepPeriod, nDo, nDone, stepLast, stepPrev;
    int32_t enabled, stepModulo, initRun, dryRun;
} orc_dyndt;
int orc_set_dyndt(void* h, const orc_dyndt* d);
int orc_get_dyndt(void* h, orc_dyndt* d);
/* DynDt::critDt_compute (pkg/dem/DynDt.cpp:61-70) on the current contacts */
int orc_dyndt_crit(void* h, double* critDt);
/* DemField::removeParticle (pkg/dem/Particle.cpp:445-493) with the slot kept as a hole;
   ParticleContainer::insert of spheres (ParticleContainer.cpp:23-72: freed ids first, FIFO) */
int orc_remove_particles(void* h, int32_t n, const int32_t* ids);
int orc_insert_spheres(void* h, int32_t n, const double* pos3, const double* vel3, const double* angVel3, const double* radius,
                       const double* mass, const double* inertia, const int32_t* mat, const uint32_t* mask, int32_t* idsOut);
int orc_set_masks(void* h, int32_t n, const int32_t* ids, const uint32_t* masks);
int orc_get_free_ids(void* h, int32_t cap, int32_t* ids, int32_t* n, int32_t* parEnd);
/* SpherePack::makeCloud (lib/sphere-pack/SpherePack.cpp:135-248), sequential, no hSize;
   *state: boost::minstd_rand state (1..2^31-2), advanced by the draws */
int orc_make_cloud(const double* mn3, const double* mx3, double rMean, double rRelFuzz, int32_t num, int32_t periodic, double porosity,
                   const double* psdSizes, const double* psdCumm, int32_t nPsd, int32_t distributeMass, uint32_t* state, int64_t cap,
                   double* c3, double* r, int64_t* nOut, double* appliedPsdScaling);
/* pack.regularOrtho / regularHexa (py/pack.py:323-350); kind 0 inAlignedBox, 1 inSphere, 2 inCylinder */
int orc_regular_pack(int32_t hexa, int32_t kind, const double* params, double radius, double gap, int64_t cap, double* c3, int64_t* nOut);
/* Scene.time, Scene.dt */
int orc_get_time(void* h, double* time, double* dt);

/* run n full timesteps (Scene::doOneStep) */
int orc_step(void* h, int32_t n);

int orc_get_nodes(void* h, double* pos3, double* ori4, double* vel3, double* angVel3,
                  double* force3, double* torque3);
int64_t orc_num_contacts(void* h);
/* contacts sorted by (min id, max id); idA/idB in functor order */
int orc_get_contacts(void* h, int64_t cap, int32_t* idA, int32_t* idB, int32_t* cellDist3,
                     uint8_t* isReal, double* forceLocal3, double* contPt3, double* trsf9,
                     double* uN, int64_t* nOut);
/* the real contacts of a checkpoint replace the contact map (phys5: FrictPhys kn, kt, tanPhi,
   -, -; HertzPhys K, R, tanPhi, kt0, alpha_sqrtMK); the next step runs a full collider run */
int orc_set_contacts(void* h, int64_t n, const int32_t* idA, const int32_t* idB, const int32_t* cellDist3,
                     const double* trsf9, const double* forceLocal3, const double* phys5, const double* lens2,
                     const double* contPt3);
/* phys5 and lens2 of the contacts in orc_get_contacts order (NaN where not real) */
int orc_get_contact_phys(void* h, int64_t cap, double* phys5, double* lens2);
/* step counter and time of a checkpoint */
int orc_set_time(void* h, int64_t step, double time);
/* out[ORC_E_N] values, created[ORC_E_N] 0/1 */
int orc_get_energy(void* h, double* out, int32_t* created);
/* stats: step, nFullRuns, verletDist(actual) */
int orc_get_stats(void* h, int64_t* step, int64_t* nFullRuns, double* verletDist);
/* chunks of the last parallel (chunked) insertion sort, -1 if every sort so far was serial,
   and the most passes one sort needed */
int orc_get_sort_stats(void* h, int64_t* sortChunks, int64_t* sortPasses);
/* current cell hSize (9, row-major) */
int orc_get_cell(void* h, double* hSize9, double* gradV9);

/* Brute-force O(N^2) potential-contact predicate (closed-interval 3-axis AABB
   overlap of the collider's stored bounds, plus Collider::mayCollide) on the
   current state; fills sorted (min,max) pairs.  Used to pin the insertion-sort
   restatement. */
int64_t orc_bruteforce_pairs(void* h, int64_t cap, int32_t* idMin, int32_t* idMax);

/* closed form of DemFuncs::pWaveDt over all spheres (pkg/dem/Funcs.cpp:116-142) */
double orc_pwave_dt(void* h);

#ifdef __cplusplus
}
#endif
