/* CPU FP64 oracle of Woo's DEM step (test infrastructure only).
 *
 * This library is a plain-C++ restatement of the reference algorithm for the
 * DEM step path. It is the checker for the HIP engine and the CPU baseline
 * timed beside it; it is never linked into, loaded by, or called from the
 * product library. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Reference files restated (paths relative to the woodem/woo tree):
 *   src/core/Scene.cpp:366-467            step driver (engine order, dt/nextDt, step++)
 *   src/dem/Collision.cpp:21-184          mayCollide, aabbIsDirty, updateAabb, setVerletDist
 *   src/dem/InsertionSortCollider.cpp:35-613 + .hpp:86-218  SAP with Verlet-enlarged AABBs
 *   src/dem/Sphere.cpp:26-117             mass/inertia, Bo1_Sphere_Aabb, Cg2_Sphere_Sphere_L6Geom
 *   src/dem/Wall.cpp:25-84                Bo1_Wall_Aabb, Cg2_Wall_Sphere_L6Geom
 *   src/dem/Facet.cpp:74-118,161-177,275-298,333-390  Facet normal/outer vectors, velocity
 *                                         interpolation, Bo1_Facet_Aabb, Cg2_Facet_Sphere_L6Geom
 *   src/dem/Facet.cpp:301-328             In2_Facet (IntraForce): contact forces spread to facet nodes
 *   src/dem/Facet.cpp:180-265             computeNeighborAngles, adjustBoundaryContactGeom (n21lim);
 *                                         the vertex branch (interpolation between two edges,
 *                                         Facet.cpp:228-236): parity unpinned (no closed form;
 *                                         the edge branch and the neighbour angles are pinned)
 *   src/supp/base/CompUtils.cpp:59-65     closestSegmentPt
 *   src/supp/base/CompUtils.cpp:94-112    triangleBarycentrics
 *   src/dem/L6Geom.cpp:25-216             handleSpheresLikeContact
 *   src/dem/FrictMat.cpp:20-44            Cp2_FrictMat_FrictPhys
 *   src/dem/IdealElPl.cpp:15-125          Law2_L6Geom_FrictPhys_IdealElPl (with rolling/twisting resistance)
 *   src/dem/Hertz.cpp:34-210              Cp2_HertzMat_HertzPhys, Law2_L6Geom_HertzPhys_DMT (DMT branch)
 *   src/dem/ContactLoop.cpp:93-276        ContactLoop::run, applyForceUninodal
 *   src/dem/Contact.cpp:24-34             getForceTorqueBranch
 *   src/dem/ContactContainer.cpp:18-200   add/remove/requestRemoval/removePending
 *   src/dem/Leapfrog.cpp:15-297           ForceResetter, Leapfrog (spherical integrator)
 *   src/dem/DynDt.cpp:16-113              DynDt
 *   src/dem/Funcs.cpp:84-114,151-175      coordNumber, porosity (solid fraction), pWaveDt
 *   src/core/Cell.hpp:97                  intrShiftPos
 */
#ifndef WOO_ORACLE_H
#define WOO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

enum { ORC_SPHERE = 0, ORC_WALL = 1, ORC_FACET = 2 };
enum { ORC_FRICTMAT = 0, ORC_HERTZMAT = 1 };
enum { ORC_LAW_IDEALELPL = 0, ORC_LAW_HERTZDMT = 1 };
enum { ORC_PHYS_FRICT = 0, ORC_PHYS_HERTZ = 1 };
/* engine codes for orc_run */
enum {
  ORC_E_FORCERESETTER = 1, ORC_E_COLLIDER = 2, ORC_E_CONTACTLOOP = 3, ORC_E_LEAPFROG = 4, ORC_E_DYNDT = 5,
  ORC_E_INTRAFORCE = 6  /* IntraForce([In2_Facet]) */
};
/* node fields for orc_get / orc_set */
enum { ORC_F_POS = 0, ORC_F_VEL = 1, ORC_F_ANGVEL = 2, ORC_F_ORI = 3, ORC_F_FORCE = 4, ORC_F_TORQUE = 5,
       ORC_F_MASS = 6, ORC_F_INERTIA = 7 };

typedef struct {
  /* Scene */
  double dt;
  double dtSafety;
  int trsfRenorm;            /* Cg2_Any_Any_L6Geom__Base.trsfRenorm (100) */
  int deterministic;         /* 1: canonical per-particle summation order; 0: OpenMP + per-node mutex */
  int nThreads;              /* OpenMP threads, 0 = runtime default */
  /* DemField */
  double gravity[3];
  double distFactor;         /* -1: off (== 1) */
  uint32_t loneMask;         /* 2 */
  int periodic;
  double cell[3];            /* diagonal hSize */
  /* InsertionSortCollider */
  double verletDist;         /* negative: relative to min radius */
  int colliderMode;          /* 0: incremental 3-axis insertion sort (aperiodic), 1: full rebuild at refresh */
  /* ContactLoop */
  int physKind;              /* ORC_PHYS_* */
  int lawKind;               /* ORC_LAW_* */
  int applyForces;
  double poisson, en;        /* Cp2_HertzMat_HertzPhys */
  int noSlip, noBreak, noFrict; /* Law2_L6Geom_FrictPhys_IdealElPl */
  int noAttraction;          /* Law2_L6Geom_HertzPhys_DMT */
  /* Leapfrog */
  double damping;
  int reset;
  /* DynDt */
  int dynDtPeriod;
  double maxRelInc;
  /* Scene.trackEnergy */
  int trackEnergy;
  /* ContactLoop.updatePhys (ContactLoop.hpp:119): 0 never, 1 always, 2 once (then never) */
  int updatePhys;
  /* Law2_L6Geom_FrictPhys_IdealElPl rolling and twisting resistance (IdealElPl.hpp:18-20);
   * off unless relRollStiff > 0 and rollTanPhi > 0 */
  double relRollStiff, relTwistStiff, rollTanPhi;
  /* ContactLoop.evalStress (ContactLoop.hpp:117): the contact loop's stress in periodic cells */
  int evalStress;
} orc_params;

/* Scene.energy (EnergyTracker, src/core/EnergyTracker.hpp:30-61): the terms the path's engines
 * add, in this order. kinetic and elast hold the last step's value (IsResettable); the others
 * accumulate (IsIncrement). */
enum { ORC_E_KINETIC = 0, ORC_E_ELAST = 1, ORC_E_GRAV = 2, ORC_E_NONVISCDAMP = 3, ORC_E_PLAST = 4, ORC_E_VISCN = 5,
       ORC_E_VISCT = 6, ORC_E_N = 7 };
void orc_energy(const orc_scene* s, double out[7]);

void orc_default_params(orc_params* p);
orc_scene* orc_create(const orc_params* p);
void orc_destroy(orc_scene* s);
int orc_set_params(orc_scene* s, const orc_params* p);

int orc_add_material(orc_scene* s, int kind, double density, double young, double ktDivKn, double tanPhi,
                     double surfEnergy, double alpha);
/* material m edited in place (the Material object's attributes assigned from Python) */
int orc_set_material(orc_scene* s, int m, int kind, double density, double young, double ktDivKn, double tanPhi,
                     double surfEnergy, double alpha);
/* the scene's parameters as they stand (ContactLoop.updatePhys 'once' turns into 'never') */
int orc_get_params(const orc_scene* s, orc_params* out);
/* returns particle id */
int orc_add_sphere(orc_scene* s, const double pos[3], double radius, int mat, int fixed, uint32_t mask);
int orc_add_wall(orc_scene* s, const double pos[3], int axis, int sense, int mat, uint32_t mask);
/* fixed facet (3 static vertices, verts 9 doubles); fakeVel may be NULL (zero); the
 * particle's node position is vertex 0 */
int orc_add_facet(orc_scene* s, const double verts[9], const double fakeVel[3], double halfThick, int mat,
                  uint32_t mask);
/* Facet.n21lim of facet i (edge-neighbour normal dot products, NaN: no smoothing) */
int orc_set_facet_n21lim(orc_scene* s, int64_t i, const double n21lim[3]);
int orc_facet_n21lim(const orc_scene* s, int64_t i, double out[3]);
/* Facet::computeNeighborAngles for every facet (a shared node = a vertex at the same place) */
int orc_facet_neighbor_angles(orc_scene* s);
/* bulk: n free spheres (pos n*3, radius n), one material and mask; returns the first id */
int orc_add_spheres(orc_scene* s, int64_t n, const double* pos, const double* radius, int mat, uint32_t mask);
/* DemData.flags of particle i (blocked DOFs and the *Skip bits, Particle.hpp:196-199) */
int orc_set_flags(orc_scene* s, int64_t i, uint32_t flags);

int orc_num_particles(const orc_scene* s);
/* bulk node data (one node per particle), n*3 or n*4 or n doubles */
int orc_get(const orc_scene* s, int field, double* out);
int orc_set(orc_scene* s, int field, const double* in);

/* run nSteps steps with the given engine list (Scene.engines order) */
int orc_run(orc_scene* s, const int* engines, int nEngines, int nSteps);
/* facet i's node forces and torques from the last IntraForce: F of nodes 0, 1, 2 (9), then
 * T of nodes 0, 1, 2 (9) */
int orc_facet_nodes(const orc_scene* s, int64_t i, double out[18]);
int64_t orc_step(const orc_scene* s);
double orc_dt(const orc_scene* s);
double orc_verlet_dist(const orc_scene* s);
int orc_num_refresh(const orc_scene* s);

/* contacts: which=0 all (potential+real), which=1 real only.
 * ids: n*2 (pA,pB) in canonical (min id, max id) key order; cellDist n*3;
 * data (may be NULL): per contact 16 doubles:
 *   trsf col0 (3), trsf col1 (3), contact point (3), local force (3), uN, kn, kt, real(1/0) */
int64_t orc_num_contacts(const orc_scene* s, int which);
int64_t orc_get_contacts(const orc_scene* s, int which, int32_t* ids, int32_t* cellDist, double* data);

/* observables */
double orc_coord_number(const orc_scene* s, int skipFree);
double orc_solid_fraction(const orc_scene* s, const double boxMin[3], const double boxMax[3]);
double orc_kinetic_energy(const orc_scene* s);
int orc_stress(const orc_scene* s, double volume, double out[9]);
/* ContactLoop.stress (ContactLoop.cpp:117,216-229,243-244): sum over the real contacts of the
 * last contact loop of F branch^T (F = node.ori * force, branch = Contact::dPos) over the cell
 * volume, row-major; zero unless evalStress in a periodic cell */
int orc_cl_stress(const orc_scene* s, double out[9]);
double orc_pwave_dt(const orc_scene* s);
/* DynDt::critDt_stiffness on the current contact state (no side effects) */
double orc_crit_dt(const orc_scene* s);
/* brute-force enlarged-AABB overlap set at the last refresh (pairs of ids, canonical order) */
int64_t orc_brute_aabb_pairs(const orc_scene* s, int32_t* ids, int64_t cap);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
