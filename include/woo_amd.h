/* woo_amd — MI355X (gfx950) DEM step engine, C ABI.
 *
 * This is the drop-in boundary: each entry point is what one of Woo's
 * Engine::run() implementations (or a ParticleContainer / DemData accessor)
 * calls in place of its OpenMP C++ body. Plain pointers and sizes only; every
 * function returns 0 on success and a negative WDG_E_* code on failure, with
 * a message in wdg_last_error() (thread-local). No exception crosses the ABI.
 *
 * Reference interfaces replaced (woodem/woo tree):
 *   wdg_collide        InsertionSortCollider::run        src/dem/InsertionSortCollider.cpp:386-613
 *   wdg_contact_loop   ContactLoop::run                  src/dem/ContactLoop.cpp:93-268
 *                      (Cg2_Sphere_Sphere_L6Geom        src/dem/Sphere.cpp:92-117,
 *                       Cg2_Wall_Sphere_L6Geom          src/dem/Wall.cpp:50-84,
 *                       Cg2_Facet_Sphere_L6Geom         src/dem/Facet.cpp:333-390,
 *                       Cp2_FrictMat_FrictPhys          src/dem/FrictMat.cpp:20-44,
 *                       Cp2_HertzMat_HertzPhys          src/dem/Hertz.cpp:34-77,
 *                       Law2_L6Geom_FrictPhys_IdealElPl src/dem/IdealElPl.cpp:15-125,
 *                       Law2_L6Geom_HertzPhys_DMT       src/dem/Hertz.cpp:85-210)
 *   wdg_leapfrog       Leapfrog::run                     src/dem/Leapfrog.cpp:104-255
 *   wdg_force_reset    ForceResetter::run                src/dem/Leapfrog.cpp:15-26
 *   wdg_crit_dt        DynDt::critDt_stiffness           src/dem/DynDt.cpp:61-71
 *   wdg_upload_particles / wdg_upload / wdg_download
 *                      ParticleContainer::pyAppend, DemData attributes
 *                                                        src/dem/ParticleContainer.cpp:176-201, src/dem/Particle.hpp:281-292
 *   wdg_contact_count / wdg_contact_list
 *                      ContactContainer size / linView   src/dem/ContactContainer.hpp:16-141
 *   wdg_restore_contacts
 *                      ContactContainer of a loaded Scene src/core/Scene.cpp:285-294
 *   wdg_observables    DemFuncs::coordNumber/porosity    src/dem/Funcs.cpp:84-114
 *   wdg_stress         DemFuncs::stressStiffness (stress) src/dem/Funcs.cpp:183-229
 *   wdg_kinetic_energy DemData kinetic energy             src/dem/Particle.cpp:247-265
 *   wdg_set_energy_tracking / wdg_energy / wdg_energy_clear / wdg_energy_set
 *                      Scene.trackEnergy, Scene.energy (EnergyTracker)
 *                                                        src/core/Scene.hpp:204,223, src/core/EnergyTracker.hpp:30-61
 *   wdg_energy_scenes / wdg_energy_scenes_set
 *                      Scene.energy of each scene of a batch (one Scene per scene in the reference)
 *   wdg_run            Scene::doOneStep (engine loop)    src/core/Scene.cpp:366-467
 */
#ifndef WOO_AMD_H
#define WOO_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WDG_OK 0
#define WDG_E_ARG (-1)      /* invalid argument */
#define WDG_E_HIP (-2)      /* HIP runtime error */
#define WDG_E_STATE (-3)    /* call not valid in the current state */
#define WDG_E_NUMERIC (-4)  /* NaN force or similar, raised by a kernel */
#define WDG_E_GEOM (-5)     /* periodic double contact, too many periods */
#define WDG_E_CAPACITY (-6) /* a size limit was exceeded */

typedef struct wdg_ctx wdg_ctx;

/* particle shapes */
enum { WDG_SPHERE = 0, WDG_WALL = 1, WDG_FACET = 2 };
/* material kinds */
enum { WDG_FRICTMAT = 0, WDG_HERTZMAT = 1 };
/* contact models */
enum { WDG_PHYS_FRICT = 0, WDG_PHYS_HERTZ = 1 };
enum { WDG_LAW_IDEALELPL = 0, WDG_LAW_HERTZDMT = 1 };
/* node fields for wdg_upload / wdg_download */
enum {
  WDG_F_POS = 0, WDG_F_VEL = 1, WDG_F_ANGVEL = 2, WDG_F_ORI = 3, WDG_F_FORCE = 4, WDG_F_TORQUE = 5,
  WDG_F_MASS = 6, WDG_F_INERTIA = 7
};
/* Scene.energy terms, in the order wdg_energy gives them. kinetic and elast are the last
 * step's values (EnergyTracker::IsResettable); the others accumulate (IsIncrement). */
enum {
  WDG_EN_KINETIC = 0, WDG_EN_ELAST = 1, WDG_EN_GRAV = 2, WDG_EN_NONVISCDAMP = 3, WDG_EN_PLAST = 4,
  WDG_EN_VISCN = 5, WDG_EN_VISCT = 6, WDG_EN_N = 7
};
/* engine codes for wdg_run */
enum {
  WDG_E_FORCERESETTER = 1, WDG_E_COLLIDER = 2, WDG_E_CONTACTLOOP = 3, WDG_E_LEAPFROG = 4, WDG_E_DYNDT = 5,
  WDG_E_INTRAFORCE = 6  /* IntraForce([In2_Facet]), after the ContactLoop */
};

/* DemField + Scene.cell (diagonal hSize only; gradV is zero) */
typedef struct {
  double gravity[3];
  double distFactor;   /* DemField.distFactor; negative means 1 */
  uint32_t loneMask;   /* DemField.loneMask */
  int32_t periodic;    /* Scene.periodic */
  double cell[3];      /* Scene.cell.hSize diagonal */
} wdg_field;

/* FrictMat / HertzMat */
typedef struct {
  int32_t kind;
  double density, young, ktDivKn, tanPhi, surfEnergy, alpha;
} wdg_material;

/* bulk particle description, one node per particle; arrays of length n (x3 / x4 where noted) */
typedef struct {
  int64_t n;
  const int32_t* shape;       /* WDG_SPHERE | WDG_WALL | WDG_FACET */
  const double* radius;       /* Sphere.radius (ignored for walls) */
  const int32_t* wallAxis;    /* Wall.axis (ignored for spheres); may be NULL if no walls */
  const int32_t* wallSense;   /* Wall.sense */
  const int32_t* material;    /* index into wdg_set_materials */
  const uint32_t* mask;       /* Particle.mask */
  const uint32_t* flags;      /* DemData.flags: blocked DOFs bits 0-5, gravitySkip 512, dampingSkip 2048 */
  const uint8_t* integrated;  /* 1 if the node is in DemField.nodes */
  const double* pos;          /* x3 */
  const double* vel;          /* x3, may be NULL (zero) */
  const double* angVel;       /* x3, may be NULL */
  const double* ori;          /* x4 (w,x,y,z), may be NULL (identity) */
  const double* force;        /* x3, may be NULL */
  const double* torque;       /* x3, may be NULL */
  const double* mass;         /* x1 */
  const double* inertia;      /* x3 */
  /* Facet (Facet.hpp): fixed triangles with static vertices; a facet's node position is its
   * vertex 0 (pos is ignored for facets), its velocities must be zero, and it takes no force
   * from the contact loop (multinodal, ContactLoop.cpp:270-276). Aperiodic scenes only. */
  const double* facetVerts;   /* x9 (vertices 0,1,2), may be NULL if there are no facets */
  const double* fakeVel;      /* x3, Facet.fakeVel, may be NULL (zero) */
  const double* halfThick;    /* x1, Facet.halfThick, may be NULL (zero) */
  /* Batched independent scenes: the scene of each particle (0..nScenes-1, see
   * wdg_set_scenes); particles of different scenes never collide. May be NULL (all 0). */
  const int32_t* scene;
  /* x3, Facet.n21lim (src/dem/Facet.hpp:45; Facet::computeNeighborAngles, Facet.cpp:180-216):
   * per edge the dot product of the facet's normal with its edge neighbour's, NaN for none; an
   * edge or vertex contact's normal is then held within the neighbour's
   * (Facet::adjustBoundaryContactGeom, Facet.cpp:220-265). May be NULL (all NaN). */
  const double* facetN21lim;
} wdg_particles;

typedef struct {
  double verletDist;   /* InsertionSortCollider.verletDist: <0 relative to the smallest sphere radius */
  int32_t forceInit;   /* InsertionSortCollider.forceInitSort */
} wdg_collide_params;

typedef struct {
  int64_t step;        /* Scene.step (trsfRenorm uses it) */
  double dt;           /* Scene.dt */
  int32_t physKind;    /* WDG_PHYS_* */
  int32_t lawKind;     /* WDG_LAW_* */
  int32_t applyForces; /* ContactLoop.applyForces */
  int32_t trsfRenorm;  /* Cg2_Any_Any_L6Geom__Base.trsfRenorm */
  double poisson, en;  /* Cp2_HertzMat_HertzPhys.poisson, .en */
  int32_t noSlip, noBreak, noFrict; /* Law2_L6Geom_FrictPhys_IdealElPl */
  int32_t noAttraction;             /* Law2_L6Geom_HertzPhys_DMT */
  /* Law2_L6Geom_FrictPhys_IdealElPl rolling and twisting resistance (IdealElPl.hpp:18-20,
   * IdealElPl.cpp:93-123): on when relRollStiff > 0 and rollTanPhi > 0 (FP64 contact form
   * only); 0 = off */
  double relRollStiff, relTwistStiff, rollTanPhi;
} wdg_contact_params;

typedef struct {
  double dt;           /* Scene.dt */
  double damping;      /* Leapfrog.damping */
  int32_t reset;       /* Leapfrog.reset */
} wdg_leapfrog_params;

/* counters, read back by wdg_stats */
typedef struct {
  int64_t nParticles;
  int64_t nCandidates;   /* contact slots (potential pairs after the last refresh, plus carried real contacts) */
  int64_t nReal;         /* real contacts after the last contact loop (computed on request) */
  int64_t nRefresh;      /* broadphase rebuilds so far */
  int64_t nSteps;        /* collider calls so far */
  double verletDist;     /* resolved absolute Verlet distance */
  int32_t brute;         /* last refresh used the all-pairs path */
  int32_t gridCells[3];  /* cell grid of the last refresh */
  double refreshSeconds; /* host time spent in broadphase refreshes so far */
} wdg_counters;

/* one Scene::doOneStep engine list, for wdg_run */
typedef struct {
  int32_t nEngines;
  int32_t engines[8];  /* WDG_E_* codes in Scene.engines order */
  wdg_collide_params collide;
  wdg_contact_params contact;
  wdg_leapfrog_params leapfrog;
  int32_t dynDtPeriod; /* DynDt.stepPeriod */
  double maxRelInc;    /* DynDt.maxRelInc */
  double dtSafety;     /* Scene.dtSafety */
} wdg_program;

const char* wdg_last_error(void);
int wdg_ctx_create(int device, wdg_ctx** out);
int wdg_ctx_destroy(wdg_ctx* ctx);
int wdg_set_field(wdg_ctx* ctx, const wdg_field* f);
int wdg_set_materials(wdg_ctx* ctx, const wdg_material* mats, int32_t n);
int wdg_upload_particles(wdg_ctx* ctx, const wdg_particles* p);
int wdg_upload(wdg_ctx* ctx, int32_t field, const double* host);
int wdg_download(wdg_ctx* ctx, int32_t field, double* host);

int wdg_collide(wdg_ctx* ctx, const wdg_collide_params* p);
int wdg_contact_loop(wdg_ctx* ctx, const wdg_contact_params* p);
/* ContactLoop.updatePhys (src/dem/ContactLoop.hpp:117, ContactLoop.cpp:180 and 265-266):
 * WDG_UPDATE_PHYS_NEVER keeps each contact's phys from the materials it was created with (a
 * material table changed by wdg_set_materials applies to new contacts only), ALWAYS makes
 * every contact loop take the current materials, ONCE does so in the next contact loop and
 * then reads NEVER. A checkpoint's restored contacts take the current materials. */
#define WDG_UPDATE_PHYS_NEVER 0
#define WDG_UPDATE_PHYS_ALWAYS 1
#define WDG_UPDATE_PHYS_ONCE 2
int wdg_set_update_phys(wdg_ctx* ctx, int32_t mode);
int wdg_get_update_phys(wdg_ctx* ctx, int32_t* mode);
/* ContactLoop.evalStress (src/dem/ContactLoop.hpp:117, ContactLoop.cpp:117,123,216-229,243-254):
 * in a periodic cell each contact loop also sums F branch^T over its real contacts (F the
 * global contact force, branch = Contact::dPos) and divides by the cell volume; read with
 * wdg_contact_loop_stress (ContactLoop.stress, row-major; zero when not evaluated). Cell.gradV
 * is zero here, so the gradV work it feeds is zero. */
int wdg_set_eval_stress(wdg_ctx* ctx, int32_t on);
int wdg_contact_loop_stress(wdg_ctx* ctx, double out[9]);
int wdg_leapfrog(wdg_ctx* ctx, const wdg_leapfrog_params* p);
int wdg_force_reset(wdg_ctx* ctx);
/* IntraForce::run (src/dem/IntraForce.cpp:26-42) with In2_Facet (src/dem/Facet.cpp:301-328,
 * In2_Facet::go / distributeForces with
 * bary=true; src/dem/Contact.cpp:24-34 getForceTorqueBranch): each real contact of a facet
 * adds w_i F and (c - v_i) x (w_i F) to the force and torque of the facet's node i, w being
 * the barycentric coordinates (src/supp/base/CompUtils.cpp:94-112) of the contact point c
 * projected onto the facet plane. The facet nodes are massless and take no gravity, so their
 * reset value is zero: the sums kept are those of this call, made in Particle::contacts order
 * (bit-identical to a sequential loop). Read them with wdg_facet_nodes. */
int wdg_intra_force(wdg_ctx* ctx);
/* The node forces and torques of facets, by user id, from the last IntraForce (zeros before
 * the first): out n*18 doubles per facet, the force on nodes 0, 1, 2 (9), then the torque on
 * nodes 0, 1, 2 (9) (DemData.force / DemData.torque of Facet.nodes, src/dem/Particle.hpp:
 * 281-292, as ReadForce reads them before Leapfrog resets them, src/dem/Impose.cpp:143-158).
 * An id that is not a facet is WDG_E_ARG. */
int wdg_facet_nodes(wdg_ctx* ctx, const int64_t* ids, int64_t n, double* out);
int wdg_crit_dt(wdg_ctx* ctx, double* critDt);

/* which=0: potential+real (the ContactContainer view), which=1: real only.
 * ids n*2 (pA,pB), cellDist n*3, data n*16 doubles:
 *   trsf col0 (3), trsf col1 (3), contact point (3), local force (3), uN, kn, kt, real(1/0).
 * Sorted by (min id, max id). Any output pointer may be NULL. */
int wdg_contact_count(wdg_ctx* ctx, int32_t which, int64_t* n);
int wdg_contact_list(wdg_ctx* ctx, int32_t which, int32_t* ids, int32_t* cellDist, double* data);
/* Load half of a checkpoint of the contact history: the ContactContainer saved with the Scene
 * (Scene::boostSave / saveTmp, src/core/Scene.cpp:285-294; per contact Contact.geom L6Geom
 * uN and trsf, src/dem/L6Geom.hpp:18-21, the contact point Contact.geom.node, Contact.phys
 * force, src/dem/Contact.hpp:21,32). The save half is wdg_contact_list(which=1) and
 * wdg_download of the node fields. Call after wdg_upload_particles with the saved node state:
 * it drops the current contacts, runs a full collider refresh (verletDist as in
 * wdg_collide_params) and makes each listed contact real with its saved state. ids n*2 in role
 * order, cellDist n*3 (periodic cells; may be NULL), data n*16 in the wdg_contact_list layout
 * (kn, kt and the real flag are not read). FP64 contact form only. Stepping on from the
 * restored state gives the trajectory of the uninterrupted run, bit for bit. */
int wdg_restore_contacts(wdg_ctx* ctx, int64_t n, const int32_t* ids, const int32_t* cellDist, const double* data,
                         double verletDist);

/* coordNumber (skipFree=0 and 1) and solid fraction in a box */
int wdg_observables(wdg_ctx* ctx, const double boxMin[3], const double boxMax[3], double out[3]);
/* Contact form: 64 = plain FP64 (the default); 32 = contact geometry, physics and law in
 * FP32 while positions, the position/velocity update and the force sums stay FP64.
 * No counterpart in the reference (a build-only form). Changing it restarts the contact
 * history. */
int wdg_set_precision(wdg_ctx* ctx, int32_t bits);
/* Cauchy stress tensor (row-major 3x3) from the real contacts; volume <= 0 takes the
 * periodic cell's volume (an error for aperiodic scenes) */
int wdg_stress(wdg_ctx* ctx, double volume, double out[9]);
/* sum over integrated nodes of m v.v/2 + w.(I w)/2 */
int wdg_kinetic_energy(wdg_ctx* ctx, double* ek);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops): wdg_kinetic_energy with the same bits, written to dOut in
 * device memory (8 bytes, caller-owned) on the context's stream; the call does not wait for it.
 * It calls no hipMalloc or hipFree: the per-workgroup partials live in a fixed device array of
 * 2 MiB per GPU (2^18 doubles), held until the process ends, so at most 2^26 particles; more is
 * WDG_E_ARG. All calls on one GPU share that array: calls that overlap on one GPU race. */
int wdg_kinetic_energy_noalloc(wdg_ctx* ctx, double* dOut);
/* Scene.trackEnergy (src/core/Scene.hpp:204): while on, wdg_run adds the terms of
 * Leapfrog.cpp:53-98 (kinetic at the on-step velocity, grav, nonviscDamp; a periodic scene
 * from step 1), IdealElPl.cpp:62-92 (elast, plast) and Hertz.cpp:181-206 (elast, plast,
 * viscN, viscT) to the context's energy terms; each is a sum over the workgroups' partials
 * in a fixed order (not ContactContainer order: the last bits differ from a serial sum).
 * One scene per context only. Graph replay (wdg_run_graph, wdg_tune) tracks too, with the
 * same sums; a periodic scene's step 0 is then run plainly. Turning it on does not clear the
 * terms. */
int wdg_set_energy_tracking(wdg_ctx* ctx, int32_t on);
/* Scene.energy: out[WDG_EN_N] in WDG_EN_* order (synchronises the context's stream) */
int wdg_energy(wdg_ctx* ctx, double out[7]);
/* EnergyTracker.clear (src/core/EnergyTracker.hpp:60): every term to zero */
int wdg_energy_clear(wdg_ctx* ctx);
/* EnergyTracker.set / __setitem__ (src/core/EnergyTracker.hpp:46-49,57): the terms from
 * in[WDG_EN_N] in WDG_EN_* order, e.g. those of a checkpoint */
int wdg_energy_set(wdg_ctx* ctx, const double in[7]);
/* Scene.energy of each scene of a batched context of several scenes: out[WDG_EN_N * s + term]
 * (synchronises the stream); wdg_energy gives their sum over the scenes there. The terms of a
 * scene are summed in another order than when it is stepped alone, so they agree to rounding. */
int wdg_energy_scenes(wdg_ctx* ctx, double* out);
/* EnergyTracker.set of each scene (in[WDG_EN_N * nScenes]) */
int wdg_energy_scenes_set(wdg_ctx* ctx, const double* in);
int wdg_stats(wdg_ctx* ctx, wdg_counters* out);

/* run nSteps of Scene::doOneStep with a fixed engine list; *step and *dt are the
 * scene's step counter and timestep, updated in place (DynDt may change dt). */
int wdg_run(wdg_ctx* ctx, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt);

/* Initial-condition generation (the role of SpherePack::makeCloud / pack.randomPeriPack,
 * src/supp/sphere-pack/SpherePack.cpp:140-247, py/pack.py:572): pushes overlapping spheres
 * apart (radii inflated by 1+inflate while relaxing) until the largest overlap is <= tol
 * or maxIter sweeps ran. Drops all contact state afterwards. */
int wdg_relax(wdg_ctx* ctx, int32_t maxIter, double inflate, double tol, double* maxOverlap, int32_t* iters);

int wdg_sync(wdg_ctx* ctx);

/* ---- batched scenes and graph replay (no counterpart in the reference: many independent
 * Scene objects stepped together, for parametric studies).
 * wdg_set_scenes: nScenes independent scenes sharing this context, each with its own
 * Scene.dt (dt[nScenes]); the contact loop and leapfrog then take dt per particle from its
 * scene (the dt passed with the engine parameters is ignored). nScenes = 0 switches back.
 * The graph replay runs [Collider, ContactLoop, Leapfrog] in this mode; wdg_run also runs
 * DynDt, each scene's DynDt::run (DynDt.cpp:90-113) on its own critical timestep and dt.
 * With energy tracking on, each scene keeps its own Scene.energy (wdg_energy_scenes). */
int wdg_set_scenes(wdg_ctx* ctx, int32_t nScenes, const double* dt);
/* the current dt of each scene (out[nScenes]; DynDt changes them) */
int wdg_get_scene_dt(wdg_ctx* ctx, double* out);
/* DynDt::critDt_stiffness (DynDt.cpp:16-71) of each scene (out[nScenes]; inf: none) */
int wdg_crit_dt_scenes(wdg_ctx* ctx, double* out);
/* tuning of the engine: workgroups of the force kernel per group of 2048 contact slots (also
 * used by wdg_run and wdg_contact_loop), steps per captured graph (even), and for batched
 * scenes the per-scene Verlet skin: 0 = every scene uses the collider's Verlet distance;
 * k > 0 = at each refresh each scene's distance is set to cover 1.25 times what its fastest
 * sphere moves in k of its steps (its speed, plus gravity over that time), between the
 * collider's Verlet distance and the smallest radius. The skin changes only which pairs are
 * candidates, never a result. {2, 8, 0} until wdg_tune or wdg_set_tuning changes it */
typedef struct {
  int32_t cfBlocks;
  int32_t graphSteps;
  int32_t skinSteps;
} wdg_tuning;
/* wdg_run for the engine list [Collider, ContactLoop, Leapfrog(reset)] with the contact loop
 * and leapfrog of graphSteps steps captured once as a hipGraph and replayed; a step whose
 * collider would rebuild (a particle left its Verlet skin) ends the replay early on the
 * device and the host runs the rebuild. Same results as wdg_run. */
int wdg_run_graph(wdg_ctx* ctx, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt);
/* autotuner: times each candidate tuning on the live state (real simulation steps, counted
 * in *step) and keeps the fastest; out receives it, stepsTaken the steps run */
int wdg_tune(wdg_ctx* ctx, const wdg_program* prog, int64_t* step, double* dt, wdg_tuning* out, int64_t* stepsTaken);
int wdg_get_tuning(wdg_ctx* ctx, wdg_tuning* out);
int wdg_set_tuning(wdg_ctx* ctx, const wdg_tuning* t);

/* ---- learned coarse step (no counterpart in the reference): a message-passing network on
 * the contact graph of the spheres, BF16 MFMA kernels with FP32 accumulation, trained on
 * trajectories of this engine; predicts each sphere's mean acceleration over a coarse step.
 * Graph: CSR by receiver (rowOff[N+1]), send[E], rev[E] = the opposite edge; node inputs
 * x[N x 7] (velocity / 1 m/s, radius / 2.5 mm, floor overlap ov = 100 (1 - height / r) clipped
 * to +-2, then max(ov,0)^(3/2) / a and v_z max(ov,0)^(1/4) / a with a = r / 2.5 mm), edge
 * inputs ef[E x 15] (relative position / (ri+rj), overlap ov = 100 (1 - d / (ri+rj)) clipped
 * to +-2, relative
 * velocity / 1 m/s, both radii / 2.5 mm, then the normal scaled by the elastic and by the
 * viscous Hertz term per unit mass of the receiver); outputs and targets in units of 100 g
 * (981 m/s^2).
 * woo_amd/csrc/wdg_mpnn.hip (edge_inputs, node_inputs) gives the exact expressions.
 * Errors: wdg_mpnn_last_error. */
typedef struct wdg_mpnn wdg_mpnn;
const char* wdg_mpnn_last_error(void);
int wdg_mpnn_create(int device, uint64_t seed, wdg_mpnn** out);
int wdg_mpnn_destroy(wdg_mpnn* m);
int wdg_mpnn_num_params(wdg_mpnn* m, int64_t* n);
/* parameters (FP32 masters) and the last gradients, host arrays of wdg_mpnn_num_params */
int wdg_mpnn_get_params(wdg_mpnn* m, float* params, float* grads);
int wdg_mpnn_set_params(wdg_mpnn* m, const float* params);
int wdg_mpnn_set_graph(wdg_mpnn* m, int64_t N, int64_t E, const float* x, const int32_t* rowOff,
                       const int32_t* send, const int32_t* rev, const float* ef, const float* targets);
/* targets [N x 3] of the current graph, in the output order (see wdg_mpnn_forward) */
int wdg_mpnn_set_targets(wdg_mpnn* m, const float* targets);
/* the current graph: sizes, CSR, inputs (BF16 as stored, widened to float) and the output row of
 * each node; every array may be NULL */
int wdg_mpnn_get_graph(wdg_mpnn* m, int64_t* N, int64_t* E, int32_t* rowOff, int32_t* send, int32_t* rev, float* x,
                       float* ef, int64_t* nodeRow);
/* keep the current graph with its targets on the device under a slot (a training set), and make
 * a kept graph current again (its outputs then in node order) */
int wdg_mpnn_keep_graph(wdg_mpnn* m, int32_t slot);
int wdg_mpnn_use_graph(wdg_mpnn* m, int32_t slot);
/* forward pass; out [N x 3] (may be NULL): node order of the graph, or for a graph from an
 * engine context its spheres in user-id order */
int wdg_mpnn_forward(wdg_mpnn* m, float* out);
/* forward and backward on the current graph and targets: mean squared error, gradients */
int wdg_mpnn_grad(wdg_mpnn* m, double* loss);
int wdg_mpnn_adam(wdg_mpnn* m, double lr, double beta1, double beta2, double eps);
/* graph of the spheres of an (aperiodic) engine context, edges closer than cutoff (ri+rj)
 * taken from its slot list. The slot list is refreshed first when forced, or when a sphere
 * moved so far since its last refresh that it may miss an edge; a refresh made here raises the
 * context's Verlet distance to (cutoff - 1) rmax + maxDisp if that is larger, so that spheres
 * can move maxDisp before the next one. */
int wdg_mpnn_graph_from_ctx(wdg_mpnn* m, wdg_ctx* ctx, double cutoff, double floorZ, double maxDisp, int force,
                            int64_t* nNodes, int64_t* nEdges, int* rebuilt);
/* coarse step of the context's spheres by the last forward pass: v += a dT, x += v dT */
int wdg_mpnn_advance(wdg_mpnn* m, wdg_ctx* ctx, double dT);
/* In-memory snapshots of the kinematic state, the role of Scene::saveTmp / loadTmp
 * (src/core/Scene.cpp:291-294) for the device: wdg_snapshot keeps positions, velocities,
 * angular velocities, orientations, forces and torques of every particle on the device in slot
 * 0..3, and in the FP64 contact form the contact history (every real contact's L6Geom and
 * force state, keyed by the particles' user ids); wdg_restore writes them back: the slot list
 * is rebuilt at the snapshot's positions and Verlet distance and the saved contacts are real
 * again with their state, so stepping on from a restored snapshot gives the trajectory the
 * state would have had, bit for bit (FP32 form, or a snapshot without contacts: the history
 * restarts and the next collider call rebuilds, resolving its verletDist again, as after an
 * upload); wdg_snapshot_distance gives the largest distance between a sphere's position and
 * its position in the snapshot. A snapshot is valid until the next wdg_upload_particles.
 * wdg_mpnn_advance empties the contact history (the coarse step replaces the state it
 * described). */
int wdg_snapshot(wdg_ctx* ctx, int32_t slot);
int wdg_restore(wdg_ctx* ctx, int32_t slot);
int wdg_snapshot_distance(wdg_ctx* ctx, int32_t slot, double* maxDist);
/* HIP stream the engine launches on (hipStream_t), for callers that time it with events */
void* wdg_stream(wdg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
