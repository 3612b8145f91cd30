// Host side of the DEM step engine: Woo's scene objects as C++ classes.
//
// Scene, DemField, Particle/Node/DemData, the materials, the engines and the functors
// keep the reference's class and attribute names (src/core/Scene.hpp,
// src/dem/Particle.hpp:189-392, src/dem/FrictMat.hpp, src/dem/Hertz.hpp,
// src/dem/Collision.hpp, src/dem/ContactLoop.hpp, src/dem/Leapfrog.hpp,
// src/dem/DynDt.hpp). Engines run on the device through the C ABI
// (include/woo_amd.h); functors are descriptors that select the fused kernel variant
// and its parameters. Particle state lives on the host between runs and is moved
// lazily: host edits are uploaded before the next step, device results are
// downloaded on the first read after a run.
#pragma once
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <exception>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/woo_amd.h"

namespace woo {

using Vec3 = std::array<double, 3>;
using Quat = std::array<double, 4>;  // w, x, y, z
constexpr double NaN = std::numeric_limits<double>::quiet_NaN();
constexpr double Inf = std::numeric_limits<double>::infinity();

struct Object {
  virtual ~Object() = default;
  std::string label;
  virtual const char* className() const = 0;
};

// ------------------------------------------------------------------ materials (FrictMat.hpp, Hertz.hpp)
struct Material : Object {
  double density = 1000;
  const char* className() const override { return "Material"; }
};
struct ElastMat : Material {
  double young = 1e9;
  const char* className() const override { return "ElastMat"; }
};
struct FrictMat : ElastMat {
  double tanPhi = .5, ktDivKn = .2;
  const char* className() const override { return "FrictMat"; }
};
struct HertzMat : FrictMat {
  double surfEnergy = 0, alpha = 0;
  const char* className() const override { return "HertzMat"; }
};

// ------------------------------------------------------------------ shapes (Sphere.hpp, Wall.hpp)
struct Node;
struct Shape : Object {
  std::vector<std::shared_ptr<Node>> nodes;  // Shape.nodes: one node for spheres and walls
  const char* className() const override { return "Shape"; }
};
struct Sphere : Shape {
  double radius = NaN;
  const char* className() const override { return "Sphere"; }
};
struct Wall : Shape {
  int axis = 0, sense = 0;
  const char* className() const override { return "Wall"; }
};
// Facet.hpp: triangle of 3 nodes; this engine supports fixed facets with static nodes
struct Facet : Shape {
  Vec3 fakeVel{};
  double halfThick = 0;
  // Facet.hpp:45: per edge the normal's dot product with the edge neighbour's (NaN: none)
  Vec3 n21lim{NaN, NaN, NaN};
  const char* className() const override { return "Facet"; }
};

struct DemField;

// ------------------------------------------------------------------ DemData, Node, Particle (Particle.hpp)
enum : uint32_t {
  DOF_X = 1, DOF_Y = 2, DOF_Z = 4, DOF_RX = 8, DOF_RY = 16, DOF_RZ = 32, DOF_ALL = 63,
  FLAG_GRAVITY_SKIP = 512, FLAG_DAMPING_SKIP = 2048
};
struct DemData {
  Vec3 vel{}, angVel{}, force{}, torque{}, inertia{};
  double mass = 0;
  uint32_t flags = 0;
  // DemData::guessMoving (Particle.cpp:638-640)
  bool guessMoving() const {
    const bool blockedAll = (flags & DOF_ALL) == DOF_ALL;
    const bool v = vel[0] != 0 || vel[1] != 0 || vel[2] != 0, w = angVel[0] != 0 || angVel[1] != 0 || angVel[2] != 0;
    return (mass != 0. && !blockedAll) || v || w;
  }
  std::string blocked() const;
  void setBlocked(const std::string& s);
};
struct Node : Object {
  Vec3 pos{};
  Quat ori{1, 0, 0, 0};
  DemData dem;
  DemField* field = nullptr;  // set when its particle is added
  bool integrated = false;    // the node is in DemField.nodes
  const char* className() const override { return "Node"; }
};
struct Particle : Object {
  int64_t id = -1;
  std::shared_ptr<Shape> shape;
  std::shared_ptr<Material> material;
  uint32_t mask = 0b0101;
  // the particle's single node (uninodal particles only, like Particle.pos in the reference)
  Node& node() const {
    if (!shape || shape->nodes.size() != 1 || !shape->nodes[0])
      throw std::runtime_error("particle has no shape with exactly one node");
    return *shape->nodes[0];
  }
  // the node that stands for the particle on the device (node 0; a facet's vertex 0)
  Node& node0() const {
    if (!shape || shape->nodes.empty() || !shape->nodes[0]) throw std::runtime_error("particle has no nodes");
    return *shape->nodes[0];
  }
  const char* className() const override { return "Particle"; }
};

// ------------------------------------------------------------------ functors (descriptors)
struct Functor : Object {
  const char* className() const override { return "Functor"; }
};
struct BoundFunctor : Functor {};
struct Bo1_Sphere_Aabb : BoundFunctor {
  const char* className() const override { return "Bo1_Sphere_Aabb"; }
};
struct Bo1_Wall_Aabb : BoundFunctor {
  const char* className() const override { return "Bo1_Wall_Aabb"; }
};
struct Bo1_Facet_Aabb : BoundFunctor {
  const char* className() const override { return "Bo1_Facet_Aabb"; }
};
struct CGeomFunctor : Functor {
  int trsfRenorm = 100;  // Cg2_Any_Any_L6Geom__Base.trsfRenorm (L6Geom.hpp:37)
};
struct Cg2_Sphere_Sphere_L6Geom : CGeomFunctor {
  const char* className() const override { return "Cg2_Sphere_Sphere_L6Geom"; }
};
struct Cg2_Wall_Sphere_L6Geom : CGeomFunctor {
  const char* className() const override { return "Cg2_Wall_Sphere_L6Geom"; }
};
struct Cg2_Facet_Sphere_L6Geom : CGeomFunctor {
  const char* className() const override { return "Cg2_Facet_Sphere_L6Geom"; }
};
struct CPhysFunctor : Functor {};
struct Cp2_FrictMat_FrictPhys : CPhysFunctor {
  const char* className() const override { return "Cp2_FrictMat_FrictPhys"; }
};
struct Cp2_HertzMat_HertzPhys : CPhysFunctor {
  double poisson = .2, en = NaN;  // Hertz.hpp:58-59
  const char* className() const override { return "Cp2_HertzMat_HertzPhys"; }
};
struct LawFunctor : Functor {};
struct Law2_L6Geom_FrictPhys_IdealElPl : LawFunctor {
  bool noSlip = false, noBreak = false, noFrict = false;  // IdealElPl.hpp:22-24
  double relRollStiff = 0., relTwistStiff = 0., rollTanPhi = 0.;  // IdealElPl.hpp:18-20
  const char* className() const override { return "Law2_L6Geom_FrictPhys_IdealElPl"; }
};
struct Law2_L6Geom_HertzPhys_DMT : LawFunctor {
  bool noAttraction = true;  // Hertz.hpp:96
  const char* className() const override { return "Law2_L6Geom_HertzPhys_DMT"; }
};

// IntraForce functors (Facet.hpp:66-76): In2_Facet spreads a facet's contact forces to its nodes
struct IntraFunctor : Functor {};
struct In2_Facet : IntraFunctor {
  const char* className() const override { return "In2_Facet"; }
};

// ------------------------------------------------------------------ engines
struct Scene;
struct Engine : Object {
  bool dead = false;
  const char* className() const override { return "Engine"; }
  virtual double critDt(Scene&) { return Inf; }
};
struct PeriodicEngine : Engine {
  int64_t stepPeriod = 1;  // PeriodicEngine.stepPeriod (only steps are supported)
  int64_t stepLast = -1, nDone = 0;
  // PeriodicEngine::isActivated (Engine.cpp:170-196) with stepModulo and initRun (both true)
  bool isActivated(int64_t stepNow) {
    if (stepLast < 0) stepLast = stepNow;
    if (stepPeriod > 0 && (stepNow - stepLast >= stepPeriod || stepNow % stepPeriod == 0)) {
      stepLast = stepNow;
      nDone++;
      return true;
    }
    return false;
  }
};
struct InsertionSortCollider : Engine {
  std::vector<std::shared_ptr<BoundFunctor>> boundDispatcher;
  double verletDist = -.05;  // Collision.hpp:94
  bool forceInitSort = false;
  const char* className() const override { return "InsertionSortCollider"; }
};
struct ContactLoop : Engine {
  std::vector<std::shared_ptr<CGeomFunctor>> geoDisp;
  std::vector<std::shared_ptr<CPhysFunctor>> phyDisp;
  std::vector<std::shared_ptr<LawFunctor>> lawDisp;
  bool applyForces = true;  // ContactLoop.hpp:118
  // ContactLoop.hpp:119: 0 'never', 1 'always', 2 'once' (back to 'never' after one loop)
  int updatePhys = 0;
  // ContactLoop.hpp:121,123: the minDist00Sq shortcut switch and the OpenMP reordering period;
  // the device keeps its real contacts packed in slot order at every step, so neither applies
  bool dist00 = true;
  int reorderEvery = 1000;
  // ContactLoop.hpp:117,122: the contact loop's stress in periodic cells, read back after each
  // device call while evalStress is set (zero otherwise)
  bool evalStress = false;
  std::array<double, 9> stress{};
  const char* className() const override { return "ContactLoop"; }
};
struct Leapfrog : Engine {
  double damping = .2;  // Leapfrog.hpp:50-51
  bool reset = false;
  const char* className() const override { return "Leapfrog"; }
};
// IntraForce (IntraForce.hpp): dispatches IntraFunctors by shape; the device runs In2_Facet
struct IntraForce : Engine {
  std::vector<std::shared_ptr<IntraFunctor>> funcs;
  const char* className() const override { return "IntraForce"; }
};
struct ForceResetter : Engine {
  const char* className() const override { return "ForceResetter"; }
};
struct DynDt : PeriodicEngine {
  double maxRelInc = 1e-4;  // DynDt.hpp:21
  DynDt() { stepPeriod = 100; }
  const char* className() const override { return "DynDt"; }
  double critDt(Scene& S) override;
};
// runs host code between device steps (the role of woo.core.PyRunner)
struct PyRunner : PeriodicEngine {
  std::function<void(Scene&)> command;
  std::string source;  // for display
  const char* className() const override { return "PyRunner"; }
};

// ------------------------------------------------------------------ contacts (read-only view)
struct ContactView {
  int64_t id1 = -1, id2 = -1;  // pA, pB user ids
  bool real = false;
  std::array<int, 3> cellDist{};
  // L6Geom: trsf rows, contact point, uN; FrictPhys: local force, kn, kt
  std::array<double, 9> trsf{};
  Vec3 contPt{}, force{};
  double uN = NaN, kn = NaN, kt = NaN;
};

// ------------------------------------------------------------------ fields
struct Cell : Object {
  Vec3 size{1, 1, 1};  // diagonal of hSize (gradV is zero)
  const char* className() const override { return "Cell"; }
};

struct DeviceState;
struct DemField : Object {
  Vec3 gravity{0, 0, 0};
  double distFactor = -1;
  uint32_t loneMask = 0b0010;
  std::vector<std::shared_ptr<Particle>> particles;
  Scene* scene = nullptr;
  std::shared_ptr<DeviceState> dev;
  // host/device coherence
  bool structureChanged = true;  // particles, shapes, materials, masks or flags changed: full upload
  bool hostDirty = false;        // kinematic state edited on the host: upload fields
  bool deviceNewer = false;      // device ran steps since the last download
  bool contactsValid = false;
  std::vector<ContactView> contactCache;
  int64_t contactGen = 0;  // counts the rebuilds of contactCache
  // the contact of particles a and b in contactCache (either order), or null
  const ContactView* findContact(int64_t a, int64_t b);
  std::unordered_map<uint64_t, size_t> contactIndex;
  int64_t contactIndexGen = -1;
  const char* className() const override { return "DemField"; }

  // state loaded with Scene.loadTmp, given to the device at the first upload: the real
  // contacts (wdg_contact_list layout) and the resolved Verlet distance of the saved run
  struct SavedContacts {
    std::vector<int32_t> ids, cellDist;
    std::vector<double> data;
    double verletDist = NaN;
  };
  std::shared_ptr<SavedContacts> savedContacts;
  bool devTrackEnergy = false;   // wdg_set_energy_tracking state of the device context
  // the device's material table, in first-use order; pushed again before every run, so that a
  // material edited in place applies as ContactLoop.updatePhys says
  std::vector<std::shared_ptr<Material>> matOrder;

  int64_t add(const std::shared_ptr<Particle>& p, int nodes);
  // device access from another thread while the scene runs in the background is refused
  void guard() const;
  void pull();                   // device -> host, if the device is newer
  void push(Scene& S);           // host -> device, if the host changed
  void touch() { pull(); hostDirty = true; contactsValid = false; }
  void restructure() { pull(); structureChanged = true; contactsValid = false; }
  const std::vector<ContactView>& contacts();
  int64_t countReal();
  double pWaveDt() const;        // DemFuncs::pWaveDt (Funcs.cpp:151-175)
  std::array<double, 3> observables(const Vec3& lo, const Vec3& hi);
  std::array<double, 9> stress(double volume);  // DemFuncs::stressStiffness, stress part
  double kineticEnergy();
  // a copy for Scene.saveTmp: particles, shapes, nodes and materials (shared materials stay
  // shared), with the device's real contacts
  std::shared_ptr<DemField> copy();
};

struct Scene : Object, std::enable_shared_from_this<Scene> {
  std::vector<std::shared_ptr<DemField>> fields;
  std::vector<std::shared_ptr<Engine>> engines;
  double dt = NaN, nextDt = NaN, dtSafety = .9, time = 0;  // Scene.hpp:191-193
  int64_t step = 0;
  bool periodic = false;
  std::shared_ptr<Cell> cell = std::make_shared<Cell>();
  std::atomic<bool> stopRequested{false};
  int device = 0;
  bool trackEnergy = false;  // Scene.trackEnergy (Scene.hpp:204)
  // energy terms given to the device at the next run (Scene.loadTmp, EnergyTracker.set)
  std::shared_ptr<std::array<double, WDG_EN_N>> pendingEnergy;
  const char* className() const override { return "Scene"; }
  ~Scene() override;

  std::shared_ptr<DemField> dem();
  void run(int64_t steps);
  void one() { run(1); }
  void stop() { stopRequested = true; }
  // Scene.run(steps, wait=False) (Scene.cpp:34-51): the steps run on a worker thread; an
  // exception there is kept and rethrown by wait() (Scene.cpp:64-83, 94-117)
  void runAsync(int64_t steps);
  void wait();
  bool running() const { return bgRunning; }
  bool inWorker() const;  // the calling thread is this scene's background worker
  // Scene.energy (EnergyTracker): the device's terms in WDG_EN_* order
  std::array<double, WDG_EN_N> energy();
  void energyClear();
  void energySet(const std::array<double, WDG_EN_N>& v);
  // Scene.saveTmp / loadTmp (Scene.cpp:291-294, Master.cpp:280-300): a copy of the scene in a
  // named in-memory slot, and a fresh copy of it on each load
  std::shared_ptr<Scene> copy();
  void saveTmp(const std::string& slot, bool quiet);
  static std::shared_ptr<Scene> loadTmp(const std::string& slot);

 private:
  std::thread bg;
  std::atomic<bool> bgRunning{false};
  std::exception_ptr bgExc;

  void autoDt(DemField& f);
  wdg_program program(bool& hostEngines);
};

// owns the device context of a DemField
struct DeviceState {
  wdg_ctx* ctx = nullptr;
  explicit DeviceState(int device);
  ~DeviceState();
  DeviceState(const DeviceState&) = delete;
  DeviceState& operator=(const DeviceState&) = delete;
};

void check(int rc);  // throws std::runtime_error with wdg_last_error() on failure

}  // namespace woo
