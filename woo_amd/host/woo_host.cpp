// Host side of the DEM step engine: lazy host/device coherence and the step driver
// (Scene::doOneStep, src/core/Scene.cpp:366-467) over the C ABI.
#include "woo_host.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <typeinfo>

namespace woo {

void check(int rc) {
  if (rc != WDG_OK) throw std::runtime_error(std::string(wdg_last_error()));
}

// 'xyzXYZ' <-> blocked-DOF bits (DemData.blocked, Particle.hpp:210-213)
static const char DOF_CHARS[6] = {'x', 'y', 'z', 'X', 'Y', 'Z'};
std::string DemData::blocked() const {
  std::string s;
  for (int k = 0; k < 6; k++)
    if (flags & (1u << k)) s += DOF_CHARS[k];
  return s;
}
void DemData::setBlocked(const std::string& s) {
  uint32_t b = 0;
  for (char ch : s) {
    const char* p = std::find(DOF_CHARS, DOF_CHARS + 6, ch);
    if (p == DOF_CHARS + 6) throw std::invalid_argument(std::string("invalid blocked DOF '") + ch + "' (use xyzXYZ)");
    b |= 1u << (p - DOF_CHARS);
  }
  flags = (flags & ~(uint32_t)DOF_ALL) | b;
}

DeviceState::DeviceState(int device) { check(wdg_ctx_create(device, &ctx)); }
DeviceState::~DeviceState() {
  if (ctx) wdg_ctx_destroy(ctx);
}

// ------------------------------------------------------------------ DemField
// the scene whose steps the current thread runs in the background (none on other threads)
static thread_local const Scene* tlsWorkerOf = nullptr;

bool Scene::inWorker() const { return tlsWorkerOf == this; }

void DemField::guard() const {
  if (scene && scene->running() && tlsWorkerOf != scene)
    throw std::runtime_error("the scene is running in the background (call S.wait() first)");
}

int64_t DemField::add(const std::shared_ptr<Particle>& p, int nodes) {
  guard();
  if (!p || !p->shape || !p->material) throw std::invalid_argument("particle needs a material and a shape");
  if (p->id >= 0) throw std::invalid_argument("particle was already added (id " + std::to_string(p->id) + ")");
  if (auto* f = dynamic_cast<Facet*>(p->shape.get())) {
    if (f->nodes.size() != 3 || !f->nodes[0] || !f->nodes[1] || !f->nodes[2])
      throw std::invalid_argument("Facet needs 3 nodes");
    for (auto& n : f->nodes) {
      const DemData& d = n->dem;
      if ((d.flags & DOF_ALL) != DOF_ALL || d.vel != Vec3{} || d.angVel != Vec3{})
        throw std::invalid_argument("only fixed facets with static nodes are supported (use Facet.fakeVel)");
    }
  } else {
    if (p->shape->nodes.size() != 1 || !p->shape->nodes[0])
      throw std::invalid_argument("particle needs a shape with one node");
    if (!dynamic_cast<Sphere*>(p->shape.get()) && !dynamic_cast<Wall*>(p->shape.get()))
      throw std::invalid_argument(std::string("shape ") + p->shape->className() +
                                  " is not supported by the device engine");
  }
  pull();
  p->id = (int64_t)particles.size();
  for (auto& n : p->shape->nodes) n->field = this;
  // ParticleContainer::pyAppend (ParticleContainer.cpp:176-201): nodes=1 always, 0 never,
  // -1 by DemData::guessMoving; facet nodes here are fixed and static
  if (dynamic_cast<Facet*>(p->shape.get())) {
    for (auto& n : p->shape->nodes) n->integrated = false;
  } else {
    p->node().integrated = nodes == 1 || (nodes == -1 && p->node().dem.guessMoving());
  }
  particles.push_back(p);
  structureChanged = true;
  contactsValid = false;
  return p->id;
}

static const int FIELDS[8] = {WDG_F_POS, WDG_F_VEL, WDG_F_ANGVEL, WDG_F_ORI,
                              WDG_F_FORCE, WDG_F_TORQUE, WDG_F_MASS, WDG_F_INERTIA};

static double* node_field(Node& n, int f) {
  switch (f) {
    case WDG_F_POS: return n.pos.data();
    case WDG_F_VEL: return n.dem.vel.data();
    case WDG_F_ANGVEL: return n.dem.angVel.data();
    case WDG_F_ORI: return n.ori.data();
    case WDG_F_FORCE: return n.dem.force.data();
    case WDG_F_TORQUE: return n.dem.torque.data();
    case WDG_F_MASS: return &n.dem.mass;
    case WDG_F_INERTIA: return n.dem.inertia.data();
  }
  return nullptr;
}
static int field_width(int f) { return f == WDG_F_ORI ? 4 : (f == WDG_F_MASS ? 1 : 3); }

void DemField::pull() {
  if (!deviceNewer || !dev) return;
  guard();
  deviceNewer = false;
  const size_t n = particles.size();
  std::vector<double> buf;
  for (int f : FIELDS) {
    const int w = field_width(f);
    buf.resize(n * w);
    check(wdg_download(dev->ctx, f, buf.data()));
    for (size_t i = 0; i < n; i++) std::copy(&buf[i * w], &buf[i * w] + w, node_field(particles[i]->node0(), f));
  }
  // facet nodes: the force and torque of the last IntraForce (In2_Facet), zero without one
  std::vector<int64_t> fid;
  for (size_t i = 0; i < n; i++)
    if (dynamic_cast<const Facet*>(particles[i]->shape.get())) fid.push_back((int64_t)i);
  if (fid.empty()) return;
  std::vector<double> fn(18 * fid.size());
  check(wdg_facet_nodes(dev->ctx, fid.data(), (int64_t)fid.size(), fn.data()));
  for (size_t j = 0; j < fid.size(); j++) {
    auto& nodes = particles[fid[j]]->shape->nodes;
    for (int k = 0; k < 3; k++)
      for (int q = 0; q < 3; q++) {
        nodes[k]->dem.force[q] = fn[18 * j + 3 * k + q];
        nodes[k]->dem.torque[q] = fn[18 * j + 9 + 3 * k + q];
      }
  }
}

static wdg_material device_material(const Material* mat) {
  const auto* fm = dynamic_cast<const FrictMat*>(mat);
  if (!fm) throw std::runtime_error(std::string("material ") + mat->className() + " is not supported (FrictMat or HertzMat)");
  const auto* hm = dynamic_cast<const HertzMat*>(fm);
  wdg_material m{};
  m.kind = hm ? WDG_HERTZMAT : WDG_FRICTMAT;
  m.density = fm->density;
  m.young = fm->young;
  m.ktDivKn = fm->ktDivKn;
  m.tanPhi = fm->tanPhi;
  m.surfEnergy = hm ? hm->surfEnergy : 0.;
  m.alpha = hm ? hm->alpha : 0.;
  return m;
}

void DemField::push(Scene& S) {
  guard();
  if (!dev) {
    dev = std::make_shared<DeviceState>(S.device);
    structureChanged = true;
    devTrackEnergy = false;
  }
  wdg_field fd{};
  for (int k = 0; k < 3; k++) fd.gravity[k] = gravity[k];
  fd.distFactor = distFactor;
  fd.loneMask = loneMask;
  fd.periodic = S.periodic ? 1 : 0;
  for (int k = 0; k < 3; k++) fd.cell[k] = S.cell->size[k];
  check(wdg_set_field(dev->ctx, &fd));
  const size_t n = particles.size();
  if (structureChanged) {
    // materials in first-use order
    std::map<const Material*, int> midx;
    std::vector<wdg_material> mats;
    matOrder.clear();
    std::vector<int32_t> shape(n), axis(n), sense(n), mat(n);
    std::vector<uint32_t> mask(n), flags(n);
    std::vector<uint8_t> integ(n);
    std::vector<double> radius(n), bufs[8], fverts(9 * n, 0.), fvel(3 * n, 0.), fht(n, 0.), fn21(3 * n, NaN);
    for (size_t i = 0; i < n; i++) {
      const Particle& p = *particles[i];
      auto it = midx.find(p.material.get());
      if (it == midx.end()) {
        const wdg_material m = device_material(p.material.get());
        it = midx.emplace(p.material.get(), (int)mats.size()).first;
        mats.push_back(m);
        matOrder.push_back(p.material);
      }
      mat[i] = it->second;
      if (const auto* s = dynamic_cast<const Sphere*>(p.shape.get())) {
        shape[i] = WDG_SPHERE;
        radius[i] = s->radius;
      } else if (const auto* f = dynamic_cast<const Facet*>(p.shape.get())) {
        shape[i] = WDG_FACET;
        for (int v = 0; v < 3; v++)
          for (int k = 0; k < 3; k++) fverts[9 * i + 3 * v + k] = f->nodes[v]->pos[k];
        for (int k = 0; k < 3; k++) fvel[3 * i + k] = f->fakeVel[k];
        fht[i] = f->halfThick;
        for (int k = 0; k < 3; k++) fn21[3 * i + k] = f->n21lim[k];
      } else {
        const auto* w = dynamic_cast<const Wall*>(p.shape.get());
        shape[i] = WDG_WALL;
        axis[i] = w->axis;
        sense[i] = w->sense;
      }
      mask[i] = p.mask;
      flags[i] = p.node0().dem.flags;
      integ[i] = p.node0().integrated ? 1 : 0;
    }
    for (int fi = 0; fi < 8; fi++) {
      const int w = field_width(FIELDS[fi]);
      bufs[fi].resize(n * w);
      for (size_t i = 0; i < n; i++) {
        const double* src = node_field(particles[i]->node0(), FIELDS[fi]);
        std::copy(src, src + w, &bufs[fi][i * w]);
      }
    }
    check(wdg_set_materials(dev->ctx, mats.data(), (int32_t)mats.size()));
    wdg_particles pd{};
    pd.n = (int64_t)n;
    pd.shape = shape.data(); pd.radius = radius.data(); pd.wallAxis = axis.data(); pd.wallSense = sense.data();
    pd.material = mat.data(); pd.mask = mask.data(); pd.flags = flags.data(); pd.integrated = integ.data();
    pd.pos = bufs[0].data(); pd.vel = bufs[1].data(); pd.angVel = bufs[2].data(); pd.ori = bufs[3].data();
    pd.force = bufs[4].data(); pd.torque = bufs[5].data(); pd.mass = bufs[6].data(); pd.inertia = bufs[7].data();
    pd.facetVerts = fverts.data(); pd.fakeVel = fvel.data(); pd.halfThick = fht.data(); pd.facetN21lim = fn21.data();
    check(wdg_upload_particles(dev->ctx, &pd));
    structureChanged = hostDirty = false;
    if (savedContacts) {
      // a loaded copy: its real contacts come back with their history
      const auto sc = savedContacts;
      savedContacts.reset();
      check(wdg_restore_contacts(dev->ctx, (int64_t)sc->ids.size() / 2, sc->ids.data(),
                                 sc->cellDist.empty() ? nullptr : sc->cellDist.data(), sc->data.data(),
                                 sc->verletDist));
    }
  } else {
    // materials edited in place (the device ignores an unchanged table)
    std::vector<wdg_material> mats;
    for (auto& m : matOrder) mats.push_back(device_material(m.get()));
    check(wdg_set_materials(dev->ctx, mats.data(), (int32_t)mats.size()));
  }
  if (!structureChanged && hostDirty) {
    std::vector<double> buf;
    for (int f : FIELDS) {
      const int w = field_width(f);
      buf.resize(n * w);
      for (size_t i = 0; i < n; i++) {
        const double* src = node_field(particles[i]->node0(), f);
        std::copy(src, src + w, &buf[i * w]);
      }
      check(wdg_upload(dev->ctx, f, buf.data()));
    }
    hostDirty = false;
  }
}

const std::vector<ContactView>& DemField::contacts() {
  guard();
  if (contactsValid) return contactCache;
  contactCache.clear();
  if (dev && !structureChanged) {
    int64_t nc = 0;
    check(wdg_contact_count(dev->ctx, 0, &nc));
    std::vector<int32_t> ids(2 * nc), cd(3 * nc);
    std::vector<double> data(16 * nc);
    if (nc) check(wdg_contact_list(dev->ctx, 0, ids.data(), cd.data(), data.data()));
    contactCache.resize(nc);
    for (int64_t i = 0; i < nc; i++) {
      ContactView& c = contactCache[i];
      const double* d = &data[16 * i];
      c.id1 = ids[2 * i];
      c.id2 = ids[2 * i + 1];
      for (int k = 0; k < 3; k++) c.cellDist[k] = cd[3 * i + k];
      c.real = d[15] != 0.;
      if (c.real) {
        // L6Geom.trsf rows are the local axes: row 0 = col0 of the stored columns, etc.
        const double c0[3] = {d[0], d[1], d[2]}, c1[3] = {d[3], d[4], d[5]};
        const double c2[3] = {c0[1] * c1[2] - c0[2] * c1[1], c0[2] * c1[0] - c0[0] * c1[2],
                              c0[0] * c1[1] - c0[1] * c1[0]};
        for (int k = 0; k < 3; k++) {
          c.trsf[k] = c0[k];
          c.trsf[3 + k] = c1[k];
          c.trsf[6 + k] = c2[k];
        }
        c.contPt = {d[6], d[7], d[8]};
        c.force = {d[9], d[10], d[11]};
        c.uN = d[12];
        c.kn = d[13];
        c.kt = d[14];
      }
    }
  }
  contactsValid = true;
  contactGen++;
  return contactCache;
}

const ContactView* DemField::findContact(int64_t a, int64_t b) {
  const auto& v = contacts();
  if (contactIndexGen != contactGen) {
    contactIndex.clear();
    for (size_t i = 0; i < v.size(); i++)
      contactIndex[(uint64_t)std::min(v[i].id1, v[i].id2) << 32 | (uint64_t)std::max(v[i].id1, v[i].id2)] = i;
    contactIndexGen = contactGen;
  }
  auto it = contactIndex.find((uint64_t)std::min(a, b) << 32 | (uint64_t)std::max(a, b));
  return it == contactIndex.end() ? nullptr : &v[it->second];
}

int64_t DemField::countReal() {
  int64_t n = 0;
  for (const auto& c : contacts()) n += c.real ? 1 : 0;
  return n;
}

double DemField::pWaveDt() const {
  double dt = Inf;
  for (const auto& p : particles) {
    const auto* s = dynamic_cast<const Sphere*>(p->shape.get());
    const auto* m = dynamic_cast<const ElastMat*>(p->material.get());
    if (!s || !m) continue;
    dt = std::min(dt, s->radius / std::sqrt(m->young / m->density));
  }
  return dt;
}

std::array<double, 3> DemField::observables(const Vec3& lo, const Vec3& hi) {
  guard();
  if (!dev || structureChanged) throw std::runtime_error("no simulation state on the device yet (run a step first)");
  std::array<double, 3> out{};
  check(wdg_observables(dev->ctx, lo.data(), hi.data(), out.data()));
  return out;
}

std::array<double, 9> DemField::stress(double volume) {
  guard();
  if (!dev || structureChanged) throw std::runtime_error("no simulation state on the device yet (run a step first)");
  std::array<double, 9> out{};
  check(wdg_stress(dev->ctx, volume, out.data()));
  return out;
}

double DemField::kineticEnergy() {
  guard();
  if (!dev || structureChanged) throw std::runtime_error("no simulation state on the device yet (run a step first)");
  double e = 0;
  check(wdg_kinetic_energy(dev->ctx, &e));
  return e;
}

// ------------------------------------------------------------------ engines
double DynDt::critDt(Scene& S) {
  auto f = S.dem();
  if (!f || !f->dev || f->structureChanged) return Inf;
  double d = Inf;
  check(wdg_crit_dt(f->dev->ctx, &d));
  return d;
}

// ------------------------------------------------------------------ Scene
std::shared_ptr<DemField> Scene::dem() {
  for (auto& f : fields)
    if (f) return f;
  return nullptr;
}

// Scene::doOneStep (Scene.cpp:380-398): dt = dtSafety * min(critDt of engines and fields)
void Scene::autoDt(DemField& f) {
  double d = Inf;
  for (auto& e : engines)
    if (e && !e->dead) d = std::min(d, e->critDt(*this));
  d = std::min(d, f.pWaveDt());
  if (std::isinf(d)) throw std::runtime_error("Failed to obtain meaningful dt from engines and fields automatically.");
  dt = dtSafety * d;
}

template <class T>
static T* find_functor(const std::vector<std::shared_ptr<T>>& v, const char* name) {
  for (auto& f : v)
    if (f && std::string(f->className()) == name) return f.get();
  return nullptr;
}

wdg_program Scene::program(bool& hostEngines) {
  wdg_program p{};
  hostEngines = false;
  auto f = dem();
  bool anyWall = false, anySphere = false, anyFacet = false, allHertz = true;
  for (auto& q : f->particles) {
    if (dynamic_cast<Wall*>(q->shape.get())) anyWall = true;
    else if (dynamic_cast<Facet*>(q->shape.get())) anyFacet = true;
    else anySphere = true;
    if (!dynamic_cast<HertzMat*>(q->material.get())) allHertz = false;
  }
  for (auto& e : engines) {
    if (!e || e->dead) continue;
    if (p.nEngines >= 8) throw std::runtime_error("at most 8 engines are supported");
    if (auto* c = dynamic_cast<InsertionSortCollider*>(e.get())) {
      if (anySphere && !find_functor(c->boundDispatcher, "Bo1_Sphere_Aabb"))
        throw std::runtime_error("InsertionSortCollider: spheres need Bo1_Sphere_Aabb");
      if (anyWall && !find_functor(c->boundDispatcher, "Bo1_Wall_Aabb"))
        throw std::runtime_error("InsertionSortCollider: walls need Bo1_Wall_Aabb");
      if (anyFacet && !find_functor(c->boundDispatcher, "Bo1_Facet_Aabb"))
        throw std::runtime_error("InsertionSortCollider: facets need Bo1_Facet_Aabb");
      p.collide.verletDist = c->verletDist;
      p.collide.forceInit = c->forceInitSort ? 1 : 0;
      p.engines[p.nEngines++] = WDG_E_COLLIDER;
    } else if (auto* cl = dynamic_cast<ContactLoop*>(e.get())) {
      if (anySphere && !find_functor(cl->geoDisp, "Cg2_Sphere_Sphere_L6Geom"))
        throw std::runtime_error("ContactLoop: spheres need Cg2_Sphere_Sphere_L6Geom");
      if (anyWall && !find_functor(cl->geoDisp, "Cg2_Wall_Sphere_L6Geom"))
        throw std::runtime_error("ContactLoop: walls need Cg2_Wall_Sphere_L6Geom");
      if (anyFacet && !find_functor(cl->geoDisp, "Cg2_Facet_Sphere_L6Geom"))
        throw std::runtime_error("ContactLoop: facets need Cg2_Facet_Sphere_L6Geom");
      p.contact.trsfRenorm = cl->geoDisp.empty() ? 100 : cl->geoDisp[0]->trsfRenorm;
      p.contact.applyForces = cl->applyForces ? 1 : 0;
      // CPhys dispatch by material type: HertzMat pairs get HertzPhys when its functor is
      // present, FrictMat pairs FrictPhys; the law follows the phys type
      auto* hz = static_cast<Cp2_HertzMat_HertzPhys*>(find_functor(cl->phyDisp, "Cp2_HertzMat_HertzPhys"));
      auto* fr = find_functor(cl->phyDisp, "Cp2_FrictMat_FrictPhys");
      auto* dmt = static_cast<Law2_L6Geom_HertzPhys_DMT*>(find_functor(cl->lawDisp, "Law2_L6Geom_HertzPhys_DMT"));
      auto* iep =
          static_cast<Law2_L6Geom_FrictPhys_IdealElPl*>(find_functor(cl->lawDisp, "Law2_L6Geom_FrictPhys_IdealElPl"));
      if (hz && allHertz && dmt) {
        p.contact.physKind = WDG_PHYS_HERTZ;
        p.contact.lawKind = WDG_LAW_HERTZDMT;
        p.contact.poisson = hz->poisson;
        p.contact.en = hz->en;
        p.contact.noAttraction = dmt->noAttraction ? 1 : 0;
      } else if (fr && iep) {
        p.contact.physKind = WDG_PHYS_FRICT;
        p.contact.lawKind = WDG_LAW_IDEALELPL;
        p.contact.poisson = .2;
        p.contact.en = NaN;
        p.contact.noSlip = iep->noSlip;
        p.contact.noBreak = iep->noBreak;
        p.contact.noFrict = iep->noFrict;
        p.contact.relRollStiff = iep->relRollStiff;
        p.contact.relTwistStiff = iep->relTwistStiff;
        p.contact.rollTanPhi = iep->rollTanPhi;
      } else {
        throw std::runtime_error("ContactLoop: unsupported functor combination (supported: Cp2_FrictMat_FrictPhys + "
                                 "Law2_L6Geom_FrictPhys_IdealElPl, Cp2_HertzMat_HertzPhys + Law2_L6Geom_HertzPhys_DMT "
                                 "with HertzMat particles)");
      }
      p.engines[p.nEngines++] = WDG_E_CONTACTLOOP;
    } else if (auto* lf = dynamic_cast<Leapfrog*>(e.get())) {
      p.leapfrog.damping = lf->damping;
      p.leapfrog.reset = lf->reset ? 1 : 0;
      p.engines[p.nEngines++] = WDG_E_LEAPFROG;
    } else if (auto* it = dynamic_cast<IntraForce*>(e.get())) {
      if (anyFacet && !find_functor(it->funcs, "In2_Facet")) throw std::runtime_error("IntraForce: facets need In2_Facet");
      p.engines[p.nEngines++] = WDG_E_INTRAFORCE;
    } else if (dynamic_cast<ForceResetter*>(e.get())) {
      p.engines[p.nEngines++] = WDG_E_FORCERESETTER;
    } else if (auto* dd = dynamic_cast<DynDt*>(e.get())) {
      p.dynDtPeriod = (int32_t)dd->stepPeriod;
      p.maxRelInc = dd->maxRelInc;
      p.engines[p.nEngines++] = WDG_E_DYNDT;
    } else if (dynamic_cast<PyRunner*>(e.get())) {
      hostEngines = true;
    } else {
      throw std::runtime_error(std::string("engine ") + e->className() + " is not supported by the device engine");
    }
  }
  p.dtSafety = dtSafety;
  return p;
}

void Scene::run(int64_t steps) {
  auto f = dem();
  if (!f) throw std::runtime_error("Scene has no DemField");
  if (steps < 0) throw std::invalid_argument("the number of steps must be given (run(steps))");
  stopRequested = false;
  f->scene = this;
  f->push(*this);
  wdg_ctx* ctx = f->dev->ctx;
  if (f->devTrackEnergy != trackEnergy) {
    check(wdg_set_energy_tracking(ctx, trackEnergy ? 1 : 0));
    f->devTrackEnergy = trackEnergy;
  }
  if (pendingEnergy) {
    check(wdg_energy_set(ctx, pendingEnergy->data()));
    pendingEnergy.reset();
  }
  if (std::isnan(dt) || std::isinf(dt) || dt < 0) autoDt(*f);
  bool hostEngines = false;
  const wdg_program prog = program(hostEngines);
  bool hasDyn = false;
  for (int i = 0; i < prog.nEngines; i++) hasDyn = hasDyn || prog.engines[i] == WDG_E_DYNDT;
  f->contactsValid = false;
  // ContactLoop.updatePhys to the device before each device call, and back ('once' turns into
  // 'never' after one contact loop)
  ContactLoop* cl = nullptr;
  for (auto& e : engines)
    if (e && !e->dead && !cl) cl = dynamic_cast<ContactLoop*>(e.get());
  auto device_run = [&](const wdg_program* pr, int64_t k, int64_t* st, double* d) {
    if (cl) {
      check(wdg_set_update_phys(ctx, cl->updatePhys));
      check(wdg_set_eval_stress(ctx, cl->evalStress ? 1 : 0));
    }
    check(wdg_run(ctx, pr, k, st, d));
    if (cl) {
      int32_t m = 0;
      check(wdg_get_update_phys(ctx, &m));
      cl->updatePhys = m;
      bool ran = false;
      for (int e = 0; e < pr->nEngines; e++) ran = ran || pr->engines[e] == WDG_E_CONTACTLOOP;
      if (ran) check(wdg_contact_loop_stress(ctx, cl->stress.data()));
    }
  };
  if (!hostEngines && !hasDyn) {
    // one call for all steps (a background run: chunks, so that stop() ends it between two);
    // time accumulates step by step as in doOneStep
    const int64_t chunk = tlsWorkerOf == this ? 256 : steps;
    for (int64_t done = 0; done < steps && !(done > 0 && stopRequested);) {
      const int64_t k = std::min(chunk, steps - done);
      int64_t st = step;
      double d = dt;
      f->deviceNewer = true;
      device_run(&prog, k, &st, &d);
      for (int64_t s = 0; s < k; s++) time += dt;
      step = st;
      done += k;
    }
    return;
  }
  for (int64_t s = 0; s < steps && !stopRequested; s++) {
    if (!hostEngines) {
      const double d0 = dt;
      f->deviceNewer = true;
      device_run(&prog, 1, &step, &dt);  // applies DynDt's nextDt at the end of the step
      time += d0;
      continue;
    }
    // engine by engine: device engines as one-engine programs, host engines in between
    nextDt = NaN;
    int di = 0;
    for (auto& e : engines) {
      if (!e || e->dead) continue;
      if (auto* pr = dynamic_cast<PyRunner*>(e.get())) {
        if (pr->isActivated(step) && pr->command) {
          pr->command(*this);
          f->push(*this);
          f->contactsValid = false;
        }
        continue;
      }
      wdg_program one = prog;
      one.nEngines = 1;
      one.engines[0] = prog.engines[di++];
      int64_t st = step;
      double d = dt;
      f->deviceNewer = true;
      device_run(&one, 1, &st, &d);
      if (one.engines[0] == WDG_E_DYNDT && d != dt) nextDt = d;
    }
    step++;
    time += dt;
    if (!std::isnan(nextDt)) {
      dt = nextDt;
      nextDt = NaN;
    }
  }
}

Scene::~Scene() {
  if (bg.joinable()) {
    stopRequested = true;
    // the worker holds the scene; the last reference can be dropped on the worker itself
    if (bg.get_id() == std::this_thread::get_id()) bg.detach();
    else bg.join();
  }
}

void Scene::runAsync(int64_t steps) {
  if (bgRunning) throw std::runtime_error("Scene.run: already running");
  if (bg.joinable()) bg.join();
  if (!dem()) throw std::runtime_error("Scene has no DemField");
  bgExc = nullptr;
  bgRunning = true;
  // the worker keeps the scene alive until it ends
  auto self = std::static_pointer_cast<Scene>(shared_from_this());
  bg = std::thread([self, steps] {
    tlsWorkerOf = self.get();
    try {
      self->run(steps);
    } catch (...) {
      self->bgExc = std::current_exception();
    }
    tlsWorkerOf = nullptr;
    self->bgRunning = false;
  });
}

void Scene::wait() {
  if (bg.joinable() && bg.get_id() != std::this_thread::get_id()) bg.join();
  if (bgExc) {
    std::exception_ptr e = bgExc;
    bgExc = nullptr;
    std::rethrow_exception(e);
  }
}

std::array<double, WDG_EN_N> Scene::energy() {
  std::array<double, WDG_EN_N> e{};
  if (pendingEnergy) return *pendingEnergy;
  auto f = dem();
  if (!f || !f->dev) return e;
  f->guard();
  check(wdg_energy(f->dev->ctx, e.data()));
  return e;
}

void Scene::energyClear() {
  pendingEnergy.reset();
  auto f = dem();
  if (!f || !f->dev) return;
  f->guard();
  check(wdg_energy_clear(f->dev->ctx));
}

void Scene::energySet(const std::array<double, WDG_EN_N>& v) {
  auto f = dem();
  if (!f || !f->dev) {
    pendingEnergy = std::make_shared<std::array<double, WDG_EN_N>>(v);
    return;
  }
  f->guard();
  check(wdg_energy_set(f->dev->ctx, v.data()));
}

// ------------------------------------------------------------------ copies (saveTmp / loadTmp)
// a copy of the object behind p, of its exact class among T, Ts...
template <class B, class T, class... Ts>
static std::shared_ptr<B> copy_as(const std::shared_ptr<B>& p) {
  if (!p) return nullptr;
  if (typeid(*p) == typeid(T)) return std::make_shared<T>(static_cast<const T&>(*p));
  if constexpr (sizeof...(Ts) > 0) return copy_as<B, Ts...>(p);
  else throw std::runtime_error(std::string("Scene.saveTmp: cannot copy a ") + p->className());
}
template <class B>
static std::vector<std::shared_ptr<B>> copy_functors(const std::vector<std::shared_ptr<B>>& v) {
  std::vector<std::shared_ptr<B>> out;
  for (const auto& f : v) {
    std::shared_ptr<Functor> g =
        copy_as<Functor, Bo1_Sphere_Aabb, Bo1_Wall_Aabb, Bo1_Facet_Aabb, Cg2_Sphere_Sphere_L6Geom, Cg2_Wall_Sphere_L6Geom,
                Cg2_Facet_Sphere_L6Geom, Cp2_FrictMat_FrictPhys, Cp2_HertzMat_HertzPhys, Law2_L6Geom_FrictPhys_IdealElPl,
                Law2_L6Geom_HertzPhys_DMT, In2_Facet>(std::static_pointer_cast<Functor>(f));
    out.push_back(std::static_pointer_cast<B>(g));
  }
  return out;
}
static std::shared_ptr<Engine> copy_engine(const std::shared_ptr<Engine>& e) {
  auto c = copy_as<Engine, InsertionSortCollider, ContactLoop, Leapfrog, ForceResetter, DynDt, PyRunner, IntraForce>(e);
  if (auto* x = dynamic_cast<InsertionSortCollider*>(c.get())) x->boundDispatcher = copy_functors(x->boundDispatcher);
  if (auto* x = dynamic_cast<ContactLoop*>(c.get())) {
    x->geoDisp = copy_functors(x->geoDisp);
    x->phyDisp = copy_functors(x->phyDisp);
    x->lawDisp = copy_functors(x->lawDisp);
  }
  if (auto* x = dynamic_cast<IntraForce*>(c.get())) x->funcs = copy_functors(x->funcs);
  return c;
}

std::shared_ptr<DemField> DemField::copy() {
  guard();
  pull();
  auto f = std::make_shared<DemField>();
  f->label = label;
  f->gravity = gravity;
  f->distFactor = distFactor;
  f->loneMask = loneMask;
  std::map<const Material*, std::shared_ptr<Material>> mats;
  std::map<const Node*, std::shared_ptr<Node>> nodes;
  for (const auto& p : particles) {
    auto q = std::make_shared<Particle>(*p);
    auto& m = mats[p->material.get()];
    if (!m) m = copy_as<Material, HertzMat, FrictMat, ElastMat, Material>(p->material);
    q->material = m;
    q->shape = copy_as<Shape, Sphere, Wall, Facet, Shape>(p->shape);
    for (auto& n : q->shape->nodes) {
      auto& c = nodes[n.get()];
      if (!c) {
        c = std::make_shared<Node>(*n);
        c->field = f.get();
      }
      n = c;
    }
    f->particles.push_back(q);
  }
  f->structureChanged = true;
  if (dev && !structureChanged) {
    auto sc = std::make_shared<SavedContacts>();
    int64_t nc = 0;
    check(wdg_contact_count(dev->ctx, 1, &nc));
    sc->ids.resize(2 * nc);
    sc->cellDist.resize(3 * nc);
    sc->data.resize(16 * nc);
    if (nc) check(wdg_contact_list(dev->ctx, 1, sc->ids.data(), sc->cellDist.data(), sc->data.data()));
    wdg_counters st{};
    check(wdg_stats(dev->ctx, &st));
    sc->verletDist = st.verletDist;
    if (nc) f->savedContacts = sc;
  } else {
    f->savedContacts = savedContacts;
  }
  return f;
}

std::shared_ptr<Scene> Scene::copy() {
  auto f = dem();
  if (f) f->guard();
  auto S = std::make_shared<Scene>();
  S->label = label;
  S->dt = dt; S->nextDt = nextDt; S->dtSafety = dtSafety; S->time = time; S->step = step;
  S->periodic = periodic;
  S->device = device;
  S->trackEnergy = trackEnergy;
  S->cell = std::make_shared<Cell>(*cell);
  for (const auto& e : engines) S->engines.push_back(copy_engine(e));
  for (const auto& g : fields) S->fields.push_back(g ? g->copy() : nullptr);
  const auto e = energy();
  bool any = trackEnergy;
  for (double x : e) any = any || x != 0.;
  if (any) S->pendingEnergy = std::make_shared<std::array<double, WDG_EN_N>>(e);
  return S;
}

static std::mutex g_tmpMu;
static std::map<std::string, std::shared_ptr<Scene>>& tmp_store() {
  static std::map<std::string, std::shared_ptr<Scene>> m;
  return m;
}

void Scene::saveTmp(const std::string& slot, bool /*quiet*/) {
  auto c = copy();
  std::lock_guard<std::mutex> lk(g_tmpMu);
  tmp_store()[slot] = c;
}

std::shared_ptr<Scene> Scene::loadTmp(const std::string& slot) {
  std::shared_ptr<Scene> s;
  {
    std::lock_guard<std::mutex> lk(g_tmpMu);
    auto it = tmp_store().find(slot);
    if (it == tmp_store().end()) throw std::runtime_error("Scene.loadTmp: no scene saved in slot '" + slot + "'");
    s = it->second;
  }
  return s->copy();
}

}  // namespace woo
