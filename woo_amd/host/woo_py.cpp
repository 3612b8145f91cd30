// Python bindings of the host classes (module woo_amd._woo). Class, attribute and
// method names follow the reference's Python surface (py/_cxxInternal.cpp and the
// WOO_CLASS_BASE_DOC_ATTRS blocks of src/core/Scene.hpp, src/dem/Particle.hpp,
// src/dem/ParticleContainer.hpp:182-184, src/dem/ContactContainer.hpp): objects take
// their attributes as keyword arguments, vectors are returned as numpy arrays.
#include <algorithm>
#include <cmath>

#include <pybind11/eval.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "woo_host.hpp"

namespace py = pybind11;
using namespace woo;

namespace {

py::array_t<double> arr(const double* p, size_t n) {
  py::array_t<double> a(n);
  std::copy(p, p + n, a.mutable_data());
  return a;
}
Vec3 vec3(py::handle h) {
  auto v = h.cast<std::vector<double>>();
  if (v.size() != 3) throw py::value_error("expected 3 components");
  return {v[0], v[1], v[2]};
}
Quat quat(py::handle h) {
  auto v = h.cast<std::vector<double>>();
  if (v.size() != 4) throw py::value_error("expected 4 components (w, x, y, z)");
  return {v[0], v[1], v[2], v[3]};
}

// keyword arguments become attributes, as in the reference's Object constructors
template <class T>
std::shared_ptr<T> make_kw(const py::kwargs& kw) {
  auto o = std::make_shared<T>();
  py::object po = py::cast(o);
  for (auto item : kw) py::setattr(po, item.first, item.second);
  return o;
}

// state access through the node's field keeps host and device coherent
void before_read(Node& n) {
  if (n.field) n.field->pull();
}
void before_write(Node& n, bool structural) {
  if (!n.field) return;
  if (structural) n.field->restructure();
  else n.field->touch();
}

#define NODE_VEC(cls, name, expr, structural)                                                   \
  def_property(                                                                                 \
      name, [](cls& self) { Node& n = node_of(self); before_read(n); return arr((expr).data(), (expr).size()); }, \
      [](cls& self, py::handle v) { Node& n = node_of(self); before_write(n, structural); auto x = vec3(v); std::copy(x.begin(), x.end(), (expr).begin()); })

struct DemDataRef {
  std::shared_ptr<Node> node;
};
Node& node_of(DemDataRef& r) { return *r.node; }
Node& node_of(Node& n) { return n; }
Node& node_of(Particle& p) { return p.node(); }

struct ParRef {
  std::shared_ptr<DemField> f;
};
struct ConRef {
  std::shared_ptr<DemField> f;
};
// A contact as Python holds it. The reference's Contact objects are shared and updated in
// place by the engines (ContactContainer holds shared_ptr<Contact>), so a handle kept across
// steps reads the current state: each read after the field's contacts changed looks the pair
// up again; a pair no longer in the container reads as not real (a removed Contact is reset).
struct LiveContact {
  std::shared_ptr<DemField> f;
  ContactView c;
  int64_t gen = -1;
  const ContactView& view() {
    if (f && (!f->contactsValid || f->contactGen != gen)) {
      const ContactView* x = f->findContact(c.id1, c.id2);
      gen = f->contactGen;
      if (x) {
        c = *x;
      } else {
        const int64_t a = c.id1, b = c.id2;
        c = ContactView{};
        c.id1 = a;
        c.id2 = b;
      }
    }
    return c;
  }
};
struct ContactRef {
  std::shared_ptr<LiveContact> L;
  const ContactView& c() const { return L->view(); }
};
struct L6GeomRef {
  std::shared_ptr<LiveContact> L;
  const ContactView& c() const { return L->view(); }
};
struct PhysRef {
  std::shared_ptr<LiveContact> L;
  const ContactView& c() const { return L->view(); }
};

ContactRef contact_ref(const std::shared_ptr<DemField>& f, const ContactView& c) {
  return ContactRef{std::make_shared<LiveContact>(LiveContact{f, c, f->contactGen})};
}

py::object contact_or_none(const std::shared_ptr<DemField>& f, int64_t a, int64_t b) {
  const ContactView* c = f->findContact(a, b);
  return c ? py::cast(contact_ref(f, *c)) : py::none();
}

// Python code run by an engine: on a background run its Python error becomes a C++ one, kept
// by the worker and raised again by S.wait() (the Python error object stays on this thread)
template <class F>
void in_python(Scene& S, F&& f) {
  if (!S.inWorker()) {
    f();
    return;
  }
  try {
    f();
  } catch (py::error_already_set& e) {
    throw std::runtime_error(e.what());
  }
}

// Scene.energy: EnergyTracker's Python surface (EnergyTracker.hpp:30-60) over the device's terms
struct EnergyRef {
  std::shared_ptr<Scene> S;
};
static const char* const EN_NAMES[WDG_EN_N] = {"kinetic", "elast", "grav", "nonviscDamp", "plast", "viscN", "viscT"};
static int en_index(const std::string& name) {
  for (int k = 0; k < WDG_EN_N; k++)
    if (name == EN_NAMES[k]) return k;
  return -1;
}
// the terms created so far (non-zero), in name order as the reference's std::map
static std::vector<std::pair<std::string, double>> en_items(Scene& S) {
  const auto e = S.energy();
  std::vector<std::pair<std::string, double>> v;
  for (int k = 0; k < WDG_EN_N; k++)
    if (e[k] != 0.) v.emplace_back(EN_NAMES[k], e[k]);
  std::sort(v.begin(), v.end());
  return v;
}

// S.lab: labelled engines and functors (LabelMapper)
py::object labels(Scene& S) {
  py::object ns = py::module_::import("types").attr("SimpleNamespace")();
  auto put = [&](const std::shared_ptr<Object>& o) {
    if (o && !o->label.empty()) py::setattr(ns, o->label.c_str(), py::cast(o));
  };
  for (auto& e : S.engines) {
    put(e);
    if (auto* c = dynamic_cast<InsertionSortCollider*>(e.get()))
      for (auto& f : c->boundDispatcher) put(f);
    if (auto* c = dynamic_cast<ContactLoop*>(e.get())) {
      for (auto& f : c->geoDisp) put(f);
      for (auto& f : c->phyDisp) put(f);
      for (auto& f : c->lawDisp) put(f);
    }
    if (auto* c = dynamic_cast<IntraForce*>(e.get()))
      for (auto& f : c->funcs) put(f);
  }
  return ns;
}

}  // namespace

PYBIND11_MODULE(_woo, m) {
  m.doc() = "Host classes of the MI355X DEM step engine (Woo's Scene/DemField/engine surface)";

  py::class_<Object, std::shared_ptr<Object>>(m, "Object", py::dynamic_attr())
      .def_readwrite("label", &Object::label)
      .def("__repr__", [](Object& o) {
        return "<" + std::string(o.className()) + (o.label.empty() ? "" : " @" + o.label) + ">";
      });

  // ---------------------------------------------------------------- materials
  py::class_<Material, Object, std::shared_ptr<Material>>(m, "Material", py::dynamic_attr())
      .def(py::init(&make_kw<Material>))
      .def_readwrite("density", &Material::density);
  py::class_<ElastMat, Material, std::shared_ptr<ElastMat>>(m, "ElastMat", py::dynamic_attr())
      .def(py::init(&make_kw<ElastMat>))
      .def_readwrite("young", &ElastMat::young);
  py::class_<FrictMat, ElastMat, std::shared_ptr<FrictMat>>(m, "FrictMat", py::dynamic_attr())
      .def(py::init(&make_kw<FrictMat>))
      .def_readwrite("tanPhi", &FrictMat::tanPhi)
      .def_readwrite("ktDivKn", &FrictMat::ktDivKn);
  py::class_<HertzMat, FrictMat, std::shared_ptr<HertzMat>>(m, "HertzMat", py::dynamic_attr())
      .def(py::init(&make_kw<HertzMat>))
      .def_readwrite("surfEnergy", &HertzMat::surfEnergy)
      .def_readwrite("alpha", &HertzMat::alpha);

  // ---------------------------------------------------------------- nodes and DemData
  py::class_<DemDataRef>(m, "DemData")
      .NODE_VEC(DemDataRef, "vel", n.dem.vel, false)
      .NODE_VEC(DemDataRef, "angVel", n.dem.angVel, false)
      .NODE_VEC(DemDataRef, "force", n.dem.force, false)
      .NODE_VEC(DemDataRef, "torque", n.dem.torque, false)
      .NODE_VEC(DemDataRef, "inertia", n.dem.inertia, false)
      .def_property(
          "mass", [](DemDataRef& r) { before_read(*r.node); return r.node->dem.mass; },
          [](DemDataRef& r, double v) { before_write(*r.node, false); r.node->dem.mass = v; })
      .def_property(
          "blocked", [](DemDataRef& r) { return r.node->dem.blocked(); },
          [](DemDataRef& r, const std::string& s) { before_write(*r.node, true); r.node->dem.setBlocked(s); })
      .def_property(
          "flags", [](DemDataRef& r) { return r.node->dem.flags; },
          [](DemDataRef& r, uint32_t f) { before_write(*r.node, true); r.node->dem.flags = f; })
      .def_property(
          "gravitySkip", [](DemDataRef& r) { return (r.node->dem.flags & FLAG_GRAVITY_SKIP) != 0; },
          [](DemDataRef& r, bool b) {
            before_write(*r.node, true);
            r.node->dem.flags = b ? (r.node->dem.flags | FLAG_GRAVITY_SKIP) : (r.node->dem.flags & ~FLAG_GRAVITY_SKIP);
          })
      .def_property(
          "dampingSkip", [](DemDataRef& r) { return (r.node->dem.flags & FLAG_DAMPING_SKIP) != 0; },
          [](DemDataRef& r, bool b) {
            before_write(*r.node, true);
            r.node->dem.flags = b ? (r.node->dem.flags | FLAG_DAMPING_SKIP) : (r.node->dem.flags & ~FLAG_DAMPING_SKIP);
          })
      .def("guessMoving", [](DemDataRef& r) { before_read(*r.node); return r.node->dem.guessMoving(); });

  py::class_<Node, Object, std::shared_ptr<Node>>(m, "Node", py::dynamic_attr())
      .def(py::init(&make_kw<Node>))
      .NODE_VEC(Node, "pos", n.pos, false)
      .def_property(
          "ori", [](Node& n) { before_read(n); return arr(n.ori.data(), 4); },
          [](Node& n, py::handle v) { before_write(n, false); n.ori = quat(v); })
      .def_property_readonly("dem", [](std::shared_ptr<Node> n) { return DemDataRef{n}; });

  // ---------------------------------------------------------------- shapes
  py::class_<Shape, Object, std::shared_ptr<Shape>>(m, "Shape", py::dynamic_attr())
      .def_readwrite("nodes", &Shape::nodes);
  py::class_<Sphere, Shape, std::shared_ptr<Sphere>>(m, "Sphere", py::dynamic_attr())
      .def(py::init(&make_kw<Sphere>))
      .def_property(
          "radius", [](Sphere& s) { return s.radius; },
          [](Sphere& s, double r) {
            if (!s.nodes.empty() && s.nodes[0] && s.nodes[0]->field) s.nodes[0]->field->restructure();
            s.radius = r;
          });
  py::class_<Wall, Shape, std::shared_ptr<Wall>>(m, "Wall", py::dynamic_attr())
      .def(py::init(&make_kw<Wall>))
      .def_readwrite("axis", &Wall::axis)
      .def_readwrite("sense", &Wall::sense);
  auto restructure_shape = [](Shape& s) {
    if (!s.nodes.empty() && s.nodes[0] && s.nodes[0]->field) s.nodes[0]->field->restructure();
  };
  py::class_<Facet, Shape, std::shared_ptr<Facet>>(m, "Facet", py::dynamic_attr())
      .def(py::init(&make_kw<Facet>))
      .def_property(
          "fakeVel", [](Facet& f) { return arr(f.fakeVel.data(), 3); },
          [restructure_shape](Facet& f, py::handle v) { restructure_shape(f); f.fakeVel = vec3(v); })
      .def_property(
          "halfThick", [](Facet& f) { return f.halfThick; },
          [restructure_shape](Facet& f, double h) { restructure_shape(f); f.halfThick = h; })
      .def_property(
          "n21lim", [](Facet& f) { return arr(f.n21lim.data(), 3); },
          [restructure_shape](Facet& f, py::handle v) { restructure_shape(f); f.n21lim = vec3(v); })
      // Facet::computeNeighborAngles (Facet.cpp:180-216): an edge shared (both nodes) with exactly
      // one other facet of the field gets n1 . n2, the neighbour's normal taken in this facet's
      // orientation; the other edges NaN
      .def("computeNeighborAngles", [restructure_shape](Facet& f) {
        if (f.nodes.size() != 3) throw std::runtime_error("Facet needs 3 nodes");
        DemField* fld = f.nodes[0] ? f.nodes[0]->field : nullptr;
        if (!fld) throw std::runtime_error("Facet.computeNeighborAngles: the facet is not in a DemField");
        auto normal = [](const Facet& g) {
          const Vec3 &a = g.nodes[0]->pos, &b = g.nodes[1]->pos, &c = g.nodes[2]->pos;
          const Vec3 u{b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
          Vec3 n{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
          const double z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
          if (z > 0) {
            const double l = std::sqrt(z);
            for (auto& x : n) x = x / l;
          }
          return n;
        };
        restructure_shape(f);
        const Vec3 n1 = normal(f);
        for (int e0 = 0; e0 < 3; e0++) {
          const int e1 = (e0 + 1) % 3;
          int cnt = 0;
          const Facet* f2 = nullptr;
          for (const auto& p : fld->particles) {
            const auto* g = dynamic_cast<const Facet*>(p->shape.get());
            if (!g || g == &f || g->nodes.size() != 3) continue;
            bool h0 = false, h1 = false;
            for (const auto& nd : g->nodes) {
              h0 = h0 || nd == f.nodes[e0];
              h1 = h1 || nd == f.nodes[e1];
            }
            if (h0 && h1) {
              f2 = g;
              cnt++;
            }
          }
          if (cnt != 1) {
            f.n21lim[e0] = NaN;
            continue;
          }
          bool inv = false;
          for (int i = 0; i < 3; i++) {
            if (f2->nodes[i] != f.nodes[e0]) continue;
            inv = f2->nodes[(i + 1) % 3] == f.nodes[e1];
            break;
          }
          const Vec3 n2 = normal(*f2);
          const double s = inv ? -1. : 1.;
          f.n21lim[e0] = (n1[0] * (n2[0] * s) + n1[1] * (n2[1] * s)) + n1[2] * (n2[2] * s);
        }
      })
      // Facet::getNormal (Facet.cpp:74-77)
      .def("getNormal", [](Facet& f) {
        if (f.nodes.size() != 3) throw std::runtime_error("Facet needs 3 nodes");
        const Vec3 &a = f.nodes[0]->pos, &b = f.nodes[1]->pos, &c = f.nodes[2]->pos;
        const Vec3 u{b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        Vec3 n{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (auto& x : n) x /= l;
        return arr(n.data(), 3);
      });

  // ---------------------------------------------------------------- particles
  py::class_<Particle, Object, std::shared_ptr<Particle>>(m, "Particle", py::dynamic_attr())
      .def(py::init(&make_kw<Particle>))
      .def_readonly("id", &Particle::id)
      .def_readwrite("shape", &Particle::shape)
      .def_property(
          "material", [](Particle& p) { return p.material; },
          [](Particle& p, std::shared_ptr<Material> m) {
            if (p.id >= 0) p.node0().field->restructure();
            p.material = m;
          })
      .def_property(
          "mat", [](Particle& p) { return p.material; },
          [](Particle& p, std::shared_ptr<Material> m) {
            if (p.id >= 0) p.node0().field->restructure();
            p.material = m;
          })
      .def_property(
          "mask", [](Particle& p) { return p.mask; },
          [](Particle& p, uint32_t v) {
            if (p.id >= 0) p.node0().field->restructure();
            p.mask = v;
          })
      // Particle shorthands for the single node (Particle.hpp:100-120)
      .NODE_VEC(Particle, "pos", n.pos, false)
      .NODE_VEC(Particle, "vel", n.dem.vel, false)
      .NODE_VEC(Particle, "angVel", n.dem.angVel, false)
      .NODE_VEC(Particle, "f", n.dem.force, false)
      .NODE_VEC(Particle, "t", n.dem.torque, false)
      .NODE_VEC(Particle, "inertia", n.dem.inertia, false)
      .def_property(
          "ori", [](Particle& p) { before_read(p.node()); return arr(p.node().ori.data(), 4); },
          [](Particle& p, py::handle v) { before_write(p.node(), false); p.node().ori = quat(v); })
      .def_property(
          "mass", [](Particle& p) { before_read(p.node()); return p.node().dem.mass; },
          [](Particle& p, double v) { before_write(p.node(), false); p.node().dem.mass = v; })
      .def_property(
          "blocked", [](Particle& p) { return p.node().dem.blocked(); },
          [](Particle& p, const std::string& s) { before_write(p.node(), true); p.node().dem.setBlocked(s); })
      .def_property_readonly("Ek", [](Particle& p) {
        Node& n = p.node();
        before_read(n);
        double e = 0;
        for (int k = 0; k < 3; k++) e += .5 * n.dem.mass * n.dem.vel[k] * n.dem.vel[k];
        for (int k = 0; k < 3; k++) e += .5 * n.dem.inertia[k] * n.dem.angVel[k] * n.dem.angVel[k];
        return e;
      });

  // ---------------------------------------------------------------- contacts
  py::class_<L6GeomRef>(m, "L6Geom")
      .def_property_readonly("uN", [](L6GeomRef& g) { return g.c().uN; })
      .def_property_readonly("trsf", [](L6GeomRef& g) {
        py::array_t<double> a({3, 3});
        std::copy(g.c().trsf.begin(), g.c().trsf.end(), a.mutable_data());
        return a;
      })
      .def_property_readonly("contPt", [](L6GeomRef& g) { return arr(g.c().contPt.data(), 3); })
      .def_property_readonly("node", [](L6GeomRef& g) {
        auto n = std::make_shared<Node>();
        n->pos = g.c().contPt;
        return n;
      });
  py::class_<PhysRef>(m, "FrictPhys")
      .def_property_readonly("force", [](PhysRef& p) { return arr(p.c().force.data(), 3); })
      .def_property_readonly("kn", [](PhysRef& p) { return p.c().kn; })
      .def_property_readonly("kt", [](PhysRef& p) { return p.c().kt; });
  py::class_<ContactRef>(m, "Contact")
      .def_property_readonly("id1", [](ContactRef& c) { return c.c().id1; })
      .def_property_readonly("id2", [](ContactRef& c) { return c.c().id2; })
      .def_property_readonly("real", [](ContactRef& c) { return c.c().real; })
      .def_property_readonly("cellDist", [](ContactRef& c) { return c.c().cellDist; })
      .def_property_readonly("geom", [](ContactRef& c) -> py::object {
        return c.c().real ? py::cast(L6GeomRef{c.L}) : py::none();
      })
      .def_property_readonly("phys", [](ContactRef& c) -> py::object {
        return c.c().real ? py::cast(PhysRef{c.L}) : py::none();
      })
      .def("__repr__", [](ContactRef& c) {
        return "<Contact ##" + std::to_string(c.c().id1) + "+" + std::to_string(c.c().id2) +
               (c.c().real ? "" : " (potential)") + ">";
      });

  py::class_<ConRef>(m, "ContactContainer")
      .def("__len__", [](ConRef& r) { return r.f->contacts().size(); })
      .def("countReal", [](ConRef& r) { return r.f->countReal(); })
      .def("__getitem__", [](ConRef& r, int64_t i) {
        const auto& v = r.f->contacts();
        if (i < 0) i += (int64_t)v.size();
        if (i < 0 || i >= (int64_t)v.size()) throw py::index_error("contact index out of range");
        return contact_ref(r.f, v[(size_t)i]);
      })
      .def("__getitem__", [](ConRef& r, std::pair<int64_t, int64_t> ij) {
        py::object c = contact_or_none(r.f, ij.first, ij.second);
        if (c.is_none()) throw py::index_error("no contact between the particles");
        return c;
      })
      .def("exists", [](ConRef& r, int64_t a, int64_t b) { return r.f->findContact(a, b) != nullptr; })
      .def("existsReal", [](ConRef& r, int64_t a, int64_t b) {
        const ContactView* c = r.f->findContact(a, b);
        return c && c->real;
      })
      .def("__iter__", [](ConRef& r) {
        py::list l;
        for (const auto& c : r.f->contacts()) l.append(contact_ref(r.f, c));
        return py::iter(l);
      });

  py::class_<ParRef>(m, "ParticleContainer")
      .def("add", [](ParRef& r, std::shared_ptr<Particle> p, int nodes) { return r.f->add(p, nodes); },
           py::arg("par"), py::arg("nodes") = -1)
      .def("add",
           [](ParRef& r, std::vector<std::shared_ptr<Particle>> pp, int nodes) {
             py::list ids;
             for (auto& p : pp) ids.append(r.f->add(p, nodes));
             return ids;
           },
           py::arg("pars"), py::arg("nodes") = -1)
      .def("__len__", [](ParRef& r) { return r.f->particles.size(); })
      .def("__getitem__", [](ParRef& r, int64_t i) {
        const int64_t n = (int64_t)r.f->particles.size();
        if (i < 0) i += n;
        if (i < 0 || i >= n) throw py::index_error("particle id out of range");
        return r.f->particles[(size_t)i];
      })
      .def("__iter__", [](ParRef& r) { return py::iter(py::cast(r.f->particles)); });

  // ---------------------------------------------------------------- fields
  py::class_<Cell, Object, std::shared_ptr<Cell>>(m, "Cell", py::dynamic_attr())
      .def(py::init(&make_kw<Cell>))
      .def_property(
          "size", [](Cell& c) { return arr(c.size.data(), 3); }, [](Cell& c, py::handle v) { c.size = vec3(v); })
      .def("setBox", [](Cell& c, py::handle v) { c.size = vec3(v); });

  py::class_<DemField, Object, std::shared_ptr<DemField>>(m, "DemField", py::dynamic_attr())
      .def(py::init(&make_kw<DemField>))
      .def_property(
          "gravity", [](DemField& f) { return arr(f.gravity.data(), 3); },
          [](DemField& f, py::handle v) { f.gravity = vec3(v); })
      .def_readwrite("distFactor", &DemField::distFactor)
      .def_property(
          "loneMask", [](DemField& f) { return f.loneMask; },
          [](DemField& f, uint32_t v) { f.restructure(); f.loneMask = v; })
      // DemField(par=[...]): the list added to an empty field, as the reference's list to
      // ParticleContainer conversion does
      .def_property(
          "par", [](std::shared_ptr<DemField> f) { return ParRef{f}; },
          [](std::shared_ptr<DemField> f, std::vector<std::shared_ptr<Particle>> pp) {
            if (!f->particles.empty()) throw py::value_error("DemField.par is assigned only while empty (use par.add)");
            for (auto& p : pp) f->add(p, -1);
          })
      .def_property_readonly("con", [](std::shared_ptr<DemField> f) { return ConRef{f}; })
      .def_property_readonly("nodes", [](DemField& f) {
        py::list l;
        for (auto& p : f.particles)
          if (p->node0().integrated) l.append(p->shape->nodes[0]);
        return l;
      })
      .def("critDt", &DemField::pWaveDt)
      .def("observables", [](DemField& f, py::handle lo, py::handle hi) { return f.observables(vec3(lo), vec3(hi)); })
      .def("stress",
           [](DemField& f, double volume) {
             auto s = f.stress(volume);
             py::array_t<double> a({3, 3});
             std::copy(s.begin(), s.end(), a.mutable_data());
             return a;
           },
           py::arg("volume") = -1.)
      .def("kineticEnergy", &DemField::kineticEnergy)
      .def_property_readonly_static("defaultMovableMask", [](py::object) { return 0b0101; })
      .def_property_readonly_static("defaultBoundaryMask", [](py::object) { return 0b0011; })
      .def_property_readonly_static("defaultLoneMask", [](py::object) { return 0b0010; });

  // ---------------------------------------------------------------- functors
  py::class_<Functor, Object, std::shared_ptr<Functor>>(m, "Functor", py::dynamic_attr());
  py::class_<BoundFunctor, Functor, std::shared_ptr<BoundFunctor>>(m, "BoundFunctor", py::dynamic_attr());
  py::class_<Bo1_Sphere_Aabb, BoundFunctor, std::shared_ptr<Bo1_Sphere_Aabb>>(m, "Bo1_Sphere_Aabb", py::dynamic_attr())
      .def(py::init(&make_kw<Bo1_Sphere_Aabb>));
  py::class_<Bo1_Wall_Aabb, BoundFunctor, std::shared_ptr<Bo1_Wall_Aabb>>(m, "Bo1_Wall_Aabb", py::dynamic_attr())
      .def(py::init(&make_kw<Bo1_Wall_Aabb>));
  py::class_<Bo1_Facet_Aabb, BoundFunctor, std::shared_ptr<Bo1_Facet_Aabb>>(m, "Bo1_Facet_Aabb", py::dynamic_attr())
      .def(py::init(&make_kw<Bo1_Facet_Aabb>));
  py::class_<CGeomFunctor, Functor, std::shared_ptr<CGeomFunctor>>(m, "CGeomFunctor", py::dynamic_attr())
      .def_readwrite("trsfRenorm", &CGeomFunctor::trsfRenorm);
  py::class_<Cg2_Sphere_Sphere_L6Geom, CGeomFunctor, std::shared_ptr<Cg2_Sphere_Sphere_L6Geom>>(
      m, "Cg2_Sphere_Sphere_L6Geom", py::dynamic_attr())
      .def(py::init(&make_kw<Cg2_Sphere_Sphere_L6Geom>));
  py::class_<Cg2_Wall_Sphere_L6Geom, CGeomFunctor, std::shared_ptr<Cg2_Wall_Sphere_L6Geom>>(
      m, "Cg2_Wall_Sphere_L6Geom", py::dynamic_attr())
      .def(py::init(&make_kw<Cg2_Wall_Sphere_L6Geom>));
  py::class_<Cg2_Facet_Sphere_L6Geom, CGeomFunctor, std::shared_ptr<Cg2_Facet_Sphere_L6Geom>>(
      m, "Cg2_Facet_Sphere_L6Geom", py::dynamic_attr())
      .def(py::init(&make_kw<Cg2_Facet_Sphere_L6Geom>));
  py::class_<CPhysFunctor, Functor, std::shared_ptr<CPhysFunctor>>(m, "CPhysFunctor", py::dynamic_attr());
  py::class_<Cp2_FrictMat_FrictPhys, CPhysFunctor, std::shared_ptr<Cp2_FrictMat_FrictPhys>>(
      m, "Cp2_FrictMat_FrictPhys", py::dynamic_attr())
      .def(py::init(&make_kw<Cp2_FrictMat_FrictPhys>));
  py::class_<Cp2_HertzMat_HertzPhys, CPhysFunctor, std::shared_ptr<Cp2_HertzMat_HertzPhys>>(
      m, "Cp2_HertzMat_HertzPhys", py::dynamic_attr())
      .def(py::init(&make_kw<Cp2_HertzMat_HertzPhys>))
      .def_readwrite("poisson", &Cp2_HertzMat_HertzPhys::poisson)
      .def_readwrite("en", &Cp2_HertzMat_HertzPhys::en);
  py::class_<LawFunctor, Functor, std::shared_ptr<LawFunctor>>(m, "LawFunctor", py::dynamic_attr());
  py::class_<Law2_L6Geom_FrictPhys_IdealElPl, LawFunctor, std::shared_ptr<Law2_L6Geom_FrictPhys_IdealElPl>>(
      m, "Law2_L6Geom_FrictPhys_IdealElPl", py::dynamic_attr())
      .def(py::init(&make_kw<Law2_L6Geom_FrictPhys_IdealElPl>))
      .def_readwrite("noSlip", &Law2_L6Geom_FrictPhys_IdealElPl::noSlip)
      .def_readwrite("noBreak", &Law2_L6Geom_FrictPhys_IdealElPl::noBreak)
      .def_readwrite("noFrict", &Law2_L6Geom_FrictPhys_IdealElPl::noFrict)
      .def_readwrite("relRollStiff", &Law2_L6Geom_FrictPhys_IdealElPl::relRollStiff)
      .def_readwrite("relTwistStiff", &Law2_L6Geom_FrictPhys_IdealElPl::relTwistStiff)
      .def_readwrite("rollTanPhi", &Law2_L6Geom_FrictPhys_IdealElPl::rollTanPhi);
  py::class_<Law2_L6Geom_HertzPhys_DMT, LawFunctor, std::shared_ptr<Law2_L6Geom_HertzPhys_DMT>>(
      m, "Law2_L6Geom_HertzPhys_DMT", py::dynamic_attr())
      .def(py::init(&make_kw<Law2_L6Geom_HertzPhys_DMT>))
      .def_readwrite("noAttraction", &Law2_L6Geom_HertzPhys_DMT::noAttraction);

  py::class_<IntraFunctor, Functor, std::shared_ptr<IntraFunctor>>(m, "IntraFunctor", py::dynamic_attr());
  py::class_<In2_Facet, IntraFunctor, std::shared_ptr<In2_Facet>>(m, "In2_Facet", py::dynamic_attr())
      .def(py::init(&make_kw<In2_Facet>));

  // ---------------------------------------------------------------- engines
  py::class_<Engine, Object, std::shared_ptr<Engine>>(m, "Engine", py::dynamic_attr())
      .def_readwrite("dead", &Engine::dead)
      .def("critDt", [](Engine& e, Scene& S) { return e.critDt(S); });
  py::class_<PeriodicEngine, Engine, std::shared_ptr<PeriodicEngine>>(m, "PeriodicEngine", py::dynamic_attr())
      .def_readwrite("stepPeriod", &PeriodicEngine::stepPeriod)
      .def_readwrite("stepLast", &PeriodicEngine::stepLast)
      .def_readonly("nDone", &PeriodicEngine::nDone);
  // InsertionSortCollider([Bo1_*], verletDist=...) (utils.py:118)
  py::class_<InsertionSortCollider, Engine, std::shared_ptr<InsertionSortCollider>>(m, "InsertionSortCollider",
                                                                                    py::dynamic_attr())
      .def(py::init([](py::args a, py::kwargs kw) {
        auto o = make_kw<InsertionSortCollider>(kw);
        if (a.size() > 1) throw py::type_error("InsertionSortCollider takes one positional argument (bound functors)");
        if (a.size() == 1) o->boundDispatcher = a[0].cast<std::vector<std::shared_ptr<BoundFunctor>>>();
        return o;
      }))
      .def_readwrite("boundDispatcher", &InsertionSortCollider::boundDispatcher)
      .def_readwrite("verletDist", &InsertionSortCollider::verletDist)
      .def_readwrite("forceInitSort", &InsertionSortCollider::forceInitSort);
  // ContactLoop([Cg2_*], [Cp2_*], [Law2_*], applyForces=...) (ContactLoop.cpp:52-66)
  py::class_<ContactLoop, Engine, std::shared_ptr<ContactLoop>>(m, "ContactLoop", py::dynamic_attr())
      .def(py::init([](py::args a, py::kwargs kw) {
        auto o = make_kw<ContactLoop>(kw);
        if (a.size() != 0 && a.size() != 3) throw py::type_error("ContactLoop takes 0 or 3 positional arguments");
        if (a.size() == 3) {
          o->geoDisp = a[0].cast<std::vector<std::shared_ptr<CGeomFunctor>>>();
          o->phyDisp = a[1].cast<std::vector<std::shared_ptr<CPhysFunctor>>>();
          o->lawDisp = a[2].cast<std::vector<std::shared_ptr<LawFunctor>>>();
        }
        return o;
      }))
      .def_readwrite("geoDisp", &ContactLoop::geoDisp)
      .def_readwrite("phyDisp", &ContactLoop::phyDisp)
      .def_readwrite("lawDisp", &ContactLoop::lawDisp)
      .def_readwrite("applyForces", &ContactLoop::applyForces)
      // a named enum (ContactLoop.hpp:119): reads as its name, takes the name or the number
      .def_property(
          "updatePhys",
          [](const ContactLoop& c) { return std::string(c.updatePhys == 1 ? "always" : c.updatePhys == 2 ? "once" : "never"); },
          [](ContactLoop& c, py::object v) {
            static const char* names[] = {"never", "always", "once"};
            if (py::isinstance<py::str>(v)) {
              const std::string s = v.cast<std::string>();
              for (int k = 0; k < 3; k++)
                if (s == names[k]) { c.updatePhys = k; return; }
              throw py::value_error("ContactLoop.updatePhys: '" + s + "' is not one of never, always, once");
            }
            const int k = v.cast<int>();
            if (k < 0 || k > 2) throw py::value_error("ContactLoop.updatePhys: 0 (never), 1 (always) or 2 (once)");
            c.updatePhys = k;
          })
      .def_readwrite("dist00", &ContactLoop::dist00)
      .def_readwrite("reorderEvery", &ContactLoop::reorderEvery)
      .def_readwrite("evalStress", &ContactLoop::evalStress)
      .def_property_readonly("stress", [](ContactLoop& c) {
        py::array_t<double> a({3, 3});
        std::copy(c.stress.begin(), c.stress.end(), a.mutable_data());
        return a;
      });
  py::class_<Leapfrog, Engine, std::shared_ptr<Leapfrog>>(m, "Leapfrog", py::dynamic_attr())
      .def(py::init(&make_kw<Leapfrog>))
      .def_readwrite("damping", &Leapfrog::damping)
      .def_readwrite("reset", &Leapfrog::reset);
  // IntraForce([In2_*]) (IntraForce.hpp)
  py::class_<IntraForce, Engine, std::shared_ptr<IntraForce>>(m, "IntraForce", py::dynamic_attr())
      .def(py::init([](py::args a, py::kwargs kw) {
        auto o = make_kw<IntraForce>(kw);
        if (a.size() > 1) throw py::type_error("IntraForce takes one positional argument (intra functors)");
        if (a.size() == 1) o->funcs = a[0].cast<std::vector<std::shared_ptr<IntraFunctor>>>();
        return o;
      }))
      .def_readwrite("funcs", &IntraForce::funcs);
  py::class_<ForceResetter, Engine, std::shared_ptr<ForceResetter>>(m, "ForceResetter", py::dynamic_attr())
      .def(py::init(&make_kw<ForceResetter>));
  py::class_<DynDt, PeriodicEngine, std::shared_ptr<DynDt>>(m, "DynDt", py::dynamic_attr())
      .def(py::init(&make_kw<DynDt>))
      .def_readwrite("maxRelInc", &DynDt::maxRelInc);
  // PyRunner(stepPeriod, command) or PyRunner(command=..., stepPeriod=...) (Engine.cpp:198-220)
  py::class_<PyRunner, PeriodicEngine, std::shared_ptr<PyRunner>>(m, "PyRunner", py::dynamic_attr())
      .def(py::init([](py::args a, py::kwargs kw) {
        auto o = std::make_shared<PyRunner>();
        py::object po = py::cast(o);
        for (auto item : kw) py::setattr(po, item.first, item.second);
        for (auto h : a) {
          if (py::isinstance<py::int_>(h)) o->stepPeriod = h.cast<int64_t>();
          else py::setattr(po, "command", h);
        }
        return o;
      }))
      .def_property(
          "command", [](PyRunner& r) { return r.source; },
          [](PyRunner& r, py::object cmd) {
            // the engine keeps the command itself (its Python wrapper may not survive); the
            // reference is dropped with the GIL held
            std::shared_ptr<py::object> keep(new py::object(cmd), [](py::object* o) {
              py::gil_scoped_acquire gil;
              delete o;
            });
            if (py::isinstance<py::str>(cmd)) {
              r.source = cmd.cast<std::string>();
              // a string runs with S and woo in its namespace (PyRunner.command)
              r.command = [keep](Scene& S) {
                py::gil_scoped_acquire gil;
                in_python(S, [&] {
                  py::dict g;
                  g["S"] = py::cast(&S, py::return_value_policy::reference);
                  g["woo"] = py::module_::import("woo_amd.woo");
                  g["__builtins__"] = py::module_::import("builtins");
                  py::exec(keep->cast<std::string>(), g);
                });
              };
            } else {
              r.source = py::str(py::repr(cmd)).cast<std::string>();
              r.command = [keep](Scene& S) {
                py::gil_scoped_acquire gil;
                in_python(S, [&] { (*keep)(); });
              };
            }
          });

  // ---------------------------------------------------------------- Scene
  py::class_<Scene, Object, std::shared_ptr<Scene>>(m, "Scene", py::dynamic_attr())
      .def(py::init(&make_kw<Scene>))
      .def_readwrite("fields", &Scene::fields)
      .def_readwrite("engines", &Scene::engines)
      .def_readwrite("dt", &Scene::dt)
      .def_readwrite("nextDt", &Scene::nextDt)
      .def_readwrite("dtSafety", &Scene::dtSafety)
      .def_readwrite("step", &Scene::step)
      .def_readwrite("time", &Scene::time)
      .def_readwrite("periodic", &Scene::periodic)
      .def_readwrite("cell", &Scene::cell)
      .def_readwrite("device", &Scene::device)
      .def_property_readonly("dem", [](Scene& S) {
        auto f = S.dem();
        if (!f) throw py::attribute_error("Scene has no DemField");
        return f;
      })
      .def_property_readonly("lab", &labels)
      .def_readwrite("trackEnergy", &Scene::trackEnergy)
      .def_property_readonly("energy", [](std::shared_ptr<Scene> S) { return EnergyRef{S}; })
      .def(
          "run",
          [](Scene& S, int64_t steps, bool wait) {
            if (wait) S.run(steps);
            else S.runAsync(steps);
          },
          py::arg("steps"), py::arg("wait") = true)
      .def("one", [](py::object self) { self.attr("run")(1); })
      .def("wait", [](Scene& S) {
        py::gil_scoped_release nogil;
        S.wait();
      })
      .def_property_readonly("running", &Scene::running)
      .def("stop", &Scene::stop)
      .def("saveTmp", &Scene::saveTmp, py::arg("slot") = "", py::arg("quiet") = false)
      .def_static("loadTmp", &Scene::loadTmp, py::arg("slot") = "");

  py::class_<EnergyRef>(m, "EnergyTracker")
      .def("__getitem__",
           [](EnergyRef& r, const std::string& name) {
             for (const auto& it : en_items(*r.S))
               if (it.first == name) return it.second;
             throw py::key_error("Unknown energy name '" + name + "'.");
           })
      .def("__setitem__",
           [](EnergyRef& r, const std::string& name, double v) {
             const int k = en_index(name);
             if (k < 0) throw py::key_error("energy term '" + name + "' is not one this engine tracks");
             auto e = r.S->energy();
             e[k] = v;
             r.S->energySet(e);
           })
      .def("__contains__",
           [](EnergyRef& r, const std::string& name) {
             for (const auto& it : en_items(*r.S))
               if (it.first == name) return true;
             return false;
           })
      .def("__len__", [](EnergyRef& r) { return en_items(*r.S).size(); })
      .def("__iter__",
           [](EnergyRef& r) {
             py::list l;
             for (const auto& it : en_items(*r.S)) l.append(it.first);
             return py::iter(l);
           })
      .def("keys",
           [](EnergyRef& r) {
             py::list l;
             for (const auto& it : en_items(*r.S)) l.append(it.first);
             return l;
           })
      .def("items",
           [](EnergyRef& r) {
             py::list l;
             for (const auto& it : en_items(*r.S)) l.append(py::make_tuple(it.first, it.second));
             return l;
           })
      .def("total",
           [](EnergyRef& r) {
             double t = 0;
             for (const auto& it : en_items(*r.S)) t += it.second;
             return t;
           })
      .def("relErr",
           [](EnergyRef& r) {
             double a = 0, t = 0;
             for (const auto& it : en_items(*r.S)) { a += std::abs(it.second); t += it.second; }
             return a > 0 ? t / a : 0.;
           })
      .def("clear", [](EnergyRef& r) { r.S->energyClear(); });
}
