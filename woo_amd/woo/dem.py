"""woo.dem: DemField, particles, materials, engines and functors of the DEM step
(src/dem/Particle.hpp, FrictMat.hpp, Hertz.hpp, Collision.hpp, ContactLoop.hpp,
Leapfrog.hpp, DynDt.hpp, Facet.hpp), plus ``Sphere.make`` / ``Wall.make`` / ``Facet.make``."""
from woo_amd._woo import (  # noqa: F401
    Bo1_Facet_Aabb, Bo1_Sphere_Aabb, Bo1_Wall_Aabb, BoundFunctor, Cg2_Facet_Sphere_L6Geom, Cg2_Sphere_Sphere_L6Geom,
    Cg2_Wall_Sphere_L6Geom, CGeomFunctor, Facet,
    Contact, ContactContainer, ContactLoop, Cp2_FrictMat_FrictPhys, Cp2_HertzMat_HertzPhys, CPhysFunctor, DemData,
    DemField, DynDt, ElastMat, ForceResetter, FrictMat, FrictPhys, Functor, HertzMat, In2_Facet, InsertionSortCollider,
    IntraForce, IntraFunctor, L6Geom,
    Law2_L6Geom_FrictPhys_IdealElPl, Law2_L6Geom_HertzPhys_DMT, LawFunctor, Leapfrog, Material, Particle,
    ParticleContainer, Shape, Sphere, Wall)


def _make_sphere(center, radius, **kw):
    from . import utils

    return utils.sphere(center, radius, **kw)


def _make_wall(position, axis, **kw):
    from . import utils

    return utils.wall(position, axis, **kw)


def _make_facet(vertices, **kw):
    from . import utils

    return utils.facet(vertices, **kw)


Sphere.make = staticmethod(_make_sphere)
Wall.make = staticmethod(_make_wall)
Facet.make = staticmethod(_make_facet)
