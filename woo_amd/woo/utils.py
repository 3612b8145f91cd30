"""woo.utils: particle constructors and the default engine list (py/utils.py:77-290)."""
import math

from woo_amd._woo import (Bo1_Facet_Aabb, Bo1_Sphere_Aabb, Bo1_Wall_Aabb, Cg2_Facet_Sphere_L6Geom,
                          Cg2_Sphere_Sphere_L6Geom, Cg2_Wall_Sphere_L6Geom, ContactLoop, Cp2_FrictMat_FrictPhys,
                          DemField, DynDt, Facet, FrictMat, InsertionSortCollider, Law2_L6Geom_FrictPhys_IdealElPl,
                          Leapfrog, Material, Node, Particle, Sphere, Wall)


def spherePWaveDt(radius, density, young):
    """P-wave critical timestep of one sphere (utils.py:77-85)."""
    return radius / math.sqrt(young / density)


def pWaveDt(S, noClumps=True):
    """Smallest sphere P-wave timestep of the scene's DemField (DemFuncs::pWaveDt)."""
    return S.dem.critDt()


def defaultMaterial():
    """FrictMat(density=1e3, young=1e7, ktDivKn=.2, tanPhi=tan(.5)) (utils.py:87-92)."""
    return FrictMat(density=1e3, young=1e7, ktDivKn=.2, tanPhi=math.tan(.5))


def _material(mat):
    if isinstance(mat, Material):
        return mat
    if callable(mat):
        return mat()
    raise TypeError("The 'mat' argument must be Material instance, or a callable returning Material.")


def _sphere_mass_inertia(radius, density):
    # Sphere::lumpMassInertia (Sphere.cpp:26-33): m = (4/3)pi r^3 rho, I = (2/5) m r^2
    m = (4 / 3.) * math.pi * (radius * radius * radius) * density
    i = 0. + 1. * (2. / 5.) * m * (radius * radius)
    return m, (i, i, i)


def sphere(center, radius, mat=defaultMaterial, fixed=False, wire=False, color=None, highlight=False,
           mask=DemField.defaultMovableMask, vel=None):
    """Sphere particle with mass and inertia from its material (utils.py:153-200)."""
    p = Particle()
    p.shape = Sphere(radius=radius)
    p.mat = _material(mat)
    n = center if isinstance(center, Node) else Node(pos=center)
    p.shape.nodes = [n]
    m, inertia = _sphere_mass_inertia(radius, p.mat.density)
    n.dem.mass = m
    n.dem.inertia = inertia
    n.dem.blocked = "xyzXYZ" if fixed else ""
    if vel:
        n.dem.vel = vel
    p.mask = mask
    return p


def wall(position, axis, sense=0, glAB=None, fixed=True, mass=0, color=None, mat=defaultMaterial, visible=True,
         mask=DemField.defaultBoundaryMask):
    """Axis-aligned infinite wall (utils.py:245-268)."""
    if not fixed and mass <= 0:
        raise ValueError("Non-fixed wall must have positive mass")
    p = Particle()
    p.shape = Wall(sense=sense, axis=axis)
    p.mat = _material(mat)
    if isinstance(position, (int, float)):
        pos = [0., 0., 0.]
        pos[axis] = position
        n = Node(pos=pos)
    elif isinstance(position, Node):
        n = position
    else:
        n = Node(pos=position)
    p.shape.nodes = [n]
    n.dem.mass = 0.
    n.dem.inertia = (0., 0., 0.)
    n.dem.blocked = "xyzXYZ" if fixed else ""
    n.dem.mass = mass
    p.mask = mask
    return p


def wallBox(box, which=(1, 1, 1, 1, 1, 1), **kw):
    """Six walls delimiting an axis-aligned box ((min), (max)) (utils.py:270-290)."""
    ret = []
    lo, hi = box
    for sense, ix, coord in ((1, 0, lo), (-1, 1, hi)):
        for ax in (0, 1, 2):
            if not which[3 * ix + ax]:
                continue
            ret.append(wall(coord[ax], sense=sense, axis=ax, **kw))
    return ret


def defaultEngines(damping=0., gravity=None, verletDist=-.05, kinSplit=False, dontCollect=False, noSlip=False,
                   noBreak=False, cp2=None, law=None, model=None, collider='sap', grid=None, dynDtPeriod=100, cpKw={},
                   lawKw={}):
    """[Leapfrog, InsertionSortCollider, ContactLoop] + [DynDt] (utils.py:94-129), with the
    functors the device engine supports."""
    if gravity:
        raise ValueError("gravity MUST NOT be specified anymore, set DemField.gravity=... instead.")
    if model is not None or collider != 'sap':
        raise NotImplementedError("only the default contact model and the 'sap' collider are supported")
    cp2 = [cp2 if cp2 else Cp2_FrictMat_FrictPhys()]
    law = [law if law else Law2_L6Geom_FrictPhys_IdealElPl(noSlip=noSlip, noBreak=noBreak, label='contactLaw')]
    for k, v in cpKw.items():
        setattr(cp2[0], k, v)
    for k, v in lawKw.items():
        setattr(law[0], k, v)
    return [
        Leapfrog(damping=damping, reset=True, label='leapfrog'),
        InsertionSortCollider([Bo1_Sphere_Aabb(), Bo1_Facet_Aabb(), Bo1_Wall_Aabb()], label='collider',
                              verletDist=verletDist),
        ContactLoop([Cg2_Sphere_Sphere_L6Geom(), Cg2_Facet_Sphere_L6Geom(), Cg2_Wall_Sphere_L6Geom()], cp2, law,
                    applyForces=True, label='contactLoop'),
    ] + ([DynDt(stepPeriod=dynDtPeriod, label='dynDt')] if dynDtPeriod > 0 else [])


def facet(vertices, fakeVel=None, halfThick=0., fixed=True, wire=True, color=None, highlight=False,
          mat=defaultMaterial, visible=True, mask=DemField.defaultBoundaryMask):
    """Triangle particle from 3 vertices (utils.py:318-344). The device engine supports fixed
    facets with static nodes; a moving surface is modelled by ``fakeVel``."""
    if not fixed:
        raise NotImplementedError("only fixed facets are supported (model moving surfaces with fakeVel)")
    p = Particle()
    nodes = [v if isinstance(v, Node) else Node(pos=v) for v in vertices]
    if len(nodes) != 3:
        raise ValueError("a facet needs 3 vertices")
    p.shape = Facet(halfThick=halfThick)
    if fakeVel is not None:
        p.shape.fakeVel = fakeVel
    p.mat = _material(mat)
    p.shape.nodes = nodes
    for n in nodes:
        n.dem.mass = 0.
        n.dem.inertia = (0., 0., 0.)
        n.dem.blocked = "xyzXYZ"
    p.mask = mask
    return p


def coordNumber(dem=None, skipFree=False, S=None):
    """Average coordination number (DemFuncs::coordNumber, Funcs.cpp:84-102)."""
    f = dem if dem is not None else S.dem
    big = 1e300
    ob = f.observables((-big,) * 3, (big,) * 3)
    return ob[1] if skipFree else ob[0]


def porosity(dem, box):
    """Solid fraction of spheres whose centres lie in box ((min), (max)) (Funcs.cpp:104-114;
    the reference's ``porosity`` returns Vs/V)."""
    return dem.observables(box[0], box[1])[2]


def stress(S, volume=-1.):
    """Cauchy stress tensor from the contacts (DemFuncs::stressStiffness, stress part);
    volume <= 0 takes the periodic cell's volume."""
    return S.dem.stress(volume)


def kineticEnergy(S):
    """Kinetic energy of the integrated nodes (DemData::getEk_any, spherical inertia)."""
    return S.dem.kineticEnergy()


# DemField.minimalEngines is defaultEngines (py/_monkey/aliases.py:102)
DemField.minimalEngines = staticmethod(defaultEngines)
