"""Inputs of the learned coarse step, on the CPU: the contact-force inputs are proportional to
the Hertz force per unit mass of the receiving sphere that the DEM engine applies
(Cp2_HertzMat_HertzPhys + Law2_L6Geom_HertzPhys_DMT, src/dem/Hertz.cpp:34-77,85-210), with one
constant for every pair of radii and every overlap, so the network only has to learn that
constant; and the host-built graph has the structure the device builder gives."""
import math

import numpy as np

from woo_amd.surrogate import (EN, POISSON, RHO, YOUNG, contact_graph, edge_inputs, node_inputs)


def hertz_acc(ri, rj, delta, vN):
    """elastic and viscous normal force of the Hertz/DMT law divided by the mass of sphere i"""
    K = (4 / 3.) / ((1 - POISSON ** 2) / YOUNG + (1 - POISSON ** 2) / YOUNG)
    R = 1. / (1. / ri + 1. / rj)
    kn0 = K * math.sqrt(R)
    mi, mj = (4 / 3.) * math.pi * ri ** 3 * RHO, (4 / 3.) * math.pi * rj ** 3 * RHO
    mbar = mi * mj / (mi + mj)
    lnE = math.log(EN)
    visc = -math.sqrt(5.) * lnE / math.sqrt(lnE * lnE + math.pi * math.pi)
    eta = visc * math.sqrt(mbar * kn0) * delta ** .25
    return kn0 * delta ** 1.5 / mi, eta * vN / mi


def test_edge_inputs_scale_as_the_hertz_law():
    rng = np.random.default_rng(1)
    ratios_e, ratios_v = [], []
    for _ in range(200):
        ri, rj = rng.uniform(2e-3, 3e-3, 2)
        frac = rng.uniform(2e-4, 2e-2)  # overlap / (ri + rj)
        d = (ri + rj) * (1 - frac)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        pi, pj = np.zeros((1, 3)), (d * n)[None, :]
        vN = rng.uniform(-1, 1)
        vi, vj = np.zeros((1, 3)), (vN * n)[None, :]
        f = edge_inputs(pi, pj, vi, vj, np.array([ri]), np.array([rj])).astype(np.float64)[0]
        ae, av = hertz_acc(ri, rj, (ri + rj) - d, vN)
        # elastic input: -n times a magnitude proportional to the elastic force per mass
        fe = f[9:12]
        assert np.allclose(fe / np.linalg.norm(fe), -n, atol=1e-6)
        ratios_e.append(ae / np.linalg.norm(fe))
        fv = f[12:15]
        assert np.allclose(fv, n * (fv @ n), atol=1e-6 * np.linalg.norm(fv) + 1e-30)
        ratios_v.append(av / (fv @ n))
    for r in (np.array(ratios_e), np.array(ratios_v)):
        assert np.all(r > 0)
        assert np.max(r) / np.min(r) - 1 < 2e-5, (np.min(r), np.max(r))


def test_floor_inputs_scale_as_the_wall_hertz_law():
    # a wall contact has lens (r, r), so R = r / 2 (Hertz.cpp:50 with L6Geom.cpp:61-67)
    K = (4 / 3.) / ((1 - POISSON ** 2) / YOUNG * 2)
    rng = np.random.default_rng(2)
    re, rv = [], []
    for _ in range(100):
        r = rng.uniform(2e-3, 3e-3)
        delta = r * rng.uniform(1e-4, 1e-2)
        vz = rng.uniform(-1, 1)
        x = node_inputs(np.array([[0., 0., vz]]), np.array([r]), np.array([r - delta])).astype(np.float64)[0]
        m = (4 / 3.) * math.pi * r ** 3 * RHO
        kn0 = K * math.sqrt(r / 2)
        re.append(kn0 * delta ** 1.5 / m / x[5])
        rv.append(math.sqrt(m * kn0) * delta ** .25 * vz / m / x[6])
    for r in (np.array(re), np.array(rv)):
        assert np.all(r > 0) and np.max(r) / np.min(r) - 1 < 2e-5


def test_no_contact_inputs_vanish_apart():
    # spheres apart: the overlap is negative and the force inputs are zero
    f = edge_inputs(np.zeros((1, 3)), np.array([[6e-3, 0, 0]]), np.zeros((1, 3)), np.array([[1., 0, 0]]),
                    np.array([2.5e-3]), np.array([2.5e-3]))[0]
    assert f[3] < 0 and np.all(f[9:15] == 0)
    x = node_inputs(np.array([[0, 0, -1.]]), np.array([2.5e-3]), np.array([3e-3]))[0]
    assert x[4] < 0 and x[5] == 0 and x[6] == 0


def test_host_graph_structure():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 0.03, size=(300, 3))
    rad = rng.uniform(2e-3, 3e-3, size=300)
    g = contact_graph(pos, np.zeros((300, 3)), rad)
    recv = np.repeat(np.arange(g.N), np.diff(g.rowOff))
    # every edge once per direction, the opposite edge found by rev, cutoff 1.1 (ri + rj)
    assert np.array_equal(recv[g.rev], g.send) and np.array_equal(g.send[g.rev], recv)
    d = np.linalg.norm(pos[g.send] - pos[recv], axis=1)
    assert np.all(d < 1.1 * (rad[g.send] + rad[recv]))
    i, j = np.triu_indices(300, 1)
    dd = np.linalg.norm(pos[i] - pos[j], axis=1)
    assert 2 * np.sum(dd < 1.1 * (rad[i] + rad[j])) == g.E
    assert g.x.shape == (300, 7) and g.ef.shape == (g.E, 15)
