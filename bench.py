#!/usr/bin/env python3
"""DEM step benchmark on one MI355X: 1,000,000 polydisperse Hertz spheres.

Workload: radius ~ U[2,4] mm, periodic cube filled to solid fraction 0.55,
velocities N(0, 0.1 m/s), HertzMat(young=1e8, tanPhi=0.4, density=3000),
Cp2_HertzMat_HertzPhys(poisson=0.25, en=0.6), Law2_L6Geom_HertzPhys_DMT, no gravity,
dt = 0.3 pWaveDt, engines [InsertionSortCollider, ContactLoop, Leapfrog(damping=0)],
plain FP64. One step = one pass of those engines over all particles.

Prints one JSON line: whole-job particle-steps/s over K timed steps (inputs resident
in HBM), with the broadphase refresh counted at its measured cost and rate; the roofline
of the dominant kernel (HIP-event timed on the engine's stream, at a fixed state that the
committed rocprofv3 counter summary was taken at) and of the whole step (BASELINE.md's
algorithmic bytes per particle-step); and the CPU oracle timed on a bounded sample of the
same workload.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "particle-steps/s, 1M Hertz spheres on 1 MI355X; % HBM roofline; speedup vs FP64"
RHO, YOUNG, TANPHI, POISSON, EN = 3000., 1e8, .4, .25, .6
ENGINES = ["Collider", "ContactLoop", "Leapfrog"]
DRUM_ENGINES = ["Collider", "ContactLoop", "IntraForce", "Leapfrog"]
# steps wdg_tune takes: one plain pair, then for 3 force-pass widths x graph lengths 4, 8, 16 a
# capturing replay and 3 timed ones (2 + 3 * 4 * (4 + 8 + 16)); the run after it is the same as
# after this many plain steps (graph replay is bit-identical to the plain loop)
TUNE_STEPS = 338
# the state the kernels are timed at, and the committed counter profile was taken at: this many
# plain steps from the initial state, then KERNEL_LAUNCHES event-timed steps
FIX_STEPS = 400
KERNEL_LAUNCHES = 20
PMC_PROFILE = os.path.join("profiles", "pmc_traffic_1M_hertz_fp64.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an arbitrary stream (the engine's own)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]

    def new(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        self.L.hipEventSynchronize(b)
        t = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(t), a, b) == 0
        return t.value


def make_workload(n, dev, seed=2):
    from woo_amd.pack import periodic_pack
    from woo_amd._native import SPHERE

    t0 = time.perf_counter()
    pos, radius, L, info = periodic_pack(n, 2e-3, 4e-3, 0.55, seed=seed, dev=dev)
    log(f"packing: n={n} L={L:.4f} m relax iterations={info['iterations']} max overlap={info['maxOverlap']:.3e} "
        f"({time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(3)
    vel = rng.normal(0., 0.1, size=(n, 3))
    mass = (4 / 3.) * math.pi * (radius * radius * radius) * RHO
    inertia = np.repeat((0. + 1. * (2. / 5.) * mass * (radius * radius))[:, None], 3, 1)
    dt = 0.3 * float(np.min(radius)) / math.sqrt(YOUNG / RHO)
    mat = dict(kind=1, density=RHO, young=YOUNG, ktDivKn=.2, tanPhi=TANPHI, surfEnergy=0., alpha=0.)
    dev.set_field(periodic=True, cell=(L, L, L))
    dev.set_materials([mat])
    dev.upload_particles(shape=np.full(n, SPHERE, np.int32), pos=pos, radius=radius, mass=mass, inertia=inertia,
                         vel=vel)
    return dict(pos=pos, radius=radius, vel=vel, mass=mass, inertia=inertia, L=L, dt=dt, mat=mat)


def contact_params():
    from woo_amd._native import PHYS_HERTZ, LAW_HERTZDMT

    return dict(physKind=PHYS_HERTZ, lawKind=LAW_HERTZDMT, poisson=POISSON, en=EN)


def algorithmic_bytes(n, slots, real):
    """Compulsory HBM bytes per launch (FP64, periodic workload), per engine.

    contact loop (k_cfilter + k_cforce): per slot lo/hi index, info word and packed
      cellDist (20 B); per particle pos, vel, angVel, radius, mass, material (92 B); per
      real contact its slot index (4 B written, 4 B read), previous state read (trsf columns
      0-1, tangential force, contact point: 88 B), new state written (trsf columns 0-1,
      local force, uN: 80 B) and the 64 B gather record written.
    k_leapfrog (force gather fused): per particle flags, adjacency offset, pos/vel/angVel/ori
      read and written, mass, inertia, refresh position (272 B; force/torque are not
      touched while they hold the reset value); per slot two adjacency entries (8 B); per
      real contact its 64 B gather record.
    """
    return {
        "contact_loop": slots * 20 + n * 92 + real * (4 + 4 + 88 + 80 + 64),
        "k_leapfrog": n * 272 + slots * 8 + real * 64,
    }


def fixed_state(wl):
    """A device at the kernel-timing state: the workload's initial state, FIX_STEPS plain steps."""
    from woo_amd._native import Device, SPHERE

    d = Device(0)
    n = len(wl["radius"])
    d.set_field(periodic=True, cell=(wl["L"],) * 3)
    d.set_materials([wl["mat"]])
    d.upload_particles(shape=np.full(n, SPHERE, np.int32), pos=wl["pos"], radius=wl["radius"], mass=wl["mass"],
                       inertia=wl["inertia"], vel=wl["vel"])
    step, _ = d.run(ENGINES, FIX_STEPS, 0, wl["dt"], contact=contact_params(), leapfrog=dict(damping=0.),
                    dynDtPeriod=0)
    d.sync()
    return d, step


def time_kernels(dev, wl, step, k):
    """Average device time of the contact loop (k_cfilter + k_cforce) and of the leapfrog (gather
    fused) over k steps, HIP events on the engine's stream; the state (slots, real contacts)
    after them."""
    ev = HipEvents()
    s = dev.stream()
    cp = contact_params()
    dt = wl["dt"]
    e = [ev.new() for _ in range(4)]
    tc = tl = 0.
    for i in range(k):
        dev.collide(-.05)
        ev.record(e[0], s)
        dev.contact_loop(step, dt, **cp)
        ev.record(e[1], s)
        ev.record(e[2], s)
        dev.leapfrog(dt, 0., True)
        ev.record(e[3], s)
        tc += ev.elapsed_ms(e[0], e[1])
        tl += ev.elapsed_ms(e[2], e[3])
        step += 1
    return tc / k * 1e-3, tl / k * 1e-3, step


def refresh_cost(dev, step, dt, k=3):
    """Average device time of a broadphase refresh (a forced collider run: cell keys, sort,
    permutation, pair pass, history merge, adjacency), HIP events on the engine's stream; one
    step between refreshes."""
    ev = HipEvents()
    s = dev.stream()
    cp = contact_params()
    a, b = ev.new(), ev.new()
    t = 0.
    for _ in range(k):
        dev.sync()
        ev.record(a, s)
        dev.collide(-.05, forceInit=True)
        ev.record(b, s)
        t += ev.elapsed_ms(a, b)
        dev.contact_loop(step, dt, **cp)
        dev.leapfrog(dt, 0., True)
        step += 1
    return t / k * 1e-3, step


def whole_step_bytes(n, slots, real, k_refresh):
    """BASELINE.md's algorithmic bytes per particle-step (FP64): 272 integrate + 80 contact-kernel
    particle read + 48 gather write + (C/N) * 280 per contact + (104 + 8 P/N) / k refresh, with the
    run's own contacts per particle C/N, candidate pairs per particle P/N and refresh interval k."""
    cn, pn = real / n, slots / n
    return 272 + 80 + 48 + cn * (8 + 176 + 96) + (104 + 8 * pn) / k_refresh, cn, pn


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, budget_s=5.):
    """The OpenMP CPU oracle (a restatement of Woo's algorithm: chunked insertion-sort collider,
    guided contact loop with per-node mutex force accumulation, guided leapfrog), compiled here
    with -march=native, on the same workload: all cores the process may use, then 1 thread; each
    the median of 3 samples of whole steps (about budget_s each), after one untimed first step
    (initial sort and contact creation, excluded as in Woo's own timings)."""
    from oracle import Oracle, native_library

    lib = native_library()
    cores = len(os.sched_getaffinity(0))
    n = len(wl["radius"])
    o = Oracle(lib=lib, dt=wl["dt"], gravity=(0, 0, 0), periodic=1, cell=(wl["L"],) * 3, physKind=1, lawKind=1,
               damping=0., poisson=POISSON, en=EN, deterministic=0, nThreads=cores)
    o.material(**wl["mat"])
    o.spheres(wl["pos"], wl["radius"], 0)
    o.set("vel", wl["vel"])
    o.set("mass", wl["mass"])
    o.set("inertia", wl["inertia"])
    t0 = time.perf_counter()
    o.run(ENGINES, 1)
    t_first = time.perf_counter() - t0
    out = {}
    for threads in (cores, 1):
        o.set_params(nThreads=threads)
        t0 = time.perf_counter()
        o.run(ENGINES, 1)
        per = time.perf_counter() - t0
        steps = max(1, int(budget_s / per))
        rates = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.run(ENGINES, steps)
            rates.append(n * steps / (time.perf_counter() - t0))
        out[threads] = dict(value=float(np.median(rates)), samples=rates, steps_per_sample=steps)
        log(f"cpu oracle, {threads} threads: {out[threads]}")
    return dict(value=out[cores]["value"], unit="particle-steps/s", cores=cores, kind="port",
                cpu=cpu_model(), os_cpu_count=os.cpu_count(), one_thread=out[1]["value"],
                samples={"all_cores": out[cores]["samples"], "one_thread": out[1]["samples"]},
                sample=(f"the same {n}-sphere workload after one untimed first step ({t_first:.1f} s); median of 3 "
                        f"samples of {out[cores]['steps_per_sample']} steps on {cores} threads (all the process may "
                        f"use) and of {out[1]['steps_per_sample']} steps on 1 thread; OpenMP oracle built with "
                        f"-march=native on this host: Woo's algorithm restated, except that its periodic collider "
                        f"rebuilds the pairs from a cell grid at each refresh instead of Woo's periodic "
                        f"insertion sort"))


def observables(dev, wl):
    """Kinetic energy, coordination number, solid fraction of the central eighth of the
    cell, and the stress tensor."""
    L = wl["L"]
    ob = dev.observables((.25 * L,) * 3, (.75 * L,) * 3)
    return dict(kinetic_energy=dev.kinetic_energy(), coord_number=ob["coordNumber"],
                solid_fraction=ob["solidFraction"], stress=dev.stress(-1.))


def fp32_drift(wl, steps=5000, every=1000):
    """The FP32 contact form against FP64 over the 1M workload's 5,000 steps (its configured
    length): both from the initial state, graph-replayed side by side, each observable's
    relative error at every `every` steps."""
    from woo_amd._native import Device, SPHERE

    n = len(wl["radius"])
    cp = contact_params()
    devs = {}
    for prec in (64, 32):
        d = Device(0)
        d.set_field(periodic=True, cell=(wl["L"],) * 3)
        d.set_materials([wl["mat"]])
        d.upload_particles(shape=np.full(n, SPHERE, np.int32), pos=wl["pos"], radius=wl["radius"], mass=wl["mass"],
                           inertia=wl["inertia"], vel=wl["vel"])
        d.set_precision(prec)
        devs[prec] = [d, 0]
    out = []
    for k in range(steps // every):
        obs = {}
        for prec, dv in devs.items():
            dv[1], _ = dv[0].run(ENGINES, every, dv[1], wl["dt"], graph=True, contact=cp, leapfrog=dict(damping=0.),
                                 dynDtPeriod=0)
            obs[prec] = observables(dv[0], wl)
        a, b = obs[64], obs[32]
        row = dict(step=devs[64][1])
        for q in ("kinetic_energy", "coord_number", "solid_fraction"):
            row[q] = abs(b[q] - a[q]) / abs(a[q])
        p64, p32 = np.trace(a["stress"]) / 3, np.trace(b["stress"]) / 3
        row["pressure"] = abs(p32 - p64) / abs(p64)
        out.append(row)
    for dv in devs.values():
        dv[0].close()
    worst = {q: max(r[q] for r in out) for q in ("kinetic_energy", "coord_number", "solid_fraction", "pressure")}
    return dict(steps=steps, rel_err_by_step=out, worst_rel_err=worst, bar=0.01,
                within_bar=bool(all(v <= 0.01 for v in worst.values())))


def fp32_form(wl, args, value64, obs64, tuning):
    """The FP32 contact form on the same initial state and step count: throughput and its
    speedup over FP64, and each observable's relative error against the FP64 run (which
    is bit-identical to the FP64 CPU oracle)."""
    from woo_amd._native import Device, SPHERE

    d = Device(0)
    n = len(wl["radius"])
    d.set_field(periodic=True, cell=(wl["L"],) * 3)
    d.set_materials([wl["mat"]])
    d.upload_particles(shape=np.full(n, SPHERE, np.int32), pos=wl["pos"], radius=wl["radius"], mass=wl["mass"],
                       inertia=wl["inertia"], vel=wl["vel"])
    d.set_precision(32)
    cp = contact_params()
    step, _ = d.run(ENGINES, args.warmup, 0, wl["dt"], contact=cp, leapfrog=dict(damping=0.), dynDtPeriod=0)
    d.sync()
    if tuning:
        # the FP64 run's tuner steps, so that both runs compare at the same step, then its tuning
        step, _ = d.run(ENGINES, tuning["tuning_steps"], step, wl["dt"], contact=cp, leapfrog=dict(damping=0.),
                        dynDtPeriod=0)
        d.set_tuning(cfBlocks=tuning["cfBlocks"], graphSteps=tuning["graphSteps"])
        d.sync()
    t0 = time.perf_counter()
    step, _ = d.run(ENGINES, args.steps, step, wl["dt"], graph=bool(args.graph), contact=cp,
                    leapfrog=dict(damping=0.), dynDtPeriod=0)
    d.sync()
    el = time.perf_counter() - t0
    v32 = n * args.steps / el
    o32 = observables(d, wl)
    d.close()
    err = {}
    for k in ("kinetic_energy", "coord_number", "solid_fraction"):
        err[k] = dict(fp64=obs64[k], fp32=o32[k], rel_err=abs(o32[k] - obs64[k]) / abs(obs64[k]))
    p64, p32 = np.trace(obs64["stress"]) / 3, np.trace(o32["stress"]) / 3
    err["pressure"] = dict(fp64=p64, fp32=p32, rel_err=abs(p32 - p64) / abs(p64))
    err["stress_max_abs_diff_over_max"] = float(np.max(np.abs(o32["stress"] - obs64["stress"])) /
                                                np.max(np.abs(obs64["stress"])))
    drift = fp32_drift(wl) if args.fp32_drift else None
    return dict(value=v32, unit="particle-steps/s", ms_per_step=el / args.steps * 1e3,
                speedup_vs_fp64=v32 / value64, steps=args.steps, warmup=args.warmup,
                dtype="f32 contact geometry/law, f64 positions and sums",
                observable_errors_vs_fp64=err, drift_vs_fp64=drift)


def measured_traffic(kernel, slots, real):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by what it reads for a copy of known size), when that summary was taken
    at the same state (contact slots and real contacts) as the launches timed here; else None."""
    path = os.path.join(ROOT, PMC_PROFILE)
    state, traffic = None, None
    try:
        for line in open(path):
            if line.startswith("state:"):
                kv = dict(x.split("=") for x in line.split(":", 1)[1].split())
                state = (int(kv["slots"]), int(kv["real"]))
            if line.startswith(f"{kernel} traffic_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
    except (OSError, ValueError, KeyError):
        return None, None
    if traffic is None or state != (slots, real):
        return None, None
    return traffic, PMC_PROFILE


def drum_workload(n, dev, seed=4):
    """Rotating drum: 1024 facets (512 circumferential strips of 2 triangles, R=0.5 m, length
    0.4 m, axis x) whose rotation at 2 rad/s is given by Facet.fakeVel (triangulated.cylinder
    with angVel, realRot=False), end walls at x=+-0.2 m, n spheres with radius
    lognormal(median 1.5 mm, sigma 0.3) clipped to [0.8, 4] mm packed into the lower part,
    velocities N(0, 0.1 m/s), gravity -z, Hertz as the periodic workload."""
    from woo_amd.pack import drum_fill
    from woo_amd._native import FACET, SPHERE, WALL

    R, length, div, omega = 0.5, 0.4, 512, 2.0
    Rin = R * math.cos(math.pi / div)
    t0 = time.perf_counter()
    pos, radius, info = drum_fill(n, Rin, length, 0.8e-3, 4e-3, 1.5e-3, 0.3, seed=seed, dev=dev)
    log(f"drum packing: n={n} fill={info['fill']:.3f} box spheres={info['boxSpheres']} kept={info['kept']} "
        f"({time.perf_counter() - t0:.1f} s)")
    ax = np.array([1., 0., 0.])
    th = np.linspace(0, 2 * math.pi, div + 1)
    verts, fake = [], []
    for k in range(div):
        a0 = (-length / 2, R * math.cos(th[k]), R * math.sin(th[k]))
        a1 = (-length / 2, R * math.cos(th[k + 1]), R * math.sin(th[k + 1]))
        b0, b1 = (length / 2,) + a0[1:], (length / 2,) + a1[1:]
        for tri in ((a0, b0, b1), (a0, b1, a1)):
            v = np.array(tri)
            nrm = np.cross(v[1] - v[0], v[2] - v[0])
            nrm /= np.linalg.norm(nrm)
            verts.append(v.reshape(9))
            fake.append(-R * omega * np.cross(nrm, ax))
    nf, nw = len(verts), 2
    N = nw + nf + n
    shape = np.concatenate([np.full(nw, WALL), np.full(nf, FACET), np.full(n, SPHERE)]).astype(np.int32)
    P = np.zeros((N, 3))
    P[0, 0], P[1, 0] = -length / 2, length / 2
    P[nw:nw + nf] = np.array(verts)[:, :3]
    P[nw + nf:] = pos
    fv = np.zeros((N, 9))
    fv[nw:nw + nf] = np.array(verts)
    fk = np.zeros((N, 3))
    fk[nw:nw + nf] = np.array(fake)
    rad = np.zeros(N)
    rad[nw + nf:] = radius
    mass = np.zeros(N)
    mass[nw + nf:] = (4 / 3.) * math.pi * radius ** 3 * RHO
    inertia = np.repeat((0. + 1. * (2. / 5.) * mass * (rad * rad))[:, None], 3, 1)
    rng = np.random.default_rng(5)
    vel = np.zeros((N, 3))
    vel[nw + nf:] = rng.normal(0., 0.1, size=(n, 3))
    flags = np.zeros(N, np.uint32)
    flags[:nw + nf] = 63
    integ = np.ones(N, np.uint8)
    integ[:nw + nf] = 0
    mask = np.full(N, 5, np.uint32)
    mask[:nw + nf] = 3
    axis = np.zeros(N, np.int32)
    sense = np.zeros(N, np.int32)
    sense[0], sense[1] = 1, -1
    dt = 0.3 * float(np.min(radius)) / math.sqrt(YOUNG / RHO)
    mat = dict(kind=1, density=RHO, young=YOUNG, ktDivKn=.2, tanPhi=TANPHI, surfEnergy=0., alpha=0.)
    arrays = dict(shape=shape, pos=P, radius=rad, wallAxis=axis, wallSense=sense, mask=mask, flags=flags,
                  integrated=integ, vel=vel, mass=mass, inertia=inertia, facetVerts=fv, fakeVel=fk)
    return dict(arrays=arrays, mat=mat, dt=dt, n=n, facets=nf, fill=info["fill"], verts=np.array(verts).reshape(-1, 3, 3),
                facet_ids=np.arange(nw, nw + nf), omega=-omega)


def drum_torque(dev, wl):
    """moment about the drum axis (x through the origin) of the mantle's node forces and torques
    from the last IntraForce (In2_Facet): what ReadForce summed over the facet nodes reads"""
    F, T = dev.facet_nodes(wl["facet_ids"])
    return float(np.sum(np.cross(wl["verts"], F)[..., 0]) + np.sum(T[..., 0]))


def drum_form(args):
    """The drum workload in the FP64 and FP32 contact forms: throughput of each, and the
    FP32 observables against FP64 (kinetic energy, coordination number)."""
    from woo_amd._native import Device

    dev = Device(0)
    wl = drum_workload(args.drum_n, dev)
    out = dict(workload=f"rotating drum, {wl['facets']} facets (fakeVel 2 rad/s) + 2 end walls, {args.drum_n} "
                        f"spheres r~lognormal(1.5 mm, 0.3) in [0.8, 4] mm, gravity, Hertz; engines "
                        f"{'+'.join(DRUM_ENGINES)}",
               n=args.drum_n, dt=wl["dt"], fill=wl["fill"], steps=args.drum_steps, warmup=args.warmup)
    cp = contact_params()
    for prec in (64, 32):
        dev.set_field(gravity=(0., 0., -9.81), periodic=False)
        dev.set_materials([wl["mat"]])
        dev.upload_particles(**wl["arrays"])
        dev.set_precision(prec)
        step, _ = dev.run(DRUM_ENGINES, args.warmup, 0, wl["dt"], contact=cp, leapfrog=dict(damping=0.),
                          dynDtPeriod=0)
        dev.sync()
        st0 = dev.stats()
        t0 = time.perf_counter()
        step, _ = dev.run(DRUM_ENGINES, args.drum_steps, step, wl["dt"], contact=cp, leapfrog=dict(damping=0.),
                          dynDtPeriod=0)
        dev.sync()
        el = time.perf_counter() - t0
        st1 = dev.stats()
        big = 1e300
        ob = dev.observables((-big,) * 3, (big,) * 3)
        M = drum_torque(dev, wl)
        out[f"fp{prec}"] = dict(value=args.drum_n * args.drum_steps / el, unit="particle-steps/s",
                                ms_per_step=el / args.drum_steps * 1e3, real_contacts=st1["nReal"],
                                contact_slots=st1["nCandidates"], refreshes=st1["nRefresh"] - st0["nRefresh"],
                                kinetic_energy=dev.kinetic_energy(), coord_number=ob["coordNumber"],
                                # the charge's moment on the mantle about +x at the last step, and the
                                # power the drive turning the mantle at omega (about +x) puts in
                                drum_torque_Nm=M, drum_power_W=-M * wl["omega"])
        log(f"drum fp{prec}: {out[f'fp{prec}']}")
        dev.set_precision(64)
    dev.close()
    a, b = out["fp64"], out["fp32"]
    out["fp32_speedup_vs_fp64"] = b["value"] / a["value"]
    out["fp32_rel_err"] = {k: abs(b[k] - a[k]) / abs(a[k]) for k in ("kinetic_energy", "coord_number")}
    return out


def batched_scenes(n_scenes=256, seed=5, r=5e-3, rho=3000.):
    """Parametric study of gravity depositions (the 2000-sphere wall-box scene): particle count
    log-uniform in [500, 50000], tanPhi uniform in [0.1, 0.9], E log-uniform in [1e6, 1e8] Pa,
    box edge 0.2*(N/2000)^(1/3) m, dt = 0.5 pWaveDt. Non-overlapping random placement: a random
    subset of the sites of a cubic lattice filling the box, each jittered within its cell.
    Scenes are laid out side by side (sorted by size) so that each occupies its own region."""
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(math.log(500), math.log(50000), size=n_scenes)).astype(np.int64)
    tanPhi = rng.uniform(.1, .9, size=n_scenes)
    E = np.exp(rng.uniform(math.log(1e6), math.log(1e8), size=n_scenes))
    scenes = []
    for i in range(n_scenes):
        box = 0.2 * (N[i] / 2000.) ** (1 / 3.)
        m = int(math.ceil(N[i] ** (1 / 3.)))
        a = box / m
        g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
        g = g[rng.permutation(len(g))[:N[i]]]
        pos = (g + 0.5) * a + rng.uniform(-1, 1, size=(N[i], 3)) * (0.49 * a - r)
        scenes.append(dict(n=int(N[i]), box=box, pos=pos, tanPhi=float(tanPhi[i]), young=float(E[i]),
                           dt=0.5 * r / math.sqrt(E[i] / rho)))
    # side by side: layers of rows, largest first
    order = np.argsort([-sc["box"] for sc in scenes], kind="stable")
    side = int(math.ceil(n_scenes ** (1 / 3.)))
    gap = 0.05
    x = y = z = 0.
    rowH = layH = 0.
    for k, i in enumerate(order):
        sc = scenes[i]
        if k % side == 0 and k:
            x = 0.
            y += rowH + gap
            rowH = 0.
        if k % (side * side) == 0 and k:
            y = 0.
            z += layH + gap
            layH = 0.
        sc["offset"] = np.array([x, y, z])
        x += sc["box"] + gap
        rowH = max(rowH, sc["box"])
        layH = max(layH, sc["box"])
    return scenes, r, rho


def batch_upload(dev, scenes, r, rho, idx=None):
    """scenes idx (all by default) into one context: spheres then the 6 walls of each scene"""
    from woo_amd._native import SPHERE, WALL

    idx = range(len(scenes)) if idx is None else idx
    parts = dict(shape=[], pos=[], radius=[], wallAxis=[], wallSense=[], material=[], mask=[], flags=[],
                 integrated=[], mass=[], inertia=[], scene=[])
    mats, dts = [], []
    for k, i in enumerate(idx):
        sc = scenes[i]
        n, off, b = sc["n"], sc["offset"], sc["box"]
        m = (4 / 3.) * math.pi * r ** 3 * rho
        wp = []
        for sense, coord in ((1, 0.), (-1, b)):
            for ax in (0, 1, 2):
                p = np.zeros(3)
                p[ax] = coord
                wp.append((p + off, ax, sense))
        parts["shape"].append(np.concatenate([np.full(n, SPHERE), np.full(6, WALL)]))
        parts["pos"].append(np.concatenate([sc["pos"] + off, np.array([w[0] for w in wp])]))
        parts["radius"].append(np.concatenate([np.full(n, r), np.zeros(6)]))
        parts["wallAxis"].append(np.concatenate([np.zeros(n), [w[1] for w in wp]]))
        parts["wallSense"].append(np.concatenate([np.zeros(n), [w[2] for w in wp]]))
        parts["material"].append(np.full(n + 6, k))
        parts["mask"].append(np.concatenate([np.full(n, 5), np.full(6, 3)]))
        parts["flags"].append(np.concatenate([np.zeros(n), np.full(6, 63)]))
        parts["integrated"].append(np.concatenate([np.ones(n), np.zeros(6)]))
        parts["mass"].append(np.concatenate([np.full(n, m), np.zeros(6)]))
        parts["inertia"].append(np.concatenate([np.full((n, 3), 0. + 1. * (2. / 5.) * m * (r * r)), np.zeros((6, 3))]))
        parts["scene"].append(np.full(n + 6, k))
        mats.append(dict(kind=0, density=rho, young=sc["young"], ktDivKn=.2, tanPhi=sc["tanPhi"], surfEnergy=0.,
                         alpha=0.))
        dts.append(sc["dt"])
    arr = {k: np.concatenate(v) for k, v in parts.items()}
    arr["inertia"] = arr["inertia"].reshape(-1, 3)
    dev.set_field(gravity=(0., 0., -9.81), periodic=False)
    dev.set_materials(mats)
    dev.upload_particles(**arr)
    dev.set_scenes(dts)
    return arr


def bed_fraction(pos, vel, scenes, idx, r):
    """per scene: the bed's solid fraction, the volume of the spheres whose centres lie below the
    90th percentile of the centre heights over the box cross-section times that height plus r
    (a few spheres still bouncing on top do not count); and the scene's largest sphere speed"""
    phi, vmax, o = [], [], 0
    for i in idx:
        sc = scenes[i]
        p = pos[o:o + sc["n"]] - sc["offset"]
        v = vel[o:o + sc["n"]]
        o += sc["n"] + 6
        h = float(np.percentile(p[:, 2], 90))
        k = int(np.sum(p[:, 2] <= h))
        phi.append(k * (4 / 3.) * math.pi * r ** 3 / (sc["box"] ** 2 * (h + r)))
        vmax.append(float(np.max(np.linalg.norm(v, axis=1))))
    return np.array(phi), np.array(vmax)


def batched_form(args):
    """256 scenes in one context, stepped by the graph runner after the autotuner; against the
    plain GPU engine running a sample of the scenes one at a time (same steps): throughput of
    both, and the sampled scenes' final states, which must be identical."""
    from woo_amd._native import Device

    scenes, r, rho = batched_scenes()
    ntot = sum(sc["n"] for sc in scenes)
    steps = args.batch_steps
    kw = dict(contact=dict(physKind=0, lawKind=0, trsfRenorm=100), leapfrog=dict(damping=.2), dynDtPeriod=0)
    dev = Device(0)
    batch_upload(dev, scenes, r, rho)
    t0 = time.perf_counter()
    if args.batch_graph:
        tun, step, taken = dev.tune(ENGINES, 0, scenes[0]["dt"], **kw)
    else:
        dev.set_tuning(cfBlocks=2, graphSteps=8, skinSteps=args.batch_skin)
        tun, step, taken = dev.tuning(), 0, 0
    t_tune = time.perf_counter() - t0
    st0 = dev.stats()
    dev.sync()
    t0 = time.perf_counter()
    step, _ = dev.run(ENGINES, steps, step, scenes[0]["dt"], graph=bool(args.batch_graph), **kw)
    dev.sync()
    el = time.perf_counter() - t0
    st1 = dev.stats()
    value = ntot * steps / el
    total = taken + steps
    pos, vel = dev.get("pos"), dev.get("vel")
    # the cost of energy tracking per scene (Scene.trackEnergy of every scene), 200 steps each,
    # after the timed window (the sampled scenes below compare the state at its end)
    energy = {}
    st = step
    for on in (False, True):
        dev.set_energy_tracking(on)
        st, _ = dev.run(ENGINES, 8, st, scenes[0]["dt"], graph=bool(args.batch_graph), **kw)
        dev.sync()
        t0 = time.perf_counter()
        st, _ = dev.run(ENGINES, 200, st, scenes[0]["dt"], graph=bool(args.batch_graph), **kw)
        dev.sync()
        energy["ms_per_step_" + ("tracking" if on else "plain")] = (time.perf_counter() - t0) / 200 * 1e3
    energy["overhead"] = energy["ms_per_step_tracking"] / energy["ms_per_step_plain"] - 1.
    es = dev.energy_scenes()
    energy["kinetic_per_scene_J"] = dict(min=float(es[:, 0].min()), max=float(es[:, 0].max()))
    dev.close()
    # serial sample: scenes spread over the sizes, each alone for the same steps
    order = np.argsort([sc["n"] for sc in scenes])
    sample = [int(order[k]) for k in np.linspace(0, len(scenes) - 1, args.batch_sample).astype(int)]
    t_ser, n_ser, same = 0., 0, True
    offs = np.cumsum([0] + [scenes[i]["n"] + 6 for i in range(len(scenes))])
    for i in sample:
        d = Device(0)
        batch_upload(d, scenes, r, rho, idx=[i])
        d.set_tuning(**tun)
        d.sync()
        t0 = time.perf_counter()
        d.run(ENGINES, total, 0, scenes[i]["dt"], **kw)
        d.sync()
        t_ser += time.perf_counter() - t0
        n_ser += scenes[i]["n"]
        same = same and bool(np.array_equal(d.get("pos"), pos[offs[i]:offs[i + 1]]))
        d.close()
    serial = n_ser * total / t_ser
    phi_b, vmax = bed_fraction(pos, vel, scenes, range(len(scenes)), r)
    settled = vmax < 0.05
    t_ref = st1["refreshSeconds"] - st0["refreshSeconds"]
    return dict(value=value, unit="particle-steps/s", scenes=len(scenes), particles=ntot, steps=steps,
                ms_per_step=el / steps * 1e3, tuning=tun, tuning_steps=taken, tuning_s=t_tune,
                refreshes=st1["nRefresh"] - st0["nRefresh"], refresh_seconds=t_ref, refresh_share=t_ref / el,
                serial_plain_engine=dict(
                    value=serial, unit="particle-steps/s", scenes=len(sample), particles=n_ser,
                    note="each sampled scene alone on the plain engine loop (wdg_run), the same steps "
                         "(tuning steps included)"),
                speedup_vs_serial=value / serial, sampled_scenes_identical=same, energy_tracking=energy,
                per_scene=dict(
                    definition="bed solid fraction: spheres whose centres lie below the 90th percentile of the "
                               "centre heights, over the box cross-section times that height plus r; settled: "
                               "largest sphere speed below 0.05 m/s. The sampled scenes run alone end in the same "
                               "positions bit for bit, so their fractions are the same",
                    solid_fraction=[round(float(x), 4) for x in phi_b],
                    settled=int(np.sum(settled)),
                    solid_fraction_settled=dict(mean=float(phi_b[settled].mean()) if settled.any() else None,
                                                min=float(phi_b[settled].min()) if settled.any() else None,
                                                max=float(phi_b[settled].max()) if settled.any() else None),
                    sample_scenes=sample))


def surrogate_form(args):
    """Learned coarse step on the granular column collapse: the network is trained on the GPU on
    DEM trajectories (this engine) of seeded tall narrow columns and of squat columns of the
    collapse's aspect ratio (graphs and targets stay on the device); then plain DEM runs the
    200,000-sphere column until the pile is at rest, and the learned step runs the same simulated
    time by coarse steps of 10 DEM steps, with a DEM re-simulation check every 50 coarse steps
    (fallback to DEM when the position error exceeds 0.1 r_min). Reported: throughput of both,
    fallbacks, and the deposit (edge, 99 % radius, height profile) of each against the initial
    column."""
    import woo_amd.surrogate as S
    from woo_amd._native import Device

    net = S.Network(seed=1)
    t0 = time.perf_counter()
    hist, ball, n_train, n_graphs = S.train_on_columns(net, n_columns=args.sur_columns, epochs=args.sur_epochs,
                                                       rounds=args.sur_rounds)
    t_train = time.perf_counter() - t0
    held, _, _ = S.column_training_set(net, 20000, n_columns=32, snapshots=20, every=1500, seed=5000)
    err_trained, err_ballistic = S.held_out_rmse(net, held)
    dev = Device(0)
    arr, rad = S.collapse_arrays()
    dt = S.dem_dt()
    center = (0.2706 / 2, 0.2706 / 2)
    prof0, edge0, p99_0 = S.deposit(arr["pos"][1:], rad, center)
    # plain DEM: one step (collider, contacts), then until the pile is at rest
    S.upload(dev, arr)
    step = S.dem_run(dev, 1, 0, dt)
    dev.sync()
    t0 = time.perf_counter()
    step_rest, ek_hist = S.collapse_to_rest(dev, dt, step, max_steps=args.sur_max_steps, log=log)
    dev.sync()
    t_dem = time.perf_counter() - t0
    p_dem = dev.get("pos")[1:]
    ek_dem = dev.kinetic_energy()
    coarse = (step_rest - 1) // S.COARSE
    # learned: the same first step, then the same simulated time in coarse steps
    S.upload(dev, arr)
    step = S.dem_run(dev, 1, 0, dt)
    dev.sync()
    # guards beside the position check: the total mechanical energy may not rise over a segment
    # by more than 2e-4 of the initial one, and a fallback makes the next segments plain DEM
    energy = lambda d: S.mechanical_energy(d, arr["mass"])
    slack = 2e-4 * energy(dev)
    t0 = time.perf_counter()
    st = S.run_learned(dev, net, rad, coarse, dt, step=step, energy=energy, slack=slack, backoff=True)
    dev.sync()
    t_sur = time.perf_counter() - t0
    p_sur = dev.get("pos")[1:]
    ek_sur = dev.kinetic_energy()
    dev.close()
    net.close()
    prof_d, edge_d, p99_d = S.deposit(p_dem, rad, center)
    prof_s, edge_s, p99_s = S.deposit(p_sur, rad, center)
    n = len(rad)
    v_dem = n * (coarse * S.COARSE) / t_dem
    v_sur = n * (coarse * S.COARSE) / t_sur
    pile = (prof_d > 0) | (prof_s > 0)
    pile[int(max(edge_d, edge_s) / 0.005) + 1:] = False
    run_d, run_s = edge_d - edge0, edge_s - edge0
    return dict(value=v_sur, unit="particle-steps/s (DEM-step equivalents)", particles=n,
                simulated_time_s=(1 + coarse * S.COARSE) * dt, dem_steps=1 + coarse * S.COARSE,
                rest=dict(rule="kinetic energy below 5 % of its peak and falling by less than 5 % over 40,000 "
                               "steps (spheres rolling away on the floor keep some)",
                          kinetic_energy_peak_J=max(e for _, e in ek_hist), kinetic_energy_dem_J=ek_dem,
                          kinetic_energy_learned_J=ek_sur),
                plain_dem=dict(value=v_dem, unit="particle-steps/s", seconds=t_dem),
                speedup_vs_plain_dem=v_sur / v_dem, seconds=t_sur, checks=st["checks"], fallbacks=st["fallbacks"],
                energy_fallbacks=st["energy_fallbacks"], dem_segments_after_fallbacks=st["dem_segments"],
                guards=dict(position_bound_m=0.2e-3, energy_slack_J=slack,
                            rule="a segment falls back to DEM when the check's position error exceeds 0.1 r_min or "
                                 "the total mechanical energy rose by more than the slack; after a fallback the "
                                 "next 1, 2, 4, ... 64 segments are plain DEM"),
                learned_fraction_of_steps=st["learned_coarse_steps"] * S.COARSE / (coarse * S.COARSE),
                learned_coarse_steps=st["learned_coarse_steps"], dem_steps_in_checks_and_fallbacks=st["dem_steps"],
                check_error_mm=dict(max=1e3 * max(st["errors"]), mean=1e3 * float(np.mean(st["errors"]))),
                deposit=dict(
                    definition="edge: outer radius of the contiguous pile at least 8 r_max high, from the column "
                               "axis, rings of 5 mm; run-out: edge at rest minus the initial column's edge",
                    initial_edge_m=edge0, edge_m=dict(dem=edge_d, learned=edge_s),
                    runout_m=dict(dem=run_d, learned=run_s,
                                  rel_err=abs(run_s - run_d) / run_d if run_d > 0 else None),
                    radius_99pct_m=dict(initial=p99_0, dem=p99_d, learned=p99_s,
                                        rel_err_of_spread=abs((p99_s - p99_0) - (p99_d - p99_0)) / (p99_d - p99_0)),
                    height_m=dict(dem=float(prof_d[0]), learned=float(prof_s[0]), initial=float(prof0[0])),
                    profile_rms_diff_m=float(np.sqrt(np.mean((prof_s[pile] - prof_d[pile]) ** 2))),
                    profile_max_abs_diff_m=float(np.max(np.abs(prof_s[pile] - prof_d[pile])))),
                training=dict(columns=args.sur_columns, squat_columns=S.SQUAT[0], spheres=n_train, graphs=n_graphs,
                              epochs=args.sur_epochs, rollout_rounds=args.sur_rounds, loss_first=hist[0],
                              loss_last=hist[-1], data_and_train_s=t_train, heldout_rmse_learned=err_trained,
                              heldout_rmse_ballistic=err_ballistic, unit="m/s^2"))


def pmc_probe(wl):
    """For rocprofv3 --pmc runs: the kernel-timing launches of the benchmark (FIX_STEPS plain steps
    from the initial state, then KERNEL_LAUNCHES steps), then a device-to-device copy of 512 MiB
    whose counters calibrate FETCH_SIZE and WRITE_SIZE. The engine is deterministic, so the probed
    launches see the state the benchmark times them at; prints that state."""
    dev, step = fixed_state(wl)
    cp, dt = contact_params(), wl["dt"]
    for _ in range(KERNEL_LAUNCHES):
        dev.collide(-.05)
        dev.contact_loop(step, dt, **cp)
        dev.leapfrog(dt, 0., True)
        step += 1
    dev.sync()
    st = dev.stats()
    # calibration: a plain copy of known size (the runtime's device-to-device copy kernel)
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    nbytes = 512 << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert L.hipMalloc(C.byref(a), nbytes) == 0 and L.hipMalloc(C.byref(b), nbytes) == 0
    assert L.hipMemset(a, 1, nbytes) == 0
    assert L.hipDeviceSynchronize() == 0
    assert L.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert L.hipDeviceSynchronize() == 0
    L.hipFree(a)
    L.hipFree(b)
    n = len(wl["radius"])
    B = algorithmic_bytes(n, st["nCandidates"], st["nReal"])
    print(json.dumps({"pmc_probe": KERNEL_LAUNCHES, "fix_steps": FIX_STEPS, "copy_bytes_each_way": nbytes,
                      "algorithmic_bytes": B, "contact_slots": st["nCandidates"], "real_contacts": st["nReal"]}),
          flush=True)
    dev.close()


def history_probe(wl):
    """The contact-history layout against a device hash keyed by the particle ids: the real
    contacts of the fixed kernel-timing state, their 11 history doubles read and 10 written per
    contact, found by index in the engine's packed layout or by linear probing in a hash of load
    factor 1/2 (woo_amd/probe/history_probe.hip); HIP events, 20 launches each."""
    dev, _ = fixed_state(wl)
    ids = dev.contacts(real_only=True, data=False)[0]
    dev.close()
    keys = np.ascontiguousarray((ids[:, 0].astype(np.uint64) << np.uint64(32)) | ids[:, 1].astype(np.uint64))
    L = C.CDLL(os.path.join(ROOT, "woo_amd", "libwdg_probe.so"))
    L.hp_run.argtypes = [C.POINTER(C.c_uint64), C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double),
                         C.POINTER(C.c_uint32)]
    out = dict(history_probe="contact history: packed stream vs hash", contacts=int(len(keys)))
    for mode, name in ((0, "packed_stream"), (1, "hash")):
        ms, miss = C.c_double(), C.c_uint32()
        rc = L.hp_run(keys.ctypes.data_as(C.POINTER(C.c_uint64)), len(keys), mode, 20, C.byref(ms), C.byref(miss))
        if rc != 0:
            raise RuntimeError("history probe failed")
        us = ms.value * 1e3
        out[name] = dict(us_per_launch=us, state_GBs=len(keys) * (88 + 80) / (us * 1e-6) / 1e9, missed=miss.value)
    out["hash_over_stream"] = out["hash"]["us_per_launch"] / out["packed_stream"]["us_per_launch"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--graph", type=int, default=1,
                    help="1: after the warm-up the autotuner picks the force-pass width and the steps per "
                         "hipGraph, and the timed steps are graph replays (bit-identical to the plain loop); "
                         "0: the plain engine loop")
    ap.add_argument("--fp32", type=int, default=1, help="also time the FP32 contact form")
    ap.add_argument("--fp32-drift", type=int, default=1,
                    help="also run FP32 and FP64 side by side for the workload's 5,000 steps, observables compared")
    ap.add_argument("--drum", type=int, default=1, help="also time the rotating-drum workload (facets)")
    ap.add_argument("--drum-n", type=int, default=4_000_000)
    ap.add_argument("--drum-steps", type=int, default=500)
    ap.add_argument("--batch", type=int, default=1, help="also time the batched 256-scene study")
    ap.add_argument("--batch-steps", type=int, default=10000)
    ap.add_argument("--batch-sample", type=int, default=8, help="scenes also run one at a time")
    ap.add_argument("--batch-graph", type=int, default=1,
                    help="0: step the batch with the plain engine loop (no tuning, no graph), e.g. under a tracer")
    ap.add_argument("--batch-skin", type=int, default=16,
                    help="per-scene skin (steps) when the batch is not tuned (--batch-graph 0)")
    ap.add_argument("--surrogate", type=int, default=1, help="also run the learned coarse step (column collapse)")
    ap.add_argument("--sur-columns", type=int, default=500)
    ap.add_argument("--sur-epochs", type=int, default=20)
    ap.add_argument("--sur-rounds", type=int, default=5, help="rounds of training on the network's own rollouts")
    ap.add_argument("--sur-max-steps", type=int, default=500000,
                    help="DEM steps at most while waiting for the collapsed pile to come to rest")
    ap.add_argument("--history-probe", type=int, default=0,
                    help="1: only the contact-history layout probe (packed stream against a hash); prints its "
                         "line, no benchmark line")
    ap.add_argument("--pmc-probe", type=int, default=0,
                    help="1: for rocprofv3 --pmc runs, only the kernel-timing launches (at their fixed state) "
                         "and a device-to-device hipMemcpy of 512 MiB whose FETCH_SIZE/WRITE_SIZE calibrate "
                         "the counters; prints no benchmark line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from woo_amd._native import Device

    dev = Device(0)
    wl = make_workload(args.n, dev)
    dt = wl["dt"]
    step = 0
    cp = contact_params()
    if args.pmc_probe:
        dev.close()
        pmc_probe(wl)
        return
    if args.history_probe:
        dev.close()
        history_probe(wl)
        return
    log(f"dt={dt:.4e} s, warmup {args.warmup} steps")
    step, _ = dev.run(ENGINES, args.warmup, step, dt, contact=cp, leapfrog=dict(damping=0.), dynDtPeriod=0)
    dev.sync()
    tuning = None
    if args.graph:
        # untimed, after the warm-up: the tuner's trial steps advance the run like any other
        tun, step, taken = dev.tune(ENGINES, step, dt, contact=cp, leapfrog=dict(damping=0.), dynDtPeriod=0)
        dev.sync()
        assert taken == TUNE_STEPS, taken
        tuning = dict(tun, tuning_steps=taken)
        log(f"tuning: {tuning}")
    st0 = dev.stats()
    log(f"after warmup: {st0}")
    t0 = time.perf_counter()
    step, _ = dev.run(ENGINES, args.steps, step, dt, graph=bool(args.graph), contact=cp, leapfrog=dict(damping=0.),
                      dynDtPeriod=0)
    dev.sync()
    el = time.perf_counter() - t0
    st1 = dev.stats()
    raw_value = args.n * args.steps / el
    r_window = st1["nRefresh"] - st0["nRefresh"]
    # refresh interval over every step run so far (warm-up, tuner, timed: `step` of them), the first
    # refresh (at step 0) aside
    k_refresh = (step - 1) / max(st1["nRefresh"] - 1, 1)
    # observables at the end of the timed steps (the FP32 form is compared there)
    obs64 = observables(dev, wl)
    c_refresh, step = refresh_cost(dev, step, dt)
    # the timed window with the refreshes it would hold at the run's rate: those it held replaced
    # by K / k refreshes of the measured cost
    el_adj = max(el + (args.steps / k_refresh - r_window) * c_refresh, 1e-12)
    value = args.n * args.steps / el_adj
    log(f"timed {args.steps} steps in {el:.4f} s ({raw_value:.4e} particle-steps/s, {r_window} refreshes); "
        f"refresh every {k_refresh:.1f} steps at {c_refresh * 1e6:.0f} us -> {value:.4e} particle-steps/s")

    fp32 = None
    if args.fp32:
        fp32 = fp32_form(wl, args, raw_value, obs64, tuning)
        log(f"fp32 form: {fp32}")

    # whole step: BASELINE.md's algorithmic bytes per particle-step at the run's state
    B_ps, cn, pn = whole_step_bytes(args.n, st1["nCandidates"], st1["nReal"], k_refresh)
    whole = dict(bytes_per_particle_step=B_ps, contacts_per_particle=cn, slots_per_particle=pn,
                 refresh_interval_steps=k_refresh, refresh_us=c_refresh * 1e6, achieved_GBs=B_ps * value / 1e9,
                 peak_GBs=HBM_PEAK / 1e9, frac=B_ps * value / HBM_PEAK,
                 formula="272 + 80 + 48 + (C/N)*280 + (104 + 8*P/N)/k (BASELINE.md), P = contact slots")

    # per-kernel device time on the engine's stream, at the fixed state of the committed profile
    kdev, kstep = fixed_state(wl)
    tc, tl, kstep = time_kernels(kdev, wl, kstep, KERNEL_LAUNCHES)
    st2 = kdev.stats()
    kdev.close()
    B = algorithmic_bytes(args.n, st2["nCandidates"], st2["nReal"])
    times = {"contact_loop": tc, "k_leapfrog": tl}
    # the roofline is quoted for the engine kernel that takes longest: the contact loop (k_cfilter
    # + k_cforce, timed together) or the leapfrog with its fused gather
    dom = max(times, key=times.get)
    achieved = B[dom] / times[dom]
    log(f"kernel times at step {FIX_STEPS}: contact loop {tc * 1e6:.1f} us ({B['contact_loop'] / tc / 1e9:.0f} GB/s), "
        f"leapfrog+gather {tl * 1e6:.1f} us ({B['k_leapfrog'] / tl / 1e9:.0f} GB/s); slots {st2['nCandidates']}, "
        f"real {st2['nReal']}")
    traffic, tsrc = measured_traffic(dom, st2["nCandidates"], st2["nReal"]) if args.n == 1_000_000 else (None, None)
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=traffic, traffic_source=tsrc, kernel=dom, bytes_per_launch=B[dom], launch_us=times[dom] * 1e6,
                state=dict(step=FIX_STEPS, launches=KERNEL_LAUNCHES, contact_slots=st2["nCandidates"],
                           real_contacts=st2["nReal"]),
                other_kernels={k: dict(bytes_per_launch=B[k], launch_us=times[k] * 1e6,
                                       achieved_GBs=B[k] / times[k] / 1e9,
                                       traffic=measured_traffic(k, st2["nCandidates"], st2["nReal"])[0])
                               for k in times if k != dom})
    # the cost of energy tracking (Scene.trackEnergy) on the same stepping, 200 steps each
    kw = dict(graph=bool(args.graph), contact=cp, leapfrog=dict(damping=0.), dynDtPeriod=0)
    energy = {}
    for on in (False, True):
        dev.set_energy_tracking(on)
        step, _ = dev.run(ENGINES, 8, step, dt, **kw)
        dev.sync()
        t0 = time.perf_counter()
        step, _ = dev.run(ENGINES, 200, step, dt, **kw)
        dev.sync()
        energy["ms_per_step_" + ("tracking" if on else "plain")] = (time.perf_counter() - t0) / 200 * 1e3
    energy["terms"] = dev.energy()
    energy["overhead"] = energy["ms_per_step_tracking"] / energy["ms_per_step_plain"] - 1.
    log(f"energy tracking: {energy}")
    drum = batch = None
    dev.close()
    if args.drum:
        drum = drum_form(args)
        log(f"drum form: {drum}")
    if args.batch:
        batch = batched_form(args)
        log(f"batched form: {batch}")
    sur = None
    if args.surrogate:
        sur = surrogate_form(args)
        log(f"surrogate form: {sur}")
    base = None
    if args.cpu_baseline:
        log("timing the CPU oracle on the same workload ...")
        base = cpu_baseline(wl)
        log(f"cpu baseline: {base}")
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el_adj / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded polydisperse periodic packing at solid fraction 0.55, Gaussian velocities",
        "config": {"workload": "1M polydisperse Hertz spheres, periodic cube, phi=0.55, FP64 (Collider, ContactLoop, "
                               "Leapfrog)", "n": args.n, "dt": dt, "cell": wl["L"],
                   "contact_slots": st1["nCandidates"], "real_contacts": st1["nReal"],
                   "refreshes_in_timed_region": r_window,
                   "timed_window": {"seconds": el, "raw_value": raw_value, "refreshes": r_window,
                                    "refresh_interval_steps": k_refresh, "refresh_us": c_refresh * 1e6,
                                    "value_rule": "value = n*K / (window seconds + (K/k - refreshes in the window) "
                                                  "* refresh seconds)"},
                   "stepping": "hipGraph replay, tuned" if args.graph else "plain engine loop",
                   "tuning": tuning},
        "roofline": roof,
        "whole_step_roofline": whole,
        "cpu_baseline": base,
        "fp32_form": fp32,
        "energy_tracking": energy,
        "drum_form": drum,
        "batched_form": batch,
        "surrogate_form": sur,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
