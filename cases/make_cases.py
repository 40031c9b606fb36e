#!/usr/bin/env python3
"""Write the case directories used by the tests and the benchmark.

Each case is a directory holding ``artracfd.case`` and ``artracfd.geo`` in
the solver's own grammar (case_loader.c:40-285, initialization.c:163-209),
plus a binary STL where a triangulated body is used (stl.c:39-79).  The
files are committed; this script documents how they were made and can
rewrite them:  python cases/make_cases.py
"""
import math
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def case_text(space, time, numerical, material, reference, init, bcs, regions=(), probes=None):
    """Assemble an artracfd.case file.

    space: ((x0,y0,z0),(x1,y1,z1),(mx,my,mz)); time: (restart,end,cfl,stepN,dataW,streamer)
    numerical: (tScheme,sScheme,multidim,jacobMean,fluxSplit,psi,ibmLayer)
    material: (mid, refMu, gState, (gx,gy,gz)); reference: (L, rho, V, T)
    init: five expressions; bcs: six entries (W,E,S,N,F,B), each a list of lines
    regions: (kind, lines) for plane/sphere/box/cylinder initializers
    probes: (points, lines, curve, force, control) with points [(x,y,z)],
      lines [((x0,y0,z0),(x1,y1,z1),samples)], curve/force counts (on when > 0)
      and control the four writing frequencies (case_loader.c:185-222)
    """
    out = ["# case written by cases/make_cases.py"]
    (x0, y0, z0), (x1, y1, z1), (mx, my, mz) = space
    out += ["space begin", f"{x0!r}, {y0!r}, {z0!r}", f"{x1!r}, {y1!r}, {z1!r}",
            f"{mx}, {my}, {mz}", "space end"]
    out += ["time begin"] + [str(v) for v in time] + ["time end"]
    out += ["numerical begin"] + [str(v) for v in numerical] + ["numerical end"]
    mid, mu, gs, g = material
    out += ["material begin", str(mid), repr(mu), str(gs), f"{g[0]}, {g[1]}, {g[2]}", "material end"]
    out += ["reference begin"] + [repr(v) for v in reference] + ["reference end"]
    out += ["initialization begin"] + list(init) + ["initialization end"]
    for name, lines in zip(("west", "east", "south", "north", "front", "back"), bcs):
        out += [f"{name} boundary begin"] + list(lines) + [f"{name} boundary end"]
    for kind, lines in regions:
        out += [f"{kind} initialization begin"] + list(lines) + [f"{kind} initialization end"]
    if probes is not None:
        points, plines, curve, force, control = probes
        out += ["probe count begin", str(len(points)), str(len(plines)), str(curve), str(force), "probe count end"]
        out += ["probe control begin"] + [str(v) for v in control] + ["probe control end"]
        if points:
            out += ["point probe begin"] + [f"{x!r}, {y!r}, {z!r}" for (x, y, z) in points] + ["point probe end"]
        if plines:
            out.append("line probe begin")
            for (a, b, m) in plines:
                out += [f"{a[0]!r}, {a[1]!r}, {a[2]!r}", f"{b[0]!r}, {b[1]!r}, {b[2]!r}", str(m)]
            out.append("line probe end")
    return "\n".join(out) + "\n"


# Temperatures are nondimensional: with reference values (1, 1, 1, 1) the gas
# constant is 1/(1.4 Ma^2) ~ 287, so T = p / (rho R) ~ 0.0035 for p ~ rho ~ 1.
STATE_STATIONARY = ("0, 0, 0, {r}, 0, 0, 0, 0, 0, 0, {rho}, {T}, {cf}, 0, 0, 0",
                    "0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0")


def geo_text(spheres=(), stls=()):
    """spheres: (x, y, z, r, T, cf[, rho]); stls: (file, r_hint, T, cf[, rho]).  A body density
    above 1e36 marks a stationary body without surface-force integration; a finite one (the body
    still does not move while the phase interaction is off) has its surface force integrated
    (initialization.c:235-256)."""
    out = ["count begin", str(len(spheres)), str(len(stls)), "count end"]
    if spheres:
        out.append("sphere state begin")
        for sp in spheres:
            (x, y, z, r, T, cf), rho = sp[:6], (sp[6] if len(sp) > 6 else "1e37")
            out.append(f"{x}, {y}, {z}, {r}, 0, 0, 0, 0, 0, 0, {rho}, {T}, {cf}, 0, 0, 0")
            out.append(STATE_STATIONARY[1])
        out.append("sphere state end")
    if stls:
        out.append("polyhedron geometry begin")
        out += [st[0] for st in stls]
        out.append("polyhedron geometry end")
        out.append("polyhedron state begin")
        for st in stls:
            (_, r, T, cf), rho = st[:4], (st[4] if len(st) > 4 else "1e37")
            out.append(STATE_STATIONARY[0].format(r=r, T=T, cf=cf, rho=rho))
            out.append(STATE_STATIONARY[1])
        out.append("polyhedron state end")
        out.append("polyhedron transform begin")
        out += ["1, 1, 1, 0, 0, 0, 0, 0, 0" for _ in stls]
        out.append("polyhedron transform end")
    return "\n".join(out) + "\n"


def body_lines(O=(0, 0, 0), r=0.1, V=(0, 0, 0), W=(0, 0, 0), rho=1.0, T=-1, cf=0, mid=0,
               at=(0, 0, 0), g=(0, 0, 0), ar=(0, 0, 0), to=0):
    """The two state lines of a body in the order of data_stream.c:154-188: centre, radius,
    velocity, angular velocity, density, wall temperature (< 0 adiabatic), friction (> 0 no-slip),
    area, volume (computed), material (restitution 0, 0.25, ..., 1 for 0..4); then two ignored
    accelerations, the powered acceleration, gravity, the powered angular acceleration and the time
    the power ends (<= 0: never)."""
    f = lambda v: ", ".join(str(q) for q in v)
    return (f"{f(O)}, {r}, {f(V)}, {f(W)}, {rho}, {T}, {cf}, 0, 0, {mid}",
            f"0, 0, 0, 0, 0, 0, {f(at)}, {f(g)}, {f(ar)}, {to}")


def moving_geo_text(spheres=(), stls=()):
    """Bodies with their full state: spheres as body_lines keyword dicts, stls as (file, dict)."""
    out = ["count begin", str(len(spheres)), str(len(stls)), "count end"]
    if spheres:
        out += ["sphere state begin"] + [q for sp in spheres for q in body_lines(**sp)] + ["sphere state end"]
    if stls:
        out += ["polyhedron geometry begin"] + [st[0] for st in stls] + ["polyhedron geometry end"]
        out += ["polyhedron state begin"] + [q for st in stls for q in body_lines(**st[1])]
        out += ["polyhedron state end", "polyhedron transform begin"]
        out += ["1, 1, 1, 0, 0, 0, 0, 0, 0" for _ in stls] + ["polyhedron transform end"]
    return "\n".join(out) + "\n"


def cylinder_stl(path, cx, cy, r, z0, z1, segments):
    """Closed cylinder along z: 2*segments side facets plus two fans."""
    f32 = lambda v: struct.unpack("<f", struct.pack("<f", v))[0]
    ring = [(f32(cx + r * math.cos(2 * math.pi * i / segments)),
             f32(cy + r * math.sin(2 * math.pi * i / segments))) for i in range(segments)]
    bz, tz = f32(z0), f32(z1)
    cb, ct = (f32(cx), f32(cy), bz), (f32(cx), f32(cy), tz)
    tris = []
    for i in range(segments):
        a, b = ring[i], ring[(i + 1) % segments]
        B0, B1, T0, T1 = (a[0], a[1], bz), (b[0], b[1], bz), (a[0], a[1], tz), (b[0], b[1], tz)
        tris.append((B0, B1, T1))
        tris.append((B0, T1, T0))
    for i in range(segments):
        a, b = ring[i], ring[(i + 1) % segments]
        tris.append((ct, (a[0], a[1], tz), (b[0], b[1], tz)))
    for i in range(segments):
        a, b = ring[i], ring[(i + 1) % segments]
        tris.append((cb, (b[0], b[1], bz), (a[0], a[1], bz)))
    with open(path, "wb") as fp:
        fp.write(b"cylinder written by cases/make_cases.py".ljust(80, b" "))
        fp.write(struct.pack("<I", len(tris)))
        for t in tris:
            e1 = [t[1][q] - t[0][q] for q in range(3)]
            e2 = [t[2][q] - t[0][q] for q in range(3)]
            n = (e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0])
            fp.write(struct.pack("<3f", *n))
            for v in t:
                fp.write(struct.pack("<3f", *v))
            fp.write(struct.pack("<H", 0))
    return len(tris)


def icosphere_stl(path, cx, cy, cz, r, level):
    """Closed sphere: an icosahedron whose facets are split in four `level`
    times, the new vertices pushed onto the sphere; 20 * 4^level facets."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [tuple(q / math.sqrt(sum(c * c for c in v)) for q in v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def middle(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = [(verts[a][q] + verts[b][q]) / 2.0 for q in range(3)]
                n = math.sqrt(sum(c * c for c in m))
                verts.append(tuple(c / n for c in m))
                mid[key] = len(verts) - 1
            return mid[key]
        nxt = []
        for a, b, c in faces:
            ab, bc, ca = middle(a, b), middle(b, c), middle(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    f32 = lambda v: struct.unpack("<f", struct.pack("<f", v))[0]
    pts = [(f32(cx + r * v[0]), f32(cy + r * v[1]), f32(cz + r * v[2])) for v in verts]
    with open(path, "wb") as fp:
        fp.write(b"icosphere written by cases/make_cases.py".ljust(80, b" "))
        fp.write(struct.pack("<I", len(faces)))
        for a, b, c in faces:
            t0, t1, t2 = pts[a], pts[b], pts[c]
            e1 = [t1[q] - t0[q] for q in range(3)]
            e2 = [t2[q] - t0[q] for q in range(3)]
            n = (e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0])
            fp.write(struct.pack("<3f", *n))
            for v in (t0, t1, t2):
                fp.write(struct.pack("<3f", *v))
            fp.write(struct.pack("<H", 0))
    return len(faces)


OUTFLOW = ["outflow"]
PERIODIC = ["periodic"]


def slip(T=-1):
    return ["slip wall", str(T)]


def noslip(T=-1):
    return ["noslip wall", str(T)]


def inflow(rho, u, v, w, p):
    return ["inflow", str(rho), str(u), str(v), str(w), str(p)]


def plane(point, normal, state):
    return ("plane", [f"{point[0]}, {point[1]}, {point[2]}", f"{normal[0]}, {normal[1]}, {normal[2]}"]
            + [str(s) for s in state])


def box(lo, hi, state):
    return ("box", [f"{lo[0]}, {lo[1]}, {lo[2]}", f"{hi[0]}, {hi[1]}, {hi[2]}"] + [str(s) for s in state])


def sphere(c, r, state):
    return ("sphere", [f"{c[0]}, {c[1]}, {c[2]}", str(r)] + [str(s) for s in state])


DEFAULT_NUM = (1, 1, 0, 0, 0, 0, 1)   # RK3, WENO5, split, arithmetic mean, LLF, fluid only, 1 IB layer
REF1 = (1.0, 1.0, 1.0, 1.0)
INVISCID = (0, 0.0, 0, (0, 0, 0))

# Mach 2.81 shock into rho=1.4, p=1 at rest (Rankine-Hugoniot)
CYL_POST = (5.143201433, 2.045106762, 0, 0, 9.04545)
# Mach 1.5 shock into rho=1, p=1 at rest
SV_POST = (1.862068966, 0.8216777477, 0, 0, 2.458333333)
# Gaussian isentropic vortex, strength 0.3, radius 0.075, centre (0.75, 1), alpha 0.204
_R2 = "((x-0.75)^2+(y-1)^2)/0.005625"
SV_VORTEX = ("(1-0.0315126*exp(0.408-0.408*" + _R2 + "))^2.5",
             "-4*(y-1)*exp(0.204-0.204*" + _R2 + ")",
             "4*(x-0.75)*exp(0.204-0.204*" + _R2 + ")",
             "0",
             "(1-0.0315126*exp(0.408-0.408*" + _R2 + "))^3.5")

TGV_INIT = ("1", "sin(x)*cos(y)*cos(z)", "-cos(x)*sin(y)*cos(z)", "0",
            "71.4285714+(cos(2*x)+cos(2*y))*(cos(2*z)+2)/16")
TWO_PI = 6.283185307179586


def cases():
    c = {}
    # 1. Sod shock tube, 256x4x4, t = 0.2
    c["sod"] = dict(case=case_text(((0, 0, 0), (1, 0.016, 0.016), (256, 4, 4)), (0, 0.2, 0.6, 0, 1, 0),
                                   DEFAULT_NUM, INVISCID, REF1, ("0.125", "0", "0", "0", "0.1"),
                                   [OUTFLOW, OUTFLOW, PERIODIC, PERIODIC, PERIODIC, PERIODIC],
                                   [plane((0.5, 0, 0), (-1, 0, 0), (1, 0, 0, 0, 1))]),
                    geo=geo_text())
    # 2. viscous Taylor-Green vortex, Re 1600, Ma 0.1 (refMu 1188.6 gives mu = 1/1600 at T = 1)
    for n, steps in ((16, 20), (32, 20), (64, 200), (128, 200)):
        c[f"tgv{n}"] = dict(case=case_text(((0, 0, 0), (TWO_PI, TWO_PI, TWO_PI), (n, n, n)),
                                           (0, 1e9, 0.6, steps, 1, 0), DEFAULT_NUM,
                                           (0, 1188.6, 0, (0, 0, 0)), (1, 1, 34.0297, 288.15), TGV_INIT,
                                           [PERIODIC] * 6),
                            geo=geo_text())
    # 3. Mach 2.81 shock over a stationary cylinder (2048-facet STL), inviscid
    for tag, (lo, hi, m, steps, shock_x) in {
            "cylinder": ((0, 0, 0), (2, 2, 0.016), (512, 512, 4), 1000, 0.6),
            "cyl64": ((0.5, 0.5, 0), (1.5, 1.5, 0.016), (64, 64, 4), 100, 0.8)}.items():
        c[tag] = dict(case=case_text((lo, hi, m), (0, 1e9, 0.6, steps, 1, 0), DEFAULT_NUM, INVISCID, REF1,
                                     ("1.4", "0", "0", "0", "1"),
                                     [inflow(*CYL_POST), OUTFLOW, slip(), slip(), PERIODIC, PERIODIC],
                                     [plane((shock_x, 0, 0), (-1, 0, 0), CYL_POST)]),
                      geo=geo_text(stls=[("cylinder.stl", 0.1, -1, 0)]),
                      stl=("cylinder.stl", 1.0, 1.0, 0.1, -0.05, 0.066, 512))
    # 3b. the same with every probe kind of the solver (data_probe.c:22-197): two points, two lines,
    # the body-conformal curve and the surface force, written every 10th of the run to t = 0.04, and the
    # field written twice
    c["cyl64_probes"] = dict(case=case_text(((0.5, 0.5, 0), (1.5, 1.5, 0.016), (64, 64, 4)), (0, 0.04, 0.6, 0, 2, 0),
                                            DEFAULT_NUM, INVISCID, REF1, ("1.4", "0", "0", "0", "1"),
                                            [inflow(*CYL_POST), OUTFLOW, slip(), slip(), PERIODIC, PERIODIC],
                                            [plane((0.8, 0, 0), (-1, 0, 0), CYL_POST)],
                                            probes=([(0.62, 1.0, 0.008), (1.3, 1.17, 0.0)],
                                                    [((0.55, 1.0, 0.008), (1.45, 1.0, 0.008), 40),
                                                     ((1.0, 0.5, 0.0), (1.0, 1.5, 0.016), 7)],
                                                    1, 1, (10, 10, 10, 10))),
                             geo=geo_text(stls=[("cylinder.stl", 0.1, -1, 0, 2.7)]),
                             stl=("cylinder.stl", 1.0, 1.0, 0.1, -0.05, 0.066, 512))
    # 4. viscous flow past a no-slip isothermal sphere (analytical body), every IB layer reconstructed
    c["sphere20"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (20, 20, 20)),
                                        (0, 1e9, 0.6, 20, 1, 0), (1, 1, 0, 0, 0, 0, 0),
                                        (0, 0.05, 0, (0, 0, 0)), (1, 1, 1, 1),
                                        ("1", "0.3", "0.1*sin(3*y)", "0", "1+0.1*cos(4*z)"),
                                        [inflow(1, 0.3, 0, 0, 1), OUTFLOW, noslip(0.004), slip(), OUTFLOW,
                                         OUTFLOW]),
                         geo=geo_text(spheres=[(0.05, 0.0, 0.0, 0.2, 0.0038, 1)]))
    # 4b. the same with probes: the surface force then has its viscous (no-slip) part
    c["sphere20_probes"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (20, 20, 20)),
                                               (0, 0.15, 0.6, 0, 2, 0), (1, 1, 0, 0, 0, 0, 0),
                                               (0, 0.05, 0, (0, 0, 0)), (1, 1, 1, 1),
                                               ("1", "0.3", "0.1*sin(3*y)", "0", "1+0.1*cos(4*z)"),
                                               [inflow(1, 0.3, 0, 0, 1), OUTFLOW, noslip(0.004), slip(), OUTFLOW,
                                                OUTFLOW],
                                               probes=([(0.3, 0.1, -0.05)], [((-0.5, 0.0, 0.0), (0.5, 0.0, 0.0), 30)],
                                                       1, 1, (5, 5, 5, 5))),
                                geo=geo_text(spheres=[(0.05, 0.0, 0.0, 0.2, 0.0038, 1, 2.7)]))
    # 4c. the cylinder and the sphere again with the EnSight Gold writer (ensight_writer.c:34-452): binary
    # field, geometry and body files and the transient case files
    c["cyl64_ensight"] = dict(case=case_text(((0.5, 0.5, 0), (1.5, 1.5, 0.016), (64, 64, 4)), (0, 0.04, 0.6, 0, 2, 1),
                                             DEFAULT_NUM, INVISCID, REF1, ("1.4", "0", "0", "0", "1"),
                                             [inflow(*CYL_POST), OUTFLOW, slip(), slip(), PERIODIC, PERIODIC],
                                             [plane((0.8, 0, 0), (-1, 0, 0), CYL_POST)]),
                              geo=geo_text(stls=[("cylinder.stl", 0.1, -1, 0, 2.7)]),
                              stl=("cylinder.stl", 1.0, 1.0, 0.1, -0.05, 0.066, 512))
    c["sphere20_ensight"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (20, 20, 20)),
                                                (0, 0.15, 0.6, 0, 6, 1), (1, 1, 0, 0, 0, 0, 0),
                                                (0, 0.05, 0, (0, 0, 0)), (1, 1, 1, 1),
                                                ("1", "0.3", "0.1*sin(3*y)", "0", "1+0.1*cos(4*z)"),
                                                [inflow(1, 0.3, 0, 0, 1), OUTFLOW, noslip(0.004), slip(), OUTFLOW,
                                                 OUTFLOW]),
                                 geo=geo_text(spheres=[(0.05, 0.0, 0.0, 0.2, 0.0038, 1, 2.7)]))
    # 4d. moving bodies (phase interaction on): the cylinder driven by the shock while it translates
    # and spins (psi 1: no collisions), written with probes and surface force; two spheres that meet
    # and rebound (psi 2: collisions), one powered for a while and under gravity, the other
    # isothermal and no-slip, with image reconstruction in the first ghost layer only
    c["cyl64_moving"] = dict(case=case_text(((0.5, 0.5, 0), (1.5, 1.5, 0.016), (64, 64, 4)), (0, 0.04, 0.6, 0, 2, 0),
                                            (1, 1, 0, 0, 0, 1, 0), INVISCID, REF1, ("1.4", "0", "0", "0", "1"),
                                            [inflow(*CYL_POST), OUTFLOW, slip(), slip(), PERIODIC, PERIODIC],
                                            [plane((0.8, 0, 0), (-1, 0, 0), CYL_POST)],
                                            probes=([(0.62, 1.0, 0.008)], [], 1, 1, (10, 10, 10, 10))),
                             geo=moving_geo_text(stls=[("cylinder.stl", dict(r=0.1, V=(0.8, 0.5, 0), W=(0, 0, 3),
                                                                             rho=2.7))]),
                             stl=("cylinder.stl", 1.0, 1.0, 0.1, -0.05, 0.066, 512))
    # the viscous sphere moving through the flow: translating and spinning no-slip isothermal body
    # whose surface force has its viscous part (psi 1)
    c["sphere20_moving"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (20, 20, 20)),
                                               (0, 0.15, 0.6, 0, 2, 0), (1, 1, 0, 0, 0, 1, 0),
                                               (0, 0.05, 0, (0, 0, 0)), (1, 1, 1, 1),
                                               ("1", "0.3", "0.1*sin(3*y)", "0", "1+0.1*cos(4*z)"),
                                               [inflow(1, 0.3, 0, 0, 1), OUTFLOW, noslip(0.004), slip(), OUTFLOW,
                                                OUTFLOW],
                                               probes=([(0.3, 0.1, -0.05)], [], 1, 1, (5, 5, 5, 5))),
                                geo=moving_geo_text(spheres=[dict(O=(0.05, 0, 0), r=0.2, V=(-0.6, 0.2, 0.1),
                                                                  W=(0, 0, 2), rho=2.7, T=0.0038, cf=1)]))
    # a triangulated sphere (1280 facets) translating and tumbling about a tilted axis through a
    # wavy stream (psi 1): closest-facet searches in all three directions
    c["ico_moving"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (28, 28, 28)),
                                          (0, 0.15, 0.6, 0, 2, 0), (1, 1, 0, 0, 0, 1, 0), INVISCID, REF1,
                                          ("1", "0.3", "0.1*sin(3*y)", "0", "1+0.1*cos(4*z)"),
                                          [inflow(1, 0.3, 0, 0, 1), OUTFLOW, slip(), slip(), OUTFLOW, OUTFLOW],
                                          probes=([(0.3, 0.1, -0.05)], [], 1, 1, (5, 5, 5, 5))),
                           geo=moving_geo_text(stls=[("ico.stl", dict(r=0.21, V=(0.5, -0.3, 0.2),
                                                                      W=(1.5, -2.0, 2.5), rho=2.0))]),
                           ico=("ico.stl", 0.03, -0.02, 0.01, 0.21, 3))
    # the benchmark's immersed cylinder (512x512x4) moving the same way, 20 steps
    c["cylinder_moving"] = dict(case=case_text(((0, 0, 0), (2, 2, 0.016), (512, 512, 4)), (0, 1e9, 0.6, 60, 1, 0),
                                               (1, 1, 0, 0, 0, 1, 0), INVISCID, REF1, ("1.4", "0", "0", "0", "1"),
                                               [inflow(*CYL_POST), OUTFLOW, slip(), slip(), PERIODIC, PERIODIC],
                                               [plane((0.6, 0, 0), (-1, 0, 0), CYL_POST)]),
                                geo=moving_geo_text(stls=[("cylinder.stl", dict(r=0.1, V=(0.8, 0.5, 0), W=(0, 0, 3),
                                                                                rho=2.7))]),
                                stl=("cylinder.stl", 1.0, 1.0, 0.1, -0.05, 0.066, 512))
    c["spheres_collide"] = dict(case=case_text(((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), (24, 24, 24)),
                                               (0, 0.08, 0.6, 0, 2, 0), (1, 1, 0, 0, 0, 2, 1), INVISCID, REF1,
                                               ("1", "0", "0", "0", "1"), [OUTFLOW] * 6,
                                               probes=([(0.0, 0.3, 0.0)], [], 1, 1, (5, 5, 5, 5))),
                                geo=moving_geo_text(spheres=[
                                    dict(O=(-0.2, 0, 0), r=0.15, V=(1.5, 0, 0), rho=1.5, mid=4,
                                         at=(2, 0, 0), g=(0, -1, 0), to=0.02),
                                    dict(O=(0.19, 0.02, 0), r=0.15, V=(-1.5, 0, 0), W=(0, 0, 2), rho=2.0,
                                         T=0.004, cf=0.2, mid=2)]))
    # 5. scheme variants on a small 2D Riemann-like problem (10 steps each)
    quad = [box((-1, -1, -1), (0.5, 0.5, 1), (0.5, 0.1, -0.2, 0, 0.4)),
            sphere((0.7, 0.3, 0.008), 0.15, (2.0, 0, 0, 0, 3.0))]
    base_bcs = [OUTFLOW, OUTFLOW, slip(), noslip(0.0037), PERIODIC, PERIODIC]
    glob = ("1+0.2*sin(6*x)*cos(4*y)", "0.1", "0.05*cos(3*x)", "0", "1")
    variants = {
        "v_base": ((1, 1, 0, 0, 0, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_weno3": ((1, 0, 0, 0, 0, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_rk2": ((0, 1, 0, 0, 0, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_roe": ((1, 1, 0, 1, 0, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_sw": ((1, 1, 0, 0, 1, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_unsplit": ((1, 1, 1, 0, 0, 0, 1), INVISCID, base_bcs, (24, 16, 4)),
        "v_mix": ((0, 0, 1, 1, 1, 0, 1), (0, 0.02, 0, (0, 0, 0)), base_bcs, (24, 16, 4)),
        "v_visc": ((1, 1, 0, 0, 0, 0, 1), (0, 0.05, 0, (0, 0, 0)), base_bcs, (24, 16, 4)),
        "v_collapse": ((1, 1, 0, 0, 0, 0, 1), INVISCID,
                       [inflow(1.5, 0.3, 0, 0, 1.2), OUTFLOW, slip(), OUTFLOW, OUTFLOW, OUTFLOW], (24, 16, 1)),
        "v_walls": ((1, 1, 0, 0, 0, 0, 1), INVISCID,
                    [slip(), noslip(), noslip(0.0033), slip(0.0036), slip(), noslip()], (16, 12, 8)),
        # gravity source term (source_term.c:21-39) with its half-step operators around the split sweeps
        "v_gravity": ((1, 1, 0, 0, 0, 0, 1), (0, 0.0, 1, (0.3, -2.0, 0.5)), base_bcs, (24, 16, 4)),
        "v_gravity2": ((0, 1, 1, 0, 0, 0, 1), (0, 0.02, 1, (-0.5, 1.5, 0)), base_bcs, (24, 16, 4)),
    }
    for tag, (num, mat, bcs, m) in variants.items():
        c[tag] = dict(case=case_text(((0, 0, 0), (1, 0.6, 0.016 if m[2] <= 4 else 0.4), m),
                                     (0, 1e9, 0.6, 10, 1, 0), num, mat, REF1, glob, bcs, quad),
                      geo=geo_text())
    # 6. shock-vortex interaction (benchmark workload at 256^3; 64^3 for tests)
    for n in (64, 128, 256):
        c[f"shockvortex{n}"] = dict(case=case_text(((0, 0, 0), (2, 2, 2), (n, n, n)), (0, 0.6, 0.6, 0, 1, 0),
                                                   DEFAULT_NUM, INVISCID, REF1, SV_VORTEX,
                                                   [inflow(*SV_POST), OUTFLOW, OUTFLOW, OUTFLOW, OUTFLOW,
                                                    OUTFLOW],
                                                   [plane((0.5, 0, 0), (-1, 0, 0), SV_POST)]),
                                    geo=geo_text())
    return c


def main():
    for tag, spec in cases().items():
        d = os.path.join(HERE, tag)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "artracfd.case"), "w") as fp:
            fp.write(spec["case"])
        with open(os.path.join(d, "artracfd.geo"), "w") as fp:
            fp.write(spec["geo"])
        if "stl" in spec:
            name, cx, cy, r, z0, z1, seg = spec["stl"]
            nf = cylinder_stl(os.path.join(d, name), cx, cy, r, z0, z1, seg)
            assert nf == 4 * seg
        if "ico" in spec:
            name, cx, cy, cz, r, level = spec["ico"]
            nf = icosphere_stl(os.path.join(d, name), cx, cy, cz, r, level)
            assert nf == 20 * 4 ** level
    for tag in sorted(cases()):
        for line in cases()[tag]["case"].splitlines():
            assert len(line) < 100, (tag, line)
    print("wrote", len(cases()), "cases")


if __name__ == "__main__":
    main()
