"""Deck generators for the workloads this path is measured and tested on.

Each function returns .peec text in the reference's deck grammar (OpenPEEC input_data.c):

  dipole_deck      centre-fed thin-wire dipole along z, 0.5 m, a = 1 mm, 2 x 16 segments,
                   100-500 MHz in 41 points (the reference's CPU-runnable case)
  plate_deck       copper plate 100 x 100 mm, 35 um, n x n grid, corner-to-corner port,
                   10 MHz - 3 GHz in 201 points (n = 64: the precision workload)
  random_wire_decks  seeded random polylines (batching workload), 20-400 segments each
  disk_quad_deck   disk panel under a quad panel, compressed (learned-preconditioner workload:
                   48 rings x 96 sectors, 32 x 16 quad, 100 MHz - 4 GHz in 101 points; the disk's
                   centre ring of 192 vertices is built as 4 sector cells, since a cell record
                   holds at most 64 polygon vertices, peec.h seg_t.pv; the reference itself stops
                   at 32 sectors)
"""
from __future__ import annotations

import math


def dipole_deck(nseg_arm: int = 16, f0: float = 1e8, f1: float = 5e8, ndiv: int = 40,
                length: float = 0.5, radius: float = 1e-3, gap: float = 1e-3) -> str:
    h = length / 2
    g = gap / 2
    return "\n".join([
        "OpenPEEC 1 0",
        "title = centre-fed dipole",
        "capacitance = 1",
        "retardation = 1",
        f"node = 1 0 0 {-g!r}",
        f"node = 2 0 0 {g!r}",
        f"wire = 0 0 {-h!r} 0 0 {-g!r} {radius!r} 5.8e7 {nseg_arm}",
        f"wire = 0 0 {g!r} 0 0 {h!r} {radius!r} 5.8e7 {nseg_arm}",
        "port = 1 2 50",
        f"frequency = {f0!r} {f1!r} {ndiv}",
        "end", ""])


def plate_deck(n: int = 64, size: float = 0.1, thick: float = 3.5e-5, f0: float = 1e7,
               f1: float = 3e9, ndiv: int = 200, groundplane: float | None = None,
               z: float = 0.0, ndivt: int = 1) -> str:
    lines = ["OpenPEEC 1 0", f"title = plate {n}x{n}", "capacitance = 1", "retardation = 1"]
    if groundplane is not None:
        lines.append(f"groundplane = {groundplane!r}")
    lines += [
        f"node = 1 0 0 {z!r}",
        f"node = 2 {size!r} {size!r} {z!r}",
        f"plate = 0 0 {z!r} {size!r} 0 0 0 {size!r} 0 {thick!r} 5.8e7 {n} {n}" + (f" {ndivt}" if ndivt > 1 else ""),
        "port = 1 2 50",
        f"frequency = {f0!r} {f1!r} {ndiv}",
        "end", ""]
    return "\n".join(lines)


class _SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def random_wire_deck(seed: int, nmin: int = 20, nmax: int = 400, box: float = 0.3,
                     f0: float = 5e7, f1: float = 2e9, ndiv: int = 63) -> str:
    """One random polyline: n = U{nmin..nmax} straight segments of U[5, 20] mm, directions
    uniform on the sphere (re-drawn up to 64 times to stay inside [0, box]^3), radius
    U[0.5, 2] mm; segment 0 carries a 1 mm delta-gap feed at its midpoint."""
    rng = _SplitMix64(seed)
    n = nmin + int(rng.uniform() * (nmax - nmin + 1))
    n = min(n, nmax)
    a = rng.uniform(0.5e-3, 2e-3)
    pts = [[box / 2, box / 2, box / 2]]
    for _ in range(n):
        L = rng.uniform(5e-3, 20e-3)
        p = pts[-1]
        for _try in range(64):
            zc = rng.uniform(-1, 1)
            ph = rng.uniform(0, 2 * math.pi)
            rr = math.sqrt(max(0.0, 1 - zc * zc))
            d = [rr * math.cos(ph), rr * math.sin(ph), zc]
            q = [p[k] + L * d[k] for k in range(3)]
            if all(0 <= q[k] <= box for k in range(3)):
                break
        else:
            q = [min(max(p[k] + L * d[k], 0.0), box) for k in range(3)]
        pts.append(q)
    # feed gap at the midpoint of segment 0
    p0, p1 = pts[0], pts[1]
    L0 = math.dist(p0, p1)
    u = [(p1[k] - p0[k]) / L0 for k in range(3)]
    mid = [(p0[k] + p1[k]) / 2 for k in range(3)]
    ga = [mid[k] - 0.5e-3 * u[k] for k in range(3)]
    gb = [mid[k] + 0.5e-3 * u[k] for k in range(3)]
    fmt = lambda v: " ".join(repr(float(x)) for x in v)
    lines = ["OpenPEEC 1 0", f"title = random wire seed {seed}", "capacitance = 1", "retardation = 1",
             f"node = 1 {fmt(ga)}", f"node = 2 {fmt(gb)}",
             f"wire = {fmt(p0)} {fmt(ga)} {a!r} 5.8e7 1",
             f"wire = {fmt(gb)} {fmt(p1)} {a!r} 5.8e7 1"]
    for k in range(1, n):
        lines.append(f"wire = {fmt(pts[k])} {fmt(pts[k + 1])} {a!r} 5.8e7 1")
    lines += ["port = 1 2 50", f"frequency = {f0!r} {f1!r} {ndiv}", "end", ""]
    return "\n".join(lines)


def disk_quad_deck(nr: int = 48, nsec: int = 96, qx: int = 32, qy: int = 16, radius: float = 0.05,
                   qw: float = 0.08, qh: float = 0.04, gap: float = 0.005, f0: float = 1e8, f1: float = 4e9,
                   ndiv: int = 100, ctol: float = 1e-6, thick: float = 3.5e-5) -> str:
    """Disk panel of radius `radius` (nr rings x nsec sectors) at z = 0 and a qw x qh quad panel
    (qx x qy) centred above it at z = gap, compressed (ACA tolerance ctol), retarded and
    capacitive; the port runs from the disk centre to the quad's centre node (qx and qy even, so
    that the centre is a grid node)."""
    if (qx % 2) or (qy % 2):
        raise ValueError("qx and qy must be even: the port node is the quad's centre")
    x0, y0 = -qw / 2, -qh / 2
    return "\n".join([
        "OpenPEEC 1 0", f"title = disk {nr}x{nsec} under quad {qx}x{qy}",
        "capacitance = 1", "retardation = 1", f"compression = 1 {ctol!r}",
        "node = 1 0 0 0", f"node = 2 0 0 {gap!r}",
        f"disk = 0 0 0 0 0 1 {radius!r} {thick!r} 5.8e7 {nr} {nsec}",
        f"quad = {x0!r} {y0!r} {gap!r}  {-x0!r} {y0!r} {gap!r}  {-x0!r} {-y0!r} {gap!r}  {x0!r} {-y0!r} {gap!r}  "
        f"{thick!r} 5.8e7 {qx} {qy}",
        "port = 1 2 50", f"frequency = {f0!r} {f1!r} {ndiv}", "end", ""])


def full_stack_deck(grid: int = 48, ndivt: int = 4, diel: tuple = (20, 20, 4), gpz: float = -1e-4,
                    f0: float = 2e7, f1: float = 3e9, nd: int = 149) -> str:
    """The full-stack configuration: a copper plate 60 x 60 mm, 35 um thick, at z = 1.6 mm on a
    grid x grid mesh split into ndivt volume layers, over an FR4-like brick (eps_r 4.4,
    60 x 60 x 1.6 mm, diel cells) that fills the gap to a ground plane, port from the plate's
    corner node to the ground node, sweep f0 .. f1 in nd + 1 points (20 MHz steps by default).

    The ground plane sits gpz (default 0.1 mm) below the brick's bottom face: with the brick on
    the image plane itself, its bottom charge cells coincide with their own images, their rows of
    P vanish and the reference's pot_fill stops on a singular matrix (potential.c:228-233)."""
    da, db, dt = diel
    lines = ["OpenPEEC 1 0", f"title = full stack {grid}x{grid}x{ndivt} over brick {da}x{db}x{dt}",
             "capacitance = 1", "retardation = 1", f"groundplane = {gpz!r}",
             "node = 1 0 0 0.0016",
             f"plate = 0 0 0.0016 0.06 0 0 0 0.06 0 3.5e-5 5.8e7 {grid} {grid} {ndivt}",
             f"dielectric = 0 0 0 0.06 0 0 0 0.06 0 0.0016 4.4 {da} {db} {dt}",
             "port = 1 0 50",
             f"frequency = {f0!r} {f1!r} {nd}",
             "end", ""]
    return "\n".join(lines)
