"""Seeded synthetic inputs in pampa's `.pmp` format.

Each generator writes a self-contained case directory (input.pmp, mesh.pmp,
one file per material) that the reference's reader format accepts unchanged,
so the product and the oracle read the very same files.

Random numbers: SplitMix64 with u = (x >> 11) * 2^-53 (a fixed, portable
mapping; no library distribution is involved), so every seed gives the same
bytes on every platform.

Workloads:
  slab_1d           1-D slab, 100 cm, 200 cells, 2 groups, 3 zones, vacuum ends
  quarter_core_2d   170 x 170 cm, 1 cm cells, 17 x 17 blocks, 4 materials, 2 groups
  hex_sn_3d         hexagonal lattice of triangular prisms, 4 groups, S_N
  batched_2d        one instance of the batched diffusion family (side, groups drawn)
  transient_core_2d small quarter core with time-step data (velocities, spectra, precursors)
  feedback_core_2d  quarter core whose fuels have temperature-dependent cross sections
"""
from __future__ import annotations

import math
import os

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def u(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def uniform(self, a: float, b: float) -> float:
        return a + (b - a) * self.u()

    def choice(self, n: int) -> int:
        return min(int(self.u() * n), n - 1)


def _fmt(v: float) -> str:
    return repr(float(v))


def write_material(path: str, G: int, sigt, sigs, D=None, nusf=None, chi=None, chid=None, vel=None,
                   lam=None, beta=None):
    """One material file; chid (delayed spectrum), vel (neutron velocity) and a
    precursor block (decay constants lam, fractions beta) are written when given."""
    lines = ["nuclear-data {", f"   energy-groups {G}", "   sigma-total",
             "   " + " ".join(_fmt(v) for v in sigt)]
    if nusf is not None:
        lines += ["   nu-sigma-fission", "   " + " ".join(_fmt(v) for v in nusf)]
    lines += ["   sigma-scattering"] + ["   " + " ".join(_fmt(v) for v in row) for row in sigs]
    if D is not None:
        lines += ["   diffusion-coefficient", "   " + " ".join(_fmt(v) for v in D)]
    if chi is not None:
        lines += ["   fission-spectrum", "   " + " ".join(_fmt(v) for v in chi)]
    if chid is not None:
        lines += ["   fission-spectrum-delayed", "   " + " ".join(_fmt(v) for v in chid)]
    if vel is not None:
        lines += ["   neutron-velocity", "   " + " ".join(_fmt(v) for v in vel)]
    lines += ["}"]
    if lam is not None:
        lines += ["precursor-data {", f"   precursor-groups {len(lam)}",
                  "   lambda", "   " + " ".join(_fmt(v) for v in lam),
                  "   beta", "   " + " ".join(_fmt(v) for v in beta), "}"]
    if nusf is not None:
        lines += ["fuel 1"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def two_group_material(rng: SplitMix64, fuel: bool, absorber_factor: float = 1.0):
    """One draw of the 2-group ranges: D_g ~ U[0.4,1.6], Sa1 ~ U[0.005,0.012],
    Sa2 ~ U[0.05,0.12] (fuel) or U[0.005,0.02] (reflector), S12 ~ U[0.015,0.03],
    nuSf = (U[0.004,0.008], U[0.08,0.14]) for fuel, chi = (1, 0).
    Draw order: D1, D2, Sa1, Sa2, S12, nuSf1, nuSf2."""
    D = [rng.uniform(0.4, 1.6), rng.uniform(0.4, 1.6)]
    sa1 = rng.uniform(0.005, 0.012)
    sa2 = rng.uniform(0.05, 0.12) if fuel else rng.uniform(0.005, 0.02)
    s12 = rng.uniform(0.015, 0.03)
    sa2 *= absorber_factor
    m = dict(G=2, sigt=[sa1 + s12, sa2], sigs=[[0.0, s12], [0.0, 0.0]], D=D)
    if fuel:
        m["nusf"] = [rng.uniform(0.004, 0.008), rng.uniform(0.08, 0.14)]
        m["chi"] = [1.0, 0.0]
    return m


def multigroup_material(rng: SplitMix64, G: int, fuel: bool, absorber_factor: float = 1.0):
    """G-group extension of the 2-group ranges (reduces to them for G = 2):
    group g has fraction t = g/(G-1) from fast to thermal; absorption and nu-fission
    ranges are interpolated geometrically between the fast and thermal ranges,
    downscatter g->g+1 ~ U[0.015,0.03] and g->g+2 ~ 0.1 x U[0.015,0.03],
    chi = (1, 0, ...) for G = 2 and (0.7, 0.3, 0, ...) beyond."""
    if G == 2:
        return two_group_material(rng, fuel, absorber_factor)

    def geo(a, b, t):
        return a * (b / a) ** t

    D, sa, nusf = [], [], []
    sigs = [[0.0] * G for _ in range(G)]
    for g in range(G):
        t = g / (G - 1)
        D.append(rng.uniform(0.4, 1.6))
        lo = geo(0.005, 0.05 if fuel else 0.005, t)
        hi = geo(0.012, 0.12 if fuel else 0.02, t)
        sa.append(rng.uniform(lo, hi) * (absorber_factor if g == G - 1 else 1.0))
        if g + 1 < G:
            sigs[g][g + 1] = rng.uniform(0.015, 0.03)
        if g + 2 < G:
            sigs[g][g + 2] = 0.1 * rng.uniform(0.015, 0.03)
        if fuel:
            nusf.append(rng.uniform(geo(0.004, 0.08, t), geo(0.008, 0.14, t)))
    sigt = [sa[g] + sum(sigs[g]) for g in range(G)]
    m = dict(G=G, sigt=sigt, sigs=sigs, D=D)
    if fuel:
        m["nusf"] = nusf
        m["chi"] = [0.7, 0.3] + [0.0] * (G - 2)
    return m


def perturb(m: dict, rng: SplitMix64, rel: float):
    """Scale every cross section and diffusion coefficient by an independent
    U[1-rel, 1+rel] factor (draw order: sigt, sigs row-major, D, nusf)."""
    out = dict(m)
    out["sigt"] = [v * rng.uniform(1 - rel, 1 + rel) for v in m["sigt"]]
    out["sigs"] = [[v * rng.uniform(1 - rel, 1 + rel) for v in row] for row in m["sigs"]]
    if "D" in m:
        out["D"] = [v * rng.uniform(1 - rel, 1 + rel) for v in m["D"]]
    if "nusf" in m:
        out["nusf"] = [v * rng.uniform(1 - rel, 1 + rel) for v in m["nusf"]]
    # total stays >= scattering out of the group (physical absorption >= 0)
    out["sigt"] = [max(t, sum(row)) for t, row in zip(out["sigt"], out["sigs"])]
    return out


def perturb_tables(sigt, sigs, D, nusf, seed: int, rel: float):
    """Cross-section tables of a served request: every total, scattering,
    diffusion and nu-fission value of every material scaled by its own
    U[1-rel, 1+rel] factor, drawn material by material in the order sigt[g],
    sigs[g][g'] (row-major), D[g], nusf[g]; then sigt[g] = max(sigt[g],
    sum_g' sigs[g][g']).  Seed 0 leaves the tables as they are.  Arrays are
    (nm, G) and (nm, G, G); returns new arrays."""
    import numpy as np
    sigt, sigs = np.array(sigt, dtype=float), np.array(sigs, dtype=float)
    D = None if D is None else np.array(D, dtype=float)
    nusf = np.array(nusf, dtype=float)
    if seed == 0 or rel == 0.0:
        return sigt, sigs, D, nusf
    rng = SplitMix64(seed)
    lo, hi = 1.0 - rel, 1.0 + rel
    nm, G = sigt.shape
    for m in range(nm):
        for g in range(G):
            sigt[m, g] = float(sigt[m, g]) * rng.uniform(lo, hi)
        for g in range(G):
            for h in range(G):
                sigs[m, g, h] = float(sigs[m, g, h]) * rng.uniform(lo, hi)
        if D is not None:
            for g in range(G):
                D[m, g] = float(D[m, g]) * rng.uniform(lo, hi)
        for g in range(G):
            nusf[m, g] = float(nusf[m, g]) * rng.uniform(lo, hi)
        for g in range(G):
            tot = 0.0
            for h in range(G):
                tot += float(sigs[m, g, h])
            sigt[m, g] = max(float(sigt[m, g]), tot)
    return sigt, sigs, D, nusf


def _write_input(d: str, names, solver_lines):
    lines = ["mesh {} mesh.pmp".format(solver_lines[0])]
    for n in names:
        lines.append(f"material {n} {n}.pmp")
    lines += [f"solver {solver_lines[1]} {{"] + ["   " + s for s in solver_lines[2:]] + ["}"]
    with open(os.path.join(d, "input.pmp"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _materials_line(vals, per_line=40):
    return "\n".join(" ".join(str(v) for v in vals[k:k + per_line])
                     for k in range(0, len(vals), per_line))


def slab_1d(d: str, seed: int = 0, cells: int = 200, length: float = 100.0):
    """1-D slab: reflector [0,20), fuel [20,80), reflector [80,100], vacuum ends.
    Materials drawn in the order left reflector, fuel, right reflector."""
    os.makedirs(d, exist_ok=True)
    rng = SplitMix64(seed)
    mats = [("reflector-left", two_group_material(rng, False)),
            ("fuel", two_group_material(rng, True)),
            ("reflector-right", two_group_material(rng, False))]
    for n, m in mats:
        write_material(os.path.join(d, n + ".pmp"), **m)
    h = length / cells
    layout = []
    for i in range(cells):
        x = (i + 0.5) * h
        layout.append(1 if x < 20.0 else (2 if x < 80.0 else 3))
    with open(os.path.join(d, "mesh.pmp"), "w") as fh:
        fh.write(f"dx -{cells}\n{_fmt(h)}\n\nbc -x vacuum\nbc +x vacuum\n\n")
        fh.write(f"materials {cells}\n{_materials_line(layout)}\n")
    _write_input(d, [n for n, _ in mats], ["cartesian", "diffusion", "energy-groups 2"])
    return d


def block_layout(rng: SplitMix64, side: int, nblocks: int = 17):
    """Block-structured material map: nblocks x nblocks blocks with edges at
    round(k*side/nblocks); outer ring of blocks = reflector (material 4), inner
    blocks uniform over {fuel A (1), fuel B (2), fuel+absorber (3), reflector (4)}.
    Returns the row-major (y outer) 1-based map."""
    edges = [round(k * side / nblocks) for k in range(nblocks + 1)]
    bmat = [[4] * nblocks for _ in range(nblocks)]
    for bj in range(nblocks):
        for bi in range(nblocks):
            if bi < nblocks - 1 and bj < nblocks - 1:
                bmat[bj][bi] = 1 + rng.choice(4)
    cell_block = []
    for k in range(nblocks):
        cell_block += [k] * (edges[k + 1] - edges[k])
    return [bmat[cell_block[j]][cell_block[i]] for j in range(side) for i in range(side)]


def quarter_core_2d(d: str, seed: int = 1, side: int = 170, pitch: float = 1.0, G: int = 2,
                    perturb_rel: float = 0.0):
    """2-D quarter core: side x side cells of `pitch` cm, 17 x 17 blocks; the
    reflective symmetry planes are x = 0 and y = 0 (the outer ring of blocks, at
    high x or high y, is reflector); vacuum on +x/+y.  Materials drawn in the order
    fuel A, fuel B, fuel+absorber (Sa_thermal x U[1.5,2.5]), reflector, then the
    block map, then (if perturb_rel > 0) per-material perturbations."""
    os.makedirs(d, exist_ok=True)
    rng = SplitMix64(seed)
    fa = multigroup_material(rng, G, True)
    fb = multigroup_material(rng, G, True)
    fac = rng.uniform(1.5, 2.5)
    fr = multigroup_material(rng, G, True, absorber_factor=fac)
    rf = multigroup_material(rng, G, False)
    layout = block_layout(rng, side)
    mats = [("fuel-a", fa), ("fuel-b", fb), ("fuel-absorber", fr), ("reflector", rf)]
    if perturb_rel > 0.0:
        mats = [(n, perturb(m, rng, perturb_rel)) for n, m in mats]
    for n, m in mats:
        write_material(os.path.join(d, n + ".pmp"), **m)
    with open(os.path.join(d, "mesh.pmp"), "w") as fh:
        fh.write(f"dx -{side}\n{_fmt(pitch)}\n\ndy -{side}\n{_fmt(pitch)}\n\n")
        fh.write("bc -x reflective\nbc +x vacuum\nbc -y reflective\nbc +y vacuum\n\n")
        fh.write(f"materials {side * side}\n{_materials_line(layout)}\n")
    _write_input(d, [n for n, _ in mats], ["cartesian", "diffusion", f"energy-groups {G}"])
    return d


def transient_core_2d(d: str, seed: int = 7, side: int = 24, G: int = 3, dt: float = 0.01, steps: int = 2):
    """Quarter core of side x side 1 cm cells (as quarter_core_2d) whose
    materials carry time-step data: every material a neutron velocity per group
    (U[1e5, 2e7] fast ... U[2e3, 1e4] thermal, geometric in the group), fuel
    materials prompt spectrum (chi of the group set), a delayed spectrum
    U-drawn and normalised (different from the prompt one), and two precursor
    groups (lambda U[0.01, 3], beta U[5e-4, 3e-3]).  The input asks for
    `steps` time steps of `dt` s.  Draw order: the quarter-core draws, then per
    material velocities, delayed spectrum, lambda, beta."""
    quarter_core_2d(d, seed=seed, side=side, G=G)
    rng = SplitMix64(seed ^ 0x5DEECE66D)
    names = ["fuel-a", "fuel-b", "fuel-absorber", "reflector"]
    # re-read the drawn base materials by redrawing them with the same seed
    base = SplitMix64(seed)
    mats = [multigroup_material(base, G, True), multigroup_material(base, G, True)]
    fac = base.uniform(1.5, 2.5)
    mats += [multigroup_material(base, G, True, absorber_factor=fac), multigroup_material(base, G, False)]
    for n, m in zip(names, mats):
        vel = [rng.uniform(1e5, 2e7) * (2e3 / 1e5) ** (g / (G - 1)) for g in range(G)]
        extra = dict(vel=vel)
        if "nusf" in m:
            raw = [rng.uniform(0.1, 1.0) for _ in range(G)]
            extra["chid"] = [v / sum(raw) for v in raw]
            extra["lam"] = [rng.uniform(0.01, 3.0) for _ in range(2)]
            extra["beta"] = [rng.uniform(5e-4, 3e-3) for _ in range(2)]
        write_material(os.path.join(d, n + ".pmp"), **m, **extra)
    with open(os.path.join(d, "input.pmp"), "a") as fh:
        fh.write(f"\ndt -{steps}\n{_fmt(dt)}\n")
    return d


def feedback_core_2d(d: str, seed: int = 9, side: int = 30, temps=(300.0, 600.0, 900.0)):
    """Quarter core (as quarter_core_2d, 2 groups) whose two plain fuels carry
    temperature-dependent data: one set per temperature in `temps`, the
    thermal absorption growing by U[2, 6] % and the fast one by U[0.5, 2] %
    per 300 K from the drawn values, nu-fission unchanged (a Doppler-like
    trend).  The absorber fuel and the reflector stay constant."""
    quarter_core_2d(d, seed=seed, side=side, G=2)
    base = SplitMix64(seed)
    fa = multigroup_material(base, 2, True)
    fb = multigroup_material(base, 2, True)
    rng = SplitMix64(seed ^ 0xFEEDBAC)
    for name, m in (("fuel-a", fa), ("fuel-b", fb)):
        k2, k1 = rng.uniform(0.02, 0.06), rng.uniform(0.005, 0.02)
        blocks = []
        for j, T in enumerate(temps):
            x = (T - temps[0]) / 300.0
            sa1 = (m["sigt"][0] - m["sigs"][0][1]) * (1.0 + k1 * x)
            sa2 = m["sigt"][1] * (1.0 + k2 * x)
            st = [sa1 + m["sigs"][0][1], sa2]
            blocks += ["   nuclear-data {", "      energy-groups 2", "      sigma-total",
                       "      " + " ".join(_fmt(v) for v in st), "      nu-sigma-fission",
                       "      " + " ".join(_fmt(v) for v in m["nusf"]), "      sigma-scattering"]
            blocks += ["      " + " ".join(_fmt(v) for v in row) for row in m["sigs"]]
            blocks += ["      diffusion-coefficient", "      " + " ".join(_fmt(v) for v in m["D"]),
                       "      fission-spectrum", "      " + " ".join(_fmt(v) for v in m["chi"]), "   }"]
        lines = ["nuclear-data-set {", f"   temperature {len(temps)}", "   " + " ".join(_fmt(t) for t in temps)]
        lines += blocks + ["}", "fuel 1"]
        with open(os.path.join(d, name + ".pmp"), "w") as fh:
            fh.write("\n".join(lines) + "\n")
    return d


BATCH_SIDES = (32, 48, 64, 96, 128, 192, 256)
BATCH_GROUPS = (2, 4, 8, 16)


def batched_instance(d: str, index: int, base_seed: int = 3_000_000, perturb_rel: float = 0.05):
    """Instance `index` of the batched family: seed base_seed + index draws the
    side from BATCH_SIDES and the group count from BATCH_GROUPS, then a quarter
    core of that size (1 cm cells) with every cross section perturbed by an
    independent U[1-rel, 1+rel] factor."""
    rng = SplitMix64(base_seed + index)
    side = BATCH_SIDES[rng.choice(len(BATCH_SIDES))]
    G = BATCH_GROUPS[rng.choice(len(BATCH_GROUPS))]
    quarter_core_2d(d, seed=rng.next(), side=side, G=G, perturb_rel=perturb_rel)
    return side, G


# ---------------------------------------------------------------------------
# hexagonal lattice of triangular prisms (S_N workload)
# ---------------------------------------------------------------------------

def hex_centers(rings: int, pitch: float):
    """Centres of a hexagonal lattice with `rings` rings around the centre
    (1 + 3 r (r + 1) hexagons), flat-to-flat `pitch`, ordered ring by ring."""
    out = [(0.0, 0.0, 0)]
    # axial directions for flat-topped... use pointy-side neighbours at 60 deg steps
    dirs = [(pitch * math.cos(math.radians(30 + 60 * k)), pitch * math.sin(math.radians(30 + 60 * k)))
            for k in range(6)]
    for r in range(1, rings + 1):
        x, y = dirs[4][0] * r, dirs[4][1] * r
        for side in range(6):
            for _ in range(r):
                out.append((x, y, r))
                x += dirs[side][0]
                y += dirs[side][1]
    return out


def hex_sn_3d(d: str, seed: int = 2, rings: int = 10, pitch: float = 1.6, layers: int = 40,
              dz: float = 5.0, G: int = 4, order: int = 6, perturb_rel: float = 0.0,
              base_seed: int | None = None):
    """Hexagonal lattice of 1 + 3 rings (rings + 1) hexagonal columns, each split
    into 6 sector triangles x 4 midpoint sub-triangles (24 CCW triangles), extruded
    into `layers` layers of `dz` cm.  G groups, downscatter-only isotropic
    scattering; vacuum everywhere; S_N `order`; pure upwind, LS off.

    Cross sections (builder choice): per material and group Sigma_t ~ U[0.2,1.2],
    c ~ U[0.5,0.95]; Sigma_s,total = c Sigma_t of which a fraction U[0.5,0.9]
    stays in the group and the rest is split equally over the lower groups (all
    of it stays in the last group); fuel: nuSf_g = 2.43 x U[0.3,0.6] x (1-c) Sigma_t,
    chi = (0.75, 0.25, 0, ...).  Materials: 3 fuels and a reflector; columns of
    the outer ring are reflector, the others draw a fuel uniformly.
    `base_seed` (if given) draws the materials, `seed` the column map and the
    perturbation, so several cases can share a base."""
    os.makedirs(d, exist_ok=True)
    rng_m = SplitMix64(base_seed if base_seed is not None else seed)
    rng = SplitMix64(seed) if base_seed is not None else rng_m

    def mat(fuel):
        sigt, sigs, nusf = [], [[0.0] * G for _ in range(G)], []
        for g in range(G):
            st = rng_m.uniform(0.2, 1.2)
            c = rng_m.uniform(0.5, 0.95)
            ss = c * st
            if g == G - 1:
                sigs[g][g] = ss
            else:
                keep = rng_m.uniform(0.5, 0.9)
                sigs[g][g] = keep * ss
                for g2 in range(g + 1, G):
                    sigs[g][g2] = (1 - keep) * ss / (G - 1 - g)
            sigt.append(st)
            if fuel:
                nusf.append(2.43 * rng_m.uniform(0.3, 0.6) * (1 - c) * st)
        m = dict(G=G, sigt=sigt, sigs=sigs)
        if fuel:
            m["nusf"] = nusf
            m["chi"] = [0.75, 0.25] + [0.0] * (G - 2)
        return m

    mats = [("fuel-1", mat(True)), ("fuel-2", mat(True)), ("fuel-3", mat(True)),
            ("reflector", mat(False))]
    if perturb_rel > 0.0:
        mats = [(n, perturb(m, rng, perturb_rel)) for n, m in mats]
    for n, m in mats:
        write_material(os.path.join(d, n + ".pmp"), **m)
    centers = hex_centers(rings, pitch)
    R = pitch / math.sqrt(3.0)  # circumradius
    pts: dict = {}
    plist = []

    def pid(x, y):
        key = (round(x, 9), round(y, 9))
        if key not in pts:
            pts[key] = len(plist)
            plist.append((x, y))
        return pts[key]

    tris = []
    colmat = []
    for (cx, cy, ring) in centers:
        m = 4 if ring == rings else 1 + rng.choice(3)
        corners = [(cx + R * math.cos(math.radians(60 * k)), cy + R * math.sin(math.radians(60 * k)))
                   for k in range(6)]
        for k in range(6):
            a = (cx, cy)
            b = corners[k]
            c = corners[(k + 1) % 6]
            ab = ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)
            bc = ((b[0] + c[0]) / 2, (b[1] + c[1]) / 2)
            ca = ((c[0] + a[0]) / 2, (c[1] + a[1]) / 2)
            for t in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)):
                tris.append([pid(*p) for p in t])
                colmat.append(m)
    with open(os.path.join(d, "mesh.pmp"), "w") as fh:
        fh.write(f"points {len(plist)}\n")
        for (x, y) in plist:
            fh.write(f"{x!r} {y!r}\n")
        fh.write(f"cells {len(tris)} {3 * len(tris)}\n")
        for t in tris:
            fh.write(f"{t[0]} {t[1]} {t[2]}\n")
        fh.write(f"\ndz -{layers}\n{_fmt(dz)}\n\n")
        fh.write("boundary exterior 0\n\n")
        fh.write("bc exterior vacuum\nbc -z vacuum\nbc +z vacuum\n\n")
        allm = colmat * layers
        fh.write(f"materials {len(allm)}\n{_materials_line(allm)}\n")
    _write_input(d, [n for n, _ in mats], ["unstructured", "sn", f"energy-groups {G}",
                                           f"order {order}", "mixed-face-interpolation 1.0",
                                           "least-squares-boundary-interpolation 0"])
    return d
