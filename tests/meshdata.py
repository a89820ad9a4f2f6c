"""Test helpers: write a mesh in the partitioned-mesh data format
(PartitionedMesh.cxx:4-263, as Mesh::writeData writes it, Mesh.cxx:408-568) and
build inputs that use boundary-condition and split materials."""
import os

import numpy as np


def write_partitioned(path, mesh, fmt=repr, ghost=1):
    """One domain: `ghost` ghost cells that no face leads to (the reader wants
    at least one).  fmt formats every real number (repr: exact; the
    reference's writer uses fixed 3 decimals)."""
    C = mesh.C
    nf = [int(n) for n in mesh.nfaces]
    F = sum(nf)
    L = ["points 1", "0.0 0.0 0.0", "", f"cells {C} {ghost} {C + ghost}", f"cell-points {C} {C}"]
    L += ["0"] * C + ["", f"cell-volumes {C}"] + [fmt(float(v)) for v in mesh.vol]
    cen = list(mesh.cen) + [mesh.cen[0]] * ghost
    L += ["", f"cell-centroids {C + ghost}"] + [" ".join(fmt(float(x)) for x in c) for c in cen]
    mats = list(mesh.cmat) + [mesh.cmat[0]] * ghost
    L += ["", f"cell-materials {C + ghost}"] + [str(int(m) + 1) for m in mats]
    L += ["", f"cell-global-indices {C + ghost}"] + [str(i) for i in range(C + ghost)]
    L += ["", f"faces {C}"] + [str(n) for n in nf]
    L += ["", f"face-areas {C} {F}"] + [" ".join(fmt(float(mesh.farea[i, f])) for f in range(nf[i]))
                                         for i in range(C)]
    L += ["", f"face-centroids {C} {3 * F}"] + [" ".join(fmt(float(x)) for x in mesh.fcen[i, f])
                                                 for i in range(C) for f in range(nf[i])]
    L += ["", f"face-normals {C} {3 * F}"] + [" ".join(fmt(float(x)) for x in mesh.fnorm[i, f])
                                               for i in range(C) for f in range(nf[i])]
    L += ["", f"face-neighbors {C} {F}"] + [" ".join(str(int(mesh.fnb[i, f])) for f in range(nf[i]))
                                             for i in range(C)]
    L += [""] + [f"boundary {b}" for b in mesh.boundaries]
    names = {1: "vacuum", 2: "reflective"}
    for b, (t, v) in sorted(mesh.bcs.items()):
        L.append(f"bc {mesh.boundaries[b]} " + (names[t] if t in names else f"robin {fmt(float(v))}"))
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def fixed3(x):
    return f"{x:.3f}"


def with_wall(d_src, d, wall_name="wall"):
    """The quarter core of d_src (side x side, reflective -x/-y, vacuum +x/+y)
    grown by one column at +x and one row at +y of a boundary-condition
    material `wall`; the mesh's +x/+y now touch wall cells only, and the solver
    puts vacuum on the wall boundary."""
    os.makedirs(d, exist_ok=True)
    mesh = open(os.path.join(d_src, "mesh.pmp")).read().split("\n")
    side = int(mesh[0].split()[1].lstrip("-"))
    i = mesh.index(next(x for x in mesh if x.startswith("materials")))
    vals = " ".join(mesh[i + 1:]).split()
    grid = np.array([int(v) for v in vals]).reshape(side, side)
    n = side + 1
    big = np.full((n, n), 5)
    big[:side, :side] = grid
    hdr = [f"dx -{n}", mesh[1], "", f"dy -{n}", mesh[4], "",
           "bc -x reflective", "bc +x vacuum", "bc -y reflective", "bc +y vacuum", ""]
    lines = hdr + [f"materials {n * n}"] + [" ".join(str(v) for v in big.reshape(-1))]
    with open(os.path.join(d, "mesh.pmp"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    inp = open(os.path.join(d_src, "input.pmp")).read().split("\n")
    out = []
    for ln in inp:
        if ln.startswith("material "):
            name = ln.split()[1]
            out.append(f"material {name} {os.path.join(d_src, name + '.pmp')}")
        elif ln.startswith("solver"):
            out.append(f"material {wall_name} {{")
            out += ["   nuclear-data {", "      energy-groups 2", "      sigma-total", "      1.0 1.0",
                    "      sigma-scattering", "      0.0 0.0", "      0.0 0.0", "      diffusion-coefficient",
                    "      1.0 1.0", "   }", "   bc 1", "}"]
            out.append(ln)
            out.append(f"   bc {wall_name} vacuum")
            out.append("   bc -x reflective")
            out.append("   bc -y reflective")
        else:
            out.append(ln)
    with open(os.path.join(d, "input.pmp"), "w") as fh:
        fh.write("\n".join(out) + "\n")
    return d
