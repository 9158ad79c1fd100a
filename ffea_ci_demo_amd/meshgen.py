"""Synthetic FFEA blob meshes: jittered cubes split into 6 quadratic tetrahedra.

A Kuhn split of each unit cube (the six tetrahedra that share the cube's main
diagonal) with 10-node tetrahedra as FFEA's ``.top`` files hold them: corner
nodes 0-3, then midside nodes on edges 01, 02, 03, 12, 13, 23 (the slot order of
tetra_element_linear.cpp:553-577). Nodes and elements are written surface-first
as Blob::load_nodes / Blob::load_topology expect (Blob.cpp:1911-2085).

Meshes are returned as numpy arrays in file units (edge = 1) and can be written
as the reference's ``.ffea`` + ``.node/.top/.surf/.mat/.stokes/.vdw/.pin`` set.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional

import numpy as np

# slot pairs of the six midside nodes (tetra_element_linear.cpp:553-577)
EDGE_SLOTS = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))

# Kuhn split: the path 000 -> e_p0 -> e_p0+e_p1 -> 111 for each axis permutation
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_PERM_SIGN = (1, -1, -1, 1, 1, -1)


@dataclasses.dataclass
class Mesh:
    pos: np.ndarray            # (Nt, 3) float64, file units
    top: np.ndarray            # (Ne, 10) int32
    num_surface_nodes: int
    num_surface_elements: int
    n: int                     # cubes per side
    surf_faces: np.ndarray     # (F, 4) int32: element, n0, n1, n2 (midside sub-triangles)
    is_linear: np.ndarray      # (Nt,) bool

    @property
    def num_nodes(self) -> int:
        return self.pos.shape[0]

    @property
    def num_elements(self) -> int:
        return self.top.shape[0]

    @property
    def num_linear_nodes(self) -> int:
        return int(self.is_linear.sum())


def _kuhn_cube_tets():
    """6 tets of the unit cube as (4, 3) corner offsets, all with positive det."""
    tets = []
    for perm, sgn in zip(_PERMS, _PERM_SIGN):
        e = np.eye(3, dtype=np.int64)
        v0 = np.zeros(3, np.int64)
        v1 = e[perm[0]]
        v2 = v1 + e[perm[1]]
        v3 = np.ones(3, np.int64)
        t = [v0, v1, v2, v3] if sgn > 0 else [v0, v1, v3, v2]
        tets.append(np.stack(t))
    return np.stack(tets)  # (6, 4, 3)


def kuhn_mesh(n: int, jitter: float = 0.1, seed: int = 0, flatten_frac: float = 0.0,
              flatten_aspect: float = 20.0) -> Mesh:
    """n^3 cubes, 6n^3 quadratic tets, interior corner nodes jittered uniformly by
    +-jitter (per axis, in edge units) from `seed`; boundary nodes not jittered.

    flatten_frac > 0 squashes about that fraction of the cubes to height
    1/flatten_aspect (see flatten_layers)."""
    n = int(n)
    N1 = n + 1
    nl = N1 ** 3
    g = np.arange(N1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    grid = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1).astype(np.float64)  # (nl, 3)

    cube = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    offs = _kuhn_cube_tets()  # (6,4,3)
    corners = cube[:, None, None, :] + offs[None, :, :, :]  # (nc,6,4,3)
    corners = corners.reshape(-1, 4, 3)
    lin = (corners[..., 0] * N1 + corners[..., 1]) * N1 + corners[..., 2]  # (ne,4)
    ne = lin.shape[0]

    # unique edges -> midside nodes
    ea = np.stack([lin[:, a] for a, b in EDGE_SLOTS], axis=1)
    eb = np.stack([lin[:, b] for a, b in EDGE_SLOTS], axis=1)
    lo = np.minimum(ea, eb).astype(np.int64)
    hi = np.maximum(ea, eb).astype(np.int64)
    key = lo * nl + hi
    ukey, inv = np.unique(key.ravel(), return_inverse=True)
    nmid = ukey.shape[0]
    mid_lo = ukey // nl
    mid_hi = ukey % nl
    top_raw = np.concatenate([lin, nl + inv.reshape(ne, 6)], axis=1)  # nodes: linear then midside

    # positions (edge units): jitter interior linear nodes
    pos_lin = grid.copy()
    boundary_lin = np.any((grid == 0) | (grid == n), axis=1)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(nl, 3))
    jit[boundary_lin] = 0.0
    pos_lin += jit
    if flatten_frac > 0:
        pos_lin = flatten_layers(pos_lin, grid, n, flatten_frac, flatten_aspect, rng)
    pos_mid = 0.5 * (pos_lin[mid_lo] + pos_lin[mid_hi])
    pos_raw = np.concatenate([pos_lin, pos_mid], axis=0)

    # surface classification on the exact grid (doubled coords for midside nodes)
    g2_lin = 2 * grid.astype(np.int64)
    g2_mid = (g2_lin[mid_lo] + g2_lin[mid_hi]) // 2   # doubled midpoint coordinates
    g2 = np.concatenate([g2_lin, g2_mid], axis=0)
    surf_node = np.any((g2 == 0) | (g2 == 2 * n), axis=1)
    is_linear_raw = np.zeros(nl + nmid, bool)
    is_linear_raw[:nl] = True

    order = np.concatenate([np.nonzero(surf_node)[0], np.nonzero(~surf_node)[0]])
    new_index = np.empty_like(order)
    new_index[order] = np.arange(order.shape[0])
    pos = pos_raw[order]
    is_linear = is_linear_raw[order]
    top = new_index[top_raw]
    num_surface_nodes = int(surf_node.sum())

    # surface elements: a face (3 corners) on one boundary plane
    cg = corners  # (ne,4,3) integer grid coords
    face_slots = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    elem_surf = np.zeros(ne, bool)
    faces = []
    for fs in face_slots:
        c = cg[:, fs, :]  # (ne,3,3)
        on = np.zeros(ne, bool)
        for ax in range(3):
            for val in (0, n):
                on |= np.all(c[:, :, ax] == val, axis=1)
        elem_surf |= on
        faces.append((fs, on))
    eorder = np.concatenate([np.nonzero(elem_surf)[0], np.nonzero(~elem_surf)[0]])
    enew = np.empty_like(eorder)
    enew[eorder] = np.arange(ne)
    top = top[eorder].astype(np.int32)
    num_surface_elements = int(elem_surf.sum())

    # .surf: 4 midside sub-triangles per boundary triangle
    slot_of_edge = {frozenset(p): 4 + k for k, p in enumerate(EDGE_SLOTS)}
    sf = []
    for fs, on in faces:
        idx = np.nonzero(on)[0]
        if idx.size == 0:
            continue
        a, b, c = fs
        mab, mbc, mca = slot_of_edge[frozenset((a, b))], slot_of_edge[frozenset((b, c))], slot_of_edge[frozenset((c, a))]
        T = top[enew[idx]]
        el = enew[idx]
        for tri in ((a, mab, mca), (mab, b, mbc), (mca, mbc, c), (mab, mbc, mca)):
            sf.append(np.stack([el, T[:, tri[0]], T[:, tri[1]], T[:, tri[2]]], axis=1))
    surf_faces = np.concatenate(sf, axis=0).astype(np.int32) if sf else np.zeros((0, 4), np.int32)
    return Mesh(pos=pos, top=top, num_surface_nodes=num_surface_nodes,
                num_surface_elements=num_surface_elements, n=n, surf_faces=surf_faces,
                is_linear=is_linear)


def flatten_layers(pos_lin, grid, n, frac, aspect, rng):
    """Squash about frac*n^3 cubes to height 1/aspect (edge units).

    The squashed cubes form a square window of node columns in one or more interior
    cube layers L (never two adjacent layers; 1 <= L <= n-2). Inside the window node
    plane L+1 is put on its grid point lowered to z = L + 1/aspect, and node plane L on
    its grid points (z = L) inside the window and one column around it: these nodes
    carry no jitter, so no tet of a squashed cube or of the wedges round it changes its
    orientation. A squashed cube is a regular unit square prism of height 1/aspect
    split into 6 tets, each of volume 1/(6 aspect) (no tet is close to degenerate), with
    longest edge over smallest height between 1.4 and 2 times `aspect`. Cubes bordering
    the window become wedges; the cubes above it become about 2 - 1/aspect tall."""
    pos = pos_lin.copy()
    target = frac * n ** 3  # cubes to squash
    n_in = max(1, n - 2)     # interior layers 1..n-2
    per_layer = n * n
    m = int(min(max(1, np.ceil(target / per_layer)), (n_in + 1) // 2))
    w = int(min(n, max(1, round(np.sqrt(target / m)))))
    # m layers from 1..n-2, pairwise non-adjacent
    choices = np.arange(1, n - 1)
    layers = []
    for L in rng.permutation(choices):
        if all(abs(int(L) - q) > 1 for q in layers):
            layers.append(int(L))
        if len(layers) == m:
            break
    x0 = int(rng.integers(0, n - w + 1))
    y0 = int(rng.integers(0, n - w + 1))
    h = 1.0 / aspect
    inwin = (grid[:, 0] >= x0) & (grid[:, 0] <= x0 + w) & (grid[:, 1] >= y0) & (grid[:, 1] <= y0 + w)
    rim = (grid[:, 0] >= x0 - 1) & (grid[:, 0] <= x0 + w + 1) & (grid[:, 1] >= y0 - 1) & (grid[:, 1] <= y0 + w + 1)
    for L in sorted(layers):
        bot = rim & (grid[:, 2] == L)
        top = inwin & (grid[:, 2] == L + 1)
        pos[bot] = grid[bot]
        pos[top] = grid[top]
        pos[top, 2] = float(L) + h
    return pos


# ---------------------------------------------------------------- file output

def _w(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_blob_files(mesh: Mesh, prefix: str, material=(1500.0, 1e-3, 1e-3, 3.6e8, 2.2e9, 1.0),
                     per_element_material: Optional[np.ndarray] = None, stokes_radius: float = 0.0,
                     pinned: Optional[np.ndarray] = None):
    """Write .node/.top/.surf/.mat/.stokes/.vdw/.pin in the reference formats."""
    nt, ne = mesh.num_nodes, mesh.num_elements
    ns, nes = mesh.num_surface_nodes, mesh.num_surface_elements
    lines = ["ffea node file", f"num_nodes {nt}", f"num_surface_nodes {ns}",
             f"num_interior_nodes {nt - ns}", "surface nodes:"]
    for i in range(nt):
        if i == ns:
            lines.append("interior nodes:")
        lines.append("%.17e %.17e %.17e" % tuple(mesh.pos[i]))
    if ns == nt:
        lines.append("interior nodes:")
    _w(prefix + ".node", lines)
    lines = ["ffea topology file", f"num_elements {ne}", f"num_surface_elements {nes}",
             f"num_interior_elements {ne - nes}", "surface elements:"]
    for i in range(ne):
        if i == nes:
            lines.append("interior elements:")
        lines.append(" ".join(str(int(v)) for v in mesh.top[i]))
    if nes == ne:
        lines.append("interior elements:")
    _w(prefix + ".top", lines)
    F = mesh.surf_faces
    lines = ["ffea surface file", f"num_surface_faces {F.shape[0]}", "faces:"]
    lines += ["%d %d %d %d" % tuple(r) for r in F]
    _w(prefix + ".surf", lines)
    lines = ["ffea material params file", f"num_elements {ne}"]
    if per_element_material is None:
        row = "%.17e %.17e %.17e %.17e %.17e %.17e" % tuple(material)
        lines += [row] * ne
    else:
        lines += ["%.17e %.17e %.17e %.17e %.17e %.17e" % tuple(r) for r in per_element_material]
    _w(prefix + ".mat", lines)
    rad = np.full(nt, stokes_radius) if np.isscalar(stokes_radius) else np.asarray(stokes_radius)
    lines = ["ffea stokes radii file", f"num_nodes {nt}"] + ["%.17e" % r for r in rad]
    _w(prefix + ".stokes", lines)
    lines = ["ffea vdw file", f"num_faces {F.shape[0]}", "vdw params:"] + ["-1"] * F.shape[0]
    _w(prefix + ".vdw", lines)
    pins = np.zeros(0, np.int64) if pinned is None else np.asarray(pinned)
    lines = ["ffea pinned nodes file", f"num_pinned_nodes {pins.size}", "pinned nodes:"] + [str(int(p)) for p in pins]
    _w(prefix + ".pin", lines)


def write_ffea_script(path, blob_prefix, *, dt=1e-14, kT=4.141945559999999e-21, num_steps=1000,
                      check=1000, rng_seed=2024, epsilon=1e-10, max_iterations_cg=10000,
                      calc_noise=1, calc_stokes=0, stokes_visc=1e-3, scale=1e-9, calc_ctforces=0,
                      ctforces_fname=None, centroid=None, num_blobs=1, centroids=None, calc_springs=0,
                      springs_fname=None, solver="CG_nomass"):
    """A `.ffea` script in the reference's <param>/<system> layout
    (FFEA_input_reader.cpp:116-196; keys of SimulationParams.cpp:239-776). With num_blobs > 1
    every blob uses the same files, blob i at centroids[i] when given."""
    base = os.path.basename(blob_prefix)
    p = ["<param>", "\t<restart = 0>", f"\t<dt = {dt!r}>", f"\t<kT = {kT!r}>", f"\t<check = {check}>",
         f"\t<num_steps = {num_steps}>", f"\t<rng_seed = {rng_seed}>",
         f"\t<trajectory_out_fname = {base}_trajectory.out>",
         f"\t<measurement_out_fname = {base}_measurement.out>",
         f"\t<epsilon = {epsilon!r}>", f"\t<max_iterations_cg = {max_iterations_cg}>",
         f"\t<calc_stokes = {calc_stokes}>", f"\t<stokes_visc = {stokes_visc!r}>",
         "\t<calc_ssint = 0>", "\t<calc_steric = 0>", f"\t<calc_noise = {calc_noise}>",
         "\t<calc_kinetics = 0>", "\t<calc_es = 0>", "\t<es_N_x = 0>", "\t<es_N_y = 0>", "\t<es_N_z = 0>",
         f"\t<calc_ctforces = {calc_ctforces}>", f"\t<calc_springs = {calc_springs}>"]
    ones = ",".join(["1"] * num_blobs)
    p += [f"\t<num_blobs = {num_blobs}>", f"\t<num_conformations = ({ones})>", f"\t<num_states = ({ones})>",
          "</param>", "", "<system>"]
    for b in range(num_blobs):
        p += ["\t<blob>", "\t\t<conformation>", "\t\t\t<motion_state = DYNAMIC>",
              f"\t\t\t<nodes = {base}.node>", f"\t\t\t<topology = {base}.top>", f"\t\t\t<surface = {base}.surf>",
              f"\t\t\t<material = {base}.mat>", f"\t\t\t<stokes = {base}.stokes>", f"\t\t\t<vdw = {base}.vdw>",
              f"\t\t\t<pin = {base}.pin>", "\t\t</conformation>", f"\t\t<solver = {solver}>",
              f"\t\t<scale = {scale!r}>"]
        c = centroids[b] if centroids is not None else (centroid if b == 0 else None)
        if c is not None:
            p.append("\t\t<centroid = (%r,%r,%r)>" % tuple(c))
        p += ["\t</blob>"]
    if ctforces_fname or springs_fname:
        p += ["\t<interactions>"]
        if ctforces_fname:
            p += ["\t\t<ctforces>", f"\t\t\t<ctforces_fname = {ctforces_fname}>", "\t\t</ctforces>"]
        if springs_fname:
            p += ["\t\t<springs>", f"\t\t\t<springs_fname = {springs_fname}>", "\t\t</springs>"]
        p += ["\t</interactions>"]
    p += ["</system>"]
    _w(path, p)


def write_springs(path, springs):
    """A springs file (World::load_springs, World.cpp:3543-3643): rows (k [N/m], l [m],
    blob 0, blob 1, conformation 0, conformation 1, node 0, node 1)."""
    lines = ["ffea springs file", f"num_springs {len(springs)}", "springs:"]
    for k, l, b0, b1, n0, n1 in springs:
        lines.append(f"{k!r} {l!r} {b0} {b1} 0 0 {n0} {n1}")
    _w(path, lines)


def write_surface_ctforces(path, sets, blob_index=0):
    """A .ctforces file of linear surface forces: rows (F [N], direction x y z, face
    indices of the .surf file), the layout Blob::load_ctforces reads (Blob.cpp:2730-2806)."""
    lines = ["ffea ctforces file", f"num_ctforces {len(sets)}", f"num_linear_surface_forces {len(sets)}",
             "linear surface forces:"]
    for F, d, faces in sets:
        lines.append(f"{F!r} {d[0]!r} {d[1]!r} {d[2]!r} {blob_index} 0 " + " ".join(str(int(f)) for f in faces))
    _w(path, lines)


def write_ctforces(path, node_forces, blob_index=0):
    """A .ctforces file of linear node forces: rows (F [N], direction x y z, node)
    (format of docs/2-UserManual/7-FFEAFiles.md:210-272)."""
    lines = ["ffea ctforces file", f"num_ctforces {len(node_forces)}", f"num_linear_forces {len(node_forces)}",
             "linear forces:"]
    for F, d, node in node_forces:
        lines.append(f"{F!r} {d[0]!r} {d[1]!r} {d[2]!r} {blob_index} 0 {node}")
    _w(path, lines)


def write_full_ctforces(path, node_forces=(), rot_forces=(), surface_sets=(), blob_index=0):
    """A .ctforces file with any of its three sections (Blob::load_ctforces,
    Blob.cpp:2478-2806): linear node forces (F [N], direction, node or "all"), rotational
    forces (F [N], type "pf"/"pt", axis point [m], axis direction, node or "all") and linear
    surface forces (F [N], direction, face indices)."""
    n = len(node_forces) + len(rot_forces) + len(surface_sets)
    lines = ["ffea ctforces file", f"num_ctforces {n}", f"num_linear_forces {len(node_forces)}",
             f"num_rot_forces {len(rot_forces)}", f"num_linear_surface_forces {len(surface_sets)}"]
    if node_forces:
        lines.append("linear forces:")
        for F, d, node in node_forces:
            lines.append(f"{F!r} {d[0]!r} {d[1]!r} {d[2]!r} {blob_index} 0 {node}")
    if rot_forces:
        lines.append("rotational forces:")
        for F, typ, pt, d, node in rot_forces:
            lines.append(f"{F!r} {typ} {pt[0]!r} {pt[1]!r} {pt[2]!r} {d[0]!r} {d[1]!r} {d[2]!r} {blob_index} 0 {node}")
    if surface_sets:
        lines.append("linear surface forces:")
        for F, d, faces in surface_sets:
            lines.append(f"{F!r} {d[0]!r} {d[1]!r} {d[2]!r} {blob_index} 0 " + " ".join(str(int(f)) for f in faces))
    _w(path, lines)

