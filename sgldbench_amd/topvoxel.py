"""`.TopVoxel` voxel-model files and the boundary-condition set-up that turns
one into a solve (F, fixed DOFs, moduli) for the GPU solver.

File format (version 1.0), as src/IO_ImportTopVoxels.m:12-48 reads it and
src/IO_ExportTopVoxels.m:19-100 writes it, whitespace-separated tokens:

    #Voxel Model for SGLDBench
    Version: 1.0
    Resolution: nelx nely nelz
    Density values: 0|1
    Solid voxels: n        then n ids (or n "id density" pairs)
    Passive elements: m    then m ids
    Fixations: k           then k rows "node fixX fixY fixZ"
    Loads: l               then l rows "node Fx Fy Fz"
    [Additional Loads: 0]  (written, never read)

Voxel ids are 1-based into the (nely, nelx, nelz) column-major voxel box;
node ids are 1-based into the (nely+1, nelx+1, nelz+1) column-major node box
of the UNPADDED resolution.  After the mesh hierarchy pads the box
(FEA_VoxelBasedDiscretization.m:37-49) nodes keep their (row, col, layer) and
are renumbered in the padded box (IO_ImportTopVoxels.m:84-132).
"""
import dataclasses
import warnings

import numpy as np

from . import problems

HEADER = "#Voxel Model for SGLDBench"
COARSEST_RESOLUTION_CONTROL = 50000  # Data_GlobalVariables.m:13


class FormatError(ValueError):
    """IO_ImportTopVoxels.m's error('Incompatible Mesh Data Format!')."""


@dataclasses.dataclass
class TopVoxelModel:
    nelx: int
    nely: int
    nelz: int
    solid_voxels: np.ndarray          # int64, 1-based voxel ids, file order
    density: np.ndarray = None        # float64 per solid voxel (file order) or None
    passive: np.ndarray = None        # int64, 1-based voxel ids
    fixing: np.ndarray = None         # (k, 4) int64: node, fixX, fixY, fixZ
    loading: np.ndarray = None        # (l, 4) float64: node, Fx, Fy, Fz

    @property
    def voxels(self):
        """(nely, nelx, nelz) uint8, IO_ImportTopVoxels.m:51-53."""
        v = np.zeros(self.nelx * self.nely * self.nelz, dtype=np.uint8)
        v[self.solid_voxels - 1] = 1
        return v.reshape((self.nely, self.nelx, self.nelz), order="F")


class _Tokens:
    """fscanf-style cursor over the whitespace-separated tokens of a file."""

    def __init__(self, data):
        self.t = data.split()
        self.i = 0

    def word(self):
        if self.i >= len(self.t):
            raise FormatError("Incompatible Mesh Data Format! (unexpected end of file)")
        w = self.t[self.i].decode("ascii", "replace")
        self.i += 1
        return w

    def expect(self, word):
        if self.word() != word:
            raise FormatError("Incompatible Mesh Data Format!")

    def numbers(self, count, dtype):
        if self.i + count > len(self.t):
            raise FormatError("Incompatible Mesh Data Format! (unexpected end of file)")
        try:
            a = np.array(self.t[self.i:self.i + count], dtype=np.float64)
        except ValueError as e:
            raise FormatError("Incompatible Mesh Data Format! (%s)" % e) from None
        self.i += count
        if dtype is np.int64:
            if not np.all(a == np.round(a)):
                raise FormatError("Incompatible Mesh Data Format! (non-integer id)")
            return a.astype(np.int64)
        return a

    def integer(self):
        return int(self.numbers(1, np.int64)[0])


def read_topvoxel(path):
    """IO_ImportTopVoxels.m:12-48.  Returns a TopVoxelModel, or None with the
    reference's 'Unsupported Data!' warning for a version other than 1.0.
    Raises FormatError where the reference raises 'Incompatible Mesh Data
    Format!' (and on truncated sections, which fscanf would read short)."""
    with open(path, "rb") as f:
        tk = _Tokens(f.read())
    for _ in range(4):  # '#Voxel Model for SGLDBench' (not checked)
        tk.word()
    tk.word()           # 'Version:'
    try:
        version = float(tk.word())
    except ValueError:
        version = None
    if version != 1.0:
        warnings.warn("Unsupported Data!")
        return None
    tk.expect("Resolution:")
    nelx, nely, nelz = (tk.integer() for _ in range(3))
    if min(nelx, nely, nelz) <= 0:
        raise FormatError("Incompatible Mesh Data Format! (resolution)")
    tk.word(); tk.word()  # 'Density values:'
    with_density = tk.integer()
    tk.word(); tk.word()  # 'Solid voxels:'
    n = tk.integer()
    density = None
    if with_density:
        pairs = tk.numbers(2 * n, np.float64).reshape(n, 2)
        ids, density = pairs[:, 0], pairs[:, 1].copy()
        if not np.all(ids == np.round(ids)):
            raise FormatError("Incompatible Mesh Data Format! (non-integer id)")
        solid = ids.astype(np.int64)
    else:
        solid = tk.numbers(n, np.int64)
    nvox = nelx * nely * nelz
    if n and (solid.min() < 1 or solid.max() > nvox):
        raise FormatError("Incompatible Mesh Data Format! (voxel id out of range)")
    tk.expect("Passive")
    tk.expect("elements:")
    passive = tk.numbers(tk.integer(), np.int64)
    tk.expect("Fixations:")
    k = tk.integer()
    fixing = tk.numbers(4 * k, np.int64).reshape(k, 4) if k > 0 else np.zeros((0, 4), np.int64)
    tk.expect("Loads:")
    l = tk.integer()
    loading = tk.numbers(4 * l, np.float64).reshape(l, 4) if l > 0 else np.zeros((0, 4))
    nnode = (nelx + 1) * (nely + 1) * (nelz + 1)
    for name, ids in (("fixation", fixing[:, 0]), ("load", loading[:, 0])):
        if ids.size and (ids.min() < 1 or ids.max() > nnode):
            raise FormatError("Incompatible Mesh Data Format! (%s node id out of range)" % name)
    return TopVoxelModel(nelx, nely, nelz, solid, density, passive, fixing, loading)


def write_topvoxel(path, model, include_density=False):
    """IO_ExportTopVoxels.m:19-100 (the current writer): same sections, same
    number formats ('%d', '%6.2e' densities, '%.4e' forces), ids in the
    unpadded boxes, rows ascending by node id."""
    if include_density and model.density is None:
        warnings.warn("No Density Layout Available! Only Exports Voxel Model!")
        include_density = False
    order = np.argsort(model.solid_voxels, kind="stable")
    lines = [HEADER + " ", "Version: 1.0",
             "Resolution: %d %d %d" % (model.nelx, model.nely, model.nelz),
             "Density values: %d" % int(include_density),
             "Solid voxels: %d" % model.solid_voxels.size]
    if include_density:
        lines += ["%d %6.2e" % (i, d) for i, d in zip(model.solid_voxels[order], model.density[order])]
    else:
        lines += ["%d" % i for i in model.solid_voxels[order]]
    passive = np.sort(model.passive) if model.passive is not None else np.zeros(0, np.int64)
    lines.append("Passive elements: %d" % passive.size)
    lines += ["%d" % i for i in passive]
    fixing = model.fixing if model.fixing is not None else np.zeros((0, 4), np.int64)
    fixing = fixing[np.argsort(fixing[:, 0], kind="stable")]
    lines.append("Fixations: %d" % len(fixing))
    lines += ["%d %d %d %d" % tuple(r) for r in fixing]
    loading = model.loading if model.loading is not None else np.zeros((0, 4))
    loading = loading[np.argsort(loading[:, 0], kind="stable")]
    lines.append("Loads: %d" % len(loading))
    lines += ["%d %.4e %.4e %.4e" % (r[0], r[1], r[2], r[3]) for r in loading]
    lines.append("Additional Loads: 0")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def load_packed(path):
    """A model packed as numpy arrays (.npz): `resolution` (nelx, nely, nelz),
    `solid_bits`/`passive_bits` (np.packbits of the column-major voxel box),
    `fixing` (k, 4) and `loading_nodes` (l,) + `loading_forces` (l, 3) rows as
    the file lists them."""
    g = np.load(path)
    nelx, nely, nelz = (int(v) for v in g["resolution"])
    nvox = nelx * nely * nelz
    solid = np.flatnonzero(np.unpackbits(g["solid_bits"])[:nvox]) + 1
    passive = np.flatnonzero(np.unpackbits(g["passive_bits"])[:nvox]) + 1
    loading = np.column_stack([g["loading_nodes"].astype(np.float64), g["loading_forces"]])
    return TopVoxelModel(nelx, nely, nelz, solid.astype(np.int64), None, passive.astype(np.int64),
                         g["fixing"].astype(np.int64), loading)


def reference_num_levels(voxels, coarsest_control=COARSEST_RESOLUTION_CONTROL):
    """FEA_VoxelBasedDiscretization.m:26-37: the level count the reference
    picks (at least 3), and the multiple the resolution is padded to."""
    num = int(np.count_nonzero(voxels))
    ctl = min(coarsest_control, _mround(_mround(num / 8) / 8))
    levels = 0
    while num >= ctl:
        levels += 1
        num = _mround(num / 8)
    return levels + 1, 2 ** levels


def _mround(x):
    """MATLAB round: halves away from zero."""
    return int(np.floor(x + 0.5))


def pad_voxels(voxels, multiple):
    """FEA_VoxelBasedDiscretization.m:37-49: append void voxels (rows at the
    bottom, columns right, layers back) up to multiples of `multiple`."""
    vox = np.asarray(voxels)
    ny, nx, nz = vox.shape
    up = lambda n: -(-n // multiple) * multiple
    out = np.zeros((up(ny), up(nx), up(nz)), dtype=np.uint8)
    out[:ny, :nx, :nz] = vox != 0
    return out


def remap_nodes(nodes, unpadded, padded):
    """IO_ImportTopVoxels.m:84-132 (AdaptBCExternalMdl): 1-based node ids of
    the (nely+1, nelx+1, nelz+1) box -> 1-based ids of the padded node box,
    keeping (row, col, layer)."""
    ny, nx, _ = unpadded
    NY, NX, _ = padded
    g = np.asarray(nodes, dtype=np.int64) - 1
    r = g % (ny + 1)
    c = (g // (ny + 1)) % (nx + 1)
    l = g // ((ny + 1) * (nx + 1))
    return 1 + r + (NY + 1) * c + (NY + 1) * (NX + 1) * l


def remap_voxels(ids, unpadded, padded):
    """IO_ImportTopVoxels.m:135-153 (AdaptPassiveElementsExternalMdl)."""
    ny, nx, _ = unpadded
    NY, NX, _ = padded
    g = np.asarray(ids, dtype=np.int64) - 1
    r = g % ny
    c = (g // ny) % nx
    l = g // (ny * nx)
    return 1 + r + NY * c + NY * NX * l


def _unique_rows(rows):
    """[~, i] = unique(rows(:,1)); rows(i,:): the first row of each id,
    ascending by id (MATLAB's unique returns first occurrences)."""
    _, i = np.unique(rows[:, 0], return_index=True)
    return rows[i]


def setup(model, num_levels=None, multiple=None):
    """From a TopVoxelModel to the solve inputs, IO_ImportTopVoxels.m:51-82
    then FEA_ApplyBoundaryCondition.m:9-41:

    * voxels padded to multiples of 2^(num_levels-1) (or of `multiple`);
    * fixations/loads de-duplicated, sorted, remapped to the padded box, and
      required to lie on boundary nodes (nodes with < 8 solid voxels; the
      reference indexes nodesOnBoundary with them and fails otherwise);
    * fixations at loaded nodes dropped; F at the loaded nodes; a DOF fixed
      where its flag equals 1;
    * passive voxels remapped to 0-based compressed element indices;
    * E = 1 per element (the file holds the shape, not a design), density
      (if present) in compressed element order.

    Returns a dict like problems.cfg*(); `num_levels` defaults to the
    reference's rule.  Returns None with the reference's warning when there
    are no loads or no fixations."""
    vox0 = model.voxels
    if num_levels is None:
        num_levels, mult = reference_num_levels(vox0)
    else:
        mult = 2 ** (num_levels - 1)
    if multiple is not None:
        mult = multiple
    vox = pad_voxels(vox0, mult)
    ny, nx, nz = vox.shape
    unpadded = (model.nely, model.nelx, model.nelz)
    bpos, nodes = problems.surface_nodes(vox)         # positions into `nodes`; 0-based lattice ids
    lattice_to_bnd = np.zeros((ny + 1) * (nx + 1) * (nz + 1), dtype=np.int64)
    lattice_to_bnd[nodes[bpos]] = np.arange(1, bpos.size + 1)

    def to_boundary(rows, what):
        rows = _unique_rows(rows)
        lat = remap_nodes(rows[:, 0], unpadded, (ny, nx, nz)) - 1
        b = lattice_to_bnd[lat]
        if np.any(b == 0):
            bad = int(rows[np.flatnonzero(b == 0)[0], 0])
            raise ValueError("%s node %d is not a boundary node of the voxel model "
                             "(FEA_ApplyBoundaryCondition.m indexes nodesOnBoundary with it)" % (what, bad))
        out = rows.astype(np.float64 if rows.dtype.kind == "f" else np.int64).copy()
        out[:, 0] = b
        return out

    fixing = model.fixing if model.fixing is not None else np.zeros((0, 4), np.int64)
    loading = model.loading if model.loading is not None else np.zeros((0, 4))
    fix_b = to_boundary(fixing, "fixation") if len(fixing) else np.zeros((0, 4), np.int64)
    load_b = to_boundary(loading, "load") if len(loading) else np.zeros((0, 4))
    # FEA_ApplyBoundaryCondition.m:9-15
    fix_b = fix_b[~np.isin(fix_b[:, 0], load_b[:, 0].astype(np.int64))]
    fix_b = _unique_rows(fix_b)
    load_b = _unique_rows(load_b)
    if len(load_b) == 0:
        warnings.warn("No Loads!")
        return None
    if len(fix_b) == 0:
        warnings.warn("No Fixations!")
        return None
    n = nodes.size
    F = np.zeros((n, 3))
    F[bpos[load_b[:, 0].astype(np.int64) - 1]] = load_b[:, 1:4]
    fixed = np.zeros((n, 3), dtype=np.uint8)
    fn = bpos[fix_b[:, 0] - 1]
    fixed[fn] = (fix_b[:, 1:4] == 1).astype(np.uint8)
    # compressed element order = ascending padded voxel id
    solid_padded = np.flatnonzero(vox.ravel(order="F"))
    ele_fwd = np.full(vox.size, -1, dtype=np.int64)
    ele_fwd[solid_padded] = np.arange(solid_padded.size)
    passive = np.zeros(0, np.int64)
    if model.passive is not None and model.passive.size:
        passive = ele_fwd[remap_voxels(np.sort(model.passive), unpadded, (ny, nx, nz)) - 1]
        if np.any(passive < 0):
            raise ValueError("passive element outside the solid voxels")
    density = None
    if model.density is not None:
        density = np.empty(solid_padded.size)
        density[ele_fwd[remap_voxels(model.solid_voxels, unpadded, (ny, nx, nz)) - 1]] = model.density
    return dict(name="topvoxel%dx%dx%d" % (model.nelx, model.nely, model.nelz), voxels=vox,
                E=np.ones(solid_padded.size), F=F.ravel(), fixed=fixed.ravel(), passive=passive,
                density=density, num_levels=num_levels, omega=0.6, tol=1e-8, max_it=500)


def load_bc(path):
    """A `data/*.bc` file (IO_LoadBuiltInDatasets.m:12-13): MATLAB `load` of a
    numeric matrix with rows [boundaryNodeIdx a b c], the boundary index
    already in the voxeliser's mesh (so usable only with that mesh)."""
    a = np.loadtxt(path, ndmin=2)
    if a.shape[1] != 4:
        raise FormatError("a .bc file holds 4 columns per row")
    return a
