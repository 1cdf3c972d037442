"""Synthetic voxel problems of the benchmark workloads, in the reference's
conventions.

Voxel arrays are (ny, nx, nz) with row 0 at the top (y = ny), as
FEA_CreateVoxelizedModel.m:6 / FEA_VoxelBasedDiscretization.m:26 produce them.
Lattice node id g = row + (ny+1) col + (ny+1)(nx+1) layer; compressed nodes are
the solid lattice nodes in ascending g (FEA_VoxelBasedDiscretization.m:81-85);
DOFs are interleaved 3 per node (FEA_ApplyBoundaryCondition.m:21,27-29).
"""
import numpy as np


def solid_nodes(voxels):
    """Ascending 0-based lattice ids of the nodes of solid voxels."""
    vox = np.asarray(voxels) != 0
    ny, nx, nz = vox.shape
    node = np.zeros((ny + 1, nx + 1, nz + 1), dtype=bool)
    for dr in (0, 1):
        for dc in (0, 1):
            for dl in (0, 1):
                node[dr:dr + ny, dc:dc + nx, dl:dl + nz] |= vox
    return np.flatnonzero(node.ravel(order="F"))


def node_coords(voxels, nodes):
    """(x, y, z) of lattice node ids (x = col, y = ny - row, z = layer)."""
    ny, nx, nz = np.asarray(voxels).shape
    g = np.asarray(nodes, dtype=np.int64)
    row = g % (ny + 1)
    col = (g // (ny + 1)) % (nx + 1)
    lay = g // ((ny + 1) * (nx + 1))
    return np.stack([col, ny - row, lay], axis=1)


def cantilever3(voxels, nelx=None):
    """'Cuboid - Cantilever 3' (FEA_BuiltInBoundaryConditions4CuboidDesignDomain.m:34-44)
    with FEA_ApplyBoundaryCondition.m:9-41: nodes at x = min x fixed in x/y/z;
    total force -1 in z spread over the nodes at x = nelx, z = min z; loaded
    nodes are dropped from the fixations.  Returns (F, fixed uint8)."""
    nodes = solid_nodes(voxels)
    xyz = node_coords(voxels, nodes)
    nx = np.asarray(voxels).shape[1] if nelx is None else nelx
    fixed_n = xyz[:, 0] == xyz[:, 0].min()
    loaded = (xyz[:, 0] == nx) & (xyz[:, 2] == xyz[:, 2].min())
    fixed_n &= ~loaded
    n = len(nodes)
    F = np.zeros(3 * n)
    F[3 * np.flatnonzero(loaded) + 2] = -1.0 / loaded.sum()
    fixed = np.zeros(3 * n, dtype=np.uint8)
    for a in range(3):
        fixed[3 * np.flatnonzero(fixed_n) + a] = 1
    return F, fixed


def simp_modulus(x, p=3.0, e0=1.0, emin=1e-9):
    """TopOpti_MaterialInterpolationSIMP.m:5: E = Emin + x^p (E0 - Emin)."""
    return emin + np.asarray(x, dtype=np.float64) ** p * (e0 - emin)


def gaussian_random_density(shape, seed, sigma=6.0, lo=0.001, hi=1.0):
    """Seeded white noise, Gaussian-smoothed (reflect, truncate 4), min-max
    scaled to [lo, hi]."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    g = gaussian_filter(w, sigma=sigma, mode="reflect", truncate=4.0)
    g = (g - g.min()) / (g.max() - g.min())
    return lo + g * (hi - lo)


def cuboid(nx, ny, nz):
    return np.ones((ny, nx, nz), dtype=np.uint8)


def cfg0():
    """64x32x32 solid cuboid, E = 1, Cantilever 3, 4 levels, omega 0.6, tol 1e-8."""
    vox = cuboid(64, 32, 32)
    F, fixed = cantilever3(vox)
    E = simp_modulus(np.ones(int(vox.sum())))
    return dict(name="cuboid64x32x32", voxels=vox, E=E, F=F, fixed=fixed, num_levels=4,
                omega=0.6, tol=1e-8, max_it=500)


CFG1_SEED = 20241
CFG3_SEED = 20243


def random_density_cuboid(nx, ny, nz, seed):
    """Density in voxel order (ny, nx, nz) column-major, as eleMapBack lists
    the elements of a full cuboid."""
    x = gaussian_random_density((ny, nx, nz), seed)
    return x.ravel(order="F")


def cfg1(nx=256, ny=128, nz=128, levels=6, seed=CFG1_SEED):
    """256x128x128 cuboid, SIMP E from a seeded smoothed random field,
    Cantilever 3, 6 levels."""
    vox = cuboid(nx, ny, nz)
    F, fixed = cantilever3(vox)
    E = simp_modulus(random_density_cuboid(nx, ny, nz, seed))
    return dict(name="cuboid%dx%dx%d-grf" % (nx, ny, nz), voxels=vox, E=E, F=F, fixed=fixed,
                num_levels=levels, omega=0.6, tol=1e-8, max_it=500)


def face_load_cases(voxels, nrhs, seed):
    """Load cases of the batched workload: each a unit total force spread
    uniformly over the solid nodes of the x = nelx face, in a direction drawn
    uniformly on the sphere (seeded); supports as Cantilever 3 (x = 0 face
    fixed in x/y/z).  Returns (F (nrhs, numDOF), fixed uint8, directions)."""
    nodes = solid_nodes(voxels)
    xyz = node_coords(voxels, nodes)
    nx = np.asarray(voxels).shape[1]
    face = np.flatnonzero(xyz[:, 0] == nx)
    fixed_n = np.flatnonzero(xyz[:, 0] == xyz[:, 0].min())
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((nrhs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = len(nodes)
    F = np.zeros((nrhs, 3 * n))
    for a in range(3):
        F[:, 3 * face + a] = d[:, a:a + 1] / len(face)
    fixed = np.zeros(3 * n, dtype=np.uint8)
    for a in range(3):
        fixed[3 * fixed_n + a] = 1
    return F, fixed, d


def cfg3(n=256, nrhs=16, levels=6, seed=CFG3_SEED):
    """n^3 cuboid, SIMP E from the GRF recipe of the 256x128x128 workload at
    n^3, 16 load cases on the x = n face (unit force, random direction),
    x = 0 face fixed."""
    vox = cuboid(n, n, n)
    F, fixed, d = face_load_cases(vox, nrhs, seed)
    E = simp_modulus(random_density_cuboid(n, n, n, seed))
    return dict(name="cuboid%d^3-grf-%drhs" % (n, nrhs), voxels=vox, E=E, F=F, fixed=fixed,
                directions=d, num_levels=levels, omega=0.6, tol=1e-8, max_it=500)


CFG4_SEED = 20244
CFG4_DIMS = (32, 48, 64, 96, 128)


def levels_for(dims, max_levels=5):
    """Multigrid levels of an instance: every axis divisible by 2^(L-1),
    at most max_levels (48 and 96 = 3 2^k allow 5)."""
    L = 1
    while L < max_levels and all(d % (2 ** L) == 0 for d in dims):
        L += 1
    return L


def cfg4(n_inst=32, seed=CFG4_SEED, dims_set=CFG4_DIMS, penal=3.0):
    """Uneven batch: n_inst cuboids with per-axis sizes drawn from dims_set
    (seeded), Cantilever 3 supports and load, each with its own seeded
    Gaussian random density in [0.001, 1] (sigma = 6 voxels scaled by the
    instance's mean edge / 256, at least 1), SIMP E = 1e-9 + x^penal."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        nx, ny, nz = (int(v) for v in rng.choice(dims_set, 3))
        vox = cuboid(nx, ny, nz)
        F, fixed = cantilever3(vox)
        sigma = max(1.0, 6.0 * (nx + ny + nz) / 3.0 / 256.0)
        x = gaussian_random_density((ny, nx, nz), seed * 1000 + i, sigma=sigma).ravel(order="F")
        out.append(dict(name="cuboid%dx%dx%d" % (nx, ny, nz), voxels=vox, x=x, E=simp_modulus(x, penal),
                        F=F, fixed=fixed, num_levels=levels_for((nx, ny, nz)), omega=0.6, tol=1e-8,
                        max_it=500, penal=penal))
    return out


CFG2_SEED = 20277  # fill close to the ~30 % the workload describes (seed scan at 1/4 scale)


def ellipsoid_union(size, n_ell, axes_range, seed, keep_below=8):
    """Union of seeded axis-aligned ellipsoids (centres uniform in the box,
    semi-axes uniform in axes_range) in a size^3 voxel box, restricted to the
    face-connected components that contain a voxel with layer < keep_below
    (a floating component would make K singular).  Returns (ny, nx, nz) uint8."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    vox = np.zeros((size, size, size), dtype=bool)
    c = rng.uniform(0, size, (n_ell, 3))
    ax = rng.uniform(axes_range[0], axes_range[1], (n_ell, 3))
    for (cy, cx, cz), (ay, bx, gz) in zip(c, ax):
        lo = np.maximum(np.floor([cy - ay, cx - bx, cz - gz]).astype(int), 0)
        hi = np.minimum(np.ceil([cy + ay, cx + bx, cz + gz]).astype(int) + 1, size)
        yy = (np.arange(lo[0], hi[0]) + 0.5 - cy) / ay
        xx = (np.arange(lo[1], hi[1]) + 0.5 - cx) / bx
        zz = (np.arange(lo[2], hi[2]) + 0.5 - cz) / gz
        m = (yy[:, None, None] ** 2 + xx[None, :, None] ** 2 + zz[None, None, :] ** 2) <= 1.0
        vox[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= m
    lab, n = ndimage.label(vox)
    keep = np.unique(lab[:, :, :keep_below])
    keep = keep[keep > 0]
    return np.isin(lab, keep).astype(np.uint8)


def surface_nodes(voxels):
    """Positions (in the solid-node list) of nodes with fewer than 8 incident
    solid voxels, as FEA_VoxelBasedDiscretization.m:92-98 marks boundary nodes."""
    vox = np.asarray(voxels) != 0
    ny, nx, nz = vox.shape
    cnt = np.zeros((ny + 1, nx + 1, nz + 1), dtype=np.uint8)
    for dr in (0, 1):
        for dc in (0, 1):
            for dl in (0, 1):
                cnt[dr:dr + ny, dc:dc + nx, dl:dl + nz] += vox
    flat = cnt.ravel(order="F")
    solid = np.flatnonzero(flat)
    return np.flatnonzero(flat[solid] < 8), solid


def cfg2(size=512, n_ell=24, axes=(40, 160), levels=7, n_loads=2000, z_fix=8, z_load=400,
         seed=CFG2_SEED):
    """Ellipsoid-union domain: nodes of active voxels with layer < z_fix fixed
    in x/y/z; n_loads seeded surface nodes above z_load carry unit forces in
    seeded random directions; E = 1."""
    vox = ellipsoid_union(size, n_ell, axes, seed, keep_below=z_fix)
    ny, nx, nz = vox.shape
    surf, solid = surface_nodes(vox)
    lay = solid // ((ny + 1) * (nx + 1))
    nn = len(solid)
    # fixed: every node of an active voxel with layer < z_fix
    fixed_node = np.zeros((ny + 1, nx + 1, nz + 1), dtype=bool)
    low = vox[:, :, :z_fix] != 0
    for dr in (0, 1):
        for dc in (0, 1):
            for dl in (0, 1):
                fixed_node[dr:dr + ny, dc:dc + nx, dl:dl + z_fix] |= low
    fixed_n = fixed_node.ravel(order="F")[solid]
    rng = np.random.default_rng(seed + 1)
    cand = surf[(lay[surf] > z_load) & ~fixed_n[surf]]
    pick = rng.choice(cand, size=min(n_loads, len(cand)), replace=False)
    d = rng.standard_normal((len(pick), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    F = np.zeros(3 * nn)
    for a in range(3):
        F[3 * pick + a] = d[:, a]
    fixed = np.zeros(3 * nn, dtype=np.uint8)
    for a in range(3):
        fixed[3 * np.flatnonzero(fixed_n) + a] = 1
    E = np.ones(int(vox.sum()))
    return dict(name="ellipsoids%d^3" % size, voxels=vox, E=E, F=F, fixed=fixed, num_levels=levels,
                omega=0.6, tol=1e-8, max_it=1000)
