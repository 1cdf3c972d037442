"""Writes tests/golden/part_R256.npz from the reference's voxel model
data/part_R256.TopVoxel (256x160x112 voxels, 1,845,040 solid, 74,546
passive, 13,042 fixations, 7,788 loads):

* the model itself, packed: solid and passive voxels as bit masks over the
  (nely, nelx, nelz) column-major box, fixation rows (node, fixX, fixY, fixZ)
  and load rows (node, Fx, Fy, Fz) as the file lists them;
* the oracle's FP64 MG-PCG solve of it (E = 1, Jacobi weight 0.6, tol 1e-8,
  `LEVELS` dyadic levels): iteration count, residual history, compliance
  F.U and U at the loaded DOFs.

The file is parsed here by a plain line reader of its own (section headers,
then one record per line), independent of sgldbench_amd.topvoxel, so the
fixture also checks that reader.  Run in a checkout that has the reference;
the tests read only the .npz.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = "/root/reference/data/part_R256.TopVoxel"
OUT = os.path.join(HERE, "part_R256.npz")
LEVELS = 6


def parse_lines(path):
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    assert lines[0].split() == ["#Voxel", "Model", "for", "SGLDBench"], lines[0]
    assert lines[1] == "Version: 1.0"
    res = [int(v) for v in lines[2].split(":")[1].split()]
    assert lines[3] == "Density values: 0"
    i = 4
    sections = {}
    for name, cols in (("Solid voxels", 1), ("Passive elements", 1), ("Fixations", 4), ("Loads", 4)):
        head, n = lines[i].rsplit(":", 1)
        assert head == name, (head, name)
        n = int(n)
        rows = [ln.split() for ln in lines[i + 1:i + 1 + n]]
        assert all(len(r) == cols for r in rows)
        sections[name] = rows
        i += 1 + n
    return res, sections


def main():
    res, sec = parse_lines(SRC)
    nelx, nely, nelz = res
    nvox = nelx * nely * nelz
    solid = np.array([int(r[0]) for r in sec["Solid voxels"]], dtype=np.int64)
    passive = np.array([int(r[0]) for r in sec["Passive elements"]], dtype=np.int64)
    fixing = np.array([[int(v) for v in r] for r in sec["Fixations"]], dtype=np.int64)
    loading = np.array([[float(v) for v in r] for r in sec["Loads"]], dtype=np.float64)
    solid_mask = np.zeros(nvox, dtype=bool)
    solid_mask[solid - 1] = True
    passive_mask = np.zeros(nvox, dtype=bool)
    passive_mask[passive - 1] = True
    assert np.all(np.diff(solid) > 0) and np.all(np.diff(passive) > 0)
    model = dict(resolution=np.array(res, dtype=np.int64),
                 solid_bits=np.packbits(solid_mask), passive_bits=np.packbits(passive_mask),
                 fixing=fixing.astype(np.int32), loading_nodes=loading[:, 0].astype(np.int32),
                 loading_forces=loading[:, 1:])

    # the oracle's solve of the model
    sys.path.insert(0, ROOT)
    from oracle import oracle
    from sgldbench_amd import topvoxel
    m = topvoxel.TopVoxelModel(nelx, nely, nelz, solid, None, passive, fixing, loading)
    p = topvoxel.setup(m, num_levels=LEVELS)
    t0 = time.time()
    mesh = oracle.Mesh(p["voxels"], LEVELS)
    mesh.set_fixed(p["fixed"])
    mesh.set_modulus(p["E"])
    mesh.assemble()
    t1 = time.time()
    U, its, hist, rc = mesh.pcg(p["F"], 0.6, 1e-8, 2000)
    t2 = time.time()
    print("oracle: setup %.1f s, pcg %.1f s, %d its, rel-res %.3e, rc %d" % (t1 - t0, t2 - t1, its, hist[-1], rc))
    assert rc == 0
    loaded = np.flatnonzero(p["F"] != 0)
    np.savez_compressed(OUT, **model, levels=np.int64(LEVELS), its=np.int64(its), hist=hist,
                        compliance=np.float64(p["F"] @ U), u_loaded_dofs=loaded.astype(np.int64),
                        u_loaded=U[loaded], u_norm=np.float64(np.linalg.norm(U)))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
