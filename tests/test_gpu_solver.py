"""Device-resident solver (lattice path) against the oracle: K.u and the
V-cycle to FP64 rounding, PCG iteration counts within one, FP64 relative
residual <= 1e-8 and compliance relative error <= 1e-9 (FP64) on cuboids,
void-containing domains and the non-dyadic hierarchy; the mixed-precision
path reaches the same FP64 true residual."""
import numpy as np
import pytest

from oracle import oracle as orc
from sgldbench_amd import problems, solver

pytestmark = pytest.mark.gpu


def holes(nx, ny, nz, seed, nholes=6, r=3):
    """Cuboid with spherical voids (connected solid, x=0 face kept)."""
    vox = problems.cuboid(nx, ny, nz)
    rng = np.random.default_rng(seed)
    yy, xx, zz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    for _ in range(nholes):
        c = rng.uniform([r, 2 * r, r], [ny - r, nx - r, nz - r])
        vox[(yy - c[0]) ** 2 + (xx - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r] = 0
    return vox


def build_pair(vox, levels, nd=0, seed=5, density=True, bc=None):
    F, fixed = problems.cantilever3(vox) if bc is None else bc
    ne = int(vox.sum())
    if density:
        x = problems.gaussian_random_density(vox.shape, seed, sigma=2.0).ravel(order="F")
        E = problems.simp_modulus(x[vox.ravel(order="F") != 0])
    else:
        E = np.ones(ne)
    g = solver.FEASolver(vox, levels, nd)
    g.set_fixed(fixed)
    g.set_modulus(E)
    g.assemble(0.6)
    m = orc.Mesh(vox, levels, nd)
    m.set_fixed(fixed)
    m.set_modulus(E)
    m.assemble()
    return g, m, F, fixed, E


CASES = {
    "cuboid32x16x16": (lambda: problems.cuboid(32, 16, 16), 4, 0),
    "ellipsoids32": (lambda: problems.cfg2(32, 12, (4, 10), 3, 20, 2, 16, seed=7), 3, 0),
    "holes32x16x16": (lambda: holes(32, 16, 16, 7), 3, 0),
    "nondyadic32x16x16": (lambda: problems.cuboid(32, 16, 16), 4, 1),
}


@pytest.fixture(scope="module", params=list(CASES))
def pair(request):
    mk, lv, nd = CASES[request.param]
    print("case:", request.param, flush=True)
    v = mk()
    if isinstance(v, dict):  # irregular domain with its own supports and loads
        return build_pair(v["voxels"], lv, nd, bc=(v["F"], v["fixed"]))
    return build_pair(v, lv, nd)


def test_hierarchy_sizes_match(pair):
    g, m, *_ = pair
    assert g.numLevels == m.num_levels
    for a, b in zip(g.levels, m.info):
        assert (a["resX"], a["resY"], a["resZ"], a["numElements"], a["numNodes"]) == \
            (b["resX"], b["resY"], b["resZ"], b["numElements"], b["numNodes"])


def test_kbyu_matches_oracle(pair):
    g, m, F, fixed, E = pair
    u = np.random.default_rng(8).standard_normal(g.numDOFs)
    y, ref = g.kbyu(u), m.kbyu(u)
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.all(y[fixed == 1] == 0)


def test_vcycle_matches_oracle(pair):
    g, m, F, fixed, E = pair
    r = np.random.default_rng(9).standard_normal(g.numDOFs)
    r[fixed == 1] = 0
    z, ref = g.vcycle(r), m.vcycle(r, 0.6)
    assert np.abs(z - ref).max() <= 1e-11 * np.abs(ref).max()


def test_pcg_fp64_matches_oracle(pair):
    g, m, F, fixed, E = pair
    U, its, rr = g.pcg(F, tol=1e-8, maxIT=500)
    Uo, itso, hist, rc = m.pcg(F, 0.6, 1e-8, 500)
    print("its gpu %d oracle %d relres %.3e" % (its, itso, rr), flush=True)
    assert rc == 0 and g.converged
    assert abs(its - itso) <= 1
    assert rr <= 1e-8
    c, co = F @ U, F @ Uo
    assert abs(c - co) <= 1e-9 * abs(co)
    ce = g.unit_compliance(U)
    assert E @ ce == pytest.approx(c, rel=1e-9)
    assert np.abs(ce - m.unit_compliance(U)).max() <= 1e-10 * np.abs(ce).max()


def test_pcg_mixed_reaches_fp64_residual(pair):
    g, m, F, fixed, E = pair
    U, its, rr = g.pcg(F, tol=1e-8, maxIT=1000, precision=solver.PREC_MIXED)
    Uo, itso, hist, rc = m.pcg(F, 0.6, 1e-10, 1000)
    st = g.stats()
    print("mixed its %d outer %d relres %.3e" % (its, st["outer"], rr), flush=True)
    assert g.converged and rr <= 1e-8
    c, co = F @ U, F @ Uo
    assert abs(c - co) <= 1e-8 * abs(co)
    assert np.linalg.norm(U - Uo) <= 1e-5 * np.linalg.norm(Uo)


def test_warm_start_and_zero_load():
    vox = problems.cuboid(16, 8, 8)
    g, m, F, fixed, E = build_pair(vox, 3)
    U, its, rr = g.pcg(F, tol=1e-8)
    U2, its2, rr2 = g.pcg(F, tol=1e-8, U0=U)
    assert its2 <= 2 and rr2 <= 1e-8
    U3, its3, rr3 = g.pcg(np.zeros_like(F), tol=1e-8)
    assert its3 == 0 and np.all(U3 == 0)


def test_maxit_reports_without_raising():
    vox = problems.cuboid(16, 8, 8)
    g, m, F, fixed, E = build_pair(vox, 3)
    U, its, rr = g.pcg(F, tol=1e-14, maxIT=3)
    assert its == 3 and not g.converged
    Uo, itso, hist, rc = m.pcg(F, 0.6, 1e-14, 3)
    assert rc == 2 and itso == 3
    assert np.linalg.norm(U - Uo) <= 1e-10 * np.linalg.norm(Uo)


def test_cfg0_fp64_against_oracle():
    p = problems.cfg0()
    g, m, F, fixed, E = build_pair(p["voxels"], 4, density=False)
    U, its, rr = g.pcg(p["F"], tol=1e-8, maxIT=500)
    Uo, itso, hist, rc = m.pcg(p["F"], 0.6, 1e-8, 500)
    print("cfg0 its gpu %d oracle %d relres %.3e c %.12e" % (its, itso, rr, p["F"] @ U), flush=True)
    assert abs(its - itso) <= 1 and rr <= 1e-8
    assert abs(p["F"] @ U - p["F"] @ Uo) <= 1e-9 * abs(p["F"] @ Uo)


@pytest.mark.parametrize("variant", ["dense", "tiled", "tiled2"])
@pytest.mark.parametrize("dims", [(38, 22, 18), (16, 8, 8)])
def test_kbyu_kernel_variants(variant, dims, monkeypatch):
    """The K.u kernels (node-centric dense Ke; tiled mode-space with one or
    two block barriers per element layer) against the oracle, on lattices
    that are not multiples of the 15-node tile, with voids and moduli
    spanning 1e-9..1."""
    if variant == "dense":
        monkeypatch.setenv("SGLD_KBYU", "dense")
    else:
        monkeypatch.delenv("SGLD_KBYU", raising=False)
    monkeypatch.setenv("SGLD_KU_BARRIERS", "2" if variant == "tiled2" else "1")
    nx, ny, nz = dims
    vox = holes(nx, ny, nz, 11, nholes=4, r=3)
    F, fixed = problems.cantilever3(vox)
    rng = np.random.default_rng(12)
    E = 10.0 ** rng.uniform(-9, 0, int(vox.sum()))
    print("case: kbyu", variant, dims, flush=True)
    g = solver.FEASolver(vox, 2)
    g.set_fixed(fixed)
    g.set_modulus(E)
    g.assemble(0.6)
    m = orc.Mesh(vox, 2)
    m.set_fixed(fixed)
    m.set_modulus(E)
    u = rng.standard_normal(g.numDOFs)
    y, ref = g.kbyu(u), m.kbyu(u)
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_kbyu_fp32_against_fp64():
    import torch
    vox = holes(38, 22, 18, 13, nholes=4, r=3)
    F, fixed = problems.cantilever3(vox)
    rng = np.random.default_rng(14)
    E = rng.uniform(0.01, 1.0, int(vox.sum()))
    g = solver.FEASolver(vox, 2)
    g.set_fixed(fixed)
    g.set_modulus(E)
    u = rng.standard_normal(g.numDOFs)
    du = torch.from_numpy(u).cuda()
    y64 = torch.zeros_like(du)
    y32 = torch.zeros_like(du)
    g.kbyu_device(du.data_ptr(), y64.data_ptr(), solver.PREC_FP64)
    g.kbyu_device(du.data_ptr(), y32.data_ptr(), solver.PREC_MIXED)
    torch.cuda.synchronize()
    a, b = y64.cpu().numpy(), y32.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)
    m = orc.Mesh(vox, 2)
    m.set_fixed(fixed)
    m.set_modulus(E)
    ref = m.kbyu(u)
    assert np.abs(a - ref).max() <= 1e-13 * np.abs(ref).max()


def test_two_level_pcg_matches_oracle():
    """Two levels: no small-level tail, and x_2.r_2 of the next rho comes from
    a separate pass over the coarsest level."""
    vox = holes(16, 8, 8, 3, nholes=2, r=2)
    g, m, F, fixed, E = build_pair(vox, 2)
    U, its, rr = g.pcg(F, tol=1e-8, maxIT=500)
    Uo, itso, _, rc = m.pcg(F, 0.6, 1e-8, 500)
    assert rc == 0 and g.converged and rr <= 1e-8
    assert abs(its - itso) <= 1
    assert abs(F @ U - F @ Uo) <= 1e-9 * abs(F @ Uo)
    U32, its32, rr32 = g.pcg(F, tol=1e-8, maxIT=1000, precision=solver.PREC_MIXED)
    assert g.converged and rr32 <= 1e-8
    assert abs(F @ U32 - F @ Uo) <= 1e-8 * abs(F @ Uo)


@pytest.mark.parametrize("levels", [4, 5])
def test_vcycle_tail_equals_level_chain(levels, monkeypatch):
    """The fused small-level tail (composite span-2^m restrictions and one
    prolongation pass) equals the level-by-level chain and the oracle."""
    vox = holes(64, 32, 32, 11, nholes=4, r=4)
    F, fixed = problems.cantilever3(vox)
    x = problems.gaussian_random_density(vox.shape, 4, sigma=2.0).ravel(order="F")
    E = problems.simp_modulus(x[vox.ravel(order="F") != 0])
    zs = []
    for tail in ("1", "0"):
        monkeypatch.setenv("SGLD_TAIL", tail)
        g = solver.FEASolver(vox, levels)
        g.set_fixed(fixed)
        g.set_modulus(E)
        g.assemble(0.6)
        r = np.random.default_rng(21).standard_normal(g.numDOFs)
        r[fixed == 1] = 0
        zs.append(g.vcycle(r))
    m = orc.Mesh(vox, levels)
    m.set_fixed(fixed)
    m.set_modulus(E)
    m.assemble()
    ref = m.vcycle(r, 0.6)
    scale = np.abs(ref).max()
    assert np.abs(zs[0] - zs[1]).max() <= 1e-12 * scale
    assert np.abs(zs[0] - ref).max() <= 1e-11 * scale


@pytest.mark.parametrize("cz", ["1", "2", "3"])
@pytest.mark.parametrize("case", ["cuboid", "holes"])
def test_tiled_restriction_matches_node_centric(case, cz, monkeypatch):
    """The LDS-tiled span-2 restriction (partial row, column and z tiles, any
    z-chunk, the tile list of an irregular domain) gives the V-cycle of the
    node-centric restriction and of the oracle."""
    monkeypatch.setenv("SGLD_RESTRICT_CZ", cz)
    vox = problems.cuboid(40, 72, 20) if case == "cuboid" else holes(40, 72, 20, 5, nholes=5, r=4)
    F, fixed = problems.cantilever3(vox)
    x = problems.gaussian_random_density(vox.shape, 6, sigma=2.0).ravel(order="F")
    E = problems.simp_modulus(x[vox.ravel(order="F") != 0])
    print("case: tiled restriction", case, "cz", cz, flush=True)
    zs = []
    for restr in ("tiled", "node-centric"):
        g = solver.FEASolver(vox, 3)
        g.set_fixed(fixed)
        g.set_modulus(E)
        g.assemble(0.6)
        g.set_schedule("tiled", 8, restr)
        r = np.random.default_rng(31).standard_normal(g.numDOFs)
        r[fixed == 1] = 0
        zs.append(g.vcycle(r))
    m = orc.Mesh(vox, 3)
    m.set_fixed(fixed)
    m.set_modulus(E)
    m.assemble()
    ref = m.vcycle(r, 0.6)
    scale = np.abs(ref).max()
    assert np.abs(zs[0] - zs[1]).max() <= 1e-12 * scale
    assert np.abs(zs[0] - ref).max() <= 1e-11 * scale


def test_mixed_forced_updates_continue_direction(monkeypatch):
    """A reliable update forced by the stretch limit (here every 8 FP32
    iterations, with a threshold the stretches never reach) keeps the search
    direction: the solve converges in about the iterations of the default
    schedule, where restarting the direction at each one would make it a
    restarted CG(8)."""
    vox = holes(32, 16, 16, 9, nholes=4, r=3)
    g, m, F, fixed, E = build_pair(vox, 3)
    U0, its0, rr0 = g.pcg(F, tol=1e-8, maxIT=2000, precision=solver.PREC_MIXED)
    assert g.converged
    monkeypatch.setenv("SGLD_MIXED_DELTA", "1e-12")
    monkeypatch.setenv("SGLD_MIXED_STRETCH", "8")
    U, its, rr = g.pcg(F, tol=1e-8, maxIT=2000, precision=solver.PREC_MIXED)
    st = g.stats()
    print("default its %d; forced every 8: its %d, updates %d" % (its0, its, st["outer"]), flush=True)
    assert g.converged and rr <= 1e-8
    assert st["outer"] >= its // 8 - 1
    assert its <= 1.25 * its0 + 8
    assert np.linalg.norm(U - U0) <= 1e-6 * np.linalg.norm(U0)


@pytest.mark.parametrize("factor", [1e-30, 1e30])
def test_mixed_load_magnitude(factor):
    """The FP32 vectors of the mixed solve are divided by one power of two near
    the initial residual norm, so loads of any magnitude FP64 holds solve to
    the same relative accuracy, and the solution scales with the load."""
    vox = holes(32, 16, 16, 5, nholes=3, r=3)
    g, m, F, fixed, E = build_pair(vox, 3)
    U1, its1, rr1 = g.pcg(F, tol=1e-8, maxIT=1000, precision=solver.PREC_MIXED)
    U, its, rr = g.pcg(F * factor, tol=1e-8, maxIT=1000, precision=solver.PREC_MIXED)
    print("factor %g: its %d (unit load %d), relres %.2e" % (factor, its, its1, rr), flush=True)
    assert g.converged and rr <= 1e-8
    assert abs(its - its1) <= 2
    assert np.linalg.norm(U - factor * U1) <= 1e-7 * np.linalg.norm(factor * U1)


def test_coarsest_size_guard():
    """A coarsest level beyond the dense factor's limit is an argument error
    that names the limit (no host or device allocation of its n^2 arrays)."""
    vox = problems.cuboid(64, 64, 64)
    F, fixed = problems.cantilever3(vox)
    g = solver.FEASolver(vox, 2)  # coarsest 32^3 elements: 107,811 lattice DOFs
    g.set_fixed(fixed)
    g.set_modulus(np.ones(int(vox.sum())))
    with pytest.raises(solver.SGLDError) as e:
        g.assemble(0.6)
    assert e.value.code == 1 and "at most 32768" in str(e.value), str(e.value)
    g.close()


@pytest.mark.parametrize("case", ["cuboid", "holes", "cuboid_4k"])
def test_coarsest_solve_is_dense_cholesky_solve(case):
    """Two-level V-cycle minus its smoothing term isolates P K_L^-1 P^T r:
    with the coarsest level's free-DOF matrix assembled densely by the oracle
    (Galerkin Ks) and solved by numpy, the GPU's triangular products agree to
    FP64 rounding; the operator is symmetric.  The factor is blocked in 64 x 64
    tiles: 9-10 tiles per edge for the small cases, 64 (4,050 free DOFs, a
    ragged last tile) for cuboid_4k."""
    vox = {"cuboid": lambda: problems.cuboid(16, 8, 8), "holes": lambda: holes(16, 8, 8, 3, nholes=2, r=2),
           "cuboid_4k": lambda: problems.cuboid(32, 16, 16)}[case]()
    print("case: coarsest dense solve", case, flush=True)
    g, m, F, fixed, E = build_pair(vox, 2)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(F.size) * (1 - fixed)
    d1 = m.get(1, "diagK")
    smooth = np.where(fixed != 0, 0.0, 2 * 0.6 * r / d1)
    coarse_gpu = g.vcycle(r) - smooth
    # dense K_L(free, free) from the oracle's level-2 element matrices
    info = m.info[1]
    eN = m.get(2, "eNodMat")
    Ks = m.get(2, "Ks")
    n2 = info["numDOFs"]
    K = np.zeros((n2, n2))
    for e in range(info["numElements"]):
        d = (3 * (eN[e] - 1)[:, None] + np.arange(3)[None, :]).ravel()
        K[np.ix_(d, d)] += Ks[:, :, e]
    free = m.get(2, "fixedDOFs") == 0
    rc = m.restrict(2, r)
    xc = np.zeros(n2)
    xc[free] = np.linalg.solve(K[np.ix_(free, free)], rc[free])
    coarse_ref = m.interpolate(2, xc) * (1 - fixed)
    assert np.max(np.abs(coarse_gpu - coarse_ref)) <= 1e-11 * np.max(np.abs(coarse_ref))
    s = rng.standard_normal(F.size) * (1 - fixed)
    a, b = s @ g.vcycle(r), r @ g.vcycle(s)
    assert abs(a - b) <= 1e-12 * abs(a)


@pytest.mark.parametrize("case", ["cuboid", "holes", "ellipsoids"])
def test_fused_residual_update_and_restriction(case, monkeypatch):
    """With the LDS-tiled first restriction the iterations fuse the residual
    update into it and alternate two residual vectors: both precisions against
    the oracle, run to run identical, and within rounding of the unfused pair
    of passes (SGLD_FUSE_RES=0 and SGLD_FUSE_RES32=0; by default both forms
    run only on large irregular domains)."""
    print("case: fused residual update", case, flush=True)
    if case == "ellipsoids":
        d = problems.cfg2(32, 12, (4, 10), 3, 20, 2, 16, seed=7)
        vox, levels, bc = d["voxels"], 3, (d["F"], d["fixed"])
    else:
        vox = problems.cuboid(48, 24, 24) if case == "cuboid" else holes(48, 24, 24, 3)
        levels, bc = 4, None
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("SGLD_FUSE_RES", fuse)
        monkeypatch.setenv("SGLD_FUSE_RES32", fuse)
        g, m, F, fixed, E = build_pair(vox, levels, bc=bc)
        g.set_schedule("tiled", 12, "tiled")  # the tiled restriction on every level
        Uo, itso, hist, rc = m.pcg(F, 0.6, 1e-8, 1000)
        assert rc == 0
        for prec in (solver.PREC_FP64, solver.PREC_MIXED):
            g.kernel_timing(True)
            U, its, rr = g.pcg(F, tol=1e-8, maxIT=1000, precision=prec)
            kt = g.kernel_times()
            g.kernel_timing(False)
            name = kt["update_r64" if prec == solver.PREC_FP64 else "update_r32"]["kernel"]
            assert ("k_update_restrict" in name) == (fuse == "1"), name
            true = np.linalg.norm(F - m.kbyu(U)) / np.linalg.norm(F)
            c, co = F @ U, F @ Uo
            print("  fuse %s precision %d: its %d (oracle %d), true rel-res %.2e" % (fuse, prec, its, itso, true),
                  flush=True)
            assert g.converged and rr <= 1e-8 and true <= 1.01e-8
            assert abs(c - co) <= (1e-9 if prec == solver.PREC_FP64 else 1e-8) * abs(co)
            if prec == solver.PREC_FP64:
                assert abs(its - itso) <= 1
            U2, its2, _ = g.pcg(F, tol=1e-8, maxIT=1000, precision=prec)
            assert its2 == its and np.array_equal(U2, U)
            out[(fuse, prec)] = (U, its)
        g.close()
    for prec in (solver.PREC_FP64, solver.PREC_MIXED):
        (U1, i1), (U0, i0) = out[("1", prec)], out[("0", prec)]
        assert abs(i1 - i0) <= 1
        assert np.linalg.norm(U1 - U0) <= 1e-7 * np.linalg.norm(U0)


def test_fused_residual_update_under_frequent_reliable_updates(monkeypatch):
    """The alternating residual vectors of the fused pass when FP32 stretches
    stop inside a batch and at odd iteration counts (delta 0.5, a forced
    update after 5 iterations): the mixed solve still reaches the FP64 true
    residual, and agrees with the unfused passes under the same schedule."""
    print("case: fused residual update, frequent reliable updates", flush=True)
    vox = holes(48, 24, 24, 3)
    monkeypatch.setenv("SGLD_MIXED_DELTA", "0.5")
    monkeypatch.setenv("SGLD_MIXED_STRETCH", "5")
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("SGLD_FUSE_RES", fuse)
        monkeypatch.setenv("SGLD_FUSE_RES32", fuse)
        g, m, F, fixed, E = build_pair(vox, 4)
        g.set_schedule("tiled", 12, "tiled")
        U, its, rr = g.pcg(F, tol=1e-8, maxIT=2000, precision=solver.PREC_MIXED)
        st = g.stats()
        true = np.linalg.norm(F - m.kbyu(U)) / np.linalg.norm(F)
        print("  fuse %s: its %d, reliable updates %d, true rel-res %.2e" % (fuse, its, st["outer"], true), flush=True)
        assert g.converged and true <= 1.01e-8 and st["outer"] >= 10
        out[fuse] = (U, its)
        g.close()
    (U1, i1), (U0, i0) = out["1"], out["0"]
    assert abs(i1 - i0) <= 2
    assert np.linalg.norm(U1 - U0) <= 1e-7 * np.linalg.norm(U0)


@pytest.mark.parametrize("case", ["cuboid", "holes"])
def test_kbyu_and_pcg_run_to_run_identical(case):
    """K.u sums every node's element contributions in a fixed order, so FP64
    K.u results and whole FP64 / mixed solves repeat bit for bit."""
    vox = problems.cuboid(96, 48, 48) if case == "cuboid" else holes(64, 32, 32, 5)
    print("case: determinism", case, flush=True)
    g, m, F, fixed, E = build_pair(vox, 4)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(F.size) * (1 - fixed)
    ys = [g.kbyu(u) for _ in range(4)]
    for y in ys[1:]:
        assert np.array_equal(y, ys[0])
    for prec in (solver.PREC_FP64, solver.PREC_MIXED):
        U1, its1, _ = g.pcg(F, tol=1e-8, maxIT=500, precision=prec)
        U2, its2, _ = g.pcg(F, tol=1e-8, maxIT=500, precision=prec)
        assert its1 == its2 and np.array_equal(U1, U2), prec


def test_contexts_with_different_element_matrices_interleave():
    """Each context holds its own element matrix: two contexts with other
    Poisson ratios and cell sizes, used alternately (and from two threads),
    each match the oracle built with their own Ke."""
    import threading
    vox = holes(24, 16, 8, 4, nholes=2, r=2)
    F, fixed = problems.cantilever3(vox)
    E = problems.simp_modulus(np.random.default_rng(9).uniform(0.05, 1.0, int(vox.sum())))
    pairs = []
    for nu, h in ((0.3, 1.0), (0.2, 0.5)):
        g = solver.FEASolver(vox, 3, poissonRatio=nu, cellSize=h)
        g.set_fixed(fixed)
        g.set_modulus(E)
        g.assemble(0.6)
        m = orc.Mesh(vox, 3, 0, nu, h)
        m.set_fixed(fixed)
        m.set_modulus(E)
        m.assemble()
        pairs.append((g, m))
    u = np.random.default_rng(10).standard_normal(F.size) * (1 - fixed)
    for _ in range(2):
        for g, m in pairs:
            yo = m.kbyu(u)
            assert np.max(np.abs(g.kbyu(u) - yo)) <= 1e-13 * np.max(np.abs(yo))
    out = {}

    def run(k):
        g, m = pairs[k]
        out[k] = [g.pcg(F, tol=1e-8, maxIT=300) for _ in range(3)]

    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, (g, m) in enumerate(pairs):
        Uo, itso, _, _ = m.pcg(F, 0.6, 1e-8, 300)
        for U, its, rr in out[k]:
            assert rr <= 1e-8 and abs(its - itso) <= 1
            assert abs(F @ U - F @ Uo) <= 1e-9 * abs(F @ Uo)


def test_chunk_sizes_change_only_rounding(monkeypatch):
    """Irregular domains visit the 16-, 32- or 64-entry chunks of the DOF
    vectors that hold a solid node (chosen when a context is created).  The
    V-cycle and K.u are entry-wise identical whatever the chunking; the PCG
    reductions sum in another order, so solves agree to rounding."""
    vox = holes(64, 32, 32, 9, nholes=8, r=5)
    F, fixed = problems.cantilever3(vox)
    E = np.ones(int(vox.sum()))
    rng = np.random.default_rng(4)
    u = rng.standard_normal(F.size) * (1 - fixed)
    out = {}
    for log in (6, 5, 4):
        monkeypatch.setenv("SGLD_CHUNK_LOG", str(log))
        print("case: chunk log", log, flush=True)
        g = solver.FEASolver(vox, 4)
        g.set_fixed(fixed)
        g.set_modulus(E)
        g.assemble(0.6)
        res = {"kbyu": g.kbyu(u), "vcycle": g.vcycle(u)}
        for prec in (solver.PREC_FP64, solver.PREC_MIXED):
            U, its, rr = g.pcg(F, tol=1e-8, maxIT=500, precision=prec)
            assert g.converged
            res[prec] = (U, its)
        g.close()
        out[log] = res
    for log in (5, 4):
        assert np.array_equal(out[log]["kbyu"], out[6]["kbyu"])
        assert np.array_equal(out[log]["vcycle"], out[6]["vcycle"])
        for prec in (solver.PREC_FP64, solver.PREC_MIXED):
            U, its = out[log][prec]
            U6, its6 = out[6][prec]
            assert abs(its - its6) <= 2, (log, prec, its, its6)
            # both meet tol 1e-8 along nearly the same iterates
            assert np.max(np.abs(U - U6)) <= 1e-6 * np.max(np.abs(U6))
