"""Many instances in one replayed graph, compliance sensitivities, and the
autotuner's schedules, each against the oracle."""
import numpy as np
import pytest

from oracle import oracle as orc
from sgldbench_amd import problems, solver

from tests.test_gpu_solver import holes

pytestmark = pytest.mark.gpu


def instance(kind, seed):
    if kind == "holes":
        vox, L = holes(32, 16, 16, seed), 3
    else:
        nx, ny, nz = kind
        vox, L = problems.cuboid(nx, ny, nz), problems.levels_for(kind, 4)
    F, fixed = problems.cantilever3(vox)
    x = problems.gaussian_random_density(vox.shape, seed, sigma=2.0).ravel(order="F")
    x = x[vox.ravel(order="F") != 0]
    return dict(voxels=vox, L=L, F=F, fixed=fixed, x=x, E=problems.simp_modulus(x))


KINDS = [(32, 16, 16), (24, 16, 8), "holes", (16, 16, 16), (40, 8, 24)]


@pytest.fixture(scope="module")
def fleet():
    print("case: multi fleet", KINDS, flush=True)
    out = []
    for i, k in enumerate(KINDS):
        d = instance(k, 30 + i)
        g = solver.FEASolver(d["voxels"], d["L"])
        g.set_fixed(d["fixed"])
        g.set_modulus(d["E"])
        g.assemble(0.6)
        m = orc.Mesh(d["voxels"], d["L"])
        m.set_fixed(d["fixed"])
        m.set_modulus(d["E"])
        m.assemble()
        ref, its, hist, rc = m.pcg(d["F"], 0.6, 1e-8, 1000)
        assert rc == 0
        out.append((g, m, d, ref, its))
    return out


@pytest.mark.parametrize("prec", [solver.PREC_FP64, solver.PREC_MIXED])
@pytest.mark.parametrize("batch", [8, 2])
def test_multi_matches_oracle(fleet, prec, batch):
    import torch
    dF = [torch.from_numpy(d["F"]).cuda() for _, _, d, _, _ in fleet]
    dU = [torch.zeros_like(f) for f in dF]
    its, rr, lat, ok = solver.multi_pcg_device([g for g, *_ in fleet], [f.data_ptr() for f in dF],
                                               [u.data_ptr() for u in dU], tol=1e-8, maxIT=1000,
                                               precision=prec, batch=batch)
    torch.cuda.synchronize()
    assert ok and np.all(lat > 0)
    for k, (g, m, d, ref, its_o) in enumerate(fleet):
        U = dU[k].cpu().numpy()
        true = np.linalg.norm(d["F"] - m.kbyu(U)) / np.linalg.norm(d["F"])
        assert true <= 1.01e-8 and rr[k] <= 1.01e-8, (k, true, rr[k])
        c, c_ref = d["F"] @ U, d["F"] @ ref
        tolc = 1e-9 if prec == solver.PREC_FP64 else 1e-8
        assert abs(c - c_ref) <= tolc * abs(c_ref), k
        if prec == solver.PREC_FP64:
            assert abs(int(its[k]) - its_o) <= 1, (k, its[k], its_o)


def test_multi_equals_single(fleet, monkeypatch):
    import torch
    monkeypatch.setenv("SGLD_MULTI_SOFT", "0")  # stopped stretches idle until their update, as in a single solve
    g, m, d, ref, _ = fleet[2]
    dF = torch.from_numpy(d["F"]).cuda()
    dU = torch.zeros_like(dF)
    its, rr, lat, ok = solver.multi_pcg_device([g], [dF.data_ptr()], [dU.data_ptr()], precision=solver.PREC_MIXED)
    U1, its1, rr1 = g.pcg(d["F"], tol=1e-8, maxIT=1000, precision=solver.PREC_MIXED)
    torch.cuda.synchronize()
    U = dU.cpu().numpy()
    assert ok and abs(int(its[0]) - its1) <= 1
    assert np.linalg.norm(U - U1) <= 1e-6 * np.linalg.norm(U1)


@pytest.mark.parametrize("prec", [solver.PREC_FP64, solver.PREC_MIXED])
def test_multi_merged_equals_single_solves(prec):
    """Every instance's iteration runs in the merged launches with its own
    grid and arguments, so each instance's solution is its single solve's
    bit for bit (one instance on the dense node-centric K.u, one irregular).
    The merged solve runs the V-cycle level by level, so the single solves
    here do too (SGLD_TAIL=0 when the contexts are created)."""
    import os
    import torch
    kinds = [(32, 16, 16), "holes", (24, 16, 8), (16, 16, 16)]
    print("case: merged multi vs single solves", kinds, prec, flush=True)
    solvers, data = [], []
    old_tail = os.environ.get("SGLD_TAIL")
    os.environ["SGLD_TAIL"] = "0"
    old_soft = os.environ.get("SGLD_MULTI_SOFT")
    os.environ["SGLD_MULTI_SOFT"] = "0"  # stopped stretches idle until their update, as in a single solve
    for i, k in enumerate(kinds):
        d = instance(k, 50 + i)
        g = solver.FEASolver(d["voxels"], d["L"])
        g.set_fixed(d["fixed"])
        g.set_modulus(d["E"])
        g.assemble(0.6)
        if i == 3:
            g.set_schedule("dense", 16, None)
        solvers.append(g)
        data.append(d)
    if old_tail is None:
        del os.environ["SGLD_TAIL"]
    else:
        os.environ["SGLD_TAIL"] = old_tail
    dF = [torch.from_numpy(d["F"]).cuda() for d in data]
    dU = [torch.zeros_like(f) for f in dF]
    its, rr, lat, ok = solver.multi_pcg_device(solvers, [f.data_ptr() for f in dF], [u.data_ptr() for u in dU],
                                               tol=1e-8, maxIT=1000, precision=prec, batch=8)
    torch.cuda.synchronize()
    if old_soft is None:
        del os.environ["SGLD_MULTI_SOFT"]
    else:
        os.environ["SGLD_MULTI_SOFT"] = old_soft
    assert ok
    for k, (g, d) in enumerate(zip(solvers, data)):
        U1, its1, rr1 = g.pcg(d["F"], tol=1e-8, maxIT=1000, precision=prec)
        U = dU[k].cpu().numpy()
        assert int(its[k]) == its1, (k, its[k], its1)
        assert np.array_equal(U, U1), (k, np.abs(U - U1).max())
    for g in solvers:
        g.close()


def test_sensitivity_matches_oracle(fleet):
    for g, m, d, ref, _ in fleet:
        penal, E0, Emin, scale = 3.0, 1.0, 1e-9, 0.37
        dc, c = g.sensitivity(ref, d["x"], penal, E0, Emin, scale)
        ce = m.unit_compliance(ref)
        want = -penal * (E0 - Emin) * d["x"] ** (penal - 1) * ce * scale
        assert np.abs(dc - want).max() <= 1e-10 * np.abs(want).max()
        assert abs(c - scale * d["E"] @ ce) <= 1e-10 * abs(scale * d["E"] @ ce)
        # at the solution, compliance = F.U
        assert abs(c / scale - d["F"] @ ref) <= 1e-8 * abs(d["F"] @ ref)


def test_multi_sensitivity_equals_single(fleet):
    """One launch over every instance's elements (full cuboids read U in
    place, the holed mesh scatters it first) gives each instance's own
    sgld_ctx_sensitivity_device results bit for bit."""
    import torch
    penal, E0, Emin, scale = 3.0, 1.0, 1e-9, 0.37
    dU = [torch.from_numpy(ref).cuda() for _, _, _, ref, _ in fleet]
    dx = [torch.from_numpy(d["x"]).cuda() for _, _, d, _, _ in fleet]
    one = [torch.zeros_like(x) for x in dx]
    many = [torch.zeros_like(x) for x in dx]
    c1 = [g.sensitivity_device(u.data_ptr(), x.data_ptr(), c.data_ptr(), penal, E0, Emin, scale)
          for (g, *_), u, x, c in zip(fleet, dU, dx, one)]
    cm = solver.multi_sensitivity_device([g for g, *_ in fleet], [u.data_ptr() for u in dU],
                                         [x.data_ptr() for x in dx], [c.data_ptr() for c in many],
                                         penal, E0, Emin, scale)
    torch.cuda.synchronize()
    for k, (g, m, d, ref, _) in enumerate(fleet):
        assert cm[k] == c1[k], (k, cm[k], c1[k])
        assert torch.equal(many[k], one[k]), k
        ce = m.unit_compliance(ref)
        want = -penal * (E0 - Emin) * d["x"] ** (penal - 1) * ce * scale
        assert np.abs(many[k].cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max(), k
        assert abs(cm[k] - scale * d["E"] @ ce) <= 1e-10 * abs(scale * d["E"] @ ce), k


@pytest.mark.parametrize("sched", [("tiled", 8, "node-centric"), ("tiled", 32, "tiled"), ("dense", 16, None),
                                   ("tiled", 0, "tiled-first")])
def test_schedules_match_oracle(sched):
    # 37 node layers: z-chunks of 8 and 32 leave a partial last chunk
    vox = holes(40, 24, 36, 5, nholes=4, r=3)
    F, fixed = problems.cantilever3(vox)
    E = problems.simp_modulus(problems.gaussian_random_density(vox.shape, 3, sigma=2.0).ravel(order="F")[vox.ravel(order="F") != 0])
    g = solver.FEASolver(vox, 3)
    g.set_fixed(fixed)
    g.set_modulus(E)
    g.assemble(0.6)
    g.set_schedule(*sched)
    m = orc.Mesh(vox, 3)
    m.set_fixed(fixed)
    m.set_modulus(E)
    m.assemble()
    u = np.random.default_rng(1).standard_normal(g.numDOFs)
    ref = m.kbyu(u)
    assert np.abs(g.kbyu(u) - ref).max() <= 1e-13 * np.abs(ref).max()
    r = m.vcycle(u, 0.6)
    assert np.abs(g.vcycle(u) - r).max() <= 1e-11 * np.abs(r).max()


def test_autotune_keeps_parity():
    vox = problems.cuboid(48, 32, 32)
    F, fixed = problems.cantilever3(vox)
    g = solver.FEASolver(vox, 4)
    g.set_fixed(fixed)
    g.set_modulus(np.ones(int(vox.sum())))
    g.assemble(0.6)
    rep = g.autotune(reps=3)
    assert rep["kbyu_variant"] in ("tiled", "dense") and rep["kbyu_ms_fp32"] > 0
    m = orc.Mesh(vox, 4)
    m.set_fixed(fixed)
    m.set_modulus(np.ones(int(vox.sum())))
    m.assemble()
    U, its, rr = g.pcg(F, tol=1e-8, maxIT=500)
    ref, its_o, hist, rc = m.pcg(F, 0.6, 1e-8, 500)
    assert abs(its - its_o) <= 1 and abs(F @ U - F @ ref) <= 1e-9 * abs(F @ ref)


def test_contexts_with_different_poisson_ratio():
    """Each context's element matrix is used even when another context with a
    different one ran in between (constant memory is re-uploaded)."""
    vox = problems.cuboid(16, 8, 8)
    F, fixed = problems.cantilever3(vox)
    E = np.ones(int(vox.sum()))
    out = {}
    for nu in (0.3, 0.2):
        g = solver.FEASolver(vox, 3, poissonRatio=nu)
        g.set_fixed(fixed)
        g.set_modulus(E)
        g.assemble(0.6)
        out[nu] = g
    u = np.random.default_rng(2).standard_normal(out[0.3].numDOFs)
    for nu in (0.3, 0.2, 0.3):
        m = orc.Mesh(vox, 3, nu=nu)
        m.set_fixed(fixed)
        m.set_modulus(E)
        ref = m.kbyu(u)
        assert np.abs(out[nu].kbyu(u) - ref).max() <= 1e-13 * np.abs(ref).max(), nu
