"""The committed rocprofv3 traffic summaries against bench.py's byte model.

bench.py fills the JSON line's `roofline.traffic` from the per-launch lines of
the committed summaries and prices `roofline.achieved` with `kernel_bytes`.
A summary whose kernel names no longer match the kernels that run would leave
`traffic` null, and measured HBM bytes below the algorithmic ones would mean a
wrong byte model or a wrong calibration factor; both fail here."""
import pytest

import bench

# 125 M-DOF ellipsoid union (512^3 box): level-1 and level-2 DOFs, elements
ELLIPSOIDS = dict(n=124_938_039, n2=16_184_880, ne=40_809_958)
# 256 x 128 x 128 cuboid: 3 * 257 * 129 * 129 DOFs, 256 * 128 * 128 elements
CUBOID = dict(n=3 * 257 * 129 * 129, n2=0, ne=256 * 128 * 128)

CASES = [
    ("cfg2", "dir32", "k_prolong_dir<float, 2, true>", ELLIPSOIDS),
    ("cfg2", "kbyu32", "k_kbyu_owner<float, true>", ELLIPSOIDS),
    ("cfg2", "update_r32", "k_update_restrict<float>", ELLIPSOIDS),
    ("cfg2", "dir64", "k_prolong_dir<double, 2, true>", ELLIPSOIDS),
    ("cfg2", "kbyu64", "k_kbyu_owner<double, true>", ELLIPSOIDS),
    ("cfg2", "update_r64", "k_update_restrict<double>", ELLIPSOIDS),
    ("cfg1", "kbyu64", "k_kbyu_owner<double, false>", CUBOID),
    ("cfg1", "kbyu32", "k_kbyu_owner<float, false>", CUBOID),
]


@pytest.mark.parametrize("workload,key,kernel,size", CASES)
def test_measured_traffic_covers_algorithmic_bytes(workload, key, kernel, size):
    traffic = bench.measured_traffic(workload, kernel)
    assert traffic is not None, (workload, kernel)
    alg = bench.kernel_bytes(key, size["n"], size["n2"], size["ne"], kernel)
    # at least the algorithmic bytes; at most the K.u halo re-reads (about 2x)
    assert 1.0 <= traffic / alg <= 2.2, (kernel, traffic / alg)


def test_kernel_bytes_of_the_direction_update():
    # 22 B per level-1 DOF in FP32 (r, p, x, p', x' and a bfloat16 weight)
    # plus x_2; 48 B in FP64 (FP64 weight)
    n, n2 = ELLIPSOIDS["n"], ELLIPSOIDS["n2"]
    assert bench.kernel_bytes("dir32", n, n2, 0) == 22 * n + 4 * n2
    assert bench.kernel_bytes("dir64", n, n2, 0) == 48 * n + 8 * n2
    assert bench.kernel_bytes("update_r32", n, n2, 0, "k_update_restrict<float>") == 14 * n + 4 * n2
    assert bench.kernel_bytes("update_r32", n, n2, 0, "k_update_r<float, true>") == 14 * n
