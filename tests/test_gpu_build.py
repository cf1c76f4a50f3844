"""The GPU QBVH builder (vgpu_qbvh_build_gpu) against the host builder and the oracle.

qbvh.BuildAccel (qbvh/build.go:28-307) partitions each split with an unstable two-pointer loop
(:111-130).  The GPU builder computes where that loop leaves every element with a closed form
(each element alone, from prefix counts); the CPU tests below check the closed form against the
loop itself on every class sequence up to length 12 and on long random ones, so a wrong stream
boundary or an off-by-one in either rank fails here.

GPU tests: node arrays, the permuted boxes / centroids / indices and the root bounds byte-identical
to vgpu_qbvh_build (itself node-for-node equal to the oracle) on random boxes, meshes, grids with
shared coordinates and signed zeros, duplicate centroids (the flat-axis pivot n/2 + 1) and sizes
either side of the GPU builder's subtree and chunk sizes (2048 and 4096 elements), for leafMax
1, 4 and 16; degenerate input and non-finite centroids fail as on the host (a NaN box coordinate
fails on the GPU only).
"""
import itertools

import numpy as np
import pytest

import oracle
import vermeer_amd as va
from vermeer_amd import scenes


def _two_pointer(right):
    """qbvh/build.go:111-130 on a class sequence (True: bin >= best): where each position ends."""
    n = len(right)
    e = list(range(n))
    left, r = 0, n - 1
    while left <= r:
        if not right[e[left]]:
            left += 1
        else:
            e[left], e[r] = e[r], e[left]
            r -= 1
    return e, left


def _closed_form(right):
    """The GPU builder's destinations (bvh_build.hip dest_of / record_rank) in numpy."""
    right = np.asarray(right, bool)
    n = len(right)
    nl = int((~right).sum())
    s = nl + (1 if nl < n and right[nl] else 0)
    rank = np.cumsum(right)                       # right elements in [0, p], inclusive
    pos = np.arange(n)
    rl = pos[right]                               # rl[k-1]: position of the k-th right element
    lr = np.zeros(nl, np.int64)
    left_pos = pos[~right]
    k = nl - (left_pos - rank[~right])            # the k-th left element counted from n - 1
    lr[k - 1] = n - 1 - left_pos
    dest = np.empty(n, np.int64)
    for p in range(n):
        r = rank[p]
        if p < s:
            dest[p] = p if not right[p] else (n - 1 if r < 2 else n - 2 - lr[r - 2])
        else:
            dest[p] = p - 1 if right[p] else rl[nl - (p - r) - 1]
    out = [None] * n
    for p in range(n):
        assert out[dest[p]] is None
        out[dest[p]] = p
    return out, nl


def test_closed_form_partition_equals_two_pointer_loop_exhaustive():
    for n in range(0, 13):
        for cls in itertools.product((False, True), repeat=n):
            assert _closed_form(cls) == _two_pointer(cls), cls


@pytest.mark.parametrize("frac", [0.02, 0.5, 0.97])
def test_closed_form_partition_equals_two_pointer_loop_random(frac):
    g = np.random.default_rng(int(frac * 100))
    for n in (100, 1000, 5000):
        cls = g.uniform(size=n) < frac
        assert _closed_form(cls) == _two_pointer(list(cls))


# ---------------------------------------------------------------- GPU


def _tri_arrays(v, f):
    V = v[f.astype(np.int64)]
    lo = np.minimum(np.minimum(V[:, 0], V[:, 1]), V[:, 2])
    hi = np.maximum(np.maximum(V[:, 0], V[:, 1]), V[:, 2])
    cent = ((V[:, 0] + V[:, 1]) + V[:, 2]) / np.float32(3)
    return (np.ascontiguousarray(np.concatenate([lo, hi], 1), np.float32), np.ascontiguousarray(cent, np.float32),
            np.arange(len(f), dtype=np.int32))


def _random_boxes(n, seed, spread=10.0, size=0.5):
    g = np.random.default_rng(seed)
    c = g.uniform(-spread, spread, (n, 3))
    e = g.uniform(0.0, size, (n, 3))
    return (np.ascontiguousarray(np.concatenate([c - e, c + e], 1), np.float32), np.ascontiguousarray(c, np.float32),
            np.arange(n, dtype=np.int32))


def _gpu_equals_host(boxes, cent, idx, leaf_max, with_oracle=False):
    b1, c1, i1 = boxes.copy(), cent.copy(), idx.copy()
    b2, c2, i2 = boxes.copy(), cent.copy(), idx.copy()
    n1, bd1 = va.build_accel(b1, c1, i1, leaf_max)
    n2, bd2 = va.build_accel_gpu(b2, c2, i2, leaf_max)
    assert len(n1) == len(n2), (len(n1), len(n2))
    if n1.tobytes() != n2.tobytes():
        bad = np.nonzero([a.tobytes() != b.tobytes() for a, b in zip(n1, n2)])[0]
        raise AssertionError(f"{len(bad)} of {len(n1)} nodes differ, first {bad[:5]}: {n1[bad[0]]} vs {n2[bad[0]]}")
    assert np.array_equal(i1, i2)
    assert b1.tobytes() == b2.tobytes() and c1.tobytes() == c2.tobytes()
    assert bd1.tobytes() == bd2.tobytes()
    if with_oracle:
        b3, c3, i3 = boxes.copy(), cent.copy(), idx.copy()
        n3, bd3 = oracle.build_qbvh(b3, c3, i3, leaf_max)
        assert n3.tobytes() == n2.tobytes() and np.array_equal(i3, i2) and bd3.tobytes() == bd2.tobytes()
    return n2


@pytest.mark.gpu
@pytest.mark.parametrize("leaf_max", [1, 4, 16])
@pytest.mark.parametrize("n", [0, 1, 3, 17, 100, 2047, 2048, 2049, 4096, 4097, 9000])
def test_gpu_builder_small_and_boundary_sizes(n, leaf_max):
    print(f"\n[case] random boxes n={n} leafMax={leaf_max}", flush=True)
    _gpu_equals_host(*_random_boxes(n, 7 + n), leaf_max, with_oracle=n <= 4097)


@pytest.mark.gpu
@pytest.mark.parametrize("leaf_max", [1, 4, 16])
def test_gpu_builder_large_random(leaf_max):
    print(f"\n[case] random boxes n=300000 leafMax={leaf_max}", flush=True)
    nodes = _gpu_equals_host(*_random_boxes(300_000, 11, spread=100.0), leaf_max)
    assert len(nodes) > 300_000 // (4 * leaf_max)


@pytest.mark.gpu
def test_gpu_builder_meshes():
    for label, (v, f) in (("sphere", scenes.uv_sphere(1.0, 256, 128)),
                          ("icosphere field", scenes.icosphere_field(512, 3)[:2])):
        for leaf_max in (4, 16):
            print(f"\n[case] mesh {label} ({len(f)} triangles) leafMax={leaf_max}", flush=True)
            _gpu_equals_host(*_tri_arrays(v, f), leaf_max)


@pytest.mark.gpu
def test_gpu_builder_grid_shared_coordinates_and_signed_zeros():
    # centroids on an integer lattice around the origin: many equal coordinates, bins with
    # exact boundaries, flat axes in small ranges, and boxes touching +0 and -0
    g = np.random.default_rng(5)
    n = 60_000
    c = g.integers(-20, 21, (n, 3)).astype(np.float32)
    c[g.uniform(size=(n, 3)) < 0.1] = np.float32(-0.0)
    e = np.float32(0.5) * g.integers(0, 3, (n, 3)).astype(np.float32)
    lo = c - e
    hi = c + e
    lo[lo == 0] = np.where(g.uniform(size=(lo == 0).sum()) < 0.5, np.float32(-0.0), np.float32(0.0))
    boxes = np.ascontiguousarray(np.concatenate([lo, hi], 1), np.float32)
    for leaf_max in (2, 16):
        print(f"\n[case] lattice n={n} leafMax={leaf_max}", flush=True)
        _gpu_equals_host(boxes, np.ascontiguousarray(c), np.arange(n, dtype=np.int32), leaf_max)


@pytest.mark.gpu
def test_gpu_builder_duplicate_centroids_take_the_flat_pivot():
    # groups of 3 identical centroids: every range of one group is flat (pivot n/2 + 1); leafMax 2,
    # as two identical centroids under leafMax 1 recurse forever
    g = np.random.default_rng(9)
    base = g.uniform(-5, 5, (5000, 3)).astype(np.float32)
    c = np.repeat(base, 3, axis=0)
    boxes = np.ascontiguousarray(np.concatenate([c - 0.1, c + 0.1], 1), np.float32)
    print("\n[case] duplicate centroids n=15000 leafMax=2", flush=True)
    _gpu_equals_host(boxes, np.ascontiguousarray(c), np.arange(len(c), dtype=np.int32), 2)


@pytest.mark.gpu
def test_gpu_builder_reports_degenerate_and_non_finite_input():
    # two coincident centroids with leafMax 1: the reference recursion never ends
    c = np.zeros((2, 3), np.float32)
    boxes = np.zeros((2, 6), np.float32)
    with pytest.raises(va.VgpuError):
        va.build_accel(boxes.copy(), c.copy(), np.arange(2, dtype=np.int32), 1)
    with pytest.raises(va.VgpuError):
        va.build_accel_gpu(boxes.copy(), c.copy(), np.arange(2, dtype=np.int32), 1)
    b, cc, i = _random_boxes(5000, 3)
    cc[1234, 1] = np.nan
    with pytest.raises(va.VgpuError):
        va.build_accel_gpu(b, cc, i, 4)
    b, cc, i = _random_boxes(5000, 4)
    b[77, 4] = np.nan
    with pytest.raises(va.VgpuError):
        va.build_accel_gpu(b, cc, i, 4)
    # an infinite box makes every SAH cost of its ranges infinite: no bin is chosen (best = -1),
    # every element goes right and the recursion repeats the range, on both builders
    b, cc, i = _random_boxes(5000, 4)
    b[77, 4] = np.inf
    with pytest.raises(va.VgpuError):
        va.build_accel(b.copy(), cc.copy(), i.copy(), 4)
    with pytest.raises(va.VgpuError):
        va.build_accel_gpu(b.copy(), cc.copy(), i.copy(), 4)
    # the builder stays usable after a failure
    _gpu_equals_host(*_random_boxes(3000, 5), 4)


@pytest.mark.gpu
def test_gpu_builder_full_icosphere_field():
    """The 4096-sphere icosphere field (5.2M triangles) at the reference's mesh leaf size, 16."""
    import time
    v, f = scenes.icosphere_field()[:2]
    boxes, cent, idx = _tri_arrays(v, f)
    print(f"\n[case] icosphere field ({len(f)} triangles) leafMax=16", flush=True)
    b1, c1, i1 = boxes.copy(), cent.copy(), idx.copy()
    b2, c2, i2 = boxes.copy(), cent.copy(), idx.copy()
    t0 = time.perf_counter()
    n1, bd1 = va.build_accel(b1, c1, i1, 16)
    t1 = time.perf_counter()
    n2, bd2 = va.build_accel_gpu(b2, c2, i2, 16)
    t2 = time.perf_counter()
    print(f"host {1e3 * (t1 - t0):.0f} ms, gpu {1e3 * (t2 - t1):.0f} ms (first call)", flush=True)
    assert n1.tobytes() == n2.tobytes() and np.array_equal(i1, i2)
    assert b1.tobytes() == b2.tobytes() and c1.tobytes() == c2.tobytes() and bd1.tobytes() == bd2.tobytes()


@pytest.mark.gpu
def test_gpu_builder_concurrent_calls_share_the_pinned_staging():
    """Builds large enough to go through the pinned staging slots (several 32 MiB chunks each way),
    run from two threads at once (ctypes releases the GIL): each equals the host builder's."""
    import threading
    cases = [_random_boxes(900_000, 21 + k, spread=100.0) for k in range(2)]
    want = []
    for boxes, cent, idx in cases:
        b, c, i = boxes.copy(), cent.copy(), idx.copy()
        nodes, bd = va.build_accel(b, c, i, 4)
        want.append((nodes.tobytes(), b.tobytes(), c.tobytes(), i.copy(), bd.tobytes()))
    got = [None, None]
    errors = []

    def work(k):
        try:
            boxes, cent, idx = cases[k]
            for _ in range(2):
                b, c, i = boxes.copy(), cent.copy(), idx.copy()
                nodes, bd = va.build_accel_gpu(b, c, i, 4)
                got[k] = (nodes.tobytes(), b.tobytes(), c.tobytes(), i, bd.tobytes())
        except Exception as e:   # reported by the main thread
            errors.append(e)

    print("\n[case] two threads, 900000 boxes each, leafMax=4", flush=True)
    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert got[k][0] == want[k][0] and got[k][1] == want[k][1] and got[k][2] == want[k][2]
        assert np.array_equal(got[k][3], want[k][3]) and got[k][4] == want[k][4]
