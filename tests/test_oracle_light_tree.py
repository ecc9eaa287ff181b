"""The oracle's light tree (buildLightTree / packLightTreeForGPU / packed
sampling and pdf) pinned by the known answers of
packages/shared-samplers/__tests__/lightTree.test.ts and
lightTreeTraversal.test.ts, and by the partition invariant: for every shading
point the leaf pmf is an integer number of 2^-24 buckets and sums to exactly 1.
"""
import numpy as np
import pytest

from oracle import oracle as O

FLOOR = 0.01
B = O.LT_BUCKETS


def point_aabb(c):
    c = np.asarray(c, np.float64)
    return c - 0.5, c + 0.5


def make_input(powers, spacing=2.0):
    """lightTreeTraversal.test.ts:43-49: emitters along +x."""
    cs = [(i * spacing, 0.0, 0.0) for i in range(len(powers))]
    mn = [point_aabb(c)[0] for c in cs]
    mx = [point_aabb(c)[1] for c in cs]
    return powers, cs, mn, mx


def build(powers, spacing=2.0, cones=None):
    p, c, mn, mx = make_input(powers, spacing)
    return O.light_tree_build(p, c, mn, mx, cones)


def pmf(packed, x, n):
    return np.array([O.light_tree_pdf(packed, x, FLOOR, e) for e in range(n)])


def lcg(seed):
    s = seed & 0xFFFFFFFF
    while True:
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        yield s / 2 ** 32


def test_single_and_pair_structure():
    """lightTree.test.ts:56-101: one emitter is one leaf; two emitters are a
    root with two leaves, root power = sum (represented: root 1.0)."""
    one = build([5.0])
    assert one.shape == (1, 16) and one[0, 0] == 0 and one[0, 2] == -1 and one[0, 3] == -1
    assert one[0, 15] == 1 and one[0, 1] == 1.0
    two = build([3.0, 7.0])
    assert two.shape == (3, 16)
    assert two[0, 0] == -1 and two[0, 1] == 1.0 and two[0, 15] == 2
    kids = [int(two[0, 2]), int(two[0, 3])]
    assert sorted(int(two[k, 0]) for k in kids) == [0, 1]
    assert np.isclose(two[kids[0], 1] + two[kids[1], 1], 1.0)
    # preorder, left first: emitter at x=0 (lower centroid) is the left child
    assert two[kids[0], 0] == 0 and kids == [1, 2]


def test_packed_layout_of_unoriented_nodes():
    """lightTree.test.ts:461-512: 16 floats per node; leaves store the
    represented power and leaf count 1; unoriented nodes pack axis 0 and
    both cosine thresholds -1."""
    t = build([1.0, 2.0, 4.0, 8.0])
    assert t.shape == (7, 16)
    leaves = t[t[:, 2] < 0]
    assert np.all(leaves[:, 15] == 1) and t[0, 15] == 4
    assert np.all(t[:, 10:13] == 0) and np.all(t[:, 13] == -1) and np.all(t[:, 14] == -1)
    # represented power = node power / max node power (the root, 15)
    for row in leaves:
        e = int(row[0])
        assert row[1] == np.float32([1.0, 2.0, 4.0, 8.0][e] / 15.0)


def test_doubling_powers_keeps_represented_tree():
    """lightTree.test.ts:127-136: scaling every power by 2 scales every node
    power by 2; the represented (normalised) tree is unchanged."""
    a = build([1.0, 3.0, 2.0, 9.0, 4.0])
    b = build([2.0, 6.0, 4.0, 18.0, 8.0])
    assert np.array_equal(a, b)


def test_colocated_centroids_split_by_power():
    """lightTree.test.ts:154-197: co-located centroids fall back to a
    power-sorted median split (descending power, then index)."""
    n = 6
    powers = [1.0, 6.0, 3.0, 5.0, 2.0, 4.0]
    c = [(0.0, 0.0, 0.0)] * n
    mn = [(-0.5, -0.5, -0.5)] * n
    mx = [(0.5, 0.5, 0.5)] * n
    t = O.light_tree_build(powers, c, mn, mx)
    assert t.shape == (2 * n - 1, 16)
    left = int(t[0, 2])

    def leaves_under(i):
        if t[i, 2] < 0:
            return [int(t[i, 0])]
        return leaves_under(int(t[i, 2])) + leaves_under(int(t[i, 3]))

    # the three brightest emitters (6, 5, 4) form the left subtree
    assert sorted(leaves_under(left)) == [1, 3, 5]


def test_exact_bucket_pmfs_at_power_extremes():
    """lightTreeTraversal.test.ts:169-180: min and max positive powers both
    keep at least one bucket, pmfs are bucket multiples summing to 1."""
    t = build([5e-324, 1.7976931348623157e308])
    p = pmf(t, [0, 0, 0], 2)
    assert np.all(p >= 1 / B) and p.sum() == 1.0
    assert np.all(p * B == np.round(p * B))


def test_boundary_buckets_select_exact_leaf_intervals():
    """lightTreeTraversal.test.ts:182-205."""
    t = build([5e-324, 1.7976931348623157e308])
    left_buckets = int(round(O.light_tree_pdf(t, [0, 0, 0], FLOOR, 0) * B))
    assert left_buckets >= 1
    e0, p0 = O.light_tree_sample(t, [0, 0, 0], FLOOR, left_buckets - 1)
    e1, p1 = O.light_tree_sample(t, [0, 0, 0], FLOOR, left_buckets)
    assert (e0, e1) == (0, 1)
    assert p0 == O.light_tree_pdf(t, [0, 0, 0], FLOOR, 0) and p1 == O.light_tree_pdf(t, [0, 0, 0], FLOOR, 1)


def test_deep_power_comb_keeps_every_leaf():
    """lightTreeTraversal.test.ts:51-99, :207-219: a 320-level comb of halving
    powers; every leaf keeps >= 1 bucket and the pmf sums to exactly 1."""
    n = 320
    powers = [2.0 ** -(i + 1) if i < n - 1 else 2.0 ** -(n - 1) for i in range(n)]
    nodes = []

    def append(first):
        idx = len(nodes)
        if first == n - 1:
            nodes.append(dict(emitter=first, power=powers[first], left=-1, right=-1))
            return idx
        nodes.append(None)
        li = len(nodes)
        nodes.append(dict(emitter=first, power=powers[first], left=-1, right=-1))
        ri = append(first + 1)
        nodes[idx] = dict(emitter=-1, power=nodes[li]["power"] + nodes[ri]["power"], left=li, right=ri)
        return idx

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(5000)
    try:
        append(0)
    finally:
        sys.setrecursionlimit(old)
    t = O.light_tree_pack_nodes(nodes)
    assert t[0, 15] == 320
    p = pmf(t, [0, 0, 0], n)
    assert np.all(p >= 1 / B) and np.all(p * B == np.round(p * B)) and p.sum() == 1.0


def test_single_leaf_tree_ignores_the_draw():
    """lightTreeTraversal.test.ts:257-270: a one-leaf tree returns (0, 1)
    whatever the draw (no RNG consumed); an empty tree returns (-1, 0)."""
    t = build([1.0])
    for bucket in (0, 123456, B - 1):
        assert O.light_tree_sample(t, [0, 0, 0], FLOOR, bucket) == (0, 1.0)
    assert O.light_tree_sample(np.zeros((0, 16), np.float32), [0, 0, 0], FLOOR, 5) == (-1, 0.0)


@pytest.mark.parametrize("x", [(0, 0, 0), (3.3, 0.2, -1), (40, 5, 5), (-7, 0, 2)])
def test_pmf_is_an_exact_partition(x):
    """lightTreeTraversal.test.ts:106-160: the pmf sums to 1 at any shading
    point (here exactly, in buckets), for 4, 16 and 64 emitters."""
    rng = np.random.default_rng(1)
    for n in (4, 16, 64):
        pw = list(rng.uniform(0.1, 10, n))
        p, c, mn, mx = make_input(pw)
        if n == 64:
            g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3) * 3.0
            c = g
            mn, mx = g - 0.5, g + 0.5
        t = O.light_tree_build(pw, c, mn, mx)
        q = pmf(t, x, n)
        assert np.all(q > 0) and np.all(q * B == np.round(q * B)) and q.sum() == 1.0


def test_sampled_pdf_equals_deterministic_pdf_and_frequencies():
    """lightTreeTraversal.test.ts:273-313: the drawn pdf is the deterministic
    pdf of the drawn leaf; empirical frequencies are within 3 SE."""
    t = build([3.0, 1.0, 5.0, 2.0])
    x = [3.3, 0.2, -1]
    g = lcg(0xFEED)
    for _ in range(200):
        e, pdf = O.light_tree_sample(t, x, FLOOR, int(next(g) * B))
        assert e >= 0 and pdf > 0 and pdf == O.light_tree_pdf(t, x, FLOOR, e)
    t = build([1.0, 4.0, 2.0, 8.0, 5.0])
    x = [4, 0, 0]
    exp = pmf(t, x, 5)
    g = lcg(0xC0FFEE)
    n = 20000
    counts = np.zeros(5)
    for _ in range(n):
        counts[O.light_tree_sample(t, x, FLOOR, int(next(g) * B))[0]] += 1
    se = np.sqrt(np.maximum(exp * (1 - exp) * n, 1))
    assert np.all(np.abs(counts - exp * n) <= 3 * se)


def test_spatial_importance():
    """lightTreeTraversal.test.ts:316-364."""
    t = build([1.0, 1.0], spacing=100)
    assert O.light_tree_pdf(t, [0, 0, 0], FLOOR, 0) > 0.9
    assert O.light_tree_pdf(t, [100, 0, 0], FLOOR, 1) > O.light_tree_pdf(t, [100, 0, 0], FLOOR, 0)
    t = build([100.0, 1.0], spacing=10)
    assert O.light_tree_pdf(t, [5, 0, 0], FLOOR, 0) > O.light_tree_pdf(t, [5, 0, 0], FLOOR, 1)
    cs = [(0, 0, 0), (5, 0, 0), (50, 0, 0)]
    t = O.light_tree_build([1.0] * 3, cs, [point_aabb(c)[0] for c in cs], [point_aabb(c)[1] for c in cs])
    near = O.light_tree_pdf(t, [2.5, 0, 0], FLOOR, 0) + O.light_tree_pdf(t, [2.5, 0, 0], FLOOR, 1)
    assert near > 2 / 3 and near + O.light_tree_pdf(t, [2.5, 0, 0], FLOOR, 2) == 1.0


def test_zero_power_pair_splits_evenly():
    """lightTreeTraversal.test.ts:367-379."""
    t = build([0.0, 0.0])
    e, pdf = O.light_tree_sample(t, [0, 0, 0], FLOOR, 12345)
    assert e >= 0 and pdf == 0.5


def test_orientation_cones():
    """lightTreeTraversal.test.ts:380-517: back-facing narrow emitters are
    culled; a lit point just in front of a tilted triangle keeps mass;
    zero-axis cones equal no cones; all-culled descents still end at a leaf."""
    cs = [(-4, 0, 0), (4, 0, 0)]
    mn = [point_aabb(c)[0] for c in cs]
    mx = [point_aabb(c)[1] for c in cs]
    narrow = ([1, 0, 0], 0.0, np.pi / 18)
    t = O.light_tree_build([10.0, 10.0], cs, mn, mx, [narrow, narrow])
    front, back = O.light_tree_pdf(t, [0, 0, 0], FLOOR, 0), O.light_tree_pdf(t, [0, 0, 0], FLOOR, 1)
    assert back < 1e-6 and front > 0.999 and front + back == 1.0
    s3 = 1 / np.sqrt(3)
    nrm = (s3, -s3, s3)
    t = O.light_tree_build([1.0, 1.0], [(0.5, 0.5, 0.5), (10, 0, 0)], [(0, 0, 0), point_aabb((10, 0, 0))[0]],
                           [(1, 1, 1), point_aabb((10, 0, 0))[1]], [(nrm, 0.0, np.pi / 2), None])
    lit = [nrm[0] * 0.05, nrm[1] * 0.05, nrm[2] * 0.05]
    assert O.light_tree_pdf(t, lit, FLOOR, 0) > 0
    a = O.light_tree_build([10.0, 10.0], cs, mn, mx)
    b = O.light_tree_build([10.0, 10.0], cs, mn, mx, [([0, 0, 0], 0.0, np.pi / 2)] * 2)
    assert np.array_equal(a, b) and O.light_tree_pdf(a, [0, 0, 0], FLOOR, 0) == 0.5
    away = ([-1, 0, 0], 0.0, np.pi / 18)
    t = O.light_tree_build([10.0, 10.0], cs, mn, mx, [away, away])
    e, pdf = O.light_tree_sample(t, [100, 0, 0], FLOOR, 777)
    assert e >= 0 and pdf > 0


def test_cone_thresholds_round_outward():
    """conservativeConeCosineThreshold (lightTree.ts:193-212): the packed
    cosine of a sharp hemisphere leaf sits at least 64 f32 steps and 2^-18
    below the rounded cosine."""
    cs = [(0, 0, 0), (9, 0, 0)]
    t = O.light_tree_build([1.0, 1.0], cs, [point_aabb(c)[0] for c in cs], [point_aabb(c)[1] for c in cs],
                           [([0, 0, 1], 0.0, np.pi / 2), None])
    leaf = t[(t[:, 0] == 0) & (t[:, 2] < 0)][0]
    assert leaf[13] <= np.float32(1.0) - 2.0 ** -18 and leaf[13] < np.nextafter(np.float32(1), np.float32(0))
    assert leaf[14] <= np.float32(np.cos(np.pi / 2)) - 2.0 ** -18
    assert np.allclose(leaf[10:13], [0, 0, 1])
