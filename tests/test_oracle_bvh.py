"""Pin the oracle's Morton codes, LBVH (Karras 2012) and traversal.

Morton codes are recomputed here with numpy float32 in the same operation
order; the hierarchy is checked against the binary radix tree built top-down
by splitting at the highest differing key bit (what Karras' bottom-up
construction must reproduce); traversal is checked against brute force over
every triangle, and against closed-form axis-aligned hits."""
import numpy as np
import pytest

import oracle_lib

LEAF = 0x80000000


def soup(n, seed, spread=10.0, size=0.5):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    return (c + rng.uniform(-size, size, (n, 3, 3))).astype(np.float32)


def expand10(v):
    v = v.astype(np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v.astype(np.uint32)


def morton_numpy(pos):
    p = pos.astype(np.float32)
    c = ((p[:, 0] + p[:, 1]) + p[:, 2]) * np.float32(1.0 / 3.0)
    lo, hi = c.min(0), c.max(0)
    ext = hi - lo
    with np.errstate(invalid="ignore", divide="ignore"):
        n = np.where(ext > 0, (c - lo) / np.where(ext > 0, ext, 1), np.float32(0))
    q = np.minimum(np.maximum(n * np.float32(1024.0), np.float32(0)), np.float32(1023)).astype(np.uint32)
    return (expand10(q[:, 0]) << 2) | (expand10(q[:, 1]) << 1) | expand10(q[:, 2])


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (1000, 2), (5000, 3)])
def test_morton_matches_numpy_restatement(n, seed):
    pos = soup(n, seed)
    codes, cb = oracle_lib.morton(pos)
    assert np.array_equal(codes, morton_numpy(pos))
    assert codes.max() < (1 << 30)


def radix_tree(keys):
    """Top-down binary radix tree over (key, index) augmented keys; returns
    {(first, last): (left, right)} with children as ("leaf", i) or (first, last)."""
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]
    out = {}

    def build(a, b):
        if a == b:
            return ("leaf", a)
        x = aug[a] ^ aug[b]
        bit = x.bit_length() - 1
        split = a
        while split < b and not ((aug[split + 1] >> bit) & 1):
            split += 1
        left = build(a, split)
        right = build(split + 1, b)
        out[(a, b)] = (left, right)
        return (a, b)

    build(0, len(keys) - 1)
    return out


def karras_to_ranges(left, right, n):
    """Map each Karras internal node to the (first, last) leaf range it covers."""
    ranges = {}

    def cover(ref):
        if ref & LEAF:
            k = ref & ~LEAF
            return (k, k)
        if ref in ranges:
            return ranges[ref]
        a = cover(int(left[ref]))
        b = cover(int(right[ref]))
        ranges[ref] = (a[0], b[1])
        return ranges[ref]

    for i in range(n - 1):
        cover(i)
    return ranges


@pytest.mark.parametrize("n,seed,dup", [(2, 1, 1), (3, 2, 1), (64, 3, 1), (500, 4, 1), (60, 5, 4), (300, 6, 10)])
def test_lbvh_equals_binary_radix_tree(n, seed, dup):
    pos = np.repeat(soup(n // dup, seed), dup, axis=0) if dup > 1 else soup(n, seed)
    n = pos.shape[0]
    r = oracle_lib.lbvh(pos, rounds=0)
    codes, order = r["codes"], r["order"]
    raw = morton_numpy(pos)
    # stable sort by code, ties by triangle index
    assert np.array_equal(order, np.argsort(raw, kind="stable"))
    assert np.array_equal(codes, raw[order])
    ranges = karras_to_ranges(r["left"], r["right"], n)
    tree = radix_tree(codes)
    assert sorted(ranges.values()) == sorted(tree.keys())
    inv = {v: k for k, v in ranges.items()}
    for i in range(n - 1):
        rng = ranges[i]
        tl, tr = tree[rng]

        def ref_of(t):
            return (LEAF | t[1]) if t[0] == "leaf" else inv[t]

        assert int(r["left"][i]) == ref_of(tl)
        assert int(r["right"][i]) == ref_of(tr)
    assert ranges[0] == (0, n - 1)  # node 0 is the root


def test_lbvh_boxes_contain_children_and_triangles():
    pos = soup(3000, 9)
    r = oracle_lib.lbvh(pos)
    n = pos.shape[0]
    lo, hi = r["lo"], r["hi"]
    tri = pos[r["order"]]
    assert np.array_equal(lo[n - 1:], tri.min(1))
    assert np.array_equal(hi[n - 1:], tri.max(1))
    for i in range(n - 1):
        for c in (int(r["left"][i]), int(r["right"][i])):
            j = (n - 1 + (c & ~LEAF)) if c & LEAF else c
            assert np.all(lo[i] <= lo[j]) and np.all(hi[i] >= hi[j])
            assert int(r["parent"][j]) == i
    # each internal box is exactly the union of its children
    for i in range(n - 1):
        js = [((n - 1 + (c & ~LEAF)) if c & LEAF else c) for c in (int(r["left"][i]), int(r["right"][i]))]
        assert np.array_equal(lo[i], np.minimum(lo[js[0]], lo[js[1]]))
        assert np.array_equal(hi[i], np.maximum(hi[js[0]], hi[js[1]]))


@pytest.mark.parametrize("seed,spread,size", [(11, 10.0, 0.5), (12, 1000.0, 0.01), (13, 0.01, 0.001)])
def test_compressed_nodes_bound_their_children(seed, spread, size):
    # The traversal's 32-byte node records, decoded here in float64 from
    # their documented layout: each child's box widened outwards by at least
    # half a step (or starting at the origin, the node's own corner), within
    # 255 steps, at the smallest power-of-two step that spans the node in
    # 253; gamma and the leaf flags name the Karras children.
    pos = soup(2000, seed, spread, size)
    pos[::7, :, 1] = pos[::7, :1, 1]                        # some flat triangles: zero extents
    r = oracle_lib.lbvh(pos)
    n = pos.shape[0]
    lo, hi, q = r["lo"].astype(np.float64), r["hi"].astype(np.float64), r["qnode"]
    assert q.shape == (n - 1, 8)
    origin = q[:, 0:3].view(np.float32).astype(np.float64)
    assert np.array_equal(origin, lo[: n - 1])
    gamma = q[:, 3] & 0x3FFFFFFF
    left = np.where(q[:, 3] & (1 << 30), LEAF | gamma, gamma).astype(np.uint32)
    right = np.where(q[:, 3] & (1 << 31), LEAF | (gamma + 1), gamma + 1).astype(np.uint32)
    assert np.array_equal(left, r["left"]) and np.array_equal(right, r["right"])
    e = np.stack([((q[:, 4] >> (8 * a)) & 0xFF).astype(np.uint8).view(np.int8) for a in range(3)], 1).astype(int)
    step = np.ldexp(1.0, e)
    ext = hi[: n - 1] - lo[: n - 1]
    assert np.all(ext <= 253 * step)
    assert np.all((e == -64) | (ext > 253 * step / 2))      # the smallest such step
    qb = q[:, 5:8].copy().view(np.uint8).reshape(n - 1, 12).astype(np.float64)
    for c, refs in ((0, r["left"]), (1, r["right"])):
        idx = np.where(refs & LEAF, n - 1 + (refs & ~np.uint32(LEAF)), refs).astype(np.int64)
        qlo, qhi = qb[:, 6 * c: 6 * c + 3], qb[:, 6 * c + 3: 6 * c + 6]
        blo, bhi = origin + qlo * step, origin + qhi * step
        assert np.all(blo <= lo[idx]) and np.all(bhi >= hi[idx])
        assert np.all((qlo == 0) | (lo[idx] - blo >= 0.5 * step))
        assert np.all(bhi - hi[idx] >= 0.5 * step)
        assert np.all(qhi <= 255)


def test_compressed_traversal_finds_the_exact_boxes_hits():
    # the traversal over the compressed records returns the hits of brute
    # force over every triangle on rays grazing the soup in every direction
    pos = soup(400, 21, spread=5.0, size=0.3)
    rays = rays_for(pos, 4000, 22)
    hits, counts = oracle_lib.Bvh(pos).closest(rays)
    ref = oracle_lib.brute_closest(pos, rays)
    assert np.array_equal(hits[:, :2].view(np.uint32), ref[:, :2].view(np.uint32))
    assert np.all(counts[:, 0] % 2 == 0)                       # two child boxes per record fetched


def sah(r, n):
    """SAH cost (1.2 per internal node box area, 1 per leaf box area) of an
    oracle hierarchy, over the root's area, in float64."""
    lo, hi = r["lo"].astype(np.float64), r["hi"].astype(np.float64)
    d = hi - lo
    area = 2 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
    return (1.2 * area[: n - 1].sum() + area[n - 1:].sum()) / area[0]


@pytest.mark.parametrize("n,seed", [(3, 31), (50, 32), (2000, 33)])
def test_restructured_hierarchy_is_a_karras_tree(n, seed):
    # The default build: the LBVH restructured by treelets, then relabelled
    # as Karras's layout of its own leaf order.  Every internal node covers
    # consecutive leaf positions, its id is an endpoint of that range by the
    # Karras rule, its children are gamma and gamma + 1, the boxes are the
    # unions of the children's, the leaves are a permutation of the
    # triangles -- and the SAH cost is lower than the plain LBVH's.
    pos = soup(n, seed)
    r, r0 = oracle_lib.lbvh(pos), oracle_lib.lbvh(pos, rounds=0)
    assert np.array_equal(np.sort(r["order"]), np.arange(n))
    tri = pos[r["order"]]
    assert np.array_equal(r["lo"][n - 1:], tri.min(1)) and np.array_equal(r["hi"][n - 1:], tri.max(1))
    ranges = karras_to_ranges(r["left"], r["right"], n)
    assert ranges[0] == (0, n - 1)
    for i in range(n - 1):
        a, b = ranges[i]
        assert i in (a, b) or i == 0
        L, R = int(r["left"][i]), int(r["right"][i])
        g = L & ~LEAF
        assert (R & ~LEAF) == g + 1 and a <= g < b
        for c, want in ((L, (a, g)), (R, (g + 1, b))):
            if c & LEAF:
                assert want[0] == want[1] == (c & ~LEAF)
            else:
                assert ranges[c] == want
                j = c
                kids = [((n - 1 + (x & ~LEAF)) if x & LEAF else x) for x in (int(r["left"][j]), int(r["right"][j]))]
                assert np.array_equal(r["lo"][j], np.minimum(r["lo"][kids[0]], r["lo"][kids[1]]))
    if n >= 50:
        assert sah(r, n) < sah(r0, n)


def all_tree_costs(boxes):
    """The least SAH cost (1.2 x area per internal node, leaf areas) over
    every binary tree of the given leaf boxes, by enumeration of splits."""
    import functools

    m = len(boxes)

    @functools.lru_cache(None)
    def best(mask):
        idx = [i for i in range(m) if mask >> i & 1]
        lo = np.min([boxes[i][0] for i in idx], 0)
        hi = np.max([boxes[i][1] for i in idx], 0)
        d = hi - lo
        a = 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
        if len(idx) == 1:
            return a
        low = mask & -mask
        out = min(best(q) + best(mask ^ q) for q in range(1, mask) if q & mask == q and q & low)
        return 1.2 * a + out

    return best((1 << m) - 1)


@pytest.mark.parametrize("seed", [41, 42, 43, 44])
def test_treelet_finds_the_cheapest_tree_of_seven_leaves(seed):
    # with 7 triangles the root's treelet is the whole hierarchy, so the
    # restructured SAH cost is the least over all 10,395 binary trees
    pos = soup(7, seed, spread=4.0, size=1.5)
    r = oracle_lib.lbvh(pos)
    tri = pos.astype(np.float64)
    boxes = [(t.min(0), t.max(0)) for t in tri]
    d = r["hi"][0].astype(np.float64) - r["lo"][0].astype(np.float64)
    root_area = 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    want = all_tree_costs(boxes) / root_area
    assert abs(sah(r, 7) - want) <= 1e-5 * want
    assert sah(oracle_lib.lbvh(pos, rounds=0), 7) >= want * (1 - 1e-5)


def rays_for(pos, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = pos.reshape(-1, 3).min(0), pos.reshape(-1, 3).max(0)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = o, 1e-4, d, 3.4e38
    r[::5, 7] = 1.5
    return r


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (50, 3), (2000, 4)])
def test_traversal_equals_brute_force(n, seed):
    pos = soup(n, seed, spread=3.0, size=1.0)
    rays = rays_for(pos, 3000, seed + 100)
    bvh = oracle_lib.Bvh(pos)
    hits, counts = bvh.closest(rays)
    brute = oracle_lib.brute_closest(pos, rays)
    assert np.array_equal(hits.view(np.uint32), brute.view(np.uint32))
    occ, _ = bvh.any(rays)
    assert np.array_equal(occ, brute[:, 1].view(np.uint32) != 0xFFFFFFFF)
    # canonical counts: 2 records per visited internal node, bounded by the tree
    assert np.all(counts[:, 0] % 2 == 0)
    assert np.all(counts[:, 0] <= 2 * max(n - 1, 0))
    assert np.all(counts[:, 1] <= n)


def test_axis_aligned_closed_form():
    # unit square in the plane z = 2 made of two triangles; rays along +z
    pos = np.array([[[0, 0, 2], [1, 0, 2], [1, 1, 2]], [[0, 0, 2], [1, 1, 2], [0, 1, 2]]], np.float32)
    xs = np.array([0.75, 0.25, 0.5, 1.5, -0.1], np.float32)
    ys = np.array([0.1, 0.6, 0.5, 0.5, 0.5], np.float32)
    rays = np.zeros((5, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = xs, ys, 0.0
    rays[:, 3], rays[:, 6], rays[:, 7] = 1e-4, 1.0, 3.4e38
    hits, _ = oracle_lib.Bvh(pos).closest(rays)
    ids = hits[:, 1].view(np.uint32)
    assert np.all(hits[:3, 0] == 2.0)
    assert ids[0] == 0 and ids[1] == 1
    assert ids[2] == 0  # on the shared diagonal both triangles hit at t = 2: lower id wins
    assert ids[3] == 0xFFFFFFFF and ids[4] == 0xFFFFFFFF
    # barycentrics reproduce the hit point: (1-u-v) v0 + u v1 + v v2
    for i in range(3):
        t = pos[ids[i]]
        u, v = hits[i, 2], hits[i, 3]
        p = (1 - u - v) * t[0] + u * t[1] + v * t[2]
        assert np.allclose(p[:2], [xs[i], ys[i]], atol=1e-6)


def test_watertight_shared_edges():
    # rays through the shared edges/vertex of a fan must hit exactly one or more
    # triangles, never fall through the cracks
    ang = np.linspace(0, 2 * np.pi, 9)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(8)], 1)
    pos = np.array([[[0, 0, 0], ring[i], ring[(i + 1) % 8]] for i in range(8)], np.float32)
    rng = np.random.default_rng(0)
    t = rng.uniform(0, 1, 4000).astype(np.float32)
    k = rng.integers(0, 8, 4000)
    pts = (ring[k] * t[:, None]).astype(np.float32)  # points on the spokes
    rays = np.zeros((4000, 8), np.float32)
    rays[:, :2] = pts[:, :2]
    rays[:, 2] = -1.0
    rays[:, 3], rays[:, 6], rays[:, 7] = 1e-4, 1.0, 3.4e38
    hits, _ = oracle_lib.Bvh(pos).closest(rays)
    assert np.all(hits[:, 1].view(np.uint32) != 0xFFFFFFFF)
