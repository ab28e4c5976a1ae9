"""CPU tests of the library's synthetic workloads (block FEM, banded batch,
AMG Poisson + prolongator), its host transpose and the block-diagonal batch
layout.  Each generator is checked against an independent construction:
scipy's k-d tree for the nearest neighbours, scipy sparse algebra for the
prolongator and the transpose, and the defining properties otherwise."""
import numpy as np
import pytest

import rtt_amd

GOLD = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def uniform_at(seed, idx):
    """splitmix64 at position idx of stream seed (Steele, Lea, Flood 2014):
    the published output function, as the library draws its points."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * GOLD
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def block_pattern(a):
    """Set of (block row, block col) of a matrix made of dense 4x4 blocks."""
    s = a.to_scipy().tocoo()
    return set(zip((s.row // 4).tolist(), (s.col // 4).tolist()))


@pytest.mark.parametrize("nodes,k", [(3000, 6), (700, 3), (9, 6)])
def test_fem_blocks_are_the_symmetrised_nearest_neighbour_graph(nodes, k):
    from scipy.spatial import cKDTree
    seed = 77
    a = rtt_amd.gen_fem_blocks(nodes, k, spatial_order=False, seed=seed)
    pts = uniform_at(seed, np.arange(3 * nodes)).reshape(nodes, 3)
    kk = min(k, nodes - 1)
    _, nb = cKDTree(pts).query(pts, k=kk + 1)
    want = {(i, i) for i in range(nodes)}
    for i in range(nodes):
        for j in nb[i][1:]:
            want.add((i, int(j)))
            want.add((int(j), i))
    assert block_pattern(a) == want
    # every block is dense 4x4; magnitudes 1e-3..1e3 with both signs
    assert a.nnz == 16 * len(want)
    mag = np.abs(a.val)
    assert mag.min() >= 1e-3 and mag.max() <= 1e3
    assert (a.val < 0).any() and (a.val > 0).any()
    rp = a.rowptr
    for r in range(0, 4 * nodes, 97):
        cols = a.col[rp[r]:rp[r + 1]]
        assert np.all(np.diff(cols) > 0)


def test_fem_blocks_spatial_order_is_a_renumbering():
    """Morton numbering relabels the same mesh: the block degree multiset is
    unchanged, and consecutive nodes are closer together than in draw order."""
    a = rtt_amd.gen_fem_blocks(4000, 6, spatial_order=False, seed=5)
    b = rtt_amd.gen_fem_blocks(4000, 6, spatial_order=True, seed=5)
    da = np.diff(a.rowptr[::4]) // 16
    db = np.diff(b.rowptr[::4]) // 16
    assert sorted(da.tolist()) == sorted(db.tolist())
    pa, pb = block_pattern(a), block_pattern(b)
    span = lambda p: np.mean([abs(i - j) for i, j in p])
    assert span(pb) < 0.2 * span(pa)
    again = rtt_amd.gen_fem_blocks(4000, 6, spatial_order=True, seed=5)
    assert np.array_equal(again.col, b.col) and np.array_equal(again.val, b.val)


def test_banded_rows_stay_in_band_with_diagonal():
    n, bw, pr = 2000, 12, 9
    a = rtt_amd.gen_banded(n, bw, pr, seed=3)
    rp = a.rowptr
    for i in range(n):
        cols = a.col[rp[i]:rp[i + 1]]
        width = min(n - 1, i + bw) - max(0, i - bw) + 1
        assert i in cols
        assert np.all(np.abs(cols - i) <= bw)
        assert np.all(np.diff(cols) > 0)
        assert min(pr, width) <= len(cols) <= min(pr, width) + 1
    assert np.all(np.abs(a.val) <= 1.0)


def test_banded_batch_shapes_follow_the_spec():
    rows, bw, pr = rtt_amd.gen_banded_batch_shapes(512)
    assert rows.min() >= 1000 and rows.max() <= 200000
    assert 8 <= bw.min() and bw.max() <= 64 and 5 <= pr.min() and pr.max() <= 30
    med = np.median(rows)
    assert 11000 < med < 23000, med
    r2, _, _ = rtt_amd.gen_banded_batch_shapes(512)
    assert np.array_equal(rows, r2)


def test_poisson7_jumps_is_the_jump_coefficient_operator():
    g = 6
    a = rtt_amd.gen_poisson7_jumps(g, 3.0, seed=11)
    S = a.to_scipy()
    assert (abs(S - S.T)).max() == 0.0          # face weights are symmetric
    d = S.diagonal()
    offsum = -(S.sum(axis=1).A1 - d)            # sum of face weights of each row
    # diagonal = face weights + the cell's own coefficient per boundary face;
    # interior cells have no boundary face
    idx = np.arange(g ** 3)
    x, y, z = idx % g, (idx // g) % g, idx // (g * g)
    interior = (x > 0) & (x < g - 1) & (y > 0) & (y < g - 1) & (z > 0) & (z < g - 1)
    assert np.allclose(d[interior], offsum[interior], rtol=1e-15)
    assert np.all(d[~interior] > offsum[~interior])
    w = -S.data[S.data < 0]
    assert w.max() / w.min() > 1e3             # jumps of several decades
    assert S.nnz == a.nnz == 7 * g ** 3 - 6 * g ** 2


def test_prolongator_matches_scipy_construction():
    import scipy.sparse as sp
    g = 7
    a = rtt_amd.gen_poisson7_jumps(g, 1.0, seed=4)
    p = rtt_amd.gen_aggregation_prolongator(a, g, 2.0 / 3.0)
    A = a.to_scipy()
    h = (g + 1) // 2
    idx = np.arange(g ** 3)
    x, y, z = idx % g, (idx // g) % g, idx // (g * g)
    agg = ((z // 2) * h + (y // 2)) * h + (x // 2)
    T = sp.csr_matrix((np.ones(g ** 3), (idx, agg)), shape=(g ** 3, h ** 3))
    P = (T - (2.0 / 3.0) * sp.diags(1.0 / A.diagonal()) @ A @ T).tocsr()
    got = p.to_scipy()
    assert got.shape == P.shape
    assert abs(got - P).max() <= 1e-15 * abs(P).max()


def test_transpose_and_block_diag():
    import scipy.sparse as sp
    mats = [rtt_amd.gen_random(m, n, k, seed=m) for m, n, k in
            ((13, 9, 3), (8, 8, 2), (1, 30, 5), (20, 5, 1), (0, 4, 0))]
    t = rtt_amd.csr_transpose(mats[0])
    assert (abs(t.to_scipy() - mats[0].to_scipy().T)).max() == 0
    big, ro, co = rtt_amd.block_diag(mats, 8)
    assert np.all(ro % 8 == 0) and np.all(co % 8 == 0)
    B = big.to_scipy()
    for p, m in enumerate(mats):
        blk = B[ro[p]:ro[p] + m.m, co[p]:co[p] + m.n]
        assert (abs(blk - m.to_scipy())).max() == 0 if m.nnz else blk.nnz == 0
    assert B.nnz == sum(m.nnz for m in mats)
    assert big.m == ro[-1] and big.n == co[-1]
