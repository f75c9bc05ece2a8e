"""GPU parity at the benchmark workloads' own shapes and at the reference's own
test and bench sizes, every output compared with the oracle bit for bit.

  * LDE + Merkle commit, 256 M31 columns of log 20 -> log 21: every LDE column
    and every tree layer;
  * batch_inverse of 2^24 uniform non-zero QM31; the reference's field.rs:41-78
    tests (2^25 M31 and 2^23 QM31 of consecutive words); its benches' 2^28
    (benches/batch_inverse.rs:11-39) for M31 and QM31;
  * the FRI chain: circle fold at log 22, then fold_line down to log 4 with a
    seeded alpha per layer, every layer;
  * DEEP quotients over 200 columns of log 21 with 3 random QM31 points and
    1..3 samples per column;
  * the mixed-size commit: 1200 columns of log 10..22 (weights 1/size), one plan,
    graph replay and stream launches;
  * the reference's benches: interpolate_columns / evaluate_polynomials of 2^16
    columns of 2^10 values at blowup 4 through the reference ABI
    (benches/interpolate_columns.rs:10-49, evaluate_columns.rs:11-63), and
    bit_reverse of 2^28 M31 and QM31 (benches/bit_reverse.rs:9-34).

Inputs come from the counter-based uniform stream that both the device
(stwo_hip_fill_uniform_m31) and the oracle generate.
"""
import gc

import numpy as np
import pytest

import oracle as O
import stwo_hip as S

pytestmark = pytest.mark.gpu
P = S.P


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.check(S.lib().stwo_hip_init(0))
    yield
    gc.collect()
    S.check(S.lib().stwo_hip_release_cached_memory())


def _device_uniform(n, seed, offset):
    v = S.BaseFieldVec.uninitialized(n)
    S.fill_uniform(v, seed, offset)
    return v


class _Slab:
    """One device allocation holding n_cols columns of 2^log_n values back to back,
    with non-owning column views (many small columns without one allocation each)."""

    def __init__(self, n_cols, log_n):
        self.buf = S.BaseFieldVec.uninitialized(n_cols << log_n)
        self.cols = [S.BaseFieldVec.view(self.buf.ptr + (c << (log_n + 2)), 1 << log_n) for c in range(n_cols)]

    def to_numpy(self, log_n):
        return self.buf.to_numpy().reshape(-1, 1 << log_n)


# ------------------------------------------------------------------ LDE + commit, 256 x 2^20 -> 2^21
def test_lde_commit_256_columns_log20():
    C, L, blow, seed = 256, 20, 1, 0x5EED0002
    n, m = 1 << L, 1 << (L + blow)
    evals = [_device_uniform(n, seed, c * n) for c in range(C)]
    coeffs = [S.BaseFieldVec.uninitialized(n) for _ in range(C)]
    ldes = [S.BaseFieldVec.uninitialized(m) for _ in range(C)]
    root = S.canonic_half_coset(L + blow)
    tw, itw = S.cached_twiddles(root)
    plan = S.CommitPlan(evals, coeffs, ldes, [L] * C, blow, tw, itw, 1 << root.log_size, autotune=False)
    plan.run(True)
    S.synchronize()

    host = [O.fill_uniform_m31(n, seed, c * n) for c in range(C)]
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    ocoef, olde = O.lde_columns(host, blow, otw, oitw)
    del host
    for c in range(C):
        assert (coeffs[c].to_numpy() == ocoef[c]).all(), f"coefficients of column {c}"
        assert (ldes[c].to_numpy() == olde[c]).all(), f"LDE of column {c}"
    del ocoef
    want = O.merkle_commit(olde)
    for k in range(plan.max_log + 1):
        assert (plan.layers[k].to_numpy().reshape(-1) == want[k]).all(), f"layer {k}"
    plan.close()


# ------------------------------------------------------------------ batch inverse
def test_batch_inverse_qm31_2e24_uniform():
    n = 1 << 24
    src = S.SecureFieldVec.uninitialized(n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(src.ptr, 4 * n, 0x5EED0002, 1 << 40))
    S.synchronize()
    host = src.to_numpy()
    nz = host.any(axis=1)
    assert nz.all(), "the seeded stream has no zero QM31 among these 2^24"
    dst = S.SecureFieldVec.uninitialized(n)
    S.batch_inverse(src, dst)
    assert (dst.to_numpy() == O.batch_inverse_qm31(host.reshape(-1)).reshape(-1, 4)).all()


def test_batch_inverse_reference_sizes():
    # field.rs:41-57: 2^25 M31 values 1..2^25
    n = 1 << 25
    vals = np.arange(1, n + 1, dtype=np.uint32)
    a, d = S.BaseFieldVec.from_numpy(vals), S.BaseFieldVec.uninitialized(n)
    S.batch_inverse(a, d)
    assert (d.to_numpy() == O.batch_inverse_m31(vals)).all()
    del a, d
    # field.rs:59-77: words 1..2^25 in chunks of four, 2^23 QM31
    q = vals.reshape(-1, 4)
    a, d = S.SecureFieldVec.from_numpy(q), S.SecureFieldVec.uninitialized(n // 4)
    S.batch_inverse(a, d)
    assert (d.to_numpy().reshape(-1) == O.batch_inverse_qm31(vals)).all()


@pytest.mark.parametrize("secure", [False, True])
def test_batch_inverse_reference_bench_2e28(secure):
    # benches/batch_inverse.rs:11-39: 2^28 elements
    n = 1 << 28
    words = 4 * n if secure else n
    src = (S.SecureFieldVec if secure else S.BaseFieldVec).uninitialized(n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(src.ptr, words, 0x5EED0006, 0))
    dst = (S.SecureFieldVec if secure else S.BaseFieldVec).uninitialized(n)
    S.batch_inverse(src, dst)
    host = O.fill_uniform_m31(words, 0x5EED0006, 0)
    # a zero element (if the stream has one) maps to zero; the oracle takes non-zero inputs
    zero = ~host.reshape(-1, 4).any(axis=1) if secure else host == 0
    fixed = host.reshape(-1, 4).copy() if secure else host.copy()
    fixed[zero] = (1, 0, 0, 0) if secure else 1
    want = O.batch_inverse_qm31(fixed.reshape(-1)).reshape(-1, 4) if secure else O.batch_inverse_m31(fixed)
    want[zero] = 0
    got = dst.to_numpy()
    assert (got == want).all()


# ------------------------------------------------------------------ bit reverse, 2^28
@pytest.mark.parametrize("secure", [False, True])
def test_bit_reverse_reference_bench_2e28(secure):
    # benches/bit_reverse.rs:9-34
    n = 1 << 28
    words = 4 * n if secure else n
    col = (S.SecureFieldVec if secure else S.BaseFieldVec).uninitialized(n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(col.ptr, words, 0x5EED0007, 0))
    S.bit_reverse_column(col)
    host = O.fill_uniform_m31(words, 0x5EED0007, 0)
    want = O.bit_reverse_qm31(host) if secure else O.bit_reverse(host)
    assert (col.to_numpy().reshape(-1) == want).all()


# ------------------------------------------------------------------ FRI chain log 22 -> 4
def test_fri_chain_log22_to_log4():
    L = 22
    rng = np.random.default_rng(0x5EED0002)
    src = S.SecureColumn([_device_uniform(1 << L, 0x5EED0002, (2 << 40) + k * (1 << L)) for k in range(4)])
    layers = [S.SecureColumn.zeros(1 << l) for l in range(L - 1, 3, -1)]
    root = S.canonic_half_coset(L)
    tw, itw = S.cached_twiddles(root)
    tree_len = 1 << root.log_size
    alphas = [S.random_qm31(rng) for _ in range(len(layers))]
    S.fold_circle_into_line_async(layers[0], src, alphas[0], itw, tree_len - len(layers[0]))
    for i in range(1, len(layers)):
        S.fold_line_async(layers[i - 1], layers[i], alphas[i], itw, tree_len - len(layers[i - 1]))
    S.synchronize()
    host = np.stack([O.fill_uniform_m31(1 << L, 0x5EED0002, (2 << 40) + k * (1 << L)) for k in range(4)])
    cur = O.fold_circle_into_line(np.zeros((4, 1 << (L - 1)), np.uint32), host, root.initial_index, root.step_size,
                                  alphas[0])
    assert (layers[0].to_numpy() == cur).all(), "circle fold"
    d = S.Coset(root.initial_index, root.step_size, root.log_size)
    for i in range(1, len(layers)):
        cur = O.fold_line(cur, d.initial_index, d.step_size, alphas[i])
        d = d.double()
        assert (layers[i].to_numpy() == cur).all(), f"line layer {i} (log {L - 1 - i})"
    assert len(layers[-1]) == 16


def test_fri_fold_chain_log22_to_log4():
    # the same chain through the fused entry point (two launches)
    L = 22
    rng = np.random.default_rng(0x5EED0002)
    src = S.SecureColumn([_device_uniform(1 << L, 0x5EED0002, (2 << 40) + k * (1 << L)) for k in range(4)])
    layers = [S.SecureColumn.zeros(1 << l) for l in range(L - 1, 3, -1)]
    root = S.canonic_half_coset(L)
    tw, itw = S.cached_twiddles(root)
    alphas = [S.random_qm31(rng) for _ in range(len(layers))]
    S.fri_fold_chain_async(src, layers, alphas, itw, 1 << root.log_size)
    S.synchronize()
    host = np.stack([O.fill_uniform_m31(1 << L, 0x5EED0002, (2 << 40) + k * (1 << L)) for k in range(4)])
    cur = O.fold_circle_into_line(np.zeros((4, 1 << (L - 1)), np.uint32), host, root.initial_index, root.step_size,
                                  alphas[0])
    assert (layers[0].to_numpy() == cur).all(), "circle fold"
    d = S.Coset(root.initial_index, root.step_size, root.log_size)
    for i in range(1, len(layers)):
        cur = O.fold_line(cur, d.initial_index, d.step_size, alphas[i])
        d = d.double()
        assert (layers[i].to_numpy() == cur).all(), f"line layer {i} (log {L - 1 - i})"


# ------------------------------------------------------------------ quotients 200 x 2^21, 3 points
def test_quotients_200_columns_log21_3_points():
    L, C = 21, 200
    n = 1 << L
    rng = np.random.default_rng(0x5EED0003)
    cols = [_device_uniform(n, 0x5EED0003, c * n) for c in range(C)]
    points = [S.random_circle_point_qm31(rng) for _ in range(3)]
    per_point = [[] for _ in range(3)]
    for c in range(C):
        for pi in sorted(rng.choice(3, size=int(rng.integers(1, 4)), replace=False)):
            per_point[pi].append((c, S.random_qm31(rng)))
    batches = [(points[i], per_point[i]) for i in range(3) if per_point[i]]
    assert len(batches) == 3 and 200 < sum(len(cv) for _, cv in batches) < 600
    rc = S.random_qm31(rng)
    out = S.SecureColumn.uninitialized(n)
    dom = S.canonic_half_coset(L)
    S.accumulate_quotients_async(dom, cols, rc, S.QuotientSamples(batches), out)
    S.synchronize()
    host = [O.fill_uniform_m31(n, 0x5EED0003, c * n) for c in range(C)]
    want = O.accumulate_quotients(dom.initial_index, dom.step_size, L, host, rc, batches)
    assert (out.to_numpy() == want).all()
    # the same through the reference ABI (synchronous, device pointer table)
    got = S.accumulate_quotients(dom, cols, rc, batches)
    assert (got.to_numpy() == want).all()


# ------------------------------------------------------------------ mixed-size commit, 1200 columns
def _mixed_log_sizes(n_cols=1200, seed=0x5EED0004):
    ks = np.arange(10, 23)
    p = 2.0 ** -ks
    return [int(k) for k in np.random.default_rng(seed).choice(ks, n_cols, p=p / p.sum())]


@pytest.mark.parametrize("use_graph", [True, False])
def test_mixed_commit_1200_columns(use_graph):
    logs = _mixed_log_sizes()
    # with weights 1/size the expected count at log 22 is 0.15 of a column: this seed's top is 21
    assert min(logs) == 10 and max(logs) == 21 and len(set(logs)) == 11
    C, blow = len(logs), 1
    evals = [_device_uniform(1 << l, 0x5EED0004, c << 22) for c, l in enumerate(logs)]
    coeffs = [S.BaseFieldVec.uninitialized(1 << l) for l in logs]
    ldes = [S.BaseFieldVec.uninitialized(2 << l) for l in logs]
    root = S.canonic_half_coset(max(logs) + blow)
    tw, itw = S.cached_twiddles(root)
    plan = S.CommitPlan(evals, coeffs, ldes, logs, blow, tw, itw, 1 << root.log_size, autotune=True)
    for rep in range(2):
        for l in ldes[::97]:
            S.fill_uniform(l, 77 + rep)  # clobber some outputs: every run must rewrite them
        plan.run(use_graph)
        S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    host = [O.fill_uniform_m31(1 << l, 0x5EED0004, c << 22) for c, l in enumerate(logs)]
    olde = [None] * C
    for k in sorted(set(logs)):
        idx = [c for c in range(C) if logs[c] == k]
        oc, ol = O.lde_columns([host[c] for c in idx], blow, otw, oitw)
        for j, c in enumerate(idx):
            olde[c] = ol[j]
            assert (coeffs[c].to_numpy() == oc[j]).all(), f"coefficients of column {c} (log {k})"
            assert (ldes[c].to_numpy() == ol[j]).all(), f"LDE of column {c} (log {k})"
    want = O.merkle_commit(olde)
    for k in range(plan.max_log + 1):
        assert (plan.layers[k].to_numpy().reshape(-1) == want[k]).all(), f"layer {k}"
    plan.close()


# ------------------------------------------------------------------ reference benches: 2^16 columns x 2^10
def test_interpolate_columns_reference_bench_shape():
    # benches/interpolate_columns.rs:10-49 through interpolate_columns (bindings.rs:128-135),
    # with distinct uniform columns in place of the bench's 2^16 copies of 0..1023
    log_n, n_cols = 10, 1 << 16
    slab = _Slab(n_cols, log_n)
    S.check(S.lib().stwo_hip_fill_uniform_m31(slab.buf.ptr, n_cols << log_n, 0x5EED0008, 0))
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    S.interpolate_columns(slab.cols, tree)
    host = O.fill_uniform_m31(n_cols << log_n, 0x5EED0008, 0).reshape(n_cols, -1)
    otw, oitw = O.precompute_twiddles(coset.initial_index, coset.step_size, coset.log_size)
    got = slab.to_numpy(log_n)
    for c0 in range(0, n_cols, 8192):
        oc, _ = O.lde_columns(list(host[c0:c0 + 8192]), 0, otw, oitw)
        assert (got[c0:c0 + 8192] == np.stack(oc)).all(), f"columns {c0}.."


def test_evaluate_polynomials_reference_bench_shape():
    # benches/evaluate_columns.rs:11-63: 2^16 polynomials of 2^10 coefficients evaluated at
    # blowup 4 (extended with zeros, then evaluate_columns, bindings.rs:145-152)
    log_n, blow, n_cols = 10, 2, 1 << 16
    m = log_n + blow
    coef = O.fill_uniform_m31(n_cols << log_n, 0x5EED0009, 0).reshape(n_cols, -1)
    ext = np.zeros((n_cols, 1 << m), np.uint32)
    ext[:, : 1 << log_n] = coef
    slab = _Slab(n_cols, m)
    S.check(S.lib().stwo_hip_memcpy(slab.buf.ptr, ext.ctypes.data, ext.nbytes))
    big = S.canonic_half_coset(m)
    tree = S.precompute_twiddles(big)
    S.evaluate_columns(slab.cols, tree)
    got = slab.to_numpy(m)
    otw, _ = O.precompute_twiddles(big.initial_index, big.step_size, big.log_size)
    for c in range(0, n_cols, 1021):
        assert (got[c] == O.evaluate(ext[c], big.initial_index, otw)).all(), f"column {c}"
    # every column: evaluation then interpolation at blowup 4 returns the zero-padded coefficients
    back = _Slab(n_cols, m)
    S.check(S.lib().stwo_hip_memcpy(back.buf.ptr, got.ctypes.data, got.nbytes))
    S.interpolate_columns(back.cols, tree)
    assert (back.to_numpy(m) == ext).all()


# ------------------------------------------------------------------ the largest columns: log 29 and 30
@pytest.mark.parametrize("log_n", [29, 30])
def test_fft_round_trip_largest_columns(log_n):
    """Columns of 2^29 and 2^30 values (byte offsets up to 2^32 - 4, every pass shape of the
    largest plan): evaluate a sparse coefficient vector into a buffer of unrelated values, then
    interpolate out of place into another; the result must be exactly the sparse vector, so a
    tile that either pass skips or misplaces shows up as nonzero words or wrong values."""
    n = 1 << log_n
    rng = np.random.default_rng(0x5EED0010 + log_n)
    idx = np.unique(np.concatenate([rng.integers(0, n, 61), [0, 1, n // 2, n - 1]])).astype(np.int64)
    vals = rng.integers(1, P, idx.size).astype(np.uint32)
    coef = S.BaseFieldVec.zeros(n)
    for i, v in zip(idx, vals):
        w = np.array([v], np.uint32)
        S.check(S.lib().stwo_hip_memcpy(coef.ptr + 4 * int(i), w.ctypes.data, 4))
    evals = S.BaseFieldVec.uninitialized(n)
    S.fill_uniform(evals, 0x5EED0011, 0)
    back = S.BaseFieldVec.uninitialized(n)
    S.fill_uniform(back, 0x5EED0012, 0)
    root = S.canonic_half_coset(log_n)
    tw, itw = S.cached_twiddles(root)
    S.evaluate_columns_async([coef], [log_n], [evals], [log_n], tw, 1 << root.log_size)
    S.interpolate_columns_oop_async([evals], [back], [log_n], itw, 1 << root.log_size)
    S.synchronize()
    del evals, coef
    got = back.to_numpy()
    del back
    nz = np.flatnonzero(got)
    assert np.array_equal(nz, idx), f"nonzero positions differ: {nz[:8]} vs {idx[:8]}"
    assert np.array_equal(got[idx], vals)
    del got
    gc.collect()
    S.check(S.lib().stwo_hip_release_twiddles())  # the log-30 trees hold 4 GiB


# ------------------------------------------------------------------ Merkle layers of 2^30 and 2^29 nodes
def _hash_rows(h, lo, count):
    """count hashes of a device hash vector from row lo (a partial device-to-host copy)."""
    out = np.empty(8 * count, np.uint32)
    S.check(S.lib().stwo_hip_memcpy(out.ctypes.data, h.ptr + 32 * lo, 32 * count))
    return out.reshape(-1, 8)


def test_merkle_layers_2e30_and_2e29_rows():
    """The largest Merkle layers: commit_on_first_layer over two columns of 2^30 rows and
    commit_on_layer_with_previous over one column of 2^29 rows; the first and last rows of
    each layer (where a column's byte offsets reach 2^32 - 4) against BLAKE2s from the oracle."""
    L = 30
    cols = [S.BaseFieldVec.uninitialized(1 << L) for _ in range(2)]
    for c, col in enumerate(cols):
        S.fill_uniform(col, 0x5EED0013, c << 31)
    top = S.commit_on_layer(L, None, cols)
    col29 = S.BaseFieldVec.uninitialized(1 << (L - 1))
    S.fill_uniform(col29, 0x5EED0014, 0)
    nxt = S.commit_on_layer(L - 1, top, [col29])
    S.synchronize()

    def leaf(vals):
        return O.blake2s_compress(np.zeros(8, np.uint32), np.array(list(vals) + [0] * (16 - len(vals)), np.uint32))

    for lo, cnt in [(0, 4), ((1 << L) - 64, 64)]:
        vals = [O.fill_uniform_m31(cnt, 0x5EED0013, (c << 31) + lo) for c in range(2)]
        got = _hash_rows(top, lo, cnt)
        for r in range(cnt):
            assert (got[r] == leaf([vals[0][r], vals[1][r]])).all(), f"leaf row {lo + r}"
    for lo, cnt in [(0, 4), ((1 << (L - 1)) - 32, 32)]:
        children = _hash_rows(top, 2 * lo, 2 * cnt)
        vals = O.fill_uniform_m31(cnt, 0x5EED0014, lo)
        got = _hash_rows(nxt, lo, cnt)
        for r in range(cnt):
            st = O.blake2s_compress(np.zeros(8, np.uint32), np.concatenate([children[2 * r], children[2 * r + 1]]))
            st = O.blake2s_compress(st, np.array([vals[r]] + [0] * 15, np.uint32))
            assert (got[r] == st).all(), f"node row {lo + r}"
    del top, nxt, cols, col29
    gc.collect()

