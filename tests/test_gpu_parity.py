"""GPU parity: every backend op through the C ABI against the oracle, bit-exact.

Mirrors the reference's own tests (stwo_gpu_backend/src/*.rs `mod tests`), with
CpuBackend replaced by the oracle; sizes are chosen so the oracle finishes in
seconds.  Full-size configurations are covered by property tests in
test_gpu_properties.py.
"""
import json
import os
import random

import numpy as np
import pytest

import oracle as O
import stwo_hip as S

pytestmark = pytest.mark.gpu
P = S.P
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    st = S.lib().stwo_hip_init(0)
    S.check(st)
    yield


def rand_m31(n, seed):
    return np.random.default_rng(seed).integers(0, P, n, dtype=np.uint32)


def rand_qm31(rng):
    return tuple(rng.randrange(P) for _ in range(4))


# ------------------------------------------------------------------ column.rs:156-186
# sizes either side of each kernel's range: 64 x 64 tiles from 2^12, 128 x 128 tiles from 2^14
# (2^14 and 2^15: every tile pairs with itself; from 2^16 distinct tiles pair up)
@pytest.mark.parametrize("log_n", [1, 5, 10, 11, 12, 13, 14, 15, 16, 17, 20, 22])
def test_bit_reverse_base_field(log_n):
    data = np.arange(1 << log_n, dtype=np.uint32)
    col = S.BaseFieldVec.from_numpy(data)
    S.bit_reverse_column(col)
    assert (col.to_numpy() == O.bit_reverse(data)).all()


# 32 x 32 tiles from 2^10, 64 x 64 tiles from 2^12 (2^12, 2^13: every tile pairs with itself)
@pytest.mark.parametrize("log_n", [3, 9, 10, 11, 12, 13, 14, 16])
def test_bit_reverse_secure_field(log_n):
    data = np.arange(1, 1 + 4 * (1 << log_n), dtype=np.uint32).reshape(-1, 4)
    v = S.SecureFieldVec.from_numpy(data)
    S.bit_reverse_column(v)
    assert (v.to_numpy().reshape(-1) == O.bit_reverse_qm31(data.reshape(-1))).all()


# ------------------------------------------------------------------ field.rs:41-78
# sizes that are multiples of 4 take the 16-byte kernel (12 and 2^16 + 4 leave partial
# blocks and threads), the others the dword kernel
@pytest.mark.parametrize("n", [1, 12, 100, 1 << 12, (1 << 16) + 4, (1 << 20) + 3])
def test_batch_inverse_basefield(n):
    src = np.arange(1, n + 1, dtype=np.uint32)
    a, d = S.BaseFieldVec.from_numpy(src), S.BaseFieldVec.uninitialized(n)
    S.batch_inverse(a, d)
    assert (d.to_numpy() == O.batch_inverse_m31(src)).all()


def test_batch_inverse_secure_field():
    n = 1 << 16
    src = np.arange(1, 4 * n + 1, dtype=np.uint32).reshape(-1, 4)
    a, d = S.SecureFieldVec.from_numpy(src), S.SecureFieldVec.uninitialized(n)
    S.batch_inverse(a, d)
    assert (d.to_numpy().reshape(-1) == O.batch_inverse_qm31(src.reshape(-1))).all()


# zeros lie outside stwo's contract for batch inversion; here inverse(0) is 0 and leaves the
# other elements' inverses intact, in both the dword (n = 6) and the 16-byte (n = 4096) kernel
@pytest.mark.parametrize("n", [6, 4096])
def test_batch_inverse_zero_is_isolated(n):
    src = np.array([3, 0, 5, 7, 0, 11] * (n // 6) + [13, 0, 17, 19][: n % 6], np.uint32)
    a, d = S.BaseFieldVec.from_numpy(src), S.BaseFieldVec.uninitialized(n)
    S.batch_inverse(a, d)
    out = d.to_numpy()
    for i in range(n):
        if src[i] == 0:
            assert out[i] == 0, i
        else:
            assert int(out[i]) * int(src[i]) % P == 1, i


# ------------------------------------------------------------------ poly.rs:216-249
def test_new_canonical_ordered():
    n = 1 << 16
    data = np.arange(n, dtype=np.uint32)
    out = S.new_canonical_ordered(S.BaseFieldVec.from_numpy(data))
    assert (out.to_numpy() == O.new_canonical_ordered(data)).all()


@pytest.mark.parametrize("log_n", [5, 6, 13, 20])
def test_precompute_twiddles(log_n):
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    tw, itw = O.precompute_twiddles(coset.initial_index, coset.step_size, coset.log_size)
    assert (tree.twiddles.to_numpy() == tw).all()
    assert (tree.itwiddles.to_numpy() == itw).all()


def test_twiddle_fixture_and_cache():
    f = FIX["interpolate_from_fib"]
    rc = f["root_coset"]
    tree = S.precompute_twiddles(S.Coset(rc["initial_index"], rc["step_size"], rc["log_size"]))
    assert tree.twiddles.to_numpy().tolist() == f["twiddles"]
    assert tree.itwiddles.to_numpy().tolist() == f["itwiddles"]
    tw, itw = S.cached_twiddles(S.Coset(rc["initial_index"], rc["step_size"], rc["log_size"]))
    host = np.zeros(64, np.uint32)
    S.lib().copy_uint32_t_vec_from_device_to_host(tw, host.ctypes.data, 64)
    assert host.tolist() == f["twiddles"]


# ------------------------------------------------------------------ poly.rs:266-396
@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 8, 11, 12, 13, 14, 16, 17])
def test_interpolate(log_n):
    vals = np.arange(1, (1 << log_n) + 1, dtype=np.uint32)
    ev = O.new_canonical_ordered(vals)
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(S.canonic_half_coset(log_n + 1))  # a larger tree serves the domain
    col = S.BaseFieldVec.from_numpy(ev)
    S.interpolate(col, tree)
    _, itw = O.precompute_twiddles(*S.canonic_half_coset(log_n + 1).__dict__.values())
    assert (col.to_numpy() == O.interpolate(ev, coset.initial_index, itw)).all()


def test_interpolate_fixtures():
    f = FIX["interpolate_from_fib"]
    rc = f["root_coset"]
    tree = S.TwiddleTree(S.Coset(rc["initial_index"], rc["step_size"], rc["log_size"]),
                         S.BaseFieldVec.from_numpy(np.array(f["twiddles"], np.uint32)),
                         S.BaseFieldVec.from_numpy(np.array(f["itwiddles"], np.uint32)))
    vals = np.array(f["values"], np.uint32)
    col = S.BaseFieldVec.from_numpy(vals)
    S.interpolate(col, tree)
    want = O.interpolate(vals, f["domain_half_coset"]["initial_index"], np.array(f["itwiddles"], np.uint32))
    assert (col.to_numpy() == want).all()
    # poly.rs:291-327
    v2 = O.new_canonical_ordered(np.array(FIX["interpolate_2_values"], np.uint32))
    col2 = S.BaseFieldVec.from_numpy(v2)
    S.interpolate(col2, tree)
    assert (col2.to_numpy() == O.interpolate(v2, 1 << 25, np.array(f["itwiddles"], np.uint32))).all()


@pytest.mark.parametrize("log_n", [1, 2, 3, 7, 12, 15, 18])
def test_evaluate(log_n):
    coeffs = rand_m31(1 << log_n, log_n)
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    col = S.BaseFieldVec.from_numpy(coeffs)
    S.evaluate(col, tree)
    tw, _ = O.precompute_twiddles(coset.initial_index, coset.step_size, coset.log_size)
    assert (col.to_numpy() == O.evaluate(coeffs, coset.initial_index, tw)).all()


@pytest.mark.parametrize("log_n", [0, 1, 2, 5, 11, 12, 16, 20])
def test_eval_at_point(log_n):
    rng = random.Random(log_n)
    coeffs = rand_m31(1 << log_n, 100 + log_n)
    px, py = rand_qm31(rng), rand_qm31(rng)  # any QM31 pair: the fold does not need the circle equation
    got = S.eval_at_point(S.BaseFieldVec.from_numpy(coeffs), px, py)
    assert got == O.eval_at_point(coeffs, px, py)


# ------------------------------------------------------------------ poly.rs:616-706
def test_interpolate_columns():
    log_n, n_cols = 9, 256
    vals = np.arange(1, (1 << log_n) + 1, dtype=np.uint32)
    ev = O.new_canonical_ordered(vals)
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    cols = [S.BaseFieldVec.from_numpy(ev) for _ in range(n_cols)]
    S.interpolate_columns(cols, tree)
    want = O.interpolate(ev, coset.initial_index, tree.itwiddles.to_numpy())
    for c in cols[:: 37]:
        assert (c.to_numpy() == want).all()


def test_evaluate_columns_blowup_4():
    log_n, n_cols, blow = 9, 256, 2
    vals = np.arange(1, (1 << log_n) + 1, dtype=np.uint32)
    big = S.canonic_half_coset(log_n + blow)
    tw, itw = O.precompute_twiddles(big.initial_index, big.step_size, big.log_size)
    coeffs = O.interpolate(O.new_canonical_ordered(vals), S.canonic_half_coset(log_n).initial_index, itw)
    tree = S.precompute_twiddles(big)
    ext = np.zeros(1 << (log_n + blow), np.uint32)
    ext[: 1 << log_n] = coeffs
    cols = [S.BaseFieldVec.from_numpy(ext) for _ in range(n_cols)]
    S.evaluate_columns(cols, tree)
    want = O.evaluate(ext, big.initial_index, tw)
    for c in cols[:: 51]:
        assert (c.to_numpy() == want).all()


@pytest.mark.parametrize("log_n,n_cols", [(15, 11), (19, 9)])
def test_interpolate_evaluate_columns_partial_band(log_n, n_cols):
    """Separate contiguous and strided passes over distinct columns whose count leaves a
    partial band of 8 columns."""
    coset = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(coset)
    itw = tree.itwiddles.to_numpy()
    evals = [rand_m31(1 << log_n, 9100 + c) for c in range(n_cols)]
    cols = [S.BaseFieldVec.from_numpy(e) for e in evals]
    S.interpolate_columns(cols, tree)
    want = [O.interpolate(e, coset.initial_index, itw) for e in evals]
    for c in range(n_cols):
        assert (cols[c].to_numpy() == want[c]).all(), f"interpolate column {c}"
    S.evaluate_columns(cols, tree)
    for c in range(n_cols):
        assert (cols[c].to_numpy() == evals[c]).all(), f"evaluate column {c}"


@pytest.mark.parametrize("log_n,blow", [(3, 1), (11, 1), (12, 1), (13, 1), (16, 1), (16, 2), (17, 3)])
def test_lde_async_matches_oracle(log_n, blow):
    """Interpolate + evaluate-from-coefficients (no extend copy) on several columns."""
    n_cols = 5
    m = log_n + blow
    root = S.canonic_half_coset(m)
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << log_n, 1000 * log_n + c) for c in range(n_cols)]
    cols = [S.BaseFieldVec.from_numpy(e) for e in evals]
    ldes = [S.BaseFieldVec.uninitialized(1 << m) for _ in range(n_cols)]
    S.interpolate_columns_async(cols, [log_n] * n_cols, itw, 1 << root.log_size)
    S.evaluate_columns_async(cols, [log_n] * n_cols, ldes, [m] * n_cols, tw, 1 << root.log_size)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    ocoef, olde = O.lde_columns(evals, blow, otw, oitw)
    for c in range(n_cols):
        assert (cols[c].to_numpy() == ocoef[c]).all()
        assert (ldes[c].to_numpy() == olde[c]).all()


@pytest.mark.parametrize("logs", [[10, 10, 10, 10], [5, 9, 9, 12, 3]])
def test_interpolate_columns_out_of_place(logs):
    """stwo_hip_interpolate_columns_oop_async: coefficients into separate buffers, sources untouched
    (one size takes the single-table upload, several sizes the per-size grouping)."""
    root = S.canonic_half_coset(max(logs))
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << l, 500 + i) for i, l in enumerate(logs)]
    src = [S.BaseFieldVec.from_numpy(e) for e in evals]
    dst = [S.BaseFieldVec.uninitialized(1 << l) for l in logs]
    S.interpolate_columns_oop_async(src, dst, logs, itw, 1 << root.log_size)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    for i, l in enumerate(logs):
        oc, _ = O.lde_columns([evals[i]], 0, otw, oitw)
        assert (dst[i].to_numpy() == oc[0]).all(), f"column {i} (log {l})"
        assert (src[i].to_numpy() == evals[i]).all(), f"source {i} changed"


# every multi-column tile size (evaluation domains of 2^3 .. 2^12) at blowups 1..3, whose
# coefficient count is a template argument of the forward pass, and blowup 4 (the runtime
# path); 7 columns leave a partial column group in every tile shape
@pytest.mark.parametrize("log_m,blow", [(m, b) for m in range(3, 13) for b in (1, 2, 3, 4) if m - b >= 1])
def test_lde_multi_column_tiles(log_m, blow):
    log_n, n_cols = log_m - blow, 7
    root = S.canonic_half_coset(log_m)
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << log_n, 500 * log_m + 10 * blow + c) for c in range(n_cols)]
    cols = [S.BaseFieldVec.from_numpy(e) for e in evals]
    ldes = [S.BaseFieldVec.uninitialized(1 << log_m) for _ in range(n_cols)]
    S.interpolate_columns_async(cols, [log_n] * n_cols, itw, 1 << root.log_size)
    S.evaluate_columns_async(cols, [log_n] * n_cols, ldes, [log_m] * n_cols, tw, 1 << root.log_size)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    ocoef, olde = O.lde_columns(evals, blow, otw, oitw)
    for c in range(n_cols):
        assert (cols[c].to_numpy() == ocoef[c]).all(), f"coefficients of column {c}"
        assert (ldes[c].to_numpy() == olde[c]).all(), f"evaluations of column {c}"


def test_lde_mixed_sizes_one_call():
    logs = [4, 13, 9, 13, 16, 4, 11]
    root = S.canonic_half_coset(17)
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << l, 77 + i) for i, l in enumerate(logs)]
    cols = [S.BaseFieldVec.from_numpy(e) for e in evals]
    ldes = [S.BaseFieldVec.uninitialized(1 << (l + 1)) for l in logs]
    S.interpolate_columns_async(cols, logs, itw, 1 << root.log_size)
    S.evaluate_columns_async(cols, logs, ldes, [l + 1 for l in logs], tw, 1 << root.log_size)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    for i, l in enumerate(logs):
        oc, ol = O.lde_columns([evals[i]], 1, otw, oitw)
        assert (cols[i].to_numpy() == oc[0]).all()
        assert (ldes[i].to_numpy() == ol[0]).all()


# ------------------------------------------------------------------ blake2s.rs:109-192
def _cols_row_times_col(n_cols, n):
    return [np.array([(r * c) % P for r in range(n)], np.uint32) for c in range(n_cols)]


def test_commit_on_first_layer_many_columns():
    log_size = 11
    cols = _cols_row_times_col(35, 1 << log_size)
    got = S.commit_on_layer(log_size, None, [S.BaseFieldVec.from_numpy(c) for c in cols])
    assert (got.to_numpy().reshape(-1) == O.merkle_commit_on_layer(log_size, None, cols)).all()


def test_commit_on_layer_with_previous():
    c12 = _cols_row_times_col(35, 1 << 12)
    c11 = _cols_row_times_col(16, 1 << 11)
    prev = S.commit_on_layer(12, None, [S.BaseFieldVec.from_numpy(c) for c in c12])
    got = S.commit_on_layer(11, prev, [S.BaseFieldVec.from_numpy(c) for c in c11])
    oprev = O.merkle_commit_on_layer(12, None, c12)
    assert (got.to_numpy().reshape(-1) == O.merkle_commit_on_layer(11, oprev, c11)).all()
    # a node layer with no columns
    got2 = S.commit_on_layer(10, got, [])
    assert (got2.to_numpy().reshape(-1) == O.merkle_commit_on_layer(10, O.merkle_commit_on_layer(11, oprev, c11), [])).all()


# trees of at most 2^12 leaves with at most 16 leaf columns are built from the leaves by
# one subtree launch: [0], [3, 1, 0], [12] * 16, [12, 5, 12, ...]; [11] * 17 is not
@pytest.mark.parametrize("sizes", [[16] * 16, [12, 5, 12, 3, 9, 9, 0, 12], [17, 17, 3], [13] * 33 + [2, 1],
                                   [18] + [14] * 40 + [16] * 3 + [11] * 20 + [9], [0], [3, 1, 0], [12] * 16,
                                   [11] * 17 + [2]])
def test_merkle_tree_mixed_degree(sizes):
    cols = [rand_m31(1 << s, 500 + i) for i, s in enumerate(sizes)]
    tree = S.MerkleTree([S.BaseFieldVec.from_numpy(c) for c in cols], sizes).commit_async()
    want = O.merkle_commit(cols)
    for k in range(tree.max_log + 1):
        assert (tree.layers[k].to_numpy().reshape(-1) == want[k]).all(), f"layer {k}"
    assert (tree.root() == want[0]).all()


# ------------------------------------------------------------------ fri.rs:99-310
def _rand_secure(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, P, (4, n), dtype=np.uint32)


@pytest.mark.parametrize("log_n", [1, 2, 5, 13, 20])
def test_fold_line(log_n):
    src = _rand_secure(1 << log_n, log_n)
    alpha = (1, 3, 5, 7)
    dom = S.canonic_half_coset(log_n + 1)  # LineDomain(CanonicCoset(log_n + 1).half_coset())
    tree = S.precompute_twiddles(dom)
    got = S.fold_line(S.SecureColumn.from_numpy(src), alpha, tree)
    assert (got.to_numpy() == O.fold_line(src, dom.initial_index, dom.step_size, alpha)).all()


def test_fold_line_more_than_once():
    log_n = 13
    src = _rand_secure(1 << log_n, 99)
    alpha = (1, 3, 5, 7)
    dom = S.canonic_half_coset(log_n + 1)
    tree = S.precompute_twiddles(dom)
    g = S.SecureColumn.from_numpy(src)
    o = src
    d = dom
    for _ in range(3):
        g = S.fold_line(g, alpha, tree)
        o = O.fold_line(o, d.initial_index, d.step_size, alpha)
        d = d.double()
    assert (g.to_numpy() == o).all()


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 13, 20])
def test_fold_circle_into_line(log_n):
    src = _rand_secure(1 << log_n, 1000 + log_n)
    dst0 = _rand_secure(1 << (log_n - 1), 2000 + log_n)
    alpha = (1, 3, 5, 7)
    hc = S.canonic_half_coset(log_n)
    tree = S.precompute_twiddles(hc)  # line_domain.coset()
    dst = S.SecureColumn.from_numpy(dst0)
    S.fold_circle_into_line(dst, S.SecureColumn.from_numpy(src), alpha, tree)
    assert (dst.to_numpy() == O.fold_circle_into_line(dst0, src, hc.initial_index, hc.step_size, alpha)).all()


@pytest.mark.parametrize("log_circle,n_layers,bigger_tree", [
    (2, 1, False), (2, 2, False), (3, 3, False), (5, 4, True), (13, 13, False), (14, 14, True),
    (16, 16, False), (17, 12, False), (20, 17, True)])
def test_fri_fold_chain(log_circle, n_layers, bigger_tree):
    # fold_circle_into_line then fold_line per layer, every coefficient known up front; the
    # cases cross the chain's launch boundaries (2^13-value block ranges, 512-block minimum)
    rng = np.random.default_rng(3000 + log_circle)
    src = _rand_secure(1 << log_circle, 3100 + log_circle)
    dst0 = _rand_secure(1 << (log_circle - 1), 3200 + log_circle)
    alphas = [tuple(int(x) for x in rng.integers(0, S.P, 4)) for _ in range(n_layers)]
    hc = S.canonic_half_coset(log_circle)
    root = S.canonic_half_coset(log_circle + 1) if bigger_tree else hc  # a larger tree serves it too
    tw, itw = S.cached_twiddles(root)
    layers = [S.SecureColumn.from_numpy(dst0)] + [S.SecureColumn.zeros(1 << (log_circle - 1 - j))
                                                   for j in range(1, n_layers)]
    S.fri_fold_chain_async(S.SecureColumn.from_numpy(src), layers, alphas, itw, 1 << root.log_size)
    S.synchronize()
    cur = O.fold_circle_into_line(dst0, src, hc.initial_index, hc.step_size, alphas[0])
    assert (layers[0].to_numpy() == cur).all(), "circle fold"
    d = S.Coset(hc.initial_index, hc.step_size, hc.log_size)
    for j in range(1, n_layers):
        cur = O.fold_line(cur, d.initial_index, d.step_size, alphas[j])
        d = d.double()
        assert (layers[j].to_numpy() == cur).all(), f"line layer {j}"


def test_fold_circle_into_line_fibonacci_fixture():
    f = FIX["fold_circle_into_line_fibonacci"]
    rc = f["root_coset"]
    tree = S.TwiddleTree(S.Coset(rc["initial_index"], rc["step_size"], rc["log_size"]),
                         S.BaseFieldVec.from_numpy(np.array(f["twiddles"], np.uint32)),
                         S.BaseFieldVec.from_numpy(np.array(f["itwiddles"], np.uint32)))
    dst0 = np.array(f["dst"], np.uint32)
    src = np.array(f["src"], np.uint32)
    dst = S.SecureColumn.from_numpy(dst0)
    S.fold_circle_into_line(dst, S.SecureColumn.from_numpy(src), f["alpha"], tree)
    hc = f["src_half_coset"]
    assert (dst.to_numpy() == O.fold_circle_into_line(dst0, src, hc["initial_index"], hc["step_size"],
                                                       f["alpha"])).all()


# ------------------------------------------------------------------ quotient.rs:132-204
def _point(rng):
    import refmath as R
    return R.random_qm31_point(rng)


def test_accumulate_quotients_reference_shape():
    # log 15 trace, blowup 1, 2 columns, 3 identical batches at one point (quotient.rs:132-204)
    rng = random.Random(15)
    log_n, blow = 15, 1
    m = log_n + blow
    root = S.canonic_half_coset(m)
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    e0 = np.arange(1 << log_n, dtype=np.uint32)
    e1 = (2 * np.arange(1 << log_n, dtype=np.uint64) % P).astype(np.uint32)
    coeffs, ldes = O.lde_columns([e0, e1], blow, otw, oitw)
    pt = _point(rng)
    a = O.eval_at_point(coeffs[0], *pt)
    b = O.eval_at_point(coeffs[1], *pt)
    batches = [(pt, [(0, a), (1, b)])] * 3
    rc = (1, 2, 3, 4)
    got = S.accumulate_quotients(root, [S.BaseFieldVec.from_numpy(c) for c in ldes], rc, batches)
    want = O.accumulate_quotients(root.initial_index, root.step_size, m, ldes, rc, batches)
    assert (got.to_numpy() == want).all()


@pytest.mark.parametrize("sizes", [[5, 1, 9, 2], [7], [3, 16], [20, 20, 20], [4, 4, 4, 4, 4], [1, 1, 1, 1, 1, 1]])
def test_accumulate_quotients_uneven_batches(sizes):
    # batches of different sizes: the reference kernel's line-coefficient offset bug case;
    # up to 4 batches take the column-major kernel, more take the batch-major one
    rng = random.Random(7 + len(sizes))
    m, n_cols = 12, 20
    root = S.canonic_half_coset(m)
    cols = [rand_m31(1 << m, 300 + i) for i in range(n_cols)]
    batches = []
    for b, size in enumerate(sizes):
        idx = rng.sample(range(n_cols), size)
        batches.append((_point(rng), [(c, rand_qm31(rng)) for c in idx]))
    rc = rand_qm31(rng)
    got = S.accumulate_quotients(root, [S.BaseFieldVec.from_numpy(c) for c in cols], rc, batches)
    want = O.accumulate_quotients(root.initial_index, root.step_size, m, cols, rc, batches)
    assert (got.to_numpy() == want).all()


# ------------------------------------------------------------------ accumulation.rs:31-101
def test_accumulation():
    n = 2 << 12
    left = S.SecureColumn.from_numpy(np.stack([np.full(n, v, np.uint32) for v in (1, 2, 3, 4)]))
    right = S.SecureColumn.from_numpy(np.stack([np.full(n, v, np.uint32) for v in (5, 6, 7, 8)]))
    S.accumulate(left, right)
    assert (left.to_numpy()[:, 0] == [6, 8, 10, 12]).all() and (left.to_numpy()[:, -1] == [6, 8, 10, 12]).all()
    ones = S.SecureColumn.from_numpy(np.ones((4, n), np.uint32))
    pm1 = S.SecureColumn.from_numpy(np.full((4, n), P - 1, np.uint32))
    S.accumulate(ones, pm1)
    assert not ones.to_numpy().any()


# ------------------------------------------------------------------ wide_fibonacci.rs:90-141
def test_wide_fib_constraint_quotients():
    rng = random.Random(3)
    n_rows, n_cols = 1 << 13, 32
    trace = [rand_m31(n_rows, 40 + i) for i in range(n_cols)]
    powers = np.array([rand_qm31(rng) for _ in range(n_cols - 2)], np.uint32)
    dinv = rand_m31(n_rows, 5)
    out = S.SecureColumn.uninitialized(n_rows)
    S.wide_fib_constraint_quotients(out, [S.BaseFieldVec.from_numpy(t) for t in trace],
                                    S.SecureFieldVec.from_numpy(powers), S.BaseFieldVec.from_numpy(dinv))
    assert (out.to_numpy() == O.wide_fib_constraint_quotients(trace, powers, dinv)).all()


# ------------------------------------------------------------------ inputs / errors
def test_uniform_generator_matches_oracle():
    col = S.BaseFieldVec.uninitialized(1 << 18)
    S.fill_uniform(col, 0x5EED0002, 12345)
    S.synchronize()
    assert (col.to_numpy() == O.fill_uniform_m31(1 << 18, 0x5EED0002, 12345)).all()


def test_errors_are_reported_not_fatal():
    col = S.BaseFieldVec.from_numpy(np.arange(12, dtype=np.uint32))
    with pytest.raises(S.StwoHipError):
        S.bit_reverse_column(col)  # not a power of two
    S.bit_reverse_column(S.BaseFieldVec.from_numpy(np.arange(16, dtype=np.uint32)))  # still usable


# ------------------------------------------------------------------ one-call LDE (fused middle pass)
@pytest.mark.parametrize("log_n,blow,n_cols,in_place", [(3, 1, 3, False), (13, 1, 2, False), (14, 1, 3, True),
                                                         (15, 2, 2, False), (17, 3, 2, True), (20, 1, 2, False),
                                                         (21, 1, 1, True), (22, 1, 1, False),
                                                         # column counts that leave a partial band of 8
                                                         (14, 1, 11, False), (17, 1, 10, True), (16, 1, 19, False),
                                                         (18, 1, 9, True)])
def test_lde_columns_fused(log_n, blow, n_cols, in_place):
    m = log_n + blow
    root = S.canonic_half_coset(m)
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << log_n, 7000 + 31 * log_n + c) for c in range(n_cols)]
    src = [S.BaseFieldVec.from_numpy(e) for e in evals]
    coeffs = src if in_place else [S.BaseFieldVec.uninitialized(1 << log_n) for _ in range(n_cols)]
    ldes = [S.BaseFieldVec.uninitialized(1 << m) for _ in range(n_cols)]
    S.lde_columns_async(src, coeffs, ldes, [log_n] * n_cols, blow, tw, itw, 1 << root.log_size)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    oc, ol = O.lde_columns(evals, blow, otw, oitw)
    for c in range(n_cols):
        assert (coeffs[c].to_numpy() == oc[c]).all()
        assert (ldes[c].to_numpy() == ol[c]).all()
        if not in_place:
            assert (src[c].to_numpy() == evals[c]).all()  # inputs untouched


def test_async_fri_chain_matches_oracle():
    log_n = 12
    src = _rand_secure(1 << log_n, 4242)
    hc = S.canonic_half_coset(log_n)
    tw, itw = S.cached_twiddles(hc)
    tree_len = 1 << hc.log_size
    rng = np.random.default_rng(1)
    alphas = [S.random_qm31(rng) for _ in range(4)]
    layers = [S.SecureColumn.zeros(1 << l) for l in range(log_n - 1, log_n - 5, -1)]
    S.fold_circle_into_line_async(layers[0], S.SecureColumn.from_numpy(src), alphas[0], itw, tree_len - len(layers[0]))
    for i in range(1, len(layers)):
        S.fold_line_async(layers[i - 1], layers[i], alphas[i], itw, tree_len - len(layers[i - 1]))
    S.synchronize()
    o = O.fold_circle_into_line(np.zeros((4, 1 << (log_n - 1)), np.uint32), src, hc.initial_index, hc.step_size,
                                alphas[0])
    assert (layers[0].to_numpy() == o).all()
    d = S.Coset(hc.initial_index, hc.step_size, hc.log_size)
    for i in range(1, len(layers)):
        o = O.fold_line(o, d.initial_index, d.step_size, alphas[i])
        d = d.double()
        assert (layers[i].to_numpy() == o).all()


@pytest.mark.parametrize("m,sizes", [(13, [4, 7, 1]), (12, [4, 7, 1, 2, 3, 5]), (13, [2, 2, 2, 2, 2, 2])])
def test_async_quotients_match_oracle(m, sizes):
    # up to 4 batches: column-major kernel; more: batch-major kernel with the column table
    # uploaded by the call.  The third case reuses the first one's cached domain tables.
    rng = np.random.default_rng(11 + len(sizes))
    n_cols = 12
    root = S.canonic_half_coset(m)
    cols = [rand_m31(1 << m, 900 + i) for i in range(n_cols)]
    pts = [S.random_circle_point_qm31(rng) for _ in range(len(sizes))]
    batches = [(pts[b], [(int(c), S.random_qm31(rng)) for c in sorted(rng.choice(n_cols, size=k, replace=False))])
               for b, k in enumerate(sizes)]
    rc = S.random_qm31(rng)
    out = S.SecureColumn.uninitialized(1 << m)
    S.accumulate_quotients_async(root, [S.BaseFieldVec.from_numpy(c) for c in cols], rc, S.QuotientSamples(batches),
                                 out)
    S.synchronize()
    want = O.accumulate_quotients(root.initial_index, root.step_size, m, cols, rc, batches)
    assert (out.to_numpy() == want).all()


# ------------------------------------------------------------------ batched commit (poly.rs:57-168 + blake2s.rs:19-97)
MIXED_LOGS = [15, 3, 10, 15, 1, 10, 10, 14, 6, 12, 15, 2]
# a tree of 2^13 leaves: the subtree launch over layers 12..0 finishes in its last block,
# with columns injected below and above the block roots (layer 5)
NARROW_LOGS = [12, 3, 1, 7, 12, 2, 9, 4]


@pytest.mark.parametrize("use_graph,autotune,logs", [(False, False, MIXED_LOGS), (True, False, MIXED_LOGS),
                                                     (True, True, MIXED_LOGS), (False, False, NARROW_LOGS),
                                                     (True, False, NARROW_LOGS)])
def test_commit_plan_mixed_sizes(use_graph, autotune, logs):
    """LDE + Merkle over columns of mixed sizes in one plan; every layer equals the
    oracle's tree over the oracle's LDE, and graph replays reproduce it."""
    blow = 1
    m = max(logs) + blow
    root = S.canonic_half_coset(m)
    tw, itw = S.cached_twiddles(root)
    evals = [rand_m31(1 << l, 8100 + i) for i, l in enumerate(logs)]
    src = [S.BaseFieldVec.from_numpy(e) for e in evals]
    coeffs = [S.BaseFieldVec.uninitialized(1 << l) for l in logs]
    ldes = [S.BaseFieldVec.uninitialized(1 << (l + blow)) for l in logs]
    plan = S.CommitPlan(src, coeffs, ldes, logs, blow, tw, itw, 1 << root.log_size, autotune=autotune)
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    want_lde = []
    for e, l in zip(evals, logs):
        d = S.canonic_half_coset(l + blow)
        t, it = O.precompute_twiddles(d.initial_index, d.step_size, d.log_size)
        want_lde.append(O.lde_columns([e], blow, t, it)[1][0])
    want = O.merkle_commit(want_lde)
    for rep in range(3):
        for l in ldes:
            S.fill_uniform(l, 99 + rep)  # clobber, so each run must rewrite everything
        plan.run(use_graph)
        S.synchronize()
        for i in range(len(logs)):
            assert (ldes[i].to_numpy() == want_lde[i]).all(), f"rep {rep} column {i}"
        for k in range(plan.max_log + 1):
            assert (plan.layers[k].to_numpy().reshape(-1) == want[k]).all(), f"rep {rep} layer {k}"
        assert (plan.root() == want[0]).all()
    plan.close()


# ------------------------------------------------------------------ batched OODS (poly.rs:84-94 over many columns)
@pytest.mark.parametrize("log_n,n_cols", [(1, 3), (5, 7), (11, 9), (12, 8), (13, 17), (16, 5), (20, 2)])
def test_eval_at_point_columns(log_n, n_cols):
    rng = random.Random(90 + log_n)
    px, py = _point(rng)
    cols = [rand_m31(1 << log_n, 9000 + 17 * log_n + c) for c in range(n_cols)]
    got = S.eval_at_point_columns([S.BaseFieldVec.from_numpy(c) for c in cols], px, py)
    for c in range(n_cols):
        assert tuple(got[c]) == O.eval_at_point(cols[c], px, py), f"column {c}"


@pytest.mark.parametrize("log_n,n_cols", [(3, 2), (14, 5)])
def test_eval_at_points_columns(log_n, n_cols):
    rng = random.Random(190 + log_n)
    points = [_point(rng) for _ in range(3)]
    cols = [rand_m31(1 << log_n, 9300 + 17 * log_n + c) for c in range(n_cols)]
    got = S.eval_at_points_columns([S.BaseFieldVec.from_numpy(c) for c in cols], points)
    for k, (px, py) in enumerate(points):
        for c in range(n_cols):
            assert tuple(got[k][c]) == O.eval_at_point(cols[c], px, py), f"point {k} column {c}"


# ------------------------------------------------------------------ Column::at / decommitment gathers
def test_gather_rows_and_auth_paths():
    sizes = [12, 12, 9, 9, 9, 5]
    cols_np = [rand_m31(1 << l, 9500 + i) for i, l in enumerate(sizes)]
    cols = [S.BaseFieldVec.from_numpy(c) for c in cols_np]
    rows = np.array([0, 4095, 17, 17, 2048], np.uint32)
    got = S.gather_rows(cols[:2], rows)
    want = np.stack([cols_np[0][rows], cols_np[1][rows]], axis=1)
    assert (got == want).all()
    with pytest.raises(S.StwoHipError):
        S.gather_rows(cols[2:5], np.array([1, 512], np.uint32))  # row 512 is past a column of 2^9
    tree = S.MerkleTree(cols, sizes)
    tree.commit_async()
    S.synchronize()
    layers = [tree.layers[k].to_numpy().reshape(-1, 8) for k in range(len(tree.layers))]
    want_layers = O.merkle_commit(cols_np)
    assert all((layers[k] == want_layers[k].reshape(-1, 8)).all() for k in range(len(layers)))
    leaves = np.array([0, 1, 4095, 1234, 3000], np.uint32)
    paths = S.merkle_auth_paths(tree.layers, leaves)
    max_log = len(layers) - 1
    for qi, leaf in enumerate(leaves):
        for j in range(max_log):
            assert (paths[qi, j] == layers[max_log - j][(int(leaf) >> j) ^ 1]).all()
    with pytest.raises(S.StwoHipError):
        S.merkle_auth_paths(tree.layers, np.array([1 << max_log], np.uint32))


# ------------------------------------------------------------------ lookups/gkr.rs, lookups/mle.rs
GKR_TEST_Y = [(7, 3, 5, 6), (1, 1, 9, 7), (3, 5, 6, 1), (1, 9, 7, 3), (5, 6, 1, 1), (9, 7, 3, 5), (6, 1, 1, 9)]


def test_gen_eq_evals_reference_input():
    # gkr.rs:46-63: y from [7, 3, 5, 6, 1, 1, 9] repeated 4 times in chunks of 4, v = 2
    got = S.gen_eq_evals(GKR_TEST_Y, (2, 0, 0, 0)).to_numpy()
    assert (got == O.gen_eq_evals(GKR_TEST_Y, (2, 0, 0, 0))).all()


@pytest.mark.parametrize("n_vars", [0, 1, 9, 10, 11, 20])
def test_gen_eq_evals_sizes(n_vars):
    rng = random.Random(600 + n_vars)
    y = [rand_qm31(rng) for _ in range(n_vars)]
    v = rand_qm31(rng)
    got = S.gen_eq_evals(y, v).to_numpy()
    assert (got == O.gen_eq_evals(y, v)).all()


@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable_reference_input(secure):
    # mle.rs:70-109: 2^8 evaluations 0..255 (QM31: words 0..1023 in chunks of 4), assignment (7, 12, 3, 2)
    a = (7, 12, 3, 2)
    if secure:
        ev = np.arange(4 << 8, dtype=np.uint32).reshape(-1, 4)
        got = S.fix_first_variable(S.SecureFieldVec.from_numpy(ev), a).to_numpy()
    else:
        ev = np.arange(1 << 8, dtype=np.uint32)
        got = S.fix_first_variable(S.BaseFieldVec.from_numpy(ev), a).to_numpy()
    assert (got == O.fix_first_variable(ev, a)).all()


@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable_large(secure):
    rng = random.Random(700)
    a = rand_qm31(rng)
    n = 1 << 20
    if secure:
        ev = rand_m31(4 * n, 701).reshape(-1, 4)
        got = S.fix_first_variable(S.SecureFieldVec.from_numpy(ev), a).to_numpy()
    else:
        ev = rand_m31(n, 702)
        got = S.fix_first_variable(S.BaseFieldVec.from_numpy(ev), a).to_numpy()
    assert (got == O.fix_first_variable(ev, a)).all()


# ------------------------------------------------------------------ size guards (checked on the host, before any launch)
def test_oversize_calls_fail_before_launch():
    L = S.lib()
    col = S.BaseFieldVec.from_numpy(np.arange(1 << 10, dtype=np.uint32))
    out = S.SecureColumn.uninitialized(1 << 10)
    samples = S.QuotientSamples([(S.random_circle_point_qm31(np.random.default_rng(1)), [(0, (1, 2, 3, 4))])])
    # no canonic coset of 2^31 points exists (it would be the whole circle group); the guard
    # rejects the size before it looks at the coset, so any (initial index, step) serves
    st = L.stwo_hip_accumulate_quotients_async(1, 2, 1 << 31,
                                               S._ptr_array([col.ptr]), 1, S.qm31((1, 0, 0, 0)), samples.n,
                                               S.C.cast(samples.points, S.vp), samples.sizes.ctypes.data_as(S.u32p),
                                               samples.idx.ctypes.data_as(S.u32p), samples.vals, out.ptrs())
    assert st == 1 and "2^30" in S.last_error()
    L.stwo_hip_clear_error()
    vals = (S.QM31 * 1)()
    st = L.stwo_hip_eval_at_point_columns(S._ptr_array([col.ptr]), 1, 31, S.qm31((1, 0, 0, 0)), S.qm31((0, 1, 0, 0)),
                                          vals)
    assert st == 1
    L.stwo_hip_clear_error()
    L.eval_at_point(col.ptr, 1 << 31, S.qm31((1, 0, 0, 0)), S.qm31((0, 1, 0, 0)))
    assert L.stwo_hip_last_status() == 1
    L.stwo_hip_clear_error()
    table = L.copy_device_pointer_vec_from_host_to_device(S._ptr_array([col.ptr]), 1)
    hashes = S.Blake2sHashVec.uninitialized(1)
    L.commit_on_first_layer(1 << 31, 1, table, hashes.ptr)
    assert L.stwo_hip_last_status() == 1
    L.stwo_hip_clear_error()
    L.cuda_free_memory(table)
    # the library still works
    S.bit_reverse_column(col)
    assert (col.to_numpy() == O.bit_reverse(np.arange(1 << 10, dtype=np.uint32))).all()


def test_quotients_without_samples_are_zero():
    # no sampled column: every numerator is zero and no column is read (none is given)
    L = S.lib()
    m = 10
    dom = S.canonic_half_coset(m)
    out = S.SecureColumn.from_numpy(np.full((4, 1 << m), 5, np.uint32))
    pts = (S.PointQM31 * 1)(S.PointQM31(S.qm31((1, 2, 3, 4)), S.qm31((5, 6, 7, 8))))
    sizes = np.zeros(1, np.uint32)
    idx = np.zeros(1, np.uint32)
    st = L.stwo_hip_accumulate_quotients_async(dom.initial_index, dom.step_size, 1 << m, None, 0,
                                               S.qm31((3, 0, 0, 0)), 1, S.C.cast(pts, S.vp),
                                               sizes.ctypes.data_as(S.u32p), idx.ctypes.data_as(S.u32p),
                                               (S.QM31 * 1)(), out.ptrs())
    assert st == 0
    S.synchronize()
    assert not out.to_numpy().any()
    assert (O.accumulate_quotients(dom.initial_index, dom.step_size, m, [], (3, 0, 0, 0),
                                   [(((1, 2, 3, 4), (5, 6, 7, 8)), [])]) == 0).all()
    # a sample naming a column when none is given is rejected
    sizes[0] = 1
    st = L.stwo_hip_accumulate_quotients_async(dom.initial_index, dom.step_size, 1 << m, None, 0,
                                               S.qm31((3, 0, 0, 0)), 1, S.C.cast(pts, S.vp),
                                               sizes.ctypes.data_as(S.u32p), idx.ctypes.data_as(S.u32p),
                                               (S.QM31 * 1)(), out.ptrs())
    assert st == 1
    L.stwo_hip_clear_error()


def test_commit_plan_autotune_with_aliased_coefficients():
    # coeffs[i] == evals[i]: the autotuner's runs must leave the evaluations in place, so the
    # first run commits the LDE of the caller's evaluations
    logs = [14, 15, 15, 12, 9]
    blow = 1
    m = max(logs) + blow
    root = S.canonic_half_coset(m)
    tw, itw = S.cached_twiddles(root)
    evals_np = [rand_m31(1 << l, 8800 + i) for i, l in enumerate(logs)]
    evals = [S.BaseFieldVec.from_numpy(e) for e in evals_np]
    ldes = [S.BaseFieldVec.uninitialized(1 << (l + blow)) for l in logs]
    plan = S.CommitPlan(evals, evals, ldes, logs, blow, tw, itw, 1 << root.log_size, autotune=True)
    for i in range(len(logs)):
        assert (evals[i].to_numpy() == evals_np[i]).all(), f"evaluations of column {i} kept"
    plan.run(True)
    S.synchronize()
    otw, oitw = O.precompute_twiddles(root.initial_index, root.step_size, root.log_size)
    want_lde = []
    for i, (e, l) in enumerate(zip(evals_np, logs)):
        oc, ol = O.lde_columns([e], blow, otw, oitw)
        assert (evals[i].to_numpy() == oc[0]).all(), f"coefficients of column {i}"
        assert (ldes[i].to_numpy() == ol[0]).all(), f"LDE of column {i}"
        want_lde.append(ol[0])
    assert (plan.root() == O.merkle_commit(want_lde)[0]).all()
    plan.close()
