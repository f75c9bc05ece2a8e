"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Each test checks the oracle against something it does not itself compute:
the reference's stored fixture (the log-6 twiddle tree, poly.rs:451-595),
Python's hashlib BLAKE2s, closed forms from the reference tests
(fri.rs:141-192, accumulation.rs:31-101, field.rs:41-78), or brute-force
evaluation of the explicit polynomial bases in tests/refmath.py.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import oracle as O
import refmath as R

P = R.P
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))


# ------------------------------------------------------------------ fields
def test_m31_edges():
    for a in (0, 1, 2, P - 1, 1 << 30, (1 << 30) + 7):
        for b in (0, 1, P - 1, 1 << 30, 12345):
            assert O.m31_mul(a, b) == a * b % P
            assert O.m31_add(a, b) == (a + b) % P
            assert O.m31_sub(a, b) == (a - b) % P
        assert O.m31_neg(a) == (-a) % P  # canonical: neg(0) = 0
        if a:
            assert O.m31_mul(a, O.m31_inv(a)) == 1


def test_extension_closed_forms():
    i = (0, 1, 0, 0)
    u = (0, 0, 1, 0)
    assert O.qm31_mul(i, i) == (P - 1, 0, 0, 0)  # i^2 = -1
    assert O.qm31_mul(u, u) == (2, 1, 0, 0)  # u^2 = 2 + i
    rng = random.Random(1)
    for _ in range(200):
        a = tuple(rng.randrange(P) for _ in range(4))
        b = tuple(rng.randrange(P) for _ in range(4))
        assert O.qm31_mul(a, b) == R.qm_mul(a, b)
        assert O.qm31_add(a, b) == R.qm_add(a, b)
        assert O.qm31_sub(a, b) == R.qm_sub(a, b)
        if a != (0, 0, 0, 0):
            assert O.qm31_mul(a, O.qm31_inv(a)) == (1, 0, 0, 0)
        e = rng.randrange(1 << 40)
        assert O.qm31_pow(a, e) == R.qm_pow(a, e)
    for _ in range(50):
        c = (rng.randrange(P), rng.randrange(P))
        ci = O.cm31_inv(c)
        assert R.cm_mul(c, ci) == (1, 0)
    # an element with a zero coordinate pair exercises neg(0)
    a = (5, 0, 0, 7)
    assert O.qm31_mul(a, O.qm31_inv(a)) == (1, 0, 0, 0)
    assert all(v < P for v in O.qm31_inv(a))


# ------------------------------------------------------------------ circle group
def test_circle_group():
    g = O.point_at_index(1)
    assert g == R.G
    assert (g[0] ** 2 + g[1] ** 2) % P == 1
    assert O.point_at_index(1 << 30) == (P - 1, 0)
    assert O.point_at_index(0) == (1, 0)
    rng = random.Random(2)
    for _ in range(50):
        a, b = rng.randrange(1 << 31), rng.randrange(1 << 31)
        assert R.pt_mul(O.point_at_index(a), O.point_at_index(b)) == O.point_at_index((a + b) % (1 << 31))
        assert O.point_at_index(a) == R.point_at_index(a)


def test_circle_domain_points():
    hi, hs, hl = R.canonic_half_coset(5)
    pts = R.circle_domain_points(hi, hs, hl)
    for i, p in enumerate(pts):
        assert O.circle_domain_at(hi, hs, hl, i) == p
    assert len(set(pts)) == 32


# ------------------------------------------------------------------ twiddles
def test_twiddle_tree_matches_reference_fixture():
    f = FIX["interpolate_from_fib"]
    rc = f["root_coset"]
    tw, itw = O.precompute_twiddles(rc["initial_index"], rc["step_size"], rc["log_size"])
    assert tw.tolist() == f["twiddles"]
    assert itw.tolist() == f["itwiddles"]
    # the root coset is CanonicCoset(7).half_coset
    assert (rc["initial_index"], rc["step_size"], rc["log_size"]) == R.canonic_half_coset(7)


# ------------------------------------------------------------------ FFT
def _brute_evals(coeffs, half_coset):
    pts = R.circle_domain_points(*half_coset)
    k = len(coeffs).bit_length() - 1
    return [R.circle_poly_eval_m31(list(map(int, coeffs)), pts[R.brev(i, k)]) for i in range(len(coeffs))]


def test_interpolate_fixture_domain_brute_force():
    f = FIX["interpolate_from_fib"]
    d = f["domain_half_coset"]
    values = np.array(f["values"], np.uint32)
    coeffs = O.interpolate(values, d["initial_index"], np.array(f["itwiddles"], np.uint32))
    assert _brute_evals(coeffs, (d["initial_index"], d["step_size"], d["log_size"])) == f["values"]
    back = O.evaluate(coeffs, d["initial_index"], np.array(f["twiddles"], np.uint32))
    assert back.tolist() == f["values"]


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5, 6, 7])
def test_fft_brute_force_canonic(log_n):
    rng = np.random.default_rng(log_n)
    hc = R.canonic_half_coset(log_n)
    root = R.canonic_half_coset(log_n + 2)  # a larger tree serves smaller domains
    tw, itw = O.precompute_twiddles(*root)
    values = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    coeffs = O.interpolate(values, hc[0], itw)
    assert _brute_evals(coeffs, hc) == values.tolist()
    assert O.evaluate(coeffs, hc[0], tw).tolist() == values.tolist()


def test_interpolate_2_fixture_roundtrip():
    # poly.rs:291-327 input (canonical order) -> bit-reversed circle evaluation
    vals = np.array(FIX["interpolate_2_values"], np.uint32)
    ev = O.new_canonical_ordered(vals)
    hc = R.canonic_half_coset(5)
    tw, itw = O.precompute_twiddles(*hc)
    coeffs = O.interpolate(ev, hc[0], itw)
    assert _brute_evals(coeffs, hc) == ev.tolist()


def test_lde_brute_force():
    rng = np.random.default_rng(9)
    log_n, blow = 5, 1
    tw, itw = O.precompute_twiddles(*R.canonic_half_coset(log_n + blow))
    cols = [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(3)]
    coeffs, ldes = O.lde_columns(cols, blow, tw, itw)
    big = R.canonic_half_coset(log_n + blow)
    pts = R.circle_domain_points(*big)
    for c, co, lde in zip(cols, coeffs, ldes):
        assert _brute_evals(co, R.canonic_half_coset(log_n)) == c.tolist()
        m = log_n + blow
        want = [R.circle_poly_eval_m31(co.tolist(), pts[R.brev(i, m)]) for i in range(1 << m)]
        assert lde.tolist() == want


@pytest.mark.parametrize("log_n", [0, 1, 2, 3, 6])
def test_eval_at_point_brute_force(log_n):
    rng = random.Random(log_n)
    coeffs = [rng.randrange(P) for _ in range(1 << log_n)]
    px, py = R.random_qm31_point(rng)
    assert O.eval_at_point(np.array(coeffs, np.uint32), px, py) == R.circle_poly_eval_qm31(coeffs, px, py)


def test_new_canonical_ordered_semantics():
    # canonical (coset-order) values land at the bit-reversed circle-domain slot of the same point
    log_n = 5
    n = 1 << log_n
    coset = R.coset_points(1 << (30 - log_n), 1 << (31 - log_n), log_n)  # CanonicCoset(n).coset
    dom = R.circle_domain_points(*R.canonic_half_coset(log_n))
    vals = np.arange(100, 100 + n, dtype=np.uint32)
    out = O.new_canonical_ordered(vals)
    for k, p in enumerate(coset):
        j = dom.index(p)
        assert out[R.brev(j, log_n)] == vals[k]


# ------------------------------------------------------------------ columns
def test_bit_reverse():
    a = np.arange(1 << 10, dtype=np.uint32)
    b = O.bit_reverse(a)
    assert all(b[i] == R.brev(i, 10) for i in range(1 << 10))
    assert (O.bit_reverse(b) == a).all()
    q = np.arange(4 * 256, dtype=np.uint32)
    qb = O.bit_reverse_qm31(q).reshape(-1, 4)
    for i in range(256):
        assert qb[i].tolist() == q.reshape(-1, 4)[R.brev(i, 8)].tolist()


def test_batch_inverse():
    a = np.arange(1, 1 + (1 << 12), dtype=np.uint32)
    inv = O.batch_inverse_m31(a)
    assert all(int(x) * int(y) % P == 1 for x, y in zip(a, inv))
    rng = np.random.default_rng(3)
    q = rng.integers(1, P, 4 * 512, dtype=np.uint32)
    qi = O.batch_inverse_qm31(q).reshape(-1, 4)
    for x, y in zip(q.reshape(-1, 4), qi):
        assert R.qm_mul(tuple(map(int, x)), tuple(map(int, y))) == (1, 0, 0, 0)


# ------------------------------------------------------------------ blake2s / merkle
@pytest.mark.parametrize("n", [0, 1, 3, 63, 64, 65, 128, 200])
def test_blake2s256_vs_hashlib(n):
    data = bytes((7 * i + 3) % 256 for i in range(n))
    assert O.blake2s256(data) == hashlib.blake2s(data, digest_size=32).digest()


def test_compress_parameters_vs_hashlib():
    # one-block message: h = IV ^ param, counter = len, last-block flag all ones
    iv = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
    for n in (5, 40, 64):
        data = bytes(range(n))
        h = list(iv)
        h[0] ^= 0x01010020
        m = np.frombuffer(data + bytes(64 - n), dtype="<u4")
        out = O.blake2s_compress(h, m, n, 0, 0xFFFFFFFF, 0)
        assert out.astype("<u4").tobytes() == hashlib.blake2s(data, digest_size=32).digest()


def _hash_node_model(children, values):
    """stwo Blake2sMerkleHasher::hash_node restated in Python: zero state,
    optional compression of left||right, then zero-padded 16-value chunks."""
    st = np.zeros(8, np.uint32)
    if children is not None:
        st = O.blake2s_compress(st, np.concatenate(children))
    vals = list(values) + [0] * ((-len(values)) % 16)
    for c in range(0, len(vals), 16):
        st = O.blake2s_compress(st, np.array(vals[c:c + 16], np.uint32))
    return st


def test_merkle_layers_reference_shapes():
    # blake2s.rs:109-171 shapes: 35 columns at log 12 then 16 columns at log 11, value row*col
    def cols(nc, n):
        return [np.array([(r * c) % P for r in range(n)], np.uint32) for c in range(nc)]
    c12, c11 = cols(35, 1 << 12), cols(16, 1 << 11)
    l12 = O.merkle_commit_on_layer(12, None, c12).reshape(-1, 8)
    l11 = O.merkle_commit_on_layer(11, l12.reshape(-1), c11).reshape(-1, 8)
    for r in (0, 1, 77, 4095):
        assert (l12[r] == _hash_node_model(None, [int(c[r]) for c in c12])).all()
    for r in (0, 5, 2047):
        assert (l11[r] == _hash_node_model((l12[2 * r], l12[2 * r + 1]), [int(c[r]) for c in c11])).all()
    # a node layer without columns, and a leaf layer without columns (no compression: zero hash)
    l10 = O.merkle_commit_on_layer(10, l11.reshape(-1), []).reshape(-1, 8)
    assert (l10[3] == _hash_node_model((l11[6], l11[7]), [])).all()
    assert (O.merkle_commit_on_layer(2, None, []) == 0).all()


def test_merkle_commit_mixed_sizes_injection():
    rng = np.random.default_rng(4)
    sizes = [6, 4, 6, 3, 4]
    cols = [rng.integers(0, P, 1 << s, dtype=np.uint32) for s in sizes]
    layers = O.merkle_commit(cols)
    # layer 4: children from layer 5 (no columns of size 5) plus the two log-4 columns in order
    l5 = layers[5].reshape(-1, 8)
    l4 = layers[4].reshape(-1, 8)
    for r in range(16):
        assert (l4[r] == _hash_node_model((l5[2 * r], l5[2 * r + 1]), [int(cols[1][r]), int(cols[4][r])])).all()
    assert layers[0].size == 8


# ------------------------------------------------------------------ FRI closed forms
def test_fold_line_closed_form():
    # fri.rs:141-192: poly = [even || odd] in the line basis, alpha = 19283,
    # domain LineDomain(Coset::half_odds(3)); folded = 2 (f_e + alpha f_o) on domain.double()
    even = [1, 2, 1, 3]
    odd = [3, 5, 4, 1]
    alpha = (19283, 0, 0, 0)
    li, ls, ll = 1 << 26, 1 << 28, 3  # half_odds(3): initial G_{2^5}, step G_{2^3}
    dom = R.coset_points(li, ls, ll)
    values = [R.line_poly_eval(even + odd, R.qm(x)) for x, _ in dom]
    values = [values[R.brev(i, ll)] for i in range(1 << ll)]
    src = np.array(values, np.uint32).T
    folded = O.fold_line(src, li, ls, alpha)
    folded = [tuple(int(v) for v in folded[:, R.brev(i, ll - 1)]) for i in range(1 << (ll - 1))]
    drp = R.coset_points(2 * li, 2 * ls, ll - 1)
    for got, (x, _) in zip(folded, drp):
        want = R.qm_scale(R.qm_add(R.line_poly_eval(even, R.qm(x)),
                                   R.qm_mul(alpha, R.line_poly_eval(odd, R.qm(x)))), 2)
        assert got == want


def test_fold_circle_into_line_closed_form():
    # a circle polynomial f = f0(x) + y f1(x) in the circle basis; folding p, -p with alpha
    # gives dst' = dst alpha^2 + 2 (f0(x) + alpha f1(x)) at the half-coset x coordinates
    rng = random.Random(5)
    log_n = 5
    n = 1 << log_n
    coeffs = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(n)]
    hc = R.canonic_half_coset(log_n)
    pts = R.circle_domain_points(*hc)
    ev = [R.circle_poly_eval_qm31(coeffs, R.qm(x), R.qm(y)) for x, y in pts]
    ev = [ev[R.brev(i, log_n)] for i in range(n)]
    alpha = tuple(rng.randrange(P) for _ in range(4))
    dst0 = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(n // 2)]
    out = O.fold_circle_into_line(np.array(dst0, np.uint32).T, np.array(ev, np.uint32).T, hc[0], hc[1], alpha)
    f0 = [coeffs[2 * j] for j in range(n // 2)]
    f1 = [coeffs[2 * j + 1] for j in range(n // 2)]

    def circ_x_poly(cs, x):  # basis factors x, pi(x), ... for bits 0.. of the half index
        k = len(cs).bit_length() - 1
        fac, cur = [], R.qm(x)
        for _ in range(k):
            fac.append(cur)
            cur = R.qm_sub(R.qm_scale(R.qm_mul(cur, cur), 2), R.qm(1))
        tot = R.qm(0)
        for idx, c in enumerate(cs):
            t = c
            for b in range(k):
                if (idx >> b) & 1:
                    t = R.qm_mul(t, fac[b])
            tot = R.qm_add(tot, t)
        return tot

    half = R.coset_points(*hc)
    a2 = R.qm_mul(alpha, alpha)
    for i in range(n // 2):
        x = half[R.brev(i, log_n - 1)][0]
        want = R.qm_add(R.qm_mul(dst0[i], a2),
                        R.qm_scale(R.qm_add(circ_x_poly(f0, x), R.qm_mul(alpha, circ_x_poly(f1, x))), 2))
        assert tuple(int(v) for v in out[:, i]) == want


# ------------------------------------------------------------------ quotients
def _interp_low_degree(col4, log_size, half_initial, itw):
    for k in range(4):
        co = O.interpolate(col4[k], half_initial, itw)
        if co[(1 << (log_size - 1)):].any():
            return False
    return True


def test_quotients_low_degree_and_broken_sample():
    rng = random.Random(6)
    log_n, blow = 5, 1
    m = log_n + blow
    tw, itw = O.precompute_twiddles(*R.canonic_half_coset(m))
    hc = R.canonic_half_coset(m)
    nrng = np.random.default_rng(6)
    trace = [nrng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(4)]
    coeffs, ldes = O.lde_columns(trace, blow, tw, itw)
    pts = [R.random_qm31_point(rng) for _ in range(3)]
    batches = []
    for b, p in enumerate(pts):
        cols = [c for c in range(4) if (c + b) % 3 != 0]
        batches.append((p, [(c, O.eval_at_point(coeffs[c], p[0], p[1])) for c in cols]))
    rc = tuple(rng.randrange(P) for _ in range(4))
    q = O.accumulate_quotients(hc[0], hc[1], m, ldes, rc, batches)
    assert _interp_low_degree(q, m, hc[0], itw)
    # one wrong sample value destroys low degree
    bad = list(batches)
    p, cv = bad[1]
    cv = list(cv)
    col, val = cv[0]
    cv[0] = (col, ((val[0] + 1) % P,) + tuple(val[1:]))
    bad[1] = (p, cv)
    q2 = O.accumulate_quotients(hc[0], hc[1], m, ldes, rc, bad)
    assert not _interp_low_degree(q2, m, hc[0], itw)


# The DEEP quotient of stwo's complex-conjugate line method, written from its
# definition rather than from the restatement's code: for domain point p,
#   Q(p) = sum_b W_b * sum_j rc^(j+1) * c_b * (f_bj(p) - L_bj(p.y)) / D_b(p)
# with c_b = conj(P_b.y) - P_b.y, L_bj the line through (P_b.y, v_bj) and its
# conjugate, D_b(p) the pair vanishing of P_b and conj(P_b) at p divided by -2u
# (the CM31 value quotients.cu:91-109 inverts), W_b = rc^(sum of |b'| over the
# batches after b) (the batch chain acc <- acc * rc^|b| + ..., quotients.cu:70,167)
# and alpha_j = rc^(j+1) within a batch (quotients.cu:76-80).
def _deep_denominator(point, p):
    d = R.qm_div(R.pair_vanishing(point, (R.qm_conj(point[0]), R.qm_conj(point[1])), (R.qm(p[0]), R.qm(p[1]))),
                 R.qm_scale(R.U, P - 2))
    assert d[2] == 0 and d[3] == 0, "the denominator lies in CM31"
    return d


def _deep_quotient_row(p, row, ldes, rc, batches):
    total = R.qm(0)
    for b, (point, cv) in enumerate(batches):
        w = R.qm_pow(rc, sum(len(c) for _, c in batches[b + 1:]))
        c_b = R.qm_sub(R.qm_conj(point[1]), point[1])
        dinv = R.qm_inv(_deep_denominator(point, p))
        for j, (col, v) in enumerate(cv):
            f = R.qm(int(ldes[col][row]))
            t = R.qm_mul(c_b, R.qm_sub(f, R.conj_line(point, v, R.qm(p[1]))))
            total = R.qm_add(total, R.qm_mul(R.qm_mul(w, R.qm_pow(rc, j + 1)), R.qm_mul(t, dinv)))
    return total


def _quotient_case(seed, layout, log_n=3, blow=1, true_samples=False):
    """layout: per batch, the list of sampled column indices."""
    rng = random.Random(seed)
    m = log_n + blow
    hc = R.canonic_half_coset(m)
    tw, itw = O.precompute_twiddles(*hc)
    n_cols = 1 + max(c for cols in layout for c in cols)
    nrng = np.random.default_rng(seed)
    trace = [nrng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(n_cols)]
    coeffs, ldes = O.lde_columns(trace, blow, tw, itw)
    batches = []
    for cols in layout:
        pt = R.random_qm31_point(rng)
        vals = [O.eval_at_point(coeffs[c], pt[0], pt[1]) if true_samples else
                tuple(rng.randrange(P) for _ in range(4)) for c in cols]
        batches.append((pt, list(zip(cols, vals))))
    rc = tuple(rng.randrange(P) for _ in range(4))
    pts = R.circle_domain_points(*hc)
    dom = [pts[R.brev(r, m)] for r in range(1 << m)]
    return hc, m, ldes, rc, batches, dom


def _quotients(case):
    hc, m, ldes, rc, batches, _ = case
    return O.accumulate_quotients(hc[0], hc[1], m, ldes, rc, batches)


def _matches_definition(q, case, rows=None):
    _, m, ldes, rc, batches, dom = case
    for r in rows if rows is not None else range(1 << m):
        if tuple(int(v) for v in q[:, r]) != _deep_quotient_row(dom[r], r, ldes, rc, batches):
            return False
    return True


QUOT_LAYOUTS = {
    "one-sample": [[0]],
    "one-batch": [[0, 2, 1]],
    "uneven": [[0, 1, 2], [1], [2, 0]],
    "four": [[1, 0], [3], [0, 1, 2, 3], [2]],
}


@pytest.mark.parametrize("name", sorted(QUOT_LAYOUTS))
@pytest.mark.parametrize("true_samples", [False, True])
def test_quotients_match_definition(name, true_samples):
    case = _quotient_case(11, QUOT_LAYOUTS[name], true_samples=true_samples)
    assert _matches_definition(_quotients(case), case)


def _perturbation_holds(case):
    # Moving one sampled value v_bj by delta changes every row by exactly
    #   -W_b rc^(j+1) c_b (delta + (conj delta - delta)(p.y - P_b.y)/c_b) / D_b(p).
    hc, m, ldes, rc, batches, dom = case
    rng = random.Random(5)
    for b, j in ((0, 2), (1, 0), (2, 1)):
        delta = tuple(rng.randrange(P) for _ in range(4))
        bad = [(pt, list(cv)) for pt, cv in batches]
        col, v = bad[b][1][j]
        bad[b][1][j] = (col, R.qm_add(v, delta))
        diff = O.accumulate_quotients(hc[0], hc[1], m, ldes, rc, bad).astype(np.int64) - _quotients(case)
        point = batches[b][0]
        c_b = R.qm_sub(R.qm_conj(point[1]), point[1])
        w = R.qm_mul(R.qm_pow(rc, sum(len(c) for _, c in batches[b + 1:])), R.qm_pow(rc, j + 1))
        for r in range(1 << m):
            y = R.qm(dom[r][1])
            dl = R.qm_add(delta, R.qm_div(R.qm_mul(R.qm_sub(R.qm_conj(delta), delta), R.qm_sub(y, point[1])), c_b))
            want = R.qm_sub(R.qm(0), R.qm_mul(w, R.qm_div(R.qm_mul(c_b, dl), _deep_denominator(point, dom[r]))))
            if tuple(int(d) % P for d in diff[:, r]) != want:
                return False
    return True


def test_quotients_sample_perturbation():
    assert _perturbation_holds(_quotient_case(12, QUOT_LAYOUTS["uneven"]))


@pytest.mark.parametrize("mut", [O.MUT_QUOT_ALPHA_POW, O.MUT_QUOT_BATCH_POW, O.MUT_QUOT_BATCH_ORDER,
                                 O.MUT_QUOT_REF_OFFSET])
def test_quotient_checks_reject_planted_mistakes(mut):
    # Each planted coefficient mistake must fail the definition check; the
    # low-degree check alone would pass the first three with true samples.
    case = _quotient_case(13, QUOT_LAYOUTS["uneven"], true_samples=True)
    with O.mutation(mut):
        q = _quotients(case)
    assert not _matches_definition(q, case, rows=range(4))
    assert _matches_definition(_quotients(case), case, rows=range(4))
    if mut != O.MUT_QUOT_REF_OFFSET:  # a storage-offset mistake may keep differences linear
        with O.mutation(mut):
            assert not _perturbation_holds(case)


def test_accumulate_closed_form():
    # accumulation.rs:31-101: (1,2,3,4)+(5,6,7,8) and 1 + (P-1) = 0
    n = 64
    left = [np.full(n, v, np.uint32) for v in (1, 2, 3, 4)]
    right = [np.full(n, v, np.uint32) for v in (5, 6, 7, 8)]
    assert O.accumulate(left, right)[:, 0].tolist() == [6, 8, 10, 12]
    ones = [np.full(n, 1, np.uint32)] * 4
    pm1 = [np.full(n, P - 1, np.uint32)] * 4
    assert not O.accumulate(ones, pm1).any()


def test_wide_fib_quotient_low_degree():
    # A trace with c_{i+2} = c_i^2 + c_{i+1}^2 has constraint quotients (over the trace
    # coset's vanishing polynomial) of degree <= n/2.  Evaluated on a 4n domain (so the
    # degree-n constraint is not aliased), their coefficients above index n vanish;
    # a single corrupted trace cell breaks that.
    log_n, n_cols = 4, 8
    n = 1 << log_n
    rng = np.random.default_rng(7)
    t = [rng.integers(0, P, n, dtype=np.uint64), rng.integers(0, P, n, dtype=np.uint64)]
    for i in range(2, n_cols):
        t.append((t[i - 2] * t[i - 2] + t[i - 1] * t[i - 1]) % P)
    trace = [c.astype(np.uint32) for c in t]
    m = log_n + 2
    tw, itw = O.precompute_twiddles(*R.canonic_half_coset(m))
    hc = R.canonic_half_coset(m)
    pts = R.circle_domain_points(*hc)

    def vanish(p):
        # CanonicCoset(log_n).coset = odd multiples of G_{2n}: x after log_n - 1 doublings
        x = p[0]
        for _ in range(log_n - 1):
            x = (2 * x * x - 1) % P
        return x
    trace_coset = R.coset_points(1 << (30 - log_n), 1 << (31 - log_n), log_n)
    assert all(vanish(p) == 0 for p in trace_coset)
    den = np.array([vanish(pts[R.brev(i, m)]) for i in range(1 << m)], np.uint32)
    assert den.all()
    dinv = O.batch_inverse_m31(den)
    alpha = (3, 1, 4, 1)
    powers = np.array([R.qm_pow(alpha, k) for k in range(n_cols - 2)], np.uint32)

    def quotient_degree_ok(tr):
        _, ext = O.lde_columns(tr, 2, tw, itw)
        out = O.wide_fib_constraint_quotients(ext, powers, dinv)
        return all(not O.interpolate(out[k], hc[0], itw)[n + 1:].any() for k in range(4))
    assert quotient_degree_ok(trace)
    bad = [c.copy() for c in trace]
    bad[5][3] = (int(bad[5][3]) + 1) % P
    assert not quotient_degree_ok(bad)


# The composition numerator is stwo's accumulator chain over the constraints
# in order, acc <- acc * alpha + C_i (PointEvaluationAccumulator: the verifier
# evaluates the composition at the OODS point this way), so constraint i
# carries alpha^(n-3-i); with powers [1, alpha, ...] that is coeff_pow[n-3-i]
# (examples/wide_fibonacci.cu:13-27).
def _wf_case(seed, n_cols=6, n=8):
    rng = np.random.default_rng(seed)
    trace = [rng.integers(0, P, n, dtype=np.uint32) for _ in range(n_cols)]
    dinv = rng.integers(1, P, n, dtype=np.uint32)
    alpha = tuple(int(v) for v in rng.integers(0, P, 4))
    powers = np.array([R.qm_pow(alpha, k) for k in range(n_cols - 2)], np.uint32)
    return trace, dinv, alpha, powers


def _wf_constraint(trace, i, r):
    a, b, c = (int(trace[i + k][r]) for k in range(3))
    return (a * a + b * b - c) % P


def _wf_horner(trace, dinv, alpha, r):
    acc = R.qm(0)
    for i in range(len(trace) - 2):
        acc = R.qm_add(R.qm_mul(acc, alpha), R.qm(_wf_constraint(trace, i, r)))
    return R.qm_scale(acc, int(dinv[r]))


def _wf_matches(out, trace, dinv, alpha):
    return all(tuple(int(v) for v in out[:, r]) == _wf_horner(trace, dinv, alpha, r)
               for r in range(trace[0].size))


def test_wide_fib_composition_matches_accumulator_chain():
    trace, dinv, alpha, powers = _wf_case(21)
    assert _wf_matches(O.wide_fib_constraint_quotients(trace, powers, dinv), trace, dinv, alpha)


def test_wide_fib_composition_perturbation():
    # Raising t_k at row r by d changes that row, and only that row, by
    #   dinv * sum over the constraints i that read t_k of alpha^(n-3-i) * dC_i.
    trace, dinv, alpha, powers = _wf_case(22)
    n_cols = len(trace)
    base = O.wide_fib_constraint_quotients(trace, powers, dinv)
    for k, r, d in ((n_cols - 1, 3, 5), (2, 0, 1), (4, 6, 1234567)):
        bad = [c.copy() for c in trace]
        bad[k][r] = (int(bad[k][r]) + d) % P
        diff = (O.wide_fib_constraint_quotients(bad, powers, dinv).astype(np.int64) - base) % P
        want = R.qm(0)
        for i in range(n_cols - 2):
            if i <= k <= i + 2:
                dc = (_wf_constraint(bad, i, r) - _wf_constraint(trace, i, r)) % P
                want = R.qm_add(want, R.qm_scale(R.qm_pow(alpha, n_cols - 3 - i), dc))
        assert tuple(int(v) for v in diff[:, r]) == R.qm_scale(want, int(dinv[r]))
        assert not np.delete(diff, r, axis=1).any()


def test_wide_fib_check_rejects_planted_power_order():
    trace, dinv, alpha, powers = _wf_case(23)
    with O.mutation(O.MUT_WF_POWER_ORDER):
        out = O.wide_fib_constraint_quotients(trace, powers, dinv)
    assert not _wf_matches(out, trace, dinv, alpha)


# ------------------------------------------------------------------ multilinear ops
# gen_eq_evals and fix_first_variable (gkr.cu, mle.cu) against the multilinear
# extension written from its definition, sum_x f(x) eq(x, r), with variable 0
# the most significant index bit.
GKR_TEST_Y = [(7, 3, 5, 6), (1, 1, 9, 7), (3, 5, 6, 1), (1, 9, 7, 3), (5, 6, 1, 1), (9, 7, 3, 5),
              (6, 1, 1, 9)]  # lookups/gkr.rs:52-58: [7, 3, 5, 6, 1, 1, 9] repeated 4 times, in chunks of 4


def test_gen_eq_evals_definition():
    two = R.qm(2)
    out = O.gen_eq_evals(GKR_TEST_Y, two)
    n = len(GKR_TEST_Y)
    for x in range(1 << n):
        assert tuple(int(v) for v in out[x]) == R.qm_mul(two, R.eq_poly(R.index_bits(x, n), GKR_TEST_Y))
    # a boolean y makes eq the indicator of the index whose bits (MSB first) are y
    yb = [R.qm(b) for b in (1, 0, 1, 1)]
    ind = O.gen_eq_evals(yb, R.qm(1))
    assert [int(v[0]) for v in ind] == [int(x == 0b1011) for x in range(16)] and not ind[:, 1:].any()
    assert O.gen_eq_evals([], (5, 6, 7, 8)).tolist() == [[5, 6, 7, 8]]


@pytest.mark.parametrize("secure", [False, True])
def test_fix_first_variable_chain_is_the_multilinear_extension(secure):
    rng = random.Random(31)
    n = 5
    if secure:
        f = np.array([[rng.randrange(P) for _ in range(4)] for _ in range(1 << n)], np.uint32)
        fv = [tuple(int(v) for v in row) for row in f]
    else:
        f = np.array([rng.randrange(P) for _ in range(1 << n)], np.uint32)
        fv = [int(v) for v in f]
    r = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(n)]
    cur = O.fix_first_variable(f, r[0])
    for k in range(1, n):
        cur = O.fix_first_variable(cur, r[k])
    assert tuple(int(v) for v in cur[0]) == R.mle_eval(fv, r)
    # assignments 0 and 1 select the low and high halves
    lo, hi = O.fix_first_variable(f, R.qm(0)), O.fix_first_variable(f, R.qm(1))
    half = 1 << (n - 1)
    if secure:
        assert (lo == f[:half]).all() and (hi == f[half:]).all()
    else:
        assert (lo[:, 0] == f[:half]).all() and (hi[:, 0] == f[half:]).all() and not lo[:, 1:].any()


def test_gen_eq_evals_weights_the_multilinear_extension():
    # sum_x gen_eq_evals(r, 1)[x] f(x) = MLE(f)(r)
    rng = random.Random(32)
    n = 4
    f = [rng.randrange(P) for _ in range(1 << n)]
    r = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(n)]
    w = O.gen_eq_evals(r, R.qm(1))
    tot = R.qm(0)
    for x in range(1 << n):
        tot = R.qm_add(tot, R.qm_scale(tuple(int(v) for v in w[x]), f[x]))
    assert tot == R.mle_eval(f, r)


# ------------------------------------------------------------------ inputs
def test_uniform_generator():
    a = O.fill_uniform_m31(1 << 16, 0x5EED0002)
    assert a.max() < P
    assert abs(a.mean() / P - 0.5) < 0.01
    assert (O.fill_uniform_m31(100, 0x5EED0002, offset=50) == a[50:150]).all()
    assert O.lib().or_uniform_m31(0x5EED0002, 7) == a[7]
