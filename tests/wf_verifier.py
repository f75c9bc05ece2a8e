"""Verifier for the wide-Fibonacci proofs of the HIP prover (test infrastructure).

It replays the Fiat-Shamir channel with hashlib's BLAKE2s-256, checks every Merkle
decommitment with the oracle's BLAKE2s compression, recomputes the DEEP quotients and
the FRI folds at the queried rows from the decommitted values, checks the OODS
relation of the composition polynomial, the proof of work and the degree of the last
FRI layer.  Field and circle arithmetic come from refmath (independent Python).

Protocol and proof layout: the header comment of stwo-gpu_amd/csrc/prover.hip and
channel.hpp; the same statement is the WideFibComponent of the reference
(stwo_gpu_backend/src/examples/wide_fibonacci.rs:90-141).
"""
import hashlib
import struct

import numpy as np

import oracle as O
import refmath as R

P = R.P
MAGIC, VERSION, TAG = 0x50575453, 1, 0x57464942


class VerificationError(Exception):
    pass


def check(cond, what):
    if not cond:
        raise VerificationError(what)


# ------------------------------------------------------------------ channel
def b2s(words):
    return list(struct.unpack("<8I", hashlib.blake2s(struct.pack(f"<{len(words)}I", *words)).digest()))


class Channel:
    def __init__(self):
        self.digest = [0] * 8
        self.n_draws = 0

    def mix(self, words):
        self.digest = b2s(self.digest + [int(w) for w in words])
        self.n_draws = 0

    def mix_qm31(self, vals):
        """Field-element vectors are mixed through their tree hash (tree_root)."""
        self.mix(tree_root([int(c) for v in vals for c in v]))

    def draw_words(self):
        out = b2s(self.digest + [self.n_draws, 0])
        self.n_draws += 1
        return out

    def draw_qm31(self):
        while True:
            w = self.draw_words()
            if all(x < 2 * P for x in w[:4]):
                return tuple(x - P if x >= P else x for x in w[:4])

    def draw_point(self):
        while True:
            t = self.draw_qm31()
            t2 = R.qm_mul(t, t)
            den = R.qm_add(R.qm(1), t2)
            if den == R.qm(0):
                continue
            di = R.qm_inv(den)
            return R.qm_mul(R.qm_sub(R.qm(1), t2), di), R.qm_mul(R.qm_add(t, t), di)

    def draw_queries(self, n, log_size):
        out = []
        while len(out) < n:
            for w in self.draw_words():
                if len(out) < n:
                    out.append(w & ((1 << log_size) - 1))
        return out


def tree_root(words):
    """16-word leaves (last zero-padded) hashed with the zero-state compression, then
    pairwise levels, an unpaired last node carried up unchanged."""
    nodes = []
    for i in range(0, len(words), 16):
        blk = list(words[i:i + 16])
        nodes.append(_compress([0] * 8, blk + [0] * (16 - len(blk))))
    while len(nodes) > 1:
        nodes = [_compress([0] * 8, nodes[i] + nodes[i + 1]) if i + 1 < len(nodes) else nodes[i]
                 for i in range(0, len(nodes), 2)]
    return nodes[0]


def pow_zeros(digest, nonce):
    h = b2s(list(digest) + [nonce & 0xFFFFFFFF, nonce >> 32])
    v = h[0] | (h[1] << 32)
    return 64 if v == 0 else (v & -v).bit_length() - 1


# ------------------------------------------------------------------ layout
class Layout:
    def __init__(self, C, log_n, Q, last_log):
        self.n_line = log_n + 1 - last_log
        off = 8
        self.roots = off
        off += 8 * (2 + self.n_line)
        self.samples = off
        off += 4 * (C + 4)
        self.last = off
        off += 4 << last_log
        self.nonce = off
        off += 2
        self.queries = off
        off += Q
        self.per_query = off
        o = 0
        self.q_trace_vals = o
        o += 2 * C
        self.q_trace_path = o
        o += 8 * log_n
        self.q_comp_vals = o
        o += 8
        self.q_comp_path = o
        o += 8 * (log_n + 1)
        self.q_line = []
        for j in range(1, self.n_line + 1):
            m = log_n + 2 - j
            self.q_line.append(o)
            o += 8 + 8 * (m - 1)
        self.query_words = o
        self.total = self.per_query + Q * self.query_words


# ------------------------------------------------------------------ merkle
def _compress(h, m):
    return [int(x) for x in O.blake2s_compress(np.array(h, np.uint32), np.array(m, np.uint32))]


def leaf_hash(values):
    h = [0] * 8
    for c in range(0, len(values), 16):
        chunk = list(values[c:c + 16])
        h = _compress(h, chunk + [0] * (16 - len(chunk)))
    return h


def node_hash(left, right):
    return _compress([0] * 8, list(left) + list(right))


def check_path(root, leaf_log, pos, vals_pos, vals_sib, siblings, what):
    """vals_* are the column values of rows pos and pos ^ 1; siblings[i] is the sibling at layer leaf_log-1-i."""
    hp, hs = leaf_hash(vals_pos), leaf_hash(vals_sib)
    node = node_hash(hp, hs) if pos % 2 == 0 else node_hash(hs, hp)
    idx = pos >> 1
    for k in range(leaf_log - 1, 0, -1):
        sib = siblings[leaf_log - 1 - k]
        node = node_hash(node, sib) if idx % 2 == 0 else node_hash(sib, node)
        idx >>= 1
    check(node == list(root), f"merkle path of {what}")


# ------------------------------------------------------------------ domains and formulas
def domain_point(log_size, row):
    """Point of CanonicCoset(log_size) at bit-reversed row `row`."""
    hi, hs, hl = R.canonic_half_coset(log_size)
    i = R.brev(row, log_size)
    half = 1 << hl
    if i < half:
        return R.point_at_index(hi + hs * i)
    x, y = R.point_at_index(hi + hs * (i - half))
    return x, (-y) % P


def conj(v):
    return (v[0], v[1], (-v[2]) % P, (-v[3]) % P)


def cm_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def cm_inv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    f = R.m31_inv(n)
    return (a[0] * f % P, (-a[1]) * f % P)


def quotient_at(p, z, samples, col_values, beta):
    """DEEP quotient of one batch (point z, sample j of column j) at the M31 point p."""
    zx, zy = z
    c = R.qm_sub(conj(zy), zy)
    num = R.qm(0)
    alpha = R.qm(1)
    for v, cv in zip(samples, col_values):
        alpha = R.qm_mul(alpha, beta)
        a = R.qm_sub(conj(v), v)
        b = R.qm_sub(R.qm_mul(v, c), R.qm_mul(a, zy))
        term = R.qm_sub(R.qm_scale(c, cv), R.qm_add(R.qm_scale(a, p[1]), b))
        num = R.qm_add(num, R.qm_mul(alpha, term))
    d = cm_sub(R.cm_mul(((zx[0] - p[0]) % P, zx[1]), (zy[2], zy[3])),
               R.cm_mul(((zy[0] - p[1]) % P, zy[1]), (zx[2], zx[3])))
    di = cm_inv(d)
    return R.qm_mul(num, (di[0], di[1], 0, 0))


def fold_pair(v_even, v_odd, t_inv, alpha):
    f0 = R.qm_add(v_even, v_odd)
    f1 = R.qm_scale(R.qm_sub(v_even, v_odd), t_inv)
    return R.qm_add(f0, R.qm_mul(alpha, f1))


def pair(pos, v_pos, v_sib):
    return (v_pos, v_sib) if pos % 2 == 0 else (v_sib, v_pos)


def line_coeffs(xs, vals):
    """Coefficients of the line polynomial through (xs[i], vals[i]) (bit-reversed line
    domain: xs[2i + 1] = -xs[2i]); index bit j selects the factor pi^j(x), so the
    coefficient at index d belongs to a term of degree d."""
    if len(vals) == 1:
        return [vals[0]]
    half = R.m31_inv(2)
    nx, f0, f1 = [], [], []
    for i in range(0, len(vals), 2):
        x = xs[i]
        check(xs[i + 1] == (-x) % P, "line domain pairing")
        f0.append(R.qm_scale(R.qm_add(vals[i], vals[i + 1]), half))
        f1.append(R.qm_scale(R.qm_sub(vals[i], vals[i + 1]), half * R.m31_inv(x) % P))
        nx.append((2 * x * x - 1) % P)
    c0, c1 = line_coeffs(nx, f0), line_coeffs(nx, f1)
    out = []
    for a, b in zip(c0, c1):
        out += [a, b]
    return out


def pi_n(x, k):
    for _ in range(k):
        x = R.qm_sub(R.qm_scale(R.qm_mul(x, x), 2), R.qm(1))
    return x


# ------------------------------------------------------------------ verifier
def verify(proof):
    """Raise VerificationError unless `proof` (uint32 words) is a valid proof; return the header."""
    w = [int(x) for x in np.asarray(proof, np.uint32)]
    magic, version, C, log_n, Q, pow_bits, last_log, n_line = w[:8]
    check(magic == MAGIC and version == VERSION, "header")
    L = Layout(C, log_n, Q, last_log)
    check(n_line == L.n_line and len(w) == L.total, "layout")

    def qm_at(off):
        return tuple(w[off:off + 4])

    def root(i):
        return w[L.roots + 8 * i:L.roots + 8 * i + 8]

    ch = Channel()
    ch.mix([C, log_n, Q, pow_bits, last_log, TAG])
    ch.mix(root(0))
    alpha = ch.draw_qm31()
    ch.mix(root(1))
    z = ch.draw_point()
    samples = [qm_at(L.samples + 4 * i) for i in range(C + 4)]
    ch.mix_qm31(samples)
    beta = ch.draw_qm31()
    a = [ch.draw_qm31()]
    for j in range(1, n_line + 1):
        ch.mix(root(1 + j))
        a.append(ch.draw_qm31())
    last = [qm_at(L.last + 4 * i) for i in range(1 << last_log)]
    ch.mix_qm31(last)
    nonce = w[L.nonce] | (w[L.nonce + 1] << 32)
    check(pow_zeros(ch.digest, nonce) >= pow_bits, "proof of work")
    ch.mix([nonce & 0xFFFFFFFF, nonce >> 32])
    queries = ch.draw_queries(Q, log_n + 2)
    check(queries == w[L.queries:L.queries + Q], "query positions")

    # OODS: E(z) Z(z) = sum_i (s_i^2 + s_{i+1}^2 - s_{i+2}) alpha^(C-3-i)
    s, e = samples[:C], samples[C:]
    ez = R.qm(0)
    for k in range(4):
        basis = tuple(1 if j == k else 0 for j in range(4))
        ez = R.qm_add(ez, R.qm_mul(e[k], basis))
    rhs = R.qm(0)
    for i in range(C - 2):
        cst = R.qm_sub(R.qm_add(R.qm_mul(s[i], s[i]), R.qm_mul(s[i + 1], s[i + 1])), s[i + 2])
        rhs = R.qm_add(rhs, R.qm_mul(cst, R.qm_pow(alpha, C - 3 - i)))
    check(R.qm_mul(ez, pi_n(z[0], log_n - 1)) == rhs, "OODS composition relation")

    # last layer: a line polynomial of degree < 2^(last_log - 1) (blowup 2)
    last_dom = log_n + 2 - n_line  # circle domain whose line domain holds the last layer
    xs = [domain_point(last_dom, 2 * i)[0] for i in range(1 << last_log)]
    coeffs = line_coeffs(xs, last)
    check(all(c == R.qm(0) for c in coeffs[1 << (last_log - 1):]), "last layer degree")

    for q, Ppos in enumerate(queries):
        base = L.per_query + q * L.query_words
        # trace rows R = P >> 1 and R ^ 1 on CanonicCoset(log_n + 1)
        Rp = Ppos >> 1
        tv0 = w[base + L.q_trace_vals:base + L.q_trace_vals + C]
        tv1 = w[base + L.q_trace_vals + C:base + L.q_trace_vals + 2 * C]
        tpath = [w[base + L.q_trace_path + 8 * i:base + L.q_trace_path + 8 * i + 8] for i in range(log_n)]
        check_path(root(0), log_n + 1, Rp, tv0, tv1, tpath, f"trace (query {q})")
        cv = [w[base + L.q_comp_vals + 4 * t:base + L.q_comp_vals + 4 * t + 4] for t in range(2)]
        cpath = [w[base + L.q_comp_path + 8 * i:base + L.q_comp_path + 8 * i + 8] for i in range(log_n + 1)]
        check_path(root(1), log_n + 2, Ppos, cv[0], cv[1], cpath, f"composition (query {q})")
        lines = []
        for j in range(1, n_line + 1):
            m = log_n + 2 - j
            o = base + L.q_line[j - 1]
            lv = [tuple(w[o:o + 4]), tuple(w[o + 4:o + 8])]
            lpath = [w[o + 8 + 8 * i:o + 16 + 8 * i] for i in range(m - 1)]
            check_path(root(1 + j), m, Ppos >> j, lv[0], lv[1], lpath, f"line layer {j} (query {q})")
            lines.append(lv)

        # quotients at the queried rows
        f2 = [quotient_at(domain_point(log_n + 2, r), z, e, vals, beta) for r, vals in ((Ppos, cv[0]), (Ppos ^ 1, cv[1]))]
        f1 = [quotient_at(domain_point(log_n + 1, r), z, s, vals, beta) for r, vals in ((Rp, tv0), (Rp ^ 1, tv1))]
        # circle fold of F2 into line 1
        ev, od = pair(Ppos, f2[0], f2[1])
        y = domain_point(log_n + 2, Ppos & ~1)[1]
        cur = fold_pair(ev, od, R.m31_inv(y), a[0])
        for j in range(1, n_line + 1):
            pj = Ppos >> j
            check(lines[j - 1][0] == cur, f"FRI layer {j} value (query {q})")
            ev, od = pair(pj, lines[j - 1][0], lines[j - 1][1])
            x = domain_point(log_n + 3 - j, 2 * (pj & ~1))[0]
            cur = fold_pair(ev, od, R.m31_inv(x), a[j])
            if j == 1:
                ev1, od1 = pair(Rp, f1[0], f1[1])
                y1 = domain_point(log_n + 1, Rp & ~1)[1]
                cur = R.qm_add(R.qm_mul(cur, R.qm_mul(a[1], a[1])), fold_pair(ev1, od1, R.m31_inv(y1), a[1]))
        check(last[Ppos >> (n_line + 1)] == cur, f"last layer value (query {q})")
    return {"n_cols": C, "log_n": log_n, "n_queries": Q, "pow_bits": pow_bits, "last_log": last_log,
            "n_line": n_line, "queries": queries}
