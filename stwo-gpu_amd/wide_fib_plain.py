"""Plain per-job wide-Fibonacci prover: the same proof as stwo_hip_wide_fib_prove, driven the
way stwo's prover drives the reference CudaBackend -- one synchronous call of the
reference-shaped ABI (include/stwo_hip.h part 1) per backend operation, Fiat-Shamir and
proof of work on the host, values for the decommitments copied back per row.

Per operation it follows the trait implementations of the reference crate:
interpolate_columns (poly.rs:57-82), extend + evaluate_polynomials (poly.rs:96-168),
MerkleOps::commit_on_layer per layer (blake2s.rs:19-97), the wide-Fibonacci composition
(examples/wide_fibonacci.rs:90-141), eval_at_point per column (poly.rs:84-94),
accumulate_quotients (quotient.rs:17-115), fold_circle_into_line / fold_line
(fri.rs:15-66); grind runs on the CPU as the reference's GrindOps does
(backend.rs:14-18).  The trace and the vanishing-polynomial inverses are generated on
the device (extension entry points) so both paths prove the same statement.

It is the baseline the proof server is measured against, and a second, independent
assembly of the protocol of csrc/prover.hip: its proofs must equal the device prover's
word for word.
"""
import hashlib
import struct

import numpy as np

import stwo_hip as S

P = S.P
MAGIC, VERSION, TAG = 0x50575453, 1, 0x57464942


# ------------------------------------------------------------------ QM31 on the host
def qm_mul(a, b):
    return S.qm_mul(a, b)


def qm_add(a, b):
    return tuple((x + y) % P for x, y in zip(a, b))


def qm_sub(a, b):
    return tuple((x - y) % P for x, y in zip(a, b))


# ------------------------------------------------------------------ channel (channel.hpp)
def _b2s(words):
    return list(struct.unpack("<8I", hashlib.blake2s(struct.pack(f"<{len(words)}I", *words)).digest()))


class Channel:
    def __init__(self):
        self.digest = [0] * 8
        self.n_draws = 0

    def mix(self, words):
        self.digest = _b2s(self.digest + [int(w) for w in words])
        self.n_draws = 0

    def draw_words(self):
        out = _b2s(self.digest + [self.n_draws, 0])
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
            t2 = qm_mul(t, t)
            den = qm_add((1, 0, 0, 0), t2)
            if den == (0, 0, 0, 0):
                continue
            di = S.qm_inv(den)
            return qm_mul(qm_sub((1, 0, 0, 0), t2), di), qm_mul(qm_add(t, t), di)

    def draw_queries(self, n, log_size):
        out = []
        while len(out) < n:
            for w in self.draw_words():
                if len(out) < n:
                    out.append(w & ((1 << log_size) - 1))
        return out


_SIGMA = [
    (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15), (14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3),
    (11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4), (7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8),
    (9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13), (2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9),
    (12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11), (13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10),
    (6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5), (10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)]
_IV = (0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19)
_M32 = 0xFFFFFFFF


def _f0(m):
    """BLAKE2s compression from the zero state, zero counter and flags (RFC 7693)."""
    v = [0] * 8 + list(_IV)

    def rot(x, c):
        return ((x >> c) | (x << (32 - c))) & _M32

    def g(a, b, c, d, x, y):
        v[a] = (v[a] + v[b] + x) & _M32
        v[d] = rot(v[d] ^ v[a], 16)
        v[c] = (v[c] + v[d]) & _M32
        v[b] = rot(v[b] ^ v[c], 12)
        v[a] = (v[a] + v[b] + y) & _M32
        v[d] = rot(v[d] ^ v[a], 8)
        v[c] = (v[c] + v[d]) & _M32
        v[b] = rot(v[b] ^ v[c], 7)
    for s in _SIGMA:
        g(0, 4, 8, 12, m[s[0]], m[s[1]])
        g(1, 5, 9, 13, m[s[2]], m[s[3]])
        g(2, 6, 10, 14, m[s[4]], m[s[5]])
        g(3, 7, 11, 15, m[s[6]], m[s[7]])
        g(0, 5, 10, 15, m[s[8]], m[s[9]])
        g(1, 6, 11, 12, m[s[10]], m[s[11]])
        g(2, 7, 8, 13, m[s[12]], m[s[13]])
        g(3, 4, 9, 14, m[s[14]], m[s[15]])
    return [v[i] ^ v[i + 8] for i in range(8)]


def tree_root(words):
    """Tree hash of a word sequence (csrc/blake2s.hpp): zero-padded 16-word leaves, pairwise
    levels, an unpaired last node carried up."""
    nodes = []
    for i in range(0, len(words), 16):
        blk = [int(w) for w in words[i:i + 16]]
        nodes.append(_f0(blk + [0] * (16 - len(blk))))
    while len(nodes) > 1:
        nodes = [_f0(nodes[i] + nodes[i + 1]) if i + 1 < len(nodes) else nodes[i] for i in range(0, len(nodes), 2)]
    return nodes[0]


def pow_ok(digest, nonce, bits):
    h = _b2s(list(digest) + [nonce & 0xFFFFFFFF, nonce >> 32])
    v = h[0] | (h[1] << 32)
    return bits == 0 or (v & ((1 << bits) - 1)) == 0


# ------------------------------------------------------------------ layout (prover.hip)
def layout(C_, log_n, Q, last_log):
    n_line = log_n + 1 - last_log
    L = {"n_line": n_line, "roots": 8}
    off = 8 + 8 * (2 + n_line)
    L["samples"] = off
    off += 4 * (C_ + 4)
    L["last"] = off
    off += 4 << last_log
    L["nonce"] = off
    off += 2
    L["queries"] = off
    off += Q
    L["per_query"] = off
    o = 0
    L["q_trace_vals"] = o
    o += 2 * C_
    L["q_trace_path"] = o
    o += 8 * log_n
    L["q_comp_vals"] = o
    o += 8
    L["q_comp_path"] = o
    o += 8 * (log_n + 1)
    L["q_line"] = []
    for j in range(1, n_line + 1):
        L["q_line"].append(o)
        o += 8 + 8 * (log_n + 2 - j - 1)
    L["query_words"] = o
    L["total"] = L["per_query"] + Q * o
    return L


# ------------------------------------------------------------------ helpers over the part-1 ABI
def _lib():
    return S.lib()


def _commit(columns_by_log, max_log):
    """MerkleProver::commit: commit_on_layer for log max_log .. 0; returns the layers."""
    layers = {}
    prev = None
    for k in range(max_log, -1, -1):
        prev = S.commit_on_layer(k, prev, columns_by_log.get(k, []))
        layers[k] = prev
    return layers


def _root(layers):
    return [int(x) for x in layers[0].to_numpy().reshape(-1)]


def _row(col, i):
    """One value of a device column (a 4-byte device-to-host copy)."""
    out = np.zeros(1, np.uint32)
    _lib().copy_uint32_t_vec_from_device_to_host(col.ptr + 4 * i, out.ctypes.data, 1)
    return int(out[0])


def _hash_row(layer, i):
    out = np.zeros(8, np.uint32)
    _lib().copy_blake_2s_hash_vec_from_device_to_host(layer.ptr + 32 * i, out.ctypes.data, 1)
    return out


def _lde(cols, log_n, tree):
    """interpolate_columns in place, then extend (zeroed buffer + copy) and evaluate_columns."""
    S.interpolate_columns(cols, tree)
    n = 1 << log_n
    ext = []
    for c in cols:
        e = S.BaseFieldVec.zeros(2 * n)
        _lib().copy_uint32_t_vec_from_device_to_device(c.ptr, e.ptr, n)
        ext.append(e)
    S.evaluate_columns(ext, tree)
    return ext


def prove(n_cols, log_n, n_queries, pow_bits, last_log, seed):
    C_ = n_cols
    n = 1 << log_n
    L = layout(C_, log_n, n_queries, last_log)
    root = S.canonic_half_coset(log_n + 2)
    tree = S.precompute_twiddles(root)

    ch = Channel()
    ch.mix([C_, log_n, n_queries, pow_bits, last_log, TAG])
    trace = [S.BaseFieldVec.uninitialized(n) for _ in range(C_)]
    S.check(_lib().stwo_hip_wide_fib_trace(S._ptr_array([c.ptr for c in trace]), C_, log_n, seed))
    S.synchronize()
    lde = _lde(trace, log_n, tree)  # trace now holds the coefficients
    layers_t = _commit({log_n + 1: lde}, log_n + 1)
    ch.mix(_root(layers_t))
    alpha = ch.draw_qm31()

    powers = S.SecureFieldVec.from_numpy(np.array([S.qm_pow(alpha, k) for k in range(C_ - 2)], np.uint32))
    dinv = S.BaseFieldVec.uninitialized(2 * n)
    S.check(_lib().stwo_hip_vanishing_inverse(dinv.ptr, log_n + 1, log_n))
    comp = S.SecureColumn.uninitialized(2 * n)
    S.wide_fib_constraint_quotients(comp, lde, powers, dinv)
    comp_lde = _lde(comp.columns, log_n + 1, tree)  # comp.columns now hold the coefficients
    layers_q = _commit({log_n + 2: comp_lde}, log_n + 2)
    ch.mix(_root(layers_q))
    zx, zy = ch.draw_point()
    samples = [S.eval_at_point(c, zx, zy) for c in trace] + [S.eval_at_point(c, zx, zy) for c in comp.columns]
    ch.mix(tree_root([w for v in samples for w in v]))
    beta = ch.draw_qm31()
    alphas = [ch.draw_qm31()]

    f1 = S.accumulate_quotients(S.canonic_half_coset(log_n + 1), lde, beta,
                                [((zx, zy), [(i, samples[i]) for i in range(C_)])])
    f2 = S.accumulate_quotients(S.canonic_half_coset(log_n + 2), comp_lde, beta,
                                [((zx, zy), [(k, samples[C_ + k]) for k in range(4)])])
    line = S.SecureColumn.zeros(2 * n)
    S.fold_circle_into_line(line, f2, alphas[0], tree)
    lines, trees_l = [line], []
    for j in range(1, L["n_line"] + 1):
        m = log_n + 2 - j
        lt = _commit({m: lines[j - 1].columns}, m)
        trees_l.append(lt)
        ch.mix(_root(lt))
        alphas.append(ch.draw_qm31())
        nxt = S.fold_line(lines[j - 1], alphas[j], tree)
        if j == 1:
            S.fold_circle_into_line(nxt, f1, alphas[1], tree)
        lines.append(nxt)
    last = lines[-1].to_numpy()
    ch.mix(tree_root([int(last[k][i]) for i in range(1 << last_log) for k in range(4)]))
    nonce = 0
    while not pow_ok(ch.digest, nonce, pow_bits):
        nonce += 1
    ch.mix([nonce & 0xFFFFFFFF, nonce >> 32])
    queries = ch.draw_queries(n_queries, log_n + 2)

    w = np.zeros(L["total"], np.uint32)
    w[:8] = [MAGIC, VERSION, C_, log_n, n_queries, pow_bits, last_log, L["n_line"]]
    roots = [_root(layers_t), _root(layers_q)] + [_root(t) for t in trees_l]
    for i, r in enumerate(roots):
        w[8 + 8 * i:16 + 8 * i] = r
    w[L["samples"]:L["samples"] + 4 * (C_ + 4)] = np.array(samples, np.uint32).reshape(-1)
    w[L["last"]:L["last"] + (4 << last_log)] = last.T.reshape(-1)
    w[L["nonce"]:L["nonce"] + 2] = [nonce & 0xFFFFFFFF, nonce >> 32]
    w[L["queries"]:L["queries"] + n_queries] = queries

    def path(layers, leaf_log, pos):
        return np.concatenate([_hash_row(layers[k], (pos >> (leaf_log - k)) ^ 1) for k in range(leaf_log - 1, 0, -1)])

    for q, Pq in enumerate(queries):
        b = L["per_query"] + q * L["query_words"]
        R = Pq >> 1
        w[b + L["q_trace_vals"]:b + L["q_trace_vals"] + C_] = [_row(c, R) for c in lde]
        w[b + L["q_trace_vals"] + C_:b + L["q_trace_vals"] + 2 * C_] = [_row(c, R ^ 1) for c in lde]
        w[b + L["q_trace_path"]:b + L["q_trace_path"] + 8 * log_n] = path(layers_t, log_n + 1, R)
        w[b + L["q_comp_vals"]:b + L["q_comp_vals"] + 4] = [_row(c, Pq) for c in comp_lde]
        w[b + L["q_comp_vals"] + 4:b + L["q_comp_vals"] + 8] = [_row(c, Pq ^ 1) for c in comp_lde]
        w[b + L["q_comp_path"]:b + L["q_comp_path"] + 8 * (log_n + 1)] = path(layers_q, log_n + 2, Pq)
        for j in range(1, L["n_line"] + 1):
            m = log_n + 2 - j
            o = b + L["q_line"][j - 1]
            pj = Pq >> j
            w[o:o + 4] = [_row(c, pj) for c in lines[j - 1].columns]
            w[o + 4:o + 8] = [_row(c, pj ^ 1) for c in lines[j - 1].columns]
            w[o + 8:o + 8 + 8 * (m - 1)] = path(trees_l[j - 1], m, pj)
    return w
