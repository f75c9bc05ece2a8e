"""Plain per-job wide-Fibonacci prover: the same proof as stwo_hip_wide_fib_prove, driven the
way stwo's prover drives the reference CudaBackend -- one synchronous call of the
reference-shaped ABI (include/stwo_hip.h part 1) per backend operation, the Fiat-Shamir
channel on the host (hashlib BLAKE2s and the library's host tree hash, in place of stwo's
Rust channel), the decommitted values and authentication paths gathered by one call per
tree.

Per operation it follows the trait implementations of the reference crate:
interpolate_columns (poly.rs:57-82), extend + evaluate_polynomials (poly.rs:96-168),
MerkleOps::commit_on_layer per layer (blake2s.rs:19-97), the wide-Fibonacci composition
(examples/wide_fibonacci.rs:90-141), eval_at_point per column (poly.rs:84-94),
accumulate_quotients (quotient.rs:17-115), fold_circle_into_line / fold_line
(fri.rs:15-66).  Proof of work uses the library's grind (the reference delegates it to
the CPU SimdBackend, backend.rs:14-18; a Python nonce loop would measure the
interpreter).  The trace and the vanishing-polynomial inverses are generated on the device
(extension entry points) so both paths prove the same statement.

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


def tree_root(words):
    """Tree hash of a word sequence (csrc/blake2s.hpp), computed on the host by the library."""
    w = np.ascontiguousarray(np.array(words, dtype=np.uint64).astype(np.uint32))
    out = np.zeros(8, np.uint32)
    S.check(S.lib().stwo_hip_host_tree_hash(w.ctypes.data_as(S.u32p), w.size, out.ctypes.data_as(S.u32p)))
    return [int(x) for x in out]


# ------------------------------------------------------------------ QM31 powers, vectorised
def _m31_mul(a, b):
    return (a.astype(np.uint64) * b.astype(np.uint64)) % P


def _qm_mul_vec(x, y):
    """Row-wise QM31 products of (n, 4) uint64 arrays: (a + b i) + (c + d i) u, u^2 = 2 + i."""
    a, b, c, d = (x[:, k] for k in range(4))
    e, f, g, h = (y[:, k] for k in range(4))
    re0 = (_m31_mul(a, e) + P * P - _m31_mul(b, f)) % P  # (a + bi)(e + fi)
    im0 = (_m31_mul(a, f) + _m31_mul(b, e)) % P
    re1 = (_m31_mul(c, g) + P * P - _m31_mul(d, h)) % P  # (c + di)(g + hi)
    im1 = (_m31_mul(c, h) + _m31_mul(d, g)) % P
    re_r = (2 * re1 + P - im1) % P  # (2 + i)(re1 + im1 i)
    im_r = (re1 + 2 * im1) % P
    re2 = (_m31_mul(a, g) + P * P - _m31_mul(b, h) + _m31_mul(c, e) + P * P - _m31_mul(d, f)) % P
    im2 = (_m31_mul(a, h) + _m31_mul(b, g) + _m31_mul(c, f) + _m31_mul(d, e)) % P
    return np.stack([(re0 + re_r) % P, (im0 + im_r) % P, re2, im2], axis=1)


def qm31_powers(alpha, n):
    """alpha^0 .. alpha^(n-1) as an (n, 4) uint32 array (doubling blocks, one vector product each)."""
    out = np.zeros((max(n, 1), 4), np.uint64)
    out[0] = (1, 0, 0, 0)
    step = np.array([alpha], np.uint64)  # alpha^(2^k)
    have = 1
    while have < n:
        take = min(have, n - have)
        out[have:have + take] = _qm_mul_vec(out[:take], np.repeat(step, take, axis=0))
        have += take
        step = _qm_mul_vec(step, step)
    return out[:n].astype(np.uint32)


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


def _rows(cols, rows):
    """values[i, c] = cols[c][rows[i]], one gather call (Column::at for the decommitment)."""
    return S.gather_rows(cols, rows)


def _paths(layers, leaf_log, leaves):
    """Siblings of layers leaf_log - 1 .. 1 on the paths of the leaves, one call per tree."""
    full = S.merkle_auth_paths([layers[k] for k in range(leaf_log + 1)], leaves)
    return full[:, 1:, :].reshape(len(leaves), -1)


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

    powers = S.SecureFieldVec.from_numpy(qm31_powers(alpha, C_ - 2))
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
    nonce = S.grind(np.array(ch.digest, np.uint32), pow_bits)
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

    # decommitments: one gather per column set and one path gather per tree, for every query
    qs = np.array(queries, np.uint32)
    tr = _rows(lde, np.concatenate([qs >> 1, (qs >> 1) ^ 1]))
    tp = _paths(layers_t, log_n + 1, qs >> 1)
    cr = _rows(comp_lde, np.concatenate([qs, qs ^ 1]))
    cp = _paths(layers_q, log_n + 2, qs)
    lr, lp = [], []
    for j in range(1, L["n_line"] + 1):
        m = log_n + 2 - j
        pj = qs >> j
        lr.append(_rows(lines[j - 1].columns, np.concatenate([pj, pj ^ 1])))
        lp.append(_paths(trees_l[j - 1], m, pj))
    Qn = len(queries)
    for q in range(Qn):
        b = L["per_query"] + q * L["query_words"]
        w[b + L["q_trace_vals"]:b + L["q_trace_vals"] + C_] = tr[q]
        w[b + L["q_trace_vals"] + C_:b + L["q_trace_vals"] + 2 * C_] = tr[Qn + q]
        w[b + L["q_trace_path"]:b + L["q_trace_path"] + 8 * log_n] = tp[q]
        w[b + L["q_comp_vals"]:b + L["q_comp_vals"] + 4] = cr[q]
        w[b + L["q_comp_vals"] + 4:b + L["q_comp_vals"] + 8] = cr[Qn + q]
        w[b + L["q_comp_path"]:b + L["q_comp_path"] + 8 * (log_n + 1)] = cp[q]
        for j in range(1, L["n_line"] + 1):
            m = log_n + 2 - j
            o = b + L["q_line"][j - 1]
            w[o:o + 4] = lr[j - 1][q]
            w[o + 4:o + 8] = lr[j - 1][Qn + q]
            w[o + 8:o + 8 + 8 * (m - 1)] = lp[j - 1][q]
    return w
