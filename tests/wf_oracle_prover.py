"""CPU prover for the wide-Fibonacci proofs, assembled from the oracle's operations
(test infrastructure: the word-for-word reference the HIP prover is compared with at
small sizes).  Same protocol and layout as stwo-gpu_amd/csrc/prover.hip; every step is
an oracle call (oracle/stwo_oracle.c) and the channel is wf_verifier.Channel.
"""
import numpy as np

import oracle as O
import refmath as R
import wf_verifier as V

P = R.P


def trace(n_cols, log_n, seed):
    n = 1 << log_n
    cols = [O.fill_uniform_m31(n, seed, 0), O.fill_uniform_m31(n, seed, n)]
    for _ in range(2, n_cols):
        a, b = cols[-2].astype(np.uint64), cols[-1].astype(np.uint64)
        cols.append(((a * a + b * b) % P).astype(np.uint32))
    return cols


def vanishing_inverse(log_ext, log_n):
    out = np.zeros(1 << log_ext, np.uint32)
    for r in range(1 << log_ext):
        x = V.domain_point(log_ext, r)[0]
        for _ in range(log_n - 1):
            x = (2 * x * x - 1) % P
        out[r] = R.m31_inv(x)
    return out


def prove(n_cols, log_n, n_queries, pow_bits, last_log, seed):
    C = n_cols
    L = V.Layout(C, log_n, n_queries, last_log)
    hi = {m: R.canonic_half_coset(m) for m in range(1, log_n + 3)}
    root = hi[log_n + 2]
    tw, itw = O.precompute_twiddles(root[0], root[1], root[2])

    ch = V.Channel()
    ch.mix([C, log_n, n_queries, pow_bits, last_log, V.TAG])
    tr = trace(C, log_n, seed)
    coeffs, lde = O.lde_columns(tr, 1, tw, itw)
    tree_t = O.merkle_commit(lde)
    ch.mix(tree_t[0])
    alpha = ch.draw_qm31()

    powers = np.array([R.qm_pow(alpha, k) for k in range(C - 2)], np.uint32)
    comp = O.wide_fib_constraint_quotients(lde, powers, vanishing_inverse(log_n + 1, log_n))
    comp_coeffs, comp_lde = O.lde_columns(list(comp), 1, tw, itw)
    tree_q = O.merkle_commit(comp_lde)
    ch.mix(tree_q[0])
    zx, zy = ch.draw_point()
    samples = [O.eval_at_point(c, zx, zy) for c in coeffs] + [O.eval_at_point(c, zx, zy) for c in comp_coeffs]
    ch.mix_qm31(samples)
    beta = ch.draw_qm31()
    alphas = [ch.draw_qm31()]

    h1, h2 = hi[log_n + 1], hi[log_n + 2]
    f1 = O.accumulate_quotients(h1[0], h1[1], log_n + 1, lde, beta, [((zx, zy), [(i, samples[i]) for i in range(C)])])
    f2 = O.accumulate_quotients(h2[0], h2[1], log_n + 2, comp_lde, beta,
                                [((zx, zy), [(k, samples[C + k]) for k in range(4)])])
    lines = [O.fold_circle_into_line(np.zeros((4, 2 << log_n), np.uint32), f2, h2[0], h2[1], alphas[0])]
    trees_l = []
    for j in range(1, L.n_line + 1):
        m = log_n + 2 - j
        tree = O.merkle_commit(list(lines[j - 1]))
        trees_l.append(tree)
        ch.mix(tree[0])
        alphas.append(ch.draw_qm31())
        dom = hi[m + 1]  # the line domain of CanonicCoset(m + 1) is its half coset's x
        nxt = O.fold_line(lines[j - 1], dom[0], dom[1], alphas[j])
        if j == 1:
            nxt = O.fold_circle_into_line(nxt, f1, h1[0], h1[1], alphas[1])
        lines.append(nxt)
    last = lines[-1]
    ch.mix_qm31([tuple(int(last[k][i]) for k in range(4)) for i in range(1 << last_log)])
    nonce = 0
    while V.pow_zeros(ch.digest, nonce) < pow_bits:
        nonce += 1
    ch.mix([nonce & 0xFFFFFFFF, nonce >> 32])
    queries = ch.draw_queries(n_queries, log_n + 2)

    w = np.zeros(L.total, np.uint32)
    w[:8] = [V.MAGIC, V.VERSION, C, log_n, n_queries, pow_bits, last_log, L.n_line]
    roots = [tree_t[0], tree_q[0]] + [t[0] for t in trees_l]
    for i, r in enumerate(roots):
        w[L.roots + 8 * i:L.roots + 8 * i + 8] = r
    w[L.samples:L.samples + 4 * (C + 4)] = np.array(samples, np.uint32).reshape(-1)
    w[L.last:L.last + (4 << last_log)] = last.T.reshape(-1)
    w[L.nonce:L.nonce + 2] = [nonce & 0xFFFFFFFF, nonce >> 32]
    w[L.queries:L.queries + n_queries] = queries

    def path(tree, leaf_log, pos):
        out = []
        for k in range(leaf_log - 1, 0, -1):
            idx = (pos >> (leaf_log - k)) ^ 1
            out.append(tree[k].reshape(-1, 8)[idx])
        return np.array(out, np.uint32).reshape(-1)

    for q, Pq in enumerate(queries):
        b = L.per_query + q * L.query_words
        Rq = Pq >> 1
        w[b + L.q_trace_vals:b + L.q_trace_vals + C] = [c[Rq] for c in lde]
        w[b + L.q_trace_vals + C:b + L.q_trace_vals + 2 * C] = [c[Rq ^ 1] for c in lde]
        w[b + L.q_trace_path:b + L.q_trace_path + 8 * log_n] = path(tree_t, log_n + 1, Rq)
        w[b + L.q_comp_vals:b + L.q_comp_vals + 4] = [c[Pq] for c in comp_lde]
        w[b + L.q_comp_vals + 4:b + L.q_comp_vals + 8] = [c[Pq ^ 1] for c in comp_lde]
        w[b + L.q_comp_path:b + L.q_comp_path + 8 * (log_n + 1)] = path(tree_q, log_n + 2, Pq)
        for j in range(1, L.n_line + 1):
            m = log_n + 2 - j
            o = b + L.q_line[j - 1]
            pj = Pq >> j
            w[o:o + 4] = lines[j - 1][:, pj]
            w[o + 4:o + 8] = lines[j - 1][:, pj ^ 1]
            w[o + 8:o + 8 + 8 * (m - 1)] = path(trees_l[j - 1], m, pj)
    return w
