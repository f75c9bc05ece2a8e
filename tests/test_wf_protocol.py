"""CPU checks of the wide-Fibonacci proof protocol: the oracle-assembled prover's proofs
pass the verifier, and corrupting any part of a proof makes it fail."""
import numpy as np
import pytest

import wf_oracle_prover as OP
import wf_verifier as V


@pytest.mark.parametrize("C,log_n,Q,pow_bits,last_log,seed", [
    (3, 4, 1, 0, 2, 1), (5, 5, 3, 4, 2, 1234), (40, 6, 3, 6, 3, 2), (17, 8, 5, 8, 4, 3), (33, 9, 2, 3, 5, 5)])
def test_oracle_proofs_verify(C, log_n, Q, pow_bits, last_log, seed):
    w = OP.prove(C, log_n, Q, pow_bits, last_log, seed)
    info = V.verify(w)
    assert info["n_cols"] == C and info["log_n"] == log_n and len(info["queries"]) == Q


def test_every_section_is_bound():
    w = OP.prove(20, 6, 3, 5, 2, 77)
    L = V.Layout(20, 6, 3, 2)
    offsets = [0, L.roots, L.roots + 8, L.roots + 16, L.samples, L.samples + 4 * 20, L.last, L.nonce, L.queries]
    qb = L.per_query
    offsets += [qb + L.q_trace_vals, qb + L.q_trace_vals + 20, qb + L.q_trace_path, qb + L.q_comp_vals,
                qb + L.q_comp_vals + 4, qb + L.q_comp_path] + [qb + o for o in L.q_line] + [qb + o + 8 for o in L.q_line]
    for off in offsets + list(range(0, len(w), 13)):
        bad = w.copy()
        bad[off] ^= 1 << (off % 31)
        with pytest.raises(V.VerificationError):
            V.verify(bad)


def test_wrong_statement_rejected():
    w = OP.prove(8, 5, 2, 3, 2, 9)
    bad = w.copy()
    bad[2] = 9  # claims 9 columns
    with pytest.raises(V.VerificationError):
        V.verify(bad)


def test_pow_rule_matches_definition():
    ch = V.Channel()
    ch.mix([1, 2, 3])
    n = 0
    while V.pow_zeros(ch.digest, n) < 6:
        n += 1
    assert all(V.pow_zeros(ch.digest, k) < 6 for k in range(n))
    h = V.b2s(ch.digest + [n, 0])
    assert (h[0] & 63) == 0
