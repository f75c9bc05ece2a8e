"""CPU checks of the host-side pieces of the reference-shaped plain prover
(stwo-gpu_amd/wide_fib_plain.py): the library's host tree hash of the channel against the
verifier's independent restatement, and the vectorised QM31 powers against schoolbook
QM31 arithmetic.  Host code only: no device call."""
import numpy as np
import pytest

import refmath as R
import wf_verifier as V
import wide_fib_plain as W


@pytest.mark.parametrize("n_words", [1, 15, 16, 17, 33, 100, 2064])
def test_host_tree_hash_matches_restatement(n_words):
    words = [int(x) for x in np.random.default_rng(n_words).integers(0, 1 << 32, n_words, dtype=np.uint64)]
    assert W.tree_root(words) == list(V.tree_root(words))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 510])
def test_qm31_powers_vectorised(n):
    rng = np.random.default_rng(n)
    alpha = tuple(int(x) for x in rng.integers(0, R.P, 4))
    got = W.qm31_powers(alpha, n)
    for k in sorted({0, min(1, n - 1), n // 2, n - 1}):
        assert tuple(int(x) for x in got[k]) == tuple(R.qm_pow(alpha, k)), f"power {k}"
