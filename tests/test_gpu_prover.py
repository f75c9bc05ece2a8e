"""GPU wide-Fibonacci prover: proofs equal the oracle-assembled proofs word for word,
pass the verifier, and the proof server returns the same proof for a job alone and in
a mix of concurrent jobs."""
import ctypes

import numpy as np
import pytest

import stwo_hip as S
import wf_oracle_prover as OP
import wf_verifier as V

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.check(S.lib().stwo_hip_init(0))
    yield


@pytest.mark.parametrize("C,log_n,Q,pow_bits,last_log,seed", [
    (3, 4, 1, 0, 2, 1), (5, 5, 3, 4, 2, 1234), (16, 7, 3, 6, 2, 7), (40, 6, 3, 6, 3, 2), (17, 8, 5, 8, 4, 3),
    (33, 9, 2, 3, 5, 5), (64, 12, 3, 10, 2, 11), (100, 14, 3, 10, 3, 12)])
def test_proof_matches_oracle(C, log_n, Q, pow_bits, last_log, seed):
    p = S.wide_fib_params(C, log_n, Q, pow_bits, last_log)
    got = S.wide_fib_prove(p, seed)
    want = OP.prove(C, log_n, Q, pow_bits, last_log, seed)
    assert got.size == want.size
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first differing word {diff[:8]}"
    V.verify(got)


def test_large_proof_verifies():
    p = S.wide_fib_params(256, 18, 3, 12, 2)
    w = S.wide_fib_prove(p, 99)
    info = V.verify(w)
    assert info["n_cols"] == 256
    # deterministic
    assert (S.wide_fib_prove(p, 99) == w).all()


def test_largest_serving_proof_verifies():
    # the top of the serving range: 512 columns of 2^20 rows
    p = S.wide_fib_params(512, 20, 3, 10, 2)
    info = V.verify(S.wide_fib_prove(p, 4242))
    assert info["n_cols"] == 512


@pytest.mark.parametrize("C,log_n,Q,pow_bits,last_log,k", [
    (3, 4, 1, 0, 2, 2), (16, 7, 3, 6, 2, 5), (40, 6, 3, 6, 3, 16), (64, 12, 3, 10, 2, 8), (100, 14, 3, 10, 3, 3),
    (32, 15, 3, 8, 2, 4)])
def test_batch_proofs_match_alone(C, log_n, Q, pow_bits, last_log, k):
    # same-phase batching: every launch covers the k jobs; each proof equals the job's proof alone
    p = S.wide_fib_params(C, log_n, Q, pow_bits, last_log)
    seeds = [7000 + 31 * i for i in range(k)]
    got = S.wide_fib_prove_batch(p, seeds)
    for i, seed in enumerate(seeds):
        alone = S.wide_fib_prove(p, seed)
        diff = np.nonzero(got[i] != alone)[0]
        assert diff.size == 0, f"job {i}: first differing word {diff[:8]}"
    assert not (got[0] == got[-1]).all() or k == 1
    V.verify(got[-1])


def test_batch_proof_matches_oracle():
    p = S.wide_fib_params(17, 8, 5, 8, 4)
    got = S.wide_fib_prove_batch(p, [3, 4, 5])
    for i, seed in enumerate([3, 4, 5]):
        assert (got[i] == OP.prove(17, 8, 5, 8, 4, seed)).all(), f"job {i}"


def test_server_batches_same_shape_jobs():
    # a burst of small jobs of three shapes: batched admission proves same-shape jobs together,
    # and every result is the job's proof alone
    shapes = [(32, 12), (64, 13), (8, 14)]
    jobs = [(S.wide_fib_params(C, ln, 3, 8, 2), 2000 + i) for i in range(24) for C, ln in [shapes[i % 3]]]
    alone = {seed: S.wide_fib_prove(p, seed) for p, seed in jobs}
    srv = S.ProofServer(16 << 30, 4, 0, max_batch=8, max_batch_work=1 << 24)
    ids = {srv.submit(p, seed): seed for p, seed in jobs}
    got, spans = {}, {}
    while len(got) < len(jobs):
        r = srv.wait(5_000_000)
        assert r is not None, "server stalled"
        jid, st, proof, times = r
        assert st == 0
        got[ids[jid]] = proof
        spans[ids[jid]] = (times[1], times[2])
    srv.close()
    for seed, proof in got.items():
        assert (proof == alone[seed]).all(), f"job with seed {seed} differs from its proof alone"
    # some jobs were proved in one batch: they share their start and end times
    assert len(set(spans.values())) < len(jobs)


@pytest.mark.parametrize("bits", [0, 1, 8, 14])
def test_grind_smallest_nonce(bits):
    digest = np.arange(8, dtype=np.uint32) * 0x01010101 + bits
    got = S.grind(digest, bits)
    n = 0
    while V.pow_zeros([int(x) for x in digest], n) < bits:
        n += 1
    assert got == n


def test_server_results_match_alone():
    rng = np.random.default_rng(5)
    jobs = []
    for i in range(14):
        C = int(rng.choice([8, 32, 64, 128]))
        ln = int(rng.integers(8, 15))
        jobs.append((S.wide_fib_params(C, ln, 3, 8, 2), 1000 + i))
    alone = {seed: S.wide_fib_prove(p, seed) for p, seed in jobs}
    srv = S.ProofServer(8 << 30, 4)
    ids = {srv.submit(p, seed): seed for p, seed in jobs}
    got = {}
    while len(got) < len(jobs):
        r = srv.wait(2_000_000)
        assert r is not None, "server stalled"
        jid, st, proof, times = r
        assert st == 0
        assert times[0] <= times[1] <= times[2]
        got[ids[jid]] = proof
    srv.close()
    for seed, proof in got.items():
        assert (proof == alone[seed]).all(), f"job with seed {seed} differs from its proof alone"
    V.verify(got[1000])


def test_server_admits_by_memory_budget():
    p = S.wide_fib_params(64, 12, 3, 4, 2)
    _, need = S.wide_fib_proof_size(p)
    srv = S.ProofServer(need + need // 2, 4)  # one job at a time fits
    seeds = [7, 8, 9]
    ids = {srv.submit(p, s): s for s in seeds}
    spans = []
    for _ in seeds:
        jid, st, proof, times = srv.wait(2_000_000)
        assert st == 0
        spans.append(times[1:])
    srv.close()
    spans.sort()
    for a, b in zip(spans, spans[1:]):
        assert b[0] >= a[1] - 1e-6, "two jobs ran together although only one fits the budget"


def test_servers_back_to_back():
    """Servers opened and closed one after another in one process: buffers cached by the
    pool and staging events from an earlier server's jobs are reused by the next one."""
    shapes = [(32, 12), (64, 13), (128, 12), (32, 14)]
    want = {i: S.wide_fib_prove(S.wide_fib_params(C, ln, 3, 6, 2), 50 + i) for i, (C, ln) in enumerate(shapes)}
    for round_ in range(4):
        srv = S.ProofServer(8 << 30, 4 + round_ % 2 * 4)
        ids = {srv.submit(S.wide_fib_params(C, ln, 3, 6, 2), 50 + i): i for i, (C, ln) in enumerate(shapes)}
        for _ in shapes:
            r = srv.wait(5_000_000)
            assert r is not None, "server stalled"
            jid, st, proof, _ = r
            assert st == 0, S.last_error()
            assert (proof == want[ids[jid]]).all()
        srv.close()


def test_server_admits_by_work_limit():
    p = S.wide_fib_params(64, 12, 3, 4, 2)
    work = 64 << 12
    srv = S.ProofServer(8 << 30, 4, work + work // 2)  # one job's trace elements at a time
    seeds = [17, 18, 19]
    ids = {srv.submit(p, s): s for s in seeds}
    spans, proofs = [], {}
    for _ in seeds:
        jid, st, proof, times = srv.wait(2_000_000)
        assert st == 0
        spans.append(times[1:3])
        proofs[ids[jid]] = proof
    srv.close()
    spans.sort()
    for a, b in zip(spans, spans[1:]):
        assert b[0] >= a[1] - 1e-6, "two jobs ran together although the work limit admits one"
    assert (proofs[17] == S.wide_fib_prove(p, 17)).all()


@pytest.mark.parametrize("C,log_n,Q,pow_bits,last_log,seed", [(3, 4, 1, 0, 2, 1), (40, 6, 3, 6, 3, 2),
                                                              (33, 11, 3, 8, 2, 21), (64, 13, 2, 10, 4, 22)])
def test_plain_path_matches_device_prover(C, log_n, Q, pow_bits, last_log, seed):
    """The reference-shaped per-operation path (synchronous part-1 ABI, host channel) and the
    device prover produce the same proof."""
    import wide_fib_plain
    want = S.wide_fib_prove(S.wide_fib_params(C, log_n, Q, pow_bits, last_log), seed)
    got = wide_fib_plain.prove(C, log_n, Q, pow_bits, last_log, seed)
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first differing word {diff[:8]}"


@pytest.mark.parametrize("C,log_n,Q,pow_bits,last_log", [
    (2, 8, 3, 4, 2), (4097, 8, 3, 4, 2), (8, 3, 3, 4, 2), (8, 27, 3, 4, 2), (8, 8, 0, 4, 2), (8, 8, 65, 4, 2),
    (8, 8, 3, 33, 2), (8, 8, 3, 4, 1), (8, 8, 3, 4, 11), (8, 4, 3, 4, 5)])
def test_invalid_statements_rejected(C, log_n, Q, pow_bits, last_log):
    """Out-of-range parameters are reported as errors before any device work, and the
    library stays usable."""
    p = S.wide_fib_params(C, log_n, Q, pow_bits, last_log)
    with pytest.raises(S.StwoHipError):
        S.wide_fib_proof_size(p)
    with pytest.raises(S.StwoHipError):  # the prove entry point validates before its buffer
        S.check(S.lib().stwo_hip_wide_fib_prove(ctypes.byref(p), 1, None, 0))
    w = S.wide_fib_prove(S.wide_fib_params(3, 4, 1, 0, 2), 1)
    V.verify(w)


@pytest.mark.parametrize("slots", [0, 17])
def test_server_slot_count_rejected(slots):
    with pytest.raises(S.StwoHipError):
        S.ProofServer(1 << 30, slots)


def test_grind_bits_rejected():
    with pytest.raises(S.StwoHipError):
        S.grind(np.zeros(8, np.uint32), 41)
    assert S.grind(np.zeros(8, np.uint32), 0) == 0
