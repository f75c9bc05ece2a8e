"""stwo_hip_eval_at_points_columns_noalloc (no device allocation, fixed per-device scratch)
against the allocating stwo_hip_eval_at_points_columns, bit for bit, on two calls in a row
(the second smaller, so stale scratch contents from the first would show), plus the size
guard."""
import random

import numpy as np
import pytest

import oracle as O
import stwo_hip as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    S.check(S.lib().stwo_hip_init(0))
    yield


def _point(rng):
    return S.random_circle_point_qm31(np.random.default_rng(rng.getrandbits(32)))


def _cols(n_cols, log_n, seed):
    rng = np.random.default_rng(seed)
    host = [rng.integers(0, S.P, 1 << log_n, dtype=np.uint32) for _ in range(n_cols)]
    return host, [S.BaseFieldVec.from_numpy(h) for h in host]


def test_noalloc_matches_allocating_entry_point_twice():
    rng = random.Random(0x5EED0005)
    # call 1: 64 columns of log 16 at 3 points (matrix-vector path)
    host1, cols1 = _cols(64, 16, 11)
    pts1 = [_point(rng) for _ in range(3)]
    got1 = S.eval_at_points_columns(cols1, pts1, noalloc=True)
    want1 = S.eval_at_points_columns(cols1, pts1)
    assert got1 == want1
    for c in (0, 63):
        assert tuple(got1[2][c]) == O.eval_at_point(host1[c], *pts1[2])
    # call 2, smaller: 5 columns of log 9 at 2 points (LDS fold path); the scratch still holds call 1
    host2, cols2 = _cols(5, 9, 12)
    pts2 = [_point(rng) for _ in range(2)]
    got2 = S.eval_at_points_columns(cols2, pts2, noalloc=True)
    want2 = S.eval_at_points_columns(cols2, pts2)
    assert got2 == want2
    for c in range(5):
        assert tuple(got2[1][c]) == O.eval_at_point(host2[c], *pts2[1])


def test_noalloc_rejects_requests_beyond_its_scratch():
    L = S.lib()
    col = S.BaseFieldVec.from_numpy(np.arange(1 << 10, dtype=np.uint32))
    # 512 columns of log 24 need about 33.6 MB of partial sums: refused before any copy or launch
    table = S._ptr_array([col.ptr] * 512)
    pts = (S.PointQM31 * 1)(S.PointQM31(S.qm31((1, 0, 0, 0)), S.qm31((0, 1, 0, 0))))
    out = (S.QM31 * 512)()
    st = L.stwo_hip_eval_at_points_columns_noalloc(table, 512, 24, S.C.cast(pts, S.vp), 1, S.C.cast(out, S.vp))
    assert st == 1 and "scratch" in S.last_error()
    L.stwo_hip_clear_error()
    # the library and the fixed scratch still work
    rng = random.Random(3)
    host, cols = _cols(3, 12, 13)
    p = [_point(rng)]
    assert S.eval_at_points_columns(cols, p, noalloc=True) == S.eval_at_points_columns(cols, p)
