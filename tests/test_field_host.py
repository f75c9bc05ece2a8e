"""The device field arithmetic (stwo-gpu_amd/csrc/field.hpp: the FFT butterflies' mul_dbl
with its doubled twiddle, reduce62/63, qmul, and the FRI folds' qmat_of, fold_mat,
fold_mat_acc, reduce64, sub3) built for the host and checked against the pure-Python
field model in refmath.py, at the extreme values where its 64-bit sums are tightest (every
coordinate P - 1) and at random values.

fold_mat computes (a + b) + alpha t (a - b) with the product by alpha written as a 4 x 4 matrix
over M31 and four 64-bit multiply-adds per coordinate: the matrix layout (which entry is
negated, where (2 + i) alpha_y enters) and the bound 4 (P - 1)^2 + 2^32 < 2^64 are exactly what
a transposed row, a missing negation or an overflowing sum would break."""
import itertools
import os
import random
import shutil
import subprocess

import pytest

import refmath as R

P = R.P
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def field_bin(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("field") / "field_host_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-I" + os.path.join(ROOT, "stwo-gpu_amd", "csrc"),
                    os.path.join(HERE, "field_host_check.cpp"), "-o", out], check=True, capture_output=True)
    return out


def run(binary, lines):
    r = subprocess.run([binary], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return r.stdout.split("\n")


def fold_ref(a, b, t, al):
    return R.qm_add(R.qm_add(a, b), R.qm_mul(al, R.qm_scale(R.qm_sub(a, b), t)))


EDGE = (0, 1, 2, P - 2, P - 1, 1 << 30)


def cases(rng, n_random):
    edge_q = [(P - 1,) * 4, (0,) * 4, (1, 0, 0, 0), (0, 0, 0, 1), (P - 1, 0, P - 1, 0), (0, P - 1, 0, P - 1),
              (1 << 30, P - 2, 1, P - 1)]
    out = []
    for a, b, al in itertools.product(edge_q, edge_q, edge_q):
        for t in (0, 1, P - 1):
            out.append((a, b, t, al))
    for _ in range(n_random):
        out.append(tuple(tuple(rng.randrange(P) for _ in range(4)) if k != 2 else rng.randrange(P) for k in range(4)))
    return out


def test_fold_mat(field_bin):
    rng = random.Random(1)
    cs = cases(rng, 2000)
    got = run(field_bin, ["fold " + " ".join(map(str, [*a, *b, t, *al])) for a, b, t, al in cs])
    for (a, b, t, al), line in zip(cs, got):
        assert tuple(map(int, line.split())) == fold_ref(a, b, t, al), (a, b, t, al)


def test_fold_mat_acc(field_bin):
    rng = random.Random(2)
    cs = cases(rng, 2000)
    accs = [(P - 1,) * 4 if i % 3 == 0 else tuple(rng.randrange(P) for _ in range(4)) for i in range(len(cs))]
    got = run(field_bin, ["facc " + " ".join(map(str, [*a, *b, t, *al, *acc])) for (a, b, t, al), acc in zip(cs, accs)])
    for ((a, b, t, al), acc), line in zip(zip(cs, accs), got):
        want = R.qm_add(R.qm_mul(acc, R.qm_mul(al, al)), fold_ref(a, b, t, al))
        assert tuple(map(int, line.split())) == want, (a, b, t, al, acc)


def test_reduce64_and_sub3(field_bin):
    rng = random.Random(3)
    ys = [0, 1, P, P + 1, 2 * P, (1 << 32) - 1, 1 << 32, (1 << 64) - 1, (1 << 64) - 2, 4 * (P - 1) ** 2 + (1 << 32) - 1]
    ys += [rng.randrange(1 << 64) for _ in range(500)]
    got = run(field_bin, [f"r64 {y >> 32} {y & 0xFFFFFFFF}" for y in ys])
    for y, line in zip(ys, got):
        assert int(line) == y % P, y
    pairs = list(itertools.product(EDGE, EDGE)) + [(rng.randrange(P), rng.randrange(P)) for _ in range(500)]
    got = run(field_bin, [f"sub3 {a} {b}" for a, b in pairs])
    for (a, b), line in zip(pairs, got):
        assert int(line) == (a - b) % P, (a, b)


def test_mul_dbl_and_reductions(field_bin):
    # mul_dbl(x, 2t) = x t for x < P, t <= P - 1 (the butterflies' product; its bound hi + lo/2 < 2P
    # is tightest at x = t = P - 1); reduce62 for products, reduce63 for sums of two products
    rng = random.Random(4)
    pairs = list(itertools.product(EDGE, EDGE)) + [(rng.randrange(P), rng.randrange(P)) for _ in range(2000)]
    got = run(field_bin, [f"mdbl {x} {t}" for x, t in pairs])
    for (x, t), line in zip(pairs, got):
        assert int(line) == x * t % P, (x, t)
    ys62 = [(P - 1) ** 2, 0, 1, P * P - 1] + [rng.randrange(P) * rng.randrange(P) for _ in range(500)]
    ys62 = [y for y in ys62 if y < 1 << 62]
    got = run(field_bin, [f"r62 {y >> 32} {y & 0xFFFFFFFF}" for y in ys62])
    for y, line in zip(ys62, got):
        assert int(line) == y % P, y
    ys63 = [2 * (P - 1) ** 2, (1 << 63) - 1] + [rng.randrange(1 << 63) for _ in range(500)]
    got = run(field_bin, [f"r63 {y >> 32} {y & 0xFFFFFFFF}" for y in ys63])
    for y, line in zip(ys63, got):
        assert int(line) == y % P, y


def test_qmul(field_bin):
    rng = random.Random(5)
    edge_q = [(P - 1,) * 4, (0,) * 4, (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (P - 1, 0, P - 1, 0)]
    cs = list(itertools.product(edge_q, edge_q)) + [
        (tuple(rng.randrange(P) for _ in range(4)), tuple(rng.randrange(P) for _ in range(4))) for _ in range(1000)]
    got = run(field_bin, ["qmul " + " ".join(map(str, [*a, *b])) for a, b in cs])
    for (a, b), line in zip(cs, got):
        assert tuple(map(int, line.split())) == R.qm_mul(a, b), (a, b)
