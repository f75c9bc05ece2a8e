"""ctypes front end of the oracle (CPU restatement of the Stwo backend ops).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never calls
into this module.

Arrays are numpy uint32.  QM31 values are 4-tuples (a, b, c, d) meaning
(a + b i) + (c + d i) u; SoA secure columns are arrays of shape (4, n).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

P = (1 << 31) - 1
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libstwo_oracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)


def build() -> None:
    """Compile the oracle in place (gcc via the committed Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    u32, u64 = C.c_uint32, C.c_uint64
    sig = {
        "or_m31_add": (u32, [u32, u32]), "or_m31_sub": (u32, [u32, u32]),
        "or_m31_neg": (u32, [u32]), "or_m31_mul": (u32, [u32, u32]), "or_m31_inv": (u32, [u32]),
        "or_qm31_add": (None, [u32p, u32p, u32p]), "or_qm31_sub": (None, [u32p, u32p, u32p]),
        "or_qm31_mul": (None, [u32p, u32p, u32p]), "or_qm31_inv": (None, [u32p, u32p]),
        "or_qm31_pow": (None, [u32p, u64, u32p]), "or_cm31_inv": (None, [u32p, u32p]),
        "or_point_at_index": (None, [u32, u32p]), "or_coset_at": (None, [u32, u32, u32, u32p]),
        "or_circle_domain_at": (None, [u32, u32, u32, u64, u32p]),
        "or_bit_reverse_u32": (None, [u32p, u64]), "or_bit_reverse_qm31": (None, [u32p, u64]),
        "or_new_canonical_ordered": (None, [u32p, u32p, u64]),
        "or_precompute_twiddles": (None, [u32, u32, u32, u32p, u32p]),
        "or_interpolate": (None, [u32p, u32, u32, u32p, u64]),
        "or_evaluate": (None, [u32p, u32, u32, u32p, u64]),
        "or_lde_columns": (None, [C.POINTER(u32p), C.POINTER(u32p), C.POINTER(u32p), u32, u32, u32,
                                  u32p, u32p, u64]),
        "or_eval_at_point": (None, [u32p, u32, u32p, u32p, u32p]),
        "or_batch_inverse_m31": (None, [u32p, u32p, u64]),
        "or_batch_inverse_qm31": (None, [u32p, u32p, u64]),
        "or_blake2s_compress": (None, [u32p, u32p, u32, u32, u32, u32]),
        "or_blake2s256": (None, [C.c_char_p, u64, C.c_char_p]),
        "or_merkle_commit_on_layer": (None, [u32, u32p, u32, C.POINTER(u32p), u32p]),
        "or_fold_line": (None, [C.POINTER(u32p), u64, u32, u32, u32p, C.POINTER(u32p)]),
        "or_fold_circle_into_line": (None, [C.POINTER(u32p), u64, u32, u32, u32p, C.POINTER(u32p)]),
        "or_accumulate_quotients": (None, [u32, u32, u32, C.POINTER(u32p), u32p, u32, u32p, u32p,
                                           u32p, u32p, C.POINTER(u32p)]),
        "or_accumulate": (None, [u64, C.POINTER(u32p), C.POINTER(u32p)]),
        "or_wide_fib_constraint_quotients": (None, [u64, u32, C.POINTER(u32p), u32p, u32p,
                                                    C.POINTER(u32p)]),
        "or_gen_eq_evals": (None, [u32p, u32p, u32, u32p]),
        "or_fix_first_variable_base": (None, [u32p, u64, u32p, u32p]),
        "or_fix_first_variable_secure": (None, [u32p, u64, u32p, u32p]),
        "or_set_mutation": (None, [C.c_int]),
        "or_uniform_m31": (u32, [u64, u64]),
        "or_fill_uniform_m31": (None, [u32p, u64, u64, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


def _arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _ptrs(arrs):
    return (u32p * len(arrs))(*[_p(a) for a in arrs])


# ---------------------------------------------------------------- fields
def m31_mul(a, b): return lib().or_m31_mul(a, b)
def m31_add(a, b): return lib().or_m31_add(a, b)
def m31_sub(a, b): return lib().or_m31_sub(a, b)
def m31_neg(a): return lib().or_m31_neg(a)
def m31_inv(a): return lib().or_m31_inv(a)


def _qm(fn, *args):
    ins = [_arr(a) for a in args]
    out = np.zeros(4, np.uint32)
    fn(*[_p(a) for a in ins], _p(out))
    return tuple(int(v) for v in out)


def qm31_add(a, b): return _qm(lib().or_qm31_add, a, b)
def qm31_sub(a, b): return _qm(lib().or_qm31_sub, a, b)
def qm31_mul(a, b): return _qm(lib().or_qm31_mul, a, b)
def qm31_inv(a): return _qm(lib().or_qm31_inv, a)


def qm31_pow(a, e):
    out = np.zeros(4, np.uint32)
    a = _arr(a)
    lib().or_qm31_pow(_p(a), e, _p(out))
    return tuple(int(v) for v in out)


def cm31_inv(a):
    a = _arr(a)
    out = np.zeros(2, np.uint32)
    lib().or_cm31_inv(_p(a), _p(out))
    return tuple(int(v) for v in out)


# ---------------------------------------------------------------- circle
def point_at_index(i):
    out = np.zeros(2, np.uint32)
    lib().or_point_at_index(i & 0x7FFFFFFF, _p(out))
    return int(out[0]), int(out[1])


def circle_domain_at(half_initial, half_step, half_log, i):
    out = np.zeros(2, np.uint32)
    lib().or_circle_domain_at(half_initial, half_step, half_log, i, _p(out))
    return int(out[0]), int(out[1])


def canonic_half_coset(log_n):
    """CanonicCoset(log_n).half_coset as (initial_index, step_size, log_size)."""
    return 1 << (30 - log_n), (1 << (32 - log_n)) & 0x7FFFFFFF, log_n - 1


# ---------------------------------------------------------------- columns
def bit_reverse(a):
    a = _arr(a).copy()
    lib().or_bit_reverse_u32(_p(a), a.size)
    return a


def bit_reverse_qm31(a):
    a = _arr(a).copy()
    lib().or_bit_reverse_qm31(_p(a), a.size // 4)
    return a


def new_canonical_ordered(a):
    a = _arr(a)
    out = np.zeros_like(a)
    lib().or_new_canonical_ordered(_p(a), _p(out), a.size)
    return out


def precompute_twiddles(initial_index, step_size, log_size):
    n = 1 << log_size
    tw = np.zeros(n, np.uint32)
    itw = np.zeros(n, np.uint32)
    lib().or_precompute_twiddles(initial_index, step_size, log_size, _p(tw), _p(itw))
    return tw, itw


def interpolate(values, half_initial, itwiddles):
    v = _arr(values).copy()
    itw = _arr(itwiddles)
    lib().or_interpolate(_p(v), half_initial, v.size.bit_length() - 2, _p(itw), itw.size)
    return v


def evaluate(coeffs, half_initial, twiddles):
    v = _arr(coeffs).copy()
    tw = _arr(twiddles)
    lib().or_evaluate(_p(v), half_initial, v.size.bit_length() - 2, _p(tw), tw.size)
    return v


def lde_columns(evals, log_blowup, twiddles, itwiddles):
    evals = [_arr(e) for e in evals]
    n = evals[0].size
    log_n = n.bit_length() - 1
    coeffs = [np.zeros(n, np.uint32) for _ in evals]
    ldes = [np.zeros(n << log_blowup, np.uint32) for _ in evals]
    tw, itw = _arr(twiddles), _arr(itwiddles)
    lib().or_lde_columns(_ptrs(evals), _ptrs(coeffs), _ptrs(ldes), len(evals), log_n, log_blowup,
                         _p(tw), _p(itw), tw.size)
    return coeffs, ldes


def eval_at_point(coeffs, px, py):
    c = _arr(coeffs)
    out = np.zeros(4, np.uint32)
    lib().or_eval_at_point(_p(c), c.size.bit_length() - 1, _p(_arr(px)), _p(_arr(py)), _p(out))
    return tuple(int(v) for v in out)


def batch_inverse_m31(a):
    a = _arr(a)
    out = np.zeros_like(a)
    lib().or_batch_inverse_m31(_p(a), _p(out), a.size)
    return out


def batch_inverse_qm31(a):
    a = _arr(a)
    out = np.zeros_like(a)
    lib().or_batch_inverse_qm31(_p(a), _p(out), a.size // 4)
    return out


# ---------------------------------------------------------------- blake2s / merkle
def blake2s_compress(h, m, t0=0, t1=0, f0=0, f1=0):
    h = _arr(h).copy()
    m = _arr(m)
    lib().or_blake2s_compress(_p(h), _p(m), t0, t1, f0, f1)
    return h


def blake2s256(data: bytes) -> bytes:
    out = C.create_string_buffer(32)
    lib().or_blake2s256(data, len(data), out)
    return out.raw


def merkle_commit_on_layer(log_size, prev, cols):
    cols = [_arr(c) for c in cols]
    out = np.zeros((1 << log_size) * 8, np.uint32)
    keep = _arr(prev) if prev is not None else None
    prev_p = _p(keep) if keep is not None else None
    lib().or_merkle_commit_on_layer(log_size, prev_p, len(cols), _ptrs(cols) if cols else None, _p(out))
    return out


def merkle_commit(cols):
    """Full tree as stwo's MerkleProver::commit builds it: columns sorted by
    decreasing length (stable), one layer per log size from max down to 0.
    Returns layers[log_size] (root layer is layers[0])."""
    cols = [_arr(c) for c in cols]
    order = sorted(range(len(cols)), key=lambda i: -cols[i].size)
    max_log = cols[order[0]].size.bit_length() - 1
    layers = {}
    prev = None
    for log_size in range(max_log, -1, -1):
        layer_cols = [cols[i] for i in order if cols[i].size == (1 << log_size)]
        prev = merkle_commit_on_layer(log_size, prev, layer_cols)
        layers[log_size] = prev
    return layers


# ---------------------------------------------------------------- FRI
def fold_line(src, line_initial, line_step, alpha):
    src = [_arr(c) for c in src]
    n = src[0].size
    dst = [np.zeros(n // 2, np.uint32) for _ in range(4)]
    lib().or_fold_line(_ptrs(src), n, line_initial, line_step, _p(_arr(alpha)), _ptrs(dst))
    return np.stack(dst)


def fold_circle_into_line(dst, src, half_initial, half_step, alpha):
    src = [_arr(c) for c in src]
    dst = [_arr(c).copy() for c in dst]
    lib().or_fold_circle_into_line(_ptrs(src), src[0].size, half_initial, half_step,
                                   _p(_arr(alpha)), _ptrs(dst))
    return np.stack(dst)


# ---------------------------------------------------------------- quotients
def accumulate_quotients(half_initial, half_step, log_size, cols, random_coeff, batches):
    """batches: list of (point, [(column_index, value), ...]); point = (x4, y4)."""
    cols = [_arr(c) for c in cols]
    points = _arr([list(p[0]) + list(p[1]) for p, _ in batches]).reshape(-1)
    sizes = _arr([len(cv) for _, cv in batches])
    idx = _arr([c for _, cv in batches for c, _ in cv] or [0])
    vals = _arr([list(v) for _, cv in batches for _, v in cv] or [[0, 0, 0, 0]]).reshape(-1)
    n = 1 << log_size
    out = [np.zeros(n, np.uint32) for _ in range(4)]
    lib().or_accumulate_quotients(half_initial, half_step, log_size, _ptrs(cols),
                                  _p(_arr(random_coeff)), len(batches), _p(points), _p(sizes),
                                  _p(idx), _p(vals), _ptrs(out))
    return np.stack(out)


def accumulate(left, right):
    left = [_arr(c).copy() for c in left]
    right = [_arr(c) for c in right]
    lib().or_accumulate(left[0].size, _ptrs(left), _ptrs(right))
    return np.stack(left)


def wide_fib_constraint_quotients(trace, coeff_powers, denom_inv):
    trace = [_arr(c) for c in trace]
    n = trace[0].size
    cp = _arr(coeff_powers).reshape(-1)
    di = _arr(denom_inv)
    out = [np.zeros(n, np.uint32) for _ in range(4)]
    lib().or_wide_fib_constraint_quotients(n, len(trace), _ptrs(trace), _p(cp), _p(di), _ptrs(out))
    return np.stack(out)


# ---------------------------------------------------------------- multilinear ops
def gen_eq_evals(y, v):
    """y: list of QM31 4-tuples; returns (2^len(y), 4) AoS evaluations."""
    n = len(y)
    yy = _arr([list(t) for t in y] or [[0, 0, 0, 0]]).reshape(-1)
    out = np.zeros((1 << n) * 4, np.uint32)
    lib().or_gen_eq_evals(_p(_arr(v)), _p(yy), n, _p(out))
    return out.reshape(-1, 4)


def fix_first_variable(evals, a):
    """evals: M31 array (n,) or QM31 AoS array (n, 4); returns (n/2, 4)."""
    ev = _arr(evals)
    n = ev.shape[0]
    out = np.zeros((n // 2) * 4, np.uint32)
    f = lib().or_fix_first_variable_secure if ev.ndim == 2 else lib().or_fix_first_variable_base
    f(_p(np.ascontiguousarray(ev).reshape(-1)), n, _p(_arr(a)), _p(out))
    return out.reshape(-1, 4)


# ---------------------------------------------------------------- planted mistakes
MUT_QUOT_ALPHA_POW, MUT_QUOT_BATCH_POW, MUT_QUOT_BATCH_ORDER, MUT_QUOT_REF_OFFSET, \
    MUT_WF_POWER_ORDER = 1, 2, 3, 4, 5


class mutation:
    """Context manager for the oracle's pinning tests: inside it the oracle
    computes the named deliberately wrong variant (stwo_oracle.h OR_MUT_*)."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().or_set_mutation(self.m)

    def __exit__(self, *exc):
        lib().or_set_mutation(0)


# ---------------------------------------------------------------- inputs
def fill_uniform_m31(n, seed, offset=0):
    out = np.zeros(n, np.uint32)
    lib().or_fill_uniform_m31(_p(out), n, seed, offset)
    return out
