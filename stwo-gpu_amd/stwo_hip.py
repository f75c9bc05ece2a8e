"""Python binding of libstwo_hip.so (the gfx950 Stwo backend) over its C ABI.

The device-buffer classes and functions mirror the reference crate's Rust
surface (stwo_gpu_backend/src/cuda/*_vec.rs and the trait impls in
stwo_gpu_backend/src/{column,field,poly,blake2s,fri,quotient,accumulation}.rs)
so tests read like the reference's own.  Every call goes to the HIP library;
there is no CPU fallback: loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

P = (1 << 31) - 1
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libstwo_hip.so")

u32 = C.c_uint32
u64 = C.c_uint64
sz = C.c_size_t
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p


class WideFibParams(C.Structure):
    _fields_ = [("n_cols", u32), ("log_n", u32), ("n_queries", u32), ("pow_bits", u32), ("last_log", u32)]


class QM31(C.Structure):
    _fields_ = [("a", u32), ("b", u32), ("c", u32), ("d", u32)]


class PointM31(C.Structure):
    _fields_ = [("x", u32), ("y", u32)]


class PointQM31(C.Structure):
    _fields_ = [("x", QM31), ("y", QM31)]


class StwoHipError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise StwoHipError(f"{LIB_PATH} is not built (run python stwo-gpu_amd/build.py)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    pp = C.POINTER(vp)
    sig = {
        # reference FFI (bindings.rs)
        "copy_uint32_t_vec_from_device_to_host": (None, [vp, vp, u32]),
        "copy_uint32_t_vec_from_host_to_device": (vp, [vp, u32]),
        "copy_uint32_t_vec_from_device_to_device": (vp, [vp, vp, u32]),
        "cuda_malloc_uint32_t": (vp, [u32]),
        "cuda_malloc_blake_2s_hash": (vp, [sz]),
        "cuda_alloc_zeroes_uint32_t": (vp, [u32]),
        "cuda_alloc_zeroes_blake_2s_hash": (vp, [sz]),
        "cuda_free_memory": (None, [vp]),
        "bit_reverse_base_field": (None, [vp, sz]),
        "bit_reverse_secure_field": (None, [vp, sz]),
        "batch_inverse_base_field": (None, [vp, vp, sz]),
        "batch_inverse_secure_field": (None, [vp, vp, sz]),
        "sort_values_and_permute_with_bit_reverse_order": (vp, [vp, sz]),
        "precompute_twiddles": (vp, [PointM31, PointM31, sz]),
        "interpolate": (None, [u32, vp, vp, u32, u32]),
        "interpolate_columns": (None, [u32, pp, vp, u32, u32, u32]),
        "evaluate": (None, [u32, vp, vp, u32, u32]),
        "evaluate_columns": (None, [u32p, pp, vp, u32, u32, u32p]),
        "eval_at_point": (QM31, [vp, u32, QM31, QM31]),
        "fold_line": (None, [vp, sz, sz, pp, QM31, pp]),
        "fold_circle_into_line": (None, [vp, sz, sz, pp, QM31, pp]),
        "accumulate": (None, [u32, pp, pp]),
        "commit_on_first_layer": (None, [sz, sz, vp, vp]),
        "commit_on_layer_with_previous": (None, [sz, sz, vp, vp, vp]),
        "copy_blake_2s_hash_vec_from_host_to_device": (vp, [vp, sz]),
        "copy_blake_2s_hash_vec_from_device_to_host": (None, [vp, vp, sz]),
        "copy_blake_2s_hash_vec_from_device_to_device": (None, [vp, vp, sz]),
        "copy_device_pointer_vec_from_host_to_device": (vp, [pp, sz]),
        "accumulate_quotients": (None, [u32, u32, u32, vp, sz, QM31, vp, u32p, u32, C.POINTER(QM31), u32p,
                                        u32, vp, vp, vp, vp, u32]),
        "evaluate_wide_fibonacci_constraint_quotients_on_domain": (None, [vp, vp, vp, vp, pp, vp, vp, u32, u32]),
        "gen_eq_evals": (None, [QM31, C.POINTER(QM31), u32, vp, u32]),
        "fix_first_variable_base_field": (None, [vp, sz, QM31, vp]),
        "fix_first_variable_secure_field": (None, [vp, sz, QM31, vp]),
        # extensions
        "stwo_hip_init": (C.c_int, [C.c_int]),
        "stwo_hip_last_status": (C.c_int, []),
        "stwo_hip_last_error": (C.c_char_p, []),
        "stwo_hip_clear_error": (None, []),
        "stwo_hip_synchronize": (C.c_int, []),
        "stwo_hip_stream": (vp, []),
        "stwo_hip_malloc": (C.c_int, [C.POINTER(vp), sz]),
        "stwo_hip_free": (C.c_int, [vp]),
        "stwo_hip_memcpy": (C.c_int, [vp, vp, sz]),
        "stwo_hip_release_cached_memory": (C.c_int, []),
        "stwo_hip_fill_uniform_m31": (C.c_int, [vp, u64, u64, u64]),
        "stwo_hip_device_copy": (C.c_int, [vp, vp, sz]),
        "stwo_hip_twiddles": (C.c_int, [u32, u32, u32, C.POINTER(vp), C.POINTER(vp)]),
        "stwo_hip_release_twiddles": (C.c_int, []),
        "stwo_hip_interpolate_columns_async": (C.c_int, [pp, u32p, u32, vp, u64]),
        "stwo_hip_evaluate_columns_async": (C.c_int, [pp, u32p, pp, u32p, u32, vp, u64]),
        "stwo_hip_interpolate_columns_oop_async": (C.c_int, [pp, pp, u32p, u32, vp, u64]),
        "stwo_hip_lde_columns_async": (C.c_int, [pp, pp, pp, u32p, u32, u32, vp, vp, u64]),
        "stwo_hip_merkle_commit_async": (C.c_int, [u32, pp, u32p, pp, u32]),
        "stwo_hip_fold_line_async": (C.c_int, [vp, sz, sz, pp, QM31, pp]),
        "stwo_hip_fold_circle_into_line_async": (C.c_int, [vp, sz, sz, pp, QM31, pp]),
        "stwo_hip_accumulate_quotients_async": (C.c_int, [u32, u32, u32, pp, u32, QM31, u32, vp, u32p, u32p,
                                                          C.POINTER(QM31), pp]),
        "stwo_hip_merkle_root": (C.c_int, [vp, vp]),
        "stwo_hip_commit_plan_create": (C.c_int, [C.POINTER(vp), u32, pp, pp, pp, u32p, u32, vp, vp, u64, pp, u32,
                                                  C.c_int]),
        "stwo_hip_commit_plan_run": (C.c_int, [vp, C.c_int]),
        "stwo_hip_commit_plan_uses_graph": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "stwo_hip_commit_plan_destroy": (C.c_int, [vp]),
        "stwo_hip_lde_schedule": (C.c_int, [u32, u32, u64, u32p, C.POINTER(C.c_int)]),
        "stwo_hip_grind": (C.c_int, [u32p, u32, C.POINTER(u64)]),
        "stwo_hip_eval_at_point_columns": (C.c_int, [pp, u32, u32, QM31, QM31, vp]),
        "stwo_hip_wide_fib_trace": (C.c_int, [pp, u32, u32, u64]),
        "stwo_hip_vanishing_inverse": (C.c_int, [vp, u32, u32]),
        "stwo_hip_wide_fib_proof_size": (C.c_int, [C.POINTER(WideFibParams), C.POINTER(u64), C.POINTER(u64)]),
        "stwo_hip_wide_fib_prove": (C.c_int, [C.POINTER(WideFibParams), u64, u32p, u64]),
        "stwo_hip_server_create": (C.c_int, [C.POINTER(vp), u64, u32]),
        "stwo_hip_server_limit_work": (C.c_int, [vp, u64]),
        "stwo_hip_gather_rows": (C.c_int, [vp, u32, u64, u32p, u32, u32p]),
        "stwo_hip_eval_at_points_columns": (C.c_int, [vp, u32, u32, vp, u32, vp]),
        "stwo_hip_fri_fold_chain_async": (C.c_int, [vp, u64, u32, vp, u32, vp, vp]),
        "stwo_hip_eval_at_points_columns_noalloc": (C.c_int, [vp, u32, u32, vp, u32, vp]),
        "stwo_hip_merkle_auth_paths": (C.c_int, [vp, u32, u32p, u32, u32p]),
        "stwo_hip_host_tree_hash": (C.c_int, [u32p, u64, u32p]),
        "stwo_hip_server_batch": (C.c_int, [vp, u32, u64]),
        "stwo_hip_wide_fib_prove_batch": (C.c_int, [C.POINTER(WideFibParams), vp, u32, u32p, u64]),
        "stwo_hip_server_submit": (C.c_int, [vp, C.POINTER(WideFibParams), u64, C.POINTER(u64)]),
        "stwo_hip_server_wait": (C.c_int, [vp, u32, C.POINTER(u64), C.POINTER(C.c_int), u32p, u64,
                                           C.POINTER(u64), C.POINTER(C.c_double)]),
        "stwo_hip_server_destroy": (C.c_int, [vp]),
        "stwo_hip_timer_begin": (C.c_int, []),
        "stwo_hip_timer_end": (C.c_int, [C.POINTER(C.c_float)]),
        "stwo_hip_kernel_timing": (C.c_int, [C.c_int]),
        "stwo_hip_kernel_stats": (C.c_int, [C.c_char_p, C.POINTER(u64), C.POINTER(C.c_double)]),
        "stwo_hip_kernel_stats_reset": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def exported_symbols():
    """Names every header declaration must resolve to (checked by the CPU tests)."""
    import re
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "stwo_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    hdr = "\n".join(l for l in hdr.splitlines() if not l.lstrip().startswith("#"))  # preprocessor lines
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", hdr)
    skip = {"if", "for", "while", "sizeof", "return"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def check(status=None):
    """Raise if the last library call failed."""
    L = lib()
    st = L.stwo_hip_last_status() if status is None else status
    if st != 0:
        msg = L.stwo_hip_last_error().decode()
        L.stwo_hip_clear_error()
        raise StwoHipError(f"stwo_hip error {st}: {msg}")


def last_error():
    """Message of the last failure on the calling thread ("" when none)."""
    return lib().stwo_hip_last_error().decode()


def _c(status):
    check(status)


def _ptr_array(ptrs):
    return (vp * len(ptrs))(*ptrs)


def qm31(v) -> QM31:
    return QM31(*[int(x) for x in v])


def from_qm31(q: QM31):
    return (q.a, q.b, q.c, q.d)


# ----------------------------------------------------------------- device buffers
class DeviceBuffer:
    """Owning device allocation (freed with cuda_free_memory, like the Rust Drop impls)."""

    def __init__(self, ptr, nbytes):
        if not ptr and nbytes:
            check()
            raise StwoHipError("allocation failed")
        self.ptr = ptr
        self.nbytes = nbytes

    def free(self):
        if self.ptr and not getattr(self, "_borrowed", False):
            lib().cuda_free_memory(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BaseFieldVec(DeviceBuffer):
    """stwo_gpu_backend/src/cuda/base_field_vec.rs:7-79: device M31 column."""

    def __init__(self, ptr, size):
        super().__init__(ptr, 4 * size)
        self.size = size

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        ptr = lib().copy_uint32_t_vec_from_host_to_device(a.ctypes.data, a.size)
        check()
        return cls(ptr, a.size)

    @classmethod
    def uninitialized(cls, size):
        ptr = lib().cuda_malloc_uint32_t(size)
        check()
        return cls(ptr, size)

    @classmethod
    def view(cls, ptr, size):
        """Non-owning view of device memory owned elsewhere (never freed through this object)."""
        v = cls(ptr, size)
        v._borrowed = True
        return v

    @classmethod
    def zeros(cls, size):
        ptr = lib().cuda_alloc_zeroes_uint32_t(size)
        check()
        return cls(ptr, size)

    def to_numpy(self):
        out = np.empty(self.size, np.uint32)
        lib().copy_uint32_t_vec_from_device_to_host(self.ptr, out.ctypes.data, self.size)
        check()
        return out

    def __len__(self):
        return self.size


class SecureFieldVec(DeviceBuffer):
    """secure_field_vec.rs:7-72: device QM31 vector, AoS (4 words per element)."""

    def __init__(self, ptr, size):
        super().__init__(ptr, 16 * size)
        self.size = size

    @classmethod
    def from_numpy(cls, a):  # shape (n, 4)
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1)
        ptr = lib().copy_uint32_t_vec_from_host_to_device(a.ctypes.data, a.size)
        check()
        return cls(ptr, a.size // 4)

    @classmethod
    def uninitialized(cls, size):
        ptr = lib().cuda_malloc_uint32_t(4 * size)
        check()
        return cls(ptr, size)

    def to_numpy(self):
        out = np.empty(4 * self.size, np.uint32)
        lib().copy_uint32_t_vec_from_device_to_host(self.ptr, out.ctypes.data, 4 * self.size)
        check()
        return out.reshape(-1, 4)


class Blake2sHashVec(DeviceBuffer):
    """blake_2s_hash_vec.rs:5-72: device vector of 32-byte hashes."""

    def __init__(self, ptr, size):
        super().__init__(ptr, 32 * size)
        self.size = size

    @classmethod
    def uninitialized(cls, size):
        ptr = lib().cuda_malloc_blake_2s_hash(size)
        check()
        return cls(ptr, size)

    def to_numpy(self):
        out = np.empty(8 * self.size, np.uint32)
        lib().copy_blake_2s_hash_vec_from_device_to_host(self.ptr, out.ctypes.data, self.size)
        check()
        return out.reshape(-1, 8)


class SecureColumn:
    """SecureColumnByCoords<CudaBackend>: four M31 device columns (secure_column.rs:6-57)."""

    def __init__(self, cols):
        self.columns = list(cols)

    @classmethod
    def from_numpy(cls, a):  # shape (4, n)
        return cls([BaseFieldVec.from_numpy(a[k]) for k in range(4)])

    @classmethod
    def uninitialized(cls, n):
        return cls([BaseFieldVec.uninitialized(n) for _ in range(4)])

    @classmethod
    def zeros(cls, n):
        return cls([BaseFieldVec.zeros(n) for _ in range(4)])

    def ptrs(self):
        return _ptr_array([c.ptr for c in self.columns])

    def to_numpy(self):
        return np.stack([c.to_numpy() for c in self.columns])

    def __len__(self):
        return self.columns[0].size


# ----------------------------------------------------------------- cosets / domains
@dataclass(frozen=True)
class Coset:
    initial_index: int
    step_size: int
    log_size: int

    def size(self):
        return 1 << self.log_size

    def double(self):
        m = (1 << 31) - 1
        return Coset((2 * self.initial_index) & m, (2 * self.step_size) & m, self.log_size - 1)


def canonic_half_coset(log_size) -> Coset:
    """CanonicCoset(log_size).half_coset (Coset::half_odds(log_size - 1))."""
    return Coset(1 << (30 - log_size), (1 << (32 - log_size)) & ((1 << 31) - 1), log_size - 1)


def canonic_coset(log_size) -> Coset:
    """CanonicCoset(log_size).coset (Coset::odds(log_size))."""
    return Coset(1 << (30 - log_size), (1 << (31 - log_size)) & ((1 << 31) - 1), log_size)


def _pt_mul(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def point_at_index(i):
    r, g = (1, 0), (2, 1268011823)
    i &= (1 << 31) - 1
    while i:
        if i & 1:
            r = _pt_mul(r, g)
        g = _pt_mul(g, g)
        i >>= 1
    return r


@dataclass
class TwiddleTree:
    """TwiddleTree<CudaBackend> (poly.rs:186-190)."""
    root_coset: Coset
    twiddles: BaseFieldVec
    itwiddles: BaseFieldVec


# ----------------------------------------------------------------- backend ops
def bit_reverse_column(col):
    if isinstance(col, SecureFieldVec):
        lib().bit_reverse_secure_field(col.ptr, col.size)
    else:
        lib().bit_reverse_base_field(col.ptr, col.size)
    check()


def batch_inverse(src, dst):
    if isinstance(src, SecureFieldVec):
        lib().batch_inverse_secure_field(src.ptr, dst.ptr, src.size)
    else:
        lib().batch_inverse_base_field(src.ptr, dst.ptr, src.size)
    check()


def new_canonical_ordered(values: BaseFieldVec) -> BaseFieldVec:
    ptr = lib().sort_values_and_permute_with_bit_reverse_order(values.ptr, values.size)
    check()
    return BaseFieldVec(ptr, values.size)


def precompute_twiddles(coset: Coset) -> TwiddleTree:
    """poly.rs:170-192: tree from the coset's initial/step points, then batch inverse."""
    ini = point_at_index(coset.initial_index)
    step = point_at_index(coset.step_size)
    ptr = lib().precompute_twiddles(PointM31(*ini), PointM31(*step), coset.size())
    check()
    tw = BaseFieldVec(ptr, coset.size())
    itw = BaseFieldVec.uninitialized(coset.size())
    batch_inverse(tw, itw)
    return TwiddleTree(coset, tw, itw)


def interpolate(values: BaseFieldVec, tree: TwiddleTree):
    lib().interpolate(values.size // 2, values.ptr, tree.itwiddles.ptr, tree.itwiddles.size, values.size)
    check()


def interpolate_columns(cols, tree: TwiddleTree):
    n = cols[0].size
    lib().interpolate_columns(n // 2, _ptr_array([c.ptr for c in cols]), tree.itwiddles.ptr, tree.itwiddles.size,
                              len(cols), n)
    check()


def evaluate(values: BaseFieldVec, tree: TwiddleTree):
    lib().evaluate(values.size // 2, values.ptr, tree.twiddles.ptr, tree.twiddles.size, values.size)
    check()


def evaluate_columns(cols, tree: TwiddleTree):
    sizes = np.array([c.size for c in cols], np.uint32)
    halves = sizes // 2
    lib().evaluate_columns(halves.ctypes.data_as(u32p), _ptr_array([c.ptr for c in cols]), tree.twiddles.ptr,
                           tree.twiddles.size, len(cols), sizes.ctypes.data_as(u32p))
    check()


def eval_at_point(coeffs: BaseFieldVec, px, py):
    r = lib().eval_at_point(coeffs.ptr, coeffs.size, qm31(px), qm31(py))
    check()
    return from_qm31(r)


def commit_on_layer(log_size, prev, columns):
    """MerkleOps::commit_on_layer (blake2s.rs:19-97)."""
    size = 1 << log_size
    out = Blake2sHashVec.uninitialized(size)
    table = lib().copy_device_pointer_vec_from_host_to_device(_ptr_array([c.ptr for c in columns] or [None]),
                                                               max(len(columns), 1))
    check()
    try:
        if prev is None:
            lib().commit_on_first_layer(size, len(columns), table, out.ptr)
        else:
            lib().commit_on_layer_with_previous(size, len(columns), table, prev.ptr, out.ptr)
        check()
    finally:
        lib().cuda_free_memory(table)
    return out


def fold_line(src: SecureColumn, alpha, tree: TwiddleTree) -> SecureColumn:
    """FriOps::fold_line (fri.rs:16-43): twiddle_offset = itwiddles.len - n."""
    n = len(src)
    dst = SecureColumn.uninitialized(n // 2)
    lib().fold_line(tree.itwiddles.ptr, tree.itwiddles.size - n, n, src.ptrs(), qm31(alpha), dst.ptrs())
    check()
    return dst


def fold_circle_into_line(dst: SecureColumn, src: SecureColumn, alpha, tree: TwiddleTree):
    """FriOps::fold_circle_into_line (fri.rs:45-66): twiddle_offset = root.size - dst.size."""
    n = len(src)
    lib().fold_circle_into_line(tree.itwiddles.ptr, tree.root_coset.size() - len(dst), n, src.ptrs(), qm31(alpha),
                                dst.ptrs())
    check()


def accumulate(left: SecureColumn, right: SecureColumn):
    lib().accumulate(len(left), left.ptrs(), right.ptrs())
    check()


def accumulate_quotients(domain_half_coset: Coset, columns, random_coeff, sample_batches) -> SecureColumn:
    """QuotientOps::accumulate_quotients (quotient.rs:17-115).
    sample_batches: list of ((px4, py4), [(column_index, value4), ...])."""
    n = 2 * domain_half_coset.size()
    out = SecureColumn.uninitialized(n)
    table = lib().copy_device_pointer_vec_from_host_to_device(_ptr_array([c.ptr for c in columns]), len(columns))
    check()
    pts = (PointQM31 * max(1, len(sample_batches)))()
    for i, ((px, py), _) in enumerate(sample_batches):
        pts[i] = PointQM31(qm31(px), qm31(py))
    idx = np.array([c for _, cv in sample_batches for c, _ in cv] or [0], np.uint32)
    vals = (QM31 * max(1, len(idx)))(*[qm31(v) for _, cv in sample_batches for _, v in cv])
    sizes = np.array([len(cv) for _, cv in sample_batches] or [0], np.uint32)
    total = int(sum(len(cv) for _, cv in sample_batches))
    try:
        lib().accumulate_quotients(domain_half_coset.initial_index, domain_half_coset.step_size, n, table,
                                   len(columns), qm31(random_coeff), C.cast(pts, vp), idx.ctypes.data_as(u32p),
                                   total, vals, sizes.ctypes.data_as(u32p), len(sample_batches),
                                   *[c.ptr for c in out.columns], 3 * total)
        check()
    finally:
        lib().cuda_free_memory(table)
    return out


def wide_fib_constraint_quotients(out: SecureColumn, trace, powers: SecureFieldVec, denom_inv: BaseFieldVec):
    lib().evaluate_wide_fibonacci_constraint_quotients_on_domain(
        *[c.ptr for c in out.columns], _ptr_array([c.ptr for c in trace]), powers.ptr, denom_inv.ptr,
        len(out), len(trace))
    check()


def gen_eq_evals(y, v) -> SecureFieldVec:
    """GkrOps::gen_eq_evals (lookups/gkr.rs:12-27): v * eq(x, y) for every x in
    {0,1}^len(y), variable 0 the most significant index bit."""
    out = SecureFieldVec.uninitialized(1 << len(y))
    ys = (QM31 * max(1, len(y)))(*[qm31(t) for t in y])
    lib().gen_eq_evals(qm31(v), ys, len(y), out.ptr, out.size)
    check()
    return out


def fix_first_variable(mle, assignment) -> SecureFieldVec:
    """MleOps<BaseField | SecureField>::fix_first_variable (lookups/mle.rs:12-57)."""
    out = SecureFieldVec.uninitialized(mle.size // 2)
    f = (lib().fix_first_variable_secure_field if isinstance(mle, SecureFieldVec)
         else lib().fix_first_variable_base_field)
    f(mle.ptr, mle.size, qm31(assignment), out.ptr)
    check()
    return out


# ----------------------------------------------------------------- extensions
def synchronize():
    _c(lib().stwo_hip_synchronize())


def cached_twiddles(coset: Coset):
    tw, itw = vp(), vp()
    _c(lib().stwo_hip_twiddles(coset.initial_index, coset.step_size, coset.log_size, C.byref(tw), C.byref(itw)))
    return tw.value, itw.value


def device_copy(dst: DeviceBuffer, src: DeviceBuffer, nbytes=None):
    _c(lib().stwo_hip_device_copy(dst.ptr, src.ptr, src.nbytes if nbytes is None else nbytes))


def fill_uniform(col: BaseFieldVec, seed, offset=0):
    _c(lib().stwo_hip_fill_uniform_m31(col.ptr, col.size, seed, offset))


def interpolate_columns_async(cols, log_sizes, itw_ptr, tree_len):
    logs = np.ascontiguousarray(log_sizes, np.uint32)
    _c(lib().stwo_hip_interpolate_columns_async(_ptr_array([c.ptr for c in cols]), logs.ctypes.data_as(u32p),
                                                len(cols), itw_ptr, tree_len))


def interpolate_columns_oop_async(src, dst, log_sizes, itw_ptr, tree_len):
    logs = np.ascontiguousarray(log_sizes, np.uint32)
    _c(lib().stwo_hip_interpolate_columns_oop_async(_ptr_array([c.ptr for c in src]), _ptr_array([c.ptr for c in dst]),
                                                    logs.ctypes.data_as(u32p), len(src), itw_ptr, tree_len))


def lde_columns_async(evals, coeffs, ldes, log_sizes, log_blowup, tw_ptr, itw_ptr, tree_len):
    logs = np.ascontiguousarray(log_sizes, np.uint32)
    _c(lib().stwo_hip_lde_columns_async(_ptr_array([c.ptr for c in evals]), _ptr_array([c.ptr for c in coeffs]),
                                        _ptr_array([c.ptr for c in ldes]), logs.ctypes.data_as(u32p), len(evals),
                                        log_blowup, tw_ptr, itw_ptr, tree_len))


def evaluate_columns_async(coeffs, coeff_logs, dst, eval_logs, tw_ptr, tree_len):
    cl = np.ascontiguousarray(coeff_logs, np.uint32)
    el = np.ascontiguousarray(eval_logs, np.uint32)
    _c(lib().stwo_hip_evaluate_columns_async(_ptr_array([c.ptr for c in coeffs]), cl.ctypes.data_as(u32p),
                                             _ptr_array([c.ptr for c in dst]), el.ctypes.data_as(u32p), len(coeffs),
                                             tw_ptr, tree_len))


class MerkleTree:
    """All layers of a Blake2s Merkle tree over columns of mixed sizes."""

    def __init__(self, cols, log_sizes=None):
        if log_sizes is None:
            log_sizes = [c.size.bit_length() - 1 for c in cols]
        self.max_log = max(log_sizes)
        self.layers = [Blake2sHashVec.uninitialized(1 << k) for k in range(self.max_log + 1)]
        self._cols = cols
        self._logs = np.ascontiguousarray(log_sizes, np.uint32)

    def commit_async(self):
        _c(lib().stwo_hip_merkle_commit_async(len(self._cols), _ptr_array([c.ptr for c in self._cols]),
                                              self._logs.ctypes.data_as(u32p),
                                              _ptr_array([l.ptr for l in self.layers]), self.max_log))
        return self

    def root(self):
        out = np.zeros(8, np.uint32)
        _c(lib().stwo_hip_merkle_root(self.layers[0].ptr, out.ctypes.data))
        return out


def fold_line_async(src: SecureColumn, dst: SecureColumn, alpha, itw_ptr, twiddle_offset):
    _c(lib().stwo_hip_fold_line_async(itw_ptr, twiddle_offset, len(src), src.ptrs(), qm31(alpha), dst.ptrs()))


def fold_circle_into_line_async(dst: SecureColumn, src: SecureColumn, alpha, itw_ptr, twiddle_offset):
    _c(lib().stwo_hip_fold_circle_into_line_async(itw_ptr, twiddle_offset, len(src), src.ptrs(), qm31(alpha),
                                                  dst.ptrs()))


def fri_fold_chain_async(circle: SecureColumn, layers, alphas, itw_ptr, tree_len):
    """layers[0] accumulates the circle fold of `circle` with alphas[0]; layers[j] = fold_line(layers[j-1],
    alphas[j]) (all coefficients known up front; two launches for a log-22 circle folded to 2^4)."""
    log_circle = len(circle).bit_length() - 1
    ptrs = (vp * (4 * len(layers)))(*[p for l in layers for p in l.ptrs()])
    al = (QM31 * len(alphas))(*[qm31(a) for a in alphas])
    _c(lib().stwo_hip_fri_fold_chain_async(itw_ptr, tree_len, log_circle, circle.ptrs(), len(layers), C.cast(al, vp),
                                           C.cast(ptrs, vp)))


class QuotientSamples:
    """Host-side flattening of ColumnSampleBatch lists (quotient.rs:55-88) for the async call."""

    def __init__(self, sample_batches):
        nb = len(sample_batches)
        self.n = nb
        self.points = (PointQM31 * max(1, nb))()
        for i, ((px, py), _) in enumerate(sample_batches):
            self.points[i] = PointQM31(qm31(px), qm31(py))
        self.sizes = np.array([len(cv) for _, cv in sample_batches] or [0], np.uint32)
        self.idx = np.array([c for _, cv in sample_batches for c, _ in cv] or [0], np.uint32)
        self.vals = (QM31 * max(1, len(self.idx)))(*[qm31(v) for _, cv in sample_batches for _, v in cv])


def column_pointers(columns):
    """Host array of the columns' device pointers, for calls repeated over the same columns."""
    return _ptr_array([c.ptr for c in columns])


def accumulate_quotients_async(domain_half_coset: Coset, columns, random_coeff, samples: QuotientSamples,
                               out: SecureColumn, col_ptrs=None):
    """col_ptrs: column_pointers(columns), built once by a caller that repeats the call."""
    n = 2 * domain_half_coset.size()
    _c(lib().stwo_hip_accumulate_quotients_async(
        domain_half_coset.initial_index, domain_half_coset.step_size, n,
        col_ptrs if col_ptrs is not None else column_pointers(columns),
        len(columns), qm31(random_coeff), samples.n, C.cast(samples.points, vp), samples.sizes.ctypes.data_as(u32p),
        samples.idx.ctypes.data_as(u32p), samples.vals, out.ptrs()))


# ----------------------------------------------------------------- host QM31 helpers
def qm_mul(p, q):
    """QM31 product of 4-tuples (a + b i) + (c + d i) u, u^2 = 2 + i."""
    def cm(x, y):
        return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)
    x1, y1, x2, y2 = (p[0], p[1]), (p[2], p[3]), (q[0], q[1]), (q[2], q[3])
    xx, yy = cm(x1, x2), cm(cm(y1, y2), (2, 1))
    a, b = cm(x1, y2), cm(y1, x2)
    return ((xx[0] + yy[0]) % P, (xx[1] + yy[1]) % P, (a[0] + b[0]) % P, (a[1] + b[1]) % P)


def qm_pow(p, e):
    r = (1, 0, 0, 0)
    while e:
        if e & 1:
            r = qm_mul(r, p)
        p = qm_mul(p, p)
        e >>= 1
    return r


def qm_inv(p):
    return qm_pow(p, P ** 4 - 2)


def random_qm31(rng):
    return tuple(int(rng.integers(0, P)) for _ in range(4))


def random_circle_point_qm31(rng):
    """Stereographic point: t -> ((1 - t^2)/(1 + t^2), 2t/(1 + t^2))."""
    while True:
        t = random_qm31(rng)
        t2 = qm_mul(t, t)
        den = ((1 + t2[0]) % P, t2[1], t2[2], t2[3])
        if den == (0, 0, 0, 0):
            continue
        inv = qm_inv(den)
        x = qm_mul(((1 - t2[0]) % P, (-t2[1]) % P, (-t2[2]) % P, (-t2[3]) % P), inv)
        y = qm_mul(tuple(2 * v % P for v in t), inv)
        return x, y


def lde_schedule(log_size, log_blowup, n_cols):
    """(layers per contiguous pass, fused strided passes) in use for this LDE size class."""
    kc, fused = C.c_uint32(), C.c_int()
    _c(lib().stwo_hip_lde_schedule(log_size, log_blowup, n_cols, C.byref(kc), C.byref(fused)))
    return kc.value, bool(fused.value)


class CommitPlan:
    """Batched LDE + Merkle commitment of columns of mixed sizes (stwo_hip_commit_plan_*)."""

    def __init__(self, evals, coeffs, ldes, log_sizes, log_blowup, tw_ptr, itw_ptr, tree_len, autotune=False):
        self.log_sizes = np.ascontiguousarray(log_sizes, np.uint32)
        self.max_log = int(self.log_sizes.max()) + log_blowup
        self.layers = [Blake2sHashVec.uninitialized(1 << k) for k in range(self.max_log + 1)]
        self._keep = (evals, coeffs, ldes)
        h = vp()
        _c(lib().stwo_hip_commit_plan_create(C.byref(h), len(evals), _ptr_array([c.ptr for c in evals]),
                                             _ptr_array([c.ptr for c in coeffs]), _ptr_array([c.ptr for c in ldes]),
                                             self.log_sizes.ctypes.data_as(u32p), log_blowup, tw_ptr, itw_ptr,
                                             tree_len, _ptr_array([l.ptr for l in self.layers]), self.max_log,
                                             1 if autotune else 0))
        self.handle = h.value

    def run(self, use_graph=True):
        _c(lib().stwo_hip_commit_plan_run(self.handle, 1 if use_graph else 0))

    def uses_graph(self):
        """Whether run(True) replays a hipGraph (False: the autotuner found stream launches faster)."""
        v = C.c_int(0)
        _c(lib().stwo_hip_commit_plan_uses_graph(self.handle, C.byref(v)))
        return bool(v.value)

    def root(self):
        out = np.zeros(8, np.uint32)
        _c(lib().stwo_hip_merkle_root(self.layers[0].ptr, out.ctypes.data))
        return out

    def close(self):
        if self.handle:
            lib().stwo_hip_commit_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- proofs
NOT_READY = 5


def eval_at_point_columns(cols, px, py):
    """Values of the polynomials with coefficients `cols` (same size) at the QM31 point (px, py)."""
    log = cols[0].size.bit_length() - 1
    out = (QM31 * len(cols))()
    _c(lib().stwo_hip_eval_at_point_columns(_ptr_array([c.ptr for c in cols]), len(cols), log, qm31(px), qm31(py),
                                            C.cast(out, vp)))
    return [from_qm31(v) for v in out]


def gather_rows(cols, rows):
    """values[i, c] = cols[c][rows[i]] (Column::at for many rows and columns at once)."""
    r = np.ascontiguousarray(rows, np.uint32)
    out = np.zeros((r.size, len(cols)), np.uint32)
    _c(lib().stwo_hip_gather_rows(_ptr_array([c.ptr for c in cols]), len(cols), cols[0].size if cols else 0,
                                  r.ctypes.data_as(u32p), r.size, out.ctypes.data_as(u32p)))
    return out


def merkle_auth_paths(layers, leaves):
    """paths[q, j] = the sibling hash (8 words) at layer max_log - j on the path of leaf leaves[q];
    layers[k] holds the 2^k hashes of layer k."""
    q = np.ascontiguousarray(leaves, np.uint32)
    max_log = len(layers) - 1
    out = np.zeros((q.size, max_log, 8), np.uint32)
    _c(lib().stwo_hip_merkle_auth_paths(_ptr_array([l.ptr for l in layers]), max_log, q.ctypes.data_as(u32p), q.size,
                                        out.ctypes.data_as(u32p)))
    return out


def eval_at_points_columns(cols, points, noalloc=False):
    """[[column c at points[k] for c] for k]: `points` is a list of QM31 circle points (px, py).
    noalloc: the entry point that allocates no device memory (fixed per-device scratch)."""
    log = cols[0].size.bit_length() - 1
    pts = (PointQM31 * len(points))(*[PointQM31(qm31(px), qm31(py)) for px, py in points])
    out = (QM31 * (len(cols) * len(points)))()
    fn = lib().stwo_hip_eval_at_points_columns_noalloc if noalloc else lib().stwo_hip_eval_at_points_columns
    _c(fn(_ptr_array([c.ptr for c in cols]), len(cols), log, C.cast(pts, vp), len(points), C.cast(out, vp)))
    return [[from_qm31(out[k * len(cols) + c]) for c in range(len(cols))] for k in range(len(points))]


def grind(digest, pow_bits):
    d = np.ascontiguousarray(digest, np.uint32)
    out = u64()
    _c(lib().stwo_hip_grind(d.ctypes.data_as(u32p), pow_bits, C.byref(out)))
    return out.value


def wide_fib_params(n_cols, log_n, n_queries=3, pow_bits=10, last_log=2):
    return WideFibParams(n_cols, log_n, n_queries, pow_bits, last_log)


def wide_fib_proof_size(params):
    w, b = u64(), u64()
    _c(lib().stwo_hip_wide_fib_proof_size(C.byref(params), C.byref(w), C.byref(b)))
    return w.value, b.value


def wide_fib_prove(params, seed):
    words, _ = wide_fib_proof_size(params)
    out = np.zeros(words, np.uint32)
    _c(lib().stwo_hip_wide_fib_prove(C.byref(params), seed, out.ctypes.data_as(u32p), words))
    return out


def wide_fib_prove_batch(params, seeds):
    """Proofs of one statement shape for several seeds, proved together (one launch per phase)."""
    words, _ = wide_fib_proof_size(params)
    sd = np.ascontiguousarray(seeds, np.uint64)
    out = np.zeros(words * sd.size, np.uint32)
    _c(lib().stwo_hip_wide_fib_prove_batch(C.byref(params), sd.ctypes.data_as(vp), sd.size, out.ctypes.data_as(u32p),
                                           out.size))
    return out.reshape(sd.size, words)


class ProofServer:
    """Continuous-batching proof server (stwo_hip_server_*).  max_batch > 1: same-shape jobs of at
    most max_batch_work trace elements are proved together as batches."""

    def __init__(self, hbm_budget_bytes, max_inflight=4, work_limit=0, max_batch=1, max_batch_work=0):
        h = vp()
        _c(lib().stwo_hip_server_create(C.byref(h), hbm_budget_bytes, max_inflight))
        self.handle = h.value
        if work_limit:
            _c(lib().stwo_hip_server_limit_work(self.handle, work_limit))
        if max_batch > 1:
            _c(lib().stwo_hip_server_batch(self.handle, max_batch, max_batch_work))

    def submit(self, params, seed):
        jid = u64()
        _c(lib().stwo_hip_server_submit(self.handle, C.byref(params), seed, C.byref(jid)))
        return jid.value

    def wait(self, timeout_us=1000):
        """(job_id, status, proof, (t_submit, t_start, t_end, host_s)) or None after the timeout."""
        jid, st, nw = u64(), C.c_int(), u64()
        if getattr(self, "_buf", None) is None:
            self._buf = np.zeros(1 << 20, np.uint32)
        buf = self._buf
        times = (C.c_double * 4)()
        r = lib().stwo_hip_server_wait(self.handle, timeout_us, C.byref(jid), C.byref(st), buf.ctypes.data_as(u32p),
                                       buf.size, C.byref(nw), times)
        if r == NOT_READY:
            return None
        _c(r)
        return jid.value, st.value, buf[:nw.value].copy(), tuple(times)

    def close(self):
        if self.handle:
            _c(lib().stwo_hip_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def timer_begin():
    _c(lib().stwo_hip_timer_begin())


def timer_end():
    ms = C.c_float()
    _c(lib().stwo_hip_timer_end(C.byref(ms)))
    return ms.value


def kernel_timing(on):
    _c(lib().stwo_hip_kernel_timing(1 if on else 0))


def kernel_stats(name):
    n, ms = u64(), C.c_double()
    _c(lib().stwo_hip_kernel_stats(name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def kernel_stats_reset():
    _c(lib().stwo_hip_kernel_stats_reset())
