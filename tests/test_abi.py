"""CPU checks of the drop-in boundary: libstwo_hip.so loads without a GPU and
exports every entry point include/stwo_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

import stwo_hip as S


def test_library_is_built_for_gfx950_only():
    assert os.path.exists(S.LIB_PATH), "run __graft_entry__.build() first"
    # the embedded code objects name their target; only gfx950 may appear
    blob = open(S.LIB_PATH, "rb").read()
    assert b"gfx950" in blob
    for other in (b"gfx942", b"gfx90a", b"gfx1100", b"sm_"):
        assert other + b"-" not in blob and other + b":" not in blob


def test_exports_every_declared_symbol():
    names = S.exported_symbols()
    assert len(names) >= 45
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


# Every extern "C" function the reference crate binds (bindings.rs:77-270), in
# declaration order, so that crate links against this library unchanged apart
# from its #[link] name.
REFERENCE_FFI = """copy_uint32_t_vec_from_device_to_host copy_uint32_t_vec_from_host_to_device
copy_uint32_t_vec_from_device_to_device cuda_malloc_uint32_t cuda_malloc_blake_2s_hash
cuda_alloc_zeroes_uint32_t cuda_alloc_zeroes_blake_2s_hash cuda_free_memory bit_reverse_base_field
bit_reverse_secure_field batch_inverse_base_field batch_inverse_secure_field
sort_values_and_permute_with_bit_reverse_order precompute_twiddles interpolate interpolate_columns
evaluate evaluate_columns eval_at_point fold_line fold_circle_into_line accumulate commit_on_first_layer
commit_on_layer_with_previous copy_blake_2s_hash_vec_from_host_to_device
copy_blake_2s_hash_vec_from_device_to_host copy_blake_2s_hash_vec_from_device_to_device
copy_device_pointer_vec_from_host_to_device accumulate_quotients gen_eq_evals
fix_first_variable_base_field fix_first_variable_secure_field
evaluate_wide_fibonacci_constraint_quotients_on_domain""".split()


def test_reference_ffi_names_present():
    assert len(REFERENCE_FFI) == 33
    names = set(S.exported_symbols())
    assert set(REFERENCE_FFI) <= names
    lib = ctypes.CDLL(S.LIB_PATH)
    assert all(hasattr(lib, n) for n in REFERENCE_FFI)


def test_reference_ffi_matches_bindings_order():
    # the header declares part 1 in the order of bindings.rs
    hdr = open(os.path.join(os.path.dirname(S.HERE), "include", "stwo_hip.h")).read()
    part1 = hdr[hdr.index("Part 1: reference FFI"):hdr.index("Part 2: extensions")]
    pos = [re.search(r"\b%s\(" % n, part1).start() for n in REFERENCE_FFI]
    assert pos == sorted(pos)


def test_python_binding_declares_all():
    lib = S.lib()
    for n in S.exported_symbols():
        assert getattr(lib, n).restype is not None or n in (
            "copy_uint32_t_vec_from_device_to_host", "cuda_free_memory", "bit_reverse_base_field",
            "bit_reverse_secure_field", "batch_inverse_base_field", "batch_inverse_secure_field",
            "interpolate", "interpolate_columns", "evaluate", "evaluate_columns", "fold_line",
            "fold_circle_into_line", "accumulate", "commit_on_first_layer", "commit_on_layer_with_previous",
            "copy_blake_2s_hash_vec_from_device_to_host", "copy_blake_2s_hash_vec_from_device_to_device",
            "accumulate_quotients", "evaluate_wide_fibonacci_constraint_quotients_on_domain",
            "gen_eq_evals", "fix_first_variable_base_field", "fix_first_variable_secure_field",
            "stwo_hip_clear_error")
