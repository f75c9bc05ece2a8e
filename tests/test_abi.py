"""CPU checks of the drop-in boundary: libstwo_hip.so loads without a GPU and
exports every entry point include/stwo_hip.h declares (no compute calls)."""
import ctypes
import os

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


def test_reference_ffi_names_present():
    # every extern "C" function of the reference's bindings.rs that is on the hot path
    ref = """copy_uint32_t_vec_from_device_to_host copy_uint32_t_vec_from_host_to_device
    copy_uint32_t_vec_from_device_to_device cuda_malloc_uint32_t cuda_malloc_blake_2s_hash
    cuda_alloc_zeroes_uint32_t cuda_alloc_zeroes_blake_2s_hash cuda_free_memory bit_reverse_base_field
    bit_reverse_secure_field batch_inverse_base_field batch_inverse_secure_field
    sort_values_and_permute_with_bit_reverse_order precompute_twiddles interpolate interpolate_columns
    evaluate evaluate_columns eval_at_point fold_line fold_circle_into_line accumulate commit_on_first_layer
    commit_on_layer_with_previous copy_blake_2s_hash_vec_from_host_to_device
    copy_blake_2s_hash_vec_from_device_to_host copy_blake_2s_hash_vec_from_device_to_device
    copy_device_pointer_vec_from_host_to_device accumulate_quotients
    evaluate_wide_fibonacci_constraint_quotients_on_domain""".split()
    names = set(S.exported_symbols())
    assert set(ref) <= names


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
            "stwo_hip_clear_error")
