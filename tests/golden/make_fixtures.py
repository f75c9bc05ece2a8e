"""Extract the numeric fixtures held by the reference's own Rust tests.

Run in the build container only (it reads the reference checkout as text):
    python tests/golden/make_fixtures.py /root/reference
Writes tests/golden/reference_fixtures.json (numbers only).

Sources (file:line in the reference repository):
  stwo_gpu_backend/src/poly.rs:291-327    test_interpolate_2 input values (log 5)
  stwo_gpu_backend/src/poly.rs:398-595    test_interpolate_from_fib: domain, values,
                                          twiddle / itwiddle tree of root coset log 6
  stwo_gpu_backend/src/fri.rs:312-937     test_fold_circle_into_line_fibonacci: dst
                                          (4 x 32), src (4 x 64), alpha, domains, tree
Only poly.rs:451-595 (the log-6 twiddle tree) stores reference *outputs*; the
other blocks are inputs whose expected values the reference computes at run time.
"""
import json
import os
import re
import sys


def fn_body(text: str, name: str) -> str:
    start = text.index(f"fn {name}(")
    nxt = text.find("#[test", start)
    return text[start: nxt if nxt >= 0 else len(text)]


def vec_blocks(body: str):
    out = []
    for m in re.finditer(r"vec!\[(.*?)\]", body, flags=re.S):
        nums = re.findall(r"(?:M31|BaseField::from)\((\d+)\)", m.group(1))
        if nums:
            out.append([int(x) for x in nums])
    return out


def coset_params(body: str):
    idx = [int(x) for x in re.findall(r"CirclePointIndex\((\d+)\)", body)]
    logs = [int(x) for x in re.findall(r"log_size:\s*(\d+)", body)]
    return idx, logs


def main(ref_root: str) -> None:
    poly = open(os.path.join(ref_root, "stwo_gpu_backend/src/poly.rs")).read()
    fri = open(os.path.join(ref_root, "stwo_gpu_backend/src/fri.rs")).read()

    b = vec_blocks(fn_body(poly, "test_interpolate_2"))
    interp2_values = b[0]

    body = fn_body(poly, "test_interpolate_from_fib")
    values, twiddles, itwiddles = vec_blocks(body)
    idx, logs = coset_params(body)
    interp_fib = {
        "domain_half_coset": {"initial_index": idx[0], "step_size": idx[1], "log_size": logs[0]},
        "values": values,
        "root_coset": {"initial_index": idx[2], "step_size": idx[3], "log_size": logs[1]},
        "twiddles": twiddles,
        "itwiddles": itwiddles,
    }

    body = fn_body(fri, "test_fold_circle_into_line_fibonacci")
    blocks = vec_blocks(body)
    idx, logs = coset_params(body)
    alpha = [int(x) for x in re.findall(r"M31\((\d+)\)", re.search(r"QM31::from_m31\((.*?)\);", body, re.S).group(1))]
    fold_fib = {
        "dst": blocks[0:4],
        "dst_line_coset": {"initial_index": idx[0], "step_size": idx[1], "log_size": logs[0]},
        "src": blocks[4:8],
        "src_half_coset": {"initial_index": idx[2], "step_size": idx[3], "log_size": logs[1]},
        "alpha": alpha,
        "root_coset": {"initial_index": idx[4], "step_size": idx[5], "log_size": logs[2]},
        "twiddles": blocks[8],
        "itwiddles": blocks[9],
    }
    out = {
        "interpolate_2_values": interp2_values,
        "interpolate_from_fib": interp_fib,
        "fold_circle_into_line_fibonacci": fold_fib,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
