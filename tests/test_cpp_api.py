"""The C++ API surface (include/gascoigne_hip.hpp): the header compiles on its
own (CPU), and the parity program tests/cpp/api_solve -- LinearMg, Gmres and
Solve through the C++ classes, checked against the oracle -- passes (GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "api_solve")
CONFIGS = os.path.join(ROOT, "gascoignelib_amd", "configs")


def test_header_compiles(tmp_path):
    src = tmp_path / "t.cc"
    src.write_text('#include "gascoigne_hip.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I" + os.path.join(ROOT, "include"), str(src)], check=True)


CASES = [
    ("square.inp", 2, 6, 1, "f64", "gmres", "none"),
    ("cube.inp", 3, 4, 1, "f64", "gmres", "none"),
    ("cube.inp", 3, 4, 200, "f64", "gmres", "lexicographic"),
    ("cube.inp", 3, 5, 1000, "f16", "fgmres", "lexicographic"),
    ("cube.inp", 3, 4, 1, "f32", "fgmres", "none"),
    ("cube.inp", 3, 4, 1, "f64", "fgmres", "lexicographic"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_cpp_api_solve(case):
    inp, dim, pre, ndirect, prec, krylov, order = case
    print(f"[api_solve {case}]", flush=True)
    assert os.path.exists(BIN), "tests/cpp/api_solve is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, os.path.join(CONFIGS, inp), str(dim), str(pre), str(ndirect), prec,
                        krylov, order], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout
