"""The C++ host facade (include/dunefem_amd.hh) and its driver
tests/cpp/inverseoperatortest.cc (after the reference's
dune/fem/solver/test/inverseoperatortest.cc)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_bin", "inverseoperatortest")


def test_facade_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cc"
    src.write_text('#include "dunefem_amd.hh"\nint main() { dunefem_amd::SolverParameter p; '
                   'p.set("fem.solver.method", "fgmres"); return p.solverMethod() == FEM_SOLVER_FGMRES ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_driver_is_built_and_links_the_library():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True, check=True).stdout
    line = [l for l in out.splitlines() if "libdunefem_amd.so" in l]
    assert line and "not found" not in line[0], out


@pytest.mark.gpu
def test_inverseoperatortest_driver_passes():
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and p.stdout.strip().endswith("PASS"), p.stdout + p.stderr
