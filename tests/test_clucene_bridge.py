"""The CLucene bridge header (exclucene_amd/clucene/gpu_conjunction_scorer.h)
compiles as C++11 and C++17 against the C ABI, and its check program is built
(the GPU run is tests/test_gpu_clucene_bridge.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "clucene_bridge_check.cpp")
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "exclucene_amd", "clucene")]


@pytest.mark.parametrize("std", ["c++11", "c++17"])
def test_header_compiles(tmp_path, std):
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "gpu_conjunction_scorer.h"\n'
        "struct Sim { float coord(int a, int b) { return float(a) / float(b); } };\n"
        "struct Base { explicit Base(Sim*) {} virtual ~Base() {} virtual bool next() = 0;\n"
        "  virtual int32_t doc() const = 0; virtual bool skipTo(int32_t) = 0; virtual float score() = 0; };\n"
        "struct Gpu : exc_clucene::ConjunctionAdapter<Base> { using ConjunctionAdapter::ConjunctionAdapter; };\n"
        "int f(Sim* s, const exc_clucene::ConjunctionBatch& b) { Gpu g(s, b, 0, {}); return g.next() ? g.doc() : -1; }\n")
    r = subprocess.run(["g++", "-std=" + std, "-Wall", "-Werror", "-fsyntax-only"] + INC + [str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_check_program_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only"] + INC + [CHECK],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
