"""The C-ABI library: loads without a GPU, exports every declared symbol, and
its headers carry only C types."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import graphpi_amd as gp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graphpi_amd.h")


def _declared():
    text = open(HEADER).read()
    return re.findall(r"GP_API\s+[\w\s\*]+?\b(gp_\w+)\s*\(", text)


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 17
    lib = ctypes.CDLL(gp.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", gp.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gp_\w+)", out))
    assert set(names) <= exported


def test_header_is_plain_c():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    for bad in ("torch", "at::", "std::", "class ", "template"):
        assert bad not in text


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "graphpi_amd.h"\nint main(void){ gp_sched_flat s; (void)s; return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-c", str(src), "-o", str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_device_calls_without_gpu_fail_cleanly():
    if gp.device_count() > 0:
        pytest.skip("a GPU is present")
    h = gp.HostGraph.complete(4)
    with pytest.raises(gp.GPError) as e:
        gp.DeviceGraph(h)
    assert e.value.code == -6


def test_cpp_api_drop_in_compiles(tmp_path):
    # a reference-style driver (tianhe/baseline_test.cpp shape) builds against
    # include/*.h and links against the library
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include "graph.h"
#include "dataloader.h"
#include "pattern.h"
#include "schedule.h"
#include "common.h"
#include "motif_generator.h"
#include <cstdio>
int main(int argc, char** argv) {
    Graph* g; DataLoader D;
    if (!D.load_complete(g, 6)) return 1;
    Pattern p(House);
    bool valid;
    Schedule s(p, valid, 1, 1, true, g->v_cnt, g->e_cnt, 20);
    if (!valid) return 2;
    MotifGenerator mg(4);
    std::printf("%d %zu\n", s.get_total_prefix_num(), mg.generate().size());
    if (argc > 1) std::printf("ans %lld\n", g->pattern_matching(s, 24));
    delete g;
    return 0;
}
''')
    exe = tmp_path / "driver"
    r = subprocess.run(["g++", "-std=c++11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        gp.LIB_PATH, "-Wl,-rpath," + os.path.dirname(gp.LIB_PATH)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split()[1] == "6"


@pytest.mark.parametrize("vertex,edge", [
    ([0, 2, 1, 3], [1, 2, 0]),          # offsets go backwards
    ([0, 1, 2, 4], [1, 0, 0, 7]),       # neighbour id out of range
    ([0, 2, 3, 4], [2, 1, 0, 0]),       # unsorted list
    ([0, 2, 3, 4], [0, 1, 0, 0]),       # self-loop
    ([0, 2, 3, 5], [1, 2, 0, 0]),       # vertex[v_cnt] != e_cnt
    ([0, 5, 5, 4], [1, 2, 0, 0]),       # offset beyond e_cnt
])
def test_upload_rejects_malformed_csr(vertex, edge):
    # the check runs before any device call, so it is the same with or without a GPU
    v = np.asarray(vertex, dtype=np.uint32)
    e = np.asarray(edge, dtype=np.int32)
    h = ctypes.c_void_p()
    rc = gp.lib().gp_graph_upload(e.ctypes.data, v.ctypes.data, 3, 4, ctypes.byref(h))
    assert rc == -1, gp.lib().gp_last_error()
    assert not h.value


def test_range_plan_cpu(tmp_path):
    # gp_count's launch plan (task ranges, slice sizing, time limit) with
    # made-up kernel times: tests/cpp/range_plan_test.cpp
    exe = tmp_path / "range_plan_test"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "graphpi_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "range_plan_test.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
