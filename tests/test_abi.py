"""CPU tests of the drop-in boundary: record layout of include/galaxy_bh.h against
the reference's frozen sizes and 144-byte fixture (runtime/src/nbody_gpu.rs:1441-1489),
and that libgalaxy_bh.so loads and exports every function the header declares."""
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from galaxy_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")))

PROBE = r"""
#include <stdio.h>
#include <string.h>
#include "galaxy_bh.h"
int main(void) {
    printf("BhSettings %zu\nGpuBody %zu\nGpuEntry %zu\nGpuCell %zu\nGpuAccel %zu\n",
           sizeof(GbhBhSettings), sizeof(GbhGpuBody), sizeof(GbhGpuEntry), sizeof(GbhGpuCell),
           sizeof(GbhGpuAccel));
    printf("EvolveSettings %zu\nGpuEvolveState %zu\nTreeSettings %zu\nTreeMeta %zu\n",
           sizeof(GbhEvolveSettings), sizeof(GbhEvolveState), sizeof(GbhTreeSettings),
           sizeof(GbhTreeMeta));
    GbhBhSettings s = {{1, 2, 3, 4}, {0.5f, 0.25f, 1.0f, 0.0f}};
    GbhGpuBody b = {{1.0f, -2.0f, 3.0f, 0.0f}, {4, 5, 6, 7}};
    GbhGpuEntry e = {{0x01234567u, 8, 9, 10}};
    GbhGpuCell c = {{1.0f, 2.0f, 3.0f, 4.0f}, {5.0f, 6.0f, 0.0f, 0.0f}, {7, 8, 9, 0},
                    {10, 11, 12, GBH_NO_CHILD}};
    unsigned char buf[144];
    memcpy(buf, &s, 32); memcpy(buf + 32, &b, 32); memcpy(buf + 64, &e, 16); memcpy(buf + 80, &c, 64);
    printf("hex ");
    for (int i = 0; i < 144; ++i) printf("%02x", buf[i]);
    printf("\n");
    return 0;
}
"""


def test_header_records_match_reference_fixture():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write(PROBE)
        subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    vals = dict(line.split(" ", 1) for line in out if line)
    for name in ["BhSettings", "GpuBody", "GpuEntry", "GpuCell", "GpuAccel", "EvolveSettings",
                 "GpuEvolveState", "TreeSettings", "TreeMeta"]:
        assert int(vals[name]) == GOLD["record_sizes"][name], name
    assert vals["hex"] == GOLD["fixture_144"]["hex"]


def test_numpy_record_views_match_fixture():
    f = GOLD["fixture_144"]
    s = np.array(f["settings"]["info"], np.uint32).tobytes() + np.array(f["settings"]["physics"], np.float32).tobytes()
    b = np.array(f["body"]["position_mass"], np.float32).tobytes() + np.array(f["body"]["meta"], np.uint32).tobytes()
    e = np.array(f["entry"]["data"], np.uint32).tobytes()
    c = f["cell"]
    cb = (np.array(c["center_half_mass"] + c["com"], np.float32).tobytes()
          + np.array(c["range_depth"] + c["children"], np.uint32).tobytes())
    assert (s + b + e + cb).hex() == f["hex"]


def test_ctypes_structs_sizes():
    assert C.sizeof(L.GbhConfig) == 32
    assert C.sizeof(L.StageTiming) == 24
    assert C.sizeof(L.BuildTiming) == 48
    assert C.sizeof(L.SimParams) == 40
    assert C.sizeof(L.Tuning) == 24


def test_library_exports_every_header_function():
    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libgalaxy_bh.so is not built (run __graft_entry__.build())")
    lib = L.lib()
    names = L.header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if " T " in line}
    assert set(names) <= exported


def test_host_side_validation_without_gpu():
    """Entry points that fail before touching the GPU return the reference's messages."""
    lib = L.lib()
    st = np.zeros((4, 8), np.float32)
    out = C.c_void_p()
    assert lib.gbh_state_create(None, st.ctypes.data_as(C.c_void_p), 4, C.byref(out), None) != 0
    assert b"null" in lib.gbh_last_error()
    bodies = np.zeros((1, 5))
    assert lib.gbh_ic_reference_disc(1, 303, 5e10, 3.0, bodies.ctypes.data_as(C.c_void_p)) != 0
    assert lib.gbh_last_error() == b"particle count must be in 2..=1000000"
    assert lib.gbh_ic_reference_collision(3, 303, bodies.ctypes.data_as(C.c_void_p)) != 0
    assert lib.gbh_last_error() == b"collision particle count must be in 4..=1000000"
    assert lib.gbh_sparse_layout_bytes(65536) < 128 * 1024 * 1024 + 65536 * 200
