"""Writes tests/golden/reference_kats.json: known-answer data of the GALAXY reference tests.

The values are transcribed from the reference's own unit tests (read as text):
  runtime/src/nbody_gpu.rs:1441-1451, nbody_parallel.rs:1000-1004  record sizes
  runtime/src/nbody_gpu.rs:1453-1489                                144-byte transfer fixture
  nbody/src/flat.rs:445-457                                         flat record sizes, Morton KATs
  nbody/src/flat.rs:460-506                                         max_depth=1 and coincident-body order
  nbody/src/flat.rs:540-564                                         Morton-range membership fixture
  nbody/src/lib.rs:564-581, flat.rs:509-589                         accuracy gates
"""
import json
import os

KATS = {
    "_source": __doc__.strip().splitlines()[0],
    "record_sizes": {"BhSettings": 32, "GpuBody": 32, "GpuEntry": 16, "GpuCell": 64, "GpuAccel": 16,
                     "EvolveSettings": 32, "GpuEvolveState": 32, "TreeSettings": 32, "TreeMeta": 32,
                     "BoundsSettings": 16, "BoundsRecord": 16, "RadixSettings": 16,
                     "MortonEntry": 8, "FlatCell": 80},
    "fixture_144": {
        "settings": {"info": [1, 2, 3, 4], "physics": [0.5, 0.25, 1.0, 0.0]},
        "body": {"position_mass": [1.0, -2.0, 3.0, 0.0], "meta": [4, 5, 6, 7]},
        "entry": {"data": [0x01234567, 8, 9, 10]},
        "cell": {"center_half_mass": [1.0, 2.0, 3.0, 4.0], "com": [5.0, 6.0, 0.0, 0.0],
                 "range_depth": [7, 8, 9, 0], "children": [10, 11, 12, 0xFFFFFFFF]},
        "hex": ("010000000200000003000000040000000000003f0000803e0000803f00000000"
                "0000803f000000c0000040400000000004000000050000000600000007000000"
                "6745230108000000090000000a0000000000803f000000400000404000008040"
                "0000a0400000c040000000000000000007000000080000000900000000000000"
                "0a0000000b0000000c000000ffffffff"),
    },
    "morton2": [[0, 0, 0], [1, 0, 1], [0, 1, 2], [1, 1, 3]],
    "morton2_order": {"a": [65535, 0], "b": [0, 65535], "a_less_than_b": True},
    "coincident": {"count": 8, "body": [0.0, 0.0, 0.0, 0.0, 1.0], "bucket": 1, "max_depth": 48,
                   "softening": 0.1, "order": list(range(8)), "tree_max_depth": 16},
    "max_depth_one": {"count": 8, "bucket": 1, "max_depth": 1, "softening": 0.1, "tree_max_depth": 1},
    "membership": {"bodies": [[0.0, 0.0, 0.0, 0.0, 1.0], [2.220446049250313e-16, 0.0, 0.0, 0.0, 2.0],
                              [1.0, 1.0, 0.0, 0.0, 3.0], [-1.0, -1.0, 0.0, 0.0, 4.0]],
                   "theta": 2.0, "bucket": 1, "max_depth": 16, "softening": 0.01,
                   "max_relative_below": 0.5},
    "gates": {"theta0_vs_direct": 2e-12, "theta_half_rms": 0.03, "theta_half_max": 0.25,
              "flat_vs_recursive": 1e-9, "root_mass": 5.0e10, "root_mass_tol": 1e-3},
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print(out)
