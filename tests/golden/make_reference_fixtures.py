"""Extracts the reference's known-answer data for this path into a fixture.

Run in a container where the pbosler/lpm checkout is mounted at
/root/reference:  python tests/golden/make_reference_fixtures.py
Writes tests/golden/reference_fixtures.json.  Only numbers are copied:
  * SWE pair-kernel golden vectors and their evaluation points
    (tests/lpm_swe_kernels_tests.cpp:20-49, 90-99, 122-127);
  * seed polyhedron coordinates (mesh_seeds/icosTriSphereSeed.dat,
    mesh_seeds/cubedSphereSeed.dat);
  * mesh-invariant tolerances (tests/lpm_polymesh_tests.cpp:76-117,
    tests/lpm_mesh_seed_tests.cpp:70-83).
"""
import json
import math
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "reference_fixtures.json")


def _array(src, name):
    m = re.search(r"^\s*const Real %s\[\d+\]\s*=\s*\{(.*?)\};" % re.escape(name),
                  src, re.S | re.M)
    return [float(v) for v in m.group(1).replace("\n", " ").split(",")]


def _commented_array(src, name):
    m = re.search(r"//\s*const Real %s\[9\] = \{(.*?)\};" % re.escape(name),
                  src, re.S)
    body = re.sub(r"//", " ", m.group(1))
    return [float(v) for v in body.replace("\n", " ").split(",")]


def _seed(fn, ncrd):
    lines = open(os.path.join(REF, "mesh_seeds", fn)).read().splitlines()
    return [[float(v) for v in l.split()] for l in lines[1:1 + ncrd]]


def main():
    src = open(os.path.join(REF, "tests", "lpm_swe_kernels_tests.cpp")).read()
    fx = {
        "source": "tests/lpm_swe_kernels_tests.cpp",
        # points: x at (lon, lat) = (pi/4, pi/8); y at (pi/6, pi/20)
        "x_lonlat": [math.pi / 4, math.pi / 8],
        "y_lonlat": [math.pi / 6, math.pi / 20],
        "eps": [0.0, 0.01],
        "kzeta_exact": _array(src, "kzeta_exact"),
        "ksigma_exact": _array(src, "ksigma_exact"),
        "grad_kzeta_exact": _array(src, "grad_kzeta_exact"),
        "grad_kzeta_exact_commented": _commented_array(src, "grad_kzeta_exact"),
        "grad_ksigma_exact": _array(src, "grad_ksigma_exact"),
        # eps = 0.01 tolerances used by the reference test (:152-168)
        "kernel_rel_tol_eps1": 20 * 0.01 ** 2,
        "grad_rel_tol_eps1": 0.12,
    }
    seeds = {
        "icos_crds": _seed("icosTriSphereSeed.dat", 32),
        "cube_crds": _seed("cubedSphereSeed.dat", 14),
    }
    invariants = {
        "icos_area_tol": 31 * 1e-14,
        "cube_area_tol": 3.5 * 1e-14,
        "seed_area_tol": 1e-14,
        "tree_lev": 3,
    }
    json.dump({"swe_kernels": fx, "seeds": seeds, "mesh": invariants},
              open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
