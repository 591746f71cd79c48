"""Regenerate tests/golden/generators_sha256.json: the seeded generators'
output arrays hashed with SHA-256 (integer arrays as little-endian int32;
floating-point arrays cast to little-endian float32, so that last-bit
differences of libm between hosts do not change the hash).

Run from the repository root: python tests/golden/make_generator_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def digest(mesh):
    out = {"nnodes": int(mesh.nnodes), "nelems": int(mesh.nelems),
           "nbcs": int(len(mesh.bcs))}
    for name, arr, kind in (("conn", mesh.conn, "<i4"),
                            ("bcs", mesh.bcs, "<i4"),
                            ("xpts", mesh.xpts, "<f4"),
                            ("data", mesh.data, "<f4"),
                            ("loads", mesh.loads, "<f4")):
        b = np.ascontiguousarray(arr, kind).tobytes()
        out[name] = hashlib.sha256(b).hexdigest()
    return out


def cases():
    from gpu_fem_amd import meshes as M
    return {
        # the workloads at full size (seed 0) and the parity meshes
        "plate10_clamped_uniform": lambda: M.plate(
            10, 10, 1.0, 1.0, thick=0.01, bc="clamped", load="uniform"),
        "plate400_ss_sine": lambda: M.plate(
            400, 400, 1.0, 1.0, E=70e9, nu=0.3, thick=0.005, bc="ss",
            load="sine", load_mag=1e3),
        "cylinder256x512_seed0": lambda: M.cylinder(256, 512, seed=0),
        "wingbox_full_seed0": lambda: M.wingbox(seed=0),
        "plate24_ss_sine": lambda: M.plate(24, 24, thick=0.005, bc="ss",
                                           load="sine"),
        "cylinder24x16_seed11": lambda: M.cylinder(24, 16, seed=11),
        "wingbox_small_seed5": lambda: M.wingbox(nc=10, nh=3, nsb=2, nribs=4,
                                                 seed=5),
    }


def served_cases():
    from gpu_fem_amd import meshes as M
    return M.served_requests(8, seed=0)


def main():
    data = {name: digest(mk()) for name, mk in cases().items()}
    data["served_requests_8_seed0"] = [
        dict(digest(m), name=m.name) for m in served_cases()]
    with open(os.path.join(HERE, "generators_sha256.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("written", os.path.join(HERE, "generators_sha256.json"))


if __name__ == "__main__":
    main()
