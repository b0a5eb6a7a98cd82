"""Writes tests/golden/ref_vectors.json from the reference's own per-call classes.

Builds oracle/_ref/ref_vectors (oracle/Makefile target `ref`, which compiles
oracle/ref_vectors.cpp against the reference checkout's headers) and runs it.
Run in a container that holds the reference checkout:

    python tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={args.ref}"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_vectors")], check=True,
                         capture_output=True, text=True).stdout
    data = json.loads(out)
    data["_meta"] = {
        "generator": "oracle/ref_vectors.cpp compiled against the reference's src/ headers (asserts enabled)",
        "classes": ["CoordFrame3", "Mapping", "CdfTable", "LambertDiffuseBsdf", "OrenNayarBsdf", "PhongBsdf",
                    "MixedBsdf", "TwoSidedBsdf", "DiffuseAreaLight", "TriangleMeshGeometry", "ThinlensCamera",
                    "EquiRectRep", "Fimage"],
    }
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print({k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main()
