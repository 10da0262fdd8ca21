"""Writes tests/golden/wire_golden.json from the reference's own serializer.

Builds oracle/_ref/wire_ref (oracle/Makefile `ref` target, compiled from the
reference's topo_opt/fem_interface.h + serialization.h) and records SHA-256
digests and sizes of:
  * the FEMInputs stream of the interface test fixture
    (topo_opt/interface_test.cpp:45-166) at resolution / scale, scale in {1,4,8}
  * the FEMOutputs stream of the interface test's fake solver on it
Run in a container where the reference sources are mounted.
"""
import hashlib
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "wire_ref")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for scale in (1, 4, 8):
            fx = os.path.join(td, f"fx{scale}.tcb")
            r = subprocess.run([REF, "fixture", fx, str(scale)], check=True, capture_output=True, text=True)
            data = open(fx, "rb").read()
            fo = os.path.join(td, f"fake{scale}.bin")
            subprocess.run([REF, "fake", fx, fo], check=True)
            fdata = open(fo, "rb").read()
            out[f"scale{scale}"] = {
                "inputs_sha256": hashlib.sha256(data).hexdigest(),
                "inputs_bytes": len(data),
                "blocks": int(r.stdout.strip().split("=")[1]),
                "fake_outputs_sha256": hashlib.sha256(fdata).hexdigest(),
                "fake_outputs_bytes": len(fdata),
            }
    with open(os.path.join(HERE, "wire_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
