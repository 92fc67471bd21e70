"""Regenerates tests/golden/ref_vectors.json from the reference's own headers.

Builds oracle/_ref/ref_vectors (oracle/ref.mk compiles oracle/ref_vectors.cpp
against include/crystal/{pred,load,utils}.h where they lie in the reference
tree) and stores its JSON output.  Needs the reference tree; the stored JSON
lets the tests run without it.
"""
import json
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[2]


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_vectors")], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    (ROOT / "tests" / "golden" / "ref_vectors.json").write_text(json.dumps(data, separators=(",", ":")) + "\n")


if __name__ == "__main__":
    main()
