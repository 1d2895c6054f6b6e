"""Regenerate tests/golden/philox_rocrand.json from rocRAND's Philox engine.

Compiles philox_rocrand.cpp (host code only) with hipcc and captures its JSON.
"""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "philox_rocrand")
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-o", exe,
                               os.path.join(HERE, "philox_rocrand.cpp")])
        out = subprocess.check_output([exe]).decode()
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
