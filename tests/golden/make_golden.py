"""Regenerates tests/golden/std_random.json by compiling and running
std_random.cpp against the C++ standard library (g++/libstdc++).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "std_random")
        subprocess.run(["g++", "-O0", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "std_random.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    with open(os.path.join(HERE, "std_random.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
