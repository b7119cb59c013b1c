"""Writes tests/golden/philox_rocrand.json: Philox4x32-10 outputs computed by
rocRAND's own implementation (rocrand_philox4x32_10.h, MIT licence, from the
ROCm install), compiled here as host code. An independent check of the
oracle's and the kernels' Philox.

Run once in a container with ROCm: python tests/golden/make_philox_fixture.py
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r'''
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  Probe p;
  unsigned c[4], k[2];
  while (scanf("%u %u %u %u %u %u", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
    uint4 r = p.run(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
    printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
'''


def main():
    rng = np.random.default_rng(20261203)
    cases = [[0, 0, 0, 0, 0, 0], [0xffffffff] * 6]
    cases += rng.integers(0, 2 ** 32, size=(62, 6), dtype=np.uint64).tolist()
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "probe.cpp")
        exe = os.path.join(td, "probe")
        with open(src, "w") as fh:
            fh.write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", src, "-o", exe], check=True)
        inp = "\n".join(" ".join(str(int(v)) for v in c) for c in cases) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    outs = [[int(v) for v in line.split()] for line in out.strip().splitlines()]
    data = [{"counter": c[:4], "key": c[4:], "out": o} for c, o in zip(cases, outs)]
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as fh:
        json.dump({"source": "rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)", "cases": data}, fh, indent=0)


if __name__ == "__main__":
    main()
