"""Writes tests/golden/philox_kat.json: Philox4x32-10 blocks computed by rocRAND's host
implementation (rocrand_device::philox4x32_10_engine, /opt/rocm/include/rocrand), for
checking the oracle's and the renderer's block function at counters that exercise all four
counter words and both key words.

rocRAND's engine(seed, subsequence, offset) starts at the 128-bit counter
subsequence * 2^64 + offset / 4 with key (seed low word, seed high word); its first four
outputs are the block at that counter. Run from the repository root:
    python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_kernel.h>
#include <cstdio>
int main(int argc, char **argv) {
  for (int i = 1; i + 2 < argc; i += 3) {
    unsigned long long seed = strtoull(argv[i], 0, 0), sub = strtoull(argv[i + 1], 0, 0),
                       off = strtoull(argv[i + 2], 0, 0);
    rocrand_device::philox4x32_10_engine e(seed, sub, off);
    for (int k = 0; k < 4; ++k) printf("%u ", e());
    printf("\n");
  }
  return 0;
}
"""

CASES = [
    (0, 0, 0),
    (0xFFFFFFFFFFFFFFFF, 0, 0),
    (0x299F31D0A4093822, 0, 0),
    (12345, 0, 4),                      # counter (1, 0, 0, 0)
    (12345, 0, 4 * 7),                  # counter (7, 0, 0, 0)
    (7, 0, 4 * (5 << 32)),              # counter (0, 5, 0, 0)
    (7, 0, 4 * ((9 << 32) + 3)),        # counter (3, 9, 0, 0)
    (0xDEADBEEF, 11, 0),                # counter (0, 0, 11, 0)
    (0xDEADBEEF, (13 << 32) + 2, 8),    # counter (2, 0, 2, 13)
    (0x1234567890ABCDEF, 65535, 4 * ((262143 << 32) + 17)),  # counter (17, 262143, 65535, 0)
]


def counter_words(sub, off):
    c = (sub << 64) + off // 4
    return [(c >> (32 * i)) & 0xFFFFFFFF for i in range(4)]


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "kat.cpp")
        exe = os.path.join(d, "kat")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "-O1", src, "-o", exe], check=True)
        args = [str(v) for case in CASES for v in case]
        lines = subprocess.run([exe, *args], check=True, capture_output=True, text=True).stdout.strip().splitlines()
    out = []
    for (seed, sub, off), line in zip(CASES, lines):
        out.append({"counter": counter_words(sub, off), "key": [seed & 0xFFFFFFFF, seed >> 32],
                    "out": [int(v) for v in line.split()]})
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump({"source": "rocrand_device::philox4x32_10_engine (ROCm 7.2 headers)", "blocks": out}, f, indent=1)
    print(f"wrote {len(out)} blocks")


if __name__ == "__main__":
    main()
