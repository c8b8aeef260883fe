"""Writes the small committed fixtures of tests/golden/.

generator.json: the first 64 SplitMix64 outputs for every seed the
workloads use, computed here with Python integers (independently of the C
oracle) from the published finaliser:
  z = seed + 0x9E3779B97F4A7C15*(i+1) mod 2^64
  z = (z ^ z>>30) * 0xBF58476D1CE4E5B9 ; z = (z ^ z>>27) * 0x94D049BB133111EB
  z ^= z>>31
reference_kats.json: the known answers of the reference's own tests
(test/*.cpp), restated as data.
"""
import json
import os

M = (1 << 64) - 1
SEEDS = [0x1001, 0x1002, 0x1003, 0x2001, 0x2002, 0x2003, 0x2004, 0x3001, 0x3002,
         0x3003, 0x4001, 0x4002, 0x4003, 0x4004, 0x5001, 0x5002, 0x5003, 0x5004,
         0x5005, 1234567]


def splitmix64(seed, i):
    z = (seed + 0x9E3779B97F4A7C15 * (i + 1)) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    gen = {"u64": {hex(s): [str(splitmix64(s, i)) for i in range(64)] for s in SEEDS}}
    with open(os.path.join(here, "generator.json"), "w") as f:
        json.dump(gen, f, indent=0)
    kats = {
        "reduce": {"float_0_to_99": 4950, "int_0_to_1586": 1258491,
                   "int_ones_1e8": 100000000,
                   "float_ones_sizes": [[1, 255], [8192, 8447]]},
        "scan": {"int_ones_exclusive_1024": "i", "int_ones_exclusive_10": "i"},
        "sha1": {"": "da39a3ee5e6b4b0d3255bfef95601890afd80709",
                 "The quick brown fox jumps over the lazy dog":
                     "2fd4e1c67a2d28fced849ee1bb76e7391b93eb12",
                 "The quick brown fox jumps over the lazy cog":
                     "de9f2c7fd25e1b3afad3e85a0bd17d9b100db4b3"},
        "sources": {"reduce": "test/ReduceTests.cpp:77-170",
                    "scan": "test/ScanTests.cpp:71-107",
                    "sha1": "test/SHA1Tests.cpp:58-77"},
    }
    with open(os.path.join(here, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    # batched configuration: 4096 segment lengths floor(10^(3+3d)),
    # d = (z>>11)*2^-53 (seed 0x5001); 512 product shapes M,N,K =
    # 64*(1+U{0..15}), U = ((z>>32)*16)>>32 from seed 0x5003 (3 draws each)
    import math
    lens = [int(math.floor(10.0 ** (3.0 + 3.0 * ((splitmix64(0x5001, i) >> 11) * 2.0 ** -53))))
            for i in range(4096)]
    u = lambda i: ((splitmix64(0x5003, i) >> 32) * 16) >> 32
    shapes = [[64 * (1 + u(3 * i)), 64 * (1 + u(3 * i + 1)), 64 * (1 + u(3 * i + 2))]
              for i in range(512)]
    with open(os.path.join(here, "batched_shapes.json"), "w") as f:
        json.dump({"segment_lengths": lens, "gemm_shapes": shapes,
                   "note": "gemm_shapes are [M, N, K]"}, f)


if __name__ == "__main__":
    main()
