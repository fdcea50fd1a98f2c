"""Writes tests/golden/golden.json.

Hand-entered values come from the reference's own text: the transition-count table
(src/dp.jl:34-37), the plan-size table (src/dp.jl:73-78), the docstring examples
(src/hafnian.jl:1103-1107, 1182-1190) and the K=4 plan derived by hand from
src/dp.jl:248-304 and src/unrolled.jl:60.  Generated values come from the pure-Python
restatement in oracle/oracle.py (plan hashes) and from brute-force matching enumeration over
Python integers (exact hafnians); neither uses the C oracle or the product library.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


def splitmix(state):
    while True:
        state = (state + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


def small_int_matrix(N, seed, lo=-3, hi=3):
    g = splitmix(seed)
    A = [[0] * N for _ in range(N)]
    for i in range(N):
        for j in range(i + 1, N):
            v = lo + next(g) % (hi - lo + 1)
            A[i][j] = A[j][i] = v
    return A


def main():
    gold = {}
    # src/dp.jl:34-37
    gold["reference_ntrans"] = {"12": 1076, "16": 10226, "20": 89665, "24": 748776,
                                "28": 6052062, "32": 47786401}
    # src/dp.jl:73-78 (MB = MiB of (ntrans + nstates + 1) Int32 words)
    gold["reference_plan_mib"] = {"24": 3.1, "26": 8.9, "28": 25.0, "30": 70.2, "32": 195.7}
    # hand-derived K = 4 plan; bit i-1 of a mask is index i
    gold["k4_plan"] = {
        "pairs": {"1,2": 1, "1,3": 2, "1,4": 3, "2,3": 4, "2,4": 5, "3,4": 6},
        "starts": [1, 1, 2, 3, 4, 7],
        "trans": [516, 517, 518, 2049, 1538, 1027],
        "levels": [1, 2, 5, 6],
    }
    # src/hafnian.jl:1103-1107 and 1182-1190
    gold["docstring"] = {"hafnian_4x4": [[[0, 1, 2, 3], [1, 0, 4, 5], [2, 4, 0, 6], [3, 5, 6, 0]], 28],
                         "hafnian_repeated": [[[1, 2], [2, 3]], [2, 2], 11]}
    # plan hashes from the pure-Python restatement of src/dp.jl:248-304
    gold["plan_sha256"] = {}
    for K in range(2, 25, 2):
        for ib in (4, 8):
            gold["plan_sha256"][f"{K}_{ib}"] = oracle.py_build_plan(K, ib).sha256()
    # exact hafnians of small integer matrices by brute-force enumeration
    gold["exact"] = []
    for N, seed in ((4, 1), (6, 2), (8, 3), (10, 4), (12, 5)):
        A = small_int_matrix(N, 1000 + seed)
        gold["exact"].append({"A": A, "haf": oracle.brute_hafnian(A)})
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(gold, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
