"""Writes the fixtures under tests/golden/ (run in the development container,
where the KGMT reference checkout is mounted read-only at /root/reference).

  obstacles/<set>.csv   copies of the reference's obstacle data files
                        (include/config/obstacles/<set>/obstacles.csv), byte
                        for byte, with their box counts in obstacles.json
  philox_kat.json       known-answer vectors of Philox4x32-10 from the
                        Random123 distribution (kat_vectors), as published

The data files are inputs, not code; nothing here executes reference code.
"""
import json
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/include/config/obstacles"
SETS = ["narrowPassage", "slot", "house", "pillars", "trees", "quadTrees", "quadNarrowPassage", "empty"]

# Random123 kat_vectors, "philox4x32 10" lines: counter(4) key(2) -> output(4)
PHILOX_KAT = [
    [[0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]],
    [[0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]],
    [[0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]],
]


def main():
    out = os.path.join(HERE, "obstacles")
    os.makedirs(out, exist_ok=True)
    counts = {}
    for s in SETS:
        src = os.path.join(REF, s, "obstacles.csv")
        dst = os.path.join(out, f"{s}.csv")
        shutil.copyfile(src, dst)
        nums = open(src).read().replace(",", " ").split()
        counts[s] = len(nums) // 6
    with open(os.path.join(out, "obstacles.json"), "w") as f:
        json.dump({"source": "include/config/obstacles/<set>/obstacles.csv", "boxes": counts}, f, indent=1)
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors (philox4x32, 10 rounds)", "vectors": PHILOX_KAT}, f, indent=1)


if __name__ == "__main__":
    main()
