"""Extract the reference's streaming golden vectors into small .npy fixtures.

Run in a container where the reference tree is mounted (default
/root/reference); the fixtures it writes are committed and the tests read only
those.  Sources (data only, nothing is executed):

  tdlbcpp/tests/StreamingSimpleTest.hpp      expected Q after streamingPull()
  tdlbcpp/tests/StreamingSimplePushTest.hpp  expected Q (and F) after streamingPush()

Both hold a 3x3x3 interior with one ghost layer (5x5x5 cells x 18 slots),
filled by ParamsCommon::fillForTest (tdlbcpp/tests/Params/ParamsCommon.hpp:167-206)
with integer tags ((i*100+j)*100+k)*100+l before streaming.

Output: tests/golden/streaming_{pull,push}_Q.npy  uint64 [5,5,5,18] (i,j,k,slot)
        tests/golden/streaming_push_F.npy         uint64 [5,5,5,3]
        tests/golden/input_debug_gridx60_numSteps20.json        the reference's sample
        tests/golden/input_debug_gridx72_IncreaseResolution.json input files (data)
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("TDLB_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

SLOT = re.compile(r"qTmp\.q\[\s*Q(\d+)\s*\]\s*=\s*(\d+)")
CELL = re.compile(r"cu\.Q\[cu\.index\((\d+),\s*(\d+),\s*(\d+)\)\]\s*=\s*qTmp;")
FORCE = re.compile(r"cu\.F\[cu\.index\((\d+),\s*(\d+),\s*(\d+)\)\]\.([xyz])\s*=\s*(\d+);")


def parse(path, n=5):
    q = np.full((n, n, n, 18), np.iinfo(np.uint64).max, dtype=np.uint64)
    f = np.full((n, n, n, 3), np.iinfo(np.uint64).max, dtype=np.uint64)
    cur = {}
    with open(path) as fh:
        for line in fh:
            for m in SLOT.finditer(line):
                cur[int(m.group(1)) - 1] = int(m.group(2))
            m = CELL.search(line)
            if m:
                if len(cur) != 18:
                    raise ValueError(f"{path}: cell {m.groups()} sets {len(cur)} slots")
                i, j, k = (int(v) for v in m.groups())
                for l, v in cur.items():
                    q[i, j, k, l] = v
                cur = {}
            m = FORCE.search(line)
            if m:
                i, j, k = (int(v) for v in m.groups()[:3])
                f[i, j, k, "xyz".index(m.group(4))] = int(m.group(5))
    if (q == np.iinfo(np.uint64).max).any():
        raise ValueError(f"{path}: not every cell was set")
    return q, f


def main():
    src = os.path.join(REF, "tdlbcpp", "tests")
    pull_q, _ = parse(os.path.join(src, "StreamingSimpleTest.hpp"))
    push_q, push_f = parse(os.path.join(src, "StreamingSimplePushTest.hpp"))
    if (push_f == np.iinfo(np.uint64).max).any():
        raise ValueError("push fixture: F not fully set")
    np.save(os.path.join(HERE, "streaming_pull_Q.npy"), pull_q)
    np.save(os.path.join(HERE, "streaming_push_Q.npy"), push_q)
    np.save(os.path.join(HERE, "streaming_push_F.npy"), push_f)
    print("wrote", pull_q.shape, push_q.shape, push_f.shape)
    # sample input files: JSON data, copied after a parse check
    import json
    for name in ("input_debug_gridx60_numSteps20.json", "input_debug_gridx72_IncreaseResolution.json"):
        with open(os.path.join(REF, name)) as fh:
            text = fh.read()
        json.loads(text)
        with open(os.path.join(HERE, name), "w") as fh:
            fh.write(text)
        print("wrote", name)
    return 0


if __name__ == "__main__":
    sys.exit(main())
