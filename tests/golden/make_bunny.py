"""Write tests/golden/bunny.npz from the reference's only data fixture.

The input is model/bunny.off, the mesh the reference driver loads
(src/main.cpp:172).  It is read here as plain OFF text, and its vertices and
faces are stored as arrays.  The script reads /root/reference, so it runs only
where the reference is mounted.  The .npz it writes is committed.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/model/bunny.off"


def read_off(path):
    with open(path) as fh:
        tokens = fh.read().split()
    assert tokens[0] == "OFF"
    nv, nf = int(tokens[1]), int(tokens[2])
    pos = 4
    V = np.array(tokens[pos:pos + 3 * nv], dtype=np.float64).reshape(nv, 3)
    pos += 3 * nv
    F = []
    for _ in range(nf):
        k = int(tokens[pos])
        F.append([int(t) for t in tokens[pos + 1:pos + 1 + k]])
        pos += 1 + k
    return V, np.array(F, dtype=np.int32)


if __name__ == "__main__":
    if not os.path.exists(SRC):
        sys.exit("reference model not mounted")
    V, F = read_off(SRC)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bunny.npz")
    np.savez_compressed(out, V=V, F=F)
    print(out, V.shape, F.shape)
