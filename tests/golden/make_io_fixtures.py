"""Writes tests/golden/c60large_cvol.json: the header fields and voxel checksums of the reference's
only volume file, pytests/volumes/C60Large.cvol, decoded here with numpy alone (64-byte header:
magic, 3 x uint64 size, 3 x float64 voxel size, int32 type, uint8 compression flag, 7 bytes padding,
renderer/volume.cpp:75-84; uchar voxels, x fastest).

The two settings files beside it (skull1.json, tooth3gauss.json) are copies of
pytests/config-files/*.json, data files of the reference's scripts."""
import json
import os
import struct

import numpy as np

SRC = "/root/reference/pytests/volumes/C60Large.cvol"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c60large_cvol.json")

raw = open(SRC, "rb").read()
magic = raw[:4].decode()
X, Y, Z = struct.unpack("<3Q", raw[4:28])
vx, vy, vz = struct.unpack("<3d", raw[28:52])
typ, = struct.unpack("<i", raw[52:56])
comp = raw[56]
vox = np.frombuffer(raw[64:], dtype=np.uint8)
assert vox.size == X * Y * Z
points = [(0, 0, 0), (64, 64, 64), (10, 70, 100), (127, 127, 127), (64, 30, 90), (90, 64, 33)]
out = {
    "magic": magic, "size": [X, Y, Z], "voxel_size": [vx, vy, vz], "type": typ, "compressed": comp,
    "byte_sum": int(vox.astype(np.int64).sum()), "nonzero": int((vox > 0).sum()),
    "min_byte": int(vox.min()), "max_byte": int(vox.max()),
    "points": [[x, y, z, int(vox[x + X * (y + Y * z)])] for x, y, z in points],
}
json.dump(out, open(OUT, "w"), indent=1)
print(out)
