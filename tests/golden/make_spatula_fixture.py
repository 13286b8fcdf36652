"""Writes tests/golden/spatula_motion_head.bin: the first 181 records (3 s at
60 fps) of the reference's knife animation, matter/animations/spatula_motion.bin
(FrameData {float pos[3]; float quat[4]}, 28 bytes each, object_spatula.hpp:7-10),
and spatula_motion.json with the full file's record count and SHA-256.  Run in
a container where the reference is mounted; the fixture is data only."""
import hashlib
import json
import os

SRC = "/root/reference/matter/animations/spatula_motion.bin"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    data = open(SRC, "rb").read()
    head = data[:181 * 28]
    with open(os.path.join(HERE, "spatula_motion_head.bin"), "wb") as f:
        f.write(head)
    with open(os.path.join(HERE, "spatula_motion.json"), "w") as f:
        json.dump({"records": len(data) // 28, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest(),
                   "head_records": 181}, f, indent=1)
    print("wrote", len(head), "bytes")
