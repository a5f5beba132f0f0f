"""Writes tests/golden/truck_cameras_head.json: the first four keyframes of
the reference's assets/cameras/truck_camera_path.json (data, unchanged: same
keys and values), read from the reference checkout when it is present."""
import json
import os
import sys

SRC = "/root/reference/assets/cameras/truck_camera_path.json"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truck_cameras_head.json")

if __name__ == "__main__":
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} not found")
    with open(SRC) as fh:
        cams = json.load(fh)
    with open(DST, "w") as fh:
        json.dump(cams[:4], fh, indent=4)
    print(f"wrote {DST}: {len(cams[:4])} of {len(cams)} cameras")
