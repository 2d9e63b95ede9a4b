#!/usr/bin/env python3
"""Copies the reference's own test data used by the FTB and trajectory tests
into tests/golden/ (run in the build container, where the reference tree is
at /root/reference; the copies travel with the repository).

* sample1 (a real Kinect-style sequence): frames 0..14 with their depth PNGs
  and the colour PNGs of frames 0, 8 and 14 -- the pair 0 -> 14 is the one
  whose accuracy the reference's Registration notebook prints, 0 -> 8 the
  one of its interactive ICP tests;
* sample2 (synthetic): frames 0..3 with depth, colour for frames 0 and 3
  (its interactive tests register 0 -> 3);
* the freiburg1_desk ground-truth / predicted trajectories, whose ATE, RTE,
  ARE and RRE the reference's metrics test pins.

frames.json is rewritten with only the copied frames (in their original
order, so indices are kept) and without the colour file of frames whose
colour image is not copied.
"""
import json
import shutil
from pathlib import Path

SRC = Path("/root/reference/test-data")
DST = Path(__file__).resolve().parent

PICK = {"sample1": (range(15), {0, 8, 14}), "sample2": (range(4), {0, 3})}


def main():
    for name, (frames, colour) in PICK.items():
        src, dst = SRC / "rgbd" / name, DST / "ftb" / name
        dst.mkdir(parents=True, exist_ok=True)
        root = json.load(open(src / "frames.json"))["root"]
        out = []
        for i in frames:
            entry = dict(root[i])
            shutil.copyfile(src / entry["depth_image"], dst / entry["depth_image"])
            if i in colour and "rgb_image" in entry:
                shutil.copyfile(src / entry["rgb_image"], dst / entry["rgb_image"])
            else:
                entry.pop("rgb_image", None)
            out.append(entry)
        json.dump({"root": out}, open(dst / "frames.json", "w"), indent=1)
    traj = DST / "trajectory"
    traj.mkdir(exist_ok=True)
    for f in ("freiburd1_desk_gt.traj", "freiburd1_desk_pred.traj"):
        shutil.copyfile(SRC / "trajectory" / f, traj / f)


if __name__ == "__main__":
    main()
