"""The FTB RGB-D dataset format (slamtb/data/ftb.py).

A directory with ``frames.json`` -- ``{"root": [{"info": FrameInfo JSON,
"depth_image": file, "rgb_image": file, "segmentation": file}, ...]}`` --
and one 16-bit depth PNG (raw sensor units, scaled by ``depth_scale``) per
frame, optionally an 8-bit RGB PNG and a 16-bit segmentation PNG. Images are
read and written with Pillow (the reference used OpenCV; the decoded arrays
are the same: depth int32, RGB uint8 [H, W, 3] in RGB order, segmentation
int16).
"""
import json
from pathlib import Path

import numpy as np
from PIL import Image

from slamtb.camera import RTCamera
from slamtb.frame import Frame, FrameInfo


def _read_png(path, color=False):
    """Depth / segmentation PNGs decode at their own bit depth (uint16 for
    the 16-bit files); colour PNGs decode to RGB uint8 [H, W, 3]."""
    with Image.open(str(path)) as img:
        return np.array(img.convert("RGB") if color else img)


class FTBDataset:
    """Frames of an FTB directory (ftb.py:15-74)."""

    def __init__(self, frames_json, frame_infos, base_path,
                 ground_truth_model_path=None):
        self.frames_json = frames_json
        self.frame_infos = frame_infos
        self.base_path = Path(base_path)
        self.ground_truth_model_path = ground_truth_model_path

    def get_info(self, idx):
        return self.frame_infos[idx]

    def set_info(self, idx, info):
        self.frame_infos[idx] = info

    def __getitem__(self, idx):
        frame_json = self.frames_json[idx]
        depth = _read_png(self.base_path / frame_json['depth_image']).astype(np.int32)
        rgb = None
        if 'rgb_image' in frame_json:
            rgb = _read_png(self.base_path / frame_json['rgb_image'], color=True)
        seg = None
        if 'segmentation' in frame_json:
            seg = _read_png(self.base_path / frame_json['segmentation']).astype(np.int16)
        return Frame(self.frame_infos[idx], depth, rgb_image=rgb, seg_image=seg)

    def __len__(self):
        return len(self.frames_json)


def load_ftb(base_path, info_file=None):
    """Open an FTB directory (ftb.py:77-101)."""
    base_path = Path(base_path)
    if info_file is None:
        info_file = base_path / "frames.json"
    with open(str(info_file), 'r') as stream:
        frames = json.load(stream)['root']
    infos = [FrameInfo.from_json(frame['info']) for frame in frames]
    return FTBDataset(frames, infos, base_path)


def load_trajectory(json_file):
    """{timestamp: RTCamera} from a JSON list of {"timestamp", "rt_cam"}
    (ftb.py:104-122)."""
    import torch
    with open(str(json_file), 'r') as stream:
        entries = json.load(stream)
    return {e['timestamp']: RTCamera(torch.tensor(e['rt_cam'])) for e in entries}


def write_ftb(base_path, dataset, prefix="frame_", max_frames=None,
              start_frame=0):
    """Write any indexable frame dataset in FTB form (ftb.py:125-180)."""
    base_path = Path(base_path)
    base_path.mkdir(parents=True, exist_ok=True)
    end = len(dataset) if max_frames is None else min(start_frame + max_frames,
                                                        len(dataset))
    frames = []
    for i in range(start_frame, end):
        frame = dataset[i]
        entry = {'info': frame.info.to_json()}
        depth_file = f"{prefix}{i:05d}_depth.png"
        Image.fromarray(np.asarray(frame.depth_image).astype(np.uint16)).save(
            str(base_path / depth_file))
        entry['depth_image'] = depth_file
        if frame.rgb_image is not None:
            rgb_file = f"{prefix}{i:05d}_rgb.png"
            Image.fromarray(np.asarray(frame.rgb_image).astype(np.uint8)).save(
                str(base_path / rgb_file))
            entry['rgb_image'] = rgb_file
        if frame.seg_image is not None:
            seg_file = f"{prefix}{i:05d}_seg.png"
            Image.fromarray(np.asarray(frame.seg_image).astype(np.uint16)).save(
                str(base_path / seg_file))
            entry['segmentation'] = seg_file
        frames.append(entry)
    with open(str(base_path / "frames.json"), 'w') as stream:
        json.dump({'root': frames}, stream, indent=1)
