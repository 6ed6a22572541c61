"""Headless frame export (reference: ChysX/src/render/batch_export.cu:21-77):
step a cloth scene on the GPU and write every frame as one merged Houdini
.bgeo (the cloth plus the scene's static boxes, one uniform colour per
piece) and one OBJ per piece.

    python -m chysx.export --scene drape --frames 120 --out output/
"""
from __future__ import annotations

import argparse
import os

import numpy as np

CLOTH_RGB = (0.85, 0.35, 0.25)
BOX_RGB = (0.6, 0.6, 0.6)


def box_piece(center, half_ext, ex, ey, ez):
    """Mesh of an oriented box: the axis-aligned box mesh at the origin,
    turned into the frame (ex, ey, ez) and moved to `center`."""
    from . import io

    p, t = io.box_mesh(np.zeros(3, np.float32), np.asarray(half_ext, np.float32))
    R = np.stack([np.asarray(ex, np.float64), np.asarray(ey, np.float64), np.asarray(ez, np.float64)], axis=1)
    return (np.asarray(center, np.float64) + p.astype(np.float64) @ R.T).astype(np.float32), t


def export_scene(spec, out_dir: str, frames: int = 120, frame_dt: float = 1.0 / 60.0, log=None) -> list[str]:
    """Steps `spec` (a chysx.scenes.ClothSpec) for `frames` frames of
    `frame_dt` (substeps of spec.dt) and writes frame_%04d.bgeo and
    piece%d_%04d.obj per frame into `out_dir`.  Returns the written paths."""
    from . import io, scenes

    os.makedirs(out_dir, exist_ok=True)
    substeps = max(1, int(round(frame_dt / spec.dt)))
    cloth = scenes.DeviceCloth(spec)
    boxes = [box_piece(*b) for b in spec.boxes]
    tris = np.ascontiguousarray(spec.tris, np.int32)
    written = []
    for frame in range(frames):
        for _ in range(substeps):
            cloth.step()
        pos = cloth.state()[0].astype(np.float32)
        pieces = [(pos, tris, CLOTH_RGB)] + [(p, t, BOX_RGB) for p, t in boxes]
        path = os.path.join(out_dir, f"frame_{frame:04d}.bgeo")
        io.write_bgeo_multi(path, pieces)
        written.append(path)
        for k, (p, t, _) in enumerate(pieces):
            path = os.path.join(out_dir, f"piece{k}_{frame:04d}.obj")
            io.save_obj(path, p, t)
            written.append(path)
        if log and ((frame + 1) % 10 == 0 or frame == 0):
            log(f"frame {frame + 1}/{frames}")
    return written


def main():
    from . import scenes

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--scene", default="drape", choices=["drape", "plain_small"])
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--frame-dt", type=float, default=1.0 / 60.0)
    ap.add_argument("--out", default="output")
    args = ap.parse_args()
    spec = getattr(scenes, args.scene)()
    export_scene(spec, args.out, args.frames, args.frame_dt, log=print)


if __name__ == "__main__":
    main()
