"""Copy three of the reference's bundled scene files into tests/golden/scenes/ as fixtures.

Source: scenes/{exhibition,forme-exposition,orfevrerie}.ron of the SyFrack tree, the bundled
scenes that combine a sun, an area emitter, a sky gradient and the height fog over a form graph
(Kelvin and linear light colours, a trap colouring ramp); scenes/studies/epic-20261014/3d/3d-01.ron (brushed metal and two concrete finishes on a ground)
and scenes/studies/epic-20261014/3d-volumes/3d-volumes-01.ron (brushed metal).  They are data (RON scene
descriptions), copied byte for byte.  Run from the repo root:
    python tests/golden/copy_scenes.py /root/reference
"""
import os
import shutil
import sys

NAMES = ("exhibition", "forme-exposition", "orfevrerie", "studies/epic-20261014/3d-volumes/3d-volumes-01", "studies/epic-20261014/3d/3d-01")


def copy(ref_root: str, out_dir: str) -> None:
    os.makedirs(out_dir, exist_ok=True)
    for name in NAMES:
        shutil.copyfile(os.path.join(ref_root, "scenes", name + ".ron"),
                        os.path.join(out_dir, os.path.basename(name) + ".ron"))


if __name__ == "__main__":
    copy(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes"))
