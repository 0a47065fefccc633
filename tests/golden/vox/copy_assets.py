"""Copies the reference's MagicaVoxel asset files (assets/models/*.vox, data
files, ~140 KB) next to this script, as fixtures for the .vox import tests.
Run from a checkout that has the reference at /root/reference."""
import shutil
from pathlib import Path

SRC = Path("/root/reference/assets/models")
HERE = Path(__file__).resolve().parent

for name in ("navigate.vox", "navigate_x.vox", "navigate_y.vox", "navigate_z.vox"):
    shutil.copyfile(SRC / name, HERE / name)
    print(name, (HERE / name).stat().st_size)
