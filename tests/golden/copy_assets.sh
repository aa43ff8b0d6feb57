#!/bin/sh
# Copies the reference's scene data files (inputs of the ingest tests) into
# tests/golden/assets: a glTF binary, three OBJ meshes and a .luz project
# (whose .luzbin the reference does not ship).
set -e
SRC=/root/reference/assets
DST=$(dirname "$0")/assets
mkdir -p "$DST"
for f in cube.glb default.luz point.obj spot.obj directional.obj; do cp "$SRC/$f" "$DST/$f"; done
