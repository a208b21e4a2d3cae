#!/bin/sh
# Copies the small model files of the reference (packages/gpu_voxels/models/) used as
# fixtures by tests/test_files.py: data files, read by the point-cloud loaders.
set -e
SRC=/root/reference/packages/gpu_voxels/models
DST=$(dirname "$0")
cp "$SRC/hollie/arm_6_link.binvox" "$DST/arm_6_link.binvox"
cp "$SRC/hollie/arm_6_link.xyz" "$DST/arm_6_link.xyz"
cp "$SRC/hollie_from_pointcloud2.pcd" "$DST/hollie_from_pointcloud2.pcd"
cp "$SRC/coordinate_system_100.binvox" "$DST/coordinate_system_100.binvox"
chmod 644 "$DST"/*.binvox "$DST"/*.xyz "$DST"/*.pcd
