#!/bin/sh
# Copies the data fixtures the reference ships (binary states it wrote with
# its CUDA build, format v2.11) into tests/golden/.  Run from the repository
# root where /root/reference is mounted.
set -e
cp /root/reference/output/cube-init.output00029.bin tests/golden/cube-init.output00029.bin
cp /root/reference/output/cube-init.status.dat tests/golden/cube-init.status.dat
