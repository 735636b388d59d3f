#!/bin/bash
# Copies the elPaSo HDF5 data files the HDF5 input/output tests use into
# tests/golden (run in a container that holds the reference at /root/reference;
# h5copy / h5dump are the HDF5 1.10 tools of /opt/conda/bin).
#  - acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5: the reference's own
#    2-element Fluid8 input fixture (test/resources/unitHDF5/material/fluid/
#    linear/elastic/), used as is
#  - acousticmedium_fluid8_2elem_AF_LIN_EQF_ISO_DIR.hdf5: the same mesh with an
#    equivalent-fluid material (test/resources/unitHDF5/material/fluid/linear/
#    viscoelastic/), a material type elPaSo-Core's factory does not create;
#    used as is
#  - egenoutput_dsg4_excerpt.hdf5: /Solution/Maps/mtxDofMap and
#    /Solution/State/vecFemStep1..2 of the solution file elPaSo wrote for the
#    DSG4 plate tutorial (doc/tutorials/FEM_tutorials/Simulations/
#    Ex_1_Analysis_of_a_rectangular_plate_DSG4/eGenOutput_*.hdf5)
#  - egenoutput_dsg4_freqstep_attrs.txt: h5dump of that file's FreqStep1 and
#    FreqStep2 attributes (their type and "%f" text)
set -euo pipefail
REF=/root/reference
H=/opt/conda/bin
OUT=$(cd "$(dirname "$0")" && pwd)
cp "$REF/elpasoCore/test/resources/unitHDF5/material/fluid/linear/elastic/acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5" "$OUT/"
chmod u+w "$OUT/acousticmedium_fluid8_2elem_AF_LIN_UAF_ISO_DIR.hdf5"
cp "$REF/elpasoCore/test/resources/unitHDF5/material/fluid/linear/viscoelastic/acousticmedium_fluid8_2elem_AF_LIN_EQF_ISO_DIR.hdf5" "$OUT/"
chmod u+w "$OUT/acousticmedium_fluid8_2elem_AF_LIN_EQF_ISO_DIR.hdf5"
SOL="$REF/doc/tutorials/FEM_tutorials/Simulations/Ex_1_Analysis_of_a_rectangular_plate_DSG4/eGenOutput_Analysis_of_a_rectangular_plate.hdf5"
EX="$OUT/egenoutput_dsg4_excerpt.hdf5"
rm -f "$EX"
for obj in /Solution/Maps/mtxDofMap /Solution/State/vecFemStep1 /Solution/State/vecFemStep2; do
  "$H/h5copy" -p -i "$SOL" -o "$EX" -s "$obj" -d "$obj"
done
"$H/h5dump" -a /Solution/State/FreqStep1 -a /Solution/State/FreqStep2 "$SOL" | sed 1d > "$OUT/egenoutput_dsg4_freqstep_attrs.txt"
