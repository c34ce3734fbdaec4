#!/bin/sh
# Writes the HDF5 loader fixtures (tests/golden/h5/*.h5) with the HDF5 C
# library.  Needs libhdf5 and its headers (HDF5_DIR, default /opt/conda); the
# committed files make the tests independent of it.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
HDF5_DIR=${HDF5_DIR:-/opt/conda}
TMP=$(mktemp -d)
gcc -O1 -o "$TMP/make_h5" "$HERE/make_h5.c" -I"$HDF5_DIR/include" -L"$HDF5_DIR/lib" -lhdf5 \
    -Wl,-rpath,"$HDF5_DIR/lib"
mkdir -p "$HERE/h5"
"$TMP/make_h5" "$HERE/h5"
rm -rf "$TMP"
ls -l "$HERE/h5"
