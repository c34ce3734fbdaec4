"""The C++ drop-in on the GPU: a program written against include/hdm_compat.hpp
(the reference's template getHausdorffDistance_CUDA_Generic<T>, hello.cu:
2025-2062) and linked with libhdm.so puts golden volumes into hipMalloc
memory, calls getHausdorffDistance_CUDA_Generic<uint8_t> and
averageHausdorffDistance<uint8_t>, and must print the fixture's HD and
average HD (sum / count in FP64).  The program is a host-only g++ build
(the HIP runtime is called through its C API), compiled in a second."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from tests.golden_cases import case, volumes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <hip/hip_runtime_api.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "hdm_compat.hpp"

// usage: prog file nx ny nz label; the file holds gold then segm, nx*ny*nz bytes each
int main(int argc, char **argv) {
  if (argc != 6) return 10;
  const int nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
  const uint8_t label = (uint8_t)atoi(argv[5]);
  const size_t n = (size_t)nx * ny * nz;
  std::vector<uint8_t> host(2 * n);
  FILE *f = fopen(argv[1], "rb");
  if (!f || fread(host.data(), 1, 2 * n, f) != 2 * n) return 11;
  fclose(f);
  uint8_t *gold = nullptr, *segm = nullptr;
  if (hipMalloc((void **)&gold, n) != hipSuccess || hipMalloc((void **)&segm, n) != hipSuccess)
    return 12;
  if (hipMemcpy(gold, host.data(), n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(segm, host.data() + n, n, hipMemcpyHostToDevice) != hipSuccess)
    return 13;
  try {
    const int hd = hdm_compat::getHausdorffDistance_CUDA_Generic<uint8_t>(
        gold, segm, nx, ny, nz, 1.0f, false, label, false);
    const double avg = hdm_compat::averageHausdorffDistance<uint8_t>(gold, segm, nx, ny, nz,
                                                                     1.0f, label);
    const int hd_half = hdm_compat::getHausdorffDistance_CUDA_Generic<uint8_t>(
        gold, segm, nx, ny, nz, 0.5f, false, label, false);
    std::printf("%d %.17g %d\n", hd, avg, hd_half);
  } catch (const hdm_compat::error &e) {
    std::printf("error %d\n", e.code);
    return 14;
  }
  (void)hipFree(gold);
  (void)hipFree(segm);
  return 0;
}
'''


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    d = tmp_path_factory.mktemp("compat_gpu")
    src = d / "compat_gpu.cpp"
    src.write_text(PROGRAM)
    exe = d / "compat_gpu"
    pkg = os.path.join(ROOT, "housdorffhdf5cmake_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", str(src),
                        "-o", str(exe), "-L", pkg, "-lhdm", f"-Wl,-rpath,{pkg}",
                        "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    return d, exe


@pytest.mark.parametrize("name", ["sphere64", "ell_160x128x96", "single_voxels", "kat_firstmeta"])
def test_cpp_drop_in_on_device_memory_matches_fixture(program, name):
    d, exe = program
    c = case(name)
    label = int(c.get("label", 1))
    g, s = volumes(c, label)
    nx, ny, nz = c["dims"]
    path = d / f"{name}.bin"
    path.write_bytes(np.concatenate([g.reshape(-1), s.reshape(-1)]).tobytes())
    print("case", name, flush=True)
    r = subprocess.run([str(exe), str(path), str(nx), str(ny), str(nz), str(label)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-2000:])
    hd, avg, hd_half = r.stdout.split()
    exact = [e for e in c["expected"] if e["robustness"] == 1.0][0]
    assert int(hd) == exact["hd"]
    want_avg = exact["sum"] / exact["count"] if exact["count"] else 0.0
    assert float(avg) == want_avg
    half = [e for e in c["expected"] if e["robustness"] == 0.5]
    if half:
        assert int(hd_half) == half[0]["hd"]
