"""The drop-in recipe for the reference's gpu-mc program builds and reaches the engine (CPU).

Follows the integration notes' recipe on a scratch copy of the reference's host files: the shim
(the C++ block of INTEGRATION.md, which replaces src/mc3d_gpu.cu), src/mc3d_main.cu and
src/mc3d_io.cu with its two CUDA runtime sections deleted (opticalprops_transfer_to_device and the
ALLOC_GPU case of opticalprops_free), compiled as HIP against include/ymc3d.h and linked against
libymc3d.so.  The program then reads a geometry and a tissue file with the reference's readers,
prints the reference's table, and calls run_3dmc_gpu, which reaches the engine: without a GPU the
engine reports that no HIP device is available.  Skipped where the reference tree or hipcc is
absent.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from ymc3d_amd import _abi, configs, io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
HIPCC = "/opt/rocm/bin/hipcc"
KEEP = ["src/mc3d_main.cu", "src/mc3d_io.cu", "src/mc3d_types.h", "src/mc3d_gpu.h", "src/mc3d_photons.h",
        "src/mc3d_rng.h", "src/mc3d_io.h", "gpumcml_rng/gpumcml_rng.h"]

pytestmark = pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "src")) and os.path.exists(HIPCC)),
                                reason="reference tree or hipcc not available")


def _delete_cuda_sections(text: str) -> str:
    a = text.index("void opticalprops_transfer_to_device")
    b = text.index("void opticalprops_free")
    text = text[:a] + text[b:]
    c = text.index("case ALLOC_GPU:")
    d = text.index("case ALLOC_HOST:")
    return text[:c] + text[d:]


@pytest.fixture(scope="module")
def gpu_mc(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    for f in KEEP:
        os.makedirs(d / os.path.dirname(f), exist_ok=True)
        shutil.copy(os.path.join(REF, f), d / f)
    io_src = d / "src" / "mc3d_io.cu"
    io_src.write_text(_delete_cuda_sections(io_src.read_text()))
    assert "cuda" not in re.sub(r"//.*", "", io_src.read_text()).replace("cuda allocated", "")
    shim = re.search(r"```cpp\n(.*?)```", open(os.path.join(ROOT, "INTEGRATION.md")).read(), re.S).group(1)
    (d / "src" / "mc3d_gpu_ymc.cpp").write_text(shim)
    lib_dir = os.path.dirname(_abi.LIB_PATH)
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-Isrc", "-Igpumcml_rng", "-I" + os.path.join(ROOT, "include"),
           "src/mc3d_main.cu", "src/mc3d_io.cu", "src/mc3d_gpu_ymc.cpp", "-L" + lib_dir, "-lymc3d",
           "-Wl,-rpath," + lib_dir, "-o", "gpu-mc"]
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return d, d / "gpu-mc"


def test_dropin_links_the_engine(gpu_mc):
    _, exe = gpu_mc
    out = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    for sym in ("ymc_run_3dmc_gpu", "ymc_multipliers_read_text", "ymc_multipliers_generate", "ymc_last_error"):
        assert sym in out
    deps = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert "libymc3d.so" in deps and "not found" not in deps
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage:" in r.stderr


def test_dropin_reads_the_files_and_calls_the_engine(gpu_mc, tmp_path):
    _, exe = gpu_mc
    scene = configs.skin_slab_sphere()
    io.write_geometry(str(tmp_path / "geometry.bin"), scene.file_values, *scene.voxel)
    io.write_optprops(str(tmp_path / "tissue.bid"), scene.records[0])
    env = dict(os.environ, YMC3D_SEED="7")
    r = subprocess.run([str(exe), str(tmp_path / "geometry.bin"), str(tmp_path / "tissue.bid"), "def.txt",
                        str(tmp_path / "run")], capture_output=True, text=True, timeout=600, env=env, cwd=tmp_path)
    assert "Read geometry:" in r.stderr
    table = [line.split() for line in r.stdout.splitlines()[1:]]
    want = np.asarray(scene.records[0], np.float64)
    assert np.allclose(np.array(table, np.float64), want, rtol=1e-6)
    if r.returncode == 0:  # a GPU is present: the run completed and wrote the reference's outputs
        drs, _ = io.read_property(str(tmp_path / "run_drs.bin"))
        assert drs.shape == (scene.shape[1], scene.shape[0])
    else:  # no device here: the call reached the engine, which said so
        assert r.returncode == 1 and "ymc3d engine:" in r.stderr and "HIP device" in r.stderr, r.stderr
