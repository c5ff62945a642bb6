"""The C ABI (CPU): the library loads without a GPU, exports every entry point include/ymc3d.h
declares, and the ctypes mirrors have the C compiler's struct layouts."""
import ctypes as C
import os
import re
import subprocess

import pytest

import ymc3d_amd as y
from ymc3d_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ymc3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(ymc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_entry_point_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(_abi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_abi.EXPORTED) == names
    assert y.load_library().ymc_abi_version() == _abi.ABI_VERSION


def test_product_library_does_not_reference_the_oracle():
    out = subprocess.run(["nm", "-D", _abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "ymc_oracle" not in out
    deps = subprocess.run(["ldd", _abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in deps


STRUCTS = {
    "ymc_geometry": _abi.Geometry, "ymc_optprops": _abi.OptProps, "ymc_params": _abi.Params,
    "ymc_totals": _abi.Totals, "ymc_tally": _abi.Tally, "ymc_photon_record": _abi.PhotonRecord,
}


def test_struct_layouts_match_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('printf("ymc_chain %zu\\n", sizeof(ymc_chain)); return 0;}')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)
    assert int(got["ymc_chain"]) == _abi.CHAIN_DTYPE.itemsize
    assert _abi.RECORD_DTYPE.itemsize == C.sizeof(_abi.PhotonRecord)
    assert _abi.TOTALS_DTYPE.itemsize == C.sizeof(_abi.Totals)


def test_engine_without_gpu_fails_loudly():
    """On a host without a GPU the engine refuses to start instead of computing anything."""
    try:
        e = y.Engine(0, y.generate_multipliers(5))
    except y.YmcError as err:
        assert err.status == _abi.ERR_DEVICE
    else:
        e.close()
        pytest.skip("a GPU is present")


def test_fma_probe_rejects_bad_arguments_before_touching_a_device():
    """The FMA-rate probe checks its arguments on the host: 0 or 9 waves per SIMD, operand counts
    outside 1..3, no iterations or no output is an argument error whether or not a GPU is present."""
    lib = y.load_library()
    tf, ms = C.c_double(), C.c_double()
    for waves, iters in ((0, 10), (9, 10), (4, 0)):
        assert lib.ymc_probe_fma_rate(0, waves, 2, iters, C.byref(tf), C.byref(ms)) == _abi.ERR_ARG
    for operands in (0, 4):
        assert lib.ymc_probe_fma_rate(0, 4, operands, 10, C.byref(tf), C.byref(ms)) == _abi.ERR_ARG
    assert lib.ymc_probe_fma_rate(0, 4, 2, 10, None, None) == _abi.ERR_ARG
