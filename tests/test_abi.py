"""The C-ABI library loads, exports every entry point of include/rvpf.h,
its structs match the ctypes mirror, and host-side validation rejects bad
input with the documented codes. No GPU needed."""
import ctypes as C
import os
import re

import pytest

from rvpf import abi, api, workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "rvpf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rvpf_\w+)\s*\(", src)))


def test_struct_sizes_match_c():
    assert workload.struct_sizes() == abi.STRUCT_SIZES


def test_library_exports_every_declared_function():
    lib = C.CDLL(abi.lib_path("librvpf.so"))
    names = declared_functions()
    assert set(names) == set(abi.RVPF_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def _frame(**kw):
    return workload.Frame(**dict(dict(width=96, height=64, bit_depth=10, lr=2, grain=True, seed=1), **kw))


def test_validate_accepts_synthetic_frames():
    for kw in [dict(), dict(bit_depth=8), dict(bit_depth=12, mono=True), dict(ssx=0, ssy=0)]:
        f = _frame(**kw)
        rc, msg = api.validate(f.params, f.host_side_info())
        assert rc == abi.RVPF_OK, msg


@pytest.mark.parametrize("field,value,code", [
    ("bit_depth", 9, abi.RVPF_E_INVALID),
    ("use_superres", 1, abi.RVPF_E_INVALID),  # frame_width is not upscaled_width * 8 / d
    ("upscaled_width", 100, abi.RVPF_E_INVALID),  # differs from frame_width without super-resolution
    ("mi_cols", 23, abi.RVPF_E_INVALID),
    ("cdef_damping", 7, abi.RVPF_E_INVALID),
    ("loop_filter_sharpness", 8, abi.RVPF_E_INVALID),
])
def test_validate_rejects(field, value, code):
    f = _frame()
    setattr(f.params, field, value)
    rc, msg = api.validate(f.params, f.host_side_info())
    assert rc == code, msg


@pytest.mark.parametrize("denom", [9, 12, 16])
def test_validate_accepts_superres(denom):
    f = _frame(width=200, superres_denom=denom)
    assert f.params.use_superres == 1 and f.params.upscaled_width == 200
    assert f.params.frame_width == (200 * 8 + denom // 2) // denom
    rc, msg = api.validate(f.params, f.host_side_info())
    assert rc == abi.RVPF_OK, msg
    f.params.frame_width += 1  # no denominator in 9..16 gives this width
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID


def test_validate_rejects_bad_arrays():
    f = _frame()
    f.params.loop_filter_level[0] = 64
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    f.params.cdef_y_sec_strength[3] = 3
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    f.params.grain.num_y_points = 15
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    f.params.grain.num_y_points = 3
    f.params.grain.point_y_value[0] = 10
    f.params.grain.point_y_value[1] = 10
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    f.params.loop_restoration_size[0] = 32
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame(ssx=0, ssy=0)
    f.params.loop_restoration_size[1] = 32
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    f.params.lr_unit_cols[0] += 1
    assert api.validate(f.params, f.host_side_info())[0] == abi.RVPF_E_INVALID
    f = _frame()
    s = f.host_side_info()
    s.mi_stride = f.params.mi_cols
    assert api.validate(f.params, s)[0] == abi.RVPF_E_INVALID
    f = _frame()
    s = f.host_side_info()
    s.lr_units[0] = None
    assert api.validate(f.params, s)[0] == abi.RVPF_E_INVALID


def test_compute_without_gpu_raises():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(RuntimeError):
        api.Context()
