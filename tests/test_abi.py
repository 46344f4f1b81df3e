"""CPU checks of the C ABI: the HIP library loads, exports every declared entry point, its config
struct layout matches the header, and configuration errors are reported without touching a GPU."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oceananigans_mi355x as om
from oceananigans_mi355x import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oceananigans_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(om_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    assert {"om_create", "om_destroy", "om_time_step", "om_rk3_substep", "om_update_state"} <= set(names)
    assert set(names) == set(_abi.SIGNATURES), "ctypes signatures and header declarations differ"


def test_library_loads_and_exports_every_declared_symbol():
    lib = _abi.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.om_abi_version() == 2


def test_config_struct_layout_matches_c():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "oceananigans_mi355x.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(om_config_t), offsetof(om_config_t, z_faces),
         offsetof(om_config_t, bc_top_value), offsetof(om_config_t, pressure_solver),
         sizeof(om_field_info_t), offsetof(om_field_info_t, stride));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    C = _abi.OmConfig
    F = _abi.OmFieldInfo
    assert vals == [ctypes.sizeof(C), C.z_faces.offset, C.bc_top_value.offset, C.pressure_solver.offset,
                    ctypes.sizeof(F), F.stride.offset]


def _config(**over):
    grid = om.RectilinearGrid(size=(8, 8, 8), extent=(8, 8, 8))
    cfg, keep = om.build_config(grid, advection=om.WENO())
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg, keep


@pytest.mark.parametrize("field,value,msg", [
    ("precision", 2, "precision"),
    ("n_tracers", 99, "n_tracers"),
])
def test_invalid_configuration_is_rejected_before_any_device_work(field, value, msg):
    lib = _abi.load_library()
    cfg, keep = _config(**{field: value})
    h = ctypes.c_void_p()
    rc = lib.om_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert msg in lib.om_last_error(None).decode()


def test_bounded_x_and_small_halo_are_rejected():
    lib = _abi.load_library()
    cfg, keep = _config()
    cfg.topology[0] = _abi.OM_BOUNDED
    h = ctypes.c_void_p()
    assert lib.om_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    cfg, keep = _config()
    cfg.H[2] = 2
    assert lib.om_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert "halo" in lib.om_last_error(None).decode()


def test_null_handle_calls_fail_cleanly():
    lib = _abi.load_library()
    assert lib.om_time_step(None, 1.0, 1) == -1
    assert lib.om_destroy(None) == -1
    assert lib.om_autotune(None, None) == -1
    kc = ctypes.c_int32()
    assert lib.om_set_tuning(None, _abi.OM_TUNE_POISSON_REFINEMENT, 1) == -1
    assert lib.om_get_tuning(None, _abi.OM_TUNE_POISSON_REFINEMENT, ctypes.byref(kc)) == -1


def test_ensemble_arguments_are_checked_without_device_work():
    lib = _abi.load_library()
    out = ctypes.c_void_p()
    null_members = (ctypes.c_void_p * 2)(None, None)
    assert lib.om_ensemble_create(null_members, 2, ctypes.byref(out)) == -1
    assert out.value is None
    assert b"null" in lib.om_ensemble_last_error(None)
    assert lib.om_ensemble_create(null_members, 0, ctypes.byref(out)) == -1
    assert lib.om_ensemble_create(None, 1, ctypes.byref(out)) == -1
    assert lib.om_ensemble_time_step(None, 1.0, 1, 1) == -1
    assert lib.om_ensemble_time_step(None, 1.0, -1, 1) == -1
    assert lib.om_ensemble_destroy(None) == 0


def test_host_api_validation():
    grid = om.RectilinearGrid(size=(8, 8, 8), extent=(8, 8, 8))
    with pytest.raises(ValueError):
        om.WENO(order=7)
    with pytest.raises(ValueError):
        om.build_config(grid, buoyancy=om.BuoyancyTracer(), tracers=("c",))
    with pytest.raises(ValueError):
        om.build_config(grid, boundary_conditions={"w": om.FieldBoundaryConditions(top=om.FluxBoundaryCondition(1.0))})
    with pytest.raises(ValueError):
        om.FPlane()
    assert abs(om.FPlane(latitude=45).f - 2 * 7.292115e-5 * np.sin(np.pi / 4)) < 1e-18


def test_grid_nodes_and_stretched_faces():
    g = om.RectilinearGrid(size=(4, 4, 4), x=(0, 4), y=(0, 8), z=[-10, -6, -3, -1, 0])
    assert g.stretched and g.Lz == 10
    np.testing.assert_allclose(g.znodes("c"), [-8, -4.5, -2, -0.5])
    np.testing.assert_allclose(g.xnodes("f"), [0, 1, 2, 3])
    g2 = om.RectilinearGrid(size=(16, 16), x=(0, 1), y=(0, 1), topology=(om.Periodic, om.Periodic, om.Flat))
    assert g2.N == (16, 16, 1) and g2.H == (3, 3, 0)


def test_synthetic_stream_is_deterministic():
    from oceananigans_mi355x.synthetic import splitmix64, uniform
    assert int(splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF   # SplitMix64 published first output for seed 0
    a = uniform(1, (4, 3, 2), -1, 1)
    b = uniform(1, (4, 3, 2), -1, 1)
    assert np.array_equal(a, b) and a.min() >= -1 and a.max() < 1


def test_tuning_keys_match_the_header():
    import re

    from oceananigans_mi355x import _abi

    text = open(HEADER).read()
    keys = dict(re.findall(r"#define\s+(OM_TUNE_\w+)\s+(\d+)", text))
    assert keys and all(getattr(_abi, k) == int(v) for k, v in keys.items()), keys


JULIA_EXT = os.path.join(ROOT, "oceananigans.jl_amd", "julia", "OceananigansMI355XExt.jl")
# size and alignment of the Julia field types used in the mirrored structs (C ABI on x86-64)
JULIA_TYPES = {"Int32": (4, 4), "UInt32": (4, 4), "Int64": (8, 8), "Float32": (4, 4), "Float64": (8, 8), "Ptr": (8, 8)}


def _julia_struct(text, name, consts):
    """[(field, size, align)] of `struct name ... end` in the extension, in declaration order."""
    body = re.search(r"^struct %s\n(.*?)^end" % name, text, flags=re.M | re.S)
    assert body, name
    out = []
    for line in body.group(1).splitlines():
        m = re.match(r"\s*(\w+)\s*::\s*(.+?)\s*$", line)
        if not m:
            continue
        field, ty = m.groups()
        t = re.match(r"NTuple\{(\w+),\s*(\w+)\}", ty)
        if t:
            n = int(consts.get(t.group(1), t.group(1)))
            size, align = JULIA_TYPES[t.group(2)]
            out.append((field, n * size, align))
        else:
            out.append((field,) + JULIA_TYPES[re.sub(r"\{.*", "", ty)])
    return out


def _julia_offsets(fields):
    off, offs, maxa = 0, [], 1
    for name, size, align in fields:
        off = (off + align - 1) // align * align
        offs.append((name, off))
        off += size
        maxa = max(maxa, align)
    return offs, (off + maxa - 1) // maxa * maxa


def test_julia_struct_mirrors_the_header():
    """The Julia extension's OmConfig / OmFieldInfo declare the C structs' fields in the header's
    order with the header's types: every field offset and the struct size agree (gcc offsetof on
    the header; Julia isbits structs follow the C layout rules), and the constants it defines
    equal the header's."""
    text = open(JULIA_EXT).read()
    header = open(HEADER).read()
    hdefs = dict(re.findall(r"#define\s+(OM_\w+)\s+\(?(-?\d+)", header))
    consts = {"OM_MAX_FIELDS": int(hdefs["OM_MAX_TRACERS"]) + 3}
    # constants: "const A, B = Int32.((0, 1))" and "const A = Int32(3)" forms
    jl = {}
    for names, vals in re.findall(r"^const ([A-Z0-9_, \n]+?)\s*=\s*(?:Int32\.)?\(\(?([-\d, ]+)\)?\)\s*$", text, flags=re.M):
        ns = [n.strip() for n in names.replace("\n", " ").split(",") if n.strip()]
        vs = [int(v) for v in vals.split(",") if v.strip()]
        jl.update(zip(ns, vs))
    for names, vals in re.findall(r"^const ([A-Z0-9_, ]+?) = ((?:Int32\(-?\d+\),?\s*)+)$", text, flags=re.M):
        ns = [n.strip() for n in names.split(",")]
        vs = [int(v) for v in re.findall(r"Int32\((-?\d+)\)", vals)]
        jl.update(zip(ns, vs))
    jl.update({k: int(v) for k, v in re.findall(r"^const (OM_MAX_TRACERS) = (\d+)$", text, flags=re.M)})
    assert len(jl) >= 20, jl
    for k, v in jl.items():
        assert k in hdefs and int(hdefs[k]) == v, (k, v, hdefs.get(k))
    for cname, jname in (("om_config_t", "OmConfig"), ("om_field_info_t", "OmFieldInfo")):
        fields = _julia_struct(text, jname, consts)
        offs, size = _julia_offsets(fields)
        prog = "#include <stddef.h>\n#include <stdio.h>\n#include \"oceananigans_mi355x.h\"\nint main(void) {\n"
        prog += '  printf("%%zu\\n", sizeof(%s));\n' % cname
        for name, _ in offs:
            prog += '  printf("%%zu\\n", offsetof(%s, %s));\n' % (cname, name)
        prog += "  return 0;\n}\n"
        with tempfile.TemporaryDirectory() as d:
            c = os.path.join(d, "layout.c")
            exe = os.path.join(d, "layout")
            open(c, "w").write(prog)
            subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
            vals = [int(x) for x in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
        assert vals[0] == size, (cname, "size", vals[0], size)
        for (name, off), coff in zip(offs, vals[1:]):
            assert off == coff, (cname, name, off, coff)
