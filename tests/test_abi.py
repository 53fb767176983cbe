"""The C-ABI library loads and exports every symbol include/tc_sim.h declares;
argument errors surface as the reference's exception types before any GPU work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cases
from taremincloth_amd import _lib
from taremincloth_amd import taremin_cloth_core as T

HEADER = os.path.join(cases.ROOT, "include", "tc_sim.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_are_exported():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_version_and_error_channel():
    L = _lib.load()
    assert b"gfx950" in L.tc_version()
    assert L.tc_last_error() is not None


def test_bad_edge_index_raises_value_error():
    pos, e, f = cases.grid(3, 3, 0.1)
    bad = e.copy()
    bad[0, 1] = 99
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, bad, f)


def test_bad_shapes_raise_value_error():
    pos, e, f = cases.grid(3, 3, 0.1)
    with pytest.raises(ValueError):
        T.ClothSimulator(pos[:, :2], e, f)
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, e[:, :1], f)
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, e, f, inv_masses=np.ones(3, np.float32))
    with pytest.raises(ValueError):
        T.ClothSimulator(pos, e, np.array([[0, 1, 99]], np.uint32))


def test_host_mesh_build_rejects_bad_faces_through_abi():
    L = _lib.load()
    pos = np.zeros((3, 3), np.float32)
    e = np.array([[0, 1]], np.uint32)
    f = np.array([[0, 1, 7]], np.uint32)
    cnt = C.c_size_t()
    st = L.tc_mesh_build_arrays(pos.ctypes.data_as(C.c_void_p), 3, e.ctypes.data_as(C.c_void_p), 1,
                                f.ctypes.data_as(C.c_void_p), 1, None, None, None, 0, None, 0, C.byref(cnt))
    assert st == _lib.TC_ERR_VALUE
    assert b"faces" in L.tc_last_error()


def test_unsupported_options_raise():
    pos, e, f = cases.grid(3, 3, 0.1)
    with pytest.raises(NotImplementedError):
        T.ClothSimulator(pos, e, f, enable_coarse_constraints=True)


def test_abi_struct_layout_matches_ctypes():
    """The sizes and offsets that tc_layout.h asserts at compile time (g++ and hipcc) are those
    of the ctypes structures the Python binding passes."""
    text = open(os.path.join(cases.ROOT, "taremincloth_amd", "csrc", "tc_layout.h")).read()
    sizes = dict(re.findall(r"static_assert\(sizeof\((\w+)\) == (\d+)", text))
    offs = re.findall(r"static_assert\(offsetof\((\w+), (\w+)\) == (\d+)", text)
    classes = {"tc_sim_params": _lib.TcParams, "tc_instance": _lib.TcInstance,
               "tc_sewing_params": _lib.TcSewingParams}
    assert set(sizes) == set(classes)
    for name, cls in classes.items():
        assert C.sizeof(cls) == int(sizes[name]), name
        fields = [f for f, *_ in cls._fields_]
        asserted = [(f, int(o)) for s, f, o in offs if s == name]
        assert [f for f, _ in asserted] == fields, name   # every field, in order
        for f, o in asserted:
            assert getattr(cls, f).offset == o, (name, f)


def test_getters_are_properties_as_in_pyo3():
    """lib.rs:2927-2950 exposes the counts as #[getter] properties."""
    for name in ("num_vertices", "num_distance_constraints", "num_bending_constraints", "num_sewing_constraints"):
        assert isinstance(getattr(T.ClothSimulator, name), property), name
    for name in ("set_profiling_enabled", "take_profile", "save_profile_trace"):
        assert callable(getattr(T.ClothSimulator, name)), name


_C_SCALAR = {"int": "i32", "uint32_t": "u32", "int32_t": "i32", "float": "f32", "double": "f64", "size_t": "usize",
             "uint64_t": "u64"}


def _c_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#.*", "", text)
    out = []
    for ret, name, params in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(tc_[a-z_0-9]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ps = [p.strip() for p in params.split(",")] if params.strip() not in ("", "void") else []
        out.append((name, ret.strip(), ps))
    return out


def _rust_functions():
    text = open(os.path.join(cases.ROOT, "include", "tc_ffi.rs")).read()
    block = text[text.index('extern "C" {'):]
    out = []
    for name, params, ret in re.findall(r"pub fn (tc_[a-z_0-9]+)\((.*?)\)(?:\s*->\s*([^;]+))?;", block):
        ps = [p.split(":", 1)[1].strip() for p in params.split(",")] if params.strip() else []
        out.append((name, (ret or "").strip(), ps))
    return out


def _shape_c(p):
    """(pointer depth, outermost pointee const, scalar Rust type or None) of a C parameter or return."""
    depth = p.count("*") + p.count("[")
    if "[" in p:  # an array parameter is a pointer to its element type
        const = p.startswith("const")
    else:
        # the qualifier just before the last '*' is the one on what the outermost pointer points to
        segs = p.split("*")
        const = depth > 0 and "const" in segs[-2].split()
    base = re.sub(r"\bconst\b|\*|\[.*?\]", " ", p).split()
    scalar = _C_SCALAR.get(base[0]) if base else None
    return depth, const if depth else False, scalar if depth == 0 else None


def _shape_rust(t):
    depth = t.count("*const") + t.count("*mut")
    const = t.startswith("*const")
    return depth, const if depth else False, t if depth == 0 else None


def test_rust_extern_block_matches_the_header():
    cf, rf = _c_functions(), _rust_functions()
    assert [n for n, _, _ in cf] == [n for n, _, _ in rf]
    assert sorted(n for n, _, _ in cf) == sorted(_lib.EXPORTS)
    for (name, cret, cps), (_, rret, rps) in zip(cf, rf):
        assert len(cps) == len(rps), name
        for cp, rp in zip(cps, rps):
            assert _shape_c(cp) == _shape_rust(rp), (name, cp, rp)
        if cret == "void":
            assert rret == "", name
        else:
            assert _shape_c(cret) == _shape_rust(rret), (name, cret, rret)


def test_rust_structs_and_constants_match_the_header():
    rs = open(os.path.join(cases.ROOT, "include", "tc_ffi.rs")).read()
    for rname, cls in (("TcParams", _lib.TcParams), ("TcInstance", _lib.TcInstance),
                       ("TcSewingParams", _lib.TcSewingParams)):
        body = re.search(r"pub struct %s \{(.*?)\}" % rname, rs, flags=re.S).group(1)
        assert re.findall(r"pub (\w+):", body) == [f for f, *_ in cls._fields_], rname
    hdr = open(HEADER).read()
    for name, val in re.findall(r"pub const (TC_\w+): \w+ = (\d+);", rs):
        m = re.search(r"\b%s\s*=\s*(\d+)" % name, hdr) or re.search(r"#define %s (\d+)u?" % name, hdr)
        assert m and int(m.group(1)) == int(val), name
