"""The C-ABI library loads and exports every entry point include/tsgpu.h
declares (no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import tsgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "tsgpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(tsg_[a-z0-9_]+)\s*\(", text))
    names.discard("tsg_progress_fn")
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(tsgpu.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_mirror_reference_parameters():
    # Parameters::default (lib.rs:342-359)
    p = tsgpu.Params()
    tsgpu.lib().tsg_params_default(C.byref(p))
    assert (p.nearest_neighbors, p.random_sample_locations, p.backtrack_stages) == (50, 50, 5)
    assert (p.cauchy_dispersion, p.backtrack_percent) == (1.0, 0.5)
    assert abs(p.guide_alpha - 0.8) < 1e-7 and p.seed == 0 and p.tiling_mode == 0
    assert p.random_init == tsgpu.U64_NONE


def test_version_string():
    assert b"gfx950" in tsgpu.lib().tsg_version()


def test_xform_round_trip():
    # lib.rs:642-692 coord_tx_serde
    ct = tsgpu.CoordinateTransform([1, 2, 3, 4, 5, 6], tsgpu.Dims(2, 1),
                                   [tsgpu.Dims(9001, 9002), tsgpu.Dims(20, 5)])
    data = ct.write()
    back = tsgpu.CoordinateTransform.read(data)
    assert list(back.buffer) == [1, 2, 3, 4, 5, 6]
    assert back.output_size == tsgpu.Dims(2, 1)
    assert back.original_maps == [tsgpu.Dims(9001, 9002), tsgpu.Dims(20, 5)]
    assert data[:4] == bytes([0x01, 0x00, 0x34, 0x12])


def test_autotune_space_names_tuning_knobs():
    from tsgpu.autotune import SPACE
    t = tsgpu.default_tuning()
    for knob, values in SPACE:
        assert knob in t and len(values) >= 2


def header_struct_fields(name):
    """Field names of `typedef struct name { ... } name;` in include/tsgpu.h."""
    with open(os.path.join(ROOT, "include", "tsgpu.h")) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = re.sub(r"//.*", "", decl).strip()
        if decl:
            fields.append(re.findall(r"(\w+)\s*(?:\[[^]]*\])?$", decl)[0])
    return fields


def test_python_mirrors_match_the_header_structs():
    # the ctypes mirrors must keep the header's field order (a drift would
    # shift every later field across the ABI)
    for cname, py in (("tsg_tuning", tsgpu.Tuning), ("tsg_schedule", tsgpu.Schedule), ("tsg_params", tsgpu.Params)):
        assert [n for n, _ in py._fields_] == header_struct_fields(cname), cname


def test_tuning_defaults():
    # every knob is valid, and the measured defaults are the documented ones
    t = tsgpu.default_tuning()
    assert t["flow"] == 1 and t["defer"] == 0 and t["persist"] == 0 and t["ahead"] == 1
    assert t["flow_head"] == 1024 and t["flow_head_team"] == 8
    assert t["flow_team"] in (1, 2, 4, 8) and t["team_max"] in (1, 2, 4, 8)
