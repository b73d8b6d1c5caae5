"""INTEGRATION.md's Params2Mbr (the sParamRender -> mbr_scene copy a
maintainer adds to the reference) against the library's own key table:
every line `s->FIELD = p.MEMBER;  // KEY, settings.cpp:LINE` must name the
mbr_scene field that mbr_set_setting(KEY) writes, and no two lines may
write the same field."""
import ctypes as C
import os
import re

from mandelbulber_amd import _abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _params2mbr_lines():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = text[text.index("static void Params2Mbr"):]
    body = body[:body.index("\n}\n")]
    out = []
    for line in body.split("\n"):
        m = re.match(r"\s*(?:std::snprintf\(s->(\w+),.*|s->([\w\[\].]+) = .*?;)\s*//\s*(\w+), settings\.cpp:(\d+)", line)
        if m:
            out.append((m.group(1) or m.group(2), m.group(3)))
        elif "s->palette[i]" in line:
            out.append(("palette", "palette"))
    return out


def _field_range(name):
    """Byte range of `name` (field, field[i], field[i][j], field[i].c, field.c) in mbr_scene."""
    m = re.match(r"(\w+)((?:\[\d+\])*)(?:\.(\w))?$", name)
    base, idx, sub = m.group(1), [int(i) for i in re.findall(r"\d+", m.group(2))], m.group(3)
    ft = dict(_abi.Scene._fields_)[base]
    off = getattr(_abi.Scene, base).offset
    t = ft
    for i in idx:
        t = t._type_
        off += i * C.sizeof(t)
    if sub:
        off += getattr(t, sub).offset
        t = dict(t._fields_)[sub]
    return off, off + C.sizeof(t)


def test_params2mbr_matches_key_table():
    lines = _params2mbr_lines()
    assert len(lines) == 281
    seen = set()
    for field, key in lines:
        s0 = _abi.Scene()
        assert lib().mbr_scene_defaults(C.byref(s0)) == 0
        s1 = _abi.Scene()
        C.memmove(C.byref(s1), C.byref(s0), C.sizeof(s0))
        value = b"123456 abcdef" if key == "palette" else (b"/tmp/x" if key.startswith("file_") else b"7")
        assert lib().mbr_set_setting(C.byref(s1), key.encode(), value) == 0, key
        b0 = C.string_at(C.byref(s0), C.sizeof(s0))
        b1 = C.string_at(C.byref(s1), C.sizeof(s1))
        diff = [i for i in range(len(b0)) if b0[i] != b1[i]]
        assert diff, f"{key}: mbr_set_setting changed nothing (value equal to the default?)"
        lo, hi = _field_range(field)
        assert lo <= diff[0] and diff[-1] < hi, f"{key}: writes bytes {diff[0]}..{diff[-1]}, {field} is {lo}..{hi}"
        assert field not in seen, field
        seen.add(field)
