"""Writes tests/golden/example_scenes.json: the reference's example scenes
(share/examples/*.fract) as loaded by mbr_load_settings, shrunk to 64x48,
for GPU-vs-oracle renders on a machine without the reference tree.

Each entry holds, by name, the fields of the scene (mbr_scene, the ctypes
mirror in mandelbulber_amd/_abi.py) that differ from mbr_scene_defaults, so
the fixture survives reordering of the struct; the palette is a hex string of
its 256 RGB triples.  examples.load_example() rebuilds a scene.
Post effects are switched off (they act on the compiled image, not on
the render).  Run in a container
that has the reference tree:  python tests/golden/make_example_scenes.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
EXAMPLES = "/root/reference/share/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_scenes.json")
W, H = 64, 48


def to_json(v):
    """A ctypes value (structure, array, scalar) as JSON data."""
    if hasattr(v, "_fields_"):
        return {name: to_json(getattr(v, name)) for name, _ in v._fields_}
    if isinstance(v, bytes):
        return v.decode("latin-1")
    if isinstance(v, (int, float)):
        return v
    return [to_json(x) for x in v]


def main():
    from mandelbulber_amd import scene as S
    entries = []
    for name in sorted(os.listdir(EXAMPLES)):
        if not name.endswith(".fract"):
            continue
        s = S.load(os.path.join(EXAMPLES, name))
        s.image_width, s.image_height, s.tiles = W, H, 1
        s.ssao_enabled = s.dof_enabled = 0
        got, base = to_json(s), to_json(S.defaults())
        diff = {k: v for k, v in got.items() if v != base[k] and k != "palette"}
        diff["palette"] = bytes(c for e in got["palette"] for c in (e["R"], e["G"], e["B"])).hex()
        entries.append({"name": name, "fields": diff})
    with open(OUT, "w") as f:
        json.dump(entries, f, separators=(",", ":"))
    print(f"wrote {len(entries)} scenes to {OUT}")


if __name__ == "__main__":
    main()
