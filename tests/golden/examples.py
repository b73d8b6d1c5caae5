"""Loader of example_scenes.json (written by make_example_scenes.py): the
reference's example scenes at 64x48 as mbr_scene structures."""
import json
import os

from mandelbulber_amd import scene as S

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_scenes.json")


def _assign(obj, name, value):
    cur = getattr(obj, name)
    if hasattr(cur, "_fields_"):
        for k, v in value.items():
            _assign(cur, k, v)
    elif isinstance(value, list):
        _assign_array(cur, value)
    else:
        setattr(obj, name, value)


def _assign_array(arr, values):
    for i, v in enumerate(values):
        if isinstance(v, dict):
            for k, x in v.items():
                _assign(arr[i], k, x)
        elif isinstance(v, list):
            _assign_array(arr[i], v)
        else:
            arr[i] = v


def load_examples():
    """[(file name, mbr_scene)] for every example scene."""
    out = []
    with open(PATH) as f:
        entries = json.load(f)
    for e in entries:
        s = S.defaults()
        fields = dict(e["fields"])
        pal = bytes.fromhex(fields.pop("palette"))
        for k, v in fields.items():
            _assign(s, k, v)
        for i in range(256):
            s.palette[i].R, s.palette[i].G, s.palette[i].B = pal[3 * i], pal[3 * i + 1], pal[3 * i + 2]
        out.append((e["name"], s))
    return out
