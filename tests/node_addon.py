"""Build and drive the N-API addon (break-that-denim_amd/node) from the tests: node-gyp against the
image's Node headers, and a scene exported as JSON in the camelCase layout the addon reads."""
import ctypes as C
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np

from bd import capi

ROOT = Path(__file__).resolve().parent.parent
NODE_DIR = ROOT / "break-that-denim_amd" / "node"
ADDON = NODE_DIR / "build" / "Release" / "bd.node"


def toolchain():
    return shutil.which("node") is not None and shutil.which("node-gyp") is not None


def build():
    """node-gyp against the installed headers (no download); the addon links the in-tree libbd.so."""
    subprocess.run(["node-gyp", "rebuild", "--nodedir=/usr"], cwd=NODE_DIR, check=True, capture_output=True, timeout=300)
    return ADDON


def _camel(name):
    head, *rest = name.split("_")
    return head + "".join(w.capitalize() for w in rest)


def _struct(s):
    out = {}
    for name, _ in s._fields_:
        if name.endswith("_"):
            continue
        v = getattr(s, name)
        if isinstance(v, C.Structure):
            v = _struct(v)
        elif isinstance(v, C.Array):
            v = [_struct(e) if isinstance(e, C.Structure) else e for e in v]
        out[_camel(name)] = v
    return out


def export(path, scene, params, grips=None, goals=None, steps=1):
    sb = capi.SceneBuffers(scene)
    k = sb.keep
    names = {"rest_length": "restLength", "seg_flags": "segFlags", "cut_mask": "cutMask", "fail_site": "failSite",
             "crossing_normal0": "crossingNormal0", "field_strength": "fieldStrength", "field_damage": "fieldDamage",
             "field_area": "fieldArea"}
    js = {"nodeCount": scene.node_count, "crossingCount": scene.crossing_count,
          "layers": [_struct(sb.layers[i]) for i in range(len(scene.layers))]}
    for key, arr in k.items():
        js[names.get(key, key)] = np.asarray(arr).tolist()
    doc = {"scene": js, "params": _struct(params), "steps": steps}
    if grips:
        gb = capi.GripBuffers(grips)
        doc["grips"] = {"node": gb.node.tolist(), "grip": gb.grip.tolist(), "weight": gb.weight.tolist(),
                        "offset": gb.offset.tolist(), "goal": gb.goal.tolist(), "stiffness": gb.k.tolist()}
    if goals is not None:
        doc["goals"] = [np.asarray(g, dtype=np.float64).ravel().tolist() for g in goals]
    Path(path).write_text(json.dumps(doc))
