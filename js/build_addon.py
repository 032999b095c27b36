"""Build the N-API addon js/build/sunaengine_napi.node with gcc (no node-gyp).

The addon links against sunaengine_amd/libsunaengine_amd.so through an
$ORIGIN-relative rpath; the napi_* symbols resolve against the running node.
Also writes js/build/mats.json (the material image the JS host uploads).
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build():
    out = os.path.join(HERE, "build")
    os.makedirs(out, exist_ok=True)
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=c11", "-DNODE_GYP_MODULE_NAME=sunaengine_napi",
           "-I/usr/include/node", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "addon.c"), "-o", os.path.join(out, "sunaengine_napi.node"),
           "-L", os.path.join(ROOT, "sunaengine_amd"), "-lsunaengine_amd",
           "-Wl,-rpath,$ORIGIN/../../sunaengine_amd"]
    subprocess.run(cmd, check=True)
    sys.path.insert(0, ROOT)
    from sunaengine_amd import materials
    img = materials.mats_image()
    with open(os.path.join(out, "mats.json"), "w") as f:
        json.dump({"fields": materials.FIELDS, "names": materials.NAMES,
                   "rows": [list(map(int, r)) for r in materials.MATS],
                   "image": [int(x) for x in img.reshape(-1)]}, f)
    return out


if __name__ == "__main__":
    print(build())
