"""The Node.js host (js/): CPU-side checks run here; the GPU run is marked gpu."""
import json
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ADDON = os.path.join(ROOT, "js", "build", "sunaengine_napi.node")
MATS = os.path.join(ROOT, "js", "build", "mats.json")

needs_node = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


def _expected_scenes():
    sys.path.insert(0, ROOT)
    from sunaengine_amd import scenes, tables
    out = {}
    for name, kw in (("dambreak_v1", {}), ("overdense_v1", {}), ("pool_v1", {"w": 60, "h": 20})):
        s = scenes.make_scene(name, **kw)
        im = scenes.load_image(s)
        acc = tables.digest_words(im.state.view("uint32").reshape(-1), 8, 8)
        out[name] = {"n": int(im.n), "nFluid": int(im.nFluid), "digest": tables.hex8(acc)}
    return out


@needs_node
def test_node_host_cpu_side():
    if not os.path.exists(MATS):
        pytest.skip("js/build/mats.json not built (python js/build_addon.py)")
    arg = json.dumps({"scenes": _expected_scenes()})
    r = subprocess.run(["node", os.path.join(HERE, "js", "host_cpu.js"), arg], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr


@pytest.mark.gpu
@needs_node
def test_node_host_on_gpu():
    assert os.path.exists(ADDON), "N-API addon not built"
    with open(os.path.join(HERE, "golden", "dambreak_v1.json")) as f:
        g = json.load(f)
    r = subprocess.run(["node", os.path.join(HERE, "js", "host_gpu.js"), json.dumps(g["checkpoints"])],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "OK node host" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
@needs_node
def test_node_serving_mix_equals_alone():
    """The JavaScript serving loop (js/serve.js): jobs in the mix = alone."""
    assert os.path.exists(ADDON), "N-API addon not built"
    r = subprocess.run(["node", os.path.join(HERE, "js", "serve_gpu.js")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "OK js serving mix = alone" in r.stdout, r.stdout + r.stderr


@needs_node
def test_node_tables3d_match_python():
    """js/tables3d.js builds the same 3D table image as tables3d.py."""
    sys.path.insert(0, ROOT)
    import numpy as np
    from sunaengine_amd import tables3d
    code = ("const t=require(%r);const a=t.buildLuts3d();"
            "process.stdout.write(Buffer.from(a.buffer).toString('hex'))" % os.path.join(ROOT, "js", "tables3d.js"))
    r = subprocess.run(["node", "-e", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.frombuffer(bytes.fromhex(r.stdout), dtype=np.int32)
    assert np.array_equal(got, tables3d.luts3d())


@pytest.mark.gpu
@needs_node
@pytest.mark.parametrize("name", ["stone_in_water", "freezing_pool"])
def test_node_material_scene_matches_oracle(tmp_path, name):
    """The Node host runs the material passes: same per-substep hashes and
    bond rows as the oracle (freezing_pool needs createEngine({materials: true}))."""
    assert os.path.exists(ADDON), "N-API addon not built"
    sys.path.insert(0, ROOT)
    import numpy as np
    from tests import material_scenes as ms
    from tests.helpers import oracle_world
    build, _, materials = ms.CASES[name]
    s = build()
    steps = 16
    doc = {"n": s.n, "nFluid": s.nFluid, "domW": s.domW, "domH": s.domH, "materials": materials,
           "pos": [int(v) for v in s.pos], "vel": [int(v) for v in s.vel], "matId": [int(v) for v in s.matId],
           "flags": None if s.flags is None else [int(v) for v in s.flags],
           "temp": None if s.temp is None else [int(v) for v in s.temp]}
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(doc))
    r = subprocess.run(["node", os.path.join(HERE, "js", "materials_gpu.js"), str(path), str(steps)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    o = oracle_world(s, 4)
    ref = o.step(steps, stream=True)
    want = ["".join("%08x" % int(w) for w in row) for row in ref]
    assert got["stream"] == want
    bond = o.read_bond().reshape(-1)
    nz = np.nonzero(bond)[0]
    assert got["bonds"] == nz.size > 0
    assert got["bondSum"] == int((bond[nz].astype(np.uint64) * (nz.astype(np.uint64) + 1)).sum() % 2**32)

