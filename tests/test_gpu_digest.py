"""GPU parity of the stateless state digest sn_digest_state, the reference's
digestState (src/hash/statehash.js:80-159) over device-resident words.

Pinned by the reference's own TABLE_HASH (digestState of the frozen LUT
image with strideWords = hashedWords = 1, src/engine.js:501), by the
engine's per-substep H of a scene the dambreak_v1 goldens pin, by the bond
fold of a material world's rows, by the 3D H_slot definition (stride 16,
hashed 8), and by the numpy twin on random words at ragged sizes.
"""
import numpy as np
import pytest

from sunaengine_amd import native, tables

pytestmark = pytest.mark.gpu


def _dev(a):
    from tests.helpers import DeviceWords
    return DeviceWords(a)


def _digest(words, stride, hashed, bond=None, n_fluid=None, bond_cap=0, entry="sn_digest_state"):
    t = _dev(words)
    b = None if bond is None else _dev(bond)
    n = t.words // stride if stride else 0
    return native.digest_state(t.ptr, t.words, n, stride, hashed, n_fluid,
                               None if b is None else b.ptr, bond_cap, entry=entry)


def test_table_hash_on_device():
    """The LUT image digests to the reference's frozen TABLE_HASH."""
    luts = tables.load_luts()
    print(f"digest: LUT image {luts.size} words", flush=True)
    assert tables.hex8(_digest(luts, 1, 1)) == tables.TABLE_HASH


def test_engine_state_digest_equals_engine_hash():
    """digestState over readState's words = the engine's H (dambreak_v1 at
    substep 10, itself pinned by the frozen checkpoint)."""
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    from tests.helpers import golden
    g = golden("dambreak_v1.json")
    e = Engine(max_particles=8192, max_cells=65536, iters=4)
    e.load_scene(scenes.make_scene("dambreak_v1"))
    e.step(10)
    st = e.read_state()
    print(f"digest: dambreak_v1 state n={st.shape[0]}", flush=True)
    acc = _digest(st, 8, 8)
    assert tables.hex8(acc) == g["checkpoints"]["10"]
    assert np.array_equal(acc, e.digest()["acc"])
    e.destroy()


def test_bond_fold_equals_material_world_hash():
    """With readBond's rows and bondCap 48 the digest folds bonds exactly as
    the engine's bonded hash does."""
    from tests import material_scenes as ms
    from sunaengine_amd import scenes
    from sunaengine_amd.engine import Engine
    s = ms.stone_in_water()
    im = scenes.load_image(s)
    e = Engine(max_particles=im.n, max_cells=im.cellTotal, iters=4, materials=True)
    e.load_scene(s)
    e.step(12)
    st, bd = e.read_state(), e.read_bond()
    assert np.count_nonzero(bd) > 0, "the scene must hold bonds"
    print(f"digest: stone_in_water n={st.shape[0]} bonds={np.count_nonzero(bd)}", flush=True)
    acc = _digest(st, 8, 8, bond=bd, n_fluid=im.nFluid, bond_cap=48)
    assert np.array_equal(acc, e.digest()["acc"])
    assert acc.tolist() == tables.digest_words(st, 8, 8, bond=bd, n_fluid=im.nFluid)
    # without the rows the fold does not run, as with an absent p.bond
    assert _digest(st, 8, 8).tolist() == tables.digest_words(st, 8, 8)
    e.destroy()


def test_3d_hash_is_digest_stride16():
    """The 3D H_slot is digestState over p.xyzw, v.xyzw of 16-word records."""
    from sunaengine_amd import scenes3d, tables3d
    from sunaengine_amd.engine3d import Engine3D
    s = scenes3d.column_scene(3000)
    e = Engine3D(max_particles=s.n, max_cells=s.cell_total, iters=3)
    e.load([s])
    e.step(7)
    st = e.read_state()
    print(f"digest: 3d column n={st.shape[0]}", flush=True)
    assert np.array_equal(_digest(st, 16, 8), e.digest())
    e.destroy()


@pytest.mark.parametrize("count,stride,hashed", [(0, 8, 8), (1, 8, 8), (1023, 8, 8), (1025, 3, 2),
                                                 (70001, 16, 8), (3_000_017, 8, 8), (4097, 5, 0)])
def test_random_words_match_twin(count, stride, hashed):
    rng = np.random.default_rng(count * 31 + stride)
    w = rng.integers(0, 2 ** 32, size=count * stride, dtype=np.uint64).astype(np.uint32)
    print(f"digest: random count={count} stride={stride} hashed={hashed}", flush=True)
    if count:
        got = _digest(w, stride, hashed)
    else:
        one = _dev(np.zeros(1, np.uint32))
        got = native.digest_state(one.ptr, 1, 0, stride, hashed)
    assert got.tolist() == tables.digest_words(w, stride, hashed)


def test_random_bond_rows_match_twin():
    """Bond rows with zero words inside (the fold stops at the first zero)
    and rows past n_fluid (never folded)."""
    rng = np.random.default_rng(7)
    n, nf = 5000, 4100
    st = rng.integers(0, 2 ** 32, size=n * 8, dtype=np.uint64).astype(np.uint32)
    bd = rng.integers(1, 2 ** 31, size=(n, 48), dtype=np.uint64).astype(np.uint32)
    cut = rng.integers(0, 49, size=n)
    for i in range(n):
        bd[i, cut[i]:] = 0
        if i % 7 == 0 and cut[i] > 3:
            bd[i, 2] = 0          # a zero inside: words after it are not folded
    print("digest: random bond rows", flush=True)
    got = _digest(st, 8, 8, bond=bd, n_fluid=nf, bond_cap=48)
    assert got.tolist() == tables.digest_words(st, 8, 8, bond=bd, n_fluid=nf)
