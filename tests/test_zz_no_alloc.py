"""sn_digest_state_noalloc, the state digest that allocates nothing, against
sn_digest_state (the entry point it is a variant of), through the C ABI.

Two calls in a row share the library's fixed device array; the second, with
a smaller input, must not see anything the first left there.  Both are
checked against sn_digest_state on the same words, bit-exact.
"""
import numpy as np
import pytest

from sunaengine_amd import native

pytestmark = pytest.mark.gpu


def _dev(a):
    from tests.helpers import DeviceWords
    return DeviceWords(a)


def _both(words, stride, hashed, bond=None, n_fluid=None, bond_cap=0):
    t = _dev(words)
    b = None if bond is None else _dev(bond)
    n = t.words // stride
    args = (t.ptr, t.words, n, stride, hashed, n_fluid, None if b is None else b.ptr, bond_cap)
    ref = native.digest_state(*args, entry="sn_digest_state")
    got = native.digest_state(*args, entry="sn_digest_state_noalloc")
    return ref, got


def test_noalloc_equals_digest_state_two_calls_shrinking():
    rng = np.random.default_rng(2024)
    big = rng.integers(0, 2 ** 32, size=3_000_000 * 8, dtype=np.uint64).astype(np.uint32)
    small = rng.integers(0, 2 ** 32, size=1_500 * 8, dtype=np.uint64).astype(np.uint32)
    print("noalloc: call 1, 3,000,000 particles", flush=True)
    ref1, got1 = _both(big, 8, 8)
    assert np.array_equal(got1, ref1)
    print("noalloc: call 2, 1,500 particles", flush=True)
    ref2, got2 = _both(small, 8, 8)
    assert np.array_equal(got2, ref2)
    assert not np.array_equal(got1, got2)


def test_noalloc_bond_fold_and_engine_state():
    """The bond fold and a real engine state (dambreak_v1 at substep 10,
    pinned by the reference's frozen checkpoint)."""
    from sunaengine_amd import scenes, tables
    from sunaengine_amd.engine import Engine
    from tests.helpers import golden
    rng = np.random.default_rng(5)
    n, nf = 9000, 7000
    st = rng.integers(0, 2 ** 32, size=n * 8, dtype=np.uint64).astype(np.uint32)
    bd = rng.integers(1, 2 ** 31, size=(n, 48), dtype=np.uint64).astype(np.uint32)
    bd[:, 20:] = 0
    print("noalloc: bond rows", flush=True)
    ref, got = _both(st, 8, 8, bond=bd, n_fluid=nf, bond_cap=48)
    assert np.array_equal(got, ref)
    e = Engine(max_particles=8192, max_cells=65536, iters=4)
    e.load_scene(scenes.make_scene("dambreak_v1"))
    e.step(10)
    print("noalloc: dambreak_v1 state", flush=True)
    ref, got = _both(e.read_state(), 8, 8)
    assert tables.hex8(got) == golden("dambreak_v1.json")["checkpoints"]["10"]
    assert np.array_equal(got, ref)
    e.destroy()


def test_noalloc_rejects_sizes_beyond_its_array():
    bogus = 0x1000  # never dereferenced: the size check fails first
    lim = native.SN_DIGEST_NOALLOC_MAX_PARTICLES
    with pytest.raises(native.SnError) as ei:
        native.digest_state(bogus, (lim + 1) * 8, lim + 1, 8, 8, entry="sn_digest_state_noalloc")
    assert ei.value.code == native.SN_E_CAPACITY
    assert "SN_DIGEST_NOALLOC_MAX_PARTICLES" in str(ei.value)
