"""Frozen LUT image and the MurmurHash3 pieces of the state hash (CPU)."""
import numpy as np

from sunaengine_amd import tables


def test_generator_reproduces_table_hash():
    img, latt, wcal = tables.build_tables()
    assert latt == tables.REF_LATT_SUM
    assert wcal == tables.REF_WCAL
    assert tables.table_hash(img) == tables.TABLE_HASH


def test_frozen_image_loads_and_matches_generator():
    img = tables.load_luts()
    gen, _, _ = tables.build_tables()
    assert np.array_equal(img, gen)
    # structural facts the solver relies on (src/sim.wgsl:1245-1249)
    assert img[tables.OFF_W] == 214942            # W(0), the self term
    assert (img[tables.OFF_S:tables.OFF_S + tables.TN] == 0).all()   # scorrK = 0


def test_frozen_image_hash_detects_a_flipped_entry():
    img = tables.load_luts().copy()
    img[1234] ^= 1
    assert tables.table_hash(img) != tables.TABLE_HASH


def _murmur3_x86_32(data: bytes, seed: int) -> int:
    """Textbook MurmurHash3_x86_32 built from the hash's own body step and
    finalizer, with the standard tail and length mix."""
    h = seed & 0xFFFFFFFF
    nblocks = len(data) // 4
    for i in range(nblocks):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        h = tables.mm_step(h, k)
    tail = data[4 * nblocks:]
    k1 = 0
    if len(tail) >= 3:
        k1 ^= tail[2] << 16
    if len(tail) >= 2:
        k1 ^= tail[1] << 8
    if len(tail) >= 1:
        k1 ^= tail[0]
        k1 = (k1 * 0xCC9E2D51) & 0xFFFFFFFF
        k1 = ((k1 << 15) | (k1 >> 17)) & 0xFFFFFFFF
        k1 = (k1 * 0x1B873593) & 0xFFFFFFFF
        h ^= k1
    h ^= len(data)
    return tables.fmix32(h)


def test_murmur3_smhasher_verification_value():
    """SMHasher's VerificationTest for MurmurHash3_x86_32 is 0xB0F57EE3."""
    key = bytearray()
    hashes = bytearray()
    for i in range(256):
        hashes += _murmur3_x86_32(bytes(key), 256 - i).to_bytes(4, "little")
        key.append(i)
    final = _murmur3_x86_32(bytes(hashes), 0)
    assert final == 0xB0F57EE3


def test_digest_permutation_properties():
    """H_set is invariant under a permutation of records; H_slot is not
    (src/hash/statehash.js:14-18)."""
    rng = np.random.default_rng(5)
    w = rng.integers(-2 ** 31, 2 ** 31, size=(64, 8), dtype=np.int64).astype(np.int32)
    a = tables.digest_words(w.reshape(-1), 8, 8)
    p = rng.permutation(64)
    b = tables.digest_words(w[p].reshape(-1), 8, 8)
    assert a[:4] == b[:4]
    assert a[4:] != b[4:]
