"""Kernel lookup tables (W, grad W, cohesion, adhesion, s_corr) of the 2D solver.

The five tables are generated offline in IEEE float64 and frozen as 16010
little-endian int32 words.  This module restates the reference generator
(``src/tables.js:53-111``, ``buildTables``/``packTables``) and the table hash
(``src/tables.js:136-158``), and loads the frozen image that ships with this
package (``data/luts.bin``).

The generator is the product's own; the frozen image is pinned by the
reference's ``TABLE_HASH`` (``src/engine.js:501``) and by the lattice sum and
calibration factor the reference records in ``src/tables.json`` (``lattSum``,
``wcal``).  ``load_luts`` refuses an image whose hash differs.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

ONE = 65536
H = 163840            # kernel radius, 2.5 wu (src/tables.js:23)
DX0 = 65536           # rest spacing
RHO0 = 1 << 20
PRE = 4
HS = H >> PRE         # 10240
H2S = HS * HS         # 104857600
KSH = 15
TN = (H2S >> KSH) + 2  # 3202
GF = 22
CF = 22
OFF_W, OFF_G, OFF_C, OFF_A, OFF_S = 0, TN, 2 * TN, 3 * TN, 4 * TN
LUT_WORDS = 5 * TN

TABLE_PARAMS = dict(h=2.5, rmin=0.1, scorrK=0, scorrN=4, scorrDq=0.2)

#: 4-lane MurmurHash3 digest of the frozen image (src/engine.js:501).
TABLE_HASH = "31ebf1f3e1b168a800f0f180007dd390d0dc182fd06e398adfeec08e8a62908b"
#: intermediate f64 values recorded next to the image (src/tables.json).
REF_LATT_SUM = 0.9933543964580493
REF_WCAL = 1.0066900630486426

DATA_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "luts.bin")


def _js_round(x: float) -> int:
    """ECMAScript Math.round: nearest integer, ties toward +infinity."""
    f = math.floor(x)
    return int(f + 1) if (x - f) >= 0.5 else int(f)


def _r_of_idx(i: int) -> float:
    """Centre of LUT bin i in wu (src/tables.js:45)."""
    return math.sqrt((i + 0.5) * float(2 ** KSH)) * float(2 ** PRE) / ONE


def build_tables(cfg=None):
    """Return (image int32[16010], lattSum, wcal) computed in float64.

    Follows src/tables.js:53-111 step for step: poly6 density kernel, spiky
    gradient divided by r, Akinci cohesion and adhesion splines normalised to
    their peaks, the discrete rest-lattice calibration of W, and the s_corr
    table (identically zero at scorrK = 0).
    """
    cfg = dict(TABLE_PARAMS if cfg is None else cfg)
    h = cfg["h"]
    pi = math.pi

    def poly6(r):
        return 0.0 if r >= h else (4 / (pi * h ** 8)) * (h * h - r * r) ** 3

    def spiky_dr(r):
        return 0.0 if r >= h else -(30 / (pi * h ** 5)) * (h - r) ** 2

    def coh_raw(r):
        if r >= h or r <= 0:
            return 0.0
        a = (h - r) ** 3 * r ** 3
        return a if (2 * r > h) else 2 * a - h ** 6 / 64

    coh_peak = coh_raw(h / 2)

    def adh_raw(r):
        if 2 * r > h and r < h:
            return (-4 * r * r / h + 6 * r - 2 * h) ** 0.25
        return 0.0

    adh_peak = adh_raw(0.75 * h)
    rmin = cfg["rmin"] * h

    def bin_of(dx, dy):
        ax = (abs(dx) + 8) >> PRE
        ay = (abs(dy) + 8) >> PRE
        if ax > HS or ay > HS:
            return -1
        r2 = ax * ax + ay * ay
        return -1 if r2 >= H2S else (r2 >> KSH)

    reach = -(-H // DX0)
    latt = poly6(_r_of_idx(0))
    for jy in range(-reach, reach + 1):
        for jx in range(-reach, reach + 1):
            if jx == 0 and jy == 0:
                continue
            b = bin_of(jx * DX0, jy * DX0)
            if b >= 0:
                latt += poly6(_r_of_idx(b))
    wcal = 1 / latt

    img = np.zeros(LUT_WORDS, dtype=np.int64)
    for i in range(TN):
        r = _r_of_idx(i)
        rc = max(r, rmin)
        img[OFF_W + i] = _js_round(RHO0 * wcal * poly6(r))
        img[OFF_G + i] = _js_round(RHO0 * (spiky_dr(rc) / rc) * float(2 ** (GF - 16)))
        img[OFF_C + i] = _js_round((coh_raw(rc) / coh_peak / rc) * float(2 ** CF))
        img[OFF_A + i] = _js_round((adh_raw(rc) / adh_peak / rc) * float(2 ** CF))
    wdq = poly6(cfg["scorrDq"] * h)
    for i in range(TN):
        ratio = poly6(_r_of_idx(i)) / wdq if wdq > 0 else 0.0
        v = -_js_round(cfg["scorrK"] * ratio ** cfg["scorrN"])
        img[OFF_S + i] = v
    assert img.min() >= -(2 ** 31) and img.max() < 2 ** 31
    return img.astype(np.int32), latt, wcal


# ---------------------------------------------------------------- hashing
_M32 = 0xFFFFFFFF
_SEED = (0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F)
_SALT = (0x00000000, 0x51633E2D, 0x9C8F2A17, 0xD0E5B943)


def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & _M32


def mm_step(h, k):
    """MurmurHash3_x86_32 body step (src/hash/statehash.js:33-41)."""
    k = (k * 0xCC9E2D51) & _M32
    k = _rotl(k, 15)
    k = (k * 0x1B873593) & _M32
    h ^= k
    h = _rotl(h, 13)
    return (h * 5 + 0xE6546B64) & _M32


def fmix32(h):
    """MurmurHash3 finalizer (src/hash/statehash.js:44-51)."""
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & _M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & _M32
    return h ^ (h >> 16)


def _np_mm_step(h, k):
    k = (k * np.uint64(0xCC9E2D51)) & np.uint64(_M32)
    k = ((k << np.uint64(15)) | (k >> np.uint64(17))) & np.uint64(_M32)
    k = (k * np.uint64(0x1B873593)) & np.uint64(_M32)
    h = h ^ k
    h = ((h << np.uint64(13)) | (h >> np.uint64(19))) & np.uint64(_M32)
    return (h * np.uint64(5) + np.uint64(0xE6546B64)) & np.uint64(_M32)


def _np_fmix(h):
    h = h ^ (h >> np.uint64(16))
    h = (h * np.uint64(0x85EBCA6B)) & np.uint64(_M32)
    h = h ^ (h >> np.uint64(13))
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(_M32)
    return h ^ (h >> np.uint64(16))


BOND_SALT, BOND_SALT2 = 0x5BD1E995, 0x2F5B3C17


def digest_words(words, stride: int, hashed: int, bond=None, n_fluid: int = 0, bond_cap: int = 48):
    """8-lane digest of a word array laid out `stride` words per record.

    ``digestState`` (src/hash/statehash.js:80-159): lanes 0-3 are H_set,
    lanes 4-7 are H_slot.  With `bond` (n x bond_cap words, slot order) the
    rows of records below `n_fluid` are folded in: the bond count, the
    order-free sum over the rest lengths (H_set) and over the whole words
    (H_slot), up to the first zero word.  Vectorised over records.
    """
    u = np.asarray(words).view(np.uint32).astype(np.uint64).reshape(-1, stride)
    n_rec = u.shape[0]
    acc = []
    lanes = []
    for lane in range(4):
        h = np.full(n_rec, _SEED[lane], dtype=np.uint64)
        for w in range(hashed):
            h = _np_mm_step(h, u[:, w] ^ np.uint64(_SALT[lane]))
        lanes.append(h)
    s_lanes, l_lanes = list(lanes), list(lanes)
    if bond is not None and n_fluid > 0:
        b = np.asarray(bond).view(np.uint32).astype(np.uint64).reshape(-1, bond_cap)[:n_fluid]
        live = np.cumprod(b != 0, axis=1).astype(bool)          # the dense prefix of each row
        bn = live.sum(axis=1).astype(np.uint64)
        com = (np.where(live, _np_fmix(((b >> np.uint64(17)) & np.uint64(0x1FFF)) ^ np.uint64(BOND_SALT)), 0)
               .sum(axis=1) & np.uint64(_M32))
        scom = np.where(live, _np_fmix(b ^ np.uint64(BOND_SALT2)), 0).sum(axis=1) & np.uint64(_M32)
        has = np.zeros(n_rec, bool)
        has[:n_fluid] = bn != 0
        bnf = np.zeros(n_rec, np.uint64); bnf[:n_fluid] = bn
        comf = np.zeros(n_rec, np.uint64); comf[:n_fluid] = com
        scomf = np.zeros(n_rec, np.uint64); scomf[:n_fluid] = scom
        for lane in range(4):
            t = _np_mm_step(lanes[lane], bnf)
            s_lanes[lane] = np.where(has, t ^ comf, lanes[lane])
            l_lanes[lane] = np.where(has, t ^ scomf, lanes[lane])
    n = np.uint64(hashed)
    m = np.uint64(hashed + 1)
    idx = np.arange(n_rec, dtype=np.uint64)
    for lane in range(4):
        acc.append(int(_np_fmix(s_lanes[lane] ^ n).sum()) & _M32)
    for lane in range(4):
        acc.append(int(_np_fmix(_np_mm_step(l_lanes[lane], idx) ^ m).sum()) & _M32)
    return acc


def hex8(words) -> str:
    return "".join("%08x" % (w & _M32) for w in words)


def table_hash(img) -> str:
    return hex8(digest_words(np.asarray(img, dtype=np.int32), 1, 1))


def write_frozen(path: str = DATA_PATH) -> str:
    img, latt, wcal = build_tables()
    h = table_hash(img)
    if h != TABLE_HASH:
        raise RuntimeError(f"generated LUT image hashes to {h}, expected {TABLE_HASH}")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(struct.pack("<%di" % LUT_WORDS, *[int(v) for v in img]))
    return h


def load_luts(path: str = DATA_PATH) -> np.ndarray:
    """Load and verify the frozen image; a hash mismatch is a hard error."""
    raw = np.fromfile(path, dtype="<i4")
    if raw.size != LUT_WORDS:
        raise RuntimeError(f"{path}: {raw.size} words, expected {LUT_WORDS}")
    h = table_hash(raw)
    if h != TABLE_HASH:
        raise RuntimeError(f"{path}: TABLE_HASH mismatch {h} != {TABLE_HASH}")
    return raw.astype(np.int32)


if __name__ == "__main__":
    print(write_frozen())
