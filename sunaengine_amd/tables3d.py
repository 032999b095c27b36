"""Kernel and boundary tables of the 3D solver, built in IEEE float64.

Restates ``buildLUTs`` / ``buildBallLUT`` / ``matWords`` of the reference's
3D host (``3d/app.js:85-231``).  The reference builds these at run time using
only + - * / and sqrt, which IEEE-754 rounds exactly, so the same sequence of
float64 operations here yields the same integers on every host.

Image layout (int32 words, ``3d/sim3d.wgsl:63-76``): W[2049] | G[2049] |
planeW[513] | planeG[513] | ballW[1024] | ballG[1024] | 6 materials x 4 words.
"""
from __future__ import annotations

import math

import numpy as np

LUT_N = 2049
PN = 513
BALL_N = 1024
RHO0 = 1048576
H_WU = 2.0
SPACING = 1.0
OFF_W, OFF_G = 0, LUT_N
OFF_P, OFF_PG = 2 * LUT_N, 2 * LUT_N + PN
OFF_BW, OFF_BG = 2 * LUT_N + 2 * PN, 2 * LUT_N + 2 * PN + BALL_N
OFF_MAT = 2 * LUT_N + 2 * PN + 2 * BALL_N      # 7172
MAT_WORDS = 4
BALL_R_DEFAULT = 3.4

#: 3d/app.js:180-199 (grav, rhoW, damp, f0); f0 is renderer-only
MATERIALS3D = [
    ("water", 180, 0, 65340, 20),
    ("oil", 110, 30, 65180, 35),
    ("mercury", 440, 0, 65420, 700),
    ("gas", -46, 120, 64700, 10),
    ("orbies", 88, 210, 65330, 60),
    ("goo", 92, 50, 62900, 45),
]


def _js_round(x: float) -> int:
    f = math.floor(x)
    return int(f + 1) if (x - f) >= 0.5 else int(f)


def build_luts3d(ball_r: float = BALL_R_DEFAULT) -> np.ndarray:
    N, h = LUT_N, H_WU
    h2 = h * h
    rawW = [0.0] * N
    rawG = [0.0] * N
    for ui in range(N):
        r2 = ui / 512
        if r2 >= h2:
            continue
        t = h2 - r2
        rawW[ui] = t * t * t
        r = math.sqrt(r2)
        hr = h - r
        rawG[ui] = (hr * hr) / r if r > 1e-6 else 0.0
    for ui in range(N):
        if rawG[ui] > 0:
            rawG[0] = rawG[ui]
            break
    sum_raw = 0.0
    rad = math.ceil(h / SPACING)
    for i in range(-rad, rad + 1):
        for j in range(-rad, rad + 1):
            for k in range(-rad, rad + 1):
                r2 = float(i * i + j * j + k * k) * SPACING * SPACING
                if r2 >= h2:
                    continue
                t = h2 - r2
                sum_raw += t * t * t
    SW = RHO0 / sum_raw
    SG = SW * 6 * (h2 - 1) * (h2 - 1)
    planeW = [0] * PN
    planeG = [0] * PN
    sp = SPACING
    lat = math.ceil(h / sp)
    for b in range(PN):
        dist = b / 256
        sw = 0.0
        sg = 0.0
        k = 0
        while True:
            pd = dist + k * sp
            if pd >= h:
                break
            for i in range(-lat, lat + 1):
                for j in range(-lat, lat + 1):
                    r2 = pd * pd + (i * sp) * (i * sp) + (j * sp) * (j * sp)
                    if r2 >= h2:
                        continue
                    t = h2 - r2
                    sw += t * t * t
                    sg += pd * t * t
            k += 1
        planeW[b] = _js_round(sw * SW)
        planeG[b] = _js_round(6 * sg * SW)
    img = np.zeros(OFF_MAT + len(MATERIALS3D) * MAT_WORDS, dtype=np.int64)
    for ui in range(N):
        img[OFF_W + ui] = _js_round(rawW[ui] * SW)
        img[OFF_G + ui] = _js_round(rawG[ui] * SG)
    img[OFF_P:OFF_P + PN] = planeW
    img[OFF_PG:OFF_PG + PN] = planeG
    # cursor ball, radius ball_r (3d/app.js:213-230)
    def at(tab, dist_world):
        return tab[max(0, min(512, _js_round(dist_world * 256)))]
    for b in range(BALL_N):
        r = math.sqrt(b * 4096) / 256
        depth = r - ball_r
        if depth >= h:
            continue
        d = max(depth, 0.0)
        img[OFF_BW + b] = at(planeW, d)
        img[OFF_BG + b] = min(_js_round(at(planeG, d) / r), 1 << 24) if r > 1e-6 else 0
    for m, (_, grav, rhoW, damp, f0) in enumerate(MATERIALS3D):
        img[OFF_MAT + m * 4:OFF_MAT + m * 4 + 4] = (grav, rhoW, damp, f0)
    assert img.min() >= -(2 ** 31) and img.max() < 2 ** 31
    return img.astype(np.int32)


_CACHE = {}


def luts3d(ball_r: float = BALL_R_DEFAULT) -> np.ndarray:
    if ball_r not in _CACHE:
        _CACHE[ball_r] = build_luts3d(ball_r)
    return _CACHE[ball_r]
