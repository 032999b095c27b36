"""The material table: 48 int32 words per row, MAT_MAX = 64 rows.

Rows are data.  The field order is the wire layout of the reference's
``struct Mat`` (``src/sim.wgsl:315-431``, field list ``src/engine.js:680-707``),
defaults follow ``MAT_DEFAULTS`` (``src/engine.js:715-733``), the
gammaCur = gammaCoh/4 default and the ``rampInv`` bake follow ``mat()``
(``src/engine.js:736-755``) and ``createEngine`` (``src/engine.js:3359-3376``).
The 28 authored rows restate ``MATS`` (``src/engine.js:1192-2060``); the
numbers must be the reference's for any material-dependent hash to match.
"""
from __future__ import annotations

import numpy as np

TEMP_ONE = 65536
AMBIENT = 20 * TEMP_ONE
TEMP_MIN = -33554432
TEMP_MAX = 536870912
NEVER_HOT = 2147483647
NEVER_COLD = -2147483648
MAT_WORDS = 48
MAT_MAX = 64

SOLID_NONE, SOLID_BOND, SOLID_RIGID = 0, 1, 2


def DEG(d: int) -> int:
    return d * TEMP_ONE


FIELDS = [
    "eps", "dpMax", "xsph", "phantom", "gammaCoh", "gammaCur", "adhesion", "grav",
    "cond", "coolRate", "heatCap",
    "meltPt", "meltTo", "freezePt", "freezeTo", "boilPt", "boilTo", "condPt", "condTo",
    "ignitePt", "fuel0", "burnRate", "burnHeat", "burnTo",
    "splitTo", "splitPct", "spawnTemp", "tint",
    "corrode", "corrPt", "corrTo", "fric",
    "bondK", "bondBreak", "bondReform", "bondFormV", "bondFormP", "bondMu",
    "bondGroup", "rampLo", "rampHi", "rampInv", "rampSign",
    "solidMode", "tempTarget", "tempDrive", "burstP", "sealMode",
]
assert len(FIELDS) == MAT_WORDS
F = {name: i for i, name in enumerate(FIELDS)}

DEFAULTS = dict(
    eps=40000, dpMax=16384, xsph=6000, phantom=0,
    gammaCoh=240, gammaCur=60, adhesion=0, grav=1200,
    cond=0, coolRate=0, heatCap=0,
    meltPt=NEVER_HOT, meltTo=0, freezePt=NEVER_COLD, freezeTo=0,
    boilPt=NEVER_HOT, boilTo=0, condPt=NEVER_COLD, condTo=0,
    ignitePt=NEVER_HOT, fuel0=0, burnRate=0, burnHeat=0, burnTo=0,
    splitTo=0, splitPct=0, spawnTemp=AMBIENT, tint=0x8090A0,
    corrode=0, corrPt=NEVER_HOT, corrTo=0, fric=0,
    bondK=0, bondBreak=0, bondReform=0, bondFormV=0, bondFormP=0, bondMu=0,
    bondGroup=-1, rampLo=0, rampHi=65536, rampInv=0, rampSign=0,
    solidMode=SOLID_NONE, tempTarget=0, tempDrive=0, burstP=0, sealMode=0,
)

(WATER, GEL, ICE, STEAM, OIL, FIRE, SMOKE, WOOD, ASH, LAVA, STONE, SAND, METAL,
 MOLTEN, GAS, PHOTO, CRYO, RUBBER, TAR, MERCURY, DUST, ACID, NITRO, BLAST, VOID,
 RICE, BEDROCK, ORBIES) = range(28)
NAMES = ["WATER", "GEL", "ICE", "STEAM", "OIL", "FIRE", "SMOKE", "WOOD", "ASH",
         "LAVA", "STONE", "SAND", "METAL", "MOLTEN", "GAS", "PHOTO", "CRYO",
         "RUBBER", "TAR", "MERCURY", "DUST", "ACID", "NITRO", "BLAST", "VOID",
         "RICE", "BEDROCK", "ORBIES"]


def mat(**o) -> list:
    """Named fields -> positional row, with the reference's defaulting rules."""
    for k in o:
        if k not in DEFAULTS:
            raise KeyError(f"unknown material field {k!r}")
    row = [o.get(f, DEFAULTS[f]) for f in FIELDS]
    if "gammaCoh" in o and "gammaCur" not in o:
        row[F["gammaCur"]] = int(o["gammaCoh"] / 4)   # JS `(x / 4) | 0` truncates
    span = row[F["rampHi"]] - row[F["rampLo"]]
    row[F["rampInv"]] = (2 ** 32) // span if span > 0 else 0
    return row


MATS = [
    mat(xsph=6000, gammaCoh=240, gammaCur=60, adhesion=0, grav=1200,
        cond=9000, coolRate=40, freezePt=DEG(0), freezeTo=ICE,
        boilPt=DEG(100), boilTo=STEAM, spawnTemp=DEG(20), tint=0x2A6FB0),
    mat(fric=49152, xsph=45000, gammaCoh=1700, gammaCur=425, adhesion=300, grav=1080,
        cond=0, coolRate=0, spawnTemp=DEG(20), tint=0x8CECFF,
        bondK=12288, bondBreak=262144, bondReform=1, bondFormV=2600, bondFormP=2000,
        bondMu=0, bondGroup=10, rampSign=0, solidMode=SOLID_BOND),
    mat(xsph=36000, gammaCoh=700, adhesion=120, grav=1050, cond=16000, coolRate=40,
        meltPt=DEG(2), meltTo=WATER, corrPt=9000, corrTo=WATER, fric=114688,
        spawnTemp=DEG(-20), tint=0xA8E0FF, bondK=32768, bondBreak=102400,
        bondReform=0, bondMu=0, bondGroup=5, burstP=98304,
        rampLo=DEG(-8), rampHi=DEG(2), rampSign=-1, solidMode=SOLID_RIGID),
    mat(xsph=3000, gammaCoh=40, adhesion=0, grav=-700, cond=4000, coolRate=110,
        condPt=DEG(55), condTo=WATER, spawnTemp=DEG(120), tint=0xD8E8F0),
    mat(xsph=12000, gammaCoh=300, adhesion=180, grav=900, cond=5000, coolRate=300,
        ignitePt=DEG(180), fuel0=DEG(100), burnRate=21000, burnHeat=264000,
        burnTo=SMOKE, spawnTemp=DEG(20), tint=0x6B4A1F),
    mat(xsph=3000, gammaCoh=30, adhesion=0, grav=-350, cond=12000, coolRate=900,
        ignitePt=TEMP_MIN, fuel0=DEG(60), burnRate=26000, burnHeat=882000,
        burnTo=SMOKE, spawnTemp=DEG(900), tint=0xFF7A1E),
    mat(xsph=2000, gammaCoh=25, adhesion=0, grav=-420, cond=3000, coolRate=600,
        condPt=DEG(22), condTo=VOID, splitTo=ASH, splitPct=6,
        spawnTemp=DEG(120), tint=0x5A5A62),
    mat(fric=114688, xsph=40000, gammaCoh=800, adhesion=400, grav=1200,
        cond=3000, coolRate=120, ignitePt=DEG(200), fuel0=DEG(200), burnRate=12000,
        burnHeat=93600, burnTo=SMOKE, splitTo=ASH, splitPct=26,
        corrPt=16000, corrTo=ASH, spawnTemp=DEG(20), tint=0x8A5A2B,
        bondK=26214, bondBreak=110756, bondReform=0, bondMu=0, bondGroup=4,
        burstP=49152, sealMode=0, rampLo=DEG(260), rampHi=DEG(420), rampSign=-1,
        solidMode=SOLID_RIGID),
    mat(fric=16384, xsph=2500, gammaCoh=24, gammaCur=6, adhesion=0, grav=520,
        cond=2000, coolRate=300, spawnTemp=DEG(20), tint=0x4A4844),
    mat(xsph=38000, gammaCoh=500, adhesion=300, grav=1500, cond=14000, coolRate=60,
        freezePt=DEG(700), freezeTo=STONE, spawnTemp=DEG(1200), tint=0xFF5A10),
    mat(fric=131072, xsph=40000, gammaCoh=900, adhesion=500, grav=1600,
        cond=6000, coolRate=60, meltPt=DEG(950), meltTo=LAVA,
        corrPt=14000, corrTo=SAND, spawnTemp=DEG(20), tint=0x6E6A66,
        bondK=32768, bondBreak=0, bondReform=0, bondMu=0, bondGroup=2,
        burstP=163840, rampLo=DEG(700), rampHi=DEG(950), rampSign=-1,
        solidMode=SOLID_RIGID),
    mat(fric=98304, xsph=8000, gammaCoh=60, adhesion=0, grav=1800,
        cond=4000, coolRate=100, meltPt=DEG(1200), meltTo=LAVA,
        spawnTemp=DEG(20), tint=0xD8C48A),
    mat(fric=131072, xsph=40000, gammaCoh=900, adhesion=400, grav=2000,
        cond=30000, coolRate=80, meltPt=DEG(1200), meltTo=MOLTEN,
        corrPt=12000, corrTo=GAS, spawnTemp=DEG(20), tint=0xB9C2CC,
        bondK=32768, bondBreak=0, bondReform=0, bondMu=0, bondGroup=3,
        burstP=294912, rampLo=DEG(800), rampHi=DEG(1200), rampSign=-1,
        solidMode=SOLID_RIGID),
    mat(xsph=20000, gammaCoh=700, adhesion=200, grav=2000, cond=30000, coolRate=90,
        freezePt=DEG(1050), freezeTo=METAL, spawnTemp=DEG(1400), tint=0xFFD07A),
    mat(xsph=1500, gammaCoh=15, adhesion=0, grav=-600, cond=6000, coolRate=400,
        ignitePt=DEG(120), fuel0=DEG(25), burnRate=40000, burnHeat=432000,
        burnTo=FIRE, spawnTemp=DEG(20), tint=0xA8FF9A),
    mat(fric=114688, xsph=40000, gammaCoh=700, adhesion=350, grav=1200,
        cond=2500, coolRate=120, ignitePt=DEG(140), fuel0=DEG(120), burnRate=10000,
        burnHeat=88800, burnTo=SMOKE, splitTo=ASH, splitPct=26,
        corrPt=10000, corrTo=ASH, spawnTemp=DEG(20), tint=0xFFFFFF,
        bondK=8192, bondBreak=262144, bondReform=0, bondMu=0, bondGroup=8,
        solidMode=SOLID_BOND),
    mat(xsph=2200, gammaCoh=120, adhesion=0, grav=950, cond=26000, coolRate=220,
        boilPt=DEG(-60), boilTo=WATER, spawnTemp=DEG(-180), tint=0x7FD8FF),
    mat(fric=0, eps=32768, dpMax=24576, xsph=400, gammaCoh=900, gammaCur=225,
        adhesion=0, grav=1100, cond=1200, coolRate=60, meltPt=DEG(220), meltTo=TAR,
        spawnTemp=DEG(20), tint=0x33333A, bondK=16384, bondBreak=409600,
        bondReform=0, bondMu=0, bondGroup=7, solidMode=SOLID_BOND),
    mat(fric=16384, xsph=45000, gammaCoh=650, adhesion=520, grav=1450,
        cond=4200, coolRate=150, ignitePt=DEG(230), fuel0=DEG(300), burnRate=8000,
        burnHeat=102000, burnTo=SMOKE, splitTo=ASH, splitPct=26,
        spawnTemp=DEG(20), tint=0x241F1A),
    mat(xsph=700, gammaCoh=950, gammaCur=237, adhesion=0, grav=2600,
        cond=32768, coolRate=70, freezePt=DEG(-39), freezeTo=METAL,
        boilPt=DEG(357), boilTo=SMOKE, spawnTemp=DEG(20), tint=0xC9CCD6),
    mat(xsph=1000, gammaCoh=8, gammaCur=2, adhesion=0, grav=260,
        cond=5000, coolRate=250, ignitePt=DEG(60), fuel0=DEG(18), burnRate=40000,
        burnHeat=700000, burnTo=FIRE, splitTo=SMOKE, splitPct=80,
        spawnTemp=DEG(20), tint=0xB0A58C),
    mat(xsph=4000, gammaCoh=180, gammaCur=45, adhesion=60, grav=1250,
        cond=7000, coolRate=120, boilPt=DEG(130), boilTo=SMOKE,
        corrode=65536, corrPt=20000, corrTo=WATER, spawnTemp=DEG(20), tint=0x8DFF3A),
    mat(xsph=900, gammaCoh=900, gammaCur=225, adhesion=40, grav=1700,
        cond=9000, coolRate=150, ignitePt=DEG(40), fuel0=DEG(4), burnRate=260000,
        burnHeat=1500000, burnTo=BLAST, spawnTemp=DEG(20), tint=0xE8E030),
    mat(xsph=1200, gammaCoh=-2500, gammaCur=-625, adhesion=0, grav=-250,
        cond=24000, coolRate=700, ignitePt=TEMP_MIN, fuel0=DEG(22), burnRate=44000,
        burnHeat=1400000, burnTo=FIRE, spawnTemp=DEG(1800), tint=0xFFF0A0),
    mat(phantom=1, eps=32768, dpMax=1, xsph=0, gammaCoh=0, gammaCur=0, adhesion=0,
        grav=0, cond=0, coolRate=0, fric=0, spawnTemp=DEG(20), tint=0x000000),
    mat(fric=114688, xsph=10000, gammaCoh=80, gammaCur=20, adhesion=0, grav=1500,
        cond=3000, coolRate=90, spawnTemp=DEG(20), tint=0xEEE4C6,
        bondK=16384, bondBreak=147456, bondReform=1, bondFormV=2600, bondFormP=2000,
        bondMu=0, bondGroup=6, rampLo=DEG(45), rampHi=DEG(95), rampSign=1,
        solidMode=SOLID_BOND),
    mat(fric=131072, xsph=40000, gammaCoh=900, adhesion=500, grav=2400,
        cond=0, coolRate=0, spawnTemp=DEG(20), tint=0x4A3A8C,
        bondK=32768, bondBreak=0, bondReform=0, bondMu=0, bondGroup=9,
        burstP=0, rampSign=0, solidMode=SOLID_RIGID),
    mat(fric=53248, xsph=10500, gammaCoh=70, gammaCur=110, adhesion=0, grav=1200,
        cond=0, coolRate=0, spawnTemp=DEG(20), tint=0xFF84E8),
]
assert len(MATS) == len(NAMES)


def mats_image(mats=None) -> np.ndarray:
    """Pack rows into the MAT_MAX x 48 int32 device image; the tail is zero."""
    rows = MATS if mats is None else mats
    img = np.zeros((MAT_MAX, MAT_WORDS), dtype=np.int64)
    for m, row in enumerate(rows):
        r = list(row)
        span = r[F["rampHi"]] - r[F["rampLo"]]
        r[F["rampInv"]] = (2 ** 32) // span if span > 0 else 0
        img[m, :] = r
    # words are 32-bit patterns: wrap the u32-valued fields into int32
    img = ((img + 2 ** 31) % 2 ** 32) - 2 ** 31
    return img.astype(np.int32)


def seal_row(row) -> bool:
    """src/engine.js:708-713"""
    return (row[F["bondK"]] != 0 and row[F["bondReform"]] == 0 and
            row[F["solidMode"]] != SOLID_NONE and row[F["phantom"]] == 0)


def spawn_welds(row) -> bool:
    """src/engine.js:287-290"""
    return row[F["bondGroup"]] >= 0 and row[F["bondReform"]] == 0
