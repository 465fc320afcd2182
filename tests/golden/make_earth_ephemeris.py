"""Writes the substitute Earth ephemeris used for annual parallax in the tests
and the benchmark (the reference's own table, data/earth_orbital_parallax_table.txt,
is not distributed with it).

Format: the Horizons vector-table CSV the reference parses
(src/lcbinint/model/trajectory.cpp:110-183): rows between $$SOE and $$EOE,
"JD, calendar date, X, Y, Z, VX, VY, VZ," with the Earth's heliocentric
position (AU) and velocity (AU/day) in the J2000 equatorial frame.

The Earth follows a fixed Keplerian ellipse (J2000 mean elements: a =
1.00000261 AU, e = 0.01671123, longitude of perihelion 102.93768193 deg, mean
longitude 100.46457166 deg at JD 2451545.0, mean motion 0.9856091 deg/day),
rotated from the ecliptic to the equator by the obliquity 23.43928 deg.  This
is a low-precision ephemeris (errors of order 1e-2 AU against the real Earth);
parity only needs the reference and the product to read the same table.

    python tests/golden/make_earth_ephemeris.py [out] [jd_first] [jd_last]
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

A = 1.00000261
E = 0.01671123
VARPI = math.radians(102.93768193)
L0 = math.radians(100.46457166)
N = math.radians(0.9856091)  # rad/day
EPS = math.radians(23.43928)
JD2000 = 2451545.0


def earth_state(jd: float):
    mean_anomaly = L0 + N * (jd - JD2000) - VARPI
    ecc_anomaly = mean_anomaly
    for _ in range(30):
        ecc_anomaly -= (ecc_anomaly - E * math.sin(ecc_anomaly) - mean_anomaly) / (1.0 - E * math.cos(ecc_anomaly))
    cos_e, sin_e = math.cos(ecc_anomaly), math.sin(ecc_anomaly)
    b = A * math.sqrt(1.0 - E * E)
    # perifocal position and velocity (dE/dt = n / (1 - e cos E))
    xp, yp = A * (cos_e - E), b * sin_e
    de_dt = N / (1.0 - E * cos_e)
    vxp, vyp = -A * sin_e * de_dt, b * cos_e * de_dt
    cw, sw = math.cos(VARPI), math.sin(VARPI)
    x, y = cw * xp - sw * yp, sw * xp + cw * yp
    vx, vy = cw * vxp - sw * vyp, sw * vxp + cw * vyp
    ce, se = math.cos(EPS), math.sin(EPS)
    return (x, y * ce, y * se), (vx, vy * ce, vy * se)


def write(path: Path, jd_first: float = 2459980.5, jd_last: float = 2460240.5) -> None:
    lines = ["Substitute Earth ephemeris (Keplerian J2000 mean elements), heliocentric,",
             "J2000 equatorial frame, AU and AU/day.", "$$SOE"]
    jd = jd_first
    while jd <= jd_last + 1e-9:
        (x, y, z), (vx, vy, vz) = earth_state(jd)
        lines.append(f"{jd:.9f}, A.D. JD{jd:.1f}, {x:.15E}, {y:.15E}, {z:.15E}, "
                     f"{vx:.15E}, {vy:.15E}, {vz:.15E},")
        jd += 1.0
    lines.append("$$EOE")
    path.write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    out = Path(sys.argv[1]) if len(sys.argv) > 1 else Path(__file__).with_name("earth_ephemeris_substitute.txt")
    first = float(sys.argv[2]) if len(sys.argv) > 2 else 2459980.5
    last = float(sys.argv[3]) if len(sys.argv) > 3 else 2460240.5
    write(out, first, last)
    print(f"wrote {out}")
