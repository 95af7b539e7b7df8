"""Writes tests/golden/a2_known_answers.json: per-halo quantities of the
pks2map halo loop (pks2map.f90:284-298, haloproject.f90:47) for three halos,
computed here from the physics the reference encodes, independently of the
oracle's code:

  * chi = |x| after the observer shift along the painter x axis; the redshift
    is taken at |x| before the shift (pks2map.f90:289-293);
  * z(chi) by inverting the flat-LCDM comoving distance
    chi(z) = (c/H0) int_0^z dz'/E(z'), E^2 = Om (1+z)^3 + OL, with scipy's
    adaptive quadrature and a root finder (the reference tabulates it with
    Euler steps, so the two agree to the Euler error: |dz| < 5e-5);
  * M = 4/3 pi rho0 Rth^3, rho0 = 2.775e11 Om h^2 Msun/Mpc^3;
  * mh = sqrt(Dc(z) / 200) M with the Bryan & Norman (1998) overdensity
    Dc = 18 pi^2 + 82 x - 39 x^2, x = Om(z) - 1, the 18 pi^2 term with the
    reference's float-rounded pi (cosmology.f90:8-9);
  * rh = (3 mh / (4 pi 200 rho_c(z)))^(1/3), rho_c(z) the comoving critical
    density rhoc0 (Om (1+z)^3 + OL) / (1+z)^3, rhoc0 = 2.775e11 h^2;
  * dtheta = 4 asin(rh / chi).

The halo records are float32, so M uses the float32 Rth.  Cosmology is the
hard-coded one of pks2map.f90:118-128 (Om 0.25, OL 0.75, h 0.7).
"""
import json
import math
import os

import numpy as np
from scipy import integrate, optimize

HERE = os.path.dirname(os.path.abspath(__file__))
OM, OL, H = 0.25, 0.75, 0.7
C_KMS = 2.99792458e10 / 1e5
PI_F = float(np.float32(3.14159265359))


def chi_of_z(z):
    f = lambda zz: 1.0 / math.sqrt(OM * (1 + zz) ** 3 + OL)
    return C_KMS / (100.0 * H) * integrate.quad(f, 0.0, z, epsabs=0, epsrel=1e-13)[0]


def z_of_chi(chi):
    return optimize.brentq(lambda z: chi_of_z(z) - chi, 0.0, 6.0, xtol=1e-14)


def halo(pos, rth, chihview):
    """pos = (pos1, pos2, pos3) of the record; painter axes x=pos3, y=pos2, z=pos1"""
    p = [float(np.float32(v)) for v in pos]
    rth = float(np.float32(rth))
    x, y, zc = p[2], p[1], p[0]
    chi_pre = math.sqrt(x * x + y * y + zc * zc)
    z = z_of_chi(chi_pre)
    xs = x - chihview
    chi = math.sqrt(xs * xs + y * y + zc * zc)
    rho0 = 2.775e11 * OM * H * H
    m = 4.0 / 3.0 * math.pi * rho0 * rth ** 3
    a3 = (1 + z) ** 3
    omz = OM * a3 / (OM * a3 + OL)
    xx = omz - 1
    dc = 18 * PI_F ** 2 + 82 * xx - 39 * xx ** 2
    mh = math.sqrt(dc / 200.0) * m
    rhoc = 2.775e11 * H * H * (OM * a3 + OL) / a3
    rh = (3 * mh / (4 * math.pi * 200.0 * rhoc)) ** (1.0 / 3.0)
    dth = 4 * math.asin(rh / chi)
    return {"record": [p[0], p[1], p[2], 0.0, 0.0, 0.0, rth], "chihview": chihview,
            "z": z, "M": m, "mh": mh, "rh": rh, "chi": chi, "dtheta": dth}


def rth_for_mass(m):
    rho0 = 2.775e11 * OM * H * H
    return (3 * m / (4 * math.pi * rho0)) ** (1.0 / 3.0)


def main():
    halos = [
        halo((900.0, -600.0, 1000.0), rth_for_mass(1e14), 0.0),       # chi ~ 1473 Mpc, z ~ 0.55
        halo((-150.0, 300.0, 200.0), rth_for_mass(1e15), 0.0),        # nearby and massive: a wide disc
        halo((2500.0, 400.0, -1800.0), rth_for_mass(3e13), 250.0),    # observer shifted by 250 Mpc
    ]
    with open(os.path.join(HERE, "a2_known_answers.json"), "w") as f:
        json.dump({"cosmology": {"omegam": OM, "omegal": OL, "h": H}, "halos": halos}, f, indent=1)


if __name__ == "__main__":
    main()
