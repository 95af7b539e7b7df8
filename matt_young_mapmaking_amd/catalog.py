"""Deterministic synthetic halo catalogues in the .pksc layout.

The reference ships no catalogue (the .pksc symlink points at a missing
/scratch2 path), so the benchmark workloads are drawn here:

  * direction: isotropic (mu ~ U(-1,1), phi ~ U(0,2pi)), or for the flat-sky
    workload uniform in a square patch around the painter +x axis;
  * comoving distance: uniform in comoving volume for zlo < z < zhi in a flat
    LCDM with (omegam, h) = (0.31, 0.68);
  * mass: dn/dM ~ M^-1.9 on [mlo, mhi], converted to the Lagrangian radius
    Rth = (3M / (4 pi rhobar))^(1/3), rhobar = 2.775e11 omegam h^2;
  * velocities: three Gaussians of sigma = 300 km/s.

Records are (pos1, pos2, pos3, vx, vy, vz, Rth) with pks2map's painter axes
x_p = pos3, y_p = pos2, z_p = pos1 (pks2map.f90:284-286).  Draws come in
fixed blocks of 2^20 halos, each from numpy's PCG64 seeded with
(seed, block index), so any prefix or chunk is reproducible on its own.
"""
import numpy as np

BLOCK = 1 << 20
C_KMS = 299792.458


def comoving_distance(z, omegam=0.31, h=0.68, n=20001):
    """chi(z) in Mpc for flat LCDM by Simpson's rule on a fine grid."""
    z = np.asarray(z, dtype=np.float64)
    zz = np.linspace(0.0, float(np.max(z)), n)
    e = np.sqrt(omegam * (1 + zz) ** 3 + 1 - omegam)
    f = 1.0 / e
    # cumulative Simpson on pairs of intervals, then linear in between
    dz = zz[1] - zz[0]
    cum = np.zeros_like(zz)
    cum[1:] = np.cumsum(0.5 * (f[1:] + f[:-1]) * dz)
    return (C_KMS / (100.0 * h)) * np.interp(z, zz, cum)


CONFIGS = {
    # name: n halos, geometry, mass range (Msun), nside, flat npix
    "C1": dict(n=20_000, geom="patch", mlo=1e13, mhi=1e15, nside=4096, npix=1024),
    "C2": dict(n=2_000_000, geom="sky", mlo=1e13, mhi=1e15, nside=2048, npix=1024),
    "C3": dict(n=20_000_000, geom="sky", mlo=1e12, mhi=10 ** 15.5, nside=4096, npix=1024),
    "C4": dict(n=100_000_000, geom="sky", mlo=1e12, mhi=10 ** 15.5, nside=4096, npix=1024),
    "C5": dict(n=100_000_000, geom="sky", mlo=1e12, mhi=10 ** 15.5, nside=8192, npix=1024),
}


def _block(seed, b, nb, geom, chi_lo3, chi_hi3, mlo, mhi, rhobar, patch_deg):
    rng = np.random.default_rng([seed, b])
    u_chi = rng.random(nb)
    a = rng.random(nb)
    c = rng.random(nb)
    u_m = rng.random(nb)
    vel = rng.normal(0.0, 300.0, size=(nb, 3))
    chi = np.cbrt(chi_lo3 + u_chi * (chi_hi3 - chi_lo3))
    if geom == "sky":
        mu = 2.0 * a - 1.0
        phi = 2.0 * np.pi * c
        s = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
        xp, yp, zp = chi * s * np.cos(phi), chi * s * np.sin(phi), chi * mu
    else:
        half = np.deg2rad(patch_deg / 2.0)
        st = np.sin((2.0 * a - 1.0) * half)
        sp = np.sin((2.0 * c - 1.0) * half)
        # flat-sky painter: theta = asin(x/chi), phi = asin(y/chi), needs z > 0
        xp, yp = chi * st, chi * sp
        zp = chi * np.sqrt(1.0 - st * st - sp * sp)
    g = -0.9
    m = (mlo ** g + u_m * (mhi ** g - mlo ** g)) ** (1.0 / g)
    rth = np.cbrt(3.0 * m / (4.0 * np.pi * rhobar))
    out = np.empty((nb, 7), np.float32)
    out[:, 0] = zp      # pos(1) = z_p
    out[:, 1] = yp      # pos(2) = y_p
    out[:, 2] = xp      # pos(3) = x_p
    out[:, 3:6] = vel
    out[:, 6] = rth
    return out


def make_catalog(n, geom="sky", mlo=1e13, mhi=1e15, zlo=0.05, zhi=1.2, seed=13579,
                 omegam=0.31, h=0.68, patch_deg=10.0, start=0):
    """Halos [start, start+n) of the catalogue, as an (n, 7) float32 array."""
    chi_lo, chi_hi = comoving_distance([zlo, zhi], omegam, h)
    rhobar = 2.775e11 * omegam * h * h
    out = np.empty((n, 7), np.float32)
    b0, b1 = start // BLOCK, (start + n + BLOCK - 1) // BLOCK
    for b in range(b0, b1):
        lo = max(start, b * BLOCK)
        hi = min(start + n, (b + 1) * BLOCK)
        if hi <= lo:
            continue
        blk = _block(seed, b, BLOCK, geom, chi_lo ** 3, chi_hi ** 3, mlo, mhi, rhobar, patch_deg)
        out[lo - start:hi - start] = blk[lo - b * BLOCK:hi - b * BLOCK]
    return out


def config_catalog(name, n=None, start=0, seed=13579):
    c = CONFIGS[name]
    return make_catalog(c["n"] if n is None else n, geom=c["geom"], mlo=c["mlo"], mhi=c["mhi"],
                        seed=seed, start=start)
