"""Synthetic particle beds for the DEM tests: random sequential placement in a
periodic box with a small allowed overlap, so that many pairs start in
contact, and random velocities.  Grid units as the fluid solver holds them."""
import numpy as np


def random_bed(n, box, dx=1.0 / 16, rad_range=(8.0, 8.0), overlap=0.01, vel=0.5, seed=0):
    """n particles, box (physical, cubic) -> (xc (n,3) grid units, radbdy (n,),
    velbdy (n,3)).  Pairs may overlap by up to ``overlap`` of r_a + r_b."""
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rad_range[0], rad_range[1], n)
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n:
            raise RuntimeError("bed too dense for random placement")
        p = rng.uniform(0.0, box, 3)
        if k:
            d = pos[:k] - p
            d -= box * np.round(d / box)
            dist = np.sqrt((d * d).sum(1))
            if np.any(dist < (1.0 - overlap) * (rad[:k] + rad[k]) * dx):
                continue
        pos[k] = p
        k += 1
    xc = pos / dx + 1.0
    v = rng.uniform(-vel, vel, (n, 3))
    return np.asfortranarray(xc), rad, np.asfortranarray(v)
