"""Writes the hand-built byte fixtures of tests/golden/.

three_halos.pksc: the .pksc layout of the catalogue writer
(merge_pkvd.f90:1311-1325: int32 N and float32 RTHmax at pos 1, then
7 float32 per halo (x,y,z,vx,vy,vz,Rth) at pos 9 + 28*i), written byte by
byte with struct so it does not depend on the package's own writer.
tiny_table.bin: the table layout of maptable.f90:94,142 (36-byte header,
then table(2,nrt,nmht,nchit) column-major) for nchit=3, nmht=3, nrt=4 with
element value 1000*p + 100*ir + 10*imh + ichi.
"""
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

HALOS = [
    (1200.5, -33.25, 450.0, 120.0, -75.5, 10.0, 3.5),
    (-2500.0, 1800.75, -90.125, -300.0, 260.0, 5.25, 7.0),
    (0.0, 0.0, 3000.0, 0.0, 0.0, 0.0, 12.5),
]


def main():
    with open(os.path.join(HERE, "three_halos.pksc"), "wb") as f:
        f.write(struct.pack("<if", len(HALOS), max(h[6] for h in HALOS)))
        for h in HALOS:
            f.write(struct.pack("<7f", *h))
    nchit, nmht, nrt = 3, 3, 4
    with open(os.path.join(HERE, "tiny_table.bin"), "wb") as f:
        f.write(struct.pack("<3i6f", nchit, nmht, nrt, 1e2, 7e3, 1e13, 1e16, 1e-2, 10.0))
        for ichi in range(1, nchit + 1):
            for imh in range(1, nmht + 1):
                for ir in range(1, nrt + 1):
                    for p in (1, 2):
                        f.write(struct.pack("<f", 1000 * p + 100 * ir + 10 * imh + ichi))


if __name__ == "__main__":
    main()
