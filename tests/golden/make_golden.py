"""Writes the hand-built byte fixtures of tests/golden/.

three_halos.pksc: the .pksc layout of the catalogue writer
(merge_pkvd.f90:1311-1325: int32 N and float32 RTHmax at pos 1, then
7 float32 per halo (x,y,z,vx,vy,vz,Rth) at pos 9 + 28*i), written byte by
byte with struct so it does not depend on the package's own writer.
tiny_table.bin: the table layout of maptable.f90:94,142 (36-byte header,
then table(2,nrt,nmht,nchit) column-major) for nchit=3, nmht=3, nrt=4 with
element value 1000*p + 100*ir + 10*imh + ichi.
fits_header_nside{8,64}.bin: the two header blocks (primary HDU, then the
BINTABLE extension) of a HEALPix map file as fits_s_inc.f90:787-960 lays it
out (repeat 1 for npix < 1024, else 1024; NAXIS2 = npix / repeat; NSIDE and
ORDERING cards), built card by card from the FITS standard's fixed format
(FITS 4.0 sec. 4.1-4.2): keyword in columns 1-8, "= " in 9-10, integers and
logicals right-justified to column 30, strings quoted from column 11 and
padded to 8 characters, " / " and the comment after the value, the END
card, blanks to a multiple of 2880 bytes.  The DATE value is the
placeholder 'YYYY-MM-DD'.
"""
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

HALOS = [
    (1200.5, -33.25, 450.0, 120.0, -75.5, 10.0, 3.5),
    (-2500.0, 1800.75, -90.125, -300.0, 260.0, 5.25, 7.0),
    (0.0, 0.0, 3000.0, 0.0, 0.0, 0.0, 12.5),
]


def fits_card(key, value, comment, kind):
    c = key.ljust(8) + "= "
    if kind == "str":
        c += ("'" + value.ljust(8) + "'").ljust(20)
    else:
        c += value.rjust(20)
    return (c + " / " + comment).ljust(80)[:80]


def fits_header(nside):
    npix = 12 * nside * nside
    repeat = 1 if npix < 1024 else 1024
    prim = [fits_card("SIMPLE", "T", "file does conform to FITS standard", "log"),
            fits_card("BITPIX", "32", "number of bits per data pixel", "int"),
            fits_card("NAXIS", "0", "number of data axes", "int"),
            fits_card("EXTEND", "T", "FITS dataset may contain extensions", "log"),
            fits_card("DATE", "YYYY-MM-DD", "file creation date (YYYY-MM-DD)", "str")]
    ext = [fits_card("XTENSION", "BINTABLE", "binary table extension", "str"),
           fits_card("BITPIX", "8", "8-bit bytes", "int"),
           fits_card("NAXIS", "2", "2-dimensional binary table", "int"),
           fits_card("NAXIS1", str(4 * repeat), "width of table in bytes", "int"),
           fits_card("NAXIS2", str(npix // repeat), "number of rows in table", "int"),
           fits_card("PCOUNT", "0", "size of special data area", "int"),
           fits_card("GCOUNT", "1", "one data group (required keyword)", "int"),
           fits_card("TFIELDS", "1", "number of fields in each row", "int"),
           fits_card("TTYPE1", "simulation", "label for field   1", "str"),
           fits_card("TFORM1", f"{repeat}E", "data format of field: 4-byte REAL", "str"),
           fits_card("NSIDE", str(nside), "the nside of the map", "int"),
           fits_card("ORDERING", "RING", "the nside of the map", "str")]
    out = b""
    for cards in (prim, ext):
        blk = "".join(cards) + "END".ljust(80)
        blk += " " * ((-len(blk)) % 2880)
        out += blk.encode("ascii")
    return out


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
    for ns in (8, 64):
        with open(os.path.join(HERE, f"fits_header_nside{ns}.bin"), "wb") as f:
            f.write(fits_header(ns))


if __name__ == "__main__":
    main()
