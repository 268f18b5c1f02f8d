"""A minimal FITS binary-table reader for the tests, written from the FITS
standard (2880-byte blocks of 80-character cards, big-endian columns of
TFORM types A, I, J, E, D): every binary-table extension of a file with its
header keywords, so that tables sharing an EXTNAME (the CALDB geometry file
has one INSTGEOM per instrument) stay apart."""
import numpy as np

BLOCK = 2880
_CODES = {"A": "S", "I": ">i2", "J": ">i4", "K": ">i8", "E": ">f4", "D": ">f8", "B": "u1"}


def _header(buf, pos):
    h = {}
    while True:
        blk = buf[pos:pos + BLOCK]
        pos += BLOCK
        for j in range(0, BLOCK, 80):
            card = blk[j:j + 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                return h, pos
            if card[8:10] != "= ":
                continue
            v = card[10:].strip()
            if v.startswith("'"):
                h[key] = v[1:].split("'")[0].rstrip()
            else:
                h[key] = v.split("/")[0].strip()


def read_tables(path):
    """[(header dict, {column (upper case): array})] of every BINTABLE."""
    buf = open(path, "rb").read()
    h, pos = _header(buf, 0)
    out = []
    while pos < len(buf):
        h, pos = _header(buf, pos)
        size = int(h.get("NAXIS1", 0)) * int(h.get("NAXIS2", 0))
        if h.get("XTENSION") == "BINTABLE":
            fields = []
            for k in range(1, int(h["TFIELDS"]) + 1):
                form = h[f"TFORM{k}"].strip()
                rep = int(form[:-1]) if len(form) > 1 else 1
                code = _CODES[form[-1]]
                if code == "S":
                    fields.append((h[f"TTYPE{k}"].upper(), f"S{rep}", ()))
                else:
                    fields.append((h[f"TTYPE{k}"].upper(), code, (rep,) if rep > 1 else ()))
            dt = np.dtype([(n, c, s) for n, c, s in fields])
            assert dt.itemsize == int(h["NAXIS1"]), path
            rec = np.frombuffer(buf[pos:pos + size], dtype=dt)
            cols = {}
            for n, c, _ in fields:
                col = rec[n]
                cols[n] = (np.char.strip(col.astype(str)) if c.startswith("S")
                           else col.astype(np.dtype(c).newbyteorder("=")))
            out.append((h, cols))
        pos += (size + BLOCK - 1) // BLOCK * BLOCK
    return out
