"""Writes FITS primary images (test input for the IMAGE source)."""
import numpy as np

DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", -32: ">f4", -64: ">f8"}


def write_image(path, data, bitpix=-32, cdelt=None, cd=None):
    """data (ny, nx); cdelt (c1, c2) degrees or cd (cd11, cd12, cd21, cd22)."""
    ny, nx = data.shape
    cards = [("SIMPLE", "T"), ("BITPIX", str(bitpix)), ("NAXIS", "2"), ("NAXIS1", str(nx)), ("NAXIS2", str(ny))]
    if cdelt is not None:
        cards += [("CDELT1", repr(float(cdelt[0]))), ("CDELT2", repr(float(cdelt[1])))]
    if cd is not None:
        cards += [(f"CD{k}", repr(float(v))) for k, v in zip(("1_1", "1_2", "2_1", "2_2"), cd)]
    hdr = "".join(f"{k:<8}= {v:>20}" .ljust(80) for k, v in cards) + "END".ljust(80)
    hdr = hdr.ljust((len(hdr) + 2879) // 2880 * 2880)
    body = np.ascontiguousarray(data, dtype=DTYPES[bitpix]).tobytes()
    body += b"\0" * ((-len(body)) % 2880)
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii") + body)
    return path
