"""Copies the reference's own configuration files into tests/golden/formats/
as fixtures of the format readers: lpmd/examples/ar108-1.tgz (argon108.lpmd,
an LPMD 1.0 file), lpmd/examples/proyectil.lpmd and zirconia-color.lpmd
(LPMD 2.0 text) and the first three frames of lpmd/examples/ZrO2-100conf.zlp
(a legacy ZLP 1 stream: its 10-byte header and the first three size-prefixed
frames, which decode on their own since the deflate stream runs forward).

Run from the repository root with the reference at /root/reference:
    python tests/golden/gen_format_fixtures.py
"""
import os
import struct
import tarfile

REF = "/root/reference/lpmd/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "formats")


def main():
    os.makedirs(OUT, exist_ok=True)
    with tarfile.open(os.path.join(REF, "ar108-1.tgz")) as t:
        data = t.extractfile("argon108.lpmd").read()
    open(os.path.join(OUT, "argon108.lpmd"), "wb").write(data)
    for name in ("proyectil.lpmd", "zirconia-color.lpmd"):
        open(os.path.join(OUT, name), "wb").write(open(os.path.join(REF, name), "rb").read())
    z = open(os.path.join(REF, "ZrO2-100conf.zlp"), "rb").read()
    pos = 10
    for _ in range(3):
        w = z[pos]
        size = struct.unpack(">I", z[pos + 1:pos + 5])[0]
        pos += 1 + w + size
    open(os.path.join(OUT, "ZrO2-3conf.zlp"), "wb").write(z[:pos])


if __name__ == "__main__":
    main()
