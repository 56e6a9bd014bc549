"""Write dcd_1base_head.json: the header fields and the first frame of the
reference's shipped trajectory cases/Diffusion Coefficient/1base.dcd
(4 beads, 24^3 mesh, dump every 4000 steps, timestep 0.005).  The fixture
is data only: control words, record lengths, title text and the first frame's
bytes in hex.  Run here, where the reference is mounted:

    python tests/golden/make_dcd_fixture.py
"""
import json
import os
import struct

SRC = "/root/reference/data/cases/Diffusion Coefficient/1base.dcd"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dcd_1base_head.json")


def main():
    b = open(SRC, "rb").read()
    rec1 = struct.unpack_from("<i", b, 0)[0]
    magic = b[4:8].decode()
    icntrl = list(struct.unpack_from("<20i", b, 8))
    delta = struct.unpack_from("<f", b, 8 + 36)[0]
    off = 8 + 80 + 4
    rec2 = struct.unpack_from("<i", b, off)[0]
    ntitle = struct.unpack_from("<i", b, off + 4)[0]
    titles = [b[off + 8 + 80 * t: off + 8 + 80 * (t + 1)].rstrip(b"\0 ").decode() for t in range(ntitle)]
    off += 4 + rec2 + 4
    rec3, natoms, rec3e = struct.unpack_from("<3i", b, off)
    off += 12
    frame_len = (4 + 48 + 4) + 3 * (4 + 4 * natoms + 4)
    frame0 = b[off:off + frame_len]
    out = {"source": "cases/Diffusion Coefficient/1base.dcd", "file_bytes": len(b),
           "record_lengths": [rec1, rec2, rec3], "magic": magic, "icntrl": icntrl, "delta": delta,
           "ntitle": ntitle, "title0": titles[0], "natoms": natoms, "header_bytes": off,
           "frame_bytes": frame_len, "frame0_hex": frame0.hex()}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
