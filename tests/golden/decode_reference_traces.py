"""Decode the reference's golden receiver traces into small .npz fixtures.

The reference keeps the outputs of its composite-grid test flow
(tests/test1-cmp-eci-ii, run on TACC Stampede, Oct 28 2016, Madagascar
1.8-svn) only in its SVN pristine store, as RSF files: a text header, the
bytes 0x0c 0x0c 0x04, then little-endian float32 samples.  This script
reads those files by sha1 and writes the samples plus the header values
this build uses (n1, n2, d1) to reference_traces.npz.  Data only: no
reference source is copied.

Run once in the development container:  python tests/golden/decode_reference_traces.py
"""
import os
import re

import numpy as np

PRISTINE = "/root/reference/.svn/pristine"
FILES = {
    # name: (sha1, what)
    "rec_eci_test1": ("4d96b646722443d963186f3de56b43992ac6a318", "ECI vx traces, par1.py"),
    "rec_eci_test2": ("74f046ed5e5ac3ffeeca746cc1bfae95aac1fb81", "ECI vx traces, par2.py (n_extra=1)"),
    "rec_eci_test3": ("4489fb15eda43832b79b2f808c7095431435e366", "ECI vx traces, par3.py (source near interface)"),
    "rec_ii_test1": ("2f40ff641bce684f913a3b7f6dd66bddb9e57c56", "II (imode=0) vx traces, par1.py"),
    "rec_ii_test2": ("ff2674791225c7dbd7db7965451b729eed7f0df4", "II vx traces, par2.py"),
    "rec_ii_test3": ("c3c5436eb4af3d50ec33a757cce8d408bf07bb85", "II vx traces, par3.py"),
    "geo_eci_test1": ("44dcc8881eb38391c90de36fe55b4ceaac8c9f12", "receiver xyz (m), 3 x 4"),
}


def read_rsf(sha: str):
    path = os.path.join(PRISTINE, sha[:2], sha + ".svn-base")
    raw = open(path, "rb").read()
    cut = raw.index(b"\x0c\x0c\x04")
    header = raw[:cut].decode("latin1")
    kv = {}
    for m in re.finditer(r'(\w+)=("[^"]*"|\S+)', header):
        kv[m.group(1)] = m.group(2)
    data = np.frombuffer(raw[cut + 3:], dtype="<f4").copy()
    return kv, data


def main():
    out = {}
    for name, (sha, _) in FILES.items():
        kv, data = read_rsf(sha)
        n1, n2 = int(kv["n1"]), int(kv["n2"])
        out[name] = data.reshape(n2, n1)
        if "d1" in kv:
            out[name + "_d1"] = np.float64(kv["d1"])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_traces.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
