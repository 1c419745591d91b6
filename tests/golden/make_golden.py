"""Writes the golden block fixtures from the reference's own codecs, compiled
from the reference tree into oracle/_ref by `make -C oracle`:
  parapfor_ref_blocks.npz  src/gpu/Compressed/include/ParaPFor.cpp
  newpfor_ref_blocks.npz   src/gpu/Compressed/include/NewPFor.cpp (+ Simple16.h)

Inputs are seeded; each record holds the input values, FRAC, the encoded words
and the decoded values the reference produced.  Run in the build container:
    make -C oracle && python tests/golden/make_golden.py [parapfor|newpfor]
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def newpfor():
    ref = oracle.ref_npfor_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_newpfor.so is not built")
    rng = np.random.default_rng(20241105)
    vals, fracs, words, woff, voff, decoded = [], [], [], [0], [0], []
    for k in range(240):
        n = int(rng.choice([1, 2, 3, 27, 28, 29, 64, 100, 127, 128, 129, 200, 255, 256]))
        bits = int(rng.integers(0, 28))
        v = rng.integers(0, 1 << bits, n).astype(np.uint32) if bits else np.zeros(n, np.uint32)
        if k % 2 == 0:  # planted outliers -> Simple16-coded exceptions
            m = int(rng.integers(1, max(2, n // 4)))
            idx = rng.choice(n, min(m, n), replace=False)
            v[idx] = rng.integers(1 << 8, 1 << 27, len(idx)).astype(np.uint32)
        if k % 7 == 3:  # every position an exception
            v[:] = rng.integers(1 << 20, 1 << 26, n).astype(np.uint32)
            v[: n // 2] = 1
        frac = float(rng.choice([0.0, 0.04, 0.1, 0.2, 0.32, 0.64]))
        out = np.zeros(4 * n + 64, np.uint32)
        w = ref.ref_npfor_encode(v.ctypes.data, n, ctypes.c_float(frac), out.ctypes.data)
        pad = np.zeros(w + 64, np.uint32)
        pad[:w] = out[:w]
        dec = np.zeros(n + 64, np.uint32)
        ref.ref_npfor_decode(pad.ctypes.data, n, dec.ctypes.data)
        vals.append(v); fracs.append(frac); words.append(out[:w]); decoded.append(dec[:n])
        woff.append(woff[-1] + w); voff.append(voff[-1] + n)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "newpfor_ref_blocks.npz"),
                        values=np.concatenate(vals), frac=np.array(fracs, np.float32),
                        words=np.concatenate(words), woff=np.array(woff, np.int64),
                        voff=np.array(voff, np.int64), decoded=np.concatenate(decoded))
    print("wrote %d NewPFor blocks" % len(fracs))


def main():
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_parapfor.so is not built")
    rng = np.random.default_rng(20240601)
    vals, fracs, words, woff, voff, decoded = [], [], [], [0], [0], []
    for k in range(160):
        n = int(rng.choice([1, 2, 3, 31, 32, 33, 64, 100, 127, 128, 129, 200, 255, 256]))
        bits = int(rng.integers(0, 31))
        v = rng.integers(0, 1 << bits, n).astype(np.uint32) if bits else np.zeros(n, np.uint32)
        if k % 3 == 0:  # planted outliers -> exceptions
            m = int(rng.integers(1, max(2, n // 6)))
            idx = rng.choice(n, min(m, n), replace=False)
            v[idx] = rng.integers(1 << 20, 1 << 30, len(idx)).astype(np.uint32)
        frac = float(rng.choice([0.0, 0.04, 0.1, 0.2, 0.25]))
        out = np.zeros(4 * n + 8, np.uint32)
        w = ref.ref_pfor_encode(v.ctypes.data, n, ctypes.c_float(frac), out.ctypes.data)
        pad = np.zeros(w + 4, np.uint32)
        pad[:w] = out[:w]
        dec = np.zeros(n, np.uint32)
        ref.ref_pfor_decode(pad.ctypes.data, n, dec.ctypes.data)
        vals.append(v); fracs.append(frac); words.append(out[:w]); decoded.append(dec)
        woff.append(woff[-1] + w); voff.append(voff[-1] + n)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "parapfor_ref_blocks.npz"),
                        values=np.concatenate(vals), frac=np.array(fracs, np.float32),
                        words=np.concatenate(words), woff=np.array(woff, np.int64),
                        voff=np.array(voff, np.int64), decoded=np.concatenate(decoded))
    print("wrote %d blocks" % len(fracs))


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("all", "parapfor"):
        main()
    if which in ("all", "newpfor"):
        newpfor()
