"""Writes the golden fixtures of tests/golden/ from the reference's own C.

Run in a container where the reference sources are mounted and
`make -C oracle ref` has built oracle/_ref/libcvp9_ref.so (the reference's
src/decoder/vp9_reconstruct.c + vp9_frame.c compiled in place). Outputs:
  golden_hashes.json  per-frame, per-plane SHA-256 of the reference's
                      reconstruction of generated sequences (+ generator
                      input hashes, so a drifting generator is caught)
  golden_frames.npz   complete I420 frames of small sequences
  idct_tables.json    the IDCT matrices as hex doubles, from the restated
                      formula (vp9_reconstruct.c:19-28) with this C library's
                      cos; the reference builds them inside vp9_idct_2d, so
                      the golden frames above pin their effect end to end
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from _pkgload import load_package  # noqa: E402
import pyoracle  # noqa: E402

# (name, preset, width, height, frames, seed, overrides)
HASH_CASES = [
    ("cfg0_720p", 0, 1280, 720, 6, 100, {}),
    ("cfg1_1080p", 1, 1920, 1080, 4, 101, {}),
    ("cfg2_2160p", 2, 3840, 2160, 2, 102, {}),
    ("cfg3_480p", 3, 854, 480, 4, 103, {}),
    ("cfg1_360p_skip_sharp", 1, 640, 360, 4, 104, {"skip_permille": 250, "sharpness": 5}),
    ("cfg1_keyint2_lvl0", 1, 320, 184, 4, 105, {"key_interval": 2, "filter_level": 0}),
]
FRAME_CASES = [
    ("tiny64", 1, 64, 64, 3, 200, {}),
    ("edge136x72", 1, 136, 72, 3, 201, {}),
    ("edge138x70", 2, 138, 70, 3, 202, {}),
    ("cfg0_200x120", 0, 200, 120, 3, 203, {}),
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def planes(img, w, h):
    cw, ch = (w + 1) // 2, (h + 1) // 2
    return img[:w * h], img[w * h:w * h + cw * ch], img[w * h + cw * ch:]


def gen(P, preset, w, h, n, seed, over):
    p = P.synth_params(preset, w, h, seed, **over)
    return [P.SynthFrame(p, i) for i in range(n)]


def main():
    P = load_package()
    if not pyoracle.have_ref():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    hashes = {}
    for name, preset, w, h, n, seed, over in HASH_CASES:
        frames = gen(P, preset, w, h, n, seed, over)
        out = pyoracle.run_reference(frames)
        hashes[name] = {
            "preset": preset, "width": w, "height": h, "frames": n, "seed": seed,
            "overrides": over,
            "input": [sha(np.concatenate([f.blocks().view(np.uint8), f.coeffs().view(np.uint8)]))
                      for f in frames],
            "planes": [[sha(pl) for pl in planes(img, w, h)] for img in out],
        }
        print(name, "done", flush=True)
    full = {}
    meta = {}
    for name, preset, w, h, n, seed, over in FRAME_CASES:
        frames = gen(P, preset, w, h, n, seed, over)
        out = pyoracle.run_reference(frames)
        full[name] = np.stack(out)
        meta[name] = {"preset": preset, "width": w, "height": h, "frames": n, "seed": seed,
                      "overrides": over}
    tables = {}
    for n in (4, 8, 16, 32):
        m = pyoracle.idct_matrix(n)
        tables[str(n)] = [float(x).hex() for x in m.ravel()]
    with open(os.path.join(HERE, "golden_hashes.json"), "w") as f:
        json.dump({"hash_cases": hashes, "frame_cases": meta}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "golden_frames.npz"), **full)
    with open(os.path.join(HERE, "idct_tables.json"), "w") as f:
        json.dump(tables, f)


if __name__ == "__main__":
    main()
