"""Writes tests/golden/setup_digests.json: SHA-256 digests of the oracle's setup
arrays (embedded pressure and u boxes, fluid mask, qu id map) at the production
grid 512x96x96.  The oracle builds them the reference's way (std::unordered_map
rows); the product's structured setup must reproduce them bit for bit.
Run: python tests/golden/make_setup_digests.py   (~90 s, ~20 GB of host memory)
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

KW = dict(nx=512, ny=96, nz=96, Lx=2.5, xcirc=0.5, ycirc=0.2, Um=2.25, dt=2.5e-4)


def main():
    S = O.Setup(O.default_params(**KW))
    dig = {n: hashlib.sha256(S.get(n).tobytes()).hexdigest() for n in ("Ap_vals", "Au_vals", "fluid_p", "id_qu")}
    dig_counts = {k: S.count(k) for k in S.INTS}
    out = {"512x96x96": dig, "512x96x96_counts": dig_counts, "params": KW}
    with open(os.path.join(HERE, "setup_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
