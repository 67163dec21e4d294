"""Writes sector_digests.json: SHA-256 digests of the oracle's basis arrays at full size.

Run from the repository root in the build container (the oracle restates the
reference's sector construction, basis.jl:622-777 and symmetry_basis.jl:38-255):

    python tests/golden/make_digests.py

For each sector it records the dimension and the digests of three arrays in
column order: the reference labels states(basis) (uint64), the orbit sizes
(uint32) and normalization(basis, states(basis)[i]) (float64), all
little-endian.  The GPU test compares libqsgpu's arrays with these digests.
Sectors: the J1-J2 ground-state sector L=28 nup=14 k=0 p=+1 z=+1 and the 26
momentum / parity blocks of the L=24 nup=12 batched quench.
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle  # noqa: E402


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.dtype(dtype).newbyteorder("<")).tobytes()).hexdigest()


def sector_record(L, nup, blocks):
    t = time.time()
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    states = sec.states()
    orbit, ref = sec.norms()
    rec = dict(L=L, nup=nup, blocks=blocks, Ns=int(sec.Ns), states=digest(states, np.uint64),
               orbit_size=digest(orbit, np.uint32), ref_norm=digest(ref, np.float64),
               orbit_sum=int(orbit.astype(np.int64).sum()), first_states=[int(s) for s in states[:4]])
    print("%s %s Ns=%d %.1fs" % (L, blocks, sec.Ns, time.time() - t), flush=True)
    return rec


def main():
    out = []
    out.append(sector_record(28, 14, dict(kblock=0, pblock=1, zblock=1)))
    for k in range(24):
        if k in (0, 12):
            out.append(sector_record(24, 12, dict(kblock=k, pblock=1)))
            out.append(sector_record(24, 12, dict(kblock=k, pblock=-1)))
        else:
            out.append(sector_record(24, 12, dict(kblock=k)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sector_digests.json")
    with open(path, "w") as f:
        json.dump({"_source": "tests/golden/make_digests.py (the oracle's sector construction)",
                   "sectors": out}, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
