"""Writes the golden fixtures in this directory.

hash_vectors.json  -- identity and murmur3 hashes of seeded keys of every
                      length 0..64, computed by the C oracle and required to
                      equal the independent Python restatement.
config1_digest.json -- config-1 (u64->u64, 2^20 slots, seed 1) results:
                      counts and SHA-256 digests of both get passes and of
                      the final flag array, computed by the C oracle and
                      required to equal the Python restatement.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "flash-map_amd")]

from flashmap_amd.workloads import Stream, config1  # noqa: E402
from oracle import pyoracle  # noqa: E402
from oracle.binding import OracleMap, hash_bytes  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def hash_vectors():
    st = Stream(42)
    rows = []
    for n in range(65):
        for rep in range(3):
            key = bytes(st.bytes(1, 8 * ((n + 7) // 8) or 8)[0][:n])
            ident, mur = hash_bytes(key, 0), hash_bytes(key, 1)
            assert ident == pyoracle.identity_hash(key)
            assert mur == pyoracle.murmur3_hash(key)
            rows.append({"key": key.hex(), "identity": f"{ident:016x}", "murmur3": f"{mur:016x}"})
    for key in [bytes(8), bytes(3), b"\xff" * 32, bytes([10] * 32)]:
        rows.append({"key": key.hex(), "identity": f"{hash_bytes(key, 0):016x}",
                     "murmur3": f"{pyoracle.murmur3_hash(key):016x}"})
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump({"source": "src/hash.rs:30-86 restated (C oracle == Python restatement)",
                   "vectors": rows}, f, indent=0)


def config1_digest():
    c = config1()
    o = OracleMap(c.capacity, 8, 8, 0)
    ins = o.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8))
    v1, f1 = o.bulk_get(c.queries.view(np.uint8))
    rem = o.bulk_remove(c.remove_keys.view(np.uint8))
    v2, f2 = o.bulk_get(c.queries.view(np.uint8))
    flags = o.export()[0]

    p = pyoracle.PyMap(c.capacity, 8, 8, 0)
    kb = [int(k).to_bytes(8, "little") for k in c.insert_keys]
    vb = [int(v).to_bytes(8, "little") for v in c.insert_vals]
    assert p.insert(list(zip(kb, vb))) == ins
    qb = [int(q).to_bytes(8, "little") for q in c.queries]

    def pyget():
        res = p.get(qb)
        vals = b"".join(r if r is not None else bytes(8) for r in res)
        found = bytes(1 if r is not None else 0 for r in res)
        return vals, found

    pv1, pf1 = pyget()
    assert pv1 == v1.tobytes() and pf1 == f1.tobytes()
    assert p.remove([int(k).to_bytes(8, "little") for k in c.remove_keys]) == rem
    pv2, pf2 = pyget()
    assert pv2 == v2.tobytes() and pf2 == f2.tobytes()
    assert np.array_equal(np.array(p.export()[0], dtype=np.uint32), flags)

    out = {
        "config": "FlashMap<u64,u64> with_capacity(2^20); SplitMix64 seed 1; 500k insert, 500k get, "
                  "100k remove, get again",
        "inserted": ins, "hits1": int(f1.sum()), "removed": rem, "hits2": int(f2.sum()),
        "get1_sha256": hashlib.sha256(v1.tobytes() + f1.tobytes()).hexdigest(),
        "get2_sha256": hashlib.sha256(v2.tobytes() + f2.tobytes()).hexdigest(),
        "flags_sha256": hashlib.sha256(flags.tobytes()).hexdigest(),
    }
    with open(os.path.join(HERE, "config1_digest.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    hash_vectors()
    config1_digest()
    print("golden fixtures written")
