"""Pins the CPU oracle (oracle/fm_oracle.c) before anything is compared with it.

Checks, each chosen so that a plausible restatement mistake fails one:
  * the reference's own known-answer tests (src/hash.rs:92-117);
  * MurmurHash3's fmix64 against the published SMHasher verification value
    of MurmurHash3_x64_128 (0x6384BA69), then the reference's murmur3 for
    keys of 0 to 47 bytes (chunk chaining, the tail branch, the length fold)
    assembled from that verified fmix64, with planted restatement mistakes
    that the expected values must each reject;
  * C vs an independent pure-Python restatement (oracle/pyoracle.py) on
    hashes, random collision-heavy histories and exhaustive short histories;
  * closed-form invariants (displacement, len == #OCCUPIED, dict model on
    tombstone-free duplicate-free histories, order-independent occupied set);
  * the 17 reference tests (tests/correctness.rs) and the anomaly histories;
  * committed golden fixtures (tests/golden/).
"""
import itertools
import json
import os

import numpy as np
import pytest

from helpers import ANOMALY_CASES, REFERENCE_TESTS, OracleAdapter, run_history
from oracle import pyoracle
from oracle.binding import OracleMap, RayonPortMap, hash_bytes, hash_many

M64 = (1 << 64) - 1
GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- hashing --
def test_identity_known_answers():
    # src/hash.rs:92-104
    assert hash_bytes(bytes([1, 2, 3, 4, 5, 6, 7, 8, 99, 99]), 0) == int.from_bytes(bytes(range(1, 9)), "little")
    assert hash_bytes(bytes([0xAB, 0xCD]), 0) == 0xCDAB
    assert hash_bytes(b"", 0) == 0


def test_murmur3_reference_unit_tests():
    # src/hash.rs:106-117
    assert hash_bytes(bytes([1, 2, 3, 4]), 1) != hash_bytes(bytes([5, 6, 7, 8]), 1)
    assert hash_bytes(bytes([10] * 32), 1) == hash_bytes(bytes([10] * 32), 1)


def _fmix64(k):
    k ^= k >> 33
    k = (k * 0xFF51AFD7ED558CCD) & M64
    k ^= k >> 33
    k = (k * 0xC4CEB9FE1A85EC53) & M64
    k ^= k >> 33
    return k


def _rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def _murmur3_x64_128(data: bytes, seed: int) -> bytes:
    """Appleby's public-domain MurmurHash3_x64_128."""
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = seed & 0xFFFFFFFF
    n = len(data)
    nb = n // 16
    for i in range(nb):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        k1 = (_rotl((k1 * c1) & M64, 31) * c2) & M64
        h1 ^= k1
        h1 = (((_rotl(h1, 27) + h2) & M64) * 5 + 0x52DCE729) & M64
        k2 = (_rotl((k2 * c2) & M64, 33) * c1) & M64
        h2 ^= k2
        h2 = (((_rotl(h2, 31) + h1) & M64) * 5 + 0x38495AB5) & M64
    tail = data[16 * nb:]
    k1 = k2 = 0
    rem = n & 15
    for i in range(rem - 1, 7, -1):
        k2 ^= tail[i] << ((i - 8) * 8)
    if rem > 8:
        k2 = (_rotl((k2 * c2) & M64, 33) * c1) & M64
        h2 ^= k2
    for i in range(min(rem, 8) - 1, -1, -1):
        k1 ^= tail[i] << (i * 8)
    if rem > 0:
        k1 = (_rotl((k1 * c1) & M64, 31) * c2) & M64
        h1 ^= k1
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = _fmix64(h1), _fmix64(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def test_fmix64_pinned_by_smhasher_verification():
    # SMHasher VerificationTest: hash keys {0..i-1} with seed 256-i, hash the
    # concatenation with seed 0, read the first 4 bytes little-endian.
    key = bytes(range(256))
    hashes = b"".join(_murmur3_x64_128(key[:i], 256 - i) for i in range(256))
    final = _murmur3_x64_128(hashes, 0)
    assert int.from_bytes(final[:4], "little") == 0x6384BA69


def test_murmur3_structure_through_verified_fmix64():
    # For an 8-byte key the reference hash is one chunk round then the
    # length avalanche (hash.rs:47-76).  The chunk mix is fmix64 without its
    # first xorshift; xorshift-33 is an involution on 64 bits, so
    # mix(k) = fmix64(k ^ (k >> 33)).  The final stage is fmix64(h ^ len).
    rng = np.random.default_rng(7)
    for k in [0, 1, 0xFFFFFFFFFFFFFFFF] + [int(x) for x in rng.integers(0, 2**63, 200, dtype=np.int64)]:
        mix = _fmix64(k ^ (k >> 33))
        h = ((0x9E3779B97F4A7C15 ^ mix) * 5 + 0x52DCE729) & M64
        assert hash_bytes(k.to_bytes(8, "little"), 1) == _fmix64(h ^ 8)
    # 0-byte key: no chunk, no tail (rem == 0 and len % 8 == 0)
    assert hash_bytes(b"", 1) == _fmix64(0x9E3779B97F4A7C15 ^ 0)


_C2 = 0xC4CEB9FE1A85EC53  # fmix64's second multiplier (verified with it above)
_C2_INV = pow(_C2, -1, 1 << 64)


def _xs33(x):
    return x ^ (x >> 33)


def _chunk_mix(k):
    # k*c1, xs33, *c2, xs33 == fmix64 after undoing its leading xs33 (an
    # involution on 64 bits)
    return _fmix64(_xs33(k))


def _tail_mix(r):
    # xs33(r*c1), the first half of fmix64: fmix64(xs33(r)) = xs33(c2 * xs33(c1*r)),
    # so xs33(c1*r) = c2^-1 * xs33(fmix64(xs33(r))) mod 2^64
    return (_C2_INV * _xs33(_fmix64(_xs33(r)))) & M64


def _murmur_via_fmix64(key: bytes, chunk_mul=5, chunk_add=0x52DCE729, tail_when_zero=True, tail_order="little",
                       chain=True, fold_len=True, tail_mix="half"):
    """The reference composition (hash.rs:43-77) assembled from the verified
    fmix64; the keyword arguments plant restatement mistakes."""
    h = 0x9E3779B97F4A7C15
    n = len(key)
    for c in range(n // 8):
        k = int.from_bytes(key[8 * c:8 * c + 8], "little")
        if chain:
            h = ((h ^ _chunk_mix(k)) * chunk_mul + chunk_add) & M64
        else:  # mistake: each chunk restarts from the seed
            h = ((0x9E3779B97F4A7C15 ^ _chunk_mix(k)) * chunk_mul + chunk_add) & M64
    tail = key[8 * (n // 8):]
    r = int.from_bytes(tail, tail_order) if tail else 0
    if (r != 0 or n % 8 != 0) if tail_when_zero else r != 0:
        if tail_mix == "half":
            h ^= _tail_mix(r)
        elif tail_mix == "full":  # mistake: the whole fmix64 on the tail
            h ^= _fmix64(r)
    return _fmix64(h ^ n) if fold_len else _fmix64(h)


MURMUR_LENGTHS = (3, 5, 12, 16, 24, 32, 33, 40, 47)


def _murmur_keys():
    rng = np.random.default_rng(29)
    keys = []
    for n in MURMUR_LENGTHS:
        keys += [bytes(rng.integers(0, 256, n, dtype=np.uint8)) for _ in range(40)]
        keys.append(bytes(n))  # all-zero key
        keys.append(bytes(rng.integers(0, 256, n - n % 8, dtype=np.uint8)) + bytes(n % 8))  # zero tail only
    return keys


def test_murmur3_tail_and_chunk_chain_through_verified_fmix64():
    # Keys of 3..47 bytes: chunk chaining (2..5 chunks), the tail branch
    # (K % 8 != 0, hash.rs:61-68) including all-zero tails, and the length
    # fold; expected values come from fmix64 (SMHasher-pinned), the xs33
    # involution and the inverse of fmix64's constant, not from the oracle.
    for key in _murmur_keys():
        exp = _murmur_via_fmix64(key)
        assert hash_bytes(key, 1) == exp, (len(key), key.hex())
        assert pyoracle.murmur3_hash(key) == exp


@pytest.mark.parametrize("mistake", [dict(chunk_mul=3), dict(chunk_add=0x52DCE728), dict(tail_order="big"),
                                     dict(chain=False), dict(fold_len=False), dict(tail_mix="none"),
                                     dict(tail_mix="full")])
def test_murmur3_expectation_rejects_planted_mistakes(mistake):
    # Each plausible restatement mistake of hash.rs:43-77, planted in the
    # fmix64-based composition, disagrees with the oracle on some key of the
    # set above, so that set pins the tail branch and the chunk chain.
    keys = _murmur_keys()
    assert any(_murmur_via_fmix64(k, **mistake) != hash_bytes(k, 1) for k in keys)


def test_murmur3_zero_tail_condition_is_unobservable():
    # The tail branch's condition (rem != 0 || len % 8 != 0, hash.rs:64) cannot
    # change a result: an all-zero tail mixes to 0 (0 * c1 = 0, xs33(0) = 0),
    # so skipping the branch for it gives the same hash.
    assert _tail_mix(0) == 0
    for k in _murmur_keys():
        assert _murmur_via_fmix64(k, tail_when_zero=False) == hash_bytes(k, 1)


def test_hash_c_vs_python_restatement():
    rng = np.random.default_rng(11)
    for n in range(0, 65):
        for _ in range(8):
            key = bytes(rng.integers(0, 256, n, dtype=np.uint8))
            assert hash_bytes(key, 0) == pyoracle.identity_hash(key)
            assert hash_bytes(key, 1) == pyoracle.murmur3_hash(key), n
    # tail with all-zero bytes still mixes (len % 8 != 0 branch of hash.rs:64)
    assert hash_bytes(bytes(3), 1) == pyoracle.murmur3_hash(bytes(3))


def test_hash_many_matches_scalar():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 256, (100, 32), dtype=np.uint8)
    for mode in (0, 1):
        hs = hash_many(keys, 32, mode)
        assert all(int(hs[i]) == hash_bytes(bytes(keys[i]), mode) for i in range(100))


def test_murmur_golden_table():
    with open(os.path.join(GOLDEN_DIR, "hash_vectors.json")) as f:
        g = json.load(f)
    for row in g["vectors"]:
        key = bytes.fromhex(row["key"])
        assert hash_bytes(key, 0) == int(row["identity"], 16)
        assert hash_bytes(key, 1) == int(row["murmur3"], 16)


# ------------------------------------------------------- table semantics --
def _py_make(cap, K, V, mode):
    return _PyAdapter(cap, K, V, mode)


class _PyAdapter:
    def __init__(self, cap, K, V, mode):
        self.m = pyoracle.PyMap(cap, K, V, mode)
        self.K, self.V = K, V

    def bulk_insert(self, keys, vals):
        k = np.ascontiguousarray(keys).view(np.uint8).reshape(-1, self.K)
        v = np.ascontiguousarray(vals).view(np.uint8).reshape(-1, self.V)
        try:
            return self.m.insert([(bytes(k[i]), bytes(v[i])) for i in range(k.shape[0])])
        except pyoracle.TableFull as e:
            import flashmap_amd as fm

            raise fm.TableFull(5, str(e), e.occupied, e.capacity, e.load_factor)

    def bulk_remove(self, keys):
        k = np.ascontiguousarray(keys).view(np.uint8).reshape(-1, self.K)
        return self.m.remove([bytes(k[i]) for i in range(k.shape[0])])

    def bulk_get(self, keys):
        k = np.ascontiguousarray(keys).view(np.uint8).reshape(-1, self.K)
        res = self.m.get([bytes(k[i]) for i in range(k.shape[0])])
        vals = np.zeros((len(res), self.V), dtype=np.uint8)
        found = np.zeros(len(res), dtype=bool)
        for i, r in enumerate(res):
            if r is not None:
                vals[i] = np.frombuffer(r, dtype=np.uint8)
                found[i] = True
        return vals, found

    def clear(self):
        self.m.clear()

    def __len__(self):
        return self.m.length

    def export_slots(self):
        f, k, v = self.m.export()
        return np.array(f, dtype=np.uint32), np.frombuffer(k, dtype=np.uint8), np.frombuffer(v, dtype=np.uint8)


@pytest.mark.parametrize("test", REFERENCE_TESTS, ids=lambda t: t.__name__[2:])
def test_reference_tests_on_oracle(test):
    test(lambda c, K, V, mode: OracleAdapter(c, K, V, mode))


@pytest.mark.parametrize("name", sorted(ANOMALY_CASES))
def test_anomaly_histories(name):
    ops, expected = ANOMALY_CASES[name]
    assert run_history(OracleAdapter(16, 8, 8, 0), ops) == expected
    assert run_history(_PyAdapter(16, 8, 8, 0), ops) == expected


def _random_ops(rng, n_ops, K, V, key_space, max_batch):
    from helpers import random_history

    return random_history(rng, 0, K, V, n_ops, max_batch, key_space)


@pytest.mark.parametrize("seed", range(12))
def test_c_oracle_vs_python_random_histories(seed):
    from helpers import apply_and_compare

    rng = np.random.default_rng(100 + seed)
    cap = [16, 32, 64, 128][seed % 4]
    K, V = [(8, 8), (8, 16), (16, 8)][seed % 3]
    mode = seed % 2
    ops = _random_ops(rng, 40, K, V, key_space=cap * 3, max_batch=cap // 2)
    apply_and_compare(OracleAdapter(cap, K, V, mode), _PyAdapter(cap, K, V, mode), ops)


def test_exhaustive_short_histories_16_slots():
    """Every 3-op history over a small alphabet, C vs Python, 16 slots."""
    from helpers import apply_and_compare

    keys = [4, 20, 36, 5, 21]
    alphabet = []
    for a, b in itertools.permutations(keys, 2):
        alphabet.append(("ins", [a, b]))
    alphabet += [("ins", [k]) for k in keys] + [("rem", [k]) for k in keys] + [("rem", [20, 20])]
    alphabet += [("ins", [20, 20])]

    def rows(ks):
        a = np.zeros((len(ks), 8), dtype=np.uint8)
        a[:, :] = np.array(ks, dtype=np.uint64).view(np.uint8).reshape(-1, 8)
        return a

    count = 0
    for hist in itertools.product(alphabet, repeat=3):
        ops = []
        for i, (op, ks) in enumerate(hist):
            if op == "ins":
                ops.append(("ins", rows(ks), rows([100 + i] * len(ks))))
            else:
                ops.append(("rem", rows(ks)))
        ops.append(("get", rows(keys)))
        apply_and_compare(OracleAdapter(16, 8, 8, 0), _PyAdapter(16, 8, 8, 0), ops, check_slots=False)
        count += 1
    assert count == len(alphabet) ** 3


def _check_invariants(o: OracleMap, mode):
    flags, keys, _ = o.export()
    cap = o.capacity
    st = flags & 0xF
    occ = np.nonzero(st == 1)[0]
    # len == #OCCUPIED (every claim/tombstone is counted exactly once)
    assert len(o) == occ.size
    # (idx - dist) & mask == home(key) for every OCCUPIED slot
    homes = hash_many(keys.reshape(cap, -1)[occ], o.K, mode) & np.uint64(cap - 1)
    dist = (flags[occ] >> 4).astype(np.uint64)
    assert np.array_equal((occ.astype(np.uint64) - dist) & np.uint64(cap - 1), homes)
    # tombstones are stored raw (dist bits cleared)
    assert (flags[st == 2] == 2).all()


@pytest.mark.parametrize("seed", range(6))
def test_invariants_random_histories(seed):
    from helpers import random_history

    rng = np.random.default_rng(200 + seed)
    cap, mode = 256, seed % 2
    o = OracleAdapter(cap, 8, 8, mode)
    for op in random_history(rng, cap, 8, 8, 60, 64, 2000):
        try:
            if op[0] == "ins":
                o.bulk_insert(op[1], op[2])
            elif op[0] == "rem":
                o.bulk_remove(op[1])
            elif op[0] == "clear":
                o.clear()
        except Exception:
            pass
        _check_invariants(o.o, mode)


def test_dict_model_on_tombstone_free_histories():
    """Without removes and without in-batch duplicates the map is a dict."""
    rng = np.random.default_rng(5)
    o = OracleAdapter(1024, 8, 8, 1)
    model = {}
    for _ in range(30):
        ks = rng.choice(5000, size=20, replace=False).astype(np.uint64)
        vs = rng.integers(0, 2**63, size=20, dtype=np.int64).astype(np.uint64)
        new = sum(int(k) not in model for k in ks)
        assert o.bulk_insert(ks, vs) == new
        model.update({int(k): int(v) for k, v in zip(ks, vs)})
    q = np.arange(5000, dtype=np.uint64)
    v, f = o.bulk_get(q)
    vv = v.view(np.uint64).reshape(-1)
    for i in range(5000):
        assert bool(f[i]) == (i in model)
        if f[i]:
            assert int(vv[i]) == model[i]


def test_occupied_set_order_independent():
    """Distinct new keys into a tombstone-free table: the occupied slot set
    does not depend on insertion order (linear-probing parking)."""
    rng = np.random.default_rng(9)
    ks = rng.choice(10_000, size=180, replace=False).astype(np.uint64)
    sets = []
    for perm in range(4):
        o = OracleMap(256, 8, 8, 0)
        order = rng.permutation(ks.size)
        o.bulk_insert(ks[order].view(np.uint8), ks[order].view(np.uint8))
        sets.append(((o.export()[0] & 0xF) == 1))
    for s in sets[1:]:
        assert np.array_equal(s, sets[0])


def test_capacity_rounding():
    # rayon_cpu.rs:64
    for req, cap in [(0, 16), (1, 16), (16, 16), (17, 32), (1000, 1024), (1 << 20, 1 << 20), (10_000_000, 1 << 24)]:
        assert OracleMap(req, 8, 8, 0).capacity == cap


def test_config1_golden_digest():
    """Config-1 results recorded in tests/golden (made by make_golden.py)."""
    import hashlib

    from flashmap_amd.workloads import config1

    with open(os.path.join(GOLDEN_DIR, "config1_digest.json")) as f:
        g = json.load(f)
    c = config1()
    o = OracleMap(c.capacity, 8, 8, 0)
    assert o.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8)) == g["inserted"]
    v1, f1 = o.bulk_get(c.queries.view(np.uint8))
    assert int(f1.sum()) == g["hits1"]
    assert hashlib.sha256(v1.tobytes() + f1.tobytes()).hexdigest() == g["get1_sha256"]
    assert o.bulk_remove(c.remove_keys.view(np.uint8)) == g["removed"]
    v2, f2 = o.bulk_get(c.queries.view(np.uint8))
    assert int(f2.sum()) == g["hits2"]
    assert hashlib.sha256(v2.tobytes() + f2.tobytes()).hexdigest() == g["get2_sha256"]
    assert hashlib.sha256(o.export()[0].tobytes()).hexdigest() == g["flags_sha256"]


def test_rayon_port_agrees_where_order_cannot_matter():
    """Multi-threaded restatement vs oracle on distinct keys with no
    tombstones: every get result is order-independent."""
    from flashmap_amd.workloads import config1

    c = config1(n_insert=100_000, n_query=100_000, n_remove=0, capacity=1 << 18)
    o = OracleMap(c.capacity, 8, 8, 0)
    p = RayonPortMap(c.capacity, 8, 8, 0, threads=4)
    assert o.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8)) == \
        p.bulk_insert(c.insert_keys.view(np.uint8), c.insert_vals.view(np.uint8))
    vo, fo = o.bulk_get(c.queries.view(np.uint8))
    vp, fp = p.bulk_get(c.queries.view(np.uint8))
    assert np.array_equal(fo, fp) and np.array_equal(vo, vp)


@pytest.mark.parametrize("structure_only", [False, True])
def test_rayon_port_copy_restores_state(structure_only):
    """The restore the CPU-baseline timing loop uses: a copy answers gets and
    len like its source, and an op on the copy leaves the source alone."""
    rng = np.random.default_rng(13)
    keys = rng.integers(0, 256, (3000, 16), dtype=np.uint8)
    vals = rng.integers(0, 256, (3000, 8), dtype=np.uint8)
    src = RayonPortMap(1 << 13, 16, 8, 0, threads=3)
    dst = RayonPortMap(1 << 13, 16, 8, 0, threads=3)
    assert src.bulk_insert(keys, vals) == 3000
    dst.bulk_insert(keys[:10], vals[:10] ^ 1)
    dst.copy_from(src, structure_only=structure_only)
    assert len(dst) == len(src) == 3000
    vs, fs = src.bulk_get(keys)
    vd, fd = dst.bulk_get(keys)
    assert fd.all() and np.array_equal(fs, fd)
    if not structure_only:
        assert np.array_equal(vs, vd)
    assert dst.bulk_remove(keys[:500]) == 500
    assert len(src) == 3000 and src.bulk_get(keys[:500])[1].all()
    with pytest.raises(ValueError):
        dst.copy_from(RayonPortMap(1 << 12, 16, 8, 0, threads=1))
