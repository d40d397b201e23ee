"""Seeded synthetic workloads for the five FlashMap configurations.

Every random draw comes from SplitMix64 (Steele, Lea, Flood 2014; Vigna's
public-domain constants), evaluated in counter form so numpy can vectorise
it: output i of the stream seeded with ``seed`` is
``mix(seed + (i + 1) * 0x9E3779B97F4A7C15)`` -- identical to calling
``next()`` i+1 times.  A ``Stream`` hands out consecutive counter ranges.

Derived draws (all recorded here so a run can be regenerated bit for bit):
  * bounded integers in [0, m), m < 2^32: ``((x >> 32) * m) >> 32``
    (multiply-shift; bias below m / 2^32);
  * uniform permutation of m items: stable argsort of m fresh draws;
  * k distinct items out of m: the first k of such a permutation;
  * random bytes: successive outputs written little-endian;
  * log-uniform integers in [a, b]:
    ``min(b, floor(exp(ln a + U * (ln(b + 1) - ln a))))`` with U = x / 2^64;
  * Zipf(s) ranks over 1..N: rejection-inversion (Hoermann & Derflinger
    1996), vectorised with re-draws for rejected lanes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start..start+count-1 of the SplitMix64 stream seeded `seed`."""
    with np.errstate(over="ignore"):
        z = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z *= GOLDEN
        z += np.uint64(seed)
        t = z >> np.uint64(30)
        z ^= t
        z *= _M1
        np.right_shift(z, np.uint64(27), out=t)
        z ^= t
        z *= _M2
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z


class Stream:
    """Consecutive draws from one seeded SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64(self.seed, self.pos, n)
        self.pos += n
        return out

    def bytes(self, n_items: int, size: int) -> np.ndarray:
        """(n_items, size) uint8 of random bytes (size a multiple of 8)."""
        words = (size + 7) // 8
        return self.u64(n_items * words).view(np.uint8).reshape(n_items, words * 8)[:, :size].copy()

    def below(self, m: int, n: int) -> np.ndarray:
        assert 0 < m < (1 << 32)
        return ((self.u64(n) >> np.uint64(32)) * np.uint64(m)) >> np.uint64(32)

    def unit(self, n: int) -> np.ndarray:
        """Uniform doubles in [0, 1) from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def permutation(self, m: int) -> np.ndarray:
        return np.argsort(self.u64(m), kind="stable")

    def sample(self, m: int, k: int) -> np.ndarray:
        """k distinct indices of range(m), in random order.  For k << m:
        2k+64 bounded draws with repeats dropped (first occurrences kept);
        otherwise (or if too few survive) the first k of a permutation."""
        if 8 * k < m < (1 << 32):
            draws = self.below(m, 2 * k + 64).astype(np.int64)
            _, first = np.unique(draws, return_index=True)
            idx = draws[np.sort(first)]
            if idx.size >= k:
                return idx[:k]
        return self.permutation(m)[:k]

    def log_uniform(self, a: int, b: int, n: int) -> np.ndarray:
        u = self.unit(n)
        v = np.floor(np.exp(math.log(a) + u * (math.log(b + 1) - math.log(a))))
        return np.minimum(b, v).astype(np.int64)

    def zipf(self, n_ranks: int, s: float, n: int) -> np.ndarray:
        """Zipf(s) ranks in 1..n_ranks (rejection-inversion)."""
        return _zipf(self, n_ranks, s, n)


def _helper1(x):  # log1p(x)/x, -> 1 at 0
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    nz = np.abs(x) > 1e-8
    out[nz] = np.log1p(x[nz]) / x[nz]
    out[~nz] = 1 - x[~nz] * (0.5 - x[~nz] * (1.0 / 3 - 0.25 * x[~nz]))
    return out


def _helper2(x):  # expm1(x)/x, -> 1 at 0
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    nz = np.abs(x) > 1e-8
    out[nz] = np.expm1(x[nz]) / x[nz]
    out[~nz] = 1 + x[~nz] * 0.5 * (1 + x[~nz] * (1.0 / 3) * (1 + 0.25 * x[~nz]))
    return out


def _zipf(st: Stream, N: int, s: float, n: int) -> np.ndarray:
    def h(x):
        return np.exp(-s * np.log(x))

    def H(x):
        lx = np.log(x)
        return _helper2((1.0 - s) * lx) * lx

    def Hinv(x):
        t = np.maximum(x * (1.0 - s), -1.0 + 1e-17)
        return np.exp(_helper1(t) * x)

    hx1 = float(H(np.array([1.5]))[0] - 1.0)
    hN = float(H(np.array([N + 0.5]))[0])
    sq = float(2.0 - Hinv(np.array([H(np.array([2.5]))[0] - h(np.array([2.0]))[0]]))[0])
    out = np.zeros(n, dtype=np.int64)
    todo = np.arange(n)
    while todo.size:
        u = hN + st.unit(todo.size) * (hx1 - hN)
        x = Hinv(u)
        k = np.clip(np.floor(x + 0.5), 1, N)
        ok = (k - x <= sq) | (u >= H(k + 0.5) - h(k))
        out[todo[ok]] = k[ok].astype(np.int64)
        todo = todo[~ok]
    return out


# ---------------------------------------------------------------------------
# Config 1: FlashMap<u64,u64> with_capacity(2^20), seed 1.
# ---------------------------------------------------------------------------
@dataclass
class Config1:
    capacity: int
    insert_keys: np.ndarray  # u64
    insert_vals: np.ndarray  # u64
    queries: np.ndarray      # u64
    remove_keys: np.ndarray  # u64


def config1(n_insert=500_000, n_query=500_000, n_remove=100_000, capacity=1 << 20, seed=1) -> Config1:
    st = Stream(seed)
    raw = st.u64(n_insert + 64)
    _, first = np.unique(raw, return_index=True)  # reject repeats, keep stream order
    keys = raw[np.sort(first)][:n_insert]
    assert keys.size == n_insert
    vals = keys ^ GOLDEN
    hits = keys[st.sample(n_insert, n_query // 2)]
    fresh = st.u64(n_query - n_query // 2 + 64)
    fresh = fresh[~np.isin(fresh, keys)][: n_query - n_query // 2]
    q = np.concatenate([hits, fresh])
    q = q[st.permutation(q.size)]
    rem = keys[st.sample(n_insert, n_remove)]
    return Config1(capacity, keys, vals, q, rem)


# ---------------------------------------------------------------------------
# Config 2: FlashMap<[u8;32],[u8;128]> with_capacity(10,000,000), seed 2.
# ---------------------------------------------------------------------------
@dataclass
class Config2:
    capacity: int
    preload_keys: np.ndarray   # (n, 32) u8
    preload_vals: np.ndarray   # (n, 128) u8
    batches: List = field(default_factory=list)  # [(keys, vals)]
    get_keys: np.ndarray = None
    remove_keys: np.ndarray = None


def config2(n_preload=8_000_000, n_batches=20, batch=100_000, n_get=1_000_000, hit=0.9,
            n_remove=200_000, capacity=10_000_000, seed=2, K=32, V=128) -> Config2:
    st = Stream(seed)
    pk = st.bytes(n_preload, K)
    pv = st.bytes(n_preload, V)
    cfg = Config2(capacity, pk, pv)
    for _ in range(n_batches):
        n_over = batch // 2
        over = pk[st.sample(n_preload, n_over)]
        new = st.bytes(batch - n_over, K)
        ks = np.concatenate([over, new])
        perm = st.permutation(batch)
        cfg.batches.append((ks[perm], st.bytes(batch, V)))
    n_hit = int(round(n_get * hit))
    gk = np.concatenate([pk[st.below(n_preload, n_hit).astype(np.int64)], st.bytes(n_get - n_hit, K)])
    cfg.get_keys = gk[st.permutation(n_get)]
    cfg.remove_keys = pk[st.sample(n_preload, n_remove)]
    return cfg


# ---------------------------------------------------------------------------
# Config 3: FlashMap<[u8;32],u64> with_capacity(50,000,000), seed 3.
# Load keys with value = row id; probes: 90% Zipf(1.0) over the inserted keys
# (rank -> key through a random permutation), 10% fresh misses, shuffled.
# ---------------------------------------------------------------------------
@dataclass
class Config3:
    capacity: int
    keys: np.ndarray     # (n, 32) u8
    row_ids: np.ndarray  # (n,) u64
    probes: np.ndarray   # (m, 32) u8
    probe_hit: np.ndarray  # (m,) bool -- True where the probe is an inserted key
    probe_row: np.ndarray = None  # (m,) int64 -- row id of an inserted probe, -1 for a miss


def config3(n_load=40_000_000, n_probe=100_000_000, zipf_s=1.0, hit=0.9, capacity=50_000_000, seed=3,
            K=32) -> Config3:
    st = Stream(seed)
    keys = st.bytes(n_load, K)
    rows = np.arange(n_load, dtype=np.uint64)
    n_hit = int(round(n_probe * hit))
    perm = st.permutation(n_load)
    ranks = st.zipf(n_load, zipf_s, n_hit) - 1
    hits = keys[perm[ranks]]
    miss = st.bytes(n_probe - n_hit, K)
    order = st.permutation(n_probe)
    probes = np.concatenate([hits, miss])[order]
    is_hit = np.concatenate([np.ones(n_hit, bool), np.zeros(n_probe - n_hit, bool)])[order]
    row = np.concatenate([perm[ranks].astype(np.int64), np.full(n_probe - n_hit, -1, np.int64)])[order]
    return Config3(capacity, keys, rows, probes, is_hit, row)


# ---------------------------------------------------------------------------
# Config 4: 512 independent FlashMap<u64,[u8;32]> tables, seed 4.
# ---------------------------------------------------------------------------
@dataclass
class TableWork:
    capacity: int           # rounded slot count
    preload_keys: np.ndarray  # u64
    preload_vals: np.ndarray  # (n, 32) u8
    ins_keys: np.ndarray
    ins_vals: np.ndarray
    get_keys: np.ndarray
    rem_keys: np.ndarray


def _round_cap(c):
    c = max(16, int(c))
    return 1 << (c - 1).bit_length()


def config4_table(t: int, seed=4, cap_lo=1 << 10, cap_hi=1 << 22, ins_hi=1 << 16, get_hi=1 << 16,
                  rem_hi=1 << 12, V=32) -> TableWork:
    """Table t of config 4 (each table has its own sub-stream seed*1000003+t)."""
    st = Stream(seed * 1000003 + t)
    cap = _round_cap(int(st.log_uniform(cap_lo, cap_hi, 1)[0]))
    lf = 0.5 + 0.35 * float(st.unit(1)[0])
    n_pre = int(lf * cap)
    # SplitMix64's output is a bijection of its counter: draws from one
    # stream never repeat, so preload and fresh keys are distinct by design.
    pre = st.u64(n_pre)
    pre_v = st.bytes(n_pre, V)
    n_ins = int(st.log_uniform(1, ins_hi, 1)[0])
    n_over = min(int(round(0.3 * n_ins)), n_pre)
    over = pre[st.sample(n_pre, n_over)] if n_over else np.zeros(0, np.uint64)
    fresh = st.u64(n_ins - n_over)
    ik = np.concatenate([over, fresh])
    ik = ik[st.permutation(ik.size)]
    iv = st.bytes(ik.size, V)
    n_get = int(st.log_uniform(1, get_hi, 1)[0])
    n_hit = int(round(0.7 * n_get))
    gk = np.concatenate([pre[st.below(max(n_pre, 1), n_hit).astype(np.int64)] if n_pre else
                         np.zeros(0, np.uint64), st.u64(n_get - n_hit)])
    gk = gk[st.permutation(gk.size)]
    n_rem = min(int(st.log_uniform(1, rem_hi, 1)[0]), n_pre)
    rk = pre[st.sample(n_pre, n_rem)] if n_rem else np.zeros(0, np.uint64)
    return TableWork(cap, pre, pre_v, ik, iv, gk, rk)


# ---------------------------------------------------------------------------
# Config 5: serving log over one FlashMap<[u8;32],u64> with_capacity(24M),
# seed 5.  Requests arrive as a Poisson stream; 70% get / 20% insert /
# 10% remove; sizes log-uniform in [64, 262144]; get keys 80% hits under
# Zipf(0.9) over the preloaded keys (random rank permutation), 20% misses;
# inserts 50% overwrites (Zipf) + 50% new keys; removes 80% uniform hits,
# 20% misses.
# ---------------------------------------------------------------------------
@dataclass
class Request:
    op: str              # "get" | "insert" | "remove"
    arrival_s: float
    keys: np.ndarray     # (n, 32) u8
    vals: np.ndarray = None  # (n,) u64 for inserts


def config5_preload(n_preload=20_000_000, capacity=24_000_000, seed=5, K=32):
    st = Stream(seed)
    keys = st.bytes(n_preload, K)
    vals = st.u64(n_preload)
    return capacity, keys, vals


def config5_requests(preload_keys: np.ndarray, rate: float, n_requests: int, seed=5,
                     size_lo=64, size_hi=262_144, zipf_s=0.9) -> List[Request]:
    st = Stream(seed * 7919 + int(rate))
    n_pre = preload_keys.shape[0]
    perm = st.permutation(n_pre)
    gaps = -np.log1p(-st.unit(n_requests)) / rate
    arrivals = np.cumsum(gaps)
    kinds = st.unit(n_requests)
    sizes = st.log_uniform(size_lo, size_hi, n_requests)
    out = []
    for r in range(n_requests):
        n = int(sizes[r])
        u = kinds[r]
        if u < 0.7:
            n_hit = int(round(0.8 * n))
            hk = preload_keys[perm[st.zipf(n_pre, zipf_s, n_hit) - 1]]
            ks = np.concatenate([hk, st.bytes(n - n_hit, preload_keys.shape[1])])
            out.append(Request("get", float(arrivals[r]), ks[st.permutation(n)]))
        elif u < 0.9:
            n_over = n // 2
            ok = preload_keys[perm[st.zipf(n_pre, zipf_s, n_over) - 1]]
            ks = np.concatenate([ok, st.bytes(n - n_over, preload_keys.shape[1])])
            out.append(Request("insert", float(arrivals[r]), ks[st.permutation(n)], st.u64(n)))
        else:
            n_hit = int(round(0.8 * n))
            hk = preload_keys[st.below(n_pre, n_hit).astype(np.int64)]
            ks = np.concatenate([hk, st.bytes(n - n_hit, preload_keys.shape[1])])
            out.append(Request("remove", float(arrivals[r]), ks[st.permutation(n)]))
    return out
