"""Independent pure-Python restatement of the reference CPU path.

TEST INFRASTRUCTURE ONLY (small cases).  Written separately from
``fm_oracle.c`` so that the two cross-check each other: hashing follows
src/hash.rs:30-86, the table follows src/rayon_cpu.rs:63-353 with each batch
applied in index order.  Slots are kept as Python tuples
``(state, dist, key_bytes, value_bytes)`` instead of packed flags, and the
flag word is only assembled on export, so a packing mistake in one
restatement does not repeat in the other.
"""

M64 = (1 << 64) - 1

EMPTY, OCCUPIED, TOMBSTONE = 0, 1, 2


def identity_hash(key: bytes) -> int:
    """src/hash.rs:31-39."""
    return int.from_bytes(key[:8], "little")


def murmur3_hash(key: bytes) -> int:
    """src/hash.rs:42-77 (wrapping u64 arithmetic)."""
    h = 0x9E3779B97F4A7C15
    n = len(key)
    full = n // 8
    for c in range(full):
        k = int.from_bytes(key[8 * c:8 * c + 8], "little")
        k = (k * 0xFF51AFD7ED558CCD) & M64
        k ^= k >> 33
        k = (k * 0xC4CEB9FE1A85EC53) & M64
        k ^= k >> 33
        h ^= k
        h = (h * 5 + 0x52DCE729) & M64
    tail = key[8 * full:]
    rem = int.from_bytes(tail, "little") if tail else 0
    if rem != 0 or n % 8 != 0:
        rem = (rem * 0xFF51AFD7ED558CCD) & M64
        rem ^= rem >> 33
        h ^= rem
    h ^= n
    h ^= h >> 33
    h = (h * 0xFF51AFD7ED558CCD) & M64
    h ^= h >> 33
    h = (h * 0xC4CEB9FE1A85EC53) & M64
    h ^= h >> 33
    return h


def hash_key(key: bytes, mode: int) -> int:
    """src/hash.rs:80-86 (mode 0 identity, 1 murmur3; hash.rs:22-27)."""
    return identity_hash(key) if mode == 0 else murmur3_hash(key)


class TableFull(Exception):
    """FlashMapError::TableFull (src/error.rs:17-24)."""

    def __init__(self, occupied, capacity):
        self.occupied = occupied
        self.capacity = capacity
        self.load_factor = occupied / capacity * 100.0
        super().__init__(f"table full: {occupied} occupied of {capacity}")


class PyMap:
    """Sequential FlashMap over a list of slot tuples."""

    def __init__(self, capacity: int, key_size: int, value_size: int, mode: int = 0):
        cap = 1
        while cap < max(capacity, 16):
            cap *= 2
        self.cap = cap
        self.K = key_size
        self.V = value_size
        self.mode = mode
        self.slots = [(EMPTY, 0, bytes(key_size), bytes(value_size))] * cap
        self.length = 0

    def _home(self, key):
        return hash_key(key, self.mode) % self.cap

    def _find(self, key):
        home = self._home(key)
        for p in range(self.cap):
            i = (home + p) % self.cap
            st, d, k, _ = self.slots[i]
            if st == EMPTY:
                return None
            if st == OCCUPIED:
                if p > d:
                    return None
                if k == key:
                    return i
        return None

    def get(self, keys):
        out = []
        for q in keys:
            i = self._find(q)
            out.append(None if i is None else self.slots[i][3])
        return out

    def insert(self, pairs):
        if self.length + len(pairs) > self.cap:
            raise TableFull(self.length, self.cap)
        new = 0
        for key, val in pairs:
            carry_k, carry_v = key, val
            home = self._home(carry_k)
            p = 0
            while True:
                if p >= self.cap:
                    break  # carried pair dropped
                i = (home + p) % self.cap
                st, d, k, v = self.slots[i]
                if st == OCCUPIED and k == carry_k:
                    self.slots[i] = (st, d, k, carry_v)
                    break
                if st == OCCUPIED and p > d:
                    self.slots[i] = (OCCUPIED, p, carry_k, carry_v)
                    carry_k, carry_v = k, v
                    home = self._home(carry_k)
                    p = d + 1
                    continue
                if st == OCCUPIED:
                    p += 1
                    continue
                self.slots[i] = (OCCUPIED, p, carry_k, carry_v)
                new += 1
                break
        self.length += new
        return new

    def remove(self, keys):
        removed = 0
        for q in keys:
            i = self._find(q)
            if i is not None:
                _, _, k, v = self.slots[i]
                self.slots[i] = (TOMBSTONE, 0, k, v)
                removed += 1
        self.length -= removed
        return removed

    def clear(self):
        self.slots = [(EMPTY, 0, k, v) for (_, _, k, v) in self.slots]
        self.length = 0

    def export(self):
        """(flags list, keys bytes, values bytes) with non-OCCUPIED zeroed."""
        flags, ks, vs = [], bytearray(), bytearray()
        for st, d, k, v in self.slots:
            flags.append(((d << 4) & 0xFFFFFFFF) | st)
            occ = st == OCCUPIED
            ks += k if occ else bytes(self.K)
            vs += v if occ else bytes(self.V)
        return flags, bytes(ks), bytes(vs)
