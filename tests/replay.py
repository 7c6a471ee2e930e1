"""Independent replay of a Ligerito proof (test infrastructure).

From a proof's bytes and hashlib's SHA-256 alone, this re-derives every
Fiat–Shamir challenge and query set of the prover's transcript, in the order
src/prover.rs:29-195 issues them (src/emulated_fs.rs:12-75 for the
transcript), and checks the proof against values computed here from the
input polynomial:

* each commitment root equals the root of the matrix the prover must have
  committed: the input, then the polynomial folded by the derived
  challenges (ligero.rs:70-95; the encoded matrix and its Merkle tree come
  from the `commit` callback, the oracle's ligero_commit in the tests);
* the opened rows are that matrix's rows at the derived queries, and the
  Merkle proof nodes verify against the root in the reference's batched
  order (merkle_tree.rs:113-188, BLAKE3 written here from its specification);
* s0 + s1 of every sumcheck triple equals the running sum: the enforced sum
  of the openings (sumcheck_polys.rs:22-72), then the previous quadratic at
  the round's challenge (ligerito.rs:143-296), glued as the verifier glues;
* where the basis is small enough to materialise (2^ell entries, ell <=
  max_basis_log), every triple equals the 0/1/3 evaluation of the folded
  polynomial against the summed basis polynomial (multilinear_poly.rs:90-131);
* yr equals the polynomial folded by every challenge, satisfies
  ligero_verify (ligero.rs:97-131), and the final sumcheck check holds
  (verify_partial, ligerito.rs:282-295), with the basis partially evaluated
  through its tensor structure rather than materialised.

Nothing here calls the C library or the oracle's prover or verifier: field
arithmetic, folds, the basis polynomials, the transcript, BLAKE3 and the
proof parser are written in this file from the reference's sources and the
mathematics.  Vectors of field elements are numpy arrays: uint64 values below
2^32 for GF(2^32), (n, 2) uint64 limb pairs (low first) for GF(2^128).
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np

U64 = np.uint64
POLY_LOW = {32: 0x8299, 128: 0x87}  # binary_field.rs:178-189


# ------------------------------------------------------------ the fields --
class Field:
    """GF(2^bits) in polynomial basis: carry-less product reduced modulo
    x^bits + POLY_LOW[bits] (binary_field.rs:52-140)."""

    def __init__(self, bits: int):
        assert bits in (32, 128)
        self.bits = bits
        self.nbytes = bits // 8
        self.low = POLY_LOW[bits]

    # scalars are Python ints
    def mul(self, a: int, b: int) -> int:
        r = 0
        while b:
            if b & 1:
                r ^= a
            a <<= 1
            b >>= 1
        for bit in range(r.bit_length() - 1, self.bits - 1, -1):
            if (r >> bit) & 1:
                r ^= ((1 << self.bits) | self.low) << (bit - self.bits)
        return r

    def inv(self, a: int) -> int:
        assert a != 0, "inverse of zero"
        # Fermat: a^(2^n - 2)
        r, e = 1, (1 << self.bits) - 2
        while e:
            if e & 1:
                r = self.mul(r, a)
            a = self.mul(a, a)
            e >>= 1
        return r

    def from_bytes(self, b: bytes) -> int:
        return int.from_bytes(b, "little")

    def to_bytes(self, v: int) -> bytes:
        return v.to_bytes(self.nbytes, "little")

    # vectors
    def vec(self, ints):
        if self.bits == 32:
            return np.array(ints, dtype=U64)
        a = np.zeros((len(ints), 2), dtype=U64)
        for i, v in enumerate(ints):
            a[i, 0] = v & 0xFFFFFFFFFFFFFFFF
            a[i, 1] = v >> 64
        return a

    def from_array(self, arr):
        """Input layout of the library and the oracle (uint32, or (n, 2)
        uint64) to this module's vectors."""
        if self.bits == 32:
            return np.asarray(arr, dtype=np.uint32).astype(U64)
        return np.ascontiguousarray(arr, dtype=U64).reshape(-1, 2).copy()

    def get(self, v, i) -> int:
        if self.bits == 32:
            return int(v[i])
        return int(v[i, 0]) | (int(v[i, 1]) << 64)

    def to_ints(self, v):
        return [self.get(v, i) for i in range(len(v))]

    def zeros(self, n):
        return np.zeros(n, dtype=U64) if self.bits == 32 else np.zeros((n, 2), dtype=U64)

    def xor_sum(self, v) -> int:
        if len(v) == 0:
            return 0
        s = np.bitwise_xor.reduce(v, axis=0)
        return int(s) if self.bits == 32 else int(s[0]) | (int(s[1]) << 64)

    # --- GF(2^32): values in uint64 lanes
    def _reduce32(self, acc):
        # x^32 = 0x8299 (x^15 + x^9 + x^7 + x^4 + x^3 + 1): fold the high
        # word twice (31 bits, then at most 14)
        for _ in range(2):
            hi = acc >> U64(32)
            acc = acc & U64(0xFFFFFFFF)
            for sh in (0, 3, 4, 7, 9, 15):
                acc ^= hi << U64(sh)
        return acc

    # --- GF(2^128): (lo, hi) limbs
    @staticmethod
    def _shl128_into4(lo, hi, i):
        z = np.zeros_like(lo)
        if i == 0:
            return lo, hi, z, z
        if i < 64:
            s, t = U64(i), U64(64 - i)
            return lo << s, (hi << s) | (lo >> t), hi >> t, z
        if i == 64:
            return z, lo, hi, z
        j = i - 64
        s, t = U64(j), U64(64 - j)
        return z, lo << s, (hi << s) | (lo >> t), hi >> t

    @staticmethod
    def _reduce256(a0, a1, a2, a3):
        # x^128 = x^7 + x^2 + x + 1: fold (a2, a3), then the <= 7 bits that
        # spill over x^128 once more
        t0, t1, t2 = a2.copy(), a3.copy(), np.zeros_like(a2)
        for k in (1, 2, 7):
            s, r = U64(k), U64(64 - k)
            t0 = t0 ^ (a2 << s)
            t1 = t1 ^ ((a3 << s) | (a2 >> r))
            t2 = t2 ^ (a3 >> r)
        a0 = a0 ^ t0
        a1 = a1 ^ t1
        a0 = a0 ^ t2 ^ (t2 << U64(1)) ^ (t2 << U64(2)) ^ (t2 << U64(7))
        return a0, a1

    def vmul_s(self, v, s: int):
        """v * s (s a scalar)."""
        if self.bits == 32:
            acc = np.zeros_like(v)
            i = 0
            while s:
                if s & 1:
                    acc ^= v << U64(i)
                s >>= 1
                i += 1
            return self._reduce32(acc)
        lo, hi = v[:, 0], v[:, 1]
        a = [np.zeros_like(lo) for _ in range(4)]
        i = 0
        while s:
            if s & 1:
                p = self._shl128_into4(lo, hi, i)
                for k in range(4):
                    a[k] ^= p[k]
            s >>= 1
            i += 1
        r0, r1 = self._reduce256(*a)
        return np.stack([r0, r1], axis=1)

    def vmul(self, u, v):
        """Elementwise u * v."""
        if self.bits == 32:
            acc = np.zeros_like(u)
            for i in range(32):
                acc ^= (u << U64(i)) * ((v >> U64(i)) & U64(1))
            return self._reduce32(acc)
        lo, hi = u[:, 0], u[:, 1]
        a = [np.zeros_like(lo) for _ in range(4)]
        for i in range(128):
            bit = (v[:, i // 64] >> U64(i % 64)) & U64(1)
            p = self._shl128_into4(lo, hi, i)
            for k in range(4):
                a[k] ^= p[k] * bit
        r0, r1 = self._reduce256(*a)
        return np.stack([r0, r1], axis=1)


# ---------------------------------------------------- reference formulas --
def lagrange_basis(F: Field, rs):
    """evaluate_lagrange_basis (utils.rs:4-27): r0 selects the index MSB."""
    layer = [1 ^ rs[0], rs[0]]
    for r in rs[1:]:
        nxt = []
        for c in layer:
            nxt += [F.mul(c, 1 ^ r), F.mul(c, r)]
        layer = nxt
    return layer


def fold(F: Field, v, r: int):
    """One MSB-first partial evaluation (multilinear_poly.rs:66-87):
    (1 + r) e[i] + r e[i + half] = e[i] + r (e[i] + e[i + half])."""
    h = len(v) // 2
    lo, hi = v[:h], v[h:]
    return lo ^ F.vmul_s(lo ^ hi, r)


def partial_eval(F: Field, v, rs):
    for r in rs:
        v = fold(F, v, r)
    return v


def sks_vks(F: Field, ell: int):
    """s_i(v_i) for i = 0..ell with v_i = 2^i, s_0(x) = x and
    s_i(x) = s_{i-1}(x)^2 + s_{i-1}(v_{i-1}) s_{i-1}(x) (eval_sk_at_vks,
    binary_fft.rs:217-246, next_s :101-107)."""
    out = [1]
    for i in range(1, ell + 1):
        # s_{i-1} at v_i, then one more recursion step
        x = 1 << i
        s = x
        for t in range(i - 1):
            s = F.mul(s, s) ^ F.mul(out[t], s)
        out.append(F.mul(s, s) ^ F.mul(out[i - 1], s))
    return out


def sks_at(F: Field, sk, x: int, ell: int):
    """[s_0(x), ..., s_{ell-1}(x)] (evaluate_scaled_basis_inplace,
    binary_fft.rs:276-301)."""
    s = [x]
    for t in range(1, ell):
        p = s[-1]
        s.append(F.mul(p, p) ^ F.mul(sk[t - 1], p))
    return s


def tensor(F: Field, scale: int, factors):
    """basis[j] = scale * prod_{t in bits(j)} factors[t] (bit t of j picks
    factor t, binary_fft.rs:294-300)."""
    v = F.vec([scale])
    for f in factors:
        v = np.concatenate([v, F.vmul_s(v, f)])
    return v


def quad_from_evals(F: Field, g0: int, g1: int, g2: int):
    """quadratic_from_evals at x = 3 (sumcheck.rs:67-81) as (a, b, c)."""
    x = 3
    num = g2 ^ g0 ^ F.mul(x, g1 ^ g0)
    den = F.mul(x, x) ^ x
    a = F.mul(num, F.inv(den))
    return a, g1 ^ g0 ^ a, g0


def quad_eval(F: Field, q, r: int) -> int:
    a, b, c = q
    return F.mul(F.mul(a, r), r) ^ F.mul(b, r) ^ c


def eval013(F: Field, f, b):
    """eval_013_product (multilinear_poly.rs:90-131): sums of f*b over the
    halves at X = 0, 1 and 3 (f(3) = 3 f1 + (1 + 3) f0)."""
    h = len(f) // 2
    f0, f1, b0, b1 = f[:h], f[h:], b[:h], b[h:]
    f3 = F.vmul_s(f1, 3) ^ F.vmul_s(f0, 2)
    b3 = F.vmul_s(b1, 3) ^ F.vmul_s(b0, 2)
    return F.xor_sum(F.vmul(f0, b0)), F.xor_sum(F.vmul(f1, b1)), F.xor_sum(F.vmul(f3, b3))


# ----------------------------------------------------------- transcript --
class Transcript:
    """FS (emulated_fs.rs:6-75) over hashlib: each squeeze hashes everything
    absorbed so far plus the u32 LE counter of every squeeze up to it."""

    def __init__(self, seed: int):
        self.h = hashlib.sha256(struct.pack("<I", seed))
        self.counter = 0

    def absorb(self, b: bytes):
        self.h.update(b)

    def squeeze(self) -> int:
        self.h.update(struct.pack("<I", self.counter))
        self.counter += 1
        return int.from_bytes(self.h.copy().digest()[:16], "little")

    def get_field(self, F: Field) -> int:
        return self.squeeze() & ((1 << F.bits) - 1)

    def get_distinct_queries(self, n: int, s: int):
        assert s <= n
        out, seen = [], set()
        while len(out) < s:
            q = (self.squeeze() & 0xFFFFFFFFFFFFFFFF) % n  # u128 as usize (64-bit)
            if q not in seen:
                seen.add(q)
                out.append(q)
        return sorted(out)


# ---------------------------------------------------------------- BLAKE3 --
_IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
_PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
_CHUNK_START, _CHUNK_END, _PARENT, _ROOT = 1, 2, 4, 8
_M32 = 0xFFFFFFFF


def _rotr(x, n):
    return ((x >> n) | (x << (32 - n))) & _M32


def _g(v, a, b, c, d, mx, my):
    v[a] = (v[a] + v[b] + mx) & _M32
    v[d] = _rotr(v[d] ^ v[a], 16)
    v[c] = (v[c] + v[d]) & _M32
    v[b] = _rotr(v[b] ^ v[c], 12)
    v[a] = (v[a] + v[b] + my) & _M32
    v[d] = _rotr(v[d] ^ v[a], 8)
    v[c] = (v[c] + v[d]) & _M32
    v[b] = _rotr(v[b] ^ v[c], 7)


def _compress(cv, block, counter, blen, flags):
    m = list(struct.unpack("<16I", block))
    v = list(cv) + _IV[:4] + [counter & _M32, counter >> 32, blen, flags]
    for rnd in range(7):
        _g(v, 0, 4, 8, 12, m[0], m[1])
        _g(v, 1, 5, 9, 13, m[2], m[3])
        _g(v, 2, 6, 10, 14, m[4], m[5])
        _g(v, 3, 7, 11, 15, m[6], m[7])
        _g(v, 0, 5, 10, 15, m[8], m[9])
        _g(v, 1, 6, 11, 12, m[10], m[11])
        _g(v, 2, 7, 8, 13, m[12], m[13])
        _g(v, 3, 4, 9, 14, m[14], m[15])
        if rnd < 6:
            m = [m[i] for i in _PERM]
    return [v[i] ^ v[i + 8] for i in range(8)] + [v[i + 8] ^ cv[i] for i in range(8)]


def _chunk_output(chunk: bytes, index: int):
    """(cv, block, counter, len, flags) of a chunk's last compression."""
    cv = _IV
    blocks = [chunk[i:i + 64] for i in range(0, len(chunk), 64)] or [b""]
    for k, blk in enumerate(blocks):
        flags = (_CHUNK_START if k == 0 else 0) | (_CHUNK_END if k == len(blocks) - 1 else 0)
        padded = blk + b"\0" * (64 - len(blk))
        if k == len(blocks) - 1:
            return cv, padded, index, len(blk), flags
        cv = _compress(cv, padded, index, 64, flags)[:8]


def _cv_of(out):
    return _compress(*out)[:8]


def _parent(left_cv, right_cv):
    return _IV, struct.pack("<16I", *(left_cv + right_cv)), 0, 64, _PARENT


def blake3(data: bytes) -> bytes:
    """BLAKE3 hash mode, 32-byte output (the BLAKE3 specification, 2020)."""
    chunks = [data[i:i + 1024] for i in range(0, len(data), 1024)] or [b""]
    stack = []
    for idx, ch in enumerate(chunks[:-1]):
        cv = _cv_of(_chunk_output(ch, idx))
        total = idx + 1
        while total & 1 == 0:
            cv = _cv_of(_parent(stack.pop(), cv))
            total >>= 1
        stack.append(cv)
    out = _chunk_output(chunks[-1], len(chunks) - 1)
    while stack:
        out = _parent(stack.pop(), _cv_of(out))
    cv, block, counter, blen, flags = out
    return struct.pack("<8I", *_compress(cv, block, counter, blen, flags | _ROOT)[:8])


def merkle_verify(root: bytes, proof, depth: int, leaves, indices) -> bool:
    """verify (merkle_tree.rs:164-188) with verify_ith_layer (:120-162):
    layer by layer, a query's sibling is the next query when both are opened,
    else the next proof node; the last query of a layer always takes a node."""
    layer = [blake3(x) for x in leaves]
    q = list(indices)
    cnt, pc = len(q), 0
    h2 = lambda a, b: blake3(a + b)  # noqa: E731  (merkle_tree.rs:113-118)
    for _ in range(depth):
        nxt, i = 0, 0
        while i < cnt:
            query = q[i]
            nxt += 1
            q[nxt - 1] = query >> 1
            if i == cnt - 1:
                pc += 1
                sib = proof[pc - 1]
                layer[nxt - 1] = h2(sib, layer[i]) if query % 2 else h2(layer[i], sib)
                break
            if query % 2:
                pc += 1
                layer[nxt - 1] = h2(proof[pc - 1], layer[i])
                i += 1
            elif q[i + 1] != query ^ 1:
                pc += 1
                layer[nxt - 1] = h2(layer[i], proof[pc - 1])
                i += 1
            else:
                layer[nxt - 1] = h2(layer[i], layer[i + 1])
                i += 2
        cnt = nxt
    return cnt == 1 and layer[0] == root


# ---------------------------------------------------------------- proofs --
def parse_proof(data: bytes, bits: int):
    """The library's canonical encoding of FinalizedLigeritoProof
    (data_structures.rs:91-103; layout in include/ligerito.h)."""
    F = Field(bits)
    pos = [0]

    def take(n):
        b = data[pos[0]:pos[0] + n]
        assert len(b) == n, "truncated proof"
        pos[0] += n
        return b

    u32 = lambda: struct.unpack("<I", take(4))[0]  # noqa: E731
    u64 = lambda: struct.unpack("<Q", take(8))[0]  # noqa: E731
    elems = lambda n: [F.from_bytes(take(F.nbytes)) for _ in range(n)]  # noqa: E731

    def rlp():
        nr, nc = u64(), u64()
        rows = [elems(nc) for _ in range(nr)]
        nn = u64()
        return {"rows": rows, "nodes": [take(32) for _ in range(nn)]}

    assert take(4) == b"LGRT" and u32() == 1 and u32() == F.nbytes
    p = {"initial_root": take(32), "initial": rlp()}
    p["rec_roots"] = [take(32) for _ in range(u64())]
    p["rec_proofs"] = [rlp() for _ in range(u64())]
    p["yr"] = elems(u64())
    p["final"] = rlp()
    p["tr"] = [tuple(elems(3)) for _ in range(u64())]
    assert pos[0] == len(data), "trailing bytes"
    return p


# ---------------------------------------------------------------- replay --
class ReplayError(AssertionError):
    pass


def _check(cond, what):
    if not cond:
        raise ReplayError(what)


def replay(proof: bytes, poly, shape, bits: int, commit, fs_seed: int = 1234, max_basis_log: int = 14):
    """Checks `proof` for input `poly` under the configuration `shape` (dict
    with recursive_steps, initial_dims, initial_k, dims, ks).  commit(v, m, n)
    returns (root, rows) of ligero_commit over the vector v (this module's
    layout) as an m x n matrix, rows(q) giving encoded row q as ints.
    Returns the derived challenges and query sets."""
    F = Field(bits)
    p = parse_proof(proof, bits)
    S = 148  # prover.rs:30
    R = shape["recursive_steps"]
    m0, n0 = shape["initial_dims"]
    dims, ks = shape["dims"], shape["ks"]
    _check(len(p["rec_roots"]) == R and len(p["rec_proofs"]) == R - 1, "proof has the wrong number of commitments")
    _check(len(p["tr"]) == 1 + sum(ks) + (R - 1), "sumcheck transcript length")
    log = {"challenges": [], "queries": []}

    def opened(rlp, root, rows, M, q, what):
        _check(rlp["rows"] == [rows(i) for i in q], f"{what}: opened rows differ from the committed rows")
        leaves = [b"".join(F.to_bytes(x) for x in r) for r in rlp["rows"]]
        _check(merkle_verify(root, rlp["nodes"], M.bit_length() - 1, leaves, q), f"{what}: Merkle proof")

    def enforced(rlp, gr, alpha):
        # sum_i alpha^i (row_i . gr) (sumcheck_polys.rs:41-63)
        tot, ap = 0, 1
        for row in rlp["rows"]:
            d = 0
            for x, g in zip(row, gr):
                d ^= F.mul(x, g)
            tot ^= F.mul(d, ap)
            ap = F.mul(ap, alpha)
        return tot

    def basis_factors(ell, q, alpha):
        sk = sks_vks(F, ell)
        ap, out = 1, []
        for qi in q:
            out.append((ap, sks_at(F, sk, qi, ell)))
            ap = F.mul(ap, alpha)
        return out

    def basis_poly(ell, fac):
        b = F.zeros(1 << ell)
        for ap, s in fac:
            b ^= tensor(F, ap, s)
        return b

    fs = Transcript(fs_seed)
    v = F.from_array(poly)
    _check(len(v) == m0 * n0, "input length")
    root0, rows0 = commit(v, m0, n0)
    _check(p["initial_root"] == root0, "initial root differs from the committed input's")
    fs.absorb(p["initial_root"])
    pe0 = [fs.get_field(F) for _ in range(shape["initial_k"])]
    log["challenges"].append(pe0)
    f = partial_eval(F, v, pe0)  # prover.rs:50-53
    root1, rows1 = commit(f, dims[0][0], dims[0][1])
    _check(p["rec_roots"][0] == root1, "first recursive root differs from the folded polynomial's")
    fs.absorb(p["rec_roots"][0])
    q = fs.get_distinct_queries(4 * m0, S)
    log["queries"].append(q)
    opened(p["initial"], root0, rows0, 4 * m0, q, "initial opening")
    alpha = fs.get_field(F)
    ell = (m0).bit_length() - 1
    fac = [basis_factors(ell, q, alpha)]
    seps = [1]
    h = enforced(p["initial"], lagrange_basis(F, pe0), alpha)
    b = basis_poly(ell, fac[0]) if ell <= max_basis_log else None
    tr = iter(p["tr"])
    g = next(tr)
    _check(g[0] ^ g[1] == h, "s0 + s1 != enforced sum (ligerito.rs:35)")
    if b is not None:
        _check(g == eval013(F, f, b), "first triple differs from eval_013_product(f, basis)")
    running = quad_from_evals(F, *g)
    fs.absorb(b"".join(F.to_bytes(x) for x in g))
    ris = []
    rows_prev, root_prev = rows1, root1
    for i in range(R):
        rs = []
        for _ in range(ks[i]):
            r = fs.get_field(F)
            rs.append(r)
            ris.append(r)
            s = quad_eval(F, running, r)
            f = fold(F, f, r)
            if b is not None:
                b = fold(F, b, r)
            g = next(tr)
            _check(g[0] ^ g[1] == s, "s0 + s1 != running sum (ligerito.rs:210)")
            if b is not None:
                _check(g == eval013(F, f, b), "round triple differs from eval_013_product")
            running = quad_from_evals(F, *g)
            fs.absorb(b"".join(F.to_bytes(x) for x in g))
        log["challenges"].append(rs)
        M = 4 * dims[i][0]
        if i == R - 1:
            # prover.rs:104-142 and verifier.rs:79-121
            _check(F.to_ints(f) == p["yr"], "yr differs from the polynomial folded by every challenge")
            fs.absorb(b"".join(F.to_bytes(x) for x in p["yr"]))
            q = fs.get_distinct_queries(M, S)
            log["queries"].append(q)
            opened(p["final"], root_prev, rows_prev, M, q, "final opening")
            # ligero_verify (ligero.rs:97-131)
            gr = lagrange_basis(F, rs)
            nl = len(p["yr"]).bit_length() - 1
            sk = sks_vks(F, nl)
            yr = F.vec(p["yr"])
            for qi, row in zip(q, p["final"]["rows"]):
                d = 0
                for x, gg in zip(row, gr):
                    d ^= F.mul(x, gg)
                e = F.xor_sum(F.vmul(yr, tensor(F, 1, sks_at(F, sk, qi, nl))))
                _check(e == d, f"ligero_verify fails at row {qi}")
            final_r = fs.get_field(F)
            log["final_r"] = final_r
            fpe = fold(F, yr, final_r)
            tot = quad_eval(F, running, final_r)
            # verify_partial: each basis polynomial partially evaluated at the
            # last challenges; a tensor's MSB factor (1, s) folds to
            # (1 + r) + r s, the remaining factors stay
            pts = ris + [final_r]
            acc = F.zeros(len(fpe))
            for sep, fc in zip(seps, fac):
                nv = len(fc[0][1])
                k = nv - (len(fpe).bit_length() - 1)
                use = pts[len(pts) - k:]
                for ap, sv in fc:
                    c = ap
                    for t, r in zip(range(nv - 1, nv - 1 - k, -1), use):
                        c = F.mul(c, (1 ^ r) ^ F.mul(r, sv[t]))
                    acc ^= tensor(F, F.mul(sep, c), sv[:nv - k])
            _check(F.xor_sum(F.vmul(fpe, acc)) == tot, "final sumcheck check (ligerito.rs:282-295)")
            return log
        # prover.rs:145-195
        nxt_root, nxt_rows = commit(f, dims[i + 1][0], dims[i + 1][1])
        _check(p["rec_roots"][i + 1] == nxt_root, f"recursive root {i + 1} differs from the folded polynomial's")
        fs.absorb(p["rec_roots"][i + 1])
        q = fs.get_distinct_queries(M, S)
        log["queries"].append(q)
        opened(p["rec_proofs"][i], root_prev, rows_prev, M, q, f"recursive opening {i}")
        a = fs.get_field(F)
        ell = dims[i][0].bit_length() - 1
        fc = basis_factors(ell, q, a)
        hn = enforced(p["rec_proofs"][i], lagrange_basis(F, rs), a)
        g = next(tr)
        _check(g[0] ^ g[1] == hn, "s0 + s1 != new enforced sum (ligerito.rs:52)")
        bn = basis_poly(ell, fc) if b is not None else None
        if bn is not None:
            _check(g == eval013(F, f, bn), "introduce_new triple differs from eval_013_product(f, new basis)")
        to_glue = quad_from_evals(F, *g)
        fs.absorb(b"".join(F.to_bytes(x) for x in g))
        beta = fs.get_field(F)
        log["challenges"].append([a, beta])
        seps.append(beta)
        fac.append(fc)
        if b is not None:
            b = b ^ F.vmul_s(bn, beta)
        running = tuple(x ^ F.mul(beta, y) for x, y in zip(running, to_glue))  # fold_quadratic (sumcheck.rs:83-94)
        rows_prev, root_prev = nxt_rows, nxt_root
    raise ReplayError("no final step (prover.rs:198)")
