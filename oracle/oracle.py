"""ORACLE — test infrastructure only.

ctypes binding of the C restatement of the reference (oracle/src).  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
may import this module, and only as the checker: the product path never
calls it.

Elements cross the binding as numpy arrays: uint16/uint32 for GF(2^16) /
GF(2^32), and (n, 2) uint64 little-endian word pairs for GF(2^128).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        SZ = C.c_size_t
        I = C.c_int
        sig = {
            "oracle_gf_mul_batch": (I, [I, P, P, P, SZ]),
            "oracle_gf_inverse": (I, [I, P, P]),
            "oracle_gf_pow": (I, [I, P, P, P]),
            "oracle_compute_twiddles": (I, [I, I, P]),
            "oracle_fft": (I, [I, P, SZ, P, I]),
            "oracle_eval_sk_at_vks": (I, [I, SZ, P]),
            "oracle_evaluate_basis": (I, [I, SZ, P, I, P, P]),
            "oracle_compute_pis": (I, [I, SZ, P, I, P]),
            "oracle_rs_encode": (I, [I, SZ, SZ, P, P, I]),
            "oracle_lagrange_basis": (I, [I, P, I, P]),
            "oracle_partial_eval": (I, [I, P, SZ, P, I, P]),
            "oracle_eval_013_product": (I, [I, P, P, SZ, P]),
            "oracle_induce": (I, [I, I, P, SZ, SZ, P, I, P, P, P, P]),
            "oracle_blake3": (None, [P, SZ, P]),
            "oracle_blake3_pieces": (None, [P, P, SZ, P]),
            "oracle_sha256": (None, [P, SZ, P]),
            "oracle_fs_new": (P, [C.c_uint32]),
            "oracle_fs_free": (None, [P]),
            "oracle_fs_absorb": (None, [P, P, SZ]),
            "oracle_fs_squeeze": (None, [P, P]),
            "oracle_fs_get_query": (SZ, [P, SZ]),
            "oracle_fs_get_distinct_queries": (I, [P, SZ, SZ, P]),
            "oracle_fs_counter": (C.c_uint32, [P]),
            "oracle_merkle_build": (I, [P, P, SZ, P, I]),
            "oracle_merkle_prove": (C.c_long, [P, P, SZ, P, SZ, P, SZ, I]),
            "oracle_merkle_verify": (I, [P, P, SZ, I, P, P, P, SZ, I]),
            "oracle_ligero_commit": (I, [I, P, SZ, SZ, P, P, P]),
            "oracle_prove": (I, [I, I, P, C.c_uint32, P, SZ, C.POINTER(SZ)]),
            "oracle_verify": (I, [I, I, P, SZ, C.c_uint32]),
            "oracle_splitmix_fill": (None, [C.c_uint64, I, SZ, P]),
            "oracle_set_threads": (None, [I]),
            "oracle_prove_timed": (I, [I, I, P, C.c_uint32, C.POINTER(C.c_double)]),
            "oracle_get_threads": (I, []),
            "oracle_config_shape": (I, [I, C.POINTER(SZ), C.POINTER(SZ), C.POINTER(I), P, P]),
            "oracle_prove_dims": (I, [I, I, SZ, SZ, I, P, P, P, C.c_uint32, P, SZ, C.POINTER(SZ)]),
            "oracle_verify_dims": (I, [I, I, I, P, I, P, P, SZ, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------- helpers --

def _dtype(bits):
    return {16: np.uint16, 32: np.uint32, 128: np.uint64}[bits]


def as_elems(x, bits):
    """Returns a C-contiguous numpy array in the element layout for `bits`."""
    if bits == 128:
        a = np.ascontiguousarray(x, dtype=np.uint64)
        return a.reshape(-1, 2)
    return np.ascontiguousarray(x, dtype=_dtype(bits))


def empty(n, bits):
    if bits == 128:
        return np.zeros((n, 2), dtype=np.uint64)
    return np.zeros(n, dtype=_dtype(bits))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def int_to_elem(v: int, bits: int):
    if bits == 128:
        return np.array([[v & (2**64 - 1), v >> 64]], dtype=np.uint64)
    return np.array([v], dtype=_dtype(bits))


def elem_to_int(a, bits) -> int:
    a = np.asarray(a)
    if bits == 128:
        a = a.reshape(-1, 2)
        return int(a[0, 0]) | (int(a[0, 1]) << 64)
    return int(a.reshape(-1)[0])


def splitmix(seed: int, n: int, bits: int):
    out = empty(n, bits)
    lib().oracle_splitmix_fill(seed, bits, n, _p(out))
    return out


# ------------------------------------------------------------ field / fft --

def mul(a, b, bits):
    a = as_elems(a, bits)
    b = as_elems(b, bits)
    n = a.shape[0]
    out = empty(n, bits)
    lib().oracle_gf_mul_batch(bits, _p(a), _p(b), _p(out), n)
    return out


def inverse(a: int, bits: int):
    x = int_to_elem(a, bits)
    out = empty(1, bits)
    rc = lib().oracle_gf_inverse(bits, _p(x), _p(out))
    if rc != 0:
        raise ZeroDivisionError("Cannot compute inverse of zero")
    return elem_to_int(out, bits)


def gf_pow(a: int, e: int, bits: int) -> int:
    x = int_to_elem(a, bits)
    ee = int_to_elem(e, bits)
    out = empty(1, bits)
    lib().oracle_gf_pow(bits, _p(x), _p(ee), _p(out))
    return elem_to_int(out, bits)


def compute_twiddles(log_n: int, bits: int):
    out = empty((1 << log_n) - 1, bits)
    lib().oracle_compute_twiddles(bits, log_n, _p(out))
    return out


def fft(v, tw, bits, inverse=False):
    v = as_elems(v, bits).copy()
    tw = as_elems(tw, bits)
    lib().oracle_fft(bits, _p(v), v.shape[0], _p(tw), 1 if inverse else 0)
    return v


def eval_sk_at_vks(n: int, bits: int):
    ns = n.bit_length() - 1
    out = empty(ns + 1, bits)
    lib().oracle_eval_sk_at_vks(bits, n, _p(out))
    return out


def evaluate_basis(basis_len: int, sks, x: int, bits: int):
    sks = as_elems(sks, bits)
    xx = int_to_elem(x, bits)
    out = empty(basis_len, bits)
    lib().oracle_evaluate_basis(bits, basis_len, _p(sks), sks.shape[0], _p(xx), _p(out))
    return out


def compute_pis(pis_len: int, sks, bits: int):
    sks = as_elems(sks, bits)
    out = empty(pis_len, bits)
    rc = lib().oracle_compute_pis(bits, pis_len, _p(sks), sks.shape[0], _p(out))
    if rc:
        raise ValueError("pis_len must be 2^(len(sks)-1)")
    return out


def rs_encode(msg, block_len: int, bits: int):
    msg = as_elems(msg, bits)
    out = empty(block_len, bits)
    rc = lib().oracle_rs_encode(bits, msg.shape[0], block_len, _p(msg), _p(out), 0)
    if rc:
        raise ValueError("bad RS parameters")
    return out


def rs_encode_non_systematic(buf, msg_len: int, bits: int):
    buf = as_elems(buf, bits).copy()
    rc = lib().oracle_rs_encode(bits, msg_len, buf.shape[0], None, _p(buf), 1)
    if rc:
        raise ValueError("bad RS parameters")
    return buf


def lagrange_basis(rs, bits):
    rs = as_elems(rs, bits)
    k = rs.shape[0]
    out = empty(1 << k, bits)
    lib().oracle_lagrange_basis(bits, _p(rs), k, _p(out))
    return out


def partial_eval(evals, rs, bits):
    e = as_elems(evals, bits)
    r = as_elems(rs, bits)
    out = empty(e.shape[0] >> r.shape[0], bits)
    lib().oracle_partial_eval(bits, _p(e), e.shape[0], _p(r), r.shape[0], _p(out))
    return out


def eval_013_product(f, g, bits):
    f = as_elems(f, bits)
    g = as_elems(g, bits)
    out = empty(3, bits)
    lib().oracle_eval_013_product(bits, _p(f), _p(g), f.shape[0], _p(out))
    return out


def induce(n, rows, v_ch, queries, alpha: int, bits):
    rows = as_elems(rows, bits)
    v = as_elems(v_ch, bits)
    q = np.ascontiguousarray(queries, dtype=np.uint64)
    n_rows = q.shape[0]
    n_cols = (rows.shape[0] // n_rows)
    al = int_to_elem(alpha, bits)
    basis = empty(1 << n, bits)
    enf = empty(1, bits)
    rc = lib().oracle_induce(bits, n, _p(rows), n_rows, n_cols, _p(v), v.shape[0], _p(q), _p(al),
                             _p(basis), _p(enf))
    if rc:
        raise ValueError("row length must be 2^len(v_challenges)")
    return basis, elem_to_int(enf, bits)


# ------------------------------------------------------- hashes / FS -------

def blake3(data: bytes) -> bytes:
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(bytes(data), max(1, len(data)))
    lib().oracle_blake3(buf, len(data), out)
    return bytes(out)


def blake3_pieces(data: bytes, pieces) -> bytes:
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(bytes(data), max(1, len(data)))
    arr = (C.c_size_t * len(pieces))(*pieces)
    lib().oracle_blake3_pieces(buf, arr, len(pieces), out)
    return bytes(out)


def sha256(data: bytes) -> bytes:
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(bytes(data), max(1, len(data)))
    lib().oracle_sha256(buf, len(data), out)
    return bytes(out)


class FS:
    """src/emulated_fs.rs:6-75"""

    def __init__(self, seed: int):
        self._h = lib().oracle_fs_new(seed)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_fs_free(self._h)
            self._h = None

    def absorb(self, data: bytes):
        buf = C.create_string_buffer(bytes(data), max(1, len(data)))
        lib().oracle_fs_absorb(self._h, buf, len(data))

    def squeeze(self) -> int:
        out = (C.c_uint8 * 16)()
        lib().oracle_fs_squeeze(self._h, out)
        return int.from_bytes(bytes(out), "little")

    def get_field(self, bits: int) -> int:
        return self.squeeze() & ((1 << bits) - 1)

    def get_query(self, n: int) -> int:
        return int(lib().oracle_fs_get_query(self._h, n))

    def get_distinct_queries(self, n: int, s: int):
        out = np.zeros(s, dtype=np.uint64)
        rc = lib().oracle_fs_get_distinct_queries(self._h, n, s, _p(out))
        if rc:
            raise ValueError("s > n: the reference would never terminate")
        return [int(x) for x in out]

    @property
    def counter(self) -> int:
        return int(lib().oracle_fs_counter(self._h))


# ----------------------------------------------------------------- Merkle --

def _leaves(leaves):
    bs = [bytes(x) for x in leaves]
    offs = np.zeros(len(bs) + 1, dtype=np.uint64)
    for i, b in enumerate(bs):
        offs[i + 1] = offs[i] + len(b)
    data = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8).copy()
    return data, offs


HASH_BLAKE3 = 0
HASH_SHA256 = 1  # only for the Julia-derived structure vectors (merkle_tree.rs:195-255)


def build_merkle_tree(leaves, hash_kind=HASH_BLAKE3):
    """src/merkle_tree.rs:22-60 -> list of layers (bytes), leaves first."""
    n = len(leaves)
    data, offs = _leaves(leaves)
    out = np.zeros(32 * (2 * n - 1), dtype=np.uint8)
    nl = lib().oracle_merkle_build(_p(data), _p(offs), n, _p(out), hash_kind)
    if nl < 0:
        raise ValueError("leaf count must be a power of two")
    layers, pos = [], 0
    for d in range(nl):
        sz = 32 * (n >> d)
        layers.append(out[pos:pos + sz].tobytes())
        pos += sz
    return layers


def merkle_prove(leaves, queries, hash_kind=HASH_BLAKE3):
    n = len(leaves)
    data, offs = _leaves(leaves)
    q = np.ascontiguousarray(queries, dtype=np.uint64)
    cap = max(1, len(queries)) * 64 + 1
    out = np.zeros(32 * cap, dtype=np.uint8)
    r = lib().oracle_merkle_prove(_p(data), _p(offs), n, _p(q), q.shape[0], _p(out), cap, hash_kind)
    if r < 0:
        raise ValueError("merkle prove failed")
    return [out[32 * i:32 * i + 32].tobytes() for i in range(r)]


def merkle_verify(root: bytes, proof, depth: int, leaves, indices, hash_kind=HASH_BLAKE3) -> bool:
    data, offs = _leaves(leaves)
    pr = np.frombuffer(b"".join(proof) + b"\0", dtype=np.uint8).copy()
    rt = np.frombuffer(bytes(root), dtype=np.uint8).copy()
    idx = np.ascontiguousarray(indices, dtype=np.uint64)
    return bool(lib().oracle_merkle_verify(_p(rt), _p(pr), len(proof), depth, _p(data), _p(offs), _p(idx),
                                           idx.shape[0], hash_kind))


# ---------------------------------------------------- commit / prove / verify

def ligero_commit(poly, m: int, n: int, bits: int, want_matrix=False, want_layers=False):
    """src/ligero.rs:70-95.  Returns (root, column-major matrix | None, layers | None)."""
    p = as_elems(poly, bits)
    root = (C.c_uint8 * 32)()
    mat = empty(4 * m * n, bits) if want_matrix else None
    M = 4 * m
    lay = np.zeros(32 * (2 * M - 1), dtype=np.uint8) if want_layers else None
    rc = lib().oracle_ligero_commit(bits, _p(p), m, n, root, _p(mat) if mat is not None else None,
                                    _p(lay) if lay is not None else None)
    if rc:
        raise ValueError(f"ligero_commit failed: {rc}")
    layers = None
    if lay is not None:
        layers, pos = [], 0
        d = 0
        while (M >> d) >= 1:
            sz = 32 * (M >> d)
            layers.append(lay[pos:pos + sz].tobytes())
            pos += sz
            d += 1
    return bytes(root), mat, layers


def prove(poly, L: int, bits: int = 32, fs_seed: int = 1234) -> bytes:
    """src/prover.rs:15-199 on the generated config for 2^L (cfg20/cfg24 at L=20/24)."""
    p = as_elems(poly, bits)
    ln = C.c_size_t(0)
    out = np.zeros(16 << 20, dtype=np.uint8)
    rc = lib().oracle_prove(bits, L, _p(p), fs_seed, _p(out), out.size, C.byref(ln))
    if rc == -11:  # larger than the first buffer: prove again into one of the right size
        out = np.zeros(ln.value, dtype=np.uint8)
        rc = lib().oracle_prove(bits, L, _p(p), fs_seed, _p(out), out.size, C.byref(ln))
    if rc:
        raise RuntimeError(f"oracle prover failed: {rc}")
    return out[: ln.value].tobytes()


def prove_timed(poly, L: int, bits: int = 32, fs_seed: int = 1234, threads: int = 0) -> float:
    """Seconds of prover() alone (config built before the clock, as
    src/main.rs:29-49 times it), on `threads` OpenMP threads (0 = default)."""
    p = as_elems(poly, bits)
    if threads:
        lib().oracle_set_threads(threads)
    t = C.c_double(0)
    rc = lib().oracle_prove_timed(bits, L, _p(p), fs_seed, C.byref(t))
    if rc:
        raise RuntimeError(f"oracle prover failed: {rc}")
    return t.value


def verify(proof: bytes, L: int, bits: int = 32, fs_seed: int = 1234) -> int:
    """src/verifier.rs:10-164.  1 accept, 0 reject, negative = reference panic site."""
    buf = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    return int(lib().oracle_verify(bits, L, _p(buf), buf.shape[0], fs_seed))


def config_shape(L: int):
    """The generated configuration for 2^L as a dict with the fields of
    ProverConfig (data_structures.rs:9-21): recursive_steps, initial_dims,
    initial_k, dims, ks."""
    im, inn, ik = C.c_size_t(0), C.c_size_t(0), C.c_int(0)
    dims = np.zeros(16, dtype=np.uint64)
    ks = np.zeros(8, dtype=np.int32)
    r = lib().oracle_config_shape(L, C.byref(im), C.byref(inn), C.byref(ik), _p(dims), _p(ks))
    if r < 0:
        raise RuntimeError(f"unsupported log size {L}")
    return {"recursive_steps": r, "initial_dims": (im.value, inn.value), "initial_k": ik.value,
            "dims": [(int(dims[2 * i]), int(dims[2 * i + 1])) for i in range(r)], "ks": [int(k) for k in ks[:r]]}


def _dims_args(shape):
    r = shape["recursive_steps"]
    dims = np.array([x for d in shape["dims"] for x in d], dtype=np.uint64)
    ks = np.array(shape["ks"], dtype=np.int32)
    assert dims.shape[0] == 2 * r and ks.shape[0] == r
    return r, dims, ks


def prove_dims(poly, shape, bits: int = 32, fs_seed: int = 1234) -> bytes:
    """prover.rs:15-199 on an explicit configuration (a dict as config_shape returns)."""
    p = as_elems(poly, bits)
    r, dims, ks = _dims_args(shape)
    im, inn = shape["initial_dims"]
    ln = C.c_size_t(0)
    out = np.zeros(16 << 20, dtype=np.uint8)
    args = (bits, r, im, inn, shape["initial_k"], _p(dims), _p(ks), _p(p), fs_seed)
    rc = lib().oracle_prove_dims(*args, _p(out), out.size, C.byref(ln))
    if rc == -11:
        out = np.zeros(ln.value, dtype=np.uint8)
        rc = lib().oracle_prove_dims(*args, _p(out), out.size, C.byref(ln))
    if rc:
        raise RuntimeError(f"oracle prover failed: {rc}")
    return out[: ln.value].tobytes()


def verify_dims(proof: bytes, shape, bits: int = 32, fs_seed: int = 1234) -> int:
    """verifier.rs:10-164 on the VerifierConfig of an explicit configuration."""
    buf = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    r, dims, ks = _dims_args(shape)
    log_dims = np.array([int(d[0]).bit_length() - 1 for d in shape["dims"]], dtype=np.int32)
    init_dim = int(shape["initial_dims"][0]).bit_length() - 1
    return int(lib().oracle_verify_dims(bits, r, init_dim, _p(log_dims), shape["initial_k"], _p(ks), _p(buf),
                                        buf.shape[0], fs_seed))
