"""Python binding of libligerito_hip.so (include/ligerito.h) mirroring the
reference crate's API names (ashutosh1206/ligerito-rust, src/lib.rs:17-31).

The library is the HIP implementation; if it is missing this module raises
instead of falling back to anything else.  Field elements are numpy arrays:
uint32 for GF(2^32), (n, 2) uint64 (low word first) for GF(2^128).
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libligerito_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ligerito.h")

LIG_OK = 0
ERRORS = {-1: "LIG_E_INVALID", -2: "LIG_E_ASSERT", -3: "LIG_E_HIP", -4: "LIG_E_OOM", -5: "LIG_E_NO_DEVICE",
          -6: "LIG_E_MALFORMED", -7: "LIG_E_CAPACITY", -8: "LIG_E_MERKLE", -9: "LIG_E_LIGERO"}

_lib = None
# The loaded library, also held by _Handle._release's default argument so
# that handles can be released while the interpreter shuts down (module
# globals are cleared before the last objects are collected).
_holder = [None]


class LigeritoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def header_symbols():
    """Every function the C header declares."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lig_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, SZ, I, U32 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint32
        PP = C.POINTER(C.c_void_p)
        sig = {
            "lig_last_error": (C.c_char_p, []),
            "lig_device_count": (I, []),
            "lig_ctx_create": (I, [I, PP]),
            "lig_ctx_destroy": (None, [P]),
            "lig_ctx_stream": (P, [P]),
            "lig_ctx_profile": (I, [P, I]),
            "lig_ctx_profile_read": (I, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
            "lig_ctx_synchronize": (I, [P]),
            "lig_bench_rate": (I, [P, I, C.POINTER(C.c_double)]),
            "lig_rs_new": (I, [P, I, SZ, SZ, PP]),
            "lig_rs_free": (None, [P]),
            "lig_rs_message_length": (SZ, [P]),
            "lig_rs_block_length": (SZ, [P]),
            "lig_rs_encode": (I, [P, P, P]),
            "lig_rs_encode_non_systematic": (I, [P, P]),
            "lig_rs_encode_non_systematic_cols_dev": (I, [P, P, SZ]),
            "lig_rs_encode_non_systematic_cols_dev_noalloc": (I, [P, P, SZ]),
            "lig_rs_twiddles": (I, [P, P]),
            "lig_rs_autotune": (I, [P, SZ, C.POINTER(I), C.POINTER(I)]),
            "lig_rs_set_tuning": (I, [P, SZ, I, I]),
            "lig_build_merkle_tree": (I, [P, P, SZ, SZ, P]),
            "lig_merkle_prove": (I, [P, SZ, P, SZ, P, SZ, C.POINTER(SZ)]),
            "lig_merkle_verify": (I, [P, P, SZ, SZ, P, SZ, P, SZ]),
            "lig_build_merkle_tree_var": (I, [P, P, P, SZ, P]),
            "lig_merkle_verify_var": (I, [P, P, SZ, SZ, P, P, P, SZ]),
            "lig_fs_new": (I, [U32, PP]),
            "lig_fs_free": (None, [P]),
            "lig_fs_absorb": (I, [P, P, SZ]),
            "lig_fs_squeeze": (I, [P, P]),
            "lig_fs_get_field": (I, [P, I, P]),
            "lig_fs_get_query": (SZ, [P, SZ]),
            "lig_fs_get_distinct_queries": (I, [P, SZ, SZ, P]),
            "lig_ligero_commit": (I, [P, P, SZ, SZ, PP]),
            "lig_ligero_commit_dev": (I, [P, P, SZ, SZ, PP]),
            "lig_witness_root": (I, [P, P]),
            "lig_witness_num_rows": (SZ, [P]),
            "lig_witness_num_cols": (SZ, [P]),
            "lig_witness_flat_mat": (I, [P, P]),
            "lig_witness_layers": (I, [P, P]),
            "lig_witness_extract_row": (I, [P, SZ, P]),
            "lig_witness_free": (None, [P]),
            "lig_prover_config_new": (I, [P, I, I, PP]),
            "lig_prover_config_free": (None, [P]),
            "lig_prover_config_from_dims": (I, [P, I, SZ, P, SZ, P, P, PP]),
            "lig_prover_config_shape": (I, [I, C.POINTER(SZ), P, C.POINTER(SZ), P, P, SZ]),
            "lig_prover_config_dims": (I, [P, C.POINTER(SZ), P, C.POINTER(SZ), P, P, SZ]),
            "lig_verifier_with_config": (I, [I, SZ, SZ, P, SZ, P, P, U32]),
            "lig_verifier_with_config_dev": (I, [P, I, SZ, SZ, P, SZ, P, P, U32]),
            "lig_prover": (I, [P, P, U32, PP]),
            "lig_prover_dev": (I, [P, P, U32, PP]),
            "lig_proof_serialize": (I, [P, P, SZ, C.POINTER(SZ)]),
            "lig_proof_deserialize": (I, [I, P, SZ, PP]),
            "lig_proof_size": (SZ, [P]),
            "lig_proof_free": (None, [P]),
            "lig_verifier": (I, [I, I, P, U32]),
            "lig_verifier_dev": (I, [P, I, I, P, U32]),
            "lig_batch_create": (I, [I, I, I, PP]),
            "lig_batch_destroy": (None, [P]),
            "lig_prove_batch": (I, [P, SZ, C.POINTER(C.c_int), PP, I, C.POINTER(U32), PP]),
            "lig_serve_start": (I, [I, I, I, C.c_uint64, PP]),
            "lig_serve_submit": (I, [P, I, P, U32, C.POINTER(C.c_uint64)]),
            "lig_serve_poll": (I, [P, C.c_uint64, PP, C.POINTER(C.c_double)]),
            "lig_serve_wait": (I, [P, C.c_uint64, PP, C.POINTER(C.c_double)]),
            "lig_serve_warm": (I, [P, C.POINTER(C.c_int), SZ, C.POINTER(SZ)]),
            "lig_serve_stop": (None, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        _holder[0] = L
    return _lib


def _check(rc):
    if rc < 0:
        raise LigeritoError(rc, lib().lig_last_error().decode())
    return rc


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def as_elems(x, bits):
    if bits == 128:
        return np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
    return np.ascontiguousarray(x, dtype=np.uint32)


def empty(n, bits):
    return np.zeros((n, 2), dtype=np.uint64) if bits == 128 else np.zeros(n, dtype=np.uint32)


def nelems(a, bits):
    return a.shape[0]


# ------------------------------------------------------------------ context --
class _Handle:
    """Base of the wrappers that own a library handle (self._h)."""

    _h = None

    def _release(self, fn_name, _holder=_holder):
        h, self._h = self._h, None
        if h and _holder[0] is not None:
            getattr(_holder[0], fn_name)(h)


class Context(_Handle):
    """Device, stream and caches (lig_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().lig_ctx_create(device, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return lib().lig_ctx_stream(self._h)

    def synchronize(self):
        _check(lib().lig_ctx_synchronize(self._h))

    def profile(self, enable: bool = True):
        """Enables per-phase HIP-event timing on the library stream (resets totals)."""
        _check(lib().lig_ctx_profile(self._h, 1 if enable else 0))

    def bench_rate(self, which: str) -> float:
        """Microbenchmarked peak: "mul_tab" (GF(2^32) products/s, LDS nibble
        tables), "mul_clmul" (general GF(2^32) products/s), "valu" (32-bit
        lane-ops/s), "mul_tab5" (GF(2^32) products/s, 5-bit-window tables),
        "lds128" (LDS bytes/s, 16-byte table reads), "mul_mask" (GF(2^32)
        products/s, bit masks over a constant's multiples in registers) or
        "gf8_logexp" (GF(2^8) products/s, log/exp tables in LDS)."""
        r = C.c_double(0)
        idx = {"mul_tab": 0, "mul_clmul": 1, "valu": 2, "mul_tab5": 3, "lds128": 4, "mul_mask": 5, "gf8_logexp": 6}[which]
        _check(lib().lig_bench_rate(self._h, idx, C.byref(r)))
        return r.value

    def profile_read(self, phase: str):
        """(total_ms, count) of a phase since profiling was enabled."""
        ms, n = C.c_double(0), C.c_uint64(0)
        _check(lib().lig_ctx_profile_read(self._h, phase.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self._h:
            self._release("lig_ctx_destroy")

    def __del__(self):
        self.close()


# --------------------------------------------------------------- Reed–Solomon --
class ReedSolomonEncoding(_Handle):
    """src/reed_solomon.rs:7-114"""

    def __init__(self, ctx: Context, message_length: int, block_length: int, bits: int = 32):
        h = C.c_void_p()
        _check(lib().lig_rs_new(ctx.handle, bits, message_length, block_length, C.byref(h)))
        self._h, self.bits, self.ctx = h, bits, ctx

    def __del__(self):
        if self._h:
            self._release("lig_rs_free")

    def message_length(self):
        return lib().lig_rs_message_length(self._h)

    def block_length(self):
        return lib().lig_rs_block_length(self._h)

    def autotune(self, n_cols: int):
        """(kmax, group) of the encode's FFT passes for n_cols columns,
        timing the candidates now if no commit of n_cols columns did."""
        k, g = C.c_int(0), C.c_int(0)
        _check(lib().lig_rs_autotune(self._h, n_cols, C.byref(k), C.byref(g)))
        return k.value, g.value

    def set_tuning(self, n_cols: int, kmax: int, group: int):
        """Fixes the launch shape of the FFT passes for n_cols columns."""
        _check(lib().lig_rs_set_tuning(self._h, n_cols, kmax, group))

    def twiddles(self):
        out = empty(self.block_length() - 1, self.bits)
        _check(lib().lig_rs_twiddles(self._h, _p(out)))
        return out

    def encode(self, message):
        m = as_elems(message, self.bits)
        if m.shape[0] != self.message_length():
            raise LigeritoError(-1, "message length mismatch (reed_solomon.rs:91)")
        out = empty(self.block_length(), self.bits)
        _check(lib().lig_rs_encode(self._h, _p(m), _p(out)))
        return out

    def encode_non_systematic_cols_dev(self, d_columns: int, n_cols: int):
        """encode_non_systematic in place on n_cols device-resident columns
        (column-major, block_length elements each), as encode_cols
        (ligero.rs:29-38) applies it."""
        _check(lib().lig_rs_encode_non_systematic_cols_dev(self._h, C.c_void_p(d_columns), n_cols))

    def encode_non_systematic_cols_dev_noalloc(self, d_columns: int, n_cols: int):
        """The same with no device allocation (a static 64 MiB scratch per
        GPU; calls that overlap on one GPU race)."""
        _check(lib().lig_rs_encode_non_systematic_cols_dev_noalloc(self._h, C.c_void_p(d_columns), n_cols))

    def encode_non_systematic(self, message):
        """Returns the encoded block (the reference encodes in place)."""
        m = as_elems(message, self.bits).copy()
        if m.shape[0] != self.block_length():
            raise LigeritoError(-1, "block length mismatch (reed_solomon.rs:105)")
        _check(lib().lig_rs_encode_non_systematic(self._h, _p(m)))
        return m


# ------------------------------------------------------------------- Merkle --
def _var_leaves(leaves):
    """A list of byte strings as one byte array plus n + 1 offsets."""
    parts = [bytes(x) for x in leaves]
    offs = np.zeros(len(parts) + 1, dtype=np.uint64)
    np.cumsum([len(x) for x in parts], out=offs[1:])
    buf = np.frombuffer(b"".join(parts) + b"\0", dtype=np.uint8)
    return buf, offs


def build_merkle_tree(ctx: Context, leaves):
    """src/merkle_tree.rs:22-60.  leaves: rows of a 2-D uint8 array (equal
    lengths, a multiple of 4 bytes) or a list of byte strings of any lengths.
    Returns the list of layers as bytes, leaves first."""
    if isinstance(leaves, np.ndarray) and leaves.ndim == 2 and leaves.shape[1] % 4 == 0:
        leaves = np.ascontiguousarray(leaves, dtype=np.uint8)
        n, ll = leaves.shape
        out = np.zeros(32 * (2 * n - 1), dtype=np.uint8)
        _check(lib().lig_build_merkle_tree(ctx.handle, _p(leaves), n, ll, _p(out)))
    else:
        buf, offs = _var_leaves(leaves)
        n = offs.shape[0] - 1
        out = np.zeros(32 * max(2 * n - 1, 1), dtype=np.uint8)
        _check(lib().lig_build_merkle_tree_var(ctx.handle, _p(buf), _p(offs), n, _p(out)))
    layers, pos, d = [], 0, 0
    while (n >> d) >= 1:
        sz = 32 * (n >> d)
        layers.append(out[pos:pos + sz].tobytes())
        pos += sz
        d += 1
    return layers


def prove(tree_layers, queries):
    """src/merkle_tree.rs:98-111"""
    flat = np.frombuffer(b"".join(tree_layers), dtype=np.uint8)
    n = len(tree_layers[0]) // 32
    q = np.ascontiguousarray(queries, dtype=np.uint64)
    cnt = C.c_size_t(0)
    _check(lib().lig_merkle_prove(_p(flat), n, _p(q), q.shape[0], None, 0, C.byref(cnt)))
    out = np.zeros(32 * max(cnt.value, 1), dtype=np.uint8)
    _check(lib().lig_merkle_prove(_p(flat), n, _p(q), q.shape[0], _p(out), cnt.value, C.byref(cnt)))
    return [out[32 * i:32 * i + 32].tobytes() for i in range(cnt.value)]


def verify(root: bytes, proof, depth: int, leaves, indices) -> bool:
    """src/merkle_tree.rs:164-188.  leaves: rows of a 2-D uint8 array or a
    list of byte strings of any lengths."""
    pr = np.frombuffer(b"".join(proof) + b"\0", dtype=np.uint8)
    rt = np.frombuffer(bytes(root), dtype=np.uint8)
    idx = np.ascontiguousarray(indices, dtype=np.uint64)
    if isinstance(leaves, np.ndarray) and leaves.ndim == 2:
        lv = np.ascontiguousarray(leaves, dtype=np.uint8)
        return bool(lib().lig_merkle_verify(_p(rt), _p(pr), len(proof), depth, _p(lv), lv.shape[1], _p(idx),
                                            idx.shape[0]))
    buf, offs = _var_leaves(leaves)
    return bool(lib().lig_merkle_verify_var(_p(rt), _p(pr), len(proof), depth, _p(buf), _p(offs), _p(idx), idx.shape[0]))


# --------------------------------------------------------------- transcript --
class FS(_Handle):
    """src/emulated_fs.rs:6-75"""

    def __init__(self, seed: int):
        h = C.c_void_p()
        _check(lib().lig_fs_new(seed, C.byref(h)))
        self._h = h

    def __del__(self):
        if self._h:
            self._release("lig_fs_free")

    def absorb(self, data: bytes):
        b = bytes(data)
        buf = C.create_string_buffer(b, max(1, len(b)))
        _check(lib().lig_fs_absorb(self._h, buf, len(b)))

    def squeeze(self) -> int:
        out = (C.c_uint8 * 16)()
        _check(lib().lig_fs_squeeze(self._h, out))
        return int.from_bytes(bytes(out), "little")

    def get_field(self, bits: int = 32) -> int:
        out = (C.c_uint8 * (bits // 8))()
        _check(lib().lig_fs_get_field(self._h, bits, out))
        return int.from_bytes(bytes(out), "little")

    def get_query(self, n: int) -> int:
        return int(lib().lig_fs_get_query(self._h, n))

    def get_distinct_queries(self, n: int, s: int):
        out = np.zeros(s, dtype=np.uint64)
        _check(lib().lig_fs_get_distinct_queries(self._h, n, s, _p(out)))
        return [int(x) for x in out]


# ------------------------------------------------------------ ligero_commit --
class Witness(_Handle):
    """RecursiveLigeroWitness (data_structures.rs:32-40), device-resident.
    Keeps its plan (and through it the context) referenced."""

    def __init__(self, h, bits, rs=None):
        self._h, self.bits, self.rs = h, bits, rs

    def __del__(self):
        if self._h:
            self._release("lig_witness_free")

    @property
    def num_rows(self):
        return lib().lig_witness_num_rows(self._h)

    @property
    def num_cols(self):
        return lib().lig_witness_num_cols(self._h)

    def root(self) -> bytes:
        out = (C.c_uint8 * 32)()
        _check(lib().lig_witness_root(self._h, out))
        return bytes(out)

    def flat_mat(self):
        out = empty(self.num_rows * self.num_cols, self.bits)
        _check(lib().lig_witness_flat_mat(self._h, _p(out)))
        return out

    def tree(self):
        n = self.num_rows
        out = np.zeros(32 * (2 * n - 1), dtype=np.uint8)
        _check(lib().lig_witness_layers(self._h, _p(out)))
        layers, pos, d = [], 0, 0
        while (n >> d) >= 1:
            sz = 32 * (n >> d)
            layers.append(out[pos:pos + sz].tobytes())
            pos += sz
            d += 1
        return layers

    def extract_row(self, row):
        out = empty(self.num_cols, self.bits)
        _check(lib().lig_witness_extract_row(self._h, row, _p(out)))
        return out


def _require_len(a, n, what):
    """The C entry points copy exactly n elements from the pointer: a shorter
    array would be read past its end (the reference panics on a wrong-length
    input instead)."""
    if a.shape[0] != n:
        raise LigeritoError(-1, f"{what}: {a.shape[0]} elements given, {n} required")


def ligero_commit(poly, m: int, n: int, rs: ReedSolomonEncoding) -> Witness:
    """src/ligero.rs:70-95"""
    p = as_elems(poly, rs.bits)
    _require_len(p, m * n, "ligero_commit (ligero.rs:11-27 reads m*n coefficients)")
    h = C.c_void_p()
    _check(lib().lig_ligero_commit(rs._h, _p(p), m, n, C.byref(h)))
    return Witness(h, rs.bits, rs)


def ligero_commit_dev(d_poly_ptr: int, m: int, n: int, rs: ReedSolomonEncoding) -> Witness:
    h = C.c_void_p()
    _check(lib().lig_ligero_commit_dev(rs._h, C.c_void_p(d_poly_ptr), m, n, C.byref(h)))
    return Witness(h, rs.bits, rs)


# ------------------------------------------------------------ prove / verify --
class ProverConfig(_Handle):
    """ProverConfig (data_structures.rs:9-21); hardcoded_config_20/24 at log_n 20/24."""

    def __init__(self, ctx: Context, log_n: int, bits: int = 32):
        h = C.c_void_p()
        _check(lib().lig_prover_config_new(ctx.handle, bits, log_n, C.byref(h)))
        self._h, self.bits, self.log_n, self.ctx = h, bits, log_n, ctx
        self.num_coeffs = 1 << log_n

    def __del__(self):
        if self._h:
            self._release("lig_prover_config_free")


def _read_shape(fn, *lead):
    r, ik = C.c_size_t(0), C.c_size_t(0)
    init = np.zeros(2, dtype=np.uint64)
    dims = np.zeros(16, dtype=np.uint64)
    ks = np.zeros(8, dtype=np.uint64)
    _check(fn(*lead, C.byref(r), _p(init), C.byref(ik), _p(dims), _p(ks), 8))
    n = r.value
    return {"recursive_steps": n, "initial_dims": (int(init[0]), int(init[1])), "initial_k": ik.value,
            "dims": [(int(dims[2 * i]), int(dims[2 * i + 1])) for i in range(n)], "ks": [int(k) for k in ks[:n]]}


def config_shape(log_n: int):
    """The generated configuration for 2^log_n coefficients (no device):
    a dict with ProverConfig's fields (data_structures.rs:9-21)."""
    return _read_shape(lib().lig_prover_config_shape, log_n)


class ProverConfigFromDims(ProverConfig):
    """ProverConfig from explicit dimensions (data_structures.rs:9-21), a dict
    with the keys config_shape returns."""

    def __init__(self, ctx: Context, shape, bits: int = 32):
        r = int(shape["recursive_steps"])
        init = np.array(shape["initial_dims"], dtype=np.uint64)
        dims = np.array([x for d in shape["dims"] for x in d] or [0], dtype=np.uint64)
        ks = np.array(shape["ks"] or [0], dtype=np.uint64)
        if len(shape["dims"]) != r or len(shape["ks"]) != r:
            raise LigeritoError(-1, "dims and ks need one entry per recursive step")
        h = C.c_void_p()
        _check(lib().lig_prover_config_from_dims(ctx.handle, bits, r, _p(init), int(shape["initial_k"]), _p(dims), _p(ks),
                                                 C.byref(h)))
        self._h, self.bits, self.ctx = h, bits, ctx
        self.num_coeffs = int(init[0]) * int(init[1])
        self.log_n = self.num_coeffs.bit_length() - 1

    def shape(self):
        return _read_shape(lib().lig_prover_config_dims, self._h)


def hardcoded_config_20(ctx, bits=32):
    return ProverConfig(ctx, 20, bits)


def hardcoded_config_24(ctx, bits=32):
    return ProverConfig(ctx, 24, bits)


class Proof(_Handle):
    """FinalizedLigeritoProof (data_structures.rs:91-103)."""

    def __init__(self, h, bits):
        self._h, self.bits = h, bits

    def __del__(self):
        if self._h:
            self._release("lig_proof_free")

    def to_bytes(self) -> bytes:
        n = C.c_size_t(0)
        _check(lib().lig_proof_serialize(self._h, None, 0, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        _check(lib().lig_proof_serialize(self._h, _p(buf), n.value, C.byref(n)))
        return buf.tobytes()

    @staticmethod
    def from_bytes(data: bytes, bits: int = 32) -> "Proof":
        buf = np.frombuffer(bytes(data), dtype=np.uint8).copy()
        h = C.c_void_p()
        _check(lib().lig_proof_deserialize(bits, _p(buf), buf.shape[0], C.byref(h)))
        return Proof(h, bits)

    def proof_size(self) -> int:
        return int(lib().lig_proof_size(self._h))


def prover(config: ProverConfig, poly, fs_seed: int = 1234) -> Proof:
    """src/prover.rs:15-199 (host-resident input)."""
    p = as_elems(poly, config.bits)
    _require_len(p, config.num_coeffs, "prover (prover.rs:15 takes 2^log_n coefficients)")
    h = C.c_void_p()
    _check(lib().lig_prover(config._h, _p(p), fs_seed, C.byref(h)))
    return Proof(h, config.bits)


def prover_dev(config: ProverConfig, d_poly_ptr: int, fs_seed: int = 1234) -> Proof:
    """prover on a device-resident polynomial (pointer to 2^log_n elements)."""
    h = C.c_void_p()
    _check(lib().lig_prover_dev(config._h, C.c_void_p(d_poly_ptr), fs_seed, C.byref(h)))
    return Proof(h, config.bits)


def verifier(log_n: int, proof: Proof, fs_seed: int = 1234) -> int:
    """src/verifier.rs:10-164: 1 accept, 0 reject, negative at the reference's panic sites."""
    return int(lib().lig_verifier(proof.bits, log_n, proof._h, fs_seed))


def verifier_with_config(shape, proof: Proof, fs_seed: int = 1234, ctx: Context = None) -> int:
    """verifier (verifier.rs:10-164) on an explicit VerifierConfig
    (data_structures.rs:23-30) given as a config_shape-style dict; with ctx
    the basis inductions run on its GPU."""
    r = int(shape["recursive_steps"])
    log_dims = np.array([int(d[0]).bit_length() - 1 for d in shape["dims"]] or [0], dtype=np.uint64)
    ks = np.array(shape["ks"] or [0], dtype=np.uint64)
    init_dim = int(shape["initial_dims"][0]).bit_length() - 1
    if ctx is None:
        return int(lib().lig_verifier_with_config(proof.bits, r, init_dim, _p(log_dims), int(shape["initial_k"]), _p(ks),
                                                  proof._h, fs_seed))
    return int(lib().lig_verifier_with_config_dev(ctx.handle, proof.bits, r, init_dim, _p(log_dims),
                                                  int(shape["initial_k"]), _p(ks), proof._h, fs_seed))


def verifier_dev(ctx: Context, log_n: int, proof: Proof, fs_seed: int = 1234) -> int:
    """verifier with the basis inductions on the GPU; same results as verifier()."""
    return int(lib().lig_verifier_dev(ctx.handle, proof.bits, log_n, proof._h, fs_seed))


# ---------------------------------------------------------------- batched --
class Batch(_Handle):
    """Many independent proves on one device (lig_batch; no reference counterpart)."""

    def __init__(self, device: int = 0, bits: int = 32, n_streams: int = 4):
        h = C.c_void_p()
        _check(lib().lig_batch_create(device, bits, n_streams, C.byref(h)))
        self._h, self.bits = h, bits

    def prove(self, log_ns, polys, fs_seeds, on_device: bool = False):
        """polys: host arrays, or device pointers (ints) when on_device."""
        n = len(log_ns)
        if len(polys) != n or len(fs_seeds) != n:
            raise LigeritoError(-1, "log_ns, polys and fs_seeds differ in length")
        ln = (C.c_int * n)(*[int(x) for x in log_ns])
        seeds = (C.c_uint32 * n)(*[int(x) for x in fs_seeds])
        keep = []
        ptrs = (C.c_void_p * n)()
        for i, pl in enumerate(polys):
            if on_device:
                ptrs[i] = int(pl)
            else:
                a = as_elems(pl, self.bits)
                _require_len(a, 1 << int(log_ns[i]), f"batch job {i}")
                keep.append(a)
                ptrs[i] = a.ctypes.data
        out = (C.c_void_p * n)()
        _check(lib().lig_prove_batch(self._h, n, ln, ptrs, 1 if on_device else 0, seeds, out))
        return [Proof(C.c_void_p(out[i]), self.bits) for i in range(n)]

    def close(self):
        if self._h:
            self._release("lig_batch_destroy")

    def __del__(self):
        self.close()


class Server(_Handle):
    """Prove server with continuous batching (lig_server; no reference counterpart)."""

    def __init__(self, device: int = 0, bits: int = 32, n_slots: int = 4, hbm_budget: int = 32 << 30):
        h = C.c_void_p()
        _check(lib().lig_serve_start(device, bits, n_slots, hbm_budget, C.byref(h)))
        self._h, self.bits = h, bits
        self._keep = {}

    def submit(self, log_n: int, poly, fs_seed: int = 1234) -> int:
        """poly: host array (kept referenced until its ticket completes) or a host pointer (int)."""
        t = C.c_uint64()
        if isinstance(poly, int):
            ptr, keep = poly, None
        else:
            keep = as_elems(poly, self.bits)
            _require_len(keep, 1 << int(log_n), "serve request")
            ptr = keep.ctypes.data
        _check(lib().lig_serve_submit(self._h, log_n, C.c_void_p(ptr), fs_seed, C.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def warm(self, log_ns) -> int:
        """Prepares every slot for requests of these sizes (configurations and
        cached device blocks, one zero-input prove per slot and size) while no
        request runs; returns the number of (slot, size) pairs prepared."""
        sizes = [int(x) for x in log_ns]
        arr = (C.c_int * max(1, len(sizes)))(*sizes)
        n = C.c_size_t(0)
        _check(lib().lig_serve_warm(self._h, arr, len(sizes), C.byref(n)))
        return n.value

    def _take(self, fn, ticket):
        h = C.c_void_p()
        lat = C.c_double(0.0)
        r = _check(fn(self._h, ticket, C.byref(h), C.byref(lat)))
        if r == 0:
            return None
        self._keep.pop(ticket, None)
        return Proof(h, self.bits), lat.value

    def poll(self, ticket: int):
        """(proof, latency_ms) once done, else None."""
        return self._take(lib().lig_serve_poll, ticket)

    def wait(self, ticket: int):
        return self._take(lib().lig_serve_wait, ticket)

    def stop(self):
        if self._h:
            self._release("lig_serve_stop")
            self._keep.clear()

    def __del__(self):
        self.stop()
