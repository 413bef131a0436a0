"""bincode 1.3 (fixint, little endian) layouts of the reference's wire types, in pure Python.

RoundRecord (src/protocol/prover.rs:14-21) = u64 round, then four ModuleElement<8>
(l_vec, r_vec, l_poly, r_poly). ModuleElement<8> (src/algebra/module.rs:95-102) = u64 len (8),
then 8 Polys, each u64 len + len x 16-byte LE Fp (src/algebra/field.rs:112-119).
Statement (src/commitment/sis.rs:23-30) = file_id[32], u64 original_len_bytes, u64 d,
com_0, g_0_seed[32].
"""
import struct

import numpy as np

K = 8


def module_bytes(m):
    """m: (K, n, 2) uint64 words"""
    m = np.ascontiguousarray(m, dtype=np.uint64)
    out = [struct.pack("<Q", m.shape[0])]
    for k in range(m.shape[0]):
        out.append(struct.pack("<Q", m.shape[1]))
        out.append(m[k].astype("<u8").tobytes())
    return b"".join(out)


def read_module(buf, pos):
    (n,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    coords = []
    for _ in range(n):
        (ln,) = struct.unpack_from("<Q", buf, pos)
        pos += 8
        arr = np.frombuffer(buf, dtype="<u8", count=2 * ln, offset=pos).reshape(ln, 2).astype(np.uint64)
        pos += 16 * ln
        coords.append(arr)
    if n != K or len({c.shape[0] for c in coords}) != 1:
        raise ValueError("module shape")
    return np.stack(coords), pos


def encode_record(rnd, l_vec, r_vec, l_poly, r_poly):
    return struct.pack("<Q", rnd) + b"".join(module_bytes(x) for x in (l_vec, r_vec, l_poly, r_poly))


def decode_record(buf):
    (rnd,) = struct.unpack_from("<Q", buf, 0)
    pos = 8
    mods = []
    for _ in range(4):
        m, pos = read_module(buf, pos)
        mods.append(m)
    return rnd, mods, pos


def split_records(blob):
    out = []
    pos = 0
    while pos < len(blob):
        _, _, used = decode_record(blob[pos:])
        out.append(blob[pos:pos + used])
        pos += used
    return out


def encode_statement(file_id, original_len, d, com0, g0_seed):
    return bytes(file_id) + struct.pack("<QQ", original_len, d) + module_bytes(com0) + bytes(g0_seed)


def decode_statement(buf):
    file_id = buf[:32]
    original_len, d = struct.unpack_from("<QQ", buf, 32)
    com0, pos = read_module(buf, 48)
    g0 = buf[pos:pos + 32]
    return {"file_id": file_id, "original_len_bytes": original_len, "d": d, "com_0": com0, "g_0_seed": g0}


def zero_module(n):
    return np.zeros((K, n, 2), dtype=np.uint64)


def pattern_bytes(size, tweak):
    """write_pattern_file (src/protocol/prover.rs:473-478): byte = (idx as u8) * 17 + tweak"""
    idx = np.arange(size, dtype=np.uint64)
    return (((idx & 0xFF) * 17 + tweak) & 0xFF).astype(np.uint8).tobytes()


def byte_len_for_depth(d):
    """prover.rs:480-487"""
    return 1 if d == 6 else ((1 << (2 * (d - 1))) + 1) * 13
