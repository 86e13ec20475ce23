"""Untrusted files: seeded mutations (truncations, byte flips, oversized and
negative size fields) of every format the host code parses -- PLY, depth
PNG, .marvin, net JSON, grid / keypoint / descriptor / patch-volume files,
pose text.  Each read must either succeed or return an error status; none may
crash.  test_host_sanitizers.py runs this file against the AddressSanitizer /
UBSan build of the library, where an out-of-bounds read or an overflow aborts
the run."""
import json
import os
import struct

import numpy as np
import pytest

import tdm
from conftest import GOLDEN
from tdm import netdef

SEED = 12345
N_MUTANTS = 120


def _mutants(data: bytes, rng, n=N_MUTANTS, fields=()):
    """Truncations, single and burst byte flips, and size fields (byte offsets
    in `fields`, int32) set to hostile values."""
    out = [b"", data[:1], data[: len(data) // 2]]
    for _ in range(n):
        kind = rng.integers(0, 4)
        b = bytearray(data)
        if kind == 0:
            b = b[: int(rng.integers(0, len(b)))]
        elif kind == 1:
            for _ in range(int(rng.integers(1, 4))):
                b[int(rng.integers(0, len(b)))] ^= 1 << int(rng.integers(0, 8))
        elif kind == 2:
            at = int(rng.integers(0, len(b)))
            span = min(int(rng.integers(1, 64)), len(b) - at)
            b[at:at + span] = rng.integers(0, 256, span, dtype=np.uint8).tobytes()
        elif fields:
            at = fields[int(rng.integers(0, len(fields)))]
            v = int(rng.choice([-1, -2**31, 2**31 - 1, 2**30, 0, 65536, 2**24 + 1]))
            b[at:at + 4] = struct.pack("<i", v)
        out.append(bytes(b))
    return out


def _try(fn, path):
    try:
        fn(path)
    except tdm.TdmError:
        pass


def _each(tmp_path, name, mutants, fn):
    p = str(tmp_path / name)
    for m in mutants:
        with open(p, "wb") as f:
            f.write(m)
        _try(fn, p)


def test_png_mutants(tmp_path):
    rng = np.random.default_rng(SEED)
    data = open(os.path.join(GOLDEN, "sample_frames", "frame-000008.depth.png"), "rb").read()
    # IHDR width / height (big-endian in PNG, hostile either way)
    _each(tmp_path, "d.png", _mutants(data, rng, fields=(16, 20)), tdm.read_png_u16)
    _each(tmp_path, "d.png", _mutants(data, rng, n=40), tdm.read_depth_png)


def test_ply_mutants(tmp_path):
    rng = np.random.default_rng(SEED + 1)
    p = str(tmp_path / "c.ply")
    tdm.write_ply_points(p, rng.random((50, 3), dtype=np.float32))
    data = open(p, "rb").read()
    text_variants = [data.replace(b"element vertex 50", b"element vertex " + v)
                     for v in (b"-1", b"2147483647", b"99999999999", b"5000000", b"0x10", b"")]
    _each(tmp_path, "m.ply", text_variants + _mutants(data, rng), tdm.read_ply_points)


def test_marvin_mutants(tmp_path):
    rng = np.random.default_rng(SEED + 2)
    p = str(tmp_path / "w.marvin")
    tdm.marvin_write(p, [("conv1.weight", rng.random((4, 1, 3, 3, 3), dtype=np.float32)),
                         ("conv1.bias", rng.random((1, 4, 1, 1, 1), dtype=np.float32))])
    data = open(p, "rb").read()
    # type byte at 0, sizeof at 1, name length at 5, ndims after the 12-byte name
    fields = (1, 5, 21, 25, 29)
    muts = _mutants(data, rng, fields=fields)
    _each(tmp_path, "m.marvin", muts, tdm.marvin_count)
    js = str(tmp_path / "net.json")
    netdef.write_net_json(js)
    net = tdm.Net(js)
    _each(tmp_path, "m.marvin", muts[:60], net.load_weights)
    net.close()


def test_net_json_mutants(tmp_path):
    rng = np.random.default_rng(SEED + 3)
    text = json.dumps(netdef.net_dict()).encode()
    muts = _mutants(text, rng)
    # well-formed JSON with hostile values
    for key, val in (("num_output", -5), ("num_output", 2**31 - 1), ("window", [0, 0, 0]),
                     ("window", [99999, 3, 3]), ("stride", [0, 1, 1]), ("num_output", "x"),
                     ("window", []), ("dim", [1, 2**30, 30, 30, 30]), ("dim", [-1, 1, 30, 30, 30])):
        d = netdef.net_dict()
        for layer in d["layers"]:
            if key in layer:
                layer[key] = val
        muts.append(json.dumps(d).encode())
    muts += [b"[" * 100_000, b"{\"layers\": " + b"[" * 5000, b"\"\\u", b"{\"a\": 1e999999}"]

    def build(path):
        tdm.Net(path).close()

    _each(tmp_path, "n.json", muts, build)


def test_binary_grid_and_descriptor_mutants(tmp_path):
    rng = np.random.default_rng(SEED + 4)
    g = str(tmp_path / "g.bin")
    tdm.write_grid_bin(g, rng.random((3, 4, 5), dtype=np.float32), np.zeros(3, np.float32),
                       np.float32(0.01), np.float32(0.05))
    _each(tmp_path, "m.bin", _mutants(open(g, "rb").read(), rng, fields=(0, 4, 8)),
          tdm.read_grid_bin)
    k = str(tmp_path / "k.bin")
    tdm.write_keypts_bin(k, rng.random((6, 3), dtype=np.float32))
    _each(tmp_path, "m.bin", _mutants(open(k, "rb").read(), rng, n=40, fields=(0,)),
          tdm.read_keypts_bin)
    d = str(tmp_path / "d.bin")
    tdm.write_desc_bin(d, rng.random((3, 512), dtype=np.float32))
    _each(tmp_path, "m.bin", _mutants(open(d, "rb").read(), rng, n=40, fields=(0, 4)),
          tdm.read_desc_bin)
    # patch volumes: float32 N, 3 float32 dims, then N x dims floats
    vol = struct.pack("<4f", 2, 3, 3, 3) + rng.random(54, dtype=np.float32).tobytes()
    hostile = [struct.pack("<4f", *v) + vol[16:] for v in
               ((2, -3, 3, 3), (1e9, 3, 3, 3), (2, 3e9, 3, 3), (float("nan"), 3, 3, 3),
                (2, 3, 3, float("inf")), (-1, 3, 3, 3), (2.5, 3, 3, 3))]
    _each(tmp_path, "m.bin", hostile + _mutants(vol, rng, n=40), tdm.read_patch_volumes)


def test_pose_text_mutants(tmp_path):
    texts = [b"", b"1 2 3", b"nan " * 16, b"1e99999 " * 16, b"\x00" * 64, b"1 2 3 4\n" * 2,
             b"-" * 1000, b"0x1p3 " * 16, b"1 " * 100_000]
    _each(tmp_path, "p.txt", texts, lambda p: tdm.load_matrix(p, 4, 4))


def test_hostile_header_sizes_fail_cleanly(tmp_path):
    """Size fields that would overflow an allocation are errors, not crashes."""
    p = str(tmp_path / "big.ply")
    open(p, "wb").write(b"ply\nformat binary_little_endian 1.0\nelement vertex 2147483647\n"
                        b"property float x\nproperty float y\nproperty float z\nend_header\n")
    with pytest.raises(tdm.TdmError):
        tdm.read_ply_points(p)
    q = str(tmp_path / "big.bin")
    open(q, "wb").write(struct.pack("<4f", 3e9, 30, 30, 30))
    with pytest.raises(tdm.TdmError):
        tdm.read_patch_volumes(q)
