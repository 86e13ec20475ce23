"""CPU checks of lib3dmatch_amd.so's C ABI: it loads, exports every function
include/tdm.h declares, and its host-only entry points (network JSON, .marvin
weights) behave as the reference's Net / Tensor code (core/marvin.hpp)."""
import ctypes
import os
import struct

import numpy as np
import pytest

import tdm
from conftest import REFERENCE
from tdm import netdef


def test_library_exports_every_header_symbol():
    L = tdm.lib()
    syms = tdm.header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings_and_version():
    L = tdm.lib()
    assert L.tdm_version() == 1
    assert L.tdm_status_string(0) == b"ok"
    assert L.tdm_status_string(tdm.ERR_NO_DEVICE) == b"no GPU device"


def test_invalid_arguments_are_reported_not_fatal():
    with pytest.raises(tdm.TdmError) as e:
        tdm.read_ply_points("/nonexistent/cloud.ply")
    assert e.value.status == tdm.ERR_IO
    L = tdm.lib()
    st = L.tdm_keypoints_to_grid(None, 3, None, ctypes.c_float(0.01), None)
    assert st == tdm.ERR_INVALID_ARG and b"null" in L.tdm_last_error()


def test_integrate_size_limits_rejected_on_the_host():
    """The TSDF kernels address a grid slice and the depth image with 32-bit
    byte offsets; larger inputs are refused before any device work."""
    L = tdm.lib()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks come first
    f = ctypes.c_float
    st = L.tdm_integrate(None, dummy, dummy, dummy, 480, 640, 70000, 70000, 1, f(0), f(0), f(0),
                         f(0.01), f(0.05), dummy, dummy, None)
    assert st == tdm.ERR_INVALID_ARG and b"slice" in L.tdm_last_error()
    st = L.tdm_integrate(None, dummy, dummy, dummy, 40000, 40000, 8, 8, 8, f(0), f(0), f(0),
                         f(0.01), f(0.05), dummy, dummy, None)
    assert st == tdm.ERR_INVALID_ARG and b"image" in L.tdm_last_error()
    st = L.tdm_integrate(None, dummy, dummy, dummy, 480, 640, 8, 8, 8, f(0), f(0), f(0),
                         f(0.0), f(0.05), dummy, dummy, None)
    assert st == tdm.ERR_INVALID_ARG


def _shapes(net):
    return [(n, net.layer_info(i)) for i, n in enumerate(net.layer_names)]


def test_generated_net_json(tmp_path):
    js = str(tmp_path / "net.json")
    netdef.write_net_json(js)
    net = tdm.Net(js)
    assert net.output_dim == 512
    got = _shapes(net)
    want = netdef.layer_shapes()
    assert [n for n, _ in got] == [w[0] for w in want]
    for (name, info), (_, cin, cout, k, d, od, pooled) in zip(got, want):
        assert (info["cin"], info["cout"], info["window"], info["in_extent"], info["out_extent"]) == (cin, cout, k, d, od)
        assert info["pool"] == (pooled is not None)
        assert info["relu"] == (name != "feat")  # no ReLU after feat
    assert got[1][1]["final_extent"] == 13  # 26^3 pooled to 13^3 (core/marvin.hpp:5560-5562)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not mounted")
def test_reference_net_json_parses_to_same_chain(tmp_path):
    ref = tdm.Net(os.path.join(REFERENCE, "core", "3dmatch-net-test.json"))
    js = str(tmp_path / "net.json")
    netdef.write_net_json(js)
    assert _shapes(ref) == _shapes(tdm.Net(js))


def test_marvin_layout_and_round_trip(tmp_path):
    p = str(tmp_path / "w.marvin")
    a = np.arange(6, dtype=np.float32).reshape(2, 3)
    b = np.array([1.5], dtype=np.float32)
    tdm.marvin_write(p, [("conv1.weight", a), ("b", b)])
    raw = open(p, "rb").read()
    # u8 type(1=float), u32 sizeof, i32 len, name, i32 ndims, i32 dims..., data
    want = (struct.pack("<BIi", 1, 4, 12) + b"conv1.weight" + struct.pack("<i", 2) +
            struct.pack("<2i", 2, 3) + a.tobytes() +
            struct.pack("<BIi", 1, 4, 1) + b"b" + struct.pack("<i", 1) + struct.pack("<i", 1) + b.tobytes())
    assert raw == want
    assert tdm.marvin_count(p) == (2, 7)


def test_marvin_half_and_double_tensors_are_converted(tmp_path):
    p = str(tmp_path / "mixed.marvin")
    h = np.array([1.0, -2.5, 65504.0, 2.0 ** -24], dtype=np.float16)
    d = np.array([0.1, 3.0], dtype=np.float64)
    with open(p, "wb") as f:
        f.write(struct.pack("<BIi", 0, 2, 1) + b"h" + struct.pack("<ii", 1, 4) + h.tobytes())
        f.write(struct.pack("<BIi", 2, 8, 1) + b"d" + struct.pack("<ii", 1, 2) + d.tobytes())
    assert tdm.marvin_count(p) == (2, 6)


def test_weights_by_name_and_mismatch_skip(tmp_path):
    js = str(tmp_path / "net.json")
    netdef.write_net_json(js)
    w = dict(netdef.he_normal_weights(5))
    p = str(tmp_path / "w.marvin")
    tensors = list(w.items())
    # a wrong-size conv3.weight is skipped (core/marvin.hpp:3321-3377) ...
    bad = [(n, (a[:, :1] if n == "conv3.weight" else a)) for n, a in tensors]
    tdm.marvin_write(p, bad)
    net = tdm.Net(js)
    net.load_weights(p)
    # ... so a forward refuses to run on the missing layer instead of using garbage
    L = tdm.lib()
    st = L.tdm_net_forward(net._h, ctypes.c_void_p(8), 1, ctypes.c_void_p(8), None)
    assert st == tdm.ERR_INVALID_ARG and b"conv3" in L.tdm_last_error()


def test_unsupported_layers_rejected(tmp_path):
    d = netdef.net_dict()
    d["layers"][2]["mode"] = "Sigmoid"
    p = str(tmp_path / "bad.json")
    import json

    json.dump(d, open(p, "w"))
    with pytest.raises(tdm.TdmError) as e:
        tdm.Net(p)
    assert e.value.status == tdm.ERR_UNSUPPORTED
    open(p, "w").write("{ \"layers\": [ }")
    with pytest.raises(tdm.TdmError) as e:
        tdm.Net(p)
    assert e.value.status == tdm.ERR_FORMAT


def test_flops_per_patch():
    assert netdef.flops_per_patch() == 6_452_240_384
    assert sum(a.size for _, a in netdef.he_normal_weights(5)) == 14_311_488


def _cli(args, cwd):
    import subprocess

    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def cli_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli_args")
    netdef.write_net_json(str(d / "tmp.json"))
    netdef.write_net_json(str(d / "3dmatch-net-test.json"))
    tdm.marvin_write(str(d / "3dmatch-weights-snapshot-137000.marvin"), netdef.he_normal_weights(5))
    return d


@pytest.mark.parametrize("tool", ["get3dmatch_descriptors", "descriptor_demo"])
def test_cli_rejects_unknown_precision(cli_dir, tool):
    args = [os.path.join(tdm.BIN_DIR, tool)]
    args += ["data.bin"] if tool == "get3dmatch_descriptors" else ["cloud.ply", "out"]
    r = _cli(args + ["--precision", "fp64"], str(cli_dir))
    assert r.returncode != 0 and "unknown precision fp64" in r.stderr


def test_get3dmatch_descriptors_rejects_wrong_volume_size(cli_dir):
    """data.bin volumes must have the net's 30^3 extent: 20^3 volumes are
    refused after reading the file, before any device buffer or launch."""
    vols = np.zeros((2, 20, 20, 20), np.float32)
    (cli_dir / "small.bin").write_bytes(struct.pack("<4f", 2, 20, 20, 20) + vols.tobytes())
    r = _cli([os.path.join(tdm.BIN_DIR, "get3dmatch_descriptors"), "small.bin"], str(cli_dir))
    assert r.returncode != 0 and "30^3" in r.stderr, r.stderr
    assert not (cli_dir / "feat.bin").exists()
