"""The 3DMatch test network as a Marvin JSON file, and seeded He-normal weights.

The layer list, names and fields are those of core/3dmatch-net-test.json
(a PlaceHolderData [batch,1,30,30,30] -> conv1..conv8 3x3x3 with ReLU,
2x2x2/2 max pool after relu2, then the 1x1x1 `feat` convolution without
ReLU, core/3dmatch-net-test.json:29-253), so the network file and the
`.marvin` tensor names are interchangeable with the reference's.
"""
from __future__ import annotations

import json

import numpy as np

# (name, num_output, window) of the convolutions, in order.
CONVS = [
    ("conv1", 64, 3), ("conv2", 64, 3), ("conv3", 128, 3), ("conv4", 128, 3),
    ("conv5", 256, 3), ("conv6", 256, 3), ("conv7", 512, 3), ("conv8", 512, 3),
    ("feat", 512, 1),
]
POOL_AFTER = "conv2"
PATCH = 30


def _conv(name, src, out, k):
    return {
        "in": [src], "type": "Convolution", "name": name, "num_output": out,
        "window": [k, k, k], "padding": [0, 0, 0], "stride": [1, 1, 1],
        "upscale": [1, 1, 1], "weight_lr_mult": 1.0, "weight_filler": "Xavier",
        "bias_lr_mult": 2.0, "bias_filler": "Constant", "bias_filler_param": 0.0,
        "train_me": True, "out": [name],
    }


def net_dict(batch: int = 50, convs=CONVS, pool_after=POOL_AFTER, patch=PATCH) -> dict:
    layers = [{"type": "PlaceHolderData", "name": "dataTest", "phase": "Testing",
               "dim": [batch, 1, patch, patch, patch], "out": ["data"]}]
    src = "data"
    for i, (name, out, k) in enumerate(convs):
        layers.append(_conv(name, src, out, k))
        src = name
        if i < len(convs) - 1:
            layers.append({"in": [name], "type": "Activation", "name": f"relu{i + 1}",
                           "mode": "ReLU", "out": [name]})
        if name == pool_after:
            layers.append({"in": [name], "type": "Pooling", "name": "pool1", "mode": "max",
                           "window": [2, 2, 2], "padding": [0, 0, 0], "stride": [2, 2, 2],
                           "out": ["pool1"]})
            src = "pool1"
    return {
        "train": {"path": "train", "solver": "SGD", "regularizer": "L2", "momentum": 0.99,
                  "weight_decay": 0.0005, "base_lr": 0.01, "lr_policy": "LR_inv",
                  "lr_gamma": 0.0001, "lr_power": 0.75, "max_iter": 1000000, "train_iter": 1,
                  "snapshot_iter": 500, "display_iter": 1, "test_iter": 5, "test_interval": 100,
                  "debug_mode": False, "GPU_solver": 0, "GPU": [0]},
        "test": {"debug_mode": False, "GPU": 0},
        "layers": layers,
    }


def write_net_json(path: str, batch: int = 50, convs=CONVS, pool_after=POOL_AFTER,
                   patch=PATCH) -> None:
    with open(path, "w") as f:
        json.dump(net_dict(batch, convs, pool_after, patch), f, indent=1)


def he_normal_weights(seed: int = 5, convs=CONVS) -> list[tuple[str, np.ndarray]]:
    """He-normal N(0, 2/fan_in) weights [O, I, k, k, k] and zero biases [1, O, 1, 1, 1]
    (the .marvin dims of core/marvin.hpp:4440-4447), fan_in = I*k^3."""
    rng = np.random.default_rng(seed)
    tensors = []
    cin = 1
    for name, cout, k in convs:
        fan_in = cin * k ** 3
        w = rng.standard_normal((cout, cin, k, k, k), dtype=np.float32) * np.float32(np.sqrt(2.0 / fan_in))
        tensors.append((f"{name}.weight", w.astype(np.float32)))
        tensors.append((f"{name}.bias", np.zeros((1, cout, 1, 1, 1), dtype=np.float32)))
        cin = cout
    return tensors


def flops_per_patch(convs=CONVS, pool_after=POOL_AFTER, patch=PATCH) -> int:
    """2 x multiply-accumulates of all convolutions for one patch."""
    total, d, cin = 0, patch, 1
    for name, cout, k in convs:
        od = d - k + 1
        total += 2 * od ** 3 * cout * cin * k ** 3
        d = od // 2 if name == pool_after else od
        cin = cout
    return total


def layer_shapes(convs=CONVS, pool_after=POOL_AFTER, patch=PATCH):
    """[(name, cin, cout, k, in_extent, out_extent, pooled_extent or None)]."""
    out, d, cin = [], patch, 1
    for name, cout, k in convs:
        od = d - k + 1
        pooled = od // 2 if name == pool_after else None
        out.append((name, cin, cout, k, d, od, pooled))
        d = pooled if pooled is not None else od
        cin = cout
    return out
