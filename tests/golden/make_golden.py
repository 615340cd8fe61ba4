#!/usr/bin/env python3
"""Writes tests/golden/gradcheck_scene_deg{0,3}.json: the reference's
finite-difference scene (tests/grad_check.cu:33-37, 72-132 of SP5555/diffsplat:
8 splats, 32x32 image, target 0.7, ortho fitter camera; degree 3 adds seeded
SH rest coefficients the reference never checked) with its FP64 outputs:
the loss, the rendered image, the analytic parameter gradient in both
gradient modes (the oracle's FP64 restatement) and the FP64 central finite
difference of the loss (h = 1e-6, with the tile lists of the unperturbed
scene), which checks the analytic gradient independently.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

from helpers import fd_grad, fp64_backward, fp64_forward, gradcheck_scene  # noqa: E402


def fixture(deg: int) -> dict:
    sc = gradcheck_scene(deg)
    n, W, H = sc["n"], sc["W"], sc["H"]
    fw = fp64_forward(sc["params"], n, deg, sc["cam"], sc["pv"], W, H, sc["target"])
    g_exact = fp64_backward(fw, sc["params"], n, deg, sc["cam"], sc["pv"], W, H, 0)["gparams"]
    g_ref = fp64_backward(fw, sc["params"], n, deg, sc["cam"], sc["pv"], W, H, 1)["gparams"]
    fd = fd_grad(sc, pairs=fw["ta"])
    return {
        "source": "SP5555/diffsplat tests/grad_check.cu:33-37, 72-132 (scene, size, "
                  "tolerance); values from the FP64 oracle (oracle/ds_oracle.cpp)",
        "n": n, "width": W, "height": H, "sh_degree": deg,
        "reference_rel_tol": 0.05, "reference_fd_step": 1e-3,
        "params": sc["params"].tolist(), "target_value": 0.7,
        "pv": sc["pv"].tolist(), "cam_pos": sc["cam"].tolist(),
        "n_isects": int(fw["ta"]["n_isects"]),
        "flatten_ids": fw["ta"]["ids"].astype(int).tolist(),
        "loss": float(fw["loss"]), "rgb": fw["comp"]["rgb"].tolist(),
        "grad_exact": g_exact.tolist(), "grad_reference": g_ref.tolist(),
        "grad_fd": fd.tolist(), "fd_step": 1e-6,
    }


def main():
    for deg in (0, 3):
        path = os.path.join(HERE, f"gradcheck_scene_deg{deg}.json")
        with open(path, "w") as f:
            json.dump(fixture(deg), f, separators=(",", ":"))
        print(path)


if __name__ == "__main__":
    main()
