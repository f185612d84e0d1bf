"""Input/output files of the C++ driver (mfcm_amd/host/mfsv_drive).

``write_input(scenario, dir)`` writes scenario.txt and the float64 arrays the
driver reads; ``run_driver(in_dir, out_dir)`` runs it; ``read_output(out_dir,
nx, ny)`` returns (h, u, v, stats).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "host", "mfsv_drive")


def write_input(sc, path, precision="f64", tile_rows=0, verbose=False):
    os.makedirs(path, exist_ok=True)
    kv = {
        "nx": sc.nx, "ny": sc.ny, "dx": repr(float(sc.dx)), "dy": repr(float(sc.dy)),
        "precision": precision, "method": sc.scheme,
        "start": 0.0, "end": repr(float(sc.t_end)), "step": repr(float(sc.step_duration)),
        "initial_dt": repr(float(sc.initial_dt)), "flags": int(sc.flags),
        "cn_target": repr(float(sc.cn_target)), "d_shallow": repr(float(sc.d_shallow)),
        "d_deep": repr(float(sc.d_deep)), "head_cols": int(sc.head_cols),
        "h_dry": repr(float(sc.h_dry)), "u_max": repr(float(sc.u_max)),
        "tile_rows": int(tile_rows), "verbose": 1 if verbose else 0,
    }
    if isinstance(sc.n_manning, np.ndarray):
        np.ascontiguousarray(sc.n_manning, np.float64).tofile(os.path.join(path, "manning_n.f64"))
        kv["n_shallow"] = kv["n_deep"] = repr(float(np.mean(sc.n_manning)))
    elif sc.n_manning is not None:
        kv["n_shallow"] = kv["n_deep"] = repr(float(sc.n_manning))
    if sc.rain_t is not None:
        kv["rain_t"] = ",".join(repr(float(x)) for x in sc.rain_t)
        kv["rain_v"] = ",".join(repr(float(x)) for x in sc.rain_v)
    if sc.storm is not None:
        for k, v in sc.storm.as_dict().items():
            kv[f"storm_{k}"] = repr(float(v))
    with open(os.path.join(path, "scenario.txt"), "w") as f:
        f.write("! mfsv_drive scenario\n")
        for k, v in kv.items():
            f.write(f"{k} = {v}\n")
    for name, arr in (("zb", sc.zb), ("h", sc.h0), ("u", sc.u0), ("v", sc.v0)):
        np.ascontiguousarray(arr, np.float64).tofile(os.path.join(path, f"{name}.f64"))


def run_driver(in_dir, out_dir, timeout=300):
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is missing: run `python -m mfcm_amd.build`")
    os.makedirs(out_dir, exist_ok=True)
    return subprocess.run([DRIVER, in_dir, out_dir], capture_output=True, text=True,
                          timeout=timeout)


def read_output(path, nx, ny):
    arrs = [np.fromfile(os.path.join(path, f"{k}.f64"), dtype=np.float64).reshape(ny, nx)
            for k in ("h", "u", "v")]
    stats = {}
    with open(os.path.join(path, "stats.txt")) as f:
        for line in f:
            k, v = (s.strip() for s in line.split("=", 1))
            stats[k] = v if k == "precision" else float(v)
    return arrs[0], arrs[1], arrs[2], stats
