"""Extract the trained EPG-X networks of MCR-MWI into plain arrays.

The reference ships the two multilayer perceptrons of its EPG-X T1 path
(MCRMWI/EPGXgen_net/MCRMWI_MLP_EPGX_RFphase50_T1M234_{phase,magn}.mat,
loaded by MCRMWI/gpuMCRMWI.m:386-387 with alpha = 0.01) as MATLAB `dlnet`
structs whose Weights and Bias are `dlarray` objects.  A MAT-file keeps the
values of such objects in its `__function_workspace__` variable: a MAT v5
stream (version and endian marks, then data elements) holding a
`FileWrapper__` cell array, whose cells are the objects' property values.
scipy's MAT v5 reader parses that stream as data; nothing in the file is
executed.  The float32 arrays come out in object order, fc1.Weights,
fc1.Bias, fc2.Weights, ...; the chain of shapes (Weights [out, in] with in =
the previous layer's out, Bias [out, 1]) is checked, so a different order
cannot pass.

Writes mcrmwi_epgx_phase.npz and mcrmwi_epgx_magn.npz beside this file, with
fc<l>_Weights / fc<l>_Bias (the layout gacelle_amd.models.mlp_pack reads)
and alpha.

    python gacelle_amd/data/extract_epgx_networks.py [path/to/MCRMWI/EPGXgen_net]
"""
from __future__ import annotations

import io
import os
import sys

import numpy as np
import scipy.io as sio
from scipy.io.matlab._mio5 import MatFile5Reader

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/MCRMWI/EPGXgen_net"
FILES = {"phase": "MCRMWI_MLP_EPGX_RFphase50_T1M234_phase.mat",
         "magn": "MCRMWI_MLP_EPGX_RFphase50_T1M234_magn.mat"}
N_OUT = {"phase": 1, "magn": 2}
ALPHA = 0.01  # MCRMWI/gpuMCRMWI.m:386-387


def workspace_arrays(path: str) -> list[np.ndarray]:
    """The float arrays held by the MAT-file's object store, in object order."""
    ws = sio.loadmat(path)["__function_workspace__"].tobytes()
    # a MAT v5 file: 116 bytes of text and an 8-byte subsystem offset, then the
    # stream's own version / endian marks (4 bytes); its 4 padding bytes are
    # dropped and its data elements follow
    stream = b" " * 116 + b"\0" * 8 + ws[:4] + ws[8:]
    r = MatFile5Reader(io.BytesIO(stream), squeeze_me=False, struct_as_record=True)
    r.initialize_read()
    r.mat_stream.seek(128)
    hdr, _ = r.read_var_header()
    cells = r.read_var_array(hdr, process=False)["MCOS"][0, 0][0]["arr"]
    out = []
    for i in range(cells.shape[0]):
        c = cells[i, 0]
        if c.dtype == object and c.size == 2:  # (metadata, value) of one array property
            v = c.ravel()[1]
            if isinstance(v, np.ndarray) and v.dtype == np.float32:
                out.append(v)
    return out


def network(path: str, n_out: int) -> dict[str, np.ndarray]:
    arrays = workspace_arrays(path)
    if len(arrays) % 2:
        raise ValueError(f"{path}: odd number of arrays ({len(arrays)})")
    out, prev = {}, 11
    for l in range(len(arrays) // 2):
        W, b = arrays[2 * l], arrays[2 * l + 1]
        # a Bias is a vector of the layer's width ([out, 1], or [1, out] for the
        # magnitude network's last layer)
        if W.ndim != 2 or W.shape[1] != prev or b.ndim != 2 or min(b.shape) != 1 or b.size != W.shape[0]:
            raise ValueError(f"{path}: layer {l + 1} shapes {W.shape} / {b.shape} break the chain from {prev}")
        out[f"fc{l + 1}_Weights"] = np.ascontiguousarray(W, np.float32)
        out[f"fc{l + 1}_Bias"] = np.ascontiguousarray(b.ravel(), np.float32)
        prev = W.shape[0]
    if prev != n_out:
        raise ValueError(f"{path}: {prev} outputs, {n_out} expected")
    out["alpha"] = np.float32(ALPHA)
    return out


def main(src: str = DEFAULT_SRC) -> None:
    for name, fname in FILES.items():
        net = network(os.path.join(src, fname), N_OUT[name])
        dst = os.path.join(HERE, f"mcrmwi_epgx_{name}.npz")
        np.savez_compressed(dst, **net)
        widths = [11] + [net[f"fc{l}_Weights"].shape[0] for l in range(1, len(net) // 2 + 1)]
        print(f"wrote {dst}: layers {widths}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC)
