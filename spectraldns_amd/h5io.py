"""Checkpoint / results files and the kill-file stop (spectralDNS/h5io/HDF5File.py:57-120).

h5py is not importable in this environment, so checkpoints are written as
NumPy ``.npz`` archives with the reference's keys: ``<name>_c.npz`` holds the
spectral checkpoint data under ``U/3D/<level>`` with ``tstep`` and ``t``;
``<name>_w.npz`` holds the physical results under ``<NAME>/3D/<tstep>``.
The kill file ``killspectraldns`` in the working directory makes ``update``
checkpoint and exit with status 1 (HDF5File.py:82-92, 107-120).
"""
import os
import sys

import numpy as np

__all__ = ["HDF5File", "read_checkpoint"]


def read_checkpoint(filename):
    """Arrays, t and tstep of a checkpoint written by HDF5File.update: the file
    ``<filename>_c.npz`` (or ``filename`` itself when it ends in .npz) as a
    dict keyed like the reference's HDF5 datasets (``U/3D/0``), with ``t``
    (float) and ``tstep`` (int).  Loads nothing executable (no pickles)."""
    fn = filename if filename.endswith(".npz") else filename + "_c.npz"
    with np.load(fn, allow_pickle=False) as f:
        out = {k: f[k] for k in f.files}
    out["t"] = float(out["t"])
    out["tstep"] = int(out["tstep"])
    return out


def _host(a):
    try:
        return a.detach().cpu().numpy()
    except AttributeError:
        return np.asarray(a)


class HDF5File:
    def __init__(self, filename, checkpoint=None, results=None):
        self.filename = filename
        self.checkpoint = checkpoint or {}
        self.results = results or {}
        self._w = {}

    def update(self, params, **kw):
        if params.tstep % params.write_result == 0 and self.results.get("data"):
            self.update_components(**kw)
            for name, val in self.results["data"].items():
                arr = val[0] if isinstance(val, (list, tuple)) else val
                self._w["%s/3D/%d" % (name, params.tstep)] = _host(arr)
            np.savez(self.filename + "_w.npz", **self._w)
        kill = self.check_if_kill()
        if params.tstep % params.checkpoint == 0 or kill:
            data = {}
            for level, fields in self.checkpoint.get("data", {}).items():
                for name, arrs in fields.items():
                    data["%s/3D/%s" % (name, level)] = _host(arrs[0])
            data["tstep"] = np.array(params.tstep)
            data["t"] = np.array(float(params.t))
            np.savez(self.filename + "_c.npz", **data)
            if kill:
                sys.exit(1)

    def update_components(self, **kw):
        pass

    def open(self):
        pass

    def close(self):
        pass

    @staticmethod
    def check_if_kill():
        """True (and the file removed) if 'killspectraldns' is in the working directory."""
        if "killspectraldns" in os.listdir(os.getcwd()):
            os.remove("killspectraldns")
            print("killspectraldns Found! Stopping simulations cleanly by checkpointing...")
            return True
        return False
