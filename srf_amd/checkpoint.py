"""OSEM checkpoints: the image estimates and the number of sub-iterations done,
under a key that names the workload, so that an interrupted reconstruction
resumes where it stopped and ends bit-identical to an uninterrupted one (the
kernels are deterministic: fixed-point backprojection, fixed reduction
orders).

The reference saves the estimate after every iteration as ./result_{i}.npy
(srfnef/func_classes/distribute.py:149-151) and has no resume.  Here a
checkpoint is one file written atomically (a temporary file, then
os.replace), read back with torch.load(weights_only=True): tensors, numbers,
strings and containers only, nothing executable.
"""
import hashlib
import os

import torch

FORMAT = 1


def tensor_digest(t):
    """sha1 of a tensor's bytes (e.g. an event permutation), for the key."""
    a = t.detach().to("cpu").contiguous()
    return hashlib.sha1(a.numpy().tobytes()).hexdigest()


def save(path, key, done, images, values=None):
    """Write a checkpoint: `key` (dict of numbers / strings / lists naming the
    workload), `done` sub-iterations, `images` (list of tensors, copied to
    host), optional `values` (list of per-scan event values)."""
    state = {"format": FORMAT, "key": dict(key), "done": int(done),
             "images": [x.detach().to("cpu").clone() for x in images]}
    if values is not None:
        state["values"] = [v.detach().to("cpu").clone() for v in values]
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = f"{path}.tmp{os.getpid()}"
    torch.save(state, tmp)
    os.replace(tmp, path)


def load(path, key):
    """The checkpoint at `path` if it exists and was written for `key`, else
    None.  A file for another workload raises ValueError (resuming it would
    silently mix two reconstructions); an unreadable file raises too."""
    if not path or not os.path.exists(path):
        return None
    state = torch.load(path, map_location="cpu", weights_only=True)
    if not isinstance(state, dict) or state.get("format") != FORMAT:
        raise ValueError(f"{path}: not an OSEM checkpoint of format {FORMAT}")
    if state.get("key") != dict(key):
        diff = sorted(k for k in set(state.get("key", {})) | set(key)
                      if state.get("key", {}).get(k) != dict(key).get(k))
        raise ValueError(f"{path}: checkpoint of another workload (differs in {diff})")
    return state


__all__ = ["save", "load", "tensor_digest"]
