"""Batched solve of independent scenes (no counterpart in the reference).

``concat`` joins the body and contact arrays of several scenes (as returned by
``SOLFEC.bodies()`` / ``SOLFEC.contacts()``) into one store with a scene id
per body; ``upload`` sends such a store to a ``Device`` and assembles it, after
which ``Device.batch_solve`` runs every scene in one persistent launch.
"""
import numpy as np

_BODY_KEYS = ("kind", "rot", "x0", "jinv", "minv", "velo", "velo0")
_CON_KEYS = ("master", "slave", "mpnt", "spnt", "base", "gap", "friction", "restitution", "cohesion", "R", "U")


def concat(scenes):
    """scenes: list of (bodies, contacts) dicts.  Returns (bodies, contacts, body_scene, contact_scene)."""
    B = {k: [] for k in _BODY_KEYS}
    Cn = {k: [] for k in _CON_KEYS}
    body_scene, con_scene = [], []
    off = 0
    for s, (b, c) in enumerate(scenes):
        nb = len(b["kind"])
        for k in _BODY_KEYS:
            B[k].append(np.asarray(b[k]))
        m = np.asarray(c["master"], np.int32) + off
        sl = np.asarray(c["slave"], np.int32)
        sl = np.where(sl >= 0, sl + off, sl).astype(np.int32)
        for k in _CON_KEYS:
            Cn[k].append(m if k == "master" else sl if k == "slave" else np.asarray(c[k]))
        body_scene.append(np.full(nb, s, np.int32))
        con_scene.append(np.full(len(c["gap"]), s, np.int32))
        off += nb
    bodies = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in B.items()}
    contacts = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in Cn.items()}
    return bodies, contacts, np.concatenate(body_scene), np.concatenate(con_scene)


def split(values, contact_scene, per_contact):
    """Per-scene views of a natural-order array with per_contact entries per contact."""
    n = len(contact_scene)
    v = np.asarray(values).reshape(n, per_contact)
    bounds = np.searchsorted(contact_scene, np.arange(contact_scene.max() + 2 if n else 1))
    return [v[bounds[s]:bounds[s + 1]] for s in range(len(bounds) - 1)]


def upload(dev, bodies, contacts, body_scene, step, R=None, U=None):
    """Upload a joined store, register its scenes and assemble it."""
    b, c = bodies, contacts
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    ns = dev.batch_scenes(body_scene)
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"] if R is None else R, c["U"] if U is None else U)
    dev.assemble(step)
    return ns
