"""State files (SOLFEC_Write_State / SOLFEC_Read_State): the body records of
BODY_Write_State (bod.c:2202-2221) and the constraint records of write_constraint
(dio.c:19-59, IOVER 4) in XDR encoding, decoded here independently with struct
('>' big-endian) to pin the field order; a round trip restores the domain.  CPU
only (detection and the half step run on the host)."""
import struct

import numpy as np
import pytest

import solfec
from scenes import lattice


def decode(path):
    b = open(path, "rb").read()
    pos = 0

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from(">" + fmt, b, pos)
        pos += struct.calcsize(">" + fmt)
        return v

    (n,) = take("I")
    magic = b[pos:pos + n].decode()
    pos += 4 * ((n + 3) // 4)
    (iover,) = take("I")
    time, step = take("dd")
    nb, nc = take("II")
    bodies, cons = [], []
    for _ in range(nb):
        kind, conf, dofs = take("iii")
        q = take(f"{conf}d")
        v = take(f"{dofs}d")
        e = take("4d") if kind == 1 else ()
        bodies.append(dict(kind=kind, rot=q[:9], x=q[9:12], velo=v, energy=e))
    for _ in range(nc):
        cid, kind = take("Ii")
        R, U, V, point, base = take("3d"), take("3d"), take("3d"), take("3d"), take("9d")
        (merit,) = take("d")
        mid, sid = take("II")
        (mat,) = take("i")            # XDR short: 4 bytes
        (ssize,) = take("i")
        state = take(f"{ssize}d")
        area, gap = take("dd")
        spair = take("ii")
        cons.append(dict(id=cid, kind=kind, R=R, U=U, V=V, point=point, base=base, merit=merit, master=mid,
                         slave=sid, mat=mat, state=state, area=area, gap=gap, spair=spair))
    assert pos == len(b)
    return magic, iover, time, step, bodies, cons


def test_state_file_layout_and_round_trip(tmp_path):
    s = lattice(125, L=0.45, seed=3)          # overlapping neighbours: contacts at the first step begin
    c, bd = s.contacts(), s.bodies()
    rng = np.random.default_rng(0)
    R = rng.normal(size=c["R"].shape)
    U = rng.normal(size=c["U"].shape)
    s.lib.SOLFEC_Set_RU(s.sol, R.ctypes.data, U.ctypes.data)
    c = s.contacts()
    path = tmp_path / "state.xdr"
    s.write_state(path)
    magic, iover, time, step, bodies, cons = decode(path)
    assert magic == "SOLFEC-1.0 STATE" and iover == 4 and step == 1e-3
    assert len(bodies) == s.nbod and len(cons) == s.ncon > 100
    for k in (0, 7, 124, 125, 130):
        b = bodies[k]
        assert b["kind"] == (0 if bd["kind"][k] == 1 else 1)        # reference numbering: OBS 0, RIG 1
        np.testing.assert_array_equal(b["rot"], bd["rot"][9 * k:9 * k + 9])
        np.testing.assert_array_equal(b["x"], bd["x"][3 * k:3 * k + 3])
        np.testing.assert_array_equal(b["velo"], bd["velo"][6 * k:6 * k + 6])
    for i in (0, 1, len(cons) - 1):
        r = cons[i]
        assert r["id"] == i + 1 and r["kind"] == 0 and r["area"] == 1.0 and r["state"] == ()
        assert (r["master"], r["slave"]) == (c["master"][i] + 1, c["slave"][i] + 1)
        np.testing.assert_array_equal(r["R"], R[3 * i:3 * i + 3])
        np.testing.assert_array_equal(r["U"], U[3 * i:3 * i + 3])
        np.testing.assert_array_equal(r["base"], c["base"][9 * i:9 * i + 9])
        assert r["gap"] == c["gap"][i]
    # into a fresh domain of the same scene, before any step
    t = lattice(125, L=0.45, seed=3, steps_begin=0)
    t.read_state(path)
    tc, tb = t.contacts(), t.bodies()
    for k in ("master", "slave", "R", "U", "base", "gap", "friction"):
        np.testing.assert_array_equal(tc[k], c[k])
    for k in ("rot", "x", "velo"):
        np.testing.assert_array_equal(tb[k], bd[k])
    np.testing.assert_allclose(tc["mpnt"], c["mpnt"], atol=1e-15)
    assert t.time == s.time


def test_state_file_mismatch_is_refused(tmp_path):
    s = lattice(64, seed=2)
    path = tmp_path / "s.xdr"
    s.write_state(path)
    t = lattice(27, seed=2, steps_begin=0)
    with pytest.raises(RuntimeError, match="does not match"):
        t.read_state(path)
    (tmp_path / "junk").write_bytes(b"\x00\x00\x00\x10not a state file")
    with pytest.raises(RuntimeError, match="not a state file"):
        t.read_state(tmp_path / "junk")
    raw = path.read_bytes()
    (tmp_path / "cut").write_bytes(raw[:len(raw) - 10])
    u = lattice(64, seed=2, steps_begin=0)
    before = u.bodies()["x"].copy()
    with pytest.raises(RuntimeError):
        u.read_state(tmp_path / "cut")
    np.testing.assert_array_equal(u.bodies()["x"], before)    # a bad file leaves the domain as it was


def test_cohesion_state_round_trip(tmp_path):
    """Contacts created at time 0 with a cohesive material carry its cohesion
    (SURFACE_MATERIAL_Transfer, sps.c:264-282); the state file holds it as the
    surface material state of size 1 (sps.c:65-69, 286-294), and a fresh domain
    reads it back into the contacts."""
    s = lattice(125, L=0.45, seed=3, cohesion=0.7)
    c = s.contacts()
    assert len(c["master"]) > 100
    np.testing.assert_array_equal(c["cohesion"], 0.7)
    path = tmp_path / "state.xdr"
    s.write_state(path)
    _, _, _, _, _, cons = decode(path)
    assert all(r["state"] == (0.7,) for r in cons)
    t = lattice(125, L=0.45, seed=3, cohesion=0.7, steps_begin=0)
    t.read_state(path)
    np.testing.assert_array_equal(t.contacts()["cohesion"], 0.7)
    # the same file read into a domain whose material is not cohesive: no cohesion (sps.c:72-80)
    u = lattice(125, L=0.45, seed=3, steps_begin=0)
    u.read_state(path)
    np.testing.assert_array_equal(u.contacts()["cohesion"], 0.0)
