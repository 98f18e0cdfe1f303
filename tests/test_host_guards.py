"""Host-layer guards (CPU only): the broad-phase grid of the host loop stays
within int range when a body has gone far from the others, and an array size
out of range is refused with an error code instead of ending the process."""
import numpy as np
import pytest

import solfec


def pair_with_far_body(far):
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    for c in ((0.0, 0.0, 0.0), (0.099, 0.0, 0.0), (far, 0.0, 0.0)):
        solfec.BODY(s, "RIGID", solfec.SPHERE(c, 0.05, 1, 1), mat)
    return s


@pytest.mark.parametrize("far", [1e3, 1e9, 3e11, 1e200])
def test_grid_with_a_distant_body(far):
    """One pair of touching spheres and a third sphere `far` metres away.  The
    grid's dimensions are ext / cell + 1 per axis (ext / cell ~ 1e10 at 1e9 m):
    sized in int arithmetic they overflowed, and the wrapped cell count asked
    for an allocation of ~1e19 bytes.  Now the grid coarsens in double until its
    count is at most 4 cells per body, and detection finds the one contact."""
    s = pair_with_far_body(far)
    assert s.step_begin() == 0, s.lib.SOLFEC_Last_Error(s.sol).decode()
    c = s.contacts()
    assert len(c["master"]) == 1
    assert sorted((int(c["master"][0]), int(c["slave"][0]))) == [0, 1]
    assert c["gap"][0] == pytest.approx(-0.001, abs=1e-12)


def test_non_finite_positions_are_an_error():
    s = pair_with_far_body(np.inf)
    assert s.step_begin() != 0
    assert "not finite" in s.lib.SOLFEC_Last_Error(s.sol).decode()


def test_state_file_with_an_absurd_contact_count_is_refused(tmp_path):
    """A state file whose header claims 2^31 - 1 contacts is refused (the
    reader checks the count against the file length before sizing anything)."""
    import struct
    s = pair_with_far_body(1.0)
    s.step_begin()
    path = tmp_path / "s.xdr"
    s.write_state(path)
    raw = bytearray(path.read_bytes())
    (n,) = struct.unpack_from(">I", raw, 0)
    at = 4 + 4 * ((n + 3) // 4) + 4 + 16 + 4       # magic, IOVER, time, step, body count
    struct.pack_into(">i", raw, at, 0x7fffffff)
    (tmp_path / "big").write_bytes(bytes(raw))
    struct.pack_into(">i", raw, at, -5)
    (tmp_path / "neg").write_bytes(bytes(raw))
    t = pair_with_far_body(1.0)
    for name in ("big", "neg"):
        with pytest.raises(RuntimeError, match="does not match"):
            t.read_state(tmp_path / name)
