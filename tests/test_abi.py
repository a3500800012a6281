"""The C-ABI library loads without a GPU, exports every symbol include/gradsolve_amd.h
declares, and rejects invalid arguments on the host before any device work."""
import os
import re

import pytest

from gradsolve_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gradsolve_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gsx_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = N.lib()
    names = _declared()
    assert "gsx_tsit5_fwd" in names and "gsx_tsit5_adjoint" in names
    for name in names:
        assert hasattr(L, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_queries():
    assert N.call("gsx_version") >= 100
    assert N.call("gsx_ckpt_rows", N.FIELD_LORENZ, 3) == 3
    assert N.call("gsx_ckpt_slots", 100, 8) == 13


def test_invalid_arguments_raise_before_launch():
    with pytest.raises(ValueError, match="not built"):
        N.call("gsx_tsit5_fwd", 99, 4, 3, 1, None, None, None, 0.0, 1.0, None, 1e-8, 1e-8,
               10, None, None, None, None, None, 0, None, 1, 0, None)
    with pytest.raises(ValueError, match="expects P=1"):
        N.call("gsx_tsit5_fwd", N.FIELD_LORENZ, 4, 3, 2, None, None, None, 0.0, 1.0, None,
               1e-8, 1e-8, 10, None, None, None, None, None, 0, None, 1, 0, None)
    with pytest.raises(ValueError, match="null"):
        N.call("gsx_tsit5_fwd", N.FIELD_LORENZ, 4, 3, 1, None, None, None, 0.0, 1.0, None,
               1e-8, 1e-8, 10, None, None, None, None, None, 0, None, 1, 0, None)
    with pytest.raises(ValueError, match="ckpt_interval"):
        N.call("gsx_tsit5_adjoint", N.FIELD_LORENZ, 4, 3, 1, None, None, 1, 0.0, 1, 4, 1, 1,
               64, 1, 1, 1, 0, None)
    with pytest.raises(ValueError, match="stiff"):
        N.call("gsx_tsit5_fwd", N.FIELD_ROBERTSON, 4, 3, 3, None, None, None, 0.0, 1.0, None,
               1e-8, 1e-8, 10, None, None, None, None, None, 0, None, 1, 0, None)
    with pytest.raises(ValueError, match="flags"):
        N.call("gsx_tsit5_fwd", N.FIELD_LORENZ, 0, 3, 1, None, None, None, 0.0, 1.0, None,
               1e-8, 1e-8, 10, None, None, None, None, None, 0, None, 1, 8, None)


@pytest.mark.parametrize("entry,field,P", [("gsx_tsit5_fwd", N.FIELD_LORENZ, 1),
                                           ("gsx_rb23_fwd", N.FIELD_ROBERTSON, 3),
                                           ("gsx_rodas5p_fwd", N.FIELD_ROBERTSON, 3)])
def test_forward_counts_are_optional_without_a_tape(entry, field, P):
    """The reference's forward handler binds (y0, p) -> (y_final, nsteps) with nothing else
    (gradsolve/cuda/kernels/tsit5_template.cu.in:127-178): without a tape the counts and the
    status may be null. A recording forward needs n_acc and status. Checked on the host's
    argument validation (the call stops at the invalid flags, after the pointer checks)."""
    fake = 4096  # never dereferenced: validation ends before any device work
    with pytest.raises(ValueError, match="flags"):
        N.call(entry, field, 4, 3, P, None, fake, fake, 0.0, 1.0, None, 1e-8, 1e-8, 10,
               fake, None, None, None, None, 0, None, 1, 64, None)
    with pytest.raises(ValueError, match="needs n_acc and status"):
        N.call(entry, field, 4, 3, P, None, fake, fake, 0.0, 1.0, None, 1e-8, 1e-8, 10,
               fake, fake, None, None, fake, 16, None, 1, 0, None)
