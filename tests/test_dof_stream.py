"""DoF vectors in the reference's binary stream format
(DiscreteFunctionDefault::write / read, dune/fem/function/common/
discretefunction_inline.hh:213-305, through StandardOutStream,
dune/fem/io/streams/standardstreams.hh:60-170).  The byte layout below is
the reference's primitive sequence: unsigned version id, int space id
(DFSpaceIdentifier, space/common/discretefunctionspace.hh:95-108), string
(unsigned length + bytes), int32 size, doubles, all little-endian."""
import struct

import numpy as np
import pytest

import dunefem_amd as fa

VERSION_2_12 = (2 << 24) | (12 << 16)


def test_space_ids_are_the_reference_enum_positions():
    # CombinedSpace, DFAdapter, DGSpace, FiniteVolumeSpace, FourierSpace,
    # GenericSpace, LagrangeSpace (6), RannacherTurekSpace, LegendreDGSpace (8)
    assert fa.SPACE_LAGRANGE == 6 and fa.SPACE_LEGENDRE_DG == 8


def test_encoded_bytes_follow_the_reference_write_sequence():
    data = fa.encode_dofs([1.5, -2.0, 0.1], name="u_h", space=fa.SPACE_LEGENDRE_DG)
    ref = (struct.pack("<I", VERSION_2_12) + struct.pack("<i", 8) + struct.pack("<I", 3) + b"u_h"
           + struct.pack("<i", 3) + struct.pack("<3d", 1.5, -2.0, 0.1))
    assert data == ref
    empty = fa.encode_dofs([], space=fa.SPACE_LAGRANGE)
    assert empty == struct.pack("<IiIi", VERSION_2_12, 6, 0, 0)


def test_round_trip_and_name():
    rng = np.random.default_rng(3)
    x = rng.normal(size=1000)
    x[3] = np.inf
    x[7] = -0.0
    data = fa.encode_dofs(x, name="solution", space=fa.SPACE_LAGRANGE)
    y, name = fa.decode_dofs(data, x.size, space=fa.SPACE_LAGRANGE)
    assert name == "solution"
    assert np.array_equal(y, x) and np.signbit(y[7])


def test_decode_checks_like_the_reference_read():
    x = np.arange(4.0)
    data = fa.encode_dofs(x, name="u")
    with pytest.raises(fa.FemError, match="different size"):
        fa.decode_dofs(data, 5)
    with pytest.raises(fa.FemError, match="different space"):
        fa.decode_dofs(data, 4, space=fa.SPACE_LAGRANGE)
    with pytest.raises(fa.FemError, match="truncated"):
        fa.decode_dofs(data[:-3], 4)
    old = struct.pack("<I", (0 << 24) | (9 << 16)) + data[4:]
    with pytest.raises(fa.FemError, match="outdated"):
        fa.decode_dofs(old, 4)
    # files before 1.5.0 carry no space id: it is not read, nor checked
    v14 = struct.pack("<I", (1 << 24) | (4 << 16)) + struct.pack("<I", 1) + b"u" + struct.pack("<i", 4) \
        + struct.pack("<4d", *x)
    y, name = fa.decode_dofs(v14, 4, space=fa.SPACE_LAGRANGE)
    assert name == "u" and np.array_equal(y, x)


@pytest.mark.gpu
def test_device_vector_write_read(ctx, tmp_path):
    rng = np.random.default_rng(4)
    for prec in (fa.F64, fa.F32):
        x = rng.normal(size=777)
        v = fa.DofVector(ctx, x.size, prec, x)
        path = tmp_path / ("v%d.bin" % prec)
        v.write(path, name="w", space=fa.SPACE_LEGENDRE_DG)
        xs = v.download()
        assert path.read_bytes() == fa.encode_dofs(xs, name="w", space=fa.SPACE_LEGENDRE_DG)
        u = fa.DofVector(ctx, x.size, prec)
        assert u.read(path) == "w"
        assert np.array_equal(u.download(), xs)
        with pytest.raises(fa.FemError):
            fa.DofVector(ctx, x.size + 1, prec).read(path)
