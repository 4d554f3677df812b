"""Volume ingestion (er_mhd_info / er_mhd_read / er_raw_read), host only.

Files are written here with numpy and read back through the library; the expected voxels
follow from the documented mapping (uint8 / uint16 kept, other integers shifted by their
minimum, floating point rescaled from [min, max] to [0, 65535])."""
import numpy as np
import pytest

import exposure_render as er
from exposure_render import capi


def _write_mhd(tmp_path, name, arr, met_type, spacing=(0.5, 0.75, 1.25), msb=False, local=False, extra=""):
    rz, ry, rx = arr.shape
    data = arr.astype(arr.dtype.newbyteorder(">" if msb else "<")).tobytes()
    header = (f"ObjectType = Image\nNDims = 3\nBinaryData = True\n"
              f"BinaryDataByteOrderMSB = {'True' if msb else 'False'}\nCompressedData = False\n"
              f"DimSize = {rx} {ry} {rz}\nElementSpacing = {spacing[0]} {spacing[1]} {spacing[2]}\n{extra}"
              f"ElementType = {met_type}\n")
    mhd = tmp_path / f"{name}.mhd"
    if local:
        mhd.write_bytes((header + "ElementDataFile = LOCAL\n").encode() + data)
    else:
        (tmp_path / f"{name}.raw").write_bytes(data)
        mhd.write_text(header + f"ElementDataFile = {name}.raw\n")
    return str(mhd)


def test_mhd_uint16_is_taken_as_is(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 65536, size=(5, 6, 7), dtype=np.uint16)
    vox, spacing, off, sc = er.load_mhd(_write_mhd(tmp_path, "u16", a, "MET_USHORT"))
    assert np.array_equal(vox, a) and (off, sc) == (0.0, 1.0)
    assert np.allclose(spacing, (0.5, 0.75, 1.25))


@pytest.mark.parametrize("local", [False, True])
def test_mhd_int16_big_endian_is_shifted_by_its_minimum(tmp_path, local):
    rng = np.random.default_rng(2)
    a = rng.integers(-1024, 3072, size=(4, 9, 3), dtype=np.int16)
    a[0, 0, 0] = -1024
    vox, _, off, sc = er.load_mhd(_write_mhd(tmp_path, "ct", a, "MET_SHORT", msb=True, local=local))
    assert off == -1024.0 and sc == 1.0
    assert np.array_equal(vox.astype(np.int32), a.astype(np.int32) + 1024)


def test_mhd_uint8_and_float(tmp_path):
    rng = np.random.default_rng(3)
    b = rng.integers(0, 256, size=(3, 4, 5), dtype=np.uint8)
    vox, _, _, _ = er.load_mhd(_write_mhd(tmp_path, "u8", b, "MET_UCHAR"))
    assert np.array_equal(vox, b.astype(np.uint16))
    f = rng.uniform(-2.0, 6.0, size=(3, 4, 5)).astype(np.float32)
    vox, _, off, sc = er.load_mhd(_write_mhd(tmp_path, "f32", f, "MET_FLOAT"))
    lo, hi = float(f.min()), float(f.max())
    assert off == pytest.approx(lo) and sc == pytest.approx(65535.0 / (hi - lo))
    want = np.floor((f.astype(np.float64) - lo) * (65535.0 / (hi - lo)) + 0.5)
    assert np.abs(vox.astype(np.float64) - want).max() <= 1.0
    assert vox.min() == 0 and vox.max() == 65535


def test_raw_with_header_bytes_and_tail_data(tmp_path):
    rng = np.random.default_rng(4)
    a = rng.integers(0, 4096, size=(6, 5, 4), dtype=np.uint16)
    p = tmp_path / "vol.raw"
    p.write_bytes(b"\x07" * 100 + a.tobytes())
    vox, off, sc = er.load_raw(str(p), (4, 5, 6), capi.ELEMENT_UINT16, header_bytes=100)
    assert np.array_equal(vox, a)
    vox, _, _ = er.load_raw(str(p), (4, 5, 6), capi.ELEMENT_UINT16, header_bytes=-1)   # data at the end
    assert np.array_equal(vox, a)


def test_ingestion_errors(tmp_path):
    a = np.zeros((2, 2, 2), dtype=np.uint16)
    with pytest.raises(er.ExposureRenderError):
        er.load_mhd(str(tmp_path / "missing.mhd"))
    with pytest.raises(er.ExposureRenderError):   # compressed
        er.load_mhd(_write_mhd(tmp_path, "z", a, "MET_USHORT", extra="CompressedData = True\n"))
    with pytest.raises(er.ExposureRenderError):   # unsupported element type
        er.load_mhd(_write_mhd(tmp_path, "l", a, "MET_LONG_LONG"))
    p = tmp_path / "short.raw"
    p.write_bytes(a.tobytes()[:-2])
    with pytest.raises(er.ExposureRenderError):   # truncated data
        er.load_raw(str(p), (2, 2, 2), capi.ELEMENT_UINT16)
    mhd = tmp_path / "nd4.mhd"
    mhd.write_text("NDims = 4\nDimSize = 2 2 2 2\nElementType = MET_USHORT\nElementDataFile = short.raw\n")
    with pytest.raises(er.ExposureRenderError):
        er.load_mhd(str(mhd))
