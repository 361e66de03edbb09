"""GRASS elevation rasters (tg_host_read_grass; the reference's GIS input,
GisRasterHdr.C:26-60, GisBinFile.C:131-228, set_from_header GisApi.C:1455-1480).
Rasters are written here from the format's description -- a positional
'key: value' header and big-endian float32 rows, raw or compressed with row
offsets and zlib-deflated rows -- and read back bit for bit; the static
coefficients of a raster read from disk equal those of the same values in
memory."""
import struct
import zlib

import numpy as np
import pytest

from discrete_adjoint_amd import problems


def _write(root, name, elev, west=1000.0, south=2000.0, res=5.0, compressed=0, offsets=4,
           stored_rows=()):
    rows, cols = elev.shape
    d = root / "loc" / "PERMANENT"
    (d / "cellhd").mkdir(parents=True, exist_ok=True)
    (d / "fcell").mkdir(parents=True, exist_ok=True)
    hdr = (f"proj:       1\nzone:       13\nnorth:      {south + rows * res}\n"
           f"south:      {south}\neast:       {west + cols * res}\nwest:       {west}\n"
           f"cols:       {cols}\nrows:       {rows}\ne-w resol:  {res}\nn-s resol:  {res}\n"
           f"format:     -1\ncompressed: {compressed}\n")
    (d / "cellhd" / name).write_text(hdr)
    raw = [struct.pack(f">{cols}f", *elev[r].tolist()) for r in range(rows)]
    if not compressed:
        data = b"".join(raw)
    else:
        body = [(b"0" + r) if i in stored_rows else (b"1" + zlib.compress(r))
                for i, r in enumerate(raw)]
        head = 1 + (rows + 1) * offsets
        ptr, p = [], head
        for b in body:
            ptr.append(p)
            p += len(b)
        ptr.append(p)
        data = bytes([offsets]) + b"".join(v.to_bytes(offsets, "big") for v in ptr) + b"".join(body)
    (d / "fcell" / name).write_bytes(data)
    return str(root)


@pytest.mark.parametrize("compressed,offsets,stored", [(0, 4, ()), (1, 4, ()), (1, 8, ()),
                                                       (1, 4, (0, 3))])
def test_read_back(tmp_path, compressed, offsets, stored):
    rng = np.random.default_rng(7)
    elev = (rng.random((9, 13)) * 800.0).astype(np.float32)
    root = _write(tmp_path, "dem", elev, compressed=compressed, offsets=offsets,
                  stored_rows=stored)
    nrows, ncols, xmin, ymax, res, got = problems.read_grass(root, "loc", "PERMANENT", "dem")
    assert (nrows, ncols, xmin, ymax, res) == (9, 13, 1000.0, 2000.0 + 9 * 5.0, 5.0)
    assert np.array_equal(got.reshape(9, 13), elev)


def test_statics_from_disk_equal_in_memory(tmp_path):
    elev = problems.fractal_dem(65, 200.0, 0.8, 3).reshape(65, 65)
    root = _write(tmp_path, "frac", elev, west=0.0, south=0.0, res=5.0, compressed=1)
    dem = problems.read_grass(root, "loc", "PERMANENT", "frac")
    ref = (65, 65, 0.0, 65 * 5.0, 5.0, np.ascontiguousarray(elev.reshape(-1)))
    a = problems.static_coefs(dem, 40, 40, 5.0, 5.0, 30.0, 30.0)
    b = problems.static_coefs(ref, 40, 40, 5.0, 5.0, 30.0, 30.0)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_bad_rasters(tmp_path):
    elev = np.ones((4, 5), np.float32)
    root = _write(tmp_path, "ok", elev, compressed=1)
    d = tmp_path / "loc" / "PERMANENT"
    # truncated data
    (d / "fcell" / "ok").write_bytes((d / "fcell" / "ok").read_bytes()[:-3])
    with pytest.raises(ValueError, match="-5"):
        problems.read_grass(root, "loc", "PERMANENT", "ok")
    # header not starting with proj
    (d / "cellhd" / "ok").write_text("zone: 1\n")
    with pytest.raises(ValueError, match="-4"):
        problems.read_grass(root, "loc", "PERMANENT", "ok")
    # resolutions that differ
    _write(tmp_path, "aniso", elev)
    h = (d / "cellhd" / "aniso").read_text().replace("n-s resol:  5.0", "n-s resol:  7.0")
    (d / "cellhd" / "aniso").write_text(h)
    with pytest.raises(ValueError, match="-4"):
        problems.read_grass(root, "loc", "PERMANENT", "aniso")
    with pytest.raises(ValueError, match="-4"):
        problems.read_grass(root, "loc", "PERMANENT", "missing")
