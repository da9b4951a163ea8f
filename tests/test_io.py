"""File I/O and the command line (biomedisa_amd.io, biomedisa_amd.__main__): TIFF
and NRRD files written and read back, files built byte by byte from the format
specifications (big-endian TIFF with horizontal differencing, strips stored out
of line, BigTIFF; NRRD with gzip and comments), the result file names of
interpolation.py:115-133, and the reference's command-line defaults."""
import gzip
import struct
import zlib

import numpy as np
import pytest

from biomedisa_amd import io as bio
from biomedisa_amd.__main__ import parser


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32, np.float32,
                                   np.float64])
@pytest.mark.parametrize("compress", [False, True])
def test_tiff_round_trip(tmp_path, dtype, compress):
    rng = np.random.default_rng(1)
    arr = (rng.random((5, 7, 9)) * 200 - (50 if np.dtype(dtype).kind != "u" else 0)).astype(dtype)
    path = str(tmp_path / "v.tif")
    bio.save_data(path, arr, compress)
    got, ext = bio.load_data(path)
    assert ext == ".tif" and got.dtype == arr.dtype
    np.testing.assert_array_equal(got, arr)


def test_bigtiff_round_trip(tmp_path):
    arr = np.arange(3 * 4 * 5, dtype=np.uint16).reshape(3, 4, 5)
    path = str(tmp_path / "b.tiff")
    bio._write_tiff(path, arr, compress=True, bigtiff=True)
    with open(path, "rb") as f:
        assert f.read(4) == b"II+\0"
    np.testing.assert_array_equal(bio.load_data(path)[0], arr)


def _ifd_be(entries, next_off, base):
    """A classic big-endian IFD at file offset base; values longer than 4 bytes
    are placed after it.  entries: (tag, type, fmt, values)."""
    n = len(entries)
    head = struct.pack(">H", n)
    extra = b""
    extra_at = base + 2 + 12 * n + 4
    body = b""
    for tag, typ, fmt, vals in entries:
        payload = struct.pack(">" + fmt * len(vals), *vals)
        if len(payload) <= 4:
            body += struct.pack(">HHI", tag, typ, len(vals)) + payload.ljust(4, b"\0")
        else:
            body += struct.pack(">HHII", tag, typ, len(vals), extra_at + len(extra))
            extra += payload
    return head + body + struct.pack(">I", next_off) + extra


def test_tiff_big_endian_predictor_strips(tmp_path):
    """Two pages of big-endian uint16 with horizontal differencing (predictor 2)
    and Deflate; the second page in two strips (offsets stored out of line)."""
    rng = np.random.default_rng(2)
    vol = rng.integers(0, 65535, (2, 6, 5)).astype(np.uint16)

    def strip(rows):
        d = rows.astype(np.int64).copy()
        d[:, 1:] = rows[:, 1:].astype(np.int64) - rows[:, :-1].astype(np.int64)
        return zlib.compress((d % 65536).astype(">u2").tobytes())

    s0 = strip(vol[0])
    s1a, s1b = strip(vol[1, :4]), strip(vol[1, 4:])
    out = bytearray(b"MM\0\x2a\0\0\0\0")
    p0 = len(out)
    out += s0 + s1a + s1b
    common = [(256, 3, "H", [5]), (257, 3, "H", [6]), (258, 3, "H", [16]), (259, 3, "H", [8]),
              (262, 3, "H", [1]), (277, 3, "H", [1]), (317, 3, "H", [2]), (339, 3, "H", [1])]
    ifd0_at = len(out)
    ifd0_len = len(_ifd_be(common + [(273, 4, "I", [p0]), (278, 3, "H", [6]), (279, 4, "I", [len(s0)])],
                           0, ifd0_at))
    ifd1_at = ifd0_at + ifd0_len
    ifd0 = _ifd_be(sorted(common + [(273, 4, "I", [p0]), (278, 3, "H", [6]), (279, 4, "I", [len(s0)])]),
                   ifd1_at, ifd0_at)
    ifd1 = _ifd_be(sorted(common + [(273, 4, "I", [p0 + len(s0), p0 + len(s0) + len(s1a)]),
                                    (278, 3, "H", [4]), (279, 4, "I", [len(s1a), len(s1b)])]), 0, ifd1_at)
    out += ifd0 + ifd1
    struct.pack_into(">I", out, 4, ifd0_at)
    path = tmp_path / "be.tif"
    path.write_bytes(bytes(out))
    got = bio.load_data(str(path))[0]
    assert got.dtype == np.uint16 and got.dtype.isnative
    np.testing.assert_array_equal(got, vol)


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32, np.float32, np.float64])
@pytest.mark.parametrize("compress", [False, True])
def test_nrrd_round_trip(tmp_path, dtype, compress):
    arr = (np.random.default_rng(3).random((4, 5, 6)) * 100).astype(dtype)
    path = str(tmp_path / "v.nrrd")
    bio.save_data(path, arr, compress)
    got, ext = bio.load_data(path)
    assert ext == ".nrrd"
    np.testing.assert_array_equal(got, arr)


def test_nrrd_hand_built_big_endian_gzip(tmp_path):
    """sizes are listed fastest axis first: 'sizes: 4 3 2' is a (2, 3, 4) array."""
    vol = np.arange(24, dtype=np.int16).reshape(2, 3, 4) - 7
    head = ("NRRD0005\n# a comment\ntype: short\ndimension: 3\nspace dimension: 3\n"
            "sizes: 4 3 2\nendian: big\nencoding: gzip\nkey:=value\n\n")
    path = tmp_path / "h.nrrd"
    path.write_bytes(head.encode() + gzip.compress(vol.astype(">i2").tobytes()))
    got = bio.load_data(str(path))[0]
    assert got.dtype == np.int16
    np.testing.assert_array_equal(got, vol)


def test_npy_npz(tmp_path):
    arr = np.arange(60, dtype=np.uint8).reshape(3, 4, 5)
    for name in ("a.npy", "a.npz"):
        bio.save_data(str(tmp_path / name), arr)
        np.testing.assert_array_equal(bio.load_data(str(tmp_path / name))[0], arr)


def test_missing_and_unsupported(tmp_path, capsys):
    got, ext = bio.load_data(str(tmp_path / "none.tif"))
    assert got is None and ext == ".tif"
    assert "Error" in capsys.readouterr().out
    (tmp_path / "x.am").write_bytes(b"# AmiraMesh")
    with pytest.raises(NotImplementedError):
        bio.load_data(str(tmp_path / "x.am"))


def test_result_paths():
    """interpolation.py:115-133: 'final.' + the image's name (without .nii.gz)
    + the label file's extension; the optional results beside it."""
    p = bio.result_paths("/data/run/img.nii.gz", ".tif")
    assert p["regular"] == "/data/run/final.img.tif"
    assert p["smooth"] == "/data/run/final.img.smooth.tif"
    assert p["smooth_cleaned"] == "/data/run/final.img.smooth.cleaned.tif"
    assert p["cleaned_filled"] == "/data/run/final.img.cleaned.filled.tif"
    assert p["uncertainty"] == "/data/run/final.img.uncertainty.tif"
    p = bio.result_paths("scan.tif", ".nii.gz")
    assert p["regular"] == "final.scan.nii.gz"
    assert p["acwe"] == "final.scan.acwe.nii.gz"
    assert p["refined"] == "final.scan.refined.nii.gz"
    assert p["uncertainty"] == "final.scan.uncertainty.tif"


def test_cli_defaults_and_flags():
    """The reference's argument defaults (interpolation.py:286-348)."""
    a = parser().parse_args(["img.tif", "lab.tif"])
    assert (a.nbrw, a.sorw, a.smooth, a.clean, a.fill, a.compression) == (10, 4000, 0, None, None, True)
    assert (a.acwe, a.acwe_alpha, a.acwe_smooth, a.acwe_steps) == (False, 1.0, 1, 3)
    assert (a.ignore, a.only, a.allaxis, a.denoise, a.uncertainty) == ("none", "all", False, False, False)
    a = parser().parse_args(["i.tif", "l.tif", "-s", "-c", "-f", "-nc", "-allx", "-d", "-u", "-rh",
                             "-i", "2,5", "-o", "1"])
    assert (a.smooth, a.clean, a.fill, a.compression) == (100, 0.1, 0.9, False)
    assert (a.allaxis, a.denoise, a.uncertainty, a.return_hits, a.ignore, a.only) == \
        (True, True, True, True, "2,5", "1")
    a = parser().parse_args(["i.tif", "l.tif", "-s", "7", "-c", "0.3", "-f", "0.5"])
    assert (a.smooth, a.clean, a.fill) == (7, 0.3, 0.5)


def test_missing_file_returns_none(tmp_path, capsys):
    from biomedisa_amd.interpolation import smart_interpolation
    assert smart_interpolation(None, None, path_to_data=str(tmp_path / "no.tif"),
                               path_to_labels=str(tmp_path / "no_l.tif")) is None
    assert "Invalid image data" in capsys.readouterr().out
