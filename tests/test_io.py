"""File I/O and the command line (biomedisa_amd.io, biomedisa_amd.__main__): TIFF
and NRRD files written and read back, files built byte by byte from the format
specifications (big-endian TIFF with horizontal differencing, strips stored out
of line, BigTIFF; NRRD with gzip and comments), the result file names of
interpolation.py:115-133, and the reference's command-line defaults.  AmiraMesh
(HxByteRLE checked against a decoder written from its definition), NIfTI-1 and
MetaImage files likewise."""
import bz2
import gzip
import os
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
    (tmp_path / "x.nc").write_bytes(b"CDF")
    with pytest.raises(ValueError):
        bio.load_data(str(tmp_path / "x.nc"))
    (tmp_path / "x.raw3d").write_bytes(b"\0" * 8)
    with pytest.raises(NotImplementedError):
        bio.load_data(str(tmp_path / "x.raw3d"))
    (tmp_path / "x.am").write_bytes(b"# AmiraMesh")
    with pytest.raises(ValueError):
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


# --- HxByteRLE, AmiraMesh, NIfTI, MetaImage ---------------------------------

def _rle_decode_py(stream, n):
    """HxByteRLE by its definition: control byte c; c & 0x80: (c & 0x7f)
    literal bytes follow; otherwise the next byte repeats c times."""
    out, i = [], 0
    while len(out) < n:
        c = stream[i]
        i += 1
        k = c & 0x7F
        if c & 0x80:
            out.extend(stream[i:i + k])
            i += k
        else:
            out.extend([stream[i]] * k)
            i += 1
    return np.array(out, np.uint8)


def test_hxbyterle_known_stream_and_round_trip():
    stream = bytes([3, 7, 0x82, 1, 2, 0x7F, 0, 0x81, 9])
    want = np.array([7, 7, 7, 1, 2] + [0] * 127 + [9], np.uint8)
    np.testing.assert_array_equal(bio._rle_decode(stream, want.size), want)
    rng = np.random.default_rng(4)
    for n in (0, 1, 2, 3, 127, 128, 129, 1000, 50000):
        runs = rng.integers(1, 300, n)
        data = np.repeat(rng.integers(0, 4, n).astype(np.uint8), runs)[:n]
        noisy = rng.random(n) < 0.2
        data[noisy] = rng.integers(0, 256, int(noisy.sum()))
        enc = bio._rle_encode(data)
        assert len(enc) <= n + (n + 126) // 127
        np.testing.assert_array_equal(_rle_decode_py(enc, n), data)
        np.testing.assert_array_equal(bio._rle_decode(enc, n), data)
    noise = rng.integers(0, 256, 10000).astype(np.uint8)
    np.testing.assert_array_equal(bio._rle_decode(bio._rle_encode(noise), noise.size), noise)
    with pytest.raises(ValueError):
        bio._rle_decode(bytes([0x85, 1, 2]), 5)          # truncated literal record
    with pytest.raises(ValueError):
        bio._rle_decode(bytes([10, 1]), 5)               # stream longer than the lattice


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32, np.float32, np.float64])
@pytest.mark.parametrize("compress", [False, True])
def test_amira_round_trip(tmp_path, dtype, compress):
    arr = (np.random.default_rng(5).random((6, 7, 8)) * 100).astype(dtype)
    arr[2:4] = 0
    path = str(tmp_path / "v.am")
    bio.save_data(path, arr, compress)
    got, ext = bio.load_data(path)
    assert ext == ".am" and got.dtype == arr.dtype
    np.testing.assert_array_equal(got, arr)


_AM_LABEL_HEADER = b"""# AmiraMesh BINARY-LITTLE-ENDIAN 2.1


define Lattice 5 4 3

Parameters {
    Materials {
        Exterior {
            id 0
        }
        Bone {
            Color 0.8 0.7 0.2,
            id 1
        }
    }
    Content "5x4x3 byte, uniform coordinates",
    CoordType "uniform"
}

Lattice { byte Labels } @1(HxByteRLE,7)

# Data section follows"""


def test_amira_hand_built_files(tmp_path):
    vol = np.zeros((3, 4, 5), np.uint8)
    vol[1, 1:3, 1:4] = 1
    # HxByteRLE byte lattice with materials, the length in the stream definition
    enc = bio._rle_encode(vol)
    head = _AM_LABEL_HEADER.replace(b"@1(HxByteRLE,7)", b"@1(HxByteRLE,%d)" % len(enc))
    (tmp_path / "l.am").write_bytes(head + b"\n@1\n" + enc + b"\n")
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "l.am"))[0], vol)
    # big-endian ushort, raw
    u = (np.arange(60, dtype=np.uint16) * 1000).reshape(3, 4, 5)
    head = (b"# AmiraMesh 3D BINARY 2.0\n\ndefine Lattice 5 4 3\n\nLattice { ushort Data } @1\n\n"
            b"# Data section follows")
    (tmp_path / "be.am").write_bytes(head + b"\n@1\n" + u.astype(">u2").tobytes() + b"\n")
    got = bio.load_data(str(tmp_path / "be.am"))[0]
    assert got.dtype == np.uint16
    np.testing.assert_array_equal(got, u)
    # HxZip float
    f = np.linspace(-1, 1, 60, dtype=np.float32).reshape(3, 4, 5)
    z = zlib.compress(f.tobytes())
    head = (b"# Avizo BINARY-LITTLE-ENDIAN 2.1\ndefine Lattice 5 4 3\nLattice { float Data } @1(HxZip,%d)\n"
            b"# Data section follows" % len(z))
    (tmp_path / "z.am").write_bytes(head + b"\n@1\n" + z + b"\n")
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "z.am"))[0], f)


def test_amira_result_keeps_the_label_header(tmp_path):
    """Results are written with the label file's header (materials kept) and the
    new encoded length, as amira_data_stream.py:476-515 does."""
    vol = np.zeros((3, 4, 5), np.uint8)
    vol[0, :2] = 1
    enc = bio._rle_encode(vol)
    head = _AM_LABEL_HEADER.replace(b"@1(HxByteRLE,7)", b"@1(HxByteRLE,%d)" % len(enc))
    (tmp_path / "l.am").write_bytes(head + b"\n@1\n" + enc + b"\n")
    header = bio.am_header(str(tmp_path / "l.am"))
    res = np.zeros((3, 4, 5), np.uint8)
    res[1:] = 1
    res[2, 3, 4] = 0
    bio.save_data(str(tmp_path / "r.am"), res, True, header=header)
    raw = (tmp_path / "r.am").read_bytes()
    assert b"Bone" in raw and b"Color 0.8 0.7 0.2" in raw
    assert b"@1(HxByteRLE,%d)" % len(bio._rle_encode(res)) in raw
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "r.am"))[0], res)
    with pytest.raises(ValueError):
        bio.save_data(str(tmp_path / "bad.am"), np.zeros((2, 2, 2), np.uint8), True, header=header)


@pytest.mark.parametrize("name", ["v.nii", "v.nii.gz"])
@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.int16, np.uint16, np.int32, np.float32, np.float64])
def test_nifti_round_trip(tmp_path, name, dtype):
    arr = (np.random.default_rng(6).random((4, 5, 6)) * 100).astype(dtype)
    path = str(tmp_path / name)
    bio.save_data(path, arr)
    got, ext = bio.load_data(path)
    assert ext == (".nii.gz" if name.endswith(".gz") else ".nii") and got.dtype == arr.dtype
    np.testing.assert_array_equal(got, arr)


def test_nifti_hand_built(tmp_path):
    """NIfTI-1: x fastest in the file; dim[1..3] = X, Y, Z; vox_offset; scl_slope
    and scl_inter applied; big-endian headers recognised by sizeof_hdr."""
    vol = np.arange(2 * 3 * 4, dtype=np.uint16).reshape(2, 3, 4)
    for bo in "<>":
        hdr = bytearray(400)
        struct.pack_into(bo + "i", hdr, 0, 348)
        struct.pack_into(bo + "8h", hdr, 40, 3, 4, 3, 2, 1, 1, 1, 1)
        struct.pack_into(bo + "2h", hdr, 70, 512, 16)
        struct.pack_into(bo + "3f", hdr, 108, 400.0, 2.0, -1.0)
        hdr[344:348] = b"n+1\0"
        (tmp_path / "h.nii").write_bytes(bytes(hdr) + vol.astype(bo + "u2").tobytes())
        got = bio.load_data(str(tmp_path / "h.nii"))[0]
        np.testing.assert_array_equal(got, vol.astype(np.float64) * 2.0 - 1.0)


@pytest.mark.parametrize("name", ["v.mha", "v.mhd"])
@pytest.mark.parametrize("compress", [False, True])
def test_metaimage_round_trip(tmp_path, name, compress):
    arr = (np.random.default_rng(7).random((4, 5, 6)) * 100).astype(np.int16)
    path = str(tmp_path / name)
    bio.save_data(path, arr, compress)
    np.testing.assert_array_equal(bio.load_data(path)[0], arr)
    if name.endswith(".mhd"):
        assert (tmp_path / ("v.zraw" if compress else "v.raw")).exists()


def test_metaimage_hand_built(tmp_path):
    vol = (np.arange(24, dtype=np.int16) - 12).reshape(2, 3, 4)
    head = (b"ObjectType = Image\nNDims = 3\nDimSize = 4 3 2\nElementType = MET_SHORT\n"
            b"BinaryDataByteOrderMSB = True\nElementDataFile = LOCAL\n")
    (tmp_path / "h.mha").write_bytes(head + vol.astype(">i2").tobytes())
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "h.mha"))[0], vol)
    (tmp_path / "d.raw").write_bytes(zlib.compress(vol.astype("<i2").tobytes()))
    head = (b"ObjectType = Image\nNDims = 3\nDimSize = 4 3 2\nElementType = MET_SHORT\n"
            b"CompressedData = True\nElementDataFile = d.raw\n")
    (tmp_path / "h.mhd").write_bytes(head)
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "h.mhd"))[0], vol)


# --- slice stacks (directories, zip archives; TIFF and PNG slices) ---------

def _png_filtered(img, filters):
    """A PNG whose rows use the given filter types, filtered here from the PNG
    specification (section 9), independent of the reader."""
    img = np.ascontiguousarray(img)
    H, W = img.shape[:2]
    raw = img.astype(img.dtype.newbyteorder(">")).view(np.uint8).reshape(H, -1).astype(np.int64)
    bpp = img.dtype.itemsize * (img.shape[2] if img.ndim == 3 else 1)
    rows = []
    for r in range(H):
        t = filters[r % len(filters)]
        out = []
        for i in range(raw.shape[1]):
            x = raw[r, i]
            a = raw[r, i - bpp] if i >= bpp else 0
            b = raw[r - 1, i] if r > 0 else 0
            c = raw[r - 1, i - bpp] if r > 0 and i >= bpp else 0
            if t == 0:
                p = 0
            elif t == 1:
                p = a
            elif t == 2:
                p = b
            elif t == 3:
                p = (a + b) // 2
            else:
                q = a + b - c
                pa, pb, pc = abs(q - a), abs(q - b), abs(q - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            out.append((x - p) % 256)
        rows.append(bytes([t]) + bytes(out))
    depth = 8 * img.dtype.itemsize
    ctype = 4 if img.ndim == 3 else 0

    def chunk(typ, body):
        return struct.pack(">I", len(body)) + typ + body + struct.pack(">I", zlib.crc32(typ + body) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b""))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_png_every_filter_type(dtype):
    rng = np.random.default_rng(8)
    img = rng.integers(0, np.iinfo(dtype).max, (7, 9)).astype(dtype)
    np.testing.assert_array_equal(bio._read_png(_png_filtered(img, [0, 1, 2, 3, 4])), img)
    np.testing.assert_array_equal(bio._read_png(_png_filtered(img, [4, 3, 2, 1])), img)
    ga = np.stack([img, img[::-1]], axis=2)            # gray + alpha: the alpha is dropped
    np.testing.assert_array_equal(bio._read_png(_png_filtered(ga, [4, 1])), img)
    np.testing.assert_array_equal(bio._read_png(bio._png_bytes(img)), img)


def test_slice_directory_natural_order(tmp_path):
    vol = (np.random.default_rng(9).random((12, 5, 6)) * 250).astype(np.uint8)
    d = tmp_path / "stack"
    d.mkdir()
    for k in range(12):
        (d / ("img_%d.png" % k)).write_bytes(_png_filtered(vol[k], [k % 5]))
    (d / ".hidden.png").write_bytes(b"not an image")
    (d / "notes.txt").write_text("ignored")
    got, ext = bio.load_data(str(d))
    assert ext == ""
    np.testing.assert_array_equal(got, vol)             # img_2 before img_10
    hdr = bio.read_header(str(d))
    assert hdr["slices"][:3] == ["img_0.png", "img_1.png", "img_2.png"] and hdr["dtype"] == np.uint8


def test_slice_zip_and_results(tmp_path):
    import zipfile
    vol = (np.random.default_rng(10).random((3, 4, 5)) * 1000).astype(np.uint16)
    with zipfile.ZipFile(tmp_path / "s.zip", "w") as z:
        for k in range(3):
            bio._write_tiff(str(tmp_path / "t.tif"), vol[k], compress=bool(k % 2))
            z.write(tmp_path / "t.tif", "slice%02d.tif" % k)
    got, ext = bio.load_data(str(tmp_path / "s.zip"))
    assert ext == ".zip"
    np.testing.assert_array_equal(got, vol)
    hdr = bio.read_header(str(tmp_path / "s.zip"))
    res = (vol > 500).astype(np.uint8)
    # results keep the label stack's names and sample type (biomedisa_helper.py:758-783)
    bio.save_data(str(tmp_path / "final.s.zip"), res, True, header=hdr)
    assert sorted(zipfile.ZipFile(tmp_path / "final.s.zip").namelist()) == ["slice00.tif", "slice01.tif",
                                                                           "slice02.tif"]
    back = bio.load_data(str(tmp_path / "final.s.zip"))[0]
    assert back.dtype == np.uint16
    np.testing.assert_array_equal(back, res)
    bio.save_data(str(tmp_path / "final.dir"), res, True, header=hdr, file_type="")
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "final.dir"))[0], res)


def test_slice_errors(tmp_path):
    d = tmp_path / "mixed"
    d.mkdir()
    (d / "a.png").write_bytes(bio._png_bytes(np.zeros((4, 5), np.uint8)))
    (d / "b.png").write_bytes(bio._png_bytes(np.zeros((4, 6), np.uint8)))
    with pytest.raises(ValueError):
        bio.load_data(str(d))
    e = tmp_path / "jpeg"
    e.mkdir()
    (e / "x.jpg").write_bytes(b"\xff\xd8\xff")
    with pytest.raises(NotImplementedError):
        bio.load_data(str(e))
    f = tmp_path / "dicom"
    f.mkdir()
    (f / "x.dcm").write_bytes(b"DICM")
    with pytest.raises(ValueError):
        bio.load_data(str(f))
    with pytest.raises(ValueError):
        bio._read_png(b"\x89PNG\r\n\x1a\n" + b"\0" * 20)


def _mrc_bytes(arr, mode, bo="<", stamp=True, ext=b""):
    """An MRC2014 file assembled field by field: words 1-4 (nx, ny, nz, mode),
    word 24 (extended header bytes), the machine stamp at byte 212."""
    nz, ny, nx = arr.shape
    hdr = bytearray(1024)
    struct.pack_into(bo + "4i", hdr, 0, nx, ny, nz, mode)
    struct.pack_into(bo + "i", hdr, 92, len(ext))
    hdr[208:212] = b"MAP "
    if stamp:
        hdr[212:214] = b"\x44\x44" if bo == "<" else b"\x11\x11"
    return bytes(hdr) + ext + arr.astype(arr.dtype.newbyteorder(bo)).tobytes()


@pytest.mark.parametrize("mode,dtype", [(0, np.int8), (1, np.int16), (2, np.float32), (6, np.uint16),
                                        (12, np.float16)])
@pytest.mark.parametrize("bo,stamp", [("<", True), (">", True), ("<", False)])
def test_mrc_read_flips_y(tmp_path, mode, dtype, bo, stamp):
    rng = np.random.default_rng(mode)
    arr = (rng.random((3, 5, 7)) * 100).astype(dtype)
    p = tmp_path / "v.mrc"
    p.write_bytes(_mrc_bytes(arr, mode, bo, stamp, ext=b"x" * 80))
    got, ext = bio.load_data(str(p))
    assert ext == ".tif"                       # results of .mrc labels are written as TIFF
    assert got.dtype == np.dtype(dtype) and got.dtype.isnative
    np.testing.assert_array_equal(got, arr[:, ::-1, :])


def test_mrc_rejects_unknown_mode_and_short_data(tmp_path):
    arr = np.zeros((2, 3, 4), np.float32)
    p = tmp_path / "v.mrc"
    p.write_bytes(_mrc_bytes(arr, 4))
    with pytest.raises(NotImplementedError):
        bio.load_data(str(p))
    p.write_bytes(_mrc_bytes(arr, 2)[:-8])
    with pytest.raises(ValueError):
        bio.load_data(str(p))


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.uint16])
def test_hdr_img_pair_round_trip(tmp_path, dtype):
    arr = (np.random.default_rng(1).random((4, 6, 9)) * 200).astype(dtype)
    p = tmp_path / "labels.hdr"
    bio.save_data(str(p), arr)
    assert (tmp_path / "labels.img").stat().st_size == arr.nbytes
    hdr = p.read_bytes()
    assert len(hdr) == 348 and hdr[344:348] == b"ni1\0"
    got, ext = bio.load_data(str(p))
    assert ext == ".hdr" and got.dtype == arr.dtype
    np.testing.assert_array_equal(got, arr)


@pytest.mark.parametrize("bo", ["<", ">"])
def test_analyze75_read(tmp_path, bo):
    """An Analyze 7.5 header (no magic) written field by field: sizeof_hdr,
    dim[8] at 40, datatype and bitpix at 70, vox_offset at 108; data in .img."""
    arr = np.arange(2 * 3 * 5, dtype=np.int16).reshape(2, 3, 5) - 7
    hdr = bytearray(348)
    struct.pack_into(bo + "i", hdr, 0, 348)
    struct.pack_into(bo + "8h", hdr, 40, 4, 5, 3, 2, 1, 0, 0, 0)
    struct.pack_into(bo + "2h", hdr, 70, 4, 16)
    struct.pack_into(bo + "f", hdr, 108, 16.0)
    (tmp_path / "a.hdr").write_bytes(bytes(hdr))
    (tmp_path / "a.img").write_bytes(b"\0" * 16 + arr.astype(arr.dtype.newbyteorder(bo)).tobytes())
    got, _ = bio.load_data(str(tmp_path / "a.hdr"))
    np.testing.assert_array_equal(got, arr)
    (tmp_path / "a.img").write_bytes(b"\0" * 20)
    with pytest.raises(ValueError):
        bio.load_data(str(tmp_path / "a.hdr"))


def _png_chunk(typ, body):
    return struct.pack(">I", len(body)) + typ + body + struct.pack(">I", zlib.crc32(typ + body) & 0xFFFFFFFF)


# R, G, B pixels and their gray values under 0.2989 R + 0.587 G + 0.114 B,
# truncated: worked by hand (white is 0.9999 x 255 = 254.97 -> 254)
_RGB_KNOWN = [((255, 0, 0), 76), ((0, 255, 0), 149), ((0, 0, 255), 29), ((255, 255, 255), 254),
              ((0, 0, 0), 0), ((10, 20, 30), 18)]


def _bmp_file(pixels_bgr_rows, bits, palette=None, top_down=False):
    """A BMP assembled field by field: 14-byte file header, 40-byte info
    header, optional palette (B, G, R, 0), rows padded to 4 bytes."""
    H = len(pixels_bgr_rows)
    W = len(pixels_bgr_rows[0]) // (bits // 8)
    stride = (W * bits // 8 + 3) // 4 * 4
    body = b"".join(bytes(r) + b"\0" * (stride - len(r)) for r in (pixels_bgr_rows if top_down
                                                                     else pixels_bgr_rows[::-1]))
    pal = b"".join(bytes(p) + b"\0" for p in (palette or []))
    off = 14 + 40 + len(pal)
    info = struct.pack("<IiiHHIIiiII", 40, W, -H if top_down else H, 1, bits, 0, len(body), 0, 0,
                       len(palette or []), 0)
    return b"BM" + struct.pack("<IHHI", off + len(body), 0, 0, off) + info + pal + body


@pytest.mark.parametrize("top_down", [False, True])
def test_bmp_24bit_gray_known_values(top_down):
    row0 = [c for rgb, _ in _RGB_KNOWN[:3] for c in rgb[::-1]]      # 3 pixels: 9 bytes, 3 of padding
    row1 = [c for rgb, _ in _RGB_KNOWN[3:] for c in rgb[::-1]]
    img = bio._read_bmp(_bmp_file([row0, row1], 24, top_down=top_down))
    assert img.dtype == np.uint8
    np.testing.assert_array_equal(img, [[76, 149, 29], [254, 0, 18]])


def test_bmp_8bit_palettes():
    idx = [[0, 1, 2, 3, 4], [4, 3, 2, 1, 0]]
    gray = [(v, v, v) for v in (0, 7, 100, 200, 255)]              # a gray palette, not a ramp
    np.testing.assert_array_equal(bio._read_bmp(_bmp_file(idx, 8, gray)),
                                  [[0, 7, 100, 200, 255], [255, 200, 100, 7, 0]])
    colour = [tuple(rgb[::-1]) for rgb, _ in _RGB_KNOWN[:5]]         # B, G, R entries
    want = [g for _, g in _RGB_KNOWN[:5]]
    np.testing.assert_array_equal(bio._read_bmp(_bmp_file(idx, 8, colour)), [want, want[::-1]])
    with pytest.raises(ValueError):
        bio._read_bmp(_bmp_file([[0, 5]], 8, gray))                   # index beyond the palette


def test_bmp_rejects_other_forms():
    with pytest.raises(NotImplementedError):
        bio._read_bmp(_bmp_file([[1, 2, 3, 4]], 32))
    with pytest.raises(ValueError):
        bio._read_bmp(b"BX" + b"\0" * 60)


def test_bmp_slices_round_trip(tmp_path):
    arr = (np.random.default_rng(3).random((4, 7, 13)) * 255).astype(np.uint8)
    names = ["s_%d.bmp" % k for k in range(4)]
    d = tmp_path / "stack"
    bio.save_data(str(d), arr, header={"slices": names, "dtype": np.uint8}, file_type="")
    assert sorted(os.listdir(d)) == sorted(names)
    got, ext = bio.load_data(str(d))
    assert ext == "" and got.dtype == np.uint8
    np.testing.assert_array_equal(got, arr)


@pytest.mark.parametrize("depth", [8, 16])
def test_png_rgb_gray_known_values(depth):
    scale = 1 if depth == 8 else 257
    rgb = np.array([[c for c, _ in _RGB_KNOWN]], np.uint32) * scale     # 1 x 6 x 3
    dt = ">u2" if depth == 16 else "u1"
    raw = b"".join(b"\0" + r.astype(dt).tobytes() for r in rgb)
    png = (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", 6, 1, depth, 2, 0, 0, 0))
           + _png_chunk(b"IDAT", zlib.compress(raw)) + _png_chunk(b"IEND", b""))
    img = bio._read_png(png)
    if depth == 8:
        np.testing.assert_array_equal(img, [[g for _, g in _RGB_KNOWN]])
    else:   # white: 0.9999 x 65535 = 65528.4 -> 65528; (2570, 5140, 7710) -> 4664.3 -> 4664
        np.testing.assert_array_equal(img, [[19588, 38469, 7470, 65528, 0, 4664]])


def _dcm_el(g, e, value, vr=None, bo="<"):
    """One data element (PS3.5 7.1): implicit VR when vr is None; explicit VR
    with a 4-byte length for OB/OW/SQ/UN, else a 2-byte length."""
    if vr is None:
        return struct.pack(bo + "HHI", g, e, len(value)) + value
    if vr in (b"OB", b"OW", b"SQ", b"UN"):
        return struct.pack(bo + "HH2sHI", g, e, vr, 0, len(value)) + value
    return struct.pack(bo + "HH2sH", g, e, vr, len(value)) + value


def _dcm_image_elements(rows, cols, alloc, stored, signed, pixels, bo="<", explicit=True, spp=1,
                        planar=0, slope=None, inter=None):
    us = lambda v: struct.pack(bo + "H", v)  # noqa: E731
    vr = (lambda v: v) if explicit else (lambda v: None)
    out = _dcm_el(0x28, 2, us(spp), vr(b"US"), bo)
    if spp == 3:
        out += _dcm_el(0x28, 6, us(planar), vr(b"US"), bo)
    out += (_dcm_el(0x28, 0x10, us(rows), vr(b"US"), bo) + _dcm_el(0x28, 0x11, us(cols), vr(b"US"), bo)
            + _dcm_el(0x28, 0x100, us(alloc), vr(b"US"), bo) + _dcm_el(0x28, 0x101, us(stored), vr(b"US"), bo)
            + _dcm_el(0x28, 0x103, us(signed), vr(b"US"), bo))
    if inter is not None:
        out += _dcm_el(0x28, 0x1052, inter, vr(b"DS"), bo)
    if slope is not None:
        out += _dcm_el(0x28, 0x1053, slope, vr(b"DS"), bo)
    return out + _dcm_el(0x7FE0, 0x10, pixels, vr(b"OW"), bo)


def _dcm_file(ts, dataset):
    meta = _dcm_el(2, 0x10, ts + (b"\0" if len(ts) % 2 else b""), b"UI")
    return b"\0" * 128 + b"DICM" + _dcm_el(2, 0, struct.pack("<I", len(meta)), b"UL") + meta + dataset


def test_dicom_implicit_bare_masks_high_bits():
    raw = np.array([[0xF123, 0x0FFF, 0x1000]], np.uint16)          # bits above BitsStored = 12 are dropped
    ds = _dcm_image_elements(1, 3, 16, 12, 0, raw.tobytes(), explicit=False)
    np.testing.assert_array_equal(bio._read_dicom(ds), [[0x0123, 0x0FFF, 0]])


def test_dicom_explicit_sequence_rescale_and_sign():
    # an undefined-length sequence whose item holds its own Rows: the walker
    # skips it, the top-level Rows (2) counts
    nested = _dcm_el(0x28, 0x10, struct.pack("<H", 99), b"US")
    item = struct.pack("<HHI", 0xFFFE, 0xE000, 0xFFFFFFFF) + nested + struct.pack("<HHI", 0xFFFE, 0xE00D, 0)
    seq = (struct.pack("<HH2sHI", 0x0008, 0x1115, b"SQ", 0, 0xFFFFFFFF) + item
           + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    raw = np.array([0x0FFF, 0x0800, 0x07FF, 0x0001], np.uint16)     # 12-bit two's complement
    ds = seq + _dcm_image_elements(2, 2, 16, 12, 1, raw.tobytes(), slope=b"2 ", inter=b"-1024")
    got = bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.1", ds))
    assert got.dtype == np.int32
    np.testing.assert_array_equal(got, np.array([[-1, -2048], [2047, 1]]) * 2 - 1024)


def test_dicom_big_endian_deflate_and_float_rescale():
    v = np.array([[1, 2, 300], [4, 5, 60000]], np.uint16)
    be = _dcm_image_elements(2, 3, 16, 16, 0, v.astype(">u2").tobytes(), bo=">")
    np.testing.assert_array_equal(bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.2", be)), v)
    le = _dcm_image_elements(2, 3, 16, 16, 0, v.tobytes(), slope=b"0.5", inter=b"1")
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    got = bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.1.99", c.compress(le) + c.flush()))
    assert got.dtype == np.float64
    np.testing.assert_array_equal(got, v * 0.5 + 1)


@pytest.mark.parametrize("planar", [0, 1])
def test_dicom_rgb_gray_known_values(planar):
    rgb = np.array([[c for c, _ in _RGB_KNOWN]], np.uint8)          # 1 x 6 x 3
    px = rgb.transpose(2, 0, 1).tobytes() if planar else rgb.tobytes()
    ds = _dcm_image_elements(1, 6, 8, 8, 0, px, spp=3, planar=planar)
    np.testing.assert_array_equal(bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.1", ds)),
                                  [[g for _, g in _RGB_KNOWN]])


def test_dicom_rejects_compressed():
    ds = _dcm_image_elements(1, 1, 8, 8, 0, b"\0\0")
    with pytest.raises(NotImplementedError):
        bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.4.50", ds))   # JPEG baseline
    enc = (_dcm_image_elements(1, 1, 8, 8, 0, b"")[:-12]
           + struct.pack("<HH2sHI", 0x7FE0, 0x10, b"OB", 0, 0xFFFFFFFF))
    with pytest.raises(NotImplementedError):
        bio._read_dicom(_dcm_file(b"1.2.840.10008.1.2.1", enc))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.int32])
def test_dicom_slices_round_trip(tmp_path, dtype):
    arr = ((np.random.default_rng(4).random((3, 5, 9)) - 0.3) * 3000).astype(dtype)
    names = ["img%02d.dcm" % k for k in range(3)]
    d = tmp_path / "dcm"
    bio.save_data(str(d), arr, header={"slices": names, "dtype": np.dtype(dtype)}, file_type="")
    got, _ = bio.load_data(str(d))
    assert got.dtype == np.dtype(dtype)
    np.testing.assert_array_equal(got, arr)


def _zarr_dir(root, meta_name, meta, chunks):
    """A Zarr store written by hand: the metadata file and {key: bytes}."""
    import json
    root.mkdir()
    (root / meta_name).write_text(json.dumps(meta))
    for key, data in chunks.items():
        p = root.joinpath(*key.split("/"))
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_bytes(data)


def test_zarr_v2_fortran_order_edges_and_fill(tmp_path):
    a = np.arange(5 * 3 * 4, dtype=np.uint16).reshape(5, 3, 4) * 3
    chunks = (3, 3, 3)          # 2 x 1 x 2 chunks, the far ones cropped
    store = {}
    for i in range(2):
        for k in range(2):
            if (i, k) == (1, 1):
                continue        # absent: holds fill_value
            block = np.zeros(chunks, np.uint16)
            part = a[i * 3:(i + 1) * 3, :, k * 3:(k + 1) * 3]
            block[:part.shape[0], :, :part.shape[2]] = part
            store["%d.0.%d" % (i, k)] = zlib.compress(block.astype("<u2").tobytes(order="F"))
    meta = {"zarr_format": 2, "shape": [5, 3, 4], "chunks": list(chunks), "dtype": "<u2",
            "compressor": {"id": "zlib", "level": 1}, "fill_value": 7, "order": "F", "filters": None}
    _zarr_dir(tmp_path / "a.zarr", ".zarray", meta, store)
    got, ext = bio.load_data(str(tmp_path / "a.zarr"))
    want = a.copy()
    want[3:, :, 3:] = 7
    assert ext == ".zarr"
    np.testing.assert_array_equal(got, want)


def test_zarr_v2_nested_keys_uncompressed(tmp_path):
    a = (np.arange(2 * 4 * 6) - 20).astype(np.int16).reshape(2, 4, 6)
    store = {"0/%d/%d" % (j, k): a[:, j * 2:(j + 1) * 2, k * 3:(k + 1) * 3].astype("<i2").tobytes()
             for j in range(2) for k in range(2)}
    meta = {"zarr_format": 2, "shape": [2, 4, 6], "chunks": [2, 2, 3], "dtype": "<i2", "compressor": None,
            "fill_value": 0, "order": "C", "filters": None, "dimension_separator": "/"}
    _zarr_dir(tmp_path / "b.zarr", ".zarray", meta, store)
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "b.zarr"))[0], a)


@pytest.mark.parametrize("enc,codec", [("default", "gzip"), ("v2", "numcodecs.zlib")])
def test_zarr_v3(tmp_path, enc, codec):
    a = (np.random.default_rng(5).random((3, 5, 2)) * 1000).astype(np.float32)
    comp = (lambda b: gzip.compress(b)) if codec == "gzip" else (lambda b: zlib.compress(b))
    store = {}
    for j in range(2):
        block = np.zeros((3, 4, 2), np.float32)
        part = a[:, j * 4:(j + 1) * 4]
        block[:, :part.shape[1]] = part
        key = ("c/0/%d/0" % j) if enc == "default" else ("0.%d.0" % j)
        store[key] = comp(block.astype(">f4").tobytes())
    meta = {"zarr_format": 3, "node_type": "array", "shape": [3, 5, 2], "data_type": "float32",
            "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [3, 4, 2]}},
            "chunk_key_encoding": {"name": enc}, "fill_value": 0.0,
            "codecs": [{"name": "bytes", "configuration": {"endian": "big"}}, {"name": codec}]}
    _zarr_dir(tmp_path / "c.zarr", "zarr.json", meta, store)
    got = bio.load_data(str(tmp_path / "c.zarr"))[0]
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, a)


def test_zarr_rejects_other_codecs(tmp_path):
    meta = {"zarr_format": 2, "shape": [1, 1, 1], "chunks": [1, 1, 1], "dtype": "|u1",
            "compressor": {"id": "blosc", "cname": "lz4"}, "fill_value": 0, "order": "C", "filters": None}
    _zarr_dir(tmp_path / "d.zarr", ".zarray", meta, {})
    with pytest.raises(NotImplementedError):
        bio.load_data(str(tmp_path / "d.zarr"))


@pytest.mark.parametrize("compress", [True, False])
def test_zarr_round_trip(tmp_path, compress):
    import json
    a = (np.random.default_rng(6).random((120, 9, 230)) * 255).astype(np.uint8)
    p = str(tmp_path / "r.zarr")
    bio.save_data(p, a, compress=compress)
    meta = json.loads(open(os.path.join(p, ".zarray")).read())
    assert meta["chunks"] == [100, 9, 100] and (meta["compressor"] is not None) == compress
    got, _ = bio.load_data(p)
    np.testing.assert_array_equal(got, a)
    assert bio.read_header(p) is None


def _cdf1_tomo():
    """A NetCDF classic (CDF-1) file laid out byte by byte after the Unidata
    format grammar: dims z=2, x=3; no global attributes; one short variable
    'tomo' over (z, x) with 6 big-endian values."""
    head = (b"CDF\x01" + struct.pack(">I", 0)                            # magic, numrecs
            + struct.pack(">II", 0x0A, 2)                                 # NC_DIMENSION, 2 dims
            + struct.pack(">I", 1) + b"z\0\0\0" + struct.pack(">I", 2)
            + struct.pack(">I", 1) + b"x\0\0\0" + struct.pack(">I", 3)
            + struct.pack(">II", 0, 0)                                    # no global attributes
            + struct.pack(">II", 0x0B, 1)                                 # NC_VARIABLE, 1 var
            + struct.pack(">I", 4) + b"tomo" + struct.pack(">III", 2, 0, 1)
            + struct.pack(">II", 0, 0) + struct.pack(">II", 3, 12))       # no attributes; short, vsize
    begin = len(head) + 4
    return head + struct.pack(">I", begin) + struct.pack(">6h", -3, 0, 7, 300, -32768, 32767)


def test_netcdf_cdf1_known_bytes(tmp_path):
    (tmp_path / "t.nc").write_bytes(_cdf1_tomo())
    got, ext = bio.load_data(str(tmp_path / "t.nc"))
    assert ext == ".nc" and got.dtype == np.int16
    np.testing.assert_array_equal(got, [[-3, 0, 7], [300, -32768, 32767]])


def test_netcdf_bz2_hdf5_and_missing_variable(tmp_path):
    import bz2
    (tmp_path / "t.nc.bz2").write_bytes(bz2.compress(_cdf1_tomo()))
    assert bio.extension(str(tmp_path / "t.nc.bz2")) == ".nc"
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "t.nc.bz2"))[0][1], [300, -32768, 32767])
    (tmp_path / "h.nc").write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(NotImplementedError):
        bio.load_data(str(tmp_path / "h.nc"))
    # a folder's blocks must hold labels, segmented or tomo
    d = tmp_path / "blocks"
    d.mkdir()
    (d / "a.nc").write_bytes(_cdf1_tomo().replace(b"tomo", b"misc"))
    with pytest.raises(ValueError):
        bio.load_data(str(d))


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32, np.float32])
def test_netcdf_write_read_and_unsigned(tmp_path, dtype):
    a = (np.random.default_rng(7).random((3, 4, 5)) * 250).astype(dtype)
    p = str(tmp_path / "r.nc")
    bio.save_data(p, a, header={"nc": "segmented", "files": [("r.nc", 3)]})
    buf = open(p, "rb").read()
    assert buf[:4] == b"CDF\x02" and b"segmented" in buf
    assert (b"_Unsigned" in buf) == (np.dtype(dtype).kind == "u")
    got, _ = bio.load_data(p)
    assert got.dtype == np.dtype(dtype)
    np.testing.assert_array_equal(got, a)


def test_netcdf_block_folder_round_trip(tmp_path):
    a = (np.random.default_rng(8).random((7, 4, 3)) * 9).astype(np.uint8)
    src = tmp_path / "labels"
    src.mkdir()
    (src / "block_10.nc").write_bytes(bio._nc_bytes(a[3:], "labels"))      # plain sort: 10 before 2
    (src / "block_2.nc").write_bytes(bio._nc_bytes(a[:3], "labels"))
    got, ext = bio.load_data(str(src))
    assert ext == ""
    np.testing.assert_array_equal(got, np.concatenate([a[3:], a[:3]]))
    hdr = bio.read_header(str(src))
    assert hdr == {"nc": "labels", "files": [("block_10.nc", 4), ("block_2.nc", 3)]}
    out = tmp_path / "final.scan"
    bio.save_data(str(out), got, header=hdr, file_type="")
    assert sorted(os.listdir(out)) == ["block_10.nc", "block_2.nc"]
    np.testing.assert_array_equal(bio.load_data(str(out))[0], got)


def test_slice_folder_with_stray_bz2_is_not_netcdf(tmp_path):
    """A folder of slices that also holds an unrelated .bz2 file is read as
    slices: only .nc and .nc.bz2 names make a NetCDF block folder."""
    arr = (np.random.default_rng(9).random((3, 5, 6)) * 255).astype(np.uint8)
    names = ["s_%d.png" % k for k in range(3)]
    d = tmp_path / "stack"
    bio.save_data(str(d), arr, header={"slices": names, "dtype": np.uint8}, file_type="")
    (d / "notes.txt.bz2").write_bytes(bz2.compress(b"not a volume"))
    got, ext = bio.load_data(str(d))
    assert ext == ""
    np.testing.assert_array_equal(got, arr)
    assert bio.read_header(str(d))["slices"] == names
    (d / "block.nc.bz2").write_bytes(bz2.compress(bio._nc_bytes(arr, "labels")))
    assert bio._is_nc_folder(str(d))


def test_netcdf_record_variable():
    """A record variable (first dimension unlimited) that is the only one:
    records are stored unpadded, back to back."""
    vals = np.arange(2 * 3, dtype=">i2").reshape(2, 3)
    head = (b"CDF\x01" + struct.pack(">I", 2) + struct.pack(">II", 0x0A, 2)
            + struct.pack(">I", 1) + b"t\0\0\0" + struct.pack(">I", 0)
            + struct.pack(">I", 1) + b"x\0\0\0" + struct.pack(">I", 3)
            + struct.pack(">II", 0, 0) + struct.pack(">II", 0x0B, 1)
            + struct.pack(">I", 3) + b"val\0" + struct.pack(">III", 2, 0, 1)
            + struct.pack(">II", 0, 0) + struct.pack(">II", 3, 8))
    buf = head + struct.pack(">I", len(head) + 4) + vals.tobytes()
    name, got = bio._nc_array(buf)
    assert name == "val"
    np.testing.assert_array_equal(got, vals)


def _rle_dicom(rows, cols, alloc, spp, segments):
    """An RLE Lossless DICOM file: explicit VR little endian data set whose
    pixel data is encapsulated (empty offset table item, one fragment holding
    the 64-byte RLE header and the given PackBits segments)."""
    head = [len(segments)] + [0] * 15
    pos = 64
    for k, s in enumerate(segments):
        head[1 + k] = pos
        pos += len(s)
    frag = struct.pack("<16I", *head) + b"".join(segments)
    frag += b"\0" * (len(frag) % 2)
    us = lambda v: struct.pack("<H", v)  # noqa: E731
    ds = (_dcm_el(0x28, 2, us(spp), b"US") + _dcm_el(0x28, 0x10, us(rows), b"US")
          + _dcm_el(0x28, 0x11, us(cols), b"US") + _dcm_el(0x28, 0x100, us(alloc), b"US")
          + _dcm_el(0x28, 0x103, us(0), b"US")
          + struct.pack("<HH2sHI", 0x7FE0, 0x10, b"OB", 0, 0xFFFFFFFF)
          + struct.pack("<HHI", 0xFFFE, 0xE000, 0) + struct.pack("<HHI", 0xFFFE, 0xE000, len(frag)) + frag
          + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    return _dcm_file(b"1.2.840.10008.1.2.5", ds)


def test_dicom_rle_16bit():
    # 2 x 3 samples 0x0102, 0x0102, 0x0102, 0x0A0B, 0x0C0D, 0x0E0F:
    # high bytes 01 01 01 0A 0C 0E = replicate 3 x 01 (header -2), literal 3 (header 2)
    # low bytes  02 02 02 0B 0D 0F likewise; a -128 no-op byte in between
    hi = bytes([0xFE, 0x01, 0x80, 0x02, 0x0A, 0x0C, 0x0E])
    lo = bytes([0xFE, 0x02, 0x02, 0x0B, 0x0D, 0x0F])
    got = bio._read_dicom(_rle_dicom(2, 3, 16, 1, [hi, lo]))
    np.testing.assert_array_equal(got, [[0x0102, 0x0102, 0x0102], [0x0A0B, 0x0C0D, 0x0E0F]])


def test_dicom_rle_rgb_planes():
    # one 8-bit segment per colour plane, 1 x 6 pixels of _RGB_KNOWN
    rgb = np.array([c for c, _ in _RGB_KNOWN], np.uint8)
    segs = [bytes([5]) + rgb[:, k].tobytes() for k in range(3)]      # literal runs of 6
    got = bio._read_dicom(_rle_dicom(1, 6, 8, 3, segs))
    np.testing.assert_array_equal(got, [[g for _, g in _RGB_KNOWN]])
    with pytest.raises(NotImplementedError):
        bio._read_dicom(_rle_dicom(1, 6, 8, 3, segs[:2]))               # segment count mismatch


def _png_raw(W, H, depth, ctype, rows, extra=b""):
    """A PNG with filter type 0 on every row, chunks assembled by hand."""
    raw = b"".join(b"\0" + bytes(r) for r in rows)
    return (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 0))
            + extra + _png_chunk(b"IDAT", zlib.compress(raw)) + _png_chunk(b"IEND", b""))


def test_png_palette_and_low_bit_gray():
    plte = _png_chunk(b"PLTE", bytes(c for rgb, _ in _RGB_KNOWN for c in rgb))
    # 2-bit indices 0 1 2 3 | 3 2 1 0 packed most significant first; 5 pixels per row -> 2 bytes
    img = bio._read_png(_png_raw(5, 2, 2, 3, [[0b00011011, 0b00000000], [0b11100100, 0b01000000]], plte))
    g = [gv for _, gv in _RGB_KNOWN]
    np.testing.assert_array_equal(img, [[g[0], g[1], g[2], g[3], g[0]], [g[3], g[2], g[1], g[0], g[1]]])
    np.testing.assert_array_equal(bio._read_png(_png_raw(3, 1, 8, 3, [[5, 4, 3]], plte)), [[g[5], g[4], g[3]]])
    # 1-bit and 4-bit gray scale to 8 bits: 1 -> 255, 4-bit 0x7 -> 119, 0xF -> 255
    np.testing.assert_array_equal(bio._read_png(_png_raw(3, 1, 1, 0, [[0b10100000]])), [[255, 0, 255]])
    np.testing.assert_array_equal(bio._read_png(_png_raw(3, 1, 4, 0, [[0x7F, 0x00]])), [[119, 255, 0]])
    with pytest.raises(NotImplementedError):
        bio._read_png(_png_raw(3, 1, 8, 3, [[0, 1, 2]], plte + _png_chunk(b"tRNS", b"\0")))
    with pytest.raises(ValueError):
        bio._read_png(_png_raw(1, 1, 8, 3, [[9]], plte))
