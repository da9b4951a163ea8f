"""Image denoising (bmx_denoise) bit-exact against the oracle's smooth_img_3x3,
smart_interpolation(denoise=True) against the oracle pipeline, and the command
line end to end: files read, results written next to the image and equal to
the in-memory call's."""
import os

import numpy as np
import pytest

import oracle.pipeline as opipe
from oracle import orc
from biomedisa_amd import _hip
from biomedisa_amd import io as bio

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 70), (33, 17, 65), (64, 64, 64)])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_denoise_bitexact(eng, shape, dtype):
    rng = np.random.default_rng(sum(shape))
    img = (rng.random(shape) * 255).astype(dtype)
    _say("denoise %s %s" % (shape, np.dtype(dtype).name))
    src = eng.upload(img)
    out = eng.alloc(img.nbytes)
    eng.denoise(src, _hip.IMG_U8 if dtype == np.uint8 else _hip.IMG_F32, shape, out)
    got = out.download(shape, dtype)
    want = orc.smooth_img(img)
    np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_smart_interpolation_denoise(dtype):
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(30, 34, 38), seed=21, n_struct=3, radius=(4, 9), every=4, allaxis=False)
    img = vol.img if dtype == np.uint8 else vol.img.astype(np.uint16) * 200
    kw = dict(nbrw=3, sorw=120, smooth=2, denoise=True)
    got = smart_interpolation(img, vol.labels, **kw)
    want = opipe.smart_interpolation(img, vol.labels, **kw)
    for k in ("regular", "smooth"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    plain = smart_interpolation(img, vol.labels, nbrw=3, sorw=120)
    assert (plain["regular"] != got["regular"]).any()


_AM_HEADER = """# AmiraMesh BINARY-LITTLE-ENDIAN 2.1

define Lattice %d %d %d

Parameters {
    Materials {
        Exterior { id 0 }
        Tissue { Color 0.1 0.6 0.3, id 1 }
    }
    Content "%dx%dx%d byte, uniform coordinates",
    CoordType "uniform"
}

Lattice { byte Labels } @1(HxByteRLE,0)

# Data section follows"""


@pytest.mark.parametrize("label_ext", [".tif", ".nrrd", ".am", ".nii.gz", ".hdr", ".mha", ".zarr", ".nc", "", ".zip"])
def test_cli_end_to_end(tmp_path, label_ext):
    from biomedisa_amd.__main__ import main
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(28, 32, 36), seed=22, n_struct=3, radius=(4, 9), every=4, allaxis=False)
    img_path = str(tmp_path / "scan.tif")
    lab_path = str(tmp_path / ("scan_labels" + label_ext))
    bio.save_data(img_path, vol.img)
    if label_ext == ".am":
        Z, Y, X = vol.labels.shape
        bio.save_data(lab_path, vol.labels, header=(_AM_HEADER % (X, Y, Z, X, Y, Z)).encode())
    elif label_ext in ("", ".zip"):
        # a stack of PNG slices, named so that natural order differs from plain order
        names = ["lab_%d.png" % k for k in range(vol.labels.shape[0])]
        bio.save_data(lab_path, vol.labels, header={"slices": names, "dtype": np.uint8}, file_type=label_ext)
    else:
        bio.save_data(lab_path, vol.labels)
    _say("cli %s" % label_ext)
    assert main([img_path, lab_path, "--nbrw", "3", "--sorw", "120", "-s", "2", "-u", "-c", "-f"]) == 0
    want = smart_interpolation(vol.img, vol.labels, nbrw=3, sorw=120, smooth=2, uncertainty=True,
                               clean=0.1, fill=0.9)
    files = bio.result_paths(img_path, label_ext)
    assert files["regular"] == str(tmp_path / ("final.scan" + label_ext))
    for k in ("regular", "smooth", "uncertainty", "cleaned", "smooth_cleaned", "filled", "cleaned_filled"):
        got = bio.load_data(files[k])[0]
        np.testing.assert_array_equal(got, want[k], err_msg=k)
    if label_ext in ("", ".zip"):
        assert bio.read_header(files["regular"])["slices"][:3] == ["lab_0.png", "lab_1.png", "lab_2.png"]
    if label_ext == ".am":
        with open(files["regular"], "rb") as f:
            assert b"Tissue { Color 0.1 0.6 0.3, id 1 }" in f.read()


def test_cli_mrc_labels(tmp_path):
    """MRC labels (int8, y flipped in the file) give TIFF results."""
    from biomedisa_amd.__main__ import main
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    from tests.test_io import _mrc_bytes
    vol = structures_volume(shape=(28, 32, 36), seed=23, n_struct=3, radius=(4, 9), every=4, allaxis=False)
    img_path = str(tmp_path / "scan.tif")
    bio.save_data(img_path, vol.img)
    (tmp_path / "scan_labels.mrc").write_bytes(_mrc_bytes(np.ascontiguousarray(vol.labels[:, ::-1]).astype(np.int8), 0))
    _say("cli .mrc labels")
    assert main([img_path, str(tmp_path / "scan_labels.mrc"), "--nbrw", "3", "--sorw", "120"]) == 0
    want = smart_interpolation(vol.img, vol.labels, nbrw=3, sorw=120)
    np.testing.assert_array_equal(bio.load_data(str(tmp_path / "final.scan.tif"))[0], want["regular"])


@pytest.mark.parametrize("suffix", [".dcm", ".bmp"])
def test_cli_slice_folder_labels(tmp_path, suffix):
    """Labels as a folder of DICOM or BMP slices: the results are folders of
    slices with the label files' names."""
    from biomedisa_amd.__main__ import main
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(24, 30, 34), seed=24, n_struct=3, radius=(4, 8), every=4, allaxis=False)
    img_path = str(tmp_path / "scan.tif")
    bio.save_data(img_path, vol.img)
    names = ["lab%03d%s" % (k, suffix) for k in range(vol.labels.shape[0])]
    bio.save_data(str(tmp_path / "labels"), vol.labels, header={"slices": names, "dtype": np.uint8}, file_type="")
    _say("cli %s slices" % suffix)
    assert main([img_path, str(tmp_path / "labels"), "--nbrw", "3", "--sorw", "120", "-s", "2"]) == 0
    want = smart_interpolation(vol.img, vol.labels, nbrw=3, sorw=120, smooth=2)
    files = bio.result_paths(img_path, "")
    for k in ("regular", "smooth"):
        assert sorted(os.listdir(files[k])) == names
        np.testing.assert_array_equal(bio.load_data(files[k])[0], want[k], err_msg=k)
