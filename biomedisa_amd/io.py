"""Volume files for the command line and the path arguments of
smart_interpolation: load_data / save_data of biomedisa_helper.py:381-843 for
the formats this package reads and writes itself.

    .tif/.tiff   multi-page TIFF, one z-plane per page (classic or BigTIFF, either
                 byte order; strips; no compression or Deflate; predictor 1/2;
                 unsigned, signed and float samples).  Written with Deflate when
                 compress is set, BigTIFF above 2 GB, as the reference's tifffile
                 call does (biomedisa_helper.py:835-843).
    .nrrd        NRRD with raw or gzip encoding.
    .am          AmiraMesh / Avizo binary lattices (either byte order; raw,
                 HxByteRLE or HxZip data; the first data stream).  Results are
                 written with the label file's header, as the reference does
                 (amira_to_np/amira_data_stream.py:476-515), or a minimal one.
    .nii/.nii.gz NIfTI-1 single files (scl_slope applied as SimpleITK does).
    .hdr         NIfTI-1 pairs and Analyze 7.5 (header in .hdr, data in .img);
                 written as a NIfTI-1 pair.
    .mrc         MRC2014 volumes (modes 0, 1, 2, 6, 12), flipped along y and
                 answered with the result type .tif, as load_data does
                 (biomedisa_helper.py:595-604); read only.
    .mha/.mhd    MetaImage, data inline or in a separate raw file, optionally
                 zlib-compressed.
    .npy/.npz    NumPy arrays (.npz: the first array).
    .nc          NetCDF classic (CDF-1, CDF-2, CDF-5), also .nc.bz2 and folders
                 of .nc blocks stacked along z (nc_reader.py:143-215): the
                 tomo, segmented or labels variable; results as CDF-2 files
                 named and sized as the label blocks.
    .zarr        Zarr v2 and v3 arrays (a directory) with no compressor, zlib or
                 gzip; written as Zarr v2, zlib level 5 when compress is set, in
                 chunks of up to 100^3, as the reference's writer configures it
                 (biomedisa_helper.py:784-833).
    directory / .zip   a stack of 2-D TIFF, PNG, BMP or DICOM slices in natural
                 file-name order (biomedisa_helper.py:505-591; colour slices
                 become gray by the reference's rgb2gray, :357-364); results
                 are written as slices with the label stack's file names.

Other extensions (NetCDF-4/HDF5, JPEG slices, JPEG-family DICOM, Zarr with
other codecs) raise
NotImplementedError.
"""
import gzip
import os
import struct
import zlib

import numpy as np

# TIFF tags used here
_W, _H, _BPS, _COMP, _PHOTO, _DESC, _OFFS, _SPP, _RPS, _COUNTS, _PLANAR, _PRED, _FMT = \
    256, 257, 258, 259, 262, 270, 273, 277, 278, 279, 284, 317, 339
_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1), 7: ("B", 1),
          8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8), 17: ("q", 8),
          18: ("Q", 8)}


def _read_tiff(path):
    with open(path, "rb") as f:
        return _parse_tiff(f.read(), path)


def _parse_tiff(buf, path="TIFF"):
    """TIFF file bytes -> (pages, rows, columns) array."""
    bo = {b"II": "<", b"MM": ">"}.get(buf[:2])
    if bo is None:
        raise ValueError("%s: not a TIFF file" % path)
    magic = struct.unpack(bo + "H", buf[2:4])[0]
    if magic == 42:
        big, off = False, struct.unpack(bo + "I", buf[4:8])[0]
    elif magic == 43:
        big, off = True, struct.unpack(bo + "Q", buf[8:16])[0]
    else:
        raise ValueError("%s: not a TIFF file" % path)
    pages = []
    while off:
        n = struct.unpack(bo + ("Q" if big else "H"), buf[off:off + (8 if big else 2)])[0]
        p = off + (8 if big else 2)
        esize = 20 if big else 12
        tags = {}
        for e in range(n):
            ent = buf[p + e * esize:p + (e + 1) * esize]
            tag, typ = struct.unpack(bo + "HH", ent[:4])
            cnt = struct.unpack(bo + ("Q" if big else "I"), ent[4:12] if big else ent[4:8])[0]
            fmt, size = _TYPES.get(typ, ("B", 1))
            total = size * cnt
            inline = ent[12:20] if big else ent[8:12]
            if total <= len(inline):
                data = inline[:total]
            else:
                at = struct.unpack(bo + ("Q" if big else "I"), inline)[0]
                data = buf[at:at + total]
            if typ == 2:
                tags[tag] = data.rstrip(b"\0").decode("latin-1")
            else:
                tags[tag] = list(struct.unpack(bo + fmt[0] * (cnt * len(fmt)), data))
        pages.append(tags)
        q = p + n * esize
        off = struct.unpack(bo + ("Q" if big else "I"), buf[q:q + (8 if big else 4)])[0]
    planes = []
    for t in pages:
        w, h = t[_W][0], t[_H][0]
        bps = t.get(_BPS, [1])[0]
        spp = t.get(_SPP, [1])[0]
        comp = t.get(_COMP, [1])[0]
        pred = t.get(_PRED, [1])[0]
        fmt = t.get(_FMT, [1])[0]
        if spp != 1:
            raise NotImplementedError("%s: %d samples per pixel" % (path, spp))
        kind = {1: "u", 2: "i", 3: "f"}[fmt]
        dt = np.dtype(bo + kind + str(bps // 8)) if bps >= 8 else None
        if dt is None:
            raise NotImplementedError("%s: %d bits per sample" % (path, bps))
        parts = []
        for o, c in zip(t[_OFFS], t[_COUNTS]):
            chunk = buf[o:o + c]
            if comp == 1:
                parts.append(chunk)
            elif comp in (8, 32946):
                parts.append(zlib.decompress(chunk))
            else:
                raise NotImplementedError("%s: TIFF compression %d" % (path, comp))
        raw = b"".join(parts)
        plane = np.frombuffer(raw[:w * h * dt.itemsize], dtype=dt).reshape(h, w)
        if pred == 2:
            plane = np.cumsum(plane, axis=1, dtype=plane.dtype)
        elif pred != 1:
            raise NotImplementedError("%s: TIFF predictor %d" % (path, pred))
        planes.append(plane.astype(dt.newbyteorder("=")))
    return np.stack(planes)


def _write_tiff(path, arr, compress=True, bigtiff=None):
    """One page per z-plane, one strip per page; BigTIFF when bigtiff is set or,
    by default, when the data exceed 2 GB."""
    with open(path, "wb") as f:
        f.write(_tiff_bytes(arr, compress, bigtiff))


def _tiff_bytes(arr, compress=True, bigtiff=None):
    arr = np.ascontiguousarray(arr)
    if arr.ndim == 2:
        arr = arr[None]
    kind = {"u": 1, "i": 2, "f": 3, "b": 1}[arr.dtype.kind]
    if arr.dtype == np.bool_:
        arr = arr.astype(np.uint8)
    arr = arr.astype(arr.dtype.newbyteorder("<"))
    big = arr.nbytes > 2e9 if bigtiff is None else bool(bigtiff)
    Z, H, W = arr.shape
    out = bytearray(b"II" + (struct.pack("<HHHQ", 43, 8, 0, 16) if big else struct.pack("<HI", 42, 8)))
    ifd_pos = []
    for z in range(Z):
        data = arr[z].tobytes()
        if compress:
            data = zlib.compress(data, 6)
        strip = len(out)
        out += data
        if len(out) % 2:
            out += b"\0"
        entries = [(_W, 4, [W]), (_H, 4, [H]), (_BPS, 3, [arr.dtype.itemsize * 8]),
                   (_COMP, 3, [8 if compress else 1]), (_PHOTO, 3, [1]),
                   (_OFFS, 16 if big else 4, [strip]), (_SPP, 3, [1]), (_RPS, 4, [H]),
                   (_COUNTS, 16 if big else 4, [len(data)]), (_PLANAR, 3, [1]), (_FMT, 3, [kind])]
        ifd_pos.append(len(out))
        if big:
            out += struct.pack("<Q", len(entries))
            for tag, typ, vals in entries:
                fmt = {3: "H", 4: "I", 16: "Q"}[typ]
                val = struct.pack("<" + fmt, vals[0]).ljust(8, b"\0")
                out += struct.pack("<HHQ", tag, typ, 1) + val
            out += struct.pack("<Q", 0)
        else:
            out += struct.pack("<H", len(entries))
            for tag, typ, vals in entries:
                fmt = {3: "H", 4: "I"}[typ]
                val = struct.pack("<" + fmt, vals[0]).ljust(4, b"\0")
                out += struct.pack("<HHI", tag, typ, 1) + val
            out += struct.pack("<I", 0)
    # link the IFDs: header -> first, each -> next
    if big:
        struct.pack_into("<Q", out, 8, ifd_pos[0])
        for a, b in zip(ifd_pos, ifd_pos[1:]):
            n = struct.unpack_from("<Q", out, a)[0]
            struct.pack_into("<Q", out, a + 8 + 20 * n, b)
    else:
        struct.pack_into("<I", out, 4, ifd_pos[0])
        for a, b in zip(ifd_pos, ifd_pos[1:]):
            n = struct.unpack_from("<H", out, a)[0]
            struct.pack_into("<I", out, a + 2 + 12 * n, b)
    return bytes(out)


_NRRD_TYPES = {"uchar": "u1", "unsigned char": "u1", "uint8": "u1", "uint8_t": "u1",
               "signed char": "i1", "int8": "i1", "int8_t": "i1",
               "ushort": "u2", "unsigned short": "u2", "uint16": "u2", "uint16_t": "u2",
               "short": "i2", "signed short": "i2", "int16": "i2", "int16_t": "i2",
               "uint": "u4", "unsigned int": "u4", "uint32": "u4", "uint32_t": "u4",
               "int": "i4", "signed int": "i4", "int32": "i4", "int32_t": "i4",
               "float": "f4", "double": "f8"}


def _read_nrrd(path):
    with open(path, "rb") as f:
        buf = f.read()
    head, sep, body = buf.partition(b"\n\n")
    lines = head.decode("latin-1").split("\n")
    if not lines[0].startswith("NRRD"):
        raise ValueError("%s: not an NRRD file" % path)
    fields = {}
    for line in lines[1:]:
        if line.startswith("#") or ":" not in line:
            continue
        k, v = line.split(":", 1)
        fields[k.strip().lower()] = v.strip().lstrip("=").strip()
    if "data file" in fields or "datafile" in fields:
        raise NotImplementedError("%s: detached NRRD data" % path)
    sizes = [int(s) for s in fields["sizes"].split()]
    kind = _NRRD_TYPES[fields["type"].lower()]
    endian = "<" if fields.get("endian", "little") == "little" else ">"
    enc = fields.get("encoding", "raw").lower()
    if enc in ("gzip", "gz"):
        body = gzip.decompress(body)
    elif enc != "raw":
        raise NotImplementedError("%s: NRRD encoding %s" % (path, enc))
    dt = np.dtype(endian + kind)
    n = int(np.prod(sizes))
    arr = np.frombuffer(body[:n * dt.itemsize], dtype=dt).reshape(sizes[::-1])
    return arr.astype(dt.newbyteorder("="))


def _write_nrrd(path, arr, compress=True):
    arr = np.ascontiguousarray(arr)
    names = {v: k for k, v in [("uint8", "u1"), ("int8", "i1"), ("uint16", "u2"), ("int16", "i2"),
                               ("uint32", "u4"), ("int32", "i4"), ("float", "f4"), ("double", "f8")]}
    key = arr.dtype.kind + str(arr.dtype.itemsize)
    head = ("NRRD0004\ntype: %s\ndimension: %d\nsizes: %s\nendian: little\nencoding: %s\n\n"
            % (names[key], arr.ndim, " ".join(str(s) for s in arr.shape[::-1]), "gzip" if compress else "raw"))
    data = arr.astype(arr.dtype.newbyteorder("<")).tobytes()
    with open(path, "wb") as f:
        f.write(head.encode("latin-1"))
        f.write(gzip.compress(data, 6) if compress else data)


_AM_TYPES = {"byte": "u1", "ubyte": "u1", "sbyte": "i1", "short": "i2", "ushort": "u2", "int": "i4",
             "uint": "u4", "float": "f4", "double": "f8"}


def _rle_decode(data, n):
    from . import _hip
    src = np.frombuffer(data, np.uint8)
    out = np.empty(n, np.uint8)
    got = _hip.load().bmx_hxbyterle_decode(src.ctypes.data, src.size, out.ctypes.data, n)
    if got != n:
        raise ValueError(_hip.load().bmx_last_error().decode(errors="replace"))
    return out


def _rle_encode(arr):
    from . import _hip
    src = np.ascontiguousarray(arr, np.uint8).ravel()
    out = np.empty(src.size + (src.size + 126) // 127, np.uint8)
    n = _hip.load().bmx_hxbyterle_encode(src.ctypes.data, src.size, out.ctypes.data, out.size)
    if n < 0:
        raise ValueError(_hip.load().bmx_last_error().decode(errors="replace"))
    return out[:n].tobytes()


def _am_split(buf):
    """-> (header bytes up to the data section, data section bytes)."""
    at = buf.find(b"\n@1\n")
    if at < 0:
        raise ValueError("AmiraMesh: no data section @1")
    return buf[:at], buf[at + 4:]


def _read_am(path):
    with open(path, "rb") as f:
        buf = f.read()
    head, body = _am_split(buf)
    text = head.decode("latin-1")
    first = text.split("\n", 1)[0]
    if "AmiraMesh" not in first and "Avizo" not in first:
        raise ValueError("%s: not an AmiraMesh file" % path)
    if "BINARY" not in first:
        raise NotImplementedError("%s: ASCII AmiraMesh" % path)
    bo = "<" if "LITTLE-ENDIAN" in first else ">"
    import re
    dims = re.search(r"define\s+Lattice\s+(\d+)\s+(\d+)\s+(\d+)", text)
    if dims is None:
        raise NotImplementedError("%s: no 3-D Lattice" % path)
    X, Y, Z = (int(v) for v in dims.groups())
    m = re.search(r"Lattice\s*\{\s*(\w+)\s+\w+\s*\}\s*@1(?:\((\w+),(\d+)\))?", text)
    if m is None:
        raise NotImplementedError("%s: no Lattice data stream @1" % path)
    typ, enc, enc_len = m.group(1), m.group(2), m.group(3)
    if typ not in _AM_TYPES:
        raise NotImplementedError("%s: lattice type %s" % (path, typ))
    dt = np.dtype(bo + _AM_TYPES[typ])
    n = X * Y * Z
    if enc is None:
        arr = np.frombuffer(body[:n * dt.itemsize], dtype=dt)
    elif enc == "HxByteRLE":
        arr = _rle_decode(body[:int(enc_len)], n * dt.itemsize).view(dt)
    elif enc == "HxZip":
        arr = np.frombuffer(zlib.decompress(body[:int(enc_len)]), dtype=dt)
    else:
        raise NotImplementedError("%s: AmiraMesh encoding %s" % (path, enc))
    if arr.size != n:
        raise ValueError("%s: data section holds %d of %d values" % (path, arr.size, n))
    return arr.reshape(Z, Y, X).astype(dt.newbyteorder("="))


def am_header(path):
    """The header of an AmiraMesh file (bytes up to its data section), as the
    reference keeps it to write results with the label file's materials."""
    with open(path, "rb") as f:
        return _am_split(f.read())[0]


def _write_am(path, arr, compress=True, header=None):
    import re
    arr = np.ascontiguousarray(arr)
    names = {"u1": "byte", "i1": "sbyte", "i2": "short", "u2": "ushort", "i4": "int", "u4": "uint",
             "f4": "float", "f8": "double"}
    key = arr.dtype.kind + str(arr.dtype.itemsize)
    Z, Y, X = arr.shape
    if header is None:
        enc = "HxByteRLE" if compress and key == "u1" else None
        head = ("# AmiraMesh BINARY-LITTLE-ENDIAN 2.1\n\n\ndefine Lattice %d %d %d\n\n"
                "Parameters {\n    Content \"%dx%dx%d %s, uniform coordinates\",\n    CoordType \"uniform\"\n}\n\n"
                "Lattice { %s Data } @1%s\n\n# Data section follows"
                % (X, Y, Z, X, Y, Z, names[key], names[key], "(HxByteRLE,0)" if enc else ""))
        head = head.encode("latin-1")
        bo = "<"
    else:
        head = bytes(header)
        text = head.decode("latin-1")
        first = text.split("\n", 1)[0]
        if "BINARY" not in first:
            raise NotImplementedError("writing ASCII AmiraMesh")
        bo = "<" if "LITTLE-ENDIAN" in first else ">"
        m = re.search(r"Lattice\s*\{\s*(\w+)\s+\w+\s*\}\s*@1(?:\((\w+),(\d+)\))?", text)
        dims = re.search(r"define\s+Lattice\s+(\d+)\s+(\d+)\s+(\d+)", text)
        if m is None or dims is None or tuple(int(v) for v in dims.groups()) != (X, Y, Z):
            raise ValueError("AmiraMesh header does not describe a %dx%dx%d lattice" % (X, Y, Z))
        if _AM_TYPES.get(m.group(1)) != key:
            arr = arr.astype(_AM_TYPES[m.group(1)])
        enc = m.group(2)
    raw = arr.astype(arr.dtype.newbyteorder(bo)).tobytes()
    if enc == "HxByteRLE":
        data = _rle_encode(np.frombuffer(raw, np.uint8))
    elif enc == "HxZip":
        data = zlib.compress(raw, 6)
    elif enc is None:
        data = raw
    else:
        raise NotImplementedError("AmiraMesh encoding %s" % enc)
    if enc is not None:
        # the encoded length in the stream's definition (amira_data_stream.py:497-503)
        head = re.sub(rb"@1\(" + enc.encode() + rb",\d+\)", b"@1(" + enc.encode() + b",%d)" % len(data), head,
                      count=1)
    with open(path, "wb") as f:
        f.write(head)
        f.write(b"\n@1\n")
        f.write(data)
        f.write(b"\n")


_NII_TYPES = {2: "u1", 4: "i2", 8: "i4", 16: "f4", 64: "f8", 256: "i1", 512: "u2", 768: "u4", 1024: "i8", 1280: "u8"}


def _read_nii(path):
    """NIfTI-1 single file (.nii, .nii.gz), or a .hdr whose data lies in the
    .img beside it: a NIfTI-1 pair (magic ni1) or Analyze 7.5 (no magic),
    whose header is read with the NIfTI-1 offsets, as NIfTI readers do: the
    fields read here sit at the same bytes."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        buf = f.read()
    if len(buf) < 348:
        raise ValueError("%s: not a NIfTI-1 file" % path)
    bo = "<" if struct.unpack("<i", buf[:4])[0] == 348 else ">"
    if struct.unpack(bo + "i", buf[:4])[0] != 348:
        raise ValueError("%s: not a NIfTI-1 file" % path)
    dim = struct.unpack(bo + "8h", buf[40:56])
    dtype_code = struct.unpack(bo + "h", buf[70:72])[0]
    vox_offset = int(struct.unpack(bo + "f", buf[108:112])[0])
    slope, inter = struct.unpack(bo + "2f", buf[112:120])
    if path.lower().endswith(".hdr"):
        if buf[344:347] == b"n+1":
            raise ValueError("%s: a single-file NIfTI header in a .hdr file" % path)
        vox_offset = max(vox_offset, 0)
        img_path = path[:-4] + (".img" if path[-4:] == ".hdr" else ".IMG")
        with open(img_path, "rb") as f:
            buf = f.read()
    elif buf[344:347] != b"n+1":
        raise ValueError("%s: not a single-file NIfTI-1 image" % path)
    if dtype_code not in _NII_TYPES:
        raise NotImplementedError("%s: NIfTI datatype %d" % (path, dtype_code))
    nd = dim[0]
    shape = [dim[k] if k <= nd else 1 for k in range(1, 4)]
    if nd > 3 and any(d > 1 for d in dim[4:nd + 1]):
        raise NotImplementedError("%s: %d-D NIfTI image" % (path, nd))
    dt = np.dtype(bo + _NII_TYPES[dtype_code])
    n = shape[0] * shape[1] * shape[2]
    if len(buf) < vox_offset + n * dt.itemsize:
        raise ValueError("%s: image data shorter than the header says" % path)
    arr = np.frombuffer(buf[vox_offset:vox_offset + n * dt.itemsize], dtype=dt)
    arr = arr.reshape(shape[2], shape[1], shape[0]).astype(dt.newbyteorder("="))
    if slope not in (0.0, 1.0) or inter != 0.0:
        arr = arr.astype(np.float64) * (slope if slope != 0.0 else 1.0) + inter
    return arr


def _write_nii(path, arr, compress=True):
    arr = np.ascontiguousarray(arr)
    codes = {v: k for k, v in _NII_TYPES.items()}
    key = arr.dtype.kind + str(arr.dtype.itemsize)
    if key not in codes:
        raise NotImplementedError("NIfTI: dtype %s" % arr.dtype)
    Z, Y, X = arr.shape
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, 3, X, Y, Z, 1, 1, 1, 1)
    struct.pack_into("<2h", hdr, 70, codes[key], arr.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    pair = path.lower().endswith(".hdr")
    struct.pack_into("<3f", hdr, 108, 0.0 if pair else 352.0, 1.0, 0.0)
    struct.pack_into("<B", hdr, 123, 10)      # xyzt_units: mm, s
    if pair:                                  # header and .img beside it
        hdr[344:348] = b"ni1\0"
        with open(path, "wb") as f:
            f.write(bytes(hdr[:348]))
        with open(path[:-4] + ".img", "wb") as f:
            f.write(arr.astype(arr.dtype.newbyteorder("<")).tobytes())
        return
    hdr[344:348] = b"n+1\0"
    data = bytes(hdr) + arr.astype(arr.dtype.newbyteorder("<")).tobytes()
    if path.endswith(".gz"):
        data = gzip.compress(data, 6)
    with open(path, "wb") as f:
        f.write(data)


_MRC_TYPES = {0: "i1", 1: "i2", 2: "f4", 6: "u2", 12: "f2"}


def _read_mrc(path):
    """MRC2014: a 1024-byte header of 32-bit words (nx, ny, nz, mode first;
    the extended header's size at word 24; the machine stamp at byte 212),
    then nz planes of ny rows of nx samples."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 1024:
        raise ValueError("%s: not an MRC file" % path)
    stamp = buf[212]
    bo = ">" if stamp == 0x11 else "<"
    if stamp not in (0x11, 0x44, 0x41):       # no stamp: take the order whose mode is known
        bo = "<" if struct.unpack("<i", buf[12:16])[0] in _MRC_TYPES else ">"
    nx, ny, nz, mode = struct.unpack(bo + "4i", buf[:16])
    next_ = struct.unpack(bo + "i", buf[92:96])[0]
    if mode not in _MRC_TYPES:
        raise NotImplementedError("%s: MRC mode %d" % (path, mode))
    if min(nx, ny, nz) < 1 or next_ < 0:
        raise ValueError("%s: not an MRC file" % path)
    dt = np.dtype(bo + _MRC_TYPES[mode])
    start, n = 1024 + next_, nx * ny * nz
    if len(buf) < start + n * dt.itemsize:
        raise ValueError("%s: MRC data shorter than the header says" % path)
    arr = np.frombuffer(buf[start:start + n * dt.itemsize], dtype=dt).reshape(nz, ny, nx)
    return np.ascontiguousarray(np.flip(arr, 1)).astype(dt.newbyteorder("="))


_ZARR3_TYPES = {"bool": "b1", "int8": "i1", "uint8": "u1", "int16": "i2", "uint16": "u2", "int32": "i4",
                "uint32": "u4", "int64": "i8", "uint64": "u8", "float32": "f4", "float64": "f8"}


def _read_zarr(path):
    """A Zarr array stored as a directory: v2 (.zarray; chunk keys i.j.k or
    i/j/k; C or F order) or v3 (zarr.json; a regular chunk grid; the default
    or v2 chunk-key encoding; a bytes codec then gzip or zlib).  Chunks on the
    far edges are stored whole and cropped; absent chunks hold fill_value."""
    import json
    if os.path.isfile(os.path.join(path, ".zarray")):
        with open(os.path.join(path, ".zarray")) as f:
            meta = json.load(f)
        dt, shape, chunks = np.dtype(meta["dtype"]), meta["shape"], meta["chunks"]
        comp = meta.get("compressor")
        codecs = [] if comp is None else [comp["id"]]
        if meta.get("filters"):
            raise NotImplementedError("%s: Zarr filters" % path)
        order, sep = meta.get("order", "C"), meta.get("dimension_separator", ".")
        key = lambda idx: sep.join(map(str, idx))  # noqa: E731
        fill = meta.get("fill_value")
    elif os.path.isfile(os.path.join(path, "zarr.json")):
        with open(os.path.join(path, "zarr.json")) as f:
            meta = json.load(f)
        if meta.get("node_type") != "array":
            raise NotImplementedError("%s: a Zarr group, not an array" % path)
        grid = meta["chunk_grid"]
        if grid.get("name") != "regular" or meta["data_type"] not in _ZARR3_TYPES:
            raise NotImplementedError("%s: Zarr chunk grid %s, data type %s" % (path, grid.get("name"),
                                                                                meta["data_type"]))
        shape, chunks = meta["shape"], grid["configuration"]["chunk_shape"]
        endian, codecs = "<", []
        for c in meta.get("codecs", []):
            if c["name"] == "bytes":
                endian = ">" if (c.get("configuration") or {}).get("endian") == "big" else "<"
            else:
                codecs.append(c["name"])
        dt = np.dtype(endian + _ZARR3_TYPES[meta["data_type"]])
        enc = meta.get("chunk_key_encoding", {"name": "default"})
        sep = (enc.get("configuration") or {}).get("separator", "/" if enc["name"] == "default" else ".")
        prefix = "c" + sep if enc["name"] == "default" else ""
        key = lambda idx: prefix + sep.join(map(str, idx))  # noqa: E731
        order, fill = "C", meta.get("fill_value")
    else:
        raise ValueError("%s: no .zarray or zarr.json" % path)
    codecs = [c.replace("numcodecs.", "") for c in codecs]
    if any(c not in ("zlib", "gzip") for c in codecs):
        raise NotImplementedError("%s: Zarr codecs %s" % (path, codecs))
    if len(shape) != len(chunks) or min(chunks, default=1) < 1:
        raise ValueError("%s: Zarr shape %s with chunks %s" % (path, shape, chunks))
    if fill == "NaN":
        fill = np.nan
    elif fill is None or isinstance(fill, str):     # no fill value, or Infinity / base64 forms
        fill = 0
    out = np.full(shape, fill, dtype=dt.newbyteorder("="))
    grid = [range(-(-s // c)) for s, c in zip(shape, chunks)]
    for idx in np.ndindex(*[len(g) for g in grid]):
        name = os.path.join(path, *key(idx).split("/"))
        if not os.path.isfile(name):
            continue
        with open(name, "rb") as f:
            data = f.read()
        for c in reversed(codecs):
            data = gzip.decompress(data) if c == "gzip" else zlib.decompress(data)
        n = int(np.prod(chunks))
        if len(data) != n * dt.itemsize:
            raise ValueError("%s: Zarr chunk %s holds %d bytes" % (path, key(idx), len(data)))
        block = np.frombuffer(data, dt).reshape(chunks, order=order)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        out[sl] = block[tuple(slice(0, t.stop - t.start) for t in sl)]
    return out


def _write_zarr(path, arr, compress=True, chunks=(100, 100, 100)):
    """Zarr v2 directory (replacing one that exists): C order, chunks of
    `chunks`, zlib level 5 when compress is set."""
    import json
    import shutil
    arr = np.ascontiguousarray(arr)
    if os.path.exists(path):
        shutil.rmtree(path)
    os.makedirs(path)
    chunks = tuple(min(c, max(s, 1)) for c, s in zip(chunks, arr.shape)) if arr.ndim == 3 else arr.shape
    meta = {"zarr_format": 2, "shape": list(arr.shape), "chunks": list(chunks), "dtype": arr.dtype.str,
            "compressor": {"id": "zlib", "level": 5} if compress else None, "fill_value": 0, "order": "C",
            "filters": None, "dimension_separator": "."}
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f)
    for idx in np.ndindex(*[-(-s // c) for s, c in zip(arr.shape, chunks)]):
        block = np.zeros(chunks, arr.dtype)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, arr.shape))
        block[tuple(slice(0, t.stop - t.start) for t in sl)] = arr[sl]
        data = block.tobytes()
        with open(os.path.join(path, ".".join(map(str, idx))), "wb") as f:
            f.write(zlib.compress(data, 5) if compress else data)


_NC_TYPES = {1: "i1", 2: "S1", 3: "i2", 4: "i4", 5: "f4", 6: "f8", 7: "u1", 8: "u2", 9: "u4", 10: "i8", 11: "u8"}


def _nc_parse(buf, name="NetCDF"):
    """NetCDF classic header (CDF-1, 64-bit offset CDF-2, 64-bit data CDF-5;
    big-endian throughout) -> (numrecs, {dim name: length},
    {var name: (dim names, type, vsize, begin, attributes)}, record size)."""
    if len(buf) < 8 or (buf[:3] == b"CDF" and buf[3] not in (1, 2, 5)):
        raise ValueError("%s: not a NetCDF file" % name)
    if buf[:3] != b"CDF":
        raise NotImplementedError("%s: not a NetCDF classic file (NetCDF-4/HDF5 is not read)" % name)
    ver = buf[3]
    cnt = ">Q" if ver == 5 else ">I"               # element counts and lengths
    off = ">Q" if ver in (2, 5) else ">I"          # variable offsets
    pos = [4]

    def take(fmt):
        v = struct.unpack_from(fmt, buf, pos[0])[0]
        pos[0] += struct.calcsize(fmt)
        return v

    def text():
        n = take(cnt)
        s = buf[pos[0]:pos[0] + n].decode("utf-8", "replace")
        pos[0] += (n + 3) // 4 * 4
        return s

    def attrs():
        tag, n = take(">I"), take(cnt)
        if tag not in (0, 0x0C):
            raise ValueError("%s: bad attribute list" % name)
        out = {}
        for _ in range(n):
            key, typ = text(), take(">I")
            k = take(cnt)
            dt = np.dtype(">" + _NC_TYPES[typ]) if typ != 2 else np.dtype("S1")
            raw = buf[pos[0]:pos[0] + k * dt.itemsize]
            pos[0] += (k * dt.itemsize + 3) // 4 * 4
            out[key] = raw.decode("utf-8", "replace") if typ == 2 else np.frombuffer(raw, dt)
        return out
    numrecs = take(cnt)
    tag, n = take(">I"), take(cnt)
    dims = []
    for _ in range(n if tag == 0x0A else 0):
        dims.append((text(), take(cnt)))
    attrs()                                        # global attributes
    tag, n = take(">I"), take(cnt)
    variables = {}
    for _ in range(n if tag == 0x0B else 0):
        vname = text()
        ids = [take(cnt) for _ in range(take(cnt))]
        vatts = attrs()
        typ, vsize, begin = take(">I"), take(cnt), take(off)
        if typ not in _NC_TYPES:
            raise NotImplementedError("%s: NetCDF type %d" % (name, typ))
        variables[vname] = ([dims[i] for i in ids], typ, vsize, begin, vatts)
    recsize = sum(v[2] for v in variables.values() if v[0] and v[0][0][1] == 0)
    if len([v for v in variables.values() if v[0] and v[0][0][1] == 0]) == 1:
        recsize = next(int(np.prod([d[1] for d in v[0][1:]])) * np.dtype(_NC_TYPES[v[1]]).itemsize
                       for v in variables.values() if v[0] and v[0][0][1] == 0)
    return numrecs, dict(dims), variables, recsize


def _nc_pick(variables, name="NetCDF", first_allowed=True):
    """The variable nc_to_np reads (nc_reader.py:162-165, 199-201): 'tomo', else
    'segmented', else 'labels', else (single files) the first one."""
    for n in ("tomo", "segmented", "labels"):
        if n in variables:
            return n
    if first_allowed and variables:
        return next(iter(variables))
    raise ValueError("%s: no labels, segmented or tomo variable" % name)


def _nc_array(buf, name="NetCDF", first_allowed=True):
    """-> (variable name, array) of one NetCDF classic file: record variables
    gathered record by record; _Unsigned, scale_factor and add_offset applied
    as netCDF4 reads them by default."""
    try:
        numrecs, _, variables, recsize = _nc_parse(buf, name)
    except (struct.error, IndexError, KeyError) as e:
        raise ValueError("%s: truncated or malformed NetCDF header (%s)" % (name, e)) from None
    vname = _nc_pick(variables, name, first_allowed)
    vdims, typ, vsize, begin, vatts = variables[vname]
    dt = np.dtype(">" + _NC_TYPES[typ])
    shape = [d[1] for d in vdims]
    if shape and shape[0] == 0:                    # a record variable
        shape[0] = numrecs
        per = int(np.prod(shape[1:])) * dt.itemsize
        if begin + (numrecs - 1) * recsize + per > len(buf) and numrecs:
            raise ValueError("%s: NetCDF data shorter than its header says" % name)
        raw = b"".join(buf[begin + r * recsize:begin + r * recsize + per] for r in range(numrecs))
    else:
        n = int(np.prod(shape)) * dt.itemsize
        if begin + n > len(buf):
            raise ValueError("%s: NetCDF data shorter than its header says" % name)
        raw = buf[begin:begin + n]
    arr = np.frombuffer(raw, dt).reshape(shape).astype(dt.newbyteorder("="))
    if str(vatts.get("_Unsigned", "")).lower() == "true" and arr.dtype.kind == "i":
        arr = arr.view(arr.dtype.str.replace("i", "u"))
    if "scale_factor" in vatts or "add_offset" in vatts:
        scale = vatts.get("scale_factor", np.ones(1))[0]
        arr = arr * scale + vatts.get("add_offset", np.zeros(1, scale.dtype))[0]
    return vname, arr


def _nc_files(path):
    """The .nc and .bz2 files of a folder, in plain sorted order (nc_reader.py:175-177)."""
    names = [os.path.join(r, n) for r, _, fs in os.walk(path) for n in fs if n.endswith((".nc", ".bz2"))]
    return sorted(names)


def _is_nc_folder(path):
    """A folder is read as NetCDF blocks when it holds a .nc or .nc.bz2 file
    (biomedisa_helper.py:527-536); a stray .bz2 file beside slices is not one."""
    return os.path.isdir(path) and any(n.endswith((".nc", ".nc.bz2")) for _, _, fs in os.walk(path) for n in fs)


def _nc_file_bytes(p):
    import bz2
    with open(p, "rb") as f:
        buf = f.read()
    return bz2.decompress(buf) if p.endswith(".bz2") else buf


def _read_nc(path):
    """A NetCDF classic file (or .nc.bz2), or a folder of them stacked along z."""
    if os.path.isdir(path):
        parts = [_nc_array(_nc_file_bytes(p), p, first_allowed=False)[1] for p in _nc_files(path)]
        if not parts:
            raise ValueError("%s: no .nc files" % path)
        return np.concatenate(parts, axis=0)
    return _nc_array(_nc_file_bytes(path), path)[1]


def _nc_bytes(arr, vname="labels"):
    """A CDF-2 file holding arr (3-D) as one variable over the dimensions
    <vname>_zdim, _ydim, _xdim; unsigned integers stored in the signed type of
    their size with _Unsigned = "true", as netCDF classic writers do."""
    arr = np.ascontiguousarray(arr)
    codes = {"i1": 1, "u1": 1, "i2": 3, "u2": 3, "i4": 4, "u4": 4, "f4": 5, "f8": 6}
    key = arr.dtype.kind + str(arr.dtype.itemsize)
    if arr.ndim != 3 or key not in codes:
        raise NotImplementedError("NetCDF results hold 3-D data of 8-64 bit, not %s %s" % (arr.dtype, arr.shape))

    def text(s):
        b = s.encode()
        return struct.pack(">I", len(b)) + b + b"\0" * (-len(b) % 4)
    dims = b"".join(text("%s_%sdim" % (vname, a)) + struct.pack(">I", n) for a, n in zip("zyx", arr.shape))
    vatt = struct.pack(">II", 0, 0)
    if arr.dtype.kind == "u":
        vatt = struct.pack(">II", 0x0C, 1) + text("_Unsigned") + struct.pack(">II", 2, 4) + b"true"
    data = arr.astype(arr.dtype.str.replace("u", "i").replace("<", ">").replace("|", ">")).tobytes()
    head_wo = (b"CDF\x02" + struct.pack(">I", 0) + struct.pack(">II", 0x0A, 3) + dims + struct.pack(">II", 0, 0)
               + struct.pack(">II", 0x0B, 1) + text(vname) + struct.pack(">IIII", 3, 0, 1, 2) + vatt
               + struct.pack(">II", codes[key], min(len(data) + (-len(data) % 4), 0xFFFFFFFF)))
    begin = len(head_wo) + 8
    return head_wo + struct.pack(">Q", begin) + data


def nc_header(path):
    """What writing NetCDF results needs from a NetCDF label file or folder:
    the variable name and, per source file, its name and z-extent."""
    files = _nc_files(path) if os.path.isdir(path) else [path]
    out = []
    for p in files:
        buf = _nc_file_bytes(p)
        vname, arr = _nc_array(buf, p, first_allowed=not os.path.isdir(path))
        out.append((os.path.basename(p), arr.shape[0]))
    return {"nc": vname, "files": out}


def _write_nc(path, arr, header=None):
    """One .nc file (path ends in .nc), or a folder of blocks named and sized
    as the label files were (nc_reader.py:85-140 writes one file per source
    block); the variable keeps the label file's name."""
    vname = (header or {}).get("nc", "labels")
    if path.endswith(".nc"):
        with open(path, "wb") as f:
            f.write(_nc_bytes(arr, vname))
        return
    os.makedirs(path, exist_ok=True)
    z = 0
    for fname, nz in header["files"]:
        fname = fname[:-4] if fname.endswith(".bz2") else fname
        with open(os.path.join(path, fname), "wb") as f:
            f.write(_nc_bytes(arr[z:z + nz], vname))
        z += nz
    if z != arr.shape[0]:
        raise ValueError("NetCDF blocks cover %d planes of %d" % (z, arr.shape[0]))


_MET_TYPES = {"MET_UCHAR": "u1", "MET_CHAR": "i1", "MET_USHORT": "u2", "MET_SHORT": "i2", "MET_UINT": "u4",
              "MET_INT": "i4", "MET_ULONG_LONG": "u8", "MET_LONG_LONG": "i8", "MET_FLOAT": "f4",
              "MET_DOUBLE": "f8"}


def _read_meta(path):
    with open(path, "rb") as f:
        buf = f.read()
    fields, pos = {}, 0
    while True:
        end = buf.find(b"\n", pos)
        if end < 0:
            raise ValueError("%s: MetaImage header without ElementDataFile" % path)
        line = buf[pos:end].decode("latin-1").strip()
        pos = end + 1
        if "=" in line:
            k, v = line.split("=", 1)
            fields[k.strip()] = v.strip()
            if k.strip() == "ElementDataFile":
                break
    dims = [int(v) for v in fields["DimSize"].split()]
    if len(dims) != 3 or int(fields.get("ElementNumberOfChannels", "1")) != 1:
        raise NotImplementedError("%s: only 3-D single-channel MetaImages" % path)
    msb = (fields.get("BinaryDataByteOrderMSB", fields.get("ElementByteOrderMSB", "False")).lower() == "true")
    dt = np.dtype((">" if msb else "<") + _MET_TYPES[fields["ElementType"]])
    src = fields["ElementDataFile"]
    if src == "LOCAL":
        body = buf[pos:]
    else:
        with open(os.path.join(os.path.dirname(path), src), "rb") as f:
            body = f.read()
    if fields.get("CompressedData", "False").lower() == "true":
        body = zlib.decompress(body)
    X, Y, Z = dims
    arr = np.frombuffer(body[:X * Y * Z * dt.itemsize], dtype=dt).reshape(Z, Y, X)
    return arr.astype(dt.newbyteorder("="))


def _write_meta(path, arr, compress=True):
    arr = np.ascontiguousarray(arr)
    names = {v: k for k, v in _MET_TYPES.items()}
    key = arr.dtype.kind + str(arr.dtype.itemsize)
    if key not in names:
        raise NotImplementedError("MetaImage: dtype %s" % arr.dtype)
    Z, Y, X = arr.shape
    data = arr.astype(arr.dtype.newbyteorder("<")).tobytes()
    if compress:
        data = zlib.compress(data, 6)
    local = path.lower().endswith(".mha")
    raw_name = os.path.splitext(os.path.basename(path))[0] + (".zraw" if compress else ".raw")
    head = ("ObjectType = Image\nNDims = 3\nBinaryData = True\nBinaryDataByteOrderMSB = False\n"
            "CompressedData = %s\n%sTransformMatrix = 1 0 0 0 1 0 0 0 1\nOffset = 0 0 0\n"
            "ElementSpacing = 1 1 1\nDimSize = %d %d %d\nElementType = %s\nElementDataFile = %s\n"
            % ("True" if compress else "False", "CompressedDataSize = %d\n" % len(data) if compress else "",
               X, Y, Z, names[key], "LOCAL" if local else raw_name))
    with open(path, "wb") as f:
        f.write(head.encode("latin-1"))
        if local:
            f.write(data)
    if not local:
        with open(os.path.join(os.path.dirname(path), raw_name), "wb") as f:
            f.write(data)


_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _rgb2gray(rgb):
    """(rows, cols, 3) R, G, B -> gray as biomedisa_helper.py:357-364 computes
    it: the weighted sum in double, truncated to the sample type."""
    return (0.2989 * rgb[:, :, 0] + 0.587 * rgb[:, :, 1] + 0.114 * rgb[:, :, 2]).astype(rgb.dtype)


def _read_png(buf, name="PNG"):
    """Non-interlaced PNG -> (rows, cols) array: grayscale (1/2/4-bit samples
    scaled to 8 bits by bit replication, as libpng's gray expansion does;
    alpha dropped), RGB, or a palette expanded to RGB; colour is made gray by
    _rgb2gray."""
    from . import _hip
    if buf[:8] != _PNG_SIG:
        raise ValueError("%s: not a PNG file" % name)
    pos, idat, ihdr, plte, trns = 8, [], None, None, False
    while pos < len(buf):
        n, typ = struct.unpack(">I4s", buf[pos:pos + 8])
        body = buf[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat.append(body)
        elif typ == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif typ == b"tRNS":
            trns = True
        elif typ == b"IEND":
            break
        pos += 12 + n
    if ihdr is None:
        raise ValueError("%s: no IHDR chunk" % name)
    W, H, depth, ctype, _, _, interlace = ihdr
    ok = {0: (1, 2, 4, 8, 16), 2: (8, 16), 3: (1, 2, 4, 8), 4: (8, 16)}
    if ctype not in ok or depth not in ok[ctype] or interlace or (ctype == 3 and (plte is None or trns)):
        raise NotImplementedError("%s: PNG colour type %d, depth %d, interlace %d" % (name, ctype, depth, interlace))
    chans = {0: 1, 2: 3, 3: 1, 4: 2}[ctype]
    bpp = max(1, chans * depth // 8)
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    rowbytes = (W * chans * depth + 7) // 8
    if raw.size < H * (rowbytes + 1):
        raise ValueError("%s: truncated image data" % name)
    out = np.empty(H * rowbytes, np.uint8)
    rc = _hip.load().bmx_png_unfilter(raw.ctypes.data, H, rowbytes, bpp, out.ctypes.data)
    if rc != 0:
        raise ValueError(_hip.load().bmx_last_error().decode(errors="replace"))
    if depth < 8:                              # packed samples, most significant first
        img = np.unpackbits(out.reshape(H, rowbytes), axis=1).reshape(H, -1, depth)[:, :W]
        img = (img * (1 << np.arange(depth - 1, -1, -1, dtype=np.uint8))).sum(axis=2, dtype=np.uint8)
        if ctype == 0:
            return img * np.uint8(255 // ((1 << depth) - 1))
        img = img[:, :, None]
    else:
        img = out.view(">u2" if depth == 16 else "u1").reshape(H, W, chans)
        img = img.astype(img.dtype.newbyteorder("="))
    if ctype == 3:
        if int(img.max()) >= len(plte):
            raise ValueError("%s: PNG index beyond its palette" % name)
        return _rgb2gray(plte[img[:, :, 0]])
    return _rgb2gray(img) if ctype == 2 else img[:, :, 0]


def _read_bmp(buf, name="BMP"):
    """Uncompressed Windows bitmap with a BITMAPINFOHEADER or later header:
    8-bit with a palette (a gray palette gives its gray values, a colour one
    goes through _rgb2gray) or 24-bit B, G, R; rows bottom-up unless the
    height is negative, each padded to 4 bytes -> (rows, cols) uint8."""
    if len(buf) < 54 or buf[:2] != b"BM":
        raise ValueError("%s: not a BMP file" % name)
    off, hsize = struct.unpack("<II", buf[10:18])
    if hsize < 40:
        raise NotImplementedError("%s: BMP core header" % name)
    W, H, _, bits, comp = struct.unpack("<iiHHI", buf[18:34])
    if comp != 0 or bits not in (8, 24):
        raise NotImplementedError("%s: BMP with %d bits, compression %d" % (name, bits, comp))
    top_down, H = H < 0, abs(H)
    if W < 1 or H < 1:
        raise ValueError("%s: BMP of %d x %d pixels" % (name, W, H))
    stride = (W * bits // 8 + 3) & ~3
    if len(buf) < off + stride * H:
        raise ValueError("%s: truncated BMP data" % name)
    rows = np.frombuffer(buf, np.uint8, stride * H, off).reshape(H, stride)[:, :W * bits // 8]
    if not top_down:
        rows = rows[::-1]
    if bits == 24:
        return _rgb2gray(np.ascontiguousarray(rows.reshape(H, W, 3)[:, :, ::-1]))
    n = struct.unpack("<I", buf[46:50])[0] or 256
    if n > 256 or 14 + hsize + 4 * n > off:
        raise ValueError("%s: BMP palette of %d entries" % (name, n))
    pal = np.frombuffer(buf, np.uint8, 4 * n, 14 + hsize).reshape(n, 4)      # B, G, R, 0
    if int(rows.max()) >= n:
        raise ValueError("%s: BMP index beyond its palette" % name)
    px = pal[rows]
    if (pal[:, 0] == pal[:, 1]).all() and (pal[:, 1] == pal[:, 2]).all():
        return np.ascontiguousarray(px[:, :, 2])
    return _rgb2gray(np.ascontiguousarray(px[:, :, 2::-1]))


def _bmp_bytes(img):
    """2-D uint8 -> 8-bit BMP with a gray-ramp palette, rows bottom-up."""
    img = np.ascontiguousarray(img)
    if img.dtype != np.uint8 or img.ndim != 2:
        raise NotImplementedError("BMP slices hold 2-D uint8 data, not %s %s" % (img.dtype, img.shape))
    H, W = img.shape
    stride = (W + 3) & ~3
    rows = np.zeros((H, stride), np.uint8)
    rows[:, :W] = img[::-1]
    ramp = np.repeat(np.arange(256, dtype=np.uint8), 4).reshape(256, 4)
    ramp[:, 3] = 0
    off = 14 + 40 + 1024
    return (b"BM" + struct.pack("<IHHI", off + rows.size, 0, 0, off)
            + struct.pack("<IiiHHIIiiII", 40, W, H, 1, 8, 0, rows.size, 2835, 2835, 256, 0)
            + ramp.tobytes() + rows.tobytes())


def _png_bytes(img, compress=True):
    img = np.ascontiguousarray(img)
    if img.dtype not in (np.uint8, np.uint16) or img.ndim != 2:
        raise NotImplementedError("PNG slices hold 2-D uint8 or uint16 data, not %s %s" % (img.dtype, img.shape))
    H, W = img.shape
    depth = 8 * img.dtype.itemsize
    rows = img.astype(img.dtype.newbyteorder(">")).view(np.uint8).reshape(H, -1)
    data = np.concatenate([np.zeros((H, 1), np.uint8), rows], axis=1).tobytes()

    def chunk(typ, body):
        return struct.pack(">I", len(body)) + typ + body + struct.pack(">I", zlib.crc32(typ + body) & 0xFFFFFFFF)
    return (_PNG_SIG + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, 0, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(data, 6 if compress else 0)) + chunk(b"IEND", b""))


def _natural_key(s):
    """File-name order of the reference (biomedisa_helper.py:377-379)."""
    import re
    return [int(t) if t.isdigit() else t.lower() for t in re.split(r"(\d+)", s)]


# DICOM: value representations with a 2-byte reserved field and a 4-byte length
# in explicit-VR encodings (PS3.5 section 7.1.2)
_DCM_LONG_VR = {b"OB", b"OD", b"OF", b"OL", b"OV", b"OW", b"SQ", b"UC", b"UN", b"UR", b"UT", b"SV", b"UV"}
_DCM_TS = {"1.2.840.10008.1.2": ("<", False), "1.2.840.10008.1.2.1": ("<", True),
           "1.2.840.10008.1.2.2": (">", True), "1.2.840.10008.1.2.1.99": ("<", True),
           "1.2.840.10008.1.2.5": ("<", True)}                 # RLE Lossless (encapsulated)
_DCM_WANT = {(0x0028, 0x0002): "spp", (0x0028, 0x0006): "planar", (0x0028, 0x0008): "frames",
             (0x0028, 0x0010): "rows", (0x0028, 0x0011): "cols", (0x0028, 0x0100): "alloc",
             (0x0028, 0x0101): "stored", (0x0028, 0x0103): "signed", (0x0028, 0x1052): "inter",
             (0x0028, 0x1053): "slope", (0x7FE0, 0x0010): "pixels"}


def _dcm_elements(buf, pos, end, bo, explicit, out, depth=0):
    """Walks data elements from pos to end (or to an item delimiter when end is
    None), keeping the values of _DCM_WANT at the top level; sequences and
    items of undefined length are walked to their delimiters.  -> end position."""
    while pos + 8 <= len(buf) and (end is None or pos < end):
        g, e = struct.unpack(bo + "HH", buf[pos:pos + 4])
        if g == 0xFFFE:                                   # item / delimiters: implicit form
            n = struct.unpack(bo + "I", buf[pos + 4:pos + 8])[0]
            pos += 8
            if e == 0xE00D or e == 0xE0DD:                # item / sequence delimitation
                return pos
            if n == 0xFFFFFFFF:
                pos = _dcm_elements(buf, pos, None, bo, explicit, out, depth + 1)
            else:
                pos += n
            continue
        vr = buf[pos + 4:pos + 6] if explicit else None
        if explicit and vr in _DCM_LONG_VR:
            n = struct.unpack(bo + "I", buf[pos + 8:pos + 12])[0]
            pos += 12
        elif explicit:
            n = struct.unpack(bo + "H", buf[pos + 6:pos + 8])[0]
            pos += 8
        else:
            n = struct.unpack(bo + "I", buf[pos + 4:pos + 8])[0]
            pos += 8
        if n == 0xFFFFFFFF and (g, e) == (0x7FE0, 0x0010) and depth == 0:
            # encapsulated pixel data: the offset table item, then fragments
            frags = []
            while pos + 8 <= len(buf):
                ig, ie, k = struct.unpack("<HHI", buf[pos:pos + 8])
                pos += 8
                if (ig, ie) == (0xFFFE, 0xE0DD):
                    break
                if (ig, ie) != (0xFFFE, 0xE000):
                    raise ValueError("DICOM: bad item in encapsulated pixel data")
                frags.append(buf[pos:pos + k])
                pos += k
            out["encapsulated"] = frags[1:]
            continue
        if n == 0xFFFFFFFF:
            pos = _dcm_elements(buf, pos, None, bo, explicit, out, depth + 1)
            continue
        key = _DCM_WANT.get((g, e)) if depth == 0 else None
        if key is not None:
            out[key] = (vr, buf[pos:pos + n])
        pos += n
    return pos


def _read_dicom(buf, name="DICOM"):
    """One DICOM image (PS3.10 file, or a bare implicit-VR data set):
    implicit or explicit VR little endian, explicit big endian, deflated
    explicit little endian or RLE Lossless; 8/16/32-bit samples with the bits above
    BitsStored masked (unsigned) or sign-extended (signed); rescale slope and
    intercept applied as GDCM does (identity: the stored type; integral:
    int32; else float64); three samples per pixel made gray by _rgb2gray."""
    pos, ts = 0, "1.2.840.10008.1.2"
    if buf[128:132] == b"DICM":
        meta = {}
        pos = 132
        while pos + 8 <= len(buf) and struct.unpack("<H", buf[pos:pos + 2])[0] == 0x0002:
            g, e = struct.unpack("<HH", buf[pos:pos + 4])
            vr = buf[pos + 4:pos + 6]
            if vr in _DCM_LONG_VR:
                n, hd = struct.unpack("<I", buf[pos + 8:pos + 12])[0], 12
            else:
                n, hd = struct.unpack("<H", buf[pos + 6:pos + 8])[0], 8
            meta[e] = buf[pos + hd:pos + hd + n]
            pos += hd + n
        ts = meta.get(0x0010, b"").rstrip(b"\0 ").decode("ascii", "replace") or ts
    if ts not in _DCM_TS:
        raise NotImplementedError("%s: DICOM transfer syntax %s (compressed)" % (name, ts))
    bo, explicit = _DCM_TS[ts]
    if ts.endswith(".99"):
        buf, pos = zlib.decompress(buf[pos:], -15), 0
    el = {}
    _dcm_elements(buf, pos, len(buf), bo, explicit, el)

    def num(key, default=None):
        if key not in el:
            if default is None:
                raise ValueError("%s: DICOM without %s" % (name, key))
            return default
        vr, raw = el[key]
        if key in ("inter", "slope", "frames"):               # DS / IS: decimal strings
            txt = raw.rstrip(b"\0 ").decode("ascii", "replace").split("\\")[0].strip()
            return float(txt) if txt else default
        return struct.unpack(bo + "H", raw[:2])[0]            # US
    rows, cols, alloc = num("rows"), num("cols"), num("alloc")
    spp, signed, frames = num("spp", 1), num("signed", 0), int(num("frames", 1))
    stored = num("stored", alloc)
    if "encapsulated" in el:
        if ts != "1.2.840.10008.1.2.5" or frames != 1 or alloc not in (8, 16, 32) or spp not in (1, 3):
            raise NotImplementedError("%s: encapsulated DICOM pixel data in %s" % (name, ts))
        el["pixels"] = (None, _dcm_rle(b"".join(el["encapsulated"]), rows * cols, spp, alloc // 8, name))
        bo, planar_rle = "<", True
    else:
        planar_rle = False
    if "pixels" not in el:
        raise ValueError("%s: DICOM without pixel data" % name)
    if alloc not in (8, 16, 32) or spp not in (1, 3) or frames != 1 or not 1 <= stored <= alloc:
        raise NotImplementedError("%s: DICOM with %d-bit samples, %d per pixel, %d frames"
                                  % (name, alloc, spp, frames))
    dt = np.dtype(bo + ("i" if signed else "u") + str(alloc // 8))
    n = rows * cols * spp
    raw = el["pixels"][1]
    if len(raw) < n * dt.itemsize:
        raise ValueError("%s: DICOM pixel data shorter than its rows and columns" % name)
    px = np.frombuffer(raw[:n * dt.itemsize], dt).astype(dt.newbyteorder("="))
    if stored < alloc:
        if signed:
            shift = alloc - stored
            px = ((px << shift) >> shift).astype(px.dtype)
        else:
            px = px & px.dtype.type((1 << stored) - 1)
    if spp == 3:
        px = (px.reshape(3, rows, cols).transpose(1, 2, 0) if num("planar", 0) or planar_rle
              else px.reshape(rows, cols, 3))
        px = _rgb2gray(np.ascontiguousarray(px))
    else:
        px = px.reshape(rows, cols)
    slope, inter = num("slope", 1.0), num("inter", 0.0)
    if slope == 1.0 and inter == 0.0:
        return px
    if float(slope).is_integer() and float(inter).is_integer():
        return (px.astype(np.int64) * int(slope) + int(inter)).astype(np.int32)
    return px.astype(np.float64) * slope + inter


def _dcm_rle(frag, npix, spp, bps, name="DICOM"):
    """RLE Lossless frame (PS3.5 annex G): a 64-byte header (segment count,
    15 offsets), then PackBits segments, most significant byte of each sample
    first and, for colour, one plane after another -> little-endian sample
    bytes (planes in sample order)."""
    if len(frag) < 64:
        raise ValueError("%s: RLE frame shorter than its header" % name)
    head = struct.unpack("<16I", frag[:64])
    nseg = head[0]
    if nseg != spp * bps or nseg > 15:
        raise NotImplementedError("%s: RLE with %d segments (want %d)" % (name, nseg, spp * bps))
    offs = list(head[1:1 + nseg]) + [len(frag)]
    segs = []
    for k in range(nseg):
        src, out, i = frag[offs[k]:offs[k + 1]], bytearray(), 0
        while i < len(src) and len(out) < npix:
            h = src[i] - 256 if src[i] > 127 else src[i]
            if h >= 0:
                out += src[i + 1:i + 2 + h]
                i += 2 + h
            elif h > -128:
                out += src[i + 1:i + 2] * (1 - h)
                i += 2
            else:
                i += 1
        if len(out) < npix:
            raise ValueError("%s: RLE segment %d decodes to %d of %d bytes" % (name, k, len(out), npix))
        segs.append(np.frombuffer(bytes(out[:npix]), np.uint8))
    # bps segments per sample, from the most significant byte down
    planes = []
    for s0 in range(0, nseg, bps):
        planes.append(np.stack(segs[s0:s0 + bps][::-1], axis=1).reshape(-1))
    return b"".join(p.tobytes() for p in planes)


def _dicom_bytes(img):
    """2-D 8/16/32-bit integers -> a minimal PS3.10 Secondary Capture file in
    explicit VR little endian, MONOCHROME2."""
    img = np.ascontiguousarray(img)
    if img.ndim != 2 or img.dtype.kind not in "iu" or img.dtype.itemsize > 4 or img.dtype == np.int8:
        raise NotImplementedError("DICOM slices hold 2-D 8/16/32-bit integers, not %s %s"
                                  % (img.dtype, img.shape))
    import uuid

    def elem(g, e, vr, value):
        if len(value) % 2:
            value += b"\0" if vr in (b"UI", b"OB") else b" "
        if vr in _DCM_LONG_VR:
            return struct.pack("<HH2sHI", g, e, vr, 0, len(value)) + value
        return struct.pack("<HH2sH", g, e, vr, len(value)) + value
    uid = ("2.25.%d" % uuid.uuid4().int).encode()
    sop_class = b"1.2.840.10008.5.1.4.1.1.7"
    meta = (elem(2, 1, b"OB", b"\x00\x01") + elem(2, 2, b"UI", sop_class) + elem(2, 3, b"UI", uid)
            + elem(2, 0x10, b"UI", b"1.2.840.10008.1.2.1"))
    meta = elem(2, 0, b"UL", struct.pack("<I", len(meta))) + meta
    bits = img.dtype.itemsize * 8
    rows, cols = img.shape
    ds = (elem(8, 0x16, b"UI", sop_class) + elem(8, 0x18, b"UI", uid)
          + elem(0x28, 2, b"US", struct.pack("<H", 1)) + elem(0x28, 4, b"CS", b"MONOCHROME2")
          + elem(0x28, 0x10, b"US", struct.pack("<H", rows)) + elem(0x28, 0x11, b"US", struct.pack("<H", cols))
          + elem(0x28, 0x100, b"US", struct.pack("<H", bits)) + elem(0x28, 0x101, b"US", struct.pack("<H", bits))
          + elem(0x28, 0x102, b"US", struct.pack("<H", bits - 1))
          + elem(0x28, 0x103, b"US", struct.pack("<H", 1 if img.dtype.kind == "i" else 0))
          + elem(0x7FE0, 0x10, b"OB" if bits == 8 else b"OW", img.astype(img.dtype.newbyteorder("<")).tobytes()))
    return b"\0" * 128 + b"DICM" + meta + ds


_SLICE_TYPES = (".png", ".tif", ".tiff", ".bmp", ".dcm", ".dicom")
_OTHER_SLICES = (".jpg", ".jpeg")


def _slice_names(path):
    """Sorted names of the slice files of a directory (full paths) or of a zip
    archive (member names)."""
    import zipfile
    if os.path.isdir(path):
        names = [os.path.join(root, n) for root, _, files in os.walk(path) for n in files]
    else:
        with zipfile.ZipFile(path) as z:
            names = [i.filename for i in z.infolist() if not i.is_dir()]
    keep = [n for n in names if not os.path.basename(n).startswith(".")
            and os.path.splitext(n)[1].lower() in _SLICE_TYPES]
    if not keep and any(os.path.splitext(n)[1].lower() in _OTHER_SLICES for n in names):
        raise NotImplementedError("%s: only TIFF and PNG slices are read" % path)
    if not keep:
        raise ValueError("%s: no image slices" % path)
    keep.sort(key=_natural_key)
    return keep


def _slice_bytes(path):
    """-> iterator over (name, file bytes) of the slices in natural order."""
    import zipfile
    names = _slice_names(path)
    if os.path.isdir(path):
        for n in names:
            with open(n, "rb") as f:
                yield n, f.read()
    else:
        with zipfile.ZipFile(path) as z:
            for n in names:
                yield n, z.read(n)


def _decode_slice(name, buf):
    ext = os.path.splitext(name)[1].lower()
    if ext == ".png":
        return _read_png(buf, name)
    if ext == ".bmp":
        return _read_bmp(buf, name)
    if ext in (".dcm", ".dicom"):
        return _read_dicom(buf, name)
    img = _parse_tiff(buf, name)
    if img.shape[0] != 1:
        raise NotImplementedError("%s: multi-page TIFF slice" % name)
    return img[0]


def _read_slices(path):
    planes, names = [], []
    for name, buf in _slice_bytes(path):
        img = _decode_slice(name, buf)
        if planes and (img.shape != planes[0].shape or img.dtype != planes[0].dtype):
            raise ValueError("%s: slice %s differs in shape or type from the first" % (path, name))
        planes.append(img)
        names.append(os.path.basename(name))
    return np.stack(planes), {"slices": names, "dtype": planes[0].dtype}


def _slice_file_bytes(name, img, compress):
    ext = os.path.splitext(name)[1].lower()
    if ext == ".png":
        return _png_bytes(img, compress)
    if ext == ".bmp":
        return _bmp_bytes(img)
    if ext in (".dcm", ".dicom"):
        return _dicom_bytes(img)
    return _tiff_bytes(img, compress)


def _write_slices(path, arr, compress=True, header=None):
    """Slices written with the label stack's file names and sample type
    (biomedisa_helper.py:758-783): a folder, or a zip archive for .zip."""
    import zipfile
    arr = np.asarray(arr)
    if header is not None:
        names, dtype = header["slices"], header["dtype"]
        if len(names) != arr.shape[0]:
            raise ValueError("%d slice names for %d planes" % (len(names), arr.shape[0]))
        arr = arr.astype(dtype)
    else:
        names = ["slice_%05d.tif" % k for k in range(arr.shape[0])]
    if path.lower().endswith(".zip"):
        with zipfile.ZipFile(path, "w") as z:
            for k, n in enumerate(names):
                z.writestr(n, _slice_file_bytes(n, arr[k], compress))
    else:
        os.makedirs(path, exist_ok=True)
        for k, n in enumerate(names):
            with open(os.path.join(path, n), "wb") as f:
                f.write(_slice_file_bytes(n, arr[k], compress))


def read_header(path):
    """What writing results in the type of `path` keeps from it: the AmiraMesh
    header, or the slice names and sample type of a slice stack; else None."""
    ext = extension(path)
    if ext == ".am":
        return am_header(path)
    if ext == ".zarr":
        return None
    if ext == ".nc" or _is_nc_folder(path):
        return nc_header(path)
    if os.path.isdir(path) or ext == ".zip":
        # the names of every slice and the sample type of the first
        slices = _slice_bytes(path)
        name, buf = next(slices)
        slices.close()
        return {"slices": [os.path.basename(n) for n in _slice_names(path)],
                "dtype": _decode_slice(name, buf).dtype}
    return None


def extension(path):
    """The file type as load_data names it (biomedisa_helper.py:390-394)."""
    ext = os.path.splitext(path)[1]
    if ext == ".gz":
        ext = ".nii.gz"
    elif ext == ".bz2":
        ext = os.path.splitext(os.path.splitext(path)[0])[1]
    return ext.lower()


def load_data(path):
    """-> (array, extension); (None, extension) after printing the error when the
    file is missing, as load_data does."""
    ext = extension(path)
    if not os.path.exists(path):
        print(f"Error: No such file or directory '{path}'")
        return None, ext
    if ext == ".zarr":
        return _read_zarr(path), ext
    if ext == ".nc" or _is_nc_folder(path):
        return _read_nc(path), "" if os.path.isdir(path) else ext
    if os.path.isdir(path):
        return _read_slices(path)[0], ""
    if ext == ".zip":
        return _read_slices(path)[0], ext
    if ext in (".tif", ".tiff"):
        return _read_tiff(path), ext
    if ext == ".nrrd":
        return _read_nrrd(path), ext
    if ext == ".am":
        return _read_am(path), ext
    if ext in (".nii", ".nii.gz", ".hdr"):
        return _read_nii(path), ext
    if ext in (".mha", ".mhd"):
        return _read_meta(path), ext
    if ext == ".mrc":
        return _read_mrc(path), ".tif"
    if ext == ".npy":
        return np.load(path, allow_pickle=False), ext
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            return z[z.files[0]], ext
    raise NotImplementedError("reading %s files is not part of this package" % ext)


def save_data(path, arr, compress=True, header=None, file_type=None):
    """Writes arr with the type file_type (default: the extension of path; ''
    is a folder of slices), as save_data(..., final_image_type) does;
    header: read_header of the label file, for .am and slice stacks."""
    ext = extension(path) if file_type is None else file_type
    if ext == ".nc" or (ext == "" and isinstance(header, dict) and "nc" in header):
        _write_nc(path, arr, header)
    elif ext in ("", ".zip"):
        _write_slices(path, arr, compress, header)
    elif ext in (".tif", ".tiff"):
        _write_tiff(path, arr, compress)
    elif ext == ".nrrd":
        _write_nrrd(path, arr, compress)
    elif ext == ".am":
        _write_am(path, arr, compress, header)
    elif ext in (".nii", ".nii.gz", ".hdr"):
        _write_nii(path, arr, compress)
    elif ext in (".mha", ".mhd"):
        _write_meta(path, arr, compress)
    elif ext == ".zarr":
        _write_zarr(path, arr, compress)
    elif ext == ".npy":
        np.save(path, arr, allow_pickle=False)
    elif ext == ".npz":
        (np.savez_compressed if compress else np.savez)(path, arr)
    else:
        raise NotImplementedError("writing %s files is not part of this package" % ext)


def result_paths(path_to_data, final_image_type):
    """The result files of interpolation.py:115-133: final.<image name> next to the
    image, with the label file's extension, and the optional results beside it."""
    base = os.path.basename(path_to_data)
    filename, ext = os.path.splitext(base)
    if ext == ".gz":
        filename = filename[:-4]
    final = os.path.join(os.path.dirname(path_to_data), "final." + filename + final_image_type)
    stem, ext = os.path.splitext(final)
    if ext == ".gz":
        stem = stem[:-4]
    return {"regular": final, "smooth": stem + ".smooth" + final_image_type,
            "smooth_cleaned": stem + ".smooth.cleaned" + final_image_type,
            "cleaned": stem + ".cleaned" + final_image_type, "filled": stem + ".filled" + final_image_type,
            "cleaned_filled": stem + ".cleaned.filled" + final_image_type,
            "refined": stem + ".refined" + final_image_type, "acwe": stem + ".acwe" + final_image_type,
            "uncertainty": stem + ".uncertainty.tif"}
