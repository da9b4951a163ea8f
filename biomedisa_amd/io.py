"""Volume files for the command line and the path arguments of
smart_interpolation: load_data / save_data of biomedisa_helper.py:381-843 for
the formats this package reads and writes itself.

    .tif/.tiff   multi-page TIFF, one z-plane per page (classic or BigTIFF, either
                 byte order; strips; no compression or Deflate; predictor 1/2;
                 unsigned, signed and float samples).  Written with Deflate when
                 compress is set, BigTIFF above 2 GB, as the reference's tifffile
                 call does (biomedisa_helper.py:835-843).
    .nrrd        NRRD with raw or gzip encoding.
    .npy/.npz    NumPy arrays (.npz: the first array).

Other extensions (.am, .nc, .nii, .mhd, .zarr, directories of slices) need
libraries this package does not carry and raise NotImplementedError.
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
        buf = f.read()
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
        raw = b""
        for o, c in zip(t[_OFFS], t[_COUNTS]):
            chunk = buf[o:o + c]
            if comp == 1:
                raw += chunk
            elif comp in (8, 32946):
                raw += zlib.decompress(chunk)
            else:
                raise NotImplementedError("%s: TIFF compression %d" % (path, comp))
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
    with open(path, "wb") as f:
        f.write(out)


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


def extension(path):
    """The file type as load_data names it (biomedisa_helper.py:390-394)."""
    ext = os.path.splitext(path)[1]
    if ext == ".gz":
        ext = ".nii.gz"
    return ext.lower()


def load_data(path):
    """-> (array, extension); (None, extension) after printing the error when the
    file is missing, as load_data does."""
    ext = extension(path)
    if not os.path.exists(path):
        print(f"Error: No such file or directory '{path}'")
        return None, ext
    if ext in (".tif", ".tiff"):
        return _read_tiff(path), ext
    if ext == ".nrrd":
        return _read_nrrd(path), ext
    if ext == ".npy":
        return np.load(path, allow_pickle=False), ext
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            return z[z.files[0]], ext
    raise NotImplementedError("reading %s files is not part of this package" % ext)


def save_data(path, arr, compress=True):
    ext = extension(path)
    if ext in (".tif", ".tiff"):
        _write_tiff(path, arr, compress)
    elif ext == ".nrrd":
        _write_nrrd(path, arr, compress)
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
