"""A small HDF5 writer and reader for flat files of float64 datasets, used by
Geodesics.savegeos / loadgeos where h5py is not installed.

Files are written in the HDF5 format's original layout, which every HDF5
reader accepts: superblock version 0 with 8-byte offsets and lengths, a root
group held in a symbol table (a version-1 B-tree node over one symbol-table
node, names in a local heap), and one version-1 object header per dataset
with dataspace, datatype (IEEE little-endian float64) and contiguous layout
messages.  Scalars are rank-0 datasets, as h5py writes them.

The reader accepts the same layout as the HDF5 library writes it by default
(superblock 0 or 1, version-1 object headers with continuation blocks,
symbol-table groups, contiguous or compact float / integer datasets), which
covers files h5py writes with its default settings.
"""

import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K = 16       # symbol-table node capacity 2K
INTERNAL_K = 16   # B-tree node capacity 2K

# object header message types
MSG_NIL, MSG_DATASPACE, MSG_DATATYPE, MSG_LAYOUT, MSG_CONTINUATION, MSG_SYMBOL_TABLE = 0x0, 0x1, 0x3, 0x8, 0x10, 0x11


def _pad8(n):
    return (n + 7) // 8 * 8


def _message(mtype, data):
    data = data + b"\0" * (_pad8(len(data)) - len(data))
    return struct.pack("<HHB3x", mtype, len(data), 0) + data


def _object_header(messages):
    body = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


def _dataspace(shape):
    return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", d) for d in shape)


def _datatype_f64():
    # class 1 (floating point), version 1; little-endian, implied msb
    # normalisation, sign bit 63; 64-bit precision, exponent at bit 52 of
    # 11 bits, mantissa at bit 0 of 52 bits, bias 1023
    return struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _layout_contiguous(addr, size):
    return struct.pack("<BBQQ", 3, 1, addr, size)


def _dataset_header(shape, addr, nbytes):
    return _object_header([_message(MSG_DATASPACE, _dataspace(shape)), _message(MSG_DATATYPE, _datatype_f64()),
                           _message(MSG_LAYOUT, _layout_contiguous(addr, nbytes))])


def write(path, datasets):
    """Write {name: array-like} as float64 datasets in the root group."""
    items = sorted((str(k), np.array(v, dtype="<f8", order="C")) for k, v in datasets.items())
    if len(items) > 2 * LEAF_K:
        raise ValueError(f"at most {2 * LEAF_K} datasets")
    names = [k for k, _ in items]
    # local heap data: "" at offset 0, then each name null-terminated, 8-aligned
    heap = bytearray(b"\0" * 8)
    offsets = {}
    for n in names:
        offsets[n] = len(heap)
        raw = n.encode() + b"\0"
        heap += raw + b"\0" * (_pad8(len(raw)) - len(raw))
    heap = bytes(heap)

    sb_size = 96
    root_addr = sb_size
    addr = root_addr + len(_object_header([_message(MSG_SYMBOL_TABLE, b"\0" * 16)]))
    btree_size = 24 + 2 * INTERNAL_K * 8 + (2 * INTERNAL_K + 1) * 8
    btree_addr = addr
    addr += btree_size
    heap_hdr_addr = addr
    addr += 32
    heap_data_addr = addr
    addr += len(heap)
    snod_size = 8 + 2 * LEAF_K * 40
    snod_addr = addr
    addr += snod_size
    # dataset object headers (all the same size), then the raw data
    oh_size = len(_dataset_header((), 0, 0))
    oh_addr = {}
    for n, a in items:
        oh_addr[n] = addr
        addr += oh_size + 8 * a.ndim
    data_addr = {}
    for n, a in items:
        data_addr[n] = addr if a.nbytes else UNDEF
        addr += _pad8(a.nbytes)
    eof = addr

    out = bytearray(eof)
    # superblock version 0, root symbol-table entry with cache type 1 (scratch
    # pad holds the B-tree and heap addresses)
    sb = SIGNATURE + struct.pack("<8B", 0, 0, 0, 0, 0, 8, 8, 0) + struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root_addr, 1, 0) + struct.pack("<QQ", btree_addr, heap_hdr_addr)
    out[0:sb_size] = sb
    root_oh = _object_header([_message(MSG_SYMBOL_TABLE, struct.pack("<QQ", btree_addr, heap_hdr_addr))])
    out[root_addr:root_addr + len(root_oh)] = root_oh
    # v1 B-tree, group node, level 0, one child: keys are heap offsets (0 and
    # the last name in the child)
    bt = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
    bt += struct.pack("<QQQ", 0, snod_addr, offsets[names[-1]] if names else 0)
    out[btree_addr:btree_addr + len(bt)] = bt
    # local heap: no free block (free-list head 1 means none)
    out[heap_hdr_addr:heap_hdr_addr + 32] = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), 1, heap_data_addr)
    out[heap_data_addr:heap_data_addr + len(heap)] = heap
    sn = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
    for n in names:
        sn += struct.pack("<QQII16x", offsets[n], oh_addr[n], 0, 0)
    out[snod_addr:snod_addr + len(sn)] = sn
    for n, a in items:
        oh = _dataset_header(a.shape, data_addr[n], a.nbytes)
        out[oh_addr[n]:oh_addr[n] + len(oh)] = oh
        if a.nbytes:
            out[data_addr[n]:data_addr[n] + a.nbytes] = a.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---------------------------------------------------------------------------
# reader
# ---------------------------------------------------------------------------

class _File:
    def __init__(self, buf):
        self.b = buf

    def u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.b, off)

    def cstr(self, off):
        end = self.b.index(b"\0", off)
        return self.b[off:end].decode()


def _messages(f, oh_addr):
    """Messages of a version-1 object header, following continuation blocks."""
    version, _, nmsg, _, size = f.u("BBHII", oh_addr)
    if version != 1:
        raise ValueError(f"object header version {version} is not supported")
    blocks = [(oh_addr + 16, size)]
    out = []
    while blocks and len(out) < nmsg:
        start, length = blocks.pop(0)
        p = start
        while p + 8 <= start + length and len(out) < nmsg:
            mtype, msize, _ = f.u("HHB", p)
            data = f.b[p + 8:p + 8 + msize]
            if mtype == MSG_CONTINUATION:
                blocks.append(struct.unpack_from("<QQ", data, 0))
            out.append((mtype, data))
            p += 8 + msize
    return out


def _dataspace_shape(data):
    # version 1: version, rank, flags, 5 reserved bytes; version 2: version,
    # rank, flags, type; the dimensions follow
    version, rank = struct.unpack_from("<BB", data, 0)
    off = 8 if version == 1 else 4
    return tuple(struct.unpack_from("<Q", data, off + 8 * i)[0] for i in range(rank))


def _dtype(data):
    cls = data[0] & 0x0F
    bits = data[1] | (data[2] << 8) | (data[3] << 16)
    size = struct.unpack_from("<I", data, 4)[0]
    order = ">" if bits & 1 else "<"
    if cls == 1:
        return np.dtype(f"{order}f{size}")
    if cls == 0:
        return np.dtype(f"{order}{'i' if bits & 0x8 else 'u'}{size}")
    raise ValueError(f"datatype class {cls} is not supported")


def _read_dataset(f, oh_addr):
    shape, dtype, raw = None, None, None
    for mtype, data in _messages(f, oh_addr):
        if mtype == MSG_DATASPACE:
            shape = _dataspace_shape(data)
        elif mtype == MSG_DATATYPE:
            dtype = _dtype(data)
        elif mtype == MSG_LAYOUT:
            version = data[0]
            if version != 3:
                raise ValueError(f"layout message version {version} is not supported")
            cls = data[1]
            if cls == 0:    # compact: size, data
                n = struct.unpack_from("<H", data, 2)[0]
                raw = bytes(data[4:4 + n])
            elif cls == 1:  # contiguous: address, size
                addr, n = struct.unpack_from("<QQ", data, 2)
                raw = b"" if addr == UNDEF else bytes(f.b[addr:addr + n])
            else:
                raise ValueError("chunked datasets are not supported")
    if shape is None or dtype is None or raw is None:
        raise ValueError("not a dataset")
    count = int(np.prod(shape)) if shape else 1
    if len(raw) < count * dtype.itemsize:   # never written: the fill value 0
        return np.zeros(shape, dtype=dtype.newbyteorder("="))
    arr = np.frombuffer(raw, dtype=dtype, count=count).astype(dtype.newbyteorder("="))
    return arr.reshape(shape) if shape else arr[0]


def _group_entries(f, btree_addr, heap_addr):
    sig, _, heap_size, _, heap_data = f.u("4sB3xQQQ", heap_addr)
    if sig != b"HEAP":
        raise ValueError("bad local heap")

    def walk(node):
        sig, ntype, level, used = f.u("4sBBH", node)
        if sig != b"TREE" or ntype != 0:
            raise ValueError("bad group B-tree node")
        p = node + 24
        for i in range(used):
            child = f.u("Q", p + 8 + 16 * i)[0]
            if level > 0:
                yield from walk(child)
                continue
            ssig, _, _, nsym = f.u("4sBBH", child)
            if ssig != b"SNOD":
                raise ValueError("bad symbol-table node")
            for k in range(nsym):
                name_off, oh = f.u("QQ", child + 8 + 40 * k)
                yield f.cstr(heap_data + name_off), oh

    return dict(walk(btree_addr))


def read(path):
    """{name: array} of the root group's datasets (scalars as numpy scalars,
    as h5py's dataset[()] returns them)."""
    with open(path, "rb") as fh:
        f = _File(fh.read())
    base = f.b.find(SIGNATURE)
    if base != 0:
        raise ValueError("not an HDF5 file (or a user block, which is not supported)")
    version, _, _, _, _, so, sl = f.u("7B", 8)
    if version > 1 or so != 8 or sl != 8:
        raise ValueError(f"superblock version {version} with {so}/{sl}-byte offsets is not supported")
    root_entry = 56 + (4 if version == 1 else 0)
    _, root_oh, cache, _, btree, heap = f.u("QQIIQQ", root_entry)
    if cache != 1:
        for mtype, data in _messages(f, root_oh):
            if mtype == MSG_SYMBOL_TABLE:
                btree, heap = struct.unpack_from("<QQ", data, 0)
                break
        else:
            raise ValueError("root group without a symbol table")
    out = {}
    for name, oh in _group_entries(f, btree, heap).items():
        out[name] = _read_dataset(f, oh)
    return out
