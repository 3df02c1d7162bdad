"""A small ISOBMFF / HEIF writer for container tests.

It assembles files from coded pictures (the generator's Annex B stream,
split into parameter sets and slices) with the box layouts the tests need:
iloc versions 0/1/2, construction method 1 (idat), several extents per item,
any property list per item, grid items, NAL length sizes 1..4.  Box syntax:
ISO/IEC 14496-12 (ftyp, meta, hdlr, pitm, iinf/infe, iref, iprp/ipco/ipma,
iloc, idat, mdat) and ISO/IEC 23008-12 (hvcC, ispe, colr, irot, imir, grid).
"""
import struct


def box(t: bytes, payload: bytes) -> bytes:
    return struct.pack(">I", 8 + len(payload)) + t + payload


def fbox(t: bytes, version: int, flags: int, payload: bytes) -> bytes:
    return box(t, bytes([version]) + flags.to_bytes(3, "big") + payload)


def split_annexb(stream: bytes):
    """(vps, sps, pps, [[slice NAL, ...] per picture]) of an Annex B stream."""
    nals = [n for n in stream.split(b"\x00\x00\x00\x01") if n]
    ps = {}
    pics = []
    for n in nals:
        t = (n[0] >> 1) & 63
        if t in (32, 33, 34):
            ps[t] = n
        elif t < 32:
            if n[2] & 0x80 or not pics:
                pics.append([])
            pics[-1].append(n)
    return ps[32], ps[33], ps[34], pics


def hvcc(vps: bytes, sps: bytes, pps: bytes, length_size: int = 4, arrays=(32, 33, 34)) -> bytes:
    head = bytearray(23)
    head[0] = 1
    head[21] = 0xFC | (length_size - 1)
    by_type = {32: vps, 33: sps, 34: pps}
    head[22] = len(arrays)
    out = bytes(head)
    for t in arrays:
        n = by_type[t]
        out += bytes([0x80 | t]) + struct.pack(">HH", 1, len(n)) + n
    return out


def sample(slices, length_size: int = 4) -> bytes:
    out = b""
    for n in slices:
        assert len(n) < (1 << (8 * length_size))
        out += len(n).to_bytes(length_size, "big") + n
    return out


def ispe(w: int, h: int) -> bytes:
    return struct.pack(">III", 0, w, h)


def nclx(matrix: int, full: bool, primaries: int = 2, transfer: int = 2) -> bytes:
    return b"nclx" + struct.pack(">HHHB", primaries, transfer, matrix, 0x80 if full else 0)


def grid_descriptor(rows: int, cols: int, w: int, h: int, wide: bool = False) -> bytes:
    if wide:
        return bytes([0, 1, rows - 1, cols - 1]) + struct.pack(">II", w, h)
    return bytes([0, 0, rows - 1, cols - 1]) + struct.pack(">HH", w, h)


class Heif:
    """Items with data, types and properties; build() lays out the file."""

    def __init__(self, brands=(b"heic", b"mif1")):
        self.brands = list(brands)
        self.props = []      # (type, payload); property index = position + 1
        self.items = []      # dicts: id, type, data, props, cm, hidden, extra (bytes extents before the data)
        self.refs = []       # (type, from, [to])
        self.pitm = 1
        self.iloc_version = 1
        self.with_meta = True

    def prop(self, t: bytes, payload: bytes) -> int:
        self.props.append((t, payload))
        return len(self.props)

    def item(self, item_id: int, t: bytes, data: bytes, props=(), cm: int = 0, hidden: bool = False,
             bogus_first_extent: bool = False, no_iloc: bool = False):
        self.items.append(dict(id=item_id, type=t, data=data, props=list(props), cm=cm, hidden=hidden,
                               bogus=bogus_first_extent, no_iloc=no_iloc))

    def _meta(self, mdat_start: int) -> bytes:
        m = fbox(b"hdlr", 0, 0, b"\0\0\0\0pict" + b"\0" * 12 + b"\0")
        m += fbox(b"pitm", 0, 0, struct.pack(">H", self.pitm))
        infes = b"".join(fbox(b"infe", 2, 1 if it["hidden"] else 0, struct.pack(">HH", it["id"], 0) + it["type"] + b"\0")
                         for it in self.items)
        m += fbox(b"iinf", 0, 0, struct.pack(">H", len(self.items)) + infes)
        if self.refs:
            r = b"".join(box(t, struct.pack(">HH", f, len(to)) + b"".join(struct.pack(">H", x) for x in to))
                         for t, f, to in self.refs)
            m += fbox(b"iref", 0, 0, r)
        ipco = b"".join(box(t, p) for t, p in self.props)
        ipma = struct.pack(">I", len(self.items))
        for it in self.items:
            ipma += struct.pack(">HB", it["id"], len(it["props"])) + bytes(it["props"])
        m += box(b"iprp", box(b"ipco", ipco) + fbox(b"ipma", 0, 0, ipma))
        # iloc with 4-byte offsets / lengths, no base offset
        v = self.iloc_version
        listed = [it for it in self.items if not it["no_iloc"]]
        loc = bytes([0x44, 0x00]) + (struct.pack(">I", len(listed)) if v == 2 else struct.pack(">H", len(listed)))
        mdat_off, idat_off = mdat_start, 0
        for it in listed:
            loc += struct.pack(">I", it["id"]) if v == 2 else struct.pack(">H", it["id"])
            if v in (1, 2):
                loc += struct.pack(">H", it["cm"])
            loc += struct.pack(">H", 0)  # data_reference_index
            ext = []
            if it["bogus"]:
                ext.append((0, 16))
            if it["cm"] == 1:
                ext.append((idat_off, len(it["data"])))
                idat_off += len(it["data"])
            else:
                ext.append((mdat_off, len(it["data"])))
                mdat_off += len(it["data"])
            loc += struct.pack(">H", len(ext)) + b"".join(struct.pack(">II", o, l) for o, l in ext)
        m += fbox(b"iloc", v, 0, loc)
        idat = b"".join(it["data"] for it in listed if it["cm"] == 1)
        if idat:
            m += box(b"idat", idat)
        return fbox(b"meta", 0, 0, m)

    def build(self) -> bytes:
        ftyp = box(b"ftyp", self.brands[0] + b"\0\0\0\0" + b"".join(self.brands))
        if not self.with_meta:
            return ftyp
        mdat_payload = b"".join(it["data"] for it in self.items if it["cm"] == 0 and not it["no_iloc"])
        size = len(self._meta(0))
        start = len(ftyp) + size + 8
        return ftyp + self._meta(start) + box(b"mdat", mdat_payload)


def single_item_file(stream: bytes, display=None, length_size: int = 4, extra_props=(), colr=None, iloc_version=1,
                     cm: int = 0, bogus_first_extent: bool = False, hvcc_arrays=(32, 33, 34)) -> bytes:
    """One hvc1 primary item holding the first picture of an Annex B stream."""
    vps, sps, pps, pics = split_annexb(stream)
    f = Heif()
    f.iloc_version = iloc_version
    props = [0x80 | f.prop(b"hvcC", hvcc(vps, sps, pps, length_size, hvcc_arrays))]
    if display is not None:
        props.append(f.prop(b"ispe", ispe(*display)))
    if colr is not None:
        props.append(f.prop(b"colr", colr))
    for t, p in extra_props:
        props.append(f.prop(t, p))
    f.item(1, b"hvc1", sample(pics[0], length_size), props, cm=cm, bogus_first_extent=bogus_first_extent)
    return f.build()


def grid_file(stream: bytes, rows: int, cols: int, out_w: int, out_h: int, tile_display=None, grid_props=(),
              drop_iloc=(), damage=(), wide: bool = False) -> bytes:
    """A grid primary item over the stream's pictures (row-major)."""
    vps, sps, pps, pics = split_annexb(stream)
    f = Heif()
    hv = f.prop(b"hvcC", hvcc(vps, sps, pps))
    tprops = [0x80 | hv]
    if tile_display is not None:
        tprops.append(f.prop(b"ispe", ispe(*tile_display)))
    n = len(pics)
    for i in range(n):
        data = sample(pics[i])
        if i in damage:
            data = data[:8] + bytes(len(data) - 8)
        f.item(i + 1, b"hvc1", data, tprops, hidden=True, no_iloc=i in drop_iloc)
    gp = [f.prop(b"ispe", ispe(out_w, out_h))]
    for t, p in grid_props:
        gp.append(f.prop(t, p))
    f.item(n + 1, b"grid", grid_descriptor(rows, cols, out_w, out_h, wide), gp, cm=1)
    f.refs.append((b"dimg", n + 1, list(range(1, n + 1))))
    f.pitm = n + 1
    return f.build()
