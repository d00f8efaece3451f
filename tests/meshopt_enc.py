"""Test-side encoders of the EXT_meshopt_compression bitstreams (vertex codec
v0, index codec v0 / v1, index sequence codec), written from the published
format description, for round trips through the host decoders.

The encoders draw among every encoding the format allows (group modes of the
vertex codec; edge-fifo, vertex-fifo, table, slow-path, restart, last +- 1
and free-index codes of the index codec; both baselines of the sequence
codec), so that a round trip exercises each decoder branch rather than
one encoder's preferences.  The expected output of a triangle stream is the
rotation of each triangle that the encoder chose (the codec keeps triangle
order and winding, not the starting vertex)."""
import numpy as np

CODEAUX_TABLE = bytes([0x00, 0x76, 0x87, 0x56, 0x67, 0x78, 0xA9, 0x86, 0x65, 0x89, 0x68, 0x98, 0x01, 0x69, 0x00, 0x00])
M32 = 0xFFFFFFFF


def zigzag8(d):
    d = np.asarray(d, np.int64)
    return (((d << 1) ^ (d >> 7)) & 0xFF).astype(np.uint8)


def zigzag32(d):
    d &= M32
    s = d >> 31
    return ((d << 1) & M32) ^ (M32 if s else 0)


def varint(v):
    out = bytearray()
    while True:
        b = v & 127
        v >>= 7
        if v:
            out.append(b | 128)
        else:
            out.append(b)
            return bytes(out)


def _group(vals, mode):
    vals = [int(x) for x in vals]
    if mode == 0:
        assert not any(vals)
        return b""
    if mode == 3:
        return bytes(vals)
    bits = 2 if mode == 1 else 4
    mask = (1 << bits) - 1
    packed = bytearray(16 * bits // 8)
    esc = bytearray()
    for i, v in enumerate(vals):
        code = v if v < mask else mask
        if code == mask:
            esc.append(v)
        bit = i * bits
        packed[bit // 8] |= code << (8 - bits - bit % 8)
    return bytes(packed) + bytes(esc)


def encode_vertex_buffer(vertices, rng=None):
    """vertices: (n, stride) uint8.  rng: None picks the smallest group
    encoding, a Generator picks a random valid one."""
    v = np.ascontiguousarray(vertices, np.uint8)
    n, stride = v.shape
    assert stride % 4 == 0 and 0 < stride <= 256
    block = min((8192 // stride) & ~15, 256)
    out = bytearray([0xA0])
    last = v[0].astype(np.int64) if n else np.zeros(stride, np.int64)
    first_vertex = last.copy()
    for f in range(0, n, block):
        blk = v[f:f + block].astype(np.int64)
        m = len(blk)
        aligned = (m + 15) & ~15
        for k in range(stride):
            col = blk[:, k]
            prev = np.r_[last[k], col[:-1]]
            d = ((col - prev + 128) & 255) - 128
            zz = np.zeros(aligned, np.uint8)
            zz[:m] = zigzag8(d)
            groups = aligned // 16
            header = bytearray((groups + 3) // 4)
            body = bytearray()
            for g in range(groups):
                vals = zz[g * 16:(g + 1) * 16]
                options = [0] if not vals.any() else []
                options += [1, 2, 3]
                if rng is None:
                    mode = min(options, key=lambda mo: len(_group(vals, mo)) if (mo or not vals.any()) else 1 << 30)
                else:
                    mode = int(rng.choice(options))
                header[g // 4] |= mode << ((g % 4) * 2)
                body += _group(vals, mode)
            out += header + body
        last = blk[-1]
    tail = max(32, stride)
    out += bytes(tail - stride) + bytes(first_vertex.astype(np.uint8))
    return bytes(out)


class _IndexState:
    def __init__(self):
        self.edge = [(M32, M32)] * 16
        self.vf = [M32] * 16
        self.eo = self.vo = 0
        self.next = 0
        self.last = 0

    def edge_at(self, fe):
        return self.edge[(self.eo - 1 - fe) & 15]

    def vfifo_back(self, i):                 # i entries back from the newest (0 = newest)
        return self.vf[(self.vo - 1 - i) & 15]

    def push_v(self, v, cond=True):
        self.vf[self.vo] = v
        self.vo = (self.vo + (1 if cond else 0)) & 15

    def push_e(self, a, b):
        self.edge[self.eo] = (a, b)
        self.eo = (self.eo + 1) & 15


def encode_index_buffer(triangles, version=1, rng=None):
    """triangles: (T, 3) u32.  Returns (stream, decoded triangles the
    stream stands for)."""
    rng = rng or np.random.default_rng(0)
    tris = np.asarray(triangles, np.int64).reshape(-1, 3)
    st = _IndexState()
    fec_max = 13 if version >= 1 else 15
    codes, data, got = bytearray(), bytearray(), []

    def free(v):
        data.extend(varint(zigzag32((v - st.last) & M32)))
        st.last = v

    for tri in tris:
        rots = [tuple(int(x) for x in np.roll(tri, -r)) for r in range(3)]
        opts = []
        for (a, b, c) in rots:
            for fe in range(15):
                if st.edge_at(fe) == (a, b):
                    fecs = [15]
                    if c == st.next:
                        fecs.append(0)
                    fecs += [i for i in range(1, fec_max) if st.vfifo_back(i) == c]
                    if version >= 1 and c == (st.last - 1) & M32:
                        fecs.append(13)
                    if version >= 1 and c == (st.last + 1) & M32:
                        fecs.append(14)
                    opts += [("edge", (a, b, c), fe, f) for f in fecs]
            # fifo-miss forms: b and c by "next" (in order), vertex fifo or free index
            for fea in (0, 15):
                if fea == 0 and a != st.next:
                    continue
                nxt = st.next + (1 if fea == 0 else 0)
                febs = []
                if b == nxt:
                    febs.append((0, nxt + 1))
                febs += [(i + 1, nxt) for i in range(14) if st.vfifo_back(i) == b]
                febs.append((15, nxt))
                for feb, n2 in febs:
                    fecs = []
                    if c == n2:
                        fecs.append(0)
                    fecs += [i + 1 for i in range(14) if st.vfifo_back(i) == c]
                    fecs.append(15)
                    for fec in fecs:
                        aux = (feb << 4) | fec
                        if fea == 0 and aux in CODEAUX_TABLE[:14]:
                            opts.append(("table", (a, b, c), CODEAUX_TABLE.index(aux), aux))
                        if aux != 0:               # aux 0 on the slow path is the restart code
                            opts.append(("slow", (a, b, c), fea, aux))
        if tuple(int(x) for x in tri) == (0, 1, 2):
            opts.append(("restart", (0, 1, 2), 0, 0))
        # half of the time prefer what a size-minded encoder picks first (an
        # edge with the next vertex); uniform choice rarely lands there
        cheap = [o for o in opts if o[0] == "edge" and o[3] == 0]
        pool = cheap if cheap and rng.random() < 0.5 else opts
        kind, (a, b, c), p, q = pool[int(rng.integers(len(pool)))]
        if kind == "edge":
            codes.append((p << 4) | q)
            if q == 0:
                st.next += 1
                st.push_v(c, True)
            elif q < fec_max:
                st.push_v(c, False)
            else:
                if q == 15:
                    free(c)
                else:
                    st.last = c
                st.push_v(c, True)
            st.push_e(c, b)
            st.push_e(a, c)
        else:
            if kind == "table":
                codes.append(0xF0 | p)
                fea, aux = 0, q
            elif kind == "restart":
                codes.append(0xFE)
                data.append(0)
                st.next = 0
                fea, aux = 0, 0
            else:
                fea, aux = p, q
                codes.append(0xFE if fea == 0 else 0xFF)
                data.append(aux)
            feb, fec = aux >> 4, aux & 15
            if fea == 0:
                st.next += 1
            if feb == 0:
                st.next += 1
            if fec == 0:
                st.next += 1
            if fea == 15:
                free(a)
            if feb == 15:
                free(b)
            if fec == 15:
                free(c)
            st.push_v(a, True)
            st.push_v(b, feb in (0, 15))
            st.push_v(c, fec in (0, 15))
            st.push_e(b, a)
            st.push_e(c, b)
            st.push_e(a, c)
        got.append((a, b, c))
    stream = bytes([0xE0 | version]) + bytes(codes) + bytes(data) + CODEAUX_TABLE
    return stream, np.asarray(got, np.uint32).reshape(-1, 3)


def encode_index_sequence(indices, rng=None):
    rng = rng or np.random.default_rng(0)
    last = [0, 0]
    out = bytearray([0xD1])
    for v in np.asarray(indices, np.int64).ravel():
        k = int(rng.integers(2))
        zz = zigzag32((int(v) - last[k]) & M32)
        assert zz < (1 << 31)
        out += varint((zz << 1) | k)
        last[k] = int(v)
    return bytes(out) + bytes(4)
