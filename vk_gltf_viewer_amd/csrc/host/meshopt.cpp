// EXT_meshopt_compression decoders for glTF buffer views (host side).
//
// The reference decodes compressed buffer views with meshoptimizer
// (CompressedBufferDataAdapter::ExecuteRange, assets.cpp:111-171): vertex
// codec for ATTRIBUTES, index codec for TRIANGLES, index sequence codec for
// INDICES, then the OCTAHEDRAL / QUATERNION / EXPONENTIAL filter.  The
// meshoptimizer submodule is not vendored with the reference; this file is
// written from the published EXT_meshopt_compression bitstream description
// (attributes: vertex codec version 0, header 0xA0; triangles: index codec
// versions 0 and 1, header 0xE0 | version; indices: sequence codec version 1,
// header 0xD1).  Return codes follow meshoptimizer's: 0 success, -1 bad header
// or version, -2 a read past the end of the stream, -3 trailing data.
// Unlike the reference (assets.cpp:148-149, where rc is ignored) every caller
// in this package checks them.
//
// Compiled with -ffp-contract=off: the filters' float expressions are
// evaluated exactly as written (no fused multiply-add), as the published
// decoder specifies them.
#include "vkgv_host.h"

#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

constexpr size_t kGroup = 16;            // byte group of the vertex codec
constexpr size_t kBlockBytes = 8192;     // vertex block budget in bytes
constexpr size_t kBlockMaxVertices = 256;
constexpr size_t kVertexTail = 32;       // minimum tail: the first vertex, zero-padded in front

inline unsigned char unzigzag8(unsigned char v) { return static_cast<unsigned char>(-(v & 1) ^ (v >> 1)); }
inline uint32_t unzigzag32(uint32_t v) { return (v >> 1) ^ (0u - (v & 1)); }

size_t vertex_block_size(size_t vertex_size)
{
    size_t n = (kBlockBytes / vertex_size) & ~(kGroup - 1);
    return n < kBlockMaxVertices ? n : kBlockMaxVertices;
}

// One byte channel of one block: a header of 2-bit group modes (four groups
// per byte, first group in the low bits) followed by the groups.  Mode 0: the
// 16 deltas are zero; 1 / 2: 2- / 4-bit codes packed high bits first, the
// all-ones code meaning "the next escape byte" (escapes follow the packed
// codes); 3: 16 literal bytes.  Returns nullptr when a read would pass `end`.
const unsigned char* decode_bytes(const unsigned char* p, const unsigned char* end, unsigned char* out, size_t n)
{
    size_t groups = n / kGroup;
    size_t header_size = (groups + 3) / 4;
    if (static_cast<size_t>(end - p) < header_size) return nullptr;
    const unsigned char* header = p;
    p += header_size;
    for (size_t g = 0; g < groups; ++g) {
        int mode = (header[g / 4] >> ((g % 4) * 2)) & 3;
        unsigned char* o = out + g * kGroup;
        if (mode == 0) {
            std::memset(o, 0, kGroup);
        } else if (mode == 3) {
            if (static_cast<size_t>(end - p) < kGroup) return nullptr;
            std::memcpy(o, p, kGroup);
            p += kGroup;
        } else {
            const int bits = mode == 1 ? 2 : 4;
            const size_t packed = kGroup * bits / 8;
            if (static_cast<size_t>(end - p) < packed) return nullptr;
            const unsigned char* esc = p + packed;
            const unsigned mask = (1u << bits) - 1;
            for (size_t i = 0; i < kGroup; ++i) {
                size_t bit = i * bits;
                unsigned code = (p[bit / 8] >> (8 - bits - bit % 8)) & mask;
                if (code == mask) {
                    if (esc >= end) return nullptr;
                    o[i] = *esc++;
                } else {
                    o[i] = static_cast<unsigned char>(code);
                }
            }
            p = esc;
        }
    }
    return p;
}

// Index codec helpers: LEB128 varint, zigzag delta from the last free index.
bool read_varint(const unsigned char*& p, const unsigned char* end, uint32_t& v)
{
    v = 0;
    for (int shift = 0; shift < 35; shift += 7) {
        if (p >= end) return false;
        unsigned char b = *p++;
        v |= static_cast<uint32_t>(b & 127) << shift;
        if (b < 128) return true;
    }
    return true;   // a fifth byte carries the top bits; its continuation bit is ignored
}

inline void put_index(void* dst, size_t i, size_t index_size, uint32_t v)
{
    if (index_size == 2) static_cast<uint16_t*>(dst)[i] = static_cast<uint16_t>(v);
    else static_cast<uint32_t*>(dst)[i] = v;
}

inline int round_signed(float v) { return static_cast<int>(v + (v >= 0.f ? 0.5f : -0.5f)); }

template <typename T>
void filter_oct(T* d, size_t count)
{
    const float maxv = static_cast<float>((1 << (sizeof(T) * 8 - 1)) - 1);
    for (size_t i = 0; i < count; ++i) {
        float x = static_cast<float>(d[i * 4 + 0]);
        float y = static_cast<float>(d[i * 4 + 1]);
        float z = static_cast<float>(d[i * 4 + 2]) - std::fabs(x) - std::fabs(y);
        float t = z >= 0.f ? 0.f : z;          // fold the lower hemisphere
        x += x >= 0.f ? t : -t;
        y += y >= 0.f ? t : -t;
        float l = std::sqrt(x * x + y * y + z * z);
        float s = maxv / l;
        d[i * 4 + 0] = static_cast<T>(round_signed(x * s));
        d[i * 4 + 1] = static_cast<T>(round_signed(y * s));
        d[i * 4 + 2] = static_cast<T>(round_signed(z * s));
    }
}

}  // namespace

extern "C" {

int vkgv_meshopt_decode_vertex_buffer(void* destination, size_t vertex_count, size_t vertex_size,
                                      const unsigned char* buffer, size_t buffer_size)
{
    if (!destination || !buffer || vertex_size == 0 || vertex_size > 256 || vertex_size % 4) return -1;
    const size_t tail = vertex_size < kVertexTail ? kVertexTail : vertex_size;
    if (buffer_size < 1 + tail) return -2;
    if (buffer[0] != 0xA0) return -1;            // vertex codec version 0 only
    const unsigned char* p = buffer + 1;
    const unsigned char* end = buffer + buffer_size;
    const unsigned char* data_end = end - tail;  // the tail is not block data
    unsigned char last[256];
    std::memcpy(last, end - vertex_size, vertex_size);
    unsigned char deltas[kBlockMaxVertices];
    unsigned char* out = static_cast<unsigned char*>(destination);
    const size_t block = vertex_block_size(vertex_size);
    for (size_t first = 0; first < vertex_count; first += block) {
        size_t n = vertex_count - first < block ? vertex_count - first : block;
        size_t aligned = (n + kGroup - 1) & ~(kGroup - 1);
        for (size_t k = 0; k < vertex_size; ++k) {
            p = decode_bytes(p, data_end, deltas, aligned);
            if (!p) return -2;
            unsigned char v = last[k];
            for (size_t i = 0; i < n; ++i) {
                v = static_cast<unsigned char>(v + unzigzag8(deltas[i]));
                out[(first + i) * vertex_size + k] = v;
            }
            last[k] = v;
        }
    }
    return p == data_end ? 0 : -3;
}

int vkgv_meshopt_decode_index_buffer(void* destination, size_t index_count, size_t index_size,
                                     const unsigned char* buffer, size_t buffer_size)
{
    if (!destination || !buffer || index_count % 3 || (index_size != 2 && index_size != 4)) return -1;
    if (buffer_size < 1 + index_count / 3 + 16) return -2;
    if ((buffer[0] & 0xF0) != 0xE0 || (buffer[0] & 0x0F) > 1) return -1;
    const int version = buffer[0] & 0x0F;
    const unsigned fec_max = version >= 1 ? 13u : 15u;
    uint32_t edge[16][2], vfifo[16];
    std::memset(edge, 0xFF, sizeof(edge));
    std::memset(vfifo, 0xFF, sizeof(vfifo));
    unsigned eo = 0, vo = 0;                   // fifo write positions
    uint32_t next = 0, last = 0;
    const unsigned char* code = buffer + 1;
    const unsigned char* data = code + index_count / 3;
    const unsigned char* data_end = buffer + buffer_size - 16;
    const unsigned char* aux_table = data_end;
    auto push_v = [&](uint32_t v, bool cond) { vfifo[vo] = v; vo = (vo + (cond ? 1 : 0)) & 15; };
    auto push_e = [&](uint32_t a, uint32_t b) { edge[eo][0] = a; edge[eo][1] = b; eo = (eo + 1) & 15; };
    auto free_index = [&](uint32_t& out) -> bool {
        uint32_t v;
        if (!read_varint(data, data_end, v)) return false;
        last += unzigzag32(v);
        out = last;
        return true;
    };
    for (size_t i = 0; i < index_count; i += 3) {
        const unsigned c0 = *code++;
        uint32_t a, b, c;
        if (c0 < 0xF0) {
            // the triangle shares the edge fe entries back in the edge fifo
            const unsigned fe = c0 >> 4, fec = c0 & 15;
            a = edge[(eo - 1 - fe) & 15][0];
            b = edge[(eo - 1 - fe) & 15][1];
            if (fec < fec_max) {
                c = fec == 0 ? next++ : vfifo[(vo - 1 - fec) & 15];
                push_v(c, fec == 0);
            } else {
                if (fec == 15) {
                    if (!free_index(c)) return -2;
                } else {                       // 13: last - 1, 14: last + 1
                    last = fec == 13 ? last - 1 : last + 1;
                    c = last;
                }
                push_v(c, true);
            }
            push_e(c, b);
            push_e(a, c);
        } else {
            unsigned aux, fea;
            if (c0 < 0xFE) {
                aux = aux_table[c0 & 15];
                fea = 0;
            } else {
                if (data >= data_end) return -2;
                aux = *data++;
                fea = c0 == 0xFE ? 0 : 15;
                if (aux == 0) next = 0;        // restart: the triangle is (0, 1, 2)
            }
            const unsigned feb = aux >> 4, fec = aux & 15;
            // a, b, c take "next" in that order, before any free index is read
            a = fea == 0 ? next++ : 0;
            b = feb == 0 ? next++ : vfifo[(vo - feb) & 15];
            c = fec == 0 ? next++ : vfifo[(vo - fec) & 15];
            if (fea == 15 && !free_index(a)) return -2;
            if (feb == 15 && !free_index(b)) return -2;
            if (fec == 15 && !free_index(c)) return -2;
            push_v(a, true);
            push_v(b, feb == 0 || feb == 15);
            push_v(c, fec == 0 || fec == 15);
            push_e(b, a);
            push_e(c, b);
            push_e(a, c);
        }
        put_index(destination, i + 0, index_size, a);
        put_index(destination, i + 1, index_size, b);
        put_index(destination, i + 2, index_size, c);
    }
    return data == data_end ? 0 : -3;
}

int vkgv_meshopt_decode_index_sequence(void* destination, size_t index_count, size_t index_size,
                                       const unsigned char* buffer, size_t buffer_size)
{
    if (!destination || !buffer || (index_size != 2 && index_size != 4)) return -1;
    if (buffer_size < 1 + index_count + 4) return -2;
    if ((buffer[0] & 0xF0) != 0xD0 || (buffer[0] & 0x0F) > 1) return -1;
    const unsigned char* data = buffer + 1;
    const unsigned char* data_end = buffer + buffer_size - 4;
    uint32_t last[2] = { 0, 0 };
    for (size_t i = 0; i < index_count; ++i) {
        uint32_t v;
        if (!read_varint(data, data_end, v)) return -2;
        // low bit: which of the two baselines the delta applies to
        uint32_t& base = last[v & 1];
        base += unzigzag32(v >> 1);
        put_index(destination, i, index_size, base);
    }
    return data == data_end ? 0 : -3;
}

void vkgv_meshopt_decode_filter_oct(void* buffer, size_t count, size_t stride)
{
    if (stride == 4) filter_oct(static_cast<int8_t*>(buffer), count);
    else if (stride == 8) filter_oct(static_cast<int16_t*>(buffer), count);
}

void vkgv_meshopt_decode_filter_quat(void* buffer, size_t count, size_t stride)
{
    if (stride != 8) return;
    int16_t* d = static_cast<int16_t*>(buffer);
    const float scale = 1.f / std::sqrt(2.f);
    for (size_t i = 0; i < count; ++i) {
        const int sf = d[i * 4 + 3] | 3;     // the scale lives in the high bits of the fourth component
        const float ss = scale / static_cast<float>(sf);
        float x = static_cast<float>(d[i * 4 + 0]) * ss;
        float y = static_cast<float>(d[i * 4 + 1]) * ss;
        float z = static_cast<float>(d[i * 4 + 2]) * ss;
        float ww = 1.f - x * x - y * y - z * z;
        float w = std::sqrt(ww >= 0.f ? ww : 0.f);
        const int qc = d[i * 4 + 3] & 3;     // which component was dropped (the largest)
        const int xf = round_signed(x * 32767.f), yf = round_signed(y * 32767.f);
        const int zf = round_signed(z * 32767.f), wf = round_signed(w * 32767.f);
        d[i * 4 + ((qc + 1) & 3)] = static_cast<int16_t>(xf);
        d[i * 4 + ((qc + 2) & 3)] = static_cast<int16_t>(yf);
        d[i * 4 + ((qc + 3) & 3)] = static_cast<int16_t>(zf);
        d[i * 4 + ((qc + 0) & 3)] = static_cast<int16_t>(wf);
    }
}

void vkgv_meshopt_decode_filter_exp(void* buffer, size_t count, size_t stride)
{
    if (stride % 4) return;
    uint32_t* d = static_cast<uint32_t*>(buffer);
    const size_t n = count * (stride / 4);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t v = d[i];
        const int32_t m = static_cast<int32_t>(v << 8) >> 8;      // 24-bit signed mantissa
        const int32_t e = static_cast<int32_t>(v) >> 24;          // 8-bit signed exponent
        uint32_t pbits = static_cast<uint32_t>(e + 127) << 23;    // 2^e
        float p;
        std::memcpy(&p, &pbits, 4);
        float r = p * static_cast<float>(m);
        std::memcpy(&d[i], &r, 4);
    }
}

int vkgv_meshopt_decode(void* destination, size_t count, size_t stride, const unsigned char* buffer,
                        size_t buffer_size, int mode, int filter)
{
    int rc;
    switch (mode) {
    case VKGV_MESHOPT_ATTRIBUTES:
        rc = vkgv_meshopt_decode_vertex_buffer(destination, count, stride, buffer, buffer_size);
        break;
    case VKGV_MESHOPT_TRIANGLES:
        rc = vkgv_meshopt_decode_index_buffer(destination, count, stride, buffer, buffer_size);
        break;
    case VKGV_MESHOPT_INDICES:
        rc = vkgv_meshopt_decode_index_sequence(destination, count, stride, buffer, buffer_size);
        break;
    default:
        return -1;
    }
    if (rc != 0) return rc;
    if (filter != VKGV_MESHOPT_FILTER_NONE && mode != VKGV_MESHOPT_ATTRIBUTES) return -1;
    if ((filter == VKGV_MESHOPT_FILTER_OCTAHEDRAL && stride != 4 && stride != 8) ||
        (filter == VKGV_MESHOPT_FILTER_QUATERNION && stride != 8))
        return -1;
    switch (filter) {
    case VKGV_MESHOPT_FILTER_NONE: break;
    case VKGV_MESHOPT_FILTER_OCTAHEDRAL: vkgv_meshopt_decode_filter_oct(destination, count, stride); break;
    case VKGV_MESHOPT_FILTER_QUATERNION: vkgv_meshopt_decode_filter_quat(destination, count, stride); break;
    case VKGV_MESHOPT_FILTER_EXPONENTIAL: vkgv_meshopt_decode_filter_exp(destination, count, stride); break;
    default: return -1;
    }
    return 0;
}

}  // extern "C"
