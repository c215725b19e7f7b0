// lscomp_container.cpp -- self-describing container for lsCOMP streams.
//
// The reference passes dims, bins and stream sizes out of band
// (lsCOMP/include/cuLSZ_entry.h:8-11); a stored or transmitted stream needs
// them.  A container holds any number of streams (one volume, or the frames
// of a batch or of the served form) with their parameters, little-endian:
//
//   header, 48 bytes:  "LSCP" | u16 version (1) | u16 header bytes (48) |
//                      u32 items | u32 index entry bytes (64) |
//                      u64 total bytes | 24 zero bytes
//   index, 64 bytes per item:
//                      u32 x, y, z | u32 element bytes (2 or 4) |
//                      u32 bins[4] | f32 pooling | u32 0 |
//                      u64 stream offset | u64 stream bytes | u64 tag
//   streams, each at a 16-byte-aligned offset (zero padding between).
//
// Host memory only: the caller copies device streams in and out.
#include <cstring>

#include "../../include/lscomp.h"

namespace {

constexpr uint32_t kHeader = 48, kEntry = 64, kVersion = 1;
constexpr unsigned char kMagic[4] = {'L', 'S', 'C', 'P'};

inline void put32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }
inline void put64(uint8_t *p, uint64_t v) { for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i)); }
inline uint32_t get32(const uint8_t *p)
{
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) v |= (uint32_t)p[i] << (8 * i);
    return v;
}
inline uint64_t get64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

inline uint64_t align16(uint64_t v) { return (v + 15) & ~15ull; }

uint64_t blocks_of(lscomp_dim3 d)
{
    return (uint64_t)d.x * (((uint64_t)d.y + 7) / 8) * (((uint64_t)d.z + 7) / 8);
}

// Parameters of one item as the codec accepts them, and a stream size
// between its header section and the worst case.
bool item_ok(const lscomp_container_item &it)
{
    if (it.dims.x == 0 || it.dims.y == 0 || it.dims.z == 0) return false;
    if (it.elem_bytes != 2 && it.elem_bytes != 4) return false;
    const lscomp_bins4 &b = it.bins;
    if (!b.x || !b.y || !b.z || !b.w || !(b.x <= b.y && b.y <= b.z && b.z <= b.w)) return false;
    if (!(it.pooling_th >= 0.0f && it.pooling_th <= 1.0f)) return false;
    const uint64_t nb = blocks_of(it.dims);
    if (nb > 0x7fffffffull) return false;   // the codec's limit per volume
    const uint64_t R = (nb + 3) / 4 * 4;
    const uint64_t worst = R + (it.elem_bytes == 2 ? 128 : 248) * nb;
    return it.stream_bytes >= R && it.stream_bytes <= worst;
}

}  // namespace

extern "C" {

size_t lscomp_container_bytes(const lscomp_container_item *items, uint32_t n_items)
{
    if (n_items && !items) return 0;
    uint64_t pos = align16(kHeader + (uint64_t)kEntry * n_items);
    for (uint32_t i = 0; i < n_items; i++) pos = align16(pos + items[i].stream_bytes);
    return (size_t)pos;
}

int lscomp_container_pack(lscomp_container_item *items, uint32_t n_items, const uint8_t *const *streams,
                          uint8_t *out, size_t capacity, size_t *out_bytes)
{
    if (!out_bytes || (n_items && (!items || !streams)) || !out) return LSCOMP_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n_items; i++)
        if (!item_ok(items[i]) || !streams[i]) return LSCOMP_ERR_INVALID_ARG;
    const size_t total = lscomp_container_bytes(items, n_items);
    *out_bytes = total;
    if (capacity < total) return LSCOMP_ERR_CAPACITY;
    std::memset(out, 0, total);
    std::memcpy(out, kMagic, 4);
    out[4] = (uint8_t)kVersion;
    out[6] = (uint8_t)kHeader;
    put32(out + 8, n_items);
    put32(out + 12, kEntry);
    put64(out + 16, total);
    uint64_t pos = align16(kHeader + (uint64_t)kEntry * n_items);
    for (uint32_t i = 0; i < n_items; i++) {
        lscomp_container_item &it = items[i];
        it.offset = pos;
        uint8_t *e = out + kHeader + (uint64_t)kEntry * i;
        put32(e, it.dims.x);
        put32(e + 4, it.dims.y);
        put32(e + 8, it.dims.z);
        put32(e + 12, it.elem_bytes);
        put32(e + 16, it.bins.x);
        put32(e + 20, it.bins.y);
        put32(e + 24, it.bins.z);
        put32(e + 28, it.bins.w);
        uint32_t pb;
        std::memcpy(&pb, &it.pooling_th, 4);
        put32(e + 32, pb);
        put64(e + 40, it.offset);
        put64(e + 48, it.stream_bytes);
        put64(e + 56, it.tag);
        std::memcpy(out + pos, streams[i], it.stream_bytes);
        pos = align16(pos + it.stream_bytes);
    }
    return LSCOMP_OK;
}

int lscomp_container_parse(const uint8_t *buf, size_t len, lscomp_container_item *items, uint32_t max_items,
                           uint32_t *n_items)
{
    if (!buf || !n_items) return LSCOMP_ERR_INVALID_ARG;
    *n_items = 0;
    if (len < kHeader || std::memcmp(buf, kMagic, 4) != 0) return LSCOMP_ERR_CORRUPT;
    const uint32_t version = (uint32_t)buf[4] | ((uint32_t)buf[5] << 8);
    const uint32_t hbytes = (uint32_t)buf[6] | ((uint32_t)buf[7] << 8);
    const uint32_t n = get32(buf + 8), ebytes = get32(buf + 12);
    const uint64_t total = get64(buf + 16);
    if (version != kVersion || hbytes != kHeader || ebytes != kEntry) return LSCOMP_ERR_CORRUPT;
    if (total > len || (uint64_t)kHeader + (uint64_t)kEntry * n > total) return LSCOMP_ERR_CORRUPT;
    *n_items = n;
    if (n > max_items) return LSCOMP_ERR_CAPACITY;
    if (n && !items) return LSCOMP_ERR_INVALID_ARG;
    const uint64_t data0 = align16(kHeader + (uint64_t)kEntry * n);
    uint64_t prev_end = data0;
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t *e = buf + kHeader + (uint64_t)kEntry * i;
        lscomp_container_item it{};
        it.dims = {get32(e), get32(e + 4), get32(e + 8)};
        it.elem_bytes = get32(e + 12);
        it.bins = {get32(e + 16), get32(e + 20), get32(e + 24), get32(e + 28)};
        const uint32_t pb = get32(e + 32);
        std::memcpy(&it.pooling_th, &pb, 4);
        it.offset = get64(e + 40);
        it.stream_bytes = get64(e + 48);
        it.tag = get64(e + 56);
        // streams in index order, 16-byte aligned, inside the container, not overlapping
        if (it.offset % 16 || it.offset < prev_end || it.offset > total || it.stream_bytes > total - it.offset ||
            !item_ok(it)) {
            *n_items = i;
            return LSCOMP_ERR_CORRUPT;
        }
        prev_end = it.offset + it.stream_bytes;
        items[i] = it;
    }
    return LSCOMP_OK;
}

}  // extern "C"
