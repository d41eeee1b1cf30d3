// PNG and JPEG decoding to RGBA8 (see image_decode.hpp for the conversions).
#include "image_decode.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <iterator>

namespace wpt {
namespace {

bool fail(std::string* error, const std::string& message) {
    if (error) *error = message;
    return false;
}

// Largest decoded image: 2^28 pixels (1 GiB of RGBA8, a 16384^2 texture).
constexpr uint64_t kMaxPixels = 1ull << 28;

// LSB-first bit reader over a byte range; reads past the end return zero
// bits and are detected by overrun().
class BitReader {
public:
    BitReader(const uint8_t* data, size_t size) : data_(data), size_(size) {}

    uint32_t peek(int count) {
        refill();
        return static_cast<uint32_t>(buffer_ & ((1ull << count) - 1));
    }
    void skip(int count) {
        buffer_ >>= count;
        bits_ -= count;
    }
    uint32_t take(int count) {
        if (count == 0) return 0;
        uint32_t v = peek(count);
        skip(count);
        return v;
    }
    void alignToByte() { skip(bits_ & 7); }
    // Bytes of the stream not yet consumed (after alignToByte).
    size_t bytePosition() const { return pos_ - static_cast<size_t>(bits_ / 8); }
    void seekByte(size_t position) {
        pos_ = position;
        buffer_ = 0;
        bits_ = 0;
    }
    bool overrun() const { return pos_ * 8 - static_cast<size_t>(bits_) > size_ * 8; }

private:
    void refill() {
        while (bits_ <= 56) {
            uint64_t byte = pos_ < size_ ? data_[pos_] : 0;
            ++pos_;
            buffer_ |= byte << bits_;
            bits_ += 8;
        }
    }
    const uint8_t* data_;
    size_t size_;
    size_t pos_ = 0;
    uint64_t buffer_ = 0;
    int bits_ = 0;
};

// Canonical Huffman decoder: a 10-bit direct table for short codes and a
// code-length walk for the rest (RFC 1951 section 3.2.2).
class Huffman {
public:
    static constexpr int kFastBits = 10;

    bool build(const uint8_t* lengths, int n) {
        std::array<uint16_t, 16> offsets{};
        counts_.fill(0);
        for (int i = 0; i < n; ++i) ++counts_[lengths[i]];
        counts_[0] = 0;
        int left = 1;
        for (int len = 1; len < 16; ++len) {
            left = left * 2 - counts_[len];
            if (left < 0) return false;  // over-subscribed
        }
        for (int len = 1; len < 15; ++len) offsets[len + 1] = static_cast<uint16_t>(offsets[len] + counts_[len]);
        for (int i = 0; i < n; ++i)
            if (lengths[i]) symbols_[offsets[lengths[i]]++] = static_cast<uint16_t>(i);
        fast_.fill(0);
        uint32_t code = 0;
        int index = 0;
        for (int len = 1; len < 16; ++len) {
            for (int k = 0; k < counts_[len]; ++k, ++index, ++code) {
                if (len > kFastBits) continue;
                uint32_t reversed = 0;
                for (int b = 0; b < len; ++b) reversed |= ((code >> b) & 1u) << (len - 1 - b);
                uint16_t entry = static_cast<uint16_t>(symbols_[index] | (len << 12));
                for (uint32_t fill = reversed; fill < (1u << kFastBits); fill += 1u << len) fast_[fill] = entry;
            }
            code <<= 1;
        }
        return true;
    }

    // Symbol, or -1 for a code outside the table.
    int decode(BitReader& in) const {
        uint16_t entry = fast_[in.peek(kFastBits)];
        if (entry) {
            in.skip(entry >> 12);
            return entry & 0xfff;
        }
        int code = 0, first = 0, index = 0;
        uint32_t bits = in.peek(15);
        for (int len = 1; len < 16; ++len) {
            code |= static_cast<int>((bits >> (len - 1)) & 1u);
            int count = counts_[len];
            if (code - first < count) {
                in.skip(len);
                return symbols_[index + code - first];
            }
            index += count;
            first = (first + count) << 1;
            code <<= 1;
        }
        return -1;
    }

private:
    std::array<uint16_t, 16> counts_{};
    std::array<uint16_t, 288> symbols_{};
    std::array<uint16_t, 1 << kFastBits> fast_{};
};

constexpr uint16_t kLengthBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83,
    99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t kLengthExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5,
    5, 0};
constexpr uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769,
    1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11,
    12, 12, 13, 13};

// Outcome of one compressed block.
enum class BlockEnd { kFailed, kEndOfBlock, kLimit };

// Decodes one compressed block.  `limit` bounds the inflated size (the
// image's filtered bytes): decoding stops once that many bytes exist, since
// data after the last row is ignored, so no stream can grow the output
// further.
BlockEnd inflateBlock(BitReader& in, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out, size_t limit,
    std::string* error) {
    auto failed = [&](const char* message) {
        fail(error, message);
        return BlockEnd::kFailed;
    };
    for (;;) {
        if (out.size() >= limit) return BlockEnd::kLimit;
        int sym = lit.decode(in);
        if (sym < 0) return failed("PNG: bad Huffman code");
        if (in.overrun()) return failed("PNG: truncated zlib stream");
        if (sym < 256) {
            out.push_back(static_cast<uint8_t>(sym));
            continue;
        }
        if (sym == 256) return BlockEnd::kEndOfBlock;
        sym -= 257;
        if (sym >= 29) return failed("PNG: bad length symbol");
        size_t length = kLengthBase[sym] + in.take(kLengthExtra[sym]);
        int d = dist.decode(in);
        if (d < 0 || d >= 30) return failed("PNG: bad distance code");
        size_t distance = kDistBase[d] + in.take(kDistExtra[d]);
        if (distance > out.size()) return failed("PNG: distance before the start of the data");
        if (in.overrun()) return failed("PNG: truncated zlib stream");
        length = std::min(length, limit - out.size());
        size_t from = out.size() - distance;
        for (size_t k = 0; k < length; ++k) out.push_back(out[from + k]);
    }
}

bool dynamicTables(BitReader& in, Huffman& lit, Huffman& dist, std::string* error) {
    static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
    int hlit = static_cast<int>(in.take(5)) + 257;
    int hdist = static_cast<int>(in.take(5)) + 1;
    int hclen = static_cast<int>(in.take(4)) + 4;
    uint8_t codeLengths[19] = {};
    for (int i = 0; i < hclen; ++i) codeLengths[order[i]] = static_cast<uint8_t>(in.take(3));
    Huffman lengthCode;
    if (!lengthCode.build(codeLengths, 19)) return fail(error, "PNG: bad code lengths");
    uint8_t lengths[288 + 32] = {};  // HLIT <= 288, HDIST <= 32
    int n = 0;
    while (n < hlit + hdist) {
        int sym = lengthCode.decode(in);
        if (sym < 0) return fail(error, "PNG: bad code lengths");
        if (sym < 16) {
            lengths[n++] = static_cast<uint8_t>(sym);
            continue;
        }
        int repeat;
        uint8_t value = 0;
        if (sym == 16) {
            if (n == 0) return fail(error, "PNG: bad code lengths");
            value = lengths[n - 1];
            repeat = 3 + static_cast<int>(in.take(2));
        } else if (sym == 17) {
            repeat = 3 + static_cast<int>(in.take(3));
        } else {
            repeat = 11 + static_cast<int>(in.take(7));
        }
        if (n + repeat > hlit + hdist) return fail(error, "PNG: bad code lengths");
        for (int k = 0; k < repeat; ++k) lengths[n++] = value;
    }
    if (in.overrun()) return fail(error, "PNG: truncated zlib stream");
    if (!lit.build(lengths, hlit) || !dist.build(lengths + hlit, hdist)) return fail(error, "PNG: bad code lengths");
    return true;
}

}  // namespace

bool inflateZlib(const uint8_t* data, size_t size, std::vector<uint8_t>& out, size_t expectedSize,
    std::string* error) {
    out.clear();
    // Deflate expands at most 1032:1, so a tiny stream cannot claim a huge
    // reservation through its header alone.
    out.reserve(std::min(expectedSize, size * 1032));
    if (size < 2) return fail(error, "PNG: truncated zlib stream");
    uint32_t cmf = data[0], flg = data[1];
    if ((cmf * 256 + flg) % 31 != 0) return fail(error, "PNG: bad zlib header");
    if (flg & 32) return fail(error, "PNG: preset dictionary not allowed");
    if ((cmf & 15) != 8) return fail(error, "PNG: bad compression method");
    BitReader in(data + 2, size - 2);
    Huffman lit, dist;
    for (bool last = false; !last;) {
        last = in.take(1) != 0;
        uint32_t type = in.take(2);
        if (type == 0) {
            in.alignToByte();
            uint32_t len = in.take(16), nlen = in.take(16);
            if ((len ^ 0xffffu) != nlen) return fail(error, "PNG: corrupt stored block");
            size_t at = in.bytePosition();
            if (at + len > size - 2) return fail(error, "PNG: truncated zlib stream");
            size_t keep = std::min<size_t>(len, expectedSize - std::min(expectedSize, out.size()));
            out.insert(out.end(), data + 2 + at, data + 2 + at + keep);
            in.seekByte(at + len);
            if (out.size() >= expectedSize) return true;
        } else if (type == 1) {
            uint8_t lengths[288 + 32];
            std::memset(lengths, 8, 144);
            std::memset(lengths + 144, 9, 112);
            std::memset(lengths + 256, 7, 24);
            std::memset(lengths + 280, 8, 8);
            std::memset(lengths + 288, 5, 32);
            lit.build(lengths, 288);
            dist.build(lengths + 288, 32);
            BlockEnd end = inflateBlock(in, lit, dist, out, expectedSize, error);
            if (end == BlockEnd::kFailed) return false;
            if (end == BlockEnd::kLimit) return true;
        } else if (type == 2) {
            if (!dynamicTables(in, lit, dist, error)) return false;
            BlockEnd end = inflateBlock(in, lit, dist, out, expectedSize, error);
            if (end == BlockEnd::kFailed) return false;
            if (end == BlockEnd::kLimit) return true;
        } else {
            return fail(error, "PNG: bad block type");
        }
        if (in.overrun()) return fail(error, "PNG: truncated zlib stream");
    }
    return true;
}

namespace {

uint32_t be32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

uint8_t paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return static_cast<uint8_t>(a);
    return static_cast<uint8_t>(pb <= pc ? b : c);
}

// Undoes the filters of a (sub)image of `rows` scanlines of `rowBytes`
// bytes each, in place in `raw` from `offset`; filterBytes is the byte
// distance to the left neighbour.  Returns the bytes consumed, 0 on error.
size_t unfilter(std::vector<uint8_t>& raw, size_t offset, uint32_t rows, size_t rowBytes, size_t filterBytes,
    std::vector<uint8_t>& pixels, std::string* error) {
    pixels.assign(static_cast<size_t>(rows) * rowBytes, 0);
    const uint8_t* src = raw.data() + offset;
    for (uint32_t y = 0; y < rows; ++y, src += rowBytes + 1) {
        uint8_t* cur = pixels.data() + static_cast<size_t>(y) * rowBytes;
        const uint8_t* prev = y ? cur - rowBytes : nullptr;
        const uint8_t* in = src + 1;
        switch (src[0]) {
        case 0:
            std::memcpy(cur, in, rowBytes);
            break;
        case 1:
            for (size_t i = 0; i < rowBytes; ++i) cur[i] = static_cast<uint8_t>(in[i] + (i >= filterBytes ? cur[i - filterBytes] : 0));
            break;
        case 2:
            for (size_t i = 0; i < rowBytes; ++i) cur[i] = static_cast<uint8_t>(in[i] + (prev ? prev[i] : 0));
            break;
        case 3:
            for (size_t i = 0; i < rowBytes; ++i) {
                int a = i >= filterBytes ? cur[i - filterBytes] : 0, b = prev ? prev[i] : 0;
                cur[i] = static_cast<uint8_t>(in[i] + ((a + b) >> 1));
            }
            break;
        case 4:
            for (size_t i = 0; i < rowBytes; ++i) {
                int a = i >= filterBytes ? cur[i - filterBytes] : 0, b = prev ? prev[i] : 0;
                int c = (prev && i >= filterBytes) ? prev[i - filterBytes] : 0;
                cur[i] = static_cast<uint8_t>(in[i] + paeth(a, b, c));
            }
            break;
        default:
            fail(error, "PNG: invalid filter");
            return 0;
        }
    }
    return static_cast<size_t>(rows) * (rowBytes + 1);
}

struct PngInfo {
    uint32_t width = 0, height = 0;
    int depth = 0, color = 0, channels = 0;
    bool interlaced = false;
    std::array<uint8_t, 1024> palette{};  // RGBA
    uint32_t paletteSize = 0;
    bool hasKey = false;
    uint16_t key[3] = {};
};

// Sample `c` of pixel x in an unfiltered row.
uint32_t sample(const uint8_t* row, uint32_t x, int c, const PngInfo& p) {
    if (p.depth == 16) {
        const uint8_t* s = row + (static_cast<size_t>(x) * p.channels + c) * 2;
        return (uint32_t(s[0]) << 8) | s[1];
    }
    if (p.depth == 8) return row[static_cast<size_t>(x) * p.channels + c];
    size_t bit = static_cast<size_t>(x) * p.depth;  // one channel below 8 bits
    return (row[bit >> 3] >> (8 - p.depth - (bit & 7))) & ((1u << p.depth) - 1);
}

void convertPixel(const uint8_t* row, uint32_t x, const PngInfo& p, uint8_t* out) {
    static const uint8_t greyScale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (p.depth == 16) return static_cast<uint8_t>(v >> 8);
        return static_cast<uint8_t>(v);
    };
    switch (p.color) {
    case 0: {
        uint32_t g = sample(row, x, 0, p);
        uint32_t v = p.depth < 8 ? g * greyScale[p.depth] : g;
        out[0] = out[1] = out[2] = to8(v);
        bool keyed = false;
        if (p.hasKey) keyed = p.depth == 16 ? g == p.key[0] : v == p.key[0];
        out[3] = keyed ? 0 : 255;
        break;
    }
    case 2: {
        uint32_t r = sample(row, x, 0, p), g = sample(row, x, 1, p), b = sample(row, x, 2, p);
        out[0] = to8(r);
        out[1] = to8(g);
        out[2] = to8(b);
        out[3] = (p.hasKey && r == p.key[0] && g == p.key[1] && b == p.key[2]) ? 0 : 255;
        break;
    }
    case 3: {
        uint32_t i = sample(row, x, 0, p);
        if (i < p.paletteSize) {
            std::memcpy(out, &p.palette[i * 4], 4);
        } else {  // index past the palette: black, opaque
            out[0] = out[1] = out[2] = 0;
            out[3] = 255;
        }
        break;
    }
    case 4:
        out[0] = out[1] = out[2] = to8(sample(row, x, 0, p));
        out[3] = to8(sample(row, x, 1, p));
        break;
    default:
        for (int c = 0; c < 4; ++c) out[c] = to8(sample(row, x, c, p));
        break;
    }
}

bool decodePng(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height,
    std::string* error) {
    PngInfo p;
    std::vector<uint8_t> compressed;
    bool sawHeader = false, sawPalette = false, sawEnd = false;
    size_t at = 8;
    while (!sawEnd) {
        if (at + 12 > size) return fail(error, "PNG: truncated file");
        uint32_t len = be32(data + at);
        uint32_t type = be32(data + at + 4);
        const uint8_t* body = data + at + 8;
        if (len > size - at - 12) return fail(error, "PNG: truncated chunk");
        auto is = [&](const char* name) { return type == be32(reinterpret_cast<const uint8_t*>(name)); };
        if (!sawHeader && !is("IHDR")) return fail(error, "PNG: first chunk is not IHDR");
        if (is("IHDR")) {
            if (sawHeader || len != 13) return fail(error, "PNG: bad IHDR");
            sawHeader = true;
            p.width = be32(body);
            p.height = be32(body + 4);
            p.depth = body[8];
            p.color = body[9];
            if (p.width == 0 || p.height == 0) return fail(error, "PNG: zero-sized image");
            if (p.width > (1u << 24) || p.height > (1u << 24) ||
                static_cast<uint64_t>(p.width) * p.height > kMaxPixels)
                return fail(error, "PNG: image too large");
            bool depthOk = false;
            switch (p.color) {
            case 0: depthOk = p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8 || p.depth == 16; p.channels = 1; break;
            case 3: depthOk = p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8; p.channels = 1; break;
            case 2: depthOk = p.depth == 8 || p.depth == 16; p.channels = 3; break;
            case 4: depthOk = p.depth == 8 || p.depth == 16; p.channels = 2; break;
            case 6: depthOk = p.depth == 8 || p.depth == 16; p.channels = 4; break;
            default: break;
            }
            if (!depthOk) return fail(error, "PNG: bad colour type or bit depth");
            if (body[10] != 0 || body[11] != 0 || body[12] > 1) return fail(error, "PNG: bad compression, filter or interlace method");
            p.interlaced = body[12] == 1;
        } else if (is("PLTE")) {
            if (len % 3 != 0 || len / 3 > 256 || len == 0) return fail(error, "PNG: bad palette");
            p.paletteSize = len / 3;
            for (uint32_t i = 0; i < p.paletteSize; ++i) {
                std::memcpy(&p.palette[i * 4], body + i * 3, 3);
                p.palette[i * 4 + 3] = 255;
            }
            sawPalette = true;
        } else if (is("tRNS")) {
            if (!compressed.empty()) return fail(error, "PNG: tRNS after IDAT");
            if (p.color == 3) {
                if (!sawPalette) return fail(error, "PNG: tRNS before PLTE");
                if (len > p.paletteSize) return fail(error, "PNG: bad tRNS length");
                for (uint32_t i = 0; i < len; ++i) p.palette[i * 4 + 3] = body[i];
            } else {
                if (p.color == 4 || p.color == 6) return fail(error, "PNG: tRNS with an alpha channel");
                if (len != static_cast<uint32_t>(p.channels) * 2) return fail(error, "PNG: bad tRNS length");
                static const uint8_t greyScale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
                p.hasKey = true;
                for (int k = 0; k < p.channels; ++k) {
                    uint16_t v = static_cast<uint16_t>((body[2 * k] << 8) | body[2 * k + 1]);
                    // stb_image keeps the low byte of 8-bit and sub-byte keys, scaled like the greys.
                    p.key[k] = p.depth == 16 ? v : static_cast<uint16_t>((v & 255) * greyScale[p.depth]);
                }
            }
        } else if (is("IDAT")) {
            if (p.color == 3 && p.paletteSize == 0) return fail(error, "PNG: palette image without PLTE");
            compressed.insert(compressed.end(), body, body + len);
        } else if (is("IEND")) {
            sawEnd = true;
        } else if (is("CgBI")) {
            return fail(error, "PNG: Apple CgBI images are not supported");
        } else if ((type & (1u << 29)) == 0) {
            return fail(error, "PNG: unknown critical chunk");
        }
        at += 12 + static_cast<size_t>(len);
    }
    if (compressed.empty()) return fail(error, "PNG: no image data");

    const size_t bitsPerPixel = static_cast<size_t>(p.channels) * p.depth;
    const size_t filterBytes = bitsPerPixel >= 8 ? bitsPerPixel / 8 : 1;
    static const uint32_t xOrigin[7] = {0, 4, 0, 2, 0, 1, 0}, yOrigin[7] = {0, 0, 4, 0, 2, 0, 1};
    static const uint32_t xStep[7] = {8, 8, 4, 4, 2, 2, 1}, yStep[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = p.interlaced ? 7 : 1;
    size_t expected = 0;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t pw = p.interlaced ? (p.width - xOrigin[pass] + xStep[pass] - 1) / xStep[pass] : p.width;
        uint32_t ph = p.interlaced ? (p.height - yOrigin[pass] + yStep[pass] - 1) / yStep[pass] : p.height;
        if (p.interlaced && (p.width <= xOrigin[pass] || p.height <= yOrigin[pass])) continue;
        expected += static_cast<size_t>(ph) * (1 + (pw * bitsPerPixel + 7) / 8);
    }
    std::vector<uint8_t> raw;
    if (!inflateZlib(compressed.data(), compressed.size(), raw, expected, error)) return false;
    if (raw.size() < expected) return fail(error, "PNG: not enough pixels");

    width = p.width;
    height = p.height;
    rgba.assign(static_cast<size_t>(width) * height * 4, 0);
    std::vector<uint8_t> pixels;
    size_t offset = 0;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t x0 = p.interlaced ? xOrigin[pass] : 0, y0 = p.interlaced ? yOrigin[pass] : 0;
        uint32_t dx = p.interlaced ? xStep[pass] : 1, dy = p.interlaced ? yStep[pass] : 1;
        if (p.width <= x0 || p.height <= y0) continue;
        uint32_t pw = (p.width - x0 + dx - 1) / dx, ph = (p.height - y0 + dy - 1) / dy;
        size_t rowBytes = (pw * bitsPerPixel + 7) / 8;
        size_t used = unfilter(raw, offset, ph, rowBytes, filterBytes, pixels, error);
        if (used == 0) return false;
        offset += used;
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t* row = pixels.data() + static_cast<size_t>(y) * rowBytes;
            for (uint32_t x = 0; x < pw; ++x)
                convertPixel(row, x, p, &rgba[(static_cast<size_t>(y0 + y * dy) * width + x0 + x * dx) * 4]);
        }
    }
    return true;
}

}  // namespace

// ---- JPEG -------------------------------------------------------------------
// Baseline and progressive Huffman JPEG (ITU T.81), 8-bit samples, 1, 3 or
// 4 components, with stb_image's arithmetic after entropy decoding: its
// integer IDCT (the IJG "islow" factorisation, constants scaled by 4096),
// its centred "fancy" 2x upsampling (triangle filters with the same
// rounding, nearest neighbour for other ratios), its reduced-precision
// YCbCr->RGB (constants x4096 << 8, the green Cb term truncated to 16
// fractional bits) and its CMYK/YCCK handling under the Adobe marker.
namespace {

constexpr uint8_t kDezigzag[64 + 15] = {0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52,
    45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
    // runs past the end of a corrupt block land on the last coefficient
    63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct JpegHuffman {
    bool present = false;
    uint16_t fast[1 << 9] = {};  // (length << 8 | symbol) + 1 for codes of <= 9 bits
    int maxcode[18] = {};        // largest code of each length, left-aligned to 16 bits (exclusive bound)
    int valptr[17] = {};
    int mincode[17] = {};
    uint8_t symbols[256] = {};

    bool build(const uint8_t counts[16], const uint8_t* values, int total) {
        std::memcpy(symbols, values, static_cast<size_t>(total));
        std::fill(std::begin(fast), std::end(fast), 0);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            for (int i = 0; i < counts[len - 1]; ++i, ++k, ++code) {
                // A code of `len` bits must stay below 1 << len; a table with
                // more codes than that is over-subscribed, and its fast-table
                // span would run past the end of `fast`.
                if (code >= (1 << len)) return false;
                if (len <= 9) {
                    int first = code << (9 - len);
                    for (int f = 0; f < (1 << (9 - len)); ++f)
                        fast[first + f] = static_cast<uint16_t>(((len << 8) | symbols[k]) + 1);
                }
            }
            maxcode[len] = code << (16 - len);
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        present = true;
        return true;
    }
};

struct JpegComponent {
    int id = 0, h = 1, v = 1, tq = 0, hd = 0, ha = 0;
    int x = 0, y = 0, w2 = 0, h2 = 0;  // sample extent; padded plane size
    int dcPred = 0;
    std::vector<uint8_t> plane;
    std::vector<int16_t> coeff;  // progressive: 64 per block, natural order
    int coeffW = 0;              // blocks per coefficient row
};

class JpegDecoder {
public:
    JpegDecoder(const uint8_t* data, size_t size) : d_(data), n_(size) {}

    bool decode(std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height, std::string* error);

private:
    // --- byte and marker access
    int byte() { return pos_ < n_ ? d_[pos_++] : -1; }
    int be16() {
        int a = byte(), b = byte();
        return (a < 0 || b < 0) ? -1 : (a << 8) | b;
    }
    // --- entropy-coded bits (MSB first, 0xFF00 stuffing; a marker ends the data)
    void fill() {
        while (bits_ <= 24) {
            int c = 0;
            if (!hitMarker_) {
                c = byte();
                if (c < 0) {
                    c = 0;
                } else if (c == 0xff) {
                    int next = byte();
                    while (next == 0xff) next = byte();  // fill bytes
                    if (next != 0) {
                        marker_ = next;
                        hitMarker_ = true;
                        c = 0;
                    }
                }
            }
            buffer_ |= static_cast<uint32_t>(c) << (24 - bits_);
            bits_ += 8;
        }
    }
    int bit() {
        fill();
        int b = static_cast<int>(buffer_ >> 31);
        buffer_ <<= 1;
        --bits_;
        return b;
    }
    int getBits(int n) {
        if (n == 0) return 0;
        fill();
        int v = static_cast<int>(buffer_ >> (32 - n));
        buffer_ <<= n;
        bits_ -= n;
        return v;
    }
    int receiveExtend(int n) {
        if (n == 0) return 0;
        int v = getBits(n);
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }
    int huffDecode(const JpegHuffman& h) {
        fill();
        uint16_t entry = h.fast[buffer_ >> (32 - 9)];
        if (entry) {
            int len = (entry - 1) >> 8;
            buffer_ <<= len;
            bits_ -= len;
            return (entry - 1) & 0xff;
        }
        uint32_t top = buffer_ >> 16;
        int len = 10;
        while (len <= 16 && static_cast<int>(top) >= h.maxcode[len]) ++len;
        if (len > 16) return -1;
        int code = static_cast<int>(top >> (16 - len));
        int index = h.valptr[len] + code - h.mincode[len];
        if (index < 0 || index >= 256) return -1;
        buffer_ <<= len;
        bits_ -= len;
        return h.symbols[index];
    }
    void resetEntropy() {
        buffer_ = 0;
        bits_ = 0;
        hitMarker_ = false;
        marker_ = -1;
        eobRun_ = 0;
        for (JpegComponent& c : comps_) c.dcPred = 0;
    }

    bool readTables(int marker, std::string* error);
    bool readFrame(std::string* error);
    bool readScan(std::string* error);
    bool decodeScan(std::string* error);
    bool decodeBlock(JpegComponent& c, int16_t* out, std::string* error);
    bool progressiveBlock(JpegComponent& c, int16_t* data, std::string* error);
    void idct(const int16_t* in, uint8_t* out, int stride);
    void output(std::vector<uint8_t>& rgba);

    const uint8_t* d_;
    size_t n_, pos_ = 0;
    uint32_t buffer_ = 0;
    int bits_ = 0;
    bool hitMarker_ = false;
    int marker_ = -1;
    JpegHuffman dc_[4], ac_[4];
    uint16_t quant_[4][64] = {};  // natural order
    bool quantPresent_[4] = {};
    std::vector<JpegComponent> comps_;
    int width_ = 0, height_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
    bool progressive_ = false, jfif_ = false;
    int adobeTransform_ = -1;
    int restartInterval_ = 0;
    std::vector<int> scan_;  // component indices of the current scan
    int ss_ = 0, se_ = 63, ah_ = 0, al_ = 0, eobRun_ = 0;
};

bool JpegDecoder::readTables(int marker, std::string* error) {
    int length = be16();
    if (length < 2 || pos_ + static_cast<size_t>(length - 2) > n_) return fail(error, "JPEG: truncated segment");
    size_t end = pos_ + static_cast<size_t>(length - 2);
    if (marker == 0xdb) {  // DQT
        while (pos_ < end) {
            int pq = byte(), precision = pq >> 4, table = pq & 15;
            if (precision > 1 || table > 3) return fail(error, "JPEG: bad DQT");
            for (int i = 0; i < 64; ++i)
                quant_[table][kDezigzag[i]] = static_cast<uint16_t>(precision ? be16() : byte());
            quantPresent_[table] = true;
        }
    } else if (marker == 0xc4) {  // DHT
        while (pos_ < end) {
            int tc = byte(), kind = tc >> 4, table = tc & 15;
            if (kind > 1 || table > 3) return fail(error, "JPEG: bad DHT");
            uint8_t counts[16];
            int total = 0;
            for (int i = 0; i < 16; ++i) {
                counts[i] = static_cast<uint8_t>(byte());
                total += counts[i];
            }
            if (total > 256 || pos_ + static_cast<size_t>(total) > end) return fail(error, "JPEG: bad DHT");
            JpegHuffman& h = kind ? ac_[table] : dc_[table];
            if (!h.build(counts, d_ + pos_, total)) return fail(error, "JPEG: bad code lengths");
            pos_ += static_cast<size_t>(total);
        }
    } else if (marker == 0xdd) {  // DRI
        if (length != 4) return fail(error, "JPEG: bad DRI length");
        restartInterval_ = be16();
    } else if (marker == 0xe0 && length >= 7) {  // APP0: JFIF
        if (std::memcmp(d_ + pos_, "JFIF\0", 5) == 0) jfif_ = true;
    } else if (marker == 0xee && length >= 14) {  // APP14: Adobe
        if (std::memcmp(d_ + pos_, "Adobe", 6) == 0) adobeTransform_ = d_[pos_ + 11];
    }
    pos_ = end;
    return true;
}

bool JpegDecoder::readFrame(std::string* error) {
    int length = be16();
    int precision = byte();
    height_ = be16();
    width_ = be16();
    int count = byte();
    if (length < 8 || precision != 8) return fail(error, "JPEG: only 8-bit samples are supported");
    if (height_ <= 0 || width_ <= 0) return fail(error, "JPEG: zero-sized image");
    if (height_ > (1 << 24) || width_ > (1 << 24) || static_cast<uint64_t>(width_) * height_ > kMaxPixels)
        return fail(error, "JPEG: image too large");
    if (count != 1 && count != 3 && count != 4) return fail(error, "JPEG: bad component count");
    if (length != 8 + 3 * count) return fail(error, "JPEG: bad SOF length");
    comps_.resize(static_cast<size_t>(count));
    for (JpegComponent& c : comps_) {
        c.id = byte();
        int hv = byte();
        c.h = hv >> 4;
        c.v = hv & 15;
        c.tq = byte();
        if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq < 0 || c.tq > 3) return fail(error, "JPEG: bad component");
        hmax_ = std::max(hmax_, c.h);
        vmax_ = std::max(vmax_, c.v);
    }
    for (const JpegComponent& c : comps_)  // stb_image's check on sampling ratios
        if (hmax_ % c.h != 0 || vmax_ % c.v != 0) return fail(error, "JPEG: bad sampling factors");
    mcux_ = (width_ + hmax_ * 8 - 1) / (hmax_ * 8);
    mcuy_ = (height_ + vmax_ * 8 - 1) / (vmax_ * 8);
    for (JpegComponent& c : comps_) {
        c.x = (width_ * c.h + hmax_ - 1) / hmax_;
        c.y = (height_ * c.v + vmax_ - 1) / vmax_;
        c.w2 = mcux_ * c.h * 8;
        c.h2 = mcuy_ * c.v * 8;
        c.plane.assign(static_cast<size_t>(c.w2) * c.h2, 0);
        if (progressive_) {
            c.coeffW = c.w2 / 8;
            c.coeff.assign(static_cast<size_t>(c.w2) * c.h2, 0);
        }
    }
    return true;
}

bool JpegDecoder::readScan(std::string* error) {
    int length = be16();
    int count = byte();
    if (count < 1 || count > 4 || count > static_cast<int>(comps_.size()) || length != 6 + 2 * count)
        return fail(error, "JPEG: bad SOS");
    scan_.clear();
    for (int i = 0; i < count; ++i) {
        int id = byte(), tables = byte();
        int which = -1;
        for (size_t k = 0; k < comps_.size(); ++k)
            if (comps_[k].id == id) which = static_cast<int>(k);
        if (which < 0) return fail(error, "JPEG: scan names an unknown component");
        comps_[which].hd = tables >> 4;
        comps_[which].ha = tables & 15;
        if (comps_[which].hd > 3 || comps_[which].ha > 3) return fail(error, "JPEG: bad table selector");
        scan_.push_back(which);
    }
    ss_ = byte();
    se_ = byte();
    int a = byte();
    ah_ = a >> 4;
    al_ = a & 15;
    if (progressive_) {
        if (ss_ > 63 || se_ > 63 || ss_ > se_ || ah_ > 13 || al_ > 13) return fail(error, "JPEG: bad progressive scan");
        if (ss_ > 0 && count != 1) return fail(error, "JPEG: interleaved AC scan");
        if (ss_ == 0 && se_ != 0) return fail(error, "JPEG: DC and AC in one progressive scan");
    } else {
        if (ss_ != 0 || ah_ != 0 || al_ != 0) return fail(error, "JPEG: bad baseline scan");
        se_ = 63;
    }
    return true;
}

bool JpegDecoder::decodeBlock(JpegComponent& c, int16_t* data, std::string* error) {
    const JpegHuffman& dc = dc_[c.hd];
    const JpegHuffman& ac = ac_[c.ha];
    if (!dc.present || !ac.present) return fail(error, "JPEG: missing Huffman table");
    const uint16_t* q = quant_[c.tq];
    std::memset(data, 0, 64 * sizeof(int16_t));
    int t = huffDecode(dc);
    if (t < 0 || t > 15) return fail(error, "JPEG: bad Huffman code");
    c.dcPred += receiveExtend(t);
    data[0] = static_cast<int16_t>(c.dcPred * q[0]);
    for (int k = 1; k < 64;) {
        int rs = huffDecode(ac);
        if (rs < 0) return fail(error, "JPEG: bad Huffman code");
        int s = rs & 15, r = rs >> 4;
        if (s == 0) {
            if (rs != 0xf0) break;  // end of block
            k += 16;
        } else {
            k += r;
            int zig = kDezigzag[k++];
            data[zig] = static_cast<int16_t>(receiveExtend(s) * q[zig]);
        }
    }
    return true;
}

// One block of a progressive scan (T.81 G.1.2): DC first/refine, AC
// first/refine with end-of-band runs; coefficients stay unquantized.
bool JpegDecoder::progressiveBlock(JpegComponent& c, int16_t* data, std::string* error) {
    if (ss_ == 0) {
        if (ah_ == 0) {
            if (!dc_[c.hd].present) return fail(error, "JPEG: missing Huffman table");
            int t = huffDecode(dc_[c.hd]);
            if (t < 0 || t > 15) return fail(error, "JPEG: bad Huffman code");
            c.dcPred += receiveExtend(t);
            data[0] = static_cast<int16_t>(c.dcPred * (1 << al_));
        } else if (bit()) {
            data[0] = static_cast<int16_t>(data[0] + (1 << al_));
        }
        return true;
    }
    const JpegHuffman& ac = ac_[c.ha];
    if (!ac.present) return fail(error, "JPEG: missing Huffman table");
    if (ah_ == 0) {
        if (eobRun_) {
            --eobRun_;
            return true;
        }
        for (int k = ss_; k <= se_;) {
            int rs = huffDecode(ac);
            if (rs < 0) return fail(error, "JPEG: bad Huffman code");
            int s = rs & 15, r = rs >> 4;
            if (s == 0) {
                if (r < 15) {
                    eobRun_ = (1 << r) - 1 + getBits(r);
                    break;
                }
                k += 16;
            } else {
                k += r;
                int zig = kDezigzag[k++];
                data[zig] = static_cast<int16_t>(receiveExtend(s) * (1 << al_));
            }
        }
        return true;
    }
    // AC refinement: one correction bit per nonzero coefficient passed,
    // new coefficients of magnitude 1 << Al placed after r zero ones.
    const int16_t p1 = static_cast<int16_t>(1 << al_);
    auto refine = [&](int16_t& coef) {
        if (bit() && (coef & p1) == 0) coef = static_cast<int16_t>(coef > 0 ? coef + p1 : coef - p1);
    };
    int k = ss_;
    if (eobRun_ == 0) {
        for (; k <= se_;) {
            int rs = huffDecode(ac);
            if (rs < 0) return fail(error, "JPEG: bad Huffman code");
            int s = rs & 15, r = rs >> 4;
            int value = 0;
            if (s == 0) {
                if (r < 15) {
                    eobRun_ = (1 << r) + getBits(r);
                    break;
                }
            } else {
                if (s != 1) return fail(error, "JPEG: bad Huffman code");
                value = bit() ? p1 : -p1;
            }
            while (k <= se_) {
                int16_t& coef = data[kDezigzag[k++]];
                if (coef != 0) {
                    refine(coef);
                } else {
                    if (r == 0) {
                        if (value) coef = static_cast<int16_t>(value);
                        break;
                    }
                    --r;
                }
            }
        }
    }
    if (eobRun_ > 0) {
        for (; k <= se_; ++k) {
            int16_t& coef = data[kDezigzag[k]];
            if (coef != 0) refine(coef);
        }
        --eobRun_;
    }
    return true;
}

bool JpegDecoder::decodeScan(std::string* error) {
    resetEntropy();
    int todo = restartInterval_ ? restartInterval_ : 0x7fffffff;
    auto restart = [&]() -> bool {  // true: continue decoding
        if (--todo > 0) return true;
        fill();
        if (!(hitMarker_ && marker_ >= 0xd0 && marker_ <= 0xd7)) return false;
        resetEntropy();
        todo = restartInterval_ ? restartInterval_ : 0x7fffffff;
        return true;
    };
    int16_t block[64];
    if (scan_.size() == 1) {
        // Non-interleaved: the component's own blocks in raster order.
        JpegComponent& c = comps_[scan_[0]];
        int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
        for (int by = 0; by < bh; ++by)
            for (int bx = 0; bx < bw; ++bx) {
                if (progressive_) {
                    if (!progressiveBlock(c, &c.coeff[(static_cast<size_t>(by) * c.coeffW + bx) * 64], error)) return false;
                } else {
                    if (!decodeBlock(c, block, error)) return false;
                    idct(block, &c.plane[static_cast<size_t>(by) * 8 * c.w2 + bx * 8], c.w2);
                }
                if (!restart()) return true;
            }
        return true;
    }
    for (int my = 0; my < mcuy_; ++my)
        for (int mx = 0; mx < mcux_; ++mx) {
            for (int which : scan_) {
                JpegComponent& c = comps_[which];
                for (int v = 0; v < c.v; ++v)
                    for (int h = 0; h < c.h; ++h) {
                        int bx = mx * c.h + h, by = my * c.v + v;
                        if (progressive_) {
                            if (!progressiveBlock(c, &c.coeff[(static_cast<size_t>(by) * c.coeffW + bx) * 64], error))
                                return false;
                        } else {
                            if (!decodeBlock(c, block, error)) return false;
                            idct(block, &c.plane[static_cast<size_t>(by) * 8 * c.w2 + bx * 8], c.w2);
                        }
                    }
            }
            if (!restart()) return true;
        }
    return true;
}

inline uint8_t clampByte(int x) { return static_cast<uint8_t>(x < 0 ? 0 : (x > 255 ? 255 : x)); }
inline uint8_t clampWide(int64_t x) { return static_cast<uint8_t>(x < 0 ? 0 : (x > 255 ? 255 : x)); }

// The IJG islow IDCT as stb_image computes it: columns keep two extra bits
// ((x + 512) >> 10), rows remove 17 bits with rounding and the +128 level
// shift folded in.
void JpegDecoder::idct(const int16_t* in, uint8_t* out, int stride) {
    auto f2f = [](double x) { return static_cast<int64_t>(static_cast<int>(x * 4096 + 0.5)); };
    static const int64_t c0 = f2f(0.5411961f), c1 = f2f(-1.847759065f), c2 = f2f(0.765366865f), c3 = f2f(1.175875602f),
                     c4 = f2f(0.298631336f), c5 = f2f(2.053119869f), c6 = f2f(3.072711026f), c7 = f2f(1.501321110f),
                     c8 = f2f(-0.899976223f), c9 = f2f(-2.562915447f), c10 = f2f(-1.961570560f),
                     c11 = f2f(-0.390180644f);
    // One 1-D pass over s[0..7]: even part (x0..x3) and odd part (t0..t3).
    auto pass = [&](const int64_t s[8], int64_t x[4], int64_t t[4]) {
        int64_t p1 = (s[2] + s[6]) * c0;
        int64_t e2 = p1 + s[6] * c1, e3 = p1 + s[2] * c2;
        int64_t e0 = (s[0] + s[4]) * 4096, e1 = (s[0] - s[4]) * 4096;
        x[0] = e0 + e3;
        x[3] = e0 - e3;
        x[1] = e1 + e2;
        x[2] = e1 - e2;
        int64_t o0 = s[7], o1 = s[5], o2 = s[3], o3 = s[1];
        int64_t q3 = o0 + o2, q4 = o1 + o3, q1 = o0 + o3, q2 = o1 + o2;
        int64_t q5 = (q3 + q4) * c3;
        o0 *= c4;
        o1 *= c5;
        o2 *= c6;
        o3 *= c7;
        q1 = q5 + q1 * c8;
        q2 = q5 + q2 * c9;
        q3 *= c10;
        q4 *= c11;
        t[3] = o3 + q1 + q4;
        t[2] = o2 + q2 + q3;
        t[1] = o1 + q2 + q4;
        t[0] = o0 + q1 + q3;
    };
    int64_t tmp[64];
    for (int col = 0; col < 8; ++col) {
        int64_t s[8], x[4], t[4];
        for (int k = 0; k < 8; ++k) s[k] = in[k * 8 + col];
        pass(s, x, t);
        for (int64_t& v : x) v += 512;
        tmp[0 * 8 + col] = (x[0] + t[3]) >> 10;
        tmp[7 * 8 + col] = (x[0] - t[3]) >> 10;
        tmp[1 * 8 + col] = (x[1] + t[2]) >> 10;
        tmp[6 * 8 + col] = (x[1] - t[2]) >> 10;
        tmp[2 * 8 + col] = (x[2] + t[1]) >> 10;
        tmp[5 * 8 + col] = (x[2] - t[1]) >> 10;
        tmp[3 * 8 + col] = (x[3] + t[0]) >> 10;
        tmp[4 * 8 + col] = (x[3] - t[0]) >> 10;
    }
    for (int row = 0; row < 8; ++row) {
        int64_t x[4], t[4];
        pass(&tmp[row * 8], x, t);
        for (int64_t& v : x) v += 65536 + (128 << 17);
        uint8_t* o = out + static_cast<size_t>(row) * stride;
        o[0] = clampWide((x[0] + t[3]) >> 17);
        o[7] = clampWide((x[0] - t[3]) >> 17);
        o[1] = clampWide((x[1] + t[2]) >> 17);
        o[6] = clampWide((x[1] - t[2]) >> 17);
        o[2] = clampWide((x[2] + t[1]) >> 17);
        o[5] = clampWide((x[2] - t[1]) >> 17);
        o[3] = clampWide((x[3] + t[0]) >> 17);
        o[4] = clampWide((x[3] - t[0]) >> 17);
    }
}

void JpegDecoder::output(std::vector<uint8_t>& rgba) {
    const int n = static_cast<int>(comps_.size());
    struct Resample {
        int hs, vs, ystep, wLores, ypos;
        const uint8_t* line0;
        const uint8_t* line1;
        std::vector<uint8_t> buffer;
    };
    std::vector<Resample> rs(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        Resample& r = rs[k];
        r.hs = hmax_ / comps_[k].h;
        r.vs = vmax_ / comps_[k].v;
        r.ystep = r.vs >> 1;
        r.wLores = (width_ + r.hs - 1) / r.hs;
        r.ypos = 0;
        r.line0 = r.line1 = comps_[k].plane.data();
        r.buffer.assign(static_cast<size_t>(width_) + 3 + 8, 0);
    }
    bool isRgb = n == 3 && ((comps_[0].id == 'R' && comps_[1].id == 'G' && comps_[2].id == 'B') ||
                            (adobeTransform_ == 0 && !jfif_));
    rgba.assign(static_cast<size_t>(width_) * height_ * 4, 0);
    const uint8_t* row[4] = {};
    for (int j = 0; j < height_; ++j) {
        for (int k = 0; k < n; ++k) {
            Resample& r = rs[k];
            bool bottom = r.ystep >= (r.vs >> 1);
            const uint8_t* nearRow = bottom ? r.line1 : r.line0;
            const uint8_t* farRow = bottom ? r.line0 : r.line1;
            uint8_t* out = r.buffer.data();
            int w = r.wLores;
            if (r.hs == 1 && r.vs == 1) {
                row[k] = nearRow;
            } else if (r.hs == 1 && r.vs == 2) {
                for (int i = 0; i < w; ++i) out[i] = static_cast<uint8_t>((3 * nearRow[i] + farRow[i] + 2) >> 2);
                row[k] = out;
            } else if (r.hs == 2 && r.vs == 1) {
                if (w == 1) {
                    out[0] = out[1] = nearRow[0];
                } else {
                    out[0] = nearRow[0];
                    out[1] = static_cast<uint8_t>((nearRow[0] * 3 + nearRow[1] + 2) >> 2);
                    int i = 1;
                    for (; i < w - 1; ++i) {
                        int m = 3 * nearRow[i] + 2;
                        out[i * 2] = static_cast<uint8_t>((m + nearRow[i - 1]) >> 2);
                        out[i * 2 + 1] = static_cast<uint8_t>((m + nearRow[i + 1]) >> 2);
                    }
                    out[i * 2] = static_cast<uint8_t>((nearRow[w - 2] * 3 + nearRow[w - 1] + 2) >> 2);
                    out[i * 2 + 1] = nearRow[w - 1];
                }
                row[k] = out;
            } else if (r.hs == 2 && r.vs == 2) {
                if (w == 1) {
                    out[0] = out[1] = static_cast<uint8_t>((3 * nearRow[0] + farRow[0] + 2) >> 2);
                } else {
                    int t1 = 3 * nearRow[0] + farRow[0];
                    out[0] = static_cast<uint8_t>((t1 + 2) >> 2);
                    for (int i = 1; i < w; ++i) {
                        int t0 = t1;
                        t1 = 3 * nearRow[i] + farRow[i];
                        out[i * 2 - 1] = static_cast<uint8_t>((3 * t0 + t1 + 8) >> 4);
                        out[i * 2] = static_cast<uint8_t>((3 * t1 + t0 + 8) >> 4);
                    }
                    out[w * 2 - 1] = static_cast<uint8_t>((t1 + 2) >> 2);
                }
                row[k] = out;
            } else {
                for (int i = 0; i < w; ++i)
                    for (int h = 0; h < r.hs; ++h) out[i * r.hs + h] = nearRow[i];
                row[k] = out;
            }
            if (++r.ystep >= r.vs) {
                r.ystep = 0;
                r.line0 = r.line1;
                if (++r.ypos < comps_[k].y) r.line1 += comps_[k].w2;
            }
        }
        uint8_t* o = &rgba[static_cast<size_t>(j) * width_ * 4];
        auto ycc = [](int y, int cb, int cr, uint8_t* px) {
            auto fixed = [](float x) { return static_cast<int>(x * 4096.0f + 0.5f) << 8; };
            int yf = (y << 20) + (1 << 19);
            cr -= 128;
            cb -= 128;
            int r = yf + cr * fixed(1.40200f);
            int g = yf + (cr * -fixed(0.71414f)) + static_cast<int>(static_cast<unsigned>(cb * -fixed(0.34414f)) & 0xffff0000u);
            int b = yf + cb * fixed(1.77200f);
            px[0] = clampByte(r >> 20);
            px[1] = clampByte(g >> 20);
            px[2] = clampByte(b >> 20);
            px[3] = 255;
        };
        auto blinn = [](int x, int y) {
            unsigned t = static_cast<unsigned>(x * y + 128);
            return static_cast<uint8_t>((t + (t >> 8)) >> 8);
        };
        for (int i = 0; i < width_; ++i, o += 4) {
            if (n == 1) {
                o[0] = o[1] = o[2] = row[0][i];
                o[3] = 255;
            } else if (n == 3) {
                if (isRgb) {
                    o[0] = row[0][i];
                    o[1] = row[1][i];
                    o[2] = row[2][i];
                    o[3] = 255;
                } else {
                    ycc(row[0][i], row[1][i], row[2][i], o);
                }
            } else if (adobeTransform_ == 0) {  // CMYK
                int m = row[3][i];
                o[0] = blinn(row[0][i], m);
                o[1] = blinn(row[1][i], m);
                o[2] = blinn(row[2][i], m);
                o[3] = 255;
            } else if (adobeTransform_ == 2) {  // YCCK
                ycc(row[0][i], row[1][i], row[2][i], o);
                int m = row[3][i];
                o[0] = blinn(255 - o[0], m);
                o[1] = blinn(255 - o[1], m);
                o[2] = blinn(255 - o[2], m);
            } else {  // YCbCr and an ignored fourth channel
                ycc(row[0][i], row[1][i], row[2][i], o);
            }
        }
    }
}

bool JpegDecoder::decode(std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height, std::string* error) {
    if (byte() != 0xff || byte() != 0xd8) return fail(error, "JPEG: missing SOI");
    bool sawFrame = false, sawScan = false;
    int marker = -1;
    for (;;) {
        if (marker < 0) {  // next marker: skip to 0xFF, then past fill bytes
            int c = byte();
            while (c >= 0 && c != 0xff) c = byte();
            while (c == 0xff) c = byte();
            if (c < 0) break;
            marker = c;
        }
        int m = marker;
        marker = -1;
        if (m == 0xd9) break;  // EOI
        if (m >= 0xd0 && m <= 0xd7) continue;  // stray restart marker
        if (m == 0xc0 || m == 0xc1 || m == 0xc2) {
            if (sawFrame) return fail(error, "JPEG: more than one frame");
            progressive_ = m == 0xc2;
            if (!readFrame(error)) return false;
            sawFrame = true;
        } else if ((m >= 0xc3 && m <= 0xcf && m != 0xc4 && m != 0xc8 && m != 0xcc)) {
            return fail(error, "JPEG: lossless, hierarchical and arithmetic-coded JPEG are not supported");
        } else if (m == 0xda) {
            if (!sawFrame) return fail(error, "JPEG: scan before frame");
            if (!readScan(error) || !decodeScan(error)) return false;
            sawScan = true;
            if (hitMarker_) marker = marker_;
        } else if (m == 0xdc) {
            return fail(error, "JPEG: DNL is not supported");
        } else {
            if (!readTables(m, error)) return false;
        }
    }
    if (!sawFrame || !sawScan) return fail(error, "JPEG: no image data");
    if (progressive_) {
        for (JpegComponent& c : comps_) {
            int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
            for (int by = 0; by < bh; ++by)
                for (int bx = 0; bx < bw; ++bx) {
                    int16_t* data = &c.coeff[(static_cast<size_t>(by) * c.coeffW + bx) * 64];
                    for (int i = 0; i < 64; ++i) data[i] = static_cast<int16_t>(data[i] * quant_[c.tq][i]);
                    idct(data, &c.plane[static_cast<size_t>(by) * 8 * c.w2 + bx * 8], c.w2);
                }
        }
    }
    output(rgba);
    width = static_cast<uint32_t>(width_);
    height = static_cast<uint32_t>(height_);
    return true;
}

}  // namespace

bool decodeImageUnchecked(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width,
    uint32_t& height, std::string* error);

bool decodeImage(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height,
    std::string* error) {
    try {
        return decodeImageUnchecked(data, size, rgba, width, height, error);
    } catch (const std::exception& e) {  // allocation failure on a corrupt header
        rgba.clear();
        return fail(error, std::string("image decoding failed: ") + e.what());
    }
}

bool decodeImageUnchecked(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width,
    uint32_t& height, std::string* error) {
    static const uint8_t pngSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size >= 8 && std::memcmp(data, pngSignature, 8) == 0)
        return decodePng(data, size, rgba, width, height, error);
    if (size >= 3 && data[0] == 0xff && data[1] == 0xd8 && data[2] == 0xff) {
        JpegDecoder jpeg(data, size);
        return jpeg.decode(rgba, width, height, error);
    }
    return fail(error, "unknown image format (PNG or JPEG expected)");
}

}  // namespace wpt
