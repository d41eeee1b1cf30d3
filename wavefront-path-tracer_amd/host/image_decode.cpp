// PNG decoding to RGBA8 (see image_decode.hpp for the conversions).
#include "image_decode.hpp"

#include <array>
#include <cstring>

namespace wpt {
namespace {

bool fail(std::string* error, const std::string& message) {
    if (error) *error = message;
    return false;
}

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

bool inflateBlock(BitReader& in, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out,
    std::string* error) {
    for (;;) {
        int sym = lit.decode(in);
        if (sym < 0) return fail(error, "PNG: bad Huffman code");
        if (sym < 256) {
            out.push_back(static_cast<uint8_t>(sym));
            continue;
        }
        if (sym == 256) return true;
        sym -= 257;
        if (sym >= 29) return fail(error, "PNG: bad length symbol");
        size_t length = kLengthBase[sym] + in.take(kLengthExtra[sym]);
        int d = dist.decode(in);
        if (d < 0 || d >= 30) return fail(error, "PNG: bad distance code");
        size_t distance = kDistBase[d] + in.take(kDistExtra[d]);
        if (distance > out.size()) return fail(error, "PNG: distance before the start of the data");
        if (in.overrun()) return fail(error, "PNG: truncated zlib stream");
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
    uint8_t lengths[286 + 32] = {};
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
    out.reserve(expectedSize);
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
            out.insert(out.end(), data + 2 + at, data + 2 + at + len);
            in.seekByte(at + len);
        } else if (type == 1) {
            uint8_t lengths[288 + 32];
            std::memset(lengths, 8, 144);
            std::memset(lengths + 144, 9, 112);
            std::memset(lengths + 256, 7, 24);
            std::memset(lengths + 280, 8, 8);
            std::memset(lengths + 288, 5, 32);
            lit.build(lengths, 288);
            dist.build(lengths + 288, 32);
            if (!inflateBlock(in, lit, dist, out, error)) return false;
        } else if (type == 2) {
            if (!dynamicTables(in, lit, dist, error) || !inflateBlock(in, lit, dist, out, error)) return false;
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
            if (p.width > (1u << 24) || p.height > (1u << 24)) return fail(error, "PNG: image too large");
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

bool decodeImage(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height,
    std::string* error) {
    static const uint8_t pngSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size >= 8 && std::memcmp(data, pngSignature, 8) == 0)
        return decodePng(data, size, rgba, width, height, error);
    if (size >= 3 && data[0] == 0xff && data[1] == 0xd8 && data[2] == 0xff)
        return fail(error, "JPEG images are not decoded by this build; convert the texture to PNG");
    return fail(error, "unknown image format (PNG expected)");
}

}  // namespace wpt
